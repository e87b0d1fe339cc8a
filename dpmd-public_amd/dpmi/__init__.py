"""dpmi: MI355X-native se_e2_a descriptor-to-force path (HIP kernels for
gfx950 behind a C ABI, with the deepmd op surface on top)."""

__version__ = "0.1.0"


def register_ops() -> None:
    """Register torch.ops.deepmd.{tabulate_fusion_se_a, prod_env_mat_a,
    prod_force_se_a, prod_virial_se_a} backed by libdpmi."""
    from . import ops

    ops.register()
