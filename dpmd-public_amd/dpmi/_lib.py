"""ctypes binding of libdpmi.so (the HIP kernels behind include/dpmi.h).

The library is built in-tree (``make -C dpmd-public_amd``) into
``dpmi/_lib/libdpmi.so``.  There is no fallback: when the library is
missing, or no gfx950 device is visible, every op raises.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libdpmi.so"

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_F = ctypes.c_float
_D = ctypes.c_double

# name -> argtypes (every entry point returns int status)
SIGNATURES = {
    "dpmi_nlist_workspace_bytes": [_P, _P, _L, _F],
    "dpmi_copy_coord_f64": [_P, _P, _P, _P, _L, _P, _P, _L, _P, _F, _P, _L, _P],
    "dpmi_build_nlist_f64": [_P, _P, _P, _P, _P, _P, _P, _L, _L, _F, _P, _I, _P, _L, _P],
    "dpmi_copy_coord_capacity": [_P, _P, _L, _F],
    "dpmi_copy_coord_nowait_f64": [_P, _P, _P, _L, _P, _P, _L, _P, _F, _P, _L, _P],
    "dpmi_build_nlist_nowait_f64": [_P] * 7 + [_L, _F, _P, _I, _P, _L, _P],
    "dpmi_nlist_status": [_P, _P, _L, _P],
    "dpmi_prod_env_mat_a_f64": [_P, _P, _P, _P, _P, _P, _L, _P, _P, _F, _F, _P, _I, _P],
    "dpmi_prod_env_mat_a_f32": [_P, _P, _P, _P, _P, _P, _L, _P, _P, _F, _F, _P, _I, _P],
    "dpmi_prod_env_mat_a_f32c64": [_P, _P, _P, _P, _P, _P, _L, _P, _P, _F, _F, _P, _I, _P],
    "dpmi_env_mat_pt_f64": [_P] * 8 + [_L, _L, _L, _L, _I, _I, _D, _D, _D, _I, _P],
    "dpmi_env_mat_pt_f32": [_P] * 8 + [_L, _L, _L, _L, _I, _I, _D, _D, _D, _I, _P],
    "dpmi_env_mat_pt_grad_f64": [_P] * 8 + [_L, _L, _L, _L, _I, _I, _D, _D, _D, _I, _P],
    "dpmi_env_mat_pt_grad_f32": [_P] * 8 + [_L, _L, _L, _L, _I, _I, _D, _D, _D, _I, _P],
    "dpmi_env_mat_pt_grad_grad_f64": [_P] * 8 + [_L, _L, _L, _L, _I, _I, _D, _D, _D, _I, _P],
    "dpmi_env_mat_pt_grad_grad_f32": [_P] * 8 + [_L, _L, _L, _L, _I, _I, _D, _D, _D, _I, _P],
    "dpmi_prod_force_se_a_f64": [_P, _P, _P, _P, _L, _L, _L, _L, _P],
    "dpmi_prod_force_se_a_f32": [_P, _P, _P, _P, _L, _L, _L, _L, _P],
    "dpmi_prod_virial_workspace_bytes": [_P, _L, _L],
    "dpmi_prod_virial_se_a_f64": [_P, _P, _P, _P, _P, _P, _L, _L, _L, _L, _P, _L, _P],
    "dpmi_prod_virial_se_a_f32": [_P, _P, _P, _P, _P, _P, _L, _L, _L, _L, _P, _L, _P],
    "dpmi_prod_virial_se_a_noalloc_f64": [_P, _P, _P, _P, _P, _P, _L, _L, _L, _L, _P],
    "dpmi_format_nlist_f64": [_P, _P, _P, _P, _P, _L, _P, _P, _P, _L, _L, _F, _P, _I, _P, _L, _P],
    "dpmi_prod_force_grad_a_f64": [_P, _P, _P, _P, _L, _L, _L, _L, _P],
    "dpmi_prod_force_grad_a_f32": [_P, _P, _P, _P, _L, _L, _L, _L, _P],
    "dpmi_prod_virial_grad_a_f64": [_P, _P, _P, _P, _P, _L, _L, _L, _P],
    "dpmi_prod_virial_grad_a_f32": [_P, _P, _P, _P, _P, _L, _L, _L, _P],
    "dpmi_force_virial_recompute_f64": [_P, _P, _P, _P, _P, _P, _P, _L, _L, _P, _F, _F, _P, _I, _P, _L, _P],
    "dpmi_force_virial_recompute_f32": [_P, _P, _P, _P, _P, _P, _P, _L, _L, _P, _F, _F, _P, _I, _P, _L, _P],
    "dpmi_force_virial_cells_workspace_bytes": [_P, _P, _L, _F, _I],
    "dpmi_force_virial_cells_f64": [_P] * 8 + [_L, _L, _P, _F, _F, _P, _I, _P, _P, _P, _L, _P],
    "dpmi_force_virial_cells_f32": [_P] * 8 + [_L, _L, _P, _F, _F, _P, _I, _P, _P, _P, _L, _P],
    "dpmi_se_a_embed_fwd_bf16": [_P, _P, _P, _P, _L, _P, _I, _I, _P, _P, _P, _P],
    "dpmi_se_a_embed_fwd_f16": [_P, _P, _P, _P, _L, _P, _I, _I, _P, _P, _P, _P],
    "dpmi_se_a_descriptor_fwd": [_P, _P, _P, _L, _I, _I, _I, _I, _I, _P],
    "dpmi_se_a_descriptor_bwd": [_P, _P, _P, _L, _I, _I, _I, _P],
    "dpmi_se_a_embed_bwd_bf16": [_P, _P, _P, _P, _P, _L, _P, _I, _I, _P, _P],
    "dpmi_se_a_embed_bwd_f16": [_P, _P, _P, _P, _P, _L, _P, _I, _I, _P, _P],
    "dpmi_se_a_section_counts": [_P, _P, _L, _P, _I, _P],
    "dpmi_se_a_embed_fwd_f64": [_P, _P, _P, _P, _P, _L, _P, _I, _I, _P, _P, _P, _P],
    "dpmi_se_a_embed_bwd_f64": [_P, _P, _P, _P, _P, _P, _L, _P, _I, _I, _P, _P],
    "dpmi_se_a_fitting_bf16": [_P, _P, _P, _P, _P, _P, _P, _P, _L, _I, _I, _P],
    "dpmi_se_a_fitting_f16": [_P, _P, _P, _P, _P, _P, _P, _P, _L, _I, _I, _P],
    "dpmi_se_a_fitting_gr_bf16": [_P, _P, _P, _P, _P, _P, _P, _P, _L, _I, _I, _I, _P],
    "dpmi_se_a_fitting_gr_f16": [_P, _P, _P, _P, _P, _P, _P, _P, _L, _I, _I, _I, _P],
    "dpmi_se_a_fitting_frag_count": [_P, _I],
    "dpmi_se_a_fitting_gr_f64": [_P, _P, _P, _I, _P, _P, _P, _P, _P, _L, _I, _I, _I, _I, _P],
    "dpmi_se_a_fitting_f64_frag_doubles": [_P, _I, _I],
    "dpmi_energy_sum_f64": [_P, _P, _L, _P],
    "dpmi_tabulate_build_f64": [_P, _P, _P, _P, _L, _P, _P, _I, _P],
    "dpmi_lr_tick": [_P, _P, _P, _P],
    "dpmi_adam_f64": [_P, _P, _P, _P, _L, _P, _D, _D, _D, _D, _P],
    "dpmi_adam_f32": [_P, _P, _P, _P, _L, _P, _D, _D, _D, _D, _P],
    "dpmi_row_sum_f64": [_P, _P, _P, _L, _L, _P],
    "dpmi_row_sum_f32": [_P, _P, _P, _L, _L, _P],
    "dpmi_tabulate_fusion_se_a_f64": [_P, _P, _P, _P, _P, _L, _L, _L, _I, _P],
    "dpmi_tabulate_fusion_se_a_f32": [_P, _P, _P, _P, _P, _L, _L, _L, _I, _P],
    "dpmi_tabulate_fusion_se_a_grad_f64": [_P, _P, _P, _P, _P, _P, _P, _L, _L, _L, _I, _P],
    "dpmi_tabulate_fusion_se_a_grad_f32": [_P, _P, _P, _P, _P, _P, _P, _L, _L, _L, _I, _P],
    "dpmi_tabulate_fusion_se_a_grad_grad_f64": [_P, _P, _P, _P, _P, _P, _P, _L, _L, _L, _I, _P],
    "dpmi_tabulate_fusion_se_a_grad_grad_f32": [_P, _P, _P, _P, _P, _P, _P, _L, _L, _L, _I, _P],
}

_lib = None


class DpmiError(RuntimeError):
    pass


def load():
    """Load libdpmi.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise DpmiError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()' or make -C dpmd-public_amd)"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.dpmi_last_error.restype = ctypes.c_char_p
        lib.dpmi_last_error.argtypes = []
        lib.dpmi_abi_version.restype = _I
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _I
        _lib = lib
    return _lib


def exported_symbols() -> list:
    return ["dpmi_last_error", "dpmi_abi_version", *SIGNATURES.keys()]


def call(name: str, *args) -> None:
    rc = getattr(load(), name)(*args)
    if rc != 0:
        msg = load().dpmi_last_error().decode(errors="replace")
        if rc == 3:
            raise DpmiCapacityError(f"{name}: {msg}")
        raise DpmiError(f"{name} failed ({rc}): {msg}")


class DpmiCapacityError(DpmiError):
    pass


def ptr(t) -> ctypes.c_void_p:
    """Device (or host) address of a contiguous tensor / numpy array."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)
