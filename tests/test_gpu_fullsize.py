"""Parity at the benchmark's full sizes: the 110,592-atom water box (env-mat
in every form against the oracle, bf16 / fp16 E+F+V against the FP64 GPU
path) and the 262,144-atom four-type alloy (bit-exact list, including the
atoms whose neighbors overflow sel)."""

import numpy as np
import pytest
import torch

from dpmi import ops, systems

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dpmi import _lib

    _lib.load()


@pytest.fixture(scope="module")
def water110k(oracle):
    s = systems.water110k()
    oc, ot, om, onl, oovf = oracle.build_nlist_pbc(s.coord, s.atype, s.box, 6.0, [46, 92])
    return s, oc, ot, om, onl, oovf


def test_water110k_nlist_bit_exact(water110k):
    s, oc, ot, om, onl, oovf = water110k
    _say("nlist water110k")
    ext_c, ext_t, mapping, nlist, ovf, _ = ops.build_nlist(torch.tensor(s.coord, device=DEV),
                                                           torch.tensor(s.atype, device=DEV), s.box, 6.0, [46, 92])
    np.testing.assert_array_equal(ext_c.cpu().numpy(), oc)
    np.testing.assert_array_equal(nlist.cpu().numpy(), onl)


def test_water110k_env_mat_all_forms(oracle, water110k):
    """f64 within 1e-10 of the oracle; f32c64 (pair difference in double,
    the rest in float: the bf16 step's env-mat) within a few float roundings
    of it, which float coordinates (ulp 7.6e-6 A at 100 A) cannot reach."""
    s, oc, ot, om, onl, _ = water110k
    _say("env_mat water110k f64 / f32c64 / f32")
    rng = np.random.default_rng(0)
    avg = np.zeros((2, 138, 4))
    avg[..., 0] = 0.05
    std = np.broadcast_to(0.5 + rng.uniform(size=(2, 1, 4)), (2, 138, 4)).copy()
    oem, oed, _ = oracle.env_mat_a(onl, oc, ot, avg, std, 6.0, 0.5, [46, 92])
    oem = oem.reshape(s.nloc, 138, 4)
    nl, c, t = (torch.as_tensor(x, device=DEV) for x in (onl, oc, ot))
    em, ed, _ = ops.env_mat_a(nl, c, t, torch.as_tensor(avg, device=DEV), torch.as_tensor(std, device=DEV), 6.0, 0.5,
                              [46, 92])
    np.testing.assert_allclose(em.cpu().numpy().reshape(oem.shape), oem, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(ed.cpu().numpy().reshape(oed.shape), oed, rtol=1e-10, atol=1e-12)
    a32, s32 = (torch.as_tensor(x, device=DEV, dtype=torch.float32) for x in (avg, std))
    scale = np.abs(oem).max()
    em_m, ed_m, _ = ops.env_mat_a(nl, c, t, a32, s32, 6.0, 0.5, [46, 92], dtype=torch.float32)
    err_m = np.abs(em_m.cpu().numpy().astype(np.float64).reshape(oem.shape) - oem).max() / scale
    em_f, _, _ = ops.env_mat_a(nl, c.float(), t, a32, s32, 6.0, 0.5, [46, 92])
    err_f = np.abs(em_f.cpu().numpy().astype(np.float64).reshape(oem.shape) - oem).max() / scale
    print(f"env-mat max error / max |em|: f32c64 {err_m:.2e}, all-float {err_f:.2e}", flush=True)
    assert err_m <= 1e-6
    assert err_f > 2 * err_m  # the difference formed in double is what keeps f32c64 accurate here
    oed_s = np.abs(oed).max()
    err_d = np.abs(ed_m.cpu().numpy().astype(np.float64).reshape(oed.shape) - oed).max() / oed_s
    assert err_d <= 2e-6


@pytest.mark.parametrize("precision,rmse_max,de_max", [("bf16", 2e-4, 2e-4), ("fp16", 5e-5, 5e-5)])
def test_water110k_low_precision_error_budget(precision, rmse_max, de_max):
    """bf16 / fp16 E+F+V at 110,592 atoms against the FP64 GPU path
    (itself within 1e-10 of the oracle at 192 atoms): force RMSE, energy per
    atom and virial, with budgets a few times the measured errors."""
    from dpmi.model import SeAEnergyModel
    from dpmi.stat import gpu_model_params

    spec = systems.WATER_SPEC
    _say(f"{precision} E+F+V water110k vs FP64")
    params = gpu_model_params(spec, systems.water192(0))
    s = systems.water110k()
    c = torch.tensor(s.coord, device=DEV)
    t = torch.tensor(s.atype, device=DEV)
    a = SeAEnergyModel.from_params(spec, params, device=DEV).compute(c, t, s.box)
    b = SeAEnergyModel.from_params(spec, params, device=DEV, precision=precision).compute(c, t, s.box)
    rmse = torch.sqrt(((a["force"] - b["force"]) ** 2).mean()).item()
    frms = torch.sqrt((a["force"] ** 2).mean()).item()
    de = abs(a["energy"].item() - b["energy"].item()) / s.nloc
    dv = (torch.abs(a["virial"] - b["virial"]).max() / torch.abs(a["virial"]).max()).item()
    print(f"water110k {precision}: force RMSE {rmse:.3e} eV/A ({rmse / frms:.2e} of RMS {frms:.3e}), "
          f"|dE|/atom {de:.3e} eV, virial max rel {dv:.2e}", flush=True)
    assert rmse <= rmse_max and rmse <= 5e-3 * frms
    assert de <= de_max
    assert dv <= 2e-3


def test_alloy262k_nlist_bit_exact_with_overflow(oracle):
    """262,144 atoms, rcut 8, sel 4 x 60: about 8.8k atoms list more than 60
    neighbors of some type; the nearest are kept (bitwise as the oracle),
    and overflow names the largest such type."""
    s = systems.alloy262k(4)
    _say("nlist alloy262k")
    sel = [60, 60, 60, 60]
    oc, ot, om, onl, oovf = oracle.build_nlist_pbc(s.coord, s.atype, s.box, 8.0, sel)
    assert (oovf >= 0).sum() > 1000
    ext_c, ext_t, mapping, nlist, ovf, _ = ops.build_nlist(torch.tensor(s.coord, device=DEV),
                                                           torch.tensor(s.atype, device=DEV), s.box, 8.0, sel)
    np.testing.assert_array_equal(ext_c.cpu().numpy(), oc)
    np.testing.assert_array_equal(mapping.cpu().numpy(), om)
    np.testing.assert_array_equal(nlist.cpu().numpy(), onl)
    np.testing.assert_array_equal(ovf.cpu().numpy(), oovf)


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_embedding_row_stream_many_windows(precision):
    """The embedding kernels on 884,736 atoms: the env-mat rows, section
    counts and types of the 110,592-atom water box tiled eight times, so each
    forward wavefront's atom range spans several 64-atom windows (one at
    110,592 atoms).  Each copy's G^T R partials equal the 110,592-atom
    launch's to the fp32 summation order (rows fall in other tiles); dE/d em
    equal bit for bit (a row's adjoint does not depend on the tile it shares
    with other atoms).  The backward splits its workgroups' atoms by rows:
    with the rows of 15 atoms in 16 dropped (counts 0), a wavefront's range
    holds about 16 times the atoms and spans several windows; the kept atoms'
    dE/d em still equal the full launch's bit for bit, the others are 0."""
    from dpmi import embed
    from dpmi.model import SeAEnergyModel
    from dpmi.stat import gpu_model_params

    spec = systems.WATER_SPEC
    _say(f"{precision} embedding row stream, water110k rows tiled 8x")
    params = gpu_model_params(spec, systems.water192(0))
    m = SeAEnergyModel.from_params(spec, params, device=DEV, precision=precision)
    s = systems.water110k()
    c = torch.tensor(s.coord, device=DEV)
    at = torch.tensor(s.atype, device=DEV).to(torch.int32)
    ext_c, ext_t, mapping, nlist, _, _, _ = m.neighbor_list(c, at, s.box)
    em = m.env_mat(nlist, ext_c, ext_t)
    counts, atype, pad_em = m._pad_args(nlist, at)
    reps = 8
    em8 = em.repeat(reps, 1, 1).contiguous()
    pad8 = (counts.repeat(reps, 1).contiguous(), atype.repeat(reps).contiguous(), pad_em)
    g1 = m.fused.forward_part(em, (counts, atype, pad_em))
    g8 = m.fused.forward_part(em8, pad8)
    gen = torch.Generator(device="cpu").manual_seed(11)
    dgr = (torch.randn(s.nloc, 128, 4, generator=gen) * 0.01).to(DEV)
    q1, q2, q3 = embed.internal_maps(m.fused.n1)
    mask = torch.zeros(128, dtype=torch.bool)
    mask[torch.as_tensor(q3)] = True
    dgr[:, ~mask.to(DEV), :] = 0
    nd1 = m.fused.backward_dgr(em, dgr, counts)
    nd8 = m.fused.backward_dgr(em8, dgr.repeat(reps, 1, 1).contiguous(), pad8[0])
    torch.cuda.synchronize()
    scale = g1.abs().max().item()
    gerr = max((g8[:, k * s.nloc:(k + 1) * s.nloc] - g1).abs().max().item() for k in range(reps)) / scale
    same = all(torch.equal(nd8[k * s.nloc:(k + 1) * s.nloc], nd1) for k in range(reps))
    print(f"{precision}: G^T R max rel difference over the copies {gerr:.2e}; dE/d em bitwise equal: {same}",
          flush=True)
    assert torch.isfinite(g8).all() and torch.isfinite(nd8).all()
    assert gerr <= 1e-5
    assert same
    keep = torch.arange(reps * s.nloc, device=DEV) % 16 == 0
    cs = torch.where(keep[:, None], pad8[0], torch.zeros_like(pad8[0]))
    nds = m.fused.backward_dgr(em8, dgr.repeat(reps, 1, 1).contiguous(), cs)
    torch.cuda.synchronize()
    kept = keep[: s.nloc]
    assert all(torch.equal(nds[k * s.nloc:(k + 1) * s.nloc][kept], nd1[kept]) for k in range(reps))
    assert (nds[~keep] == 0).all()
