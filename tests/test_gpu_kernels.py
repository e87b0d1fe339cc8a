"""Parity of the HIP kernels with the CPU oracle (and golden vectors)."""

import numpy as np
import pytest
import torch

from dpmi import ops, systems

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _lib():
    from dpmi import _lib as L

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.load()


def T(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV, dtype=dtype)


def _clustered():
    """3000 atoms spread over a 40 A box plus 600 packed into a 2.5 A ball:
    the ball's atoms have more candidates than the cell-wise search keeps
    per atom, so they go through the per-atom path."""
    rng = np.random.default_rng(11)
    spread = rng.uniform(0, 40, (3000, 3))
    v = rng.normal(size=(600, 3))
    ball = 20 + v / np.linalg.norm(v, axis=1, keepdims=True) * 2.5 * rng.uniform(0, 1, (600, 1)) ** (1 / 3)
    coord = np.concatenate([spread, ball])
    atype = rng.integers(0, 2, coord.shape[0]).astype(np.int32)
    return systems.System("clustered", coord, atype, np.diag([40.0, 40.0, 40.0]).ravel())


def _lattice_at_rc(jitter: float):
    """Simple cubic lattice, spacing 3 A, 16^3 sites, rc = 6 A: every atom has
    six neighbors at exactly rc (excluded by the strict double test), and with
    a 1e-6 A jitter pairs fall on both sides of rc within float rounding."""
    rng = np.random.default_rng(13)
    g = np.arange(16) * 3.0
    c = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    c = c + jitter * rng.uniform(-1, 1, c.shape)
    atype = rng.integers(0, 2, c.shape[0]).astype(np.int32)
    return systems.System("lattice", c, atype, np.diag([48.0, 48.0, 48.0]).ravel())


NLIST_CASES = [
    ("water192", lambda: systems.water192(0), 6.0, [46, 92]),
    # neighbors at exactly rc, and within float rounding of it
    ("lattice_at_rc", lambda: _lattice_at_rc(0.0), 6.0, [40, 40]),
    ("lattice_jitter", lambda: _lattice_at_rc(1e-6), 6.0, [40, 40]),
    # neighborhoods larger than the cell-wise search's LDS candidate buffer
    ("water12k_rcut9", lambda: systems.water12k(1), 9.0, [110, 220]),
    ("clustered", _clustered, 6.0, [200, 200]),
    ("water192_tight_sel", lambda: systems.water192(2), 6.0, [20, 40]),
    ("water12k", lambda: systems.water12k(1), 6.0, [46, 92]),
    ("alloy_triclinic_4k", lambda: systems.alloy262k(4, ncell=16), 8.0, [60, 60, 60, 60]),
    ("water110k", lambda: systems.water110k(3), 6.0, [46, 92]),
]


@pytest.mark.parametrize("name,make,rcut,sel", NLIST_CASES, ids=[c[0] for c in NLIST_CASES])
def test_nlist_bit_exact(oracle, name, make, rcut, sel):
    s = make()
    _say(f"nlist {name} nloc={s.nloc}")
    ext_c, ext_t, mapping, nlist, ovf, nl_map = ops.build_nlist(T(s.coord), T(s.atype), s.box, rcut, sel,
                                                                 mapped=True)
    oc, ot, om, onl, oovf = oracle.build_nlist_pbc(s.coord, s.atype, s.box, rcut, sel)
    np.testing.assert_array_equal(ext_c.cpu().numpy(), oc)  # ghost layout and coordinates, bitwise
    np.testing.assert_array_equal(ext_t.cpu().numpy(), ot)
    np.testing.assert_array_equal(mapping.cpu().numpy(), om)
    np.testing.assert_array_equal(nlist.cpu().numpy(), onl)
    np.testing.assert_array_equal(ovf.cpu().numpy(), oovf)
    exp_map = np.where(onl >= 0, om[np.maximum(onl, 0)], -1)
    np.testing.assert_array_equal(nl_map.cpu().numpy(), exp_map)


def test_nlist_golden_fmt(golden):
    g = golden("fmt_nlist.json")
    for case, sel in (("TestFormatNlist", [10, 10]), ("TestFormatNlistShortSel", [2, 2])):
        _say(f"nlist golden {case}")
        c = g[case]
        posi = np.array(c["posi"]).reshape(-1, 3)
        atype = np.array(c["atype"], np.int32)
        box = np.diag([13.0, 13.0, 13.0]).ravel()
        _, _, _, nlist, ovf, _ = ops.build_nlist(T(posi), T(atype), box, 6.0, sel)
        np.testing.assert_array_equal(nlist.cpu().numpy().ravel(), np.array(c["expect_nlist_cpy"]))


def test_nlist_open_system(oracle):
    _say("nlist open system")
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 9, (300, 3))
    atype = rng.integers(0, 2, 300).astype(np.int32)
    sel = [30, 30]
    ext_c, ext_t, mapping, nlist, ovf, _ = ops.build_nlist(T(pos), T(atype), None, 3.0, sel)
    _, _, _, onl, oovf = oracle.build_nlist_nopbc(pos, atype, 3.0, sel)
    assert ext_c.shape[0] == 300
    np.testing.assert_array_equal(nlist.cpu().numpy(), onl)
    np.testing.assert_array_equal(ovf.cpu().numpy(), oovf)


def test_nlist_empty():
    _say("nlist empty")
    box = np.diag([20.0, 20.0, 20.0]).ravel()
    ext_c, ext_t, mapping, nlist, ovf, _ = ops.build_nlist(
        torch.zeros((0, 3), dtype=torch.float64, device=DEV), torch.zeros(0, dtype=torch.int32, device=DEV), box,
        6.0, [4, 4])
    assert nlist.shape == (0, 8) and ext_c.shape[0] == 0


def _env_inputs(oracle, s, rcut, sel, seed=0):
    oc, ot, om, onl, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, rcut, sel)
    nnei = sum(sel)
    rng = np.random.default_rng(seed)
    avg = rng.normal(size=(len(sel), nnei, 4)) * 0.05
    std = 0.5 + rng.uniform(size=(len(sel), nnei, 4))
    return oc, ot, om, onl, avg, std


@pytest.mark.parametrize("name,make,rcut,rs,sel", [
    ("water192", lambda: systems.water192(0), 6.0, 0.5, [46, 92]),
    ("alloy_triclinic_4k", lambda: systems.alloy262k(4, ncell=16), 8.0, 2.0, [60, 60, 60, 60]),
    # rows longer than 256 slots: the kernel's per-chunk path
    ("water192_rcut9", lambda: systems.water192(0), 9.0, 0.5, [110, 220]),
])
def test_env_mat_f64(oracle, name, make, rcut, rs, sel):
    s = make()
    _say(f"env_mat f64 {name}")
    oc, ot, om, onl, avg, std = _env_inputs(oracle, s, rcut, sel)
    em, ed, rij = ops.env_mat_a(T(onl), T(oc), T(ot), T(avg), T(std), rcut, rs, sel)
    oem, oed, orij = oracle.env_mat_a(onl, oc, ot, avg, std, rcut, rs, sel)
    np.testing.assert_allclose(em.cpu().numpy(), oem, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(ed.cpu().numpy(), oed, rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(rij.cpu().numpy(), orij)


def test_env_mat_f32(oracle):
    s = systems.water192(0)
    _say("env_mat f32 water192")
    oc, ot, om, onl, avg, std = _env_inputs(oracle, s, 6.0, [46, 92])
    em, ed, rij = ops.env_mat_a(T(onl), T(oc, torch.float32), T(ot), T(avg, torch.float32),
                                T(std, torch.float32), 6.0, 0.5, [46, 92])
    oem, oed, orij = oracle.env_mat_a(onl, oc, ot, avg, std, 6.0, 0.5, [46, 92])
    np.testing.assert_allclose(em.cpu().numpy(), oem, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(ed.cpu().numpy(), oed, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("name,make,rcut,rs,sel", [
    ("water192", lambda: systems.water192(0), 6.0, 0.5, [46, 92]),
    ("alloy_triclinic_4k", lambda: systems.alloy262k(4, ncell=16), 8.0, 2.0, [60, 60, 60, 60]),
    ("water192_rcut9", lambda: systems.water192(0), 9.0, 0.5, [110, 220]),
])
def test_env_mat_f32c64(oracle, name, make, rcut, rs, sel):
    """The precision form's env-mat (double coordinates, x_j - x_i in double,
    the rest in float): within a few float roundings of the FP64 oracle, em,
    its derivative and r_ij."""
    s = make()
    _say(f"env_mat f32c64 {name}")
    oc, ot, om, onl, avg, std = _env_inputs(oracle, s, rcut, sel)
    em, ed, rij = ops.env_mat_a(T(onl), T(oc), T(ot), T(avg, torch.float32), T(std, torch.float32), rcut, rs, sel,
                                dtype=torch.float32)
    oem, oed, orij = oracle.env_mat_a(onl, oc, ot, avg.astype(np.float32).astype(np.float64),
                                      std.astype(np.float32).astype(np.float64), rcut, rs, sel)
    for got, want in ((em, oem), (ed, oed), (rij, orij)):
        got = got.cpu().numpy().astype(np.float64).reshape(want.shape)
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= 1e-6, err
    # r_ij is the double difference rounded once
    np.testing.assert_array_equal(rij.cpu().numpy().reshape(orij.shape), orij.astype(np.float32))


def test_prod_force_virial_f64(oracle):
    s = systems.water192(0)
    _say("prod_force/prod_virial f64 water192")
    oc, ot, om, onl, avg, std = _env_inputs(oracle, s, 6.0, [46, 92])
    oem, oed, orij = oracle.env_mat_a(onl, oc, ot, avg, std, 6.0, 0.5, [46, 92])
    rng = np.random.default_rng(4)
    nd = rng.normal(size=(s.nloc, 138 * 4))
    nall = oc.shape[0]
    f = ops.prod_force(T(nd), T(oed), T(onl), s.nloc, nall)
    of = oracle.prod_force_a(nd, oed, onl, nall)
    np.testing.assert_allclose(f.cpu().numpy().reshape(of.shape), of, rtol=1e-10, atol=1e-11)
    v, av = ops.prod_virial(T(nd), T(oed), T(orij), T(onl), s.nloc, nall)
    ov, oav = oracle.prod_virial_a(nd, oed, orij, onl, nall)
    np.testing.assert_allclose(v.cpu().numpy().ravel(), ov, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(av.cpu().numpy().reshape(oav.shape), oav, rtol=1e-10, atol=1e-10)


def test_prod_force_golden(golden, oracle):
    g = golden("prod_force_a.json")["TestProdForceA"]
    _say("prod_force golden")
    posi = np.array(g["posi"]).reshape(-1, 3)
    atype = np.array(g["atype"], np.int32)
    box = np.diag([13.0] * 3).ravel()
    oc, ot, om, onl, _ = oracle.build_nlist_pbc(posi, atype, box, 6.0, [5, 5])
    oem, oed, orij = oracle.env_mat_a(onl, oc, ot, np.zeros((2, 10, 4)), np.ones((2, 10, 4)), 6.0, 0.8, [5, 5])
    nd = 10 - np.arange(6 * 40) * 0.01
    f = ops.prod_force(T(np.concatenate([nd, nd])), T(np.concatenate([oed.ravel(), oed.ravel()])),
                       T(np.concatenate([onl, onl])), 6, oc.shape[0], nframes=2)
    np.testing.assert_allclose(f.cpu().numpy().ravel(), np.array(g["expected_force"] * 2), atol=1e-5)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_force_virial_recompute(oracle, dtype):
    s = systems.water12k(1)
    _say(f"force/virial recompute {dtype} water12k")
    oc, ot, om, onl, avg, std = _env_inputs(oracle, s, 6.0, [46, 92])
    oem, oed, orij = oracle.env_mat_a(onl, oc, ot, avg, std, 6.0, 0.5, [46, 92])
    rng = np.random.default_rng(5)
    nd = rng.normal(size=(s.nloc, 138, 4))
    f, v = ops.force_virial_recompute(T(nd, dtype), T(onl), T(om), T(oc), T(ot), T(std, dtype), 6.0, 0.5, [46, 92])
    nall = oc.shape[0]
    of = oracle.prod_force_a(nd.reshape(s.nloc, -1), oed, onl, nall)[0]
    of_local = np.zeros((s.nloc, 3))
    np.add.at(of_local, om, of)
    ov, _ = oracle.prod_virial_a(nd.reshape(s.nloc, -1), oed, orij, onl, nall)
    tol = 1e-10 if dtype == torch.float64 else 2e-4
    scale = np.abs(of_local).max()
    np.testing.assert_allclose(f.cpu().numpy(), of_local, rtol=tol, atol=tol * scale)
    np.testing.assert_allclose(v.cpu().numpy(), ov, rtol=tol, atol=tol * np.abs(ov).max())
    # extended-atom variant: folding by mapping gives the local result
    fe, _ = ops.force_virial_recompute(T(nd, dtype), T(onl), None, T(oc), T(ot), T(std, dtype), 6.0, 0.5,
                                       [46, 92], want_virial=False)
    np.testing.assert_allclose(fe.cpu().numpy(), of, rtol=tol, atol=tol * scale)


CELLS_CASES = [
    # ncell = 2 per side: several images of one atom in a cell's table
    ("water192", lambda: systems.water192(0), 6.0, [46, 92]),
    ("water12k", lambda: systems.water12k(1), 6.0, [46, 92]),
    ("alloy_triclinic_4k", lambda: systems.alloy262k(4, ncell=16), 8.0, [60, 60, 60, 60]),
    # atoms and cells the search handles one by one: neighbor terms by atomics
    ("clustered", _clustered, 6.0, [200, 200]),
    ("water12k_rcut9", lambda: systems.water12k(1), 9.0, [110, 220]),
    # box side just above 8 rc: cells only just wider than rc search two
    # cells out (nsearch 2) while the ghost region is one cell deep, so some
    # neighbor-cell offsets of a cell's table hold no atoms
    ("box_just_over_8rc", lambda: systems._water_like("w48", 16, 48.0001, 1, True), 6.0, [46, 92]),
]


def _poisoned_cells_workspace(s, rcut, dtype):
    """Force workspace filled with 0xFF bytes (NaN in fp32 and fp64), so a
    partial slot read before it is written shows up as NaN forces."""
    import ctypes

    from dpmi import _lib

    n = ctypes.c_int64(0)
    box = np.ascontiguousarray(np.asarray(s.box, dtype=np.float64))
    _lib.call("dpmi_force_virial_cells_workspace_bytes", ctypes.byref(n), _lib.ptr(box), s.nloc, rcut,
              int(dtype == torch.float64))
    return torch.full((int(n.value),), 255, dtype=torch.uint8, device=DEV)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("name,make,rcut,sel", CELLS_CASES, ids=[c[0] for c in CELLS_CASES])
def test_force_virial_cells(oracle, name, make, rcut, sel, dtype):
    """Cell-wise adjoint (no neighbor atomics) against the per-atom recompute
    adjoint (itself pinned to the oracle above) on the GPU's own list, and for
    FP64 also against the oracle's prod_force/prod_virial."""
    s = make()
    _say(f"force/virial cells {name} {dtype}")
    ext_c, ext_t, mapping, nlist, ovf, _, cells = ops.build_nlist(T(s.coord), T(s.atype), s.box, rcut, sel,
                                                                  cells=True)
    nnei = int(sum(sel))
    rng = np.random.default_rng(7)
    std = 0.5 + rng.uniform(size=(len(sel), nnei, 4))
    nd = rng.normal(size=(s.nloc, nnei, 4))
    f0, v0 = ops.force_virial_recompute(T(nd, dtype), nlist, mapping, ext_c, ext_t, T(std, dtype), rcut, 0.5, sel)
    f1, v1 = ops.force_virial_cells(T(nd, dtype), nlist, cells, mapping, ext_c, ext_t, T(std, dtype), rcut, 0.5,
                                    sel, s.box, workspace=_poisoned_cells_workspace(s, rcut, dtype))
    f0, v0, f1, v1 = (x.cpu().numpy() for x in (f0, v0, f1, v1))
    assert np.isfinite(f1).all() and np.isfinite(v1).all()
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    np.testing.assert_allclose(f1, f0, rtol=tol, atol=tol * np.abs(f0).max())
    np.testing.assert_allclose(v1, v0, rtol=tol, atol=tol * np.abs(v0).max())
    if dtype == torch.float64 and name in ("water12k", "box_just_over_8rc"):
        oc, ot, om, onl, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, rcut, sel)
        avg = np.zeros_like(std)
        _, oed, orij = oracle.env_mat_a(onl, oc, ot, avg, std, rcut, 0.5, sel)
        of = oracle.prod_force_a(nd.reshape(s.nloc, -1), oed, onl, oc.shape[0])[0]
        of_local = np.zeros((s.nloc, 3))
        np.add.at(of_local, om, of)
        np.testing.assert_allclose(f1, of_local, rtol=1e-10, atol=1e-10 * np.abs(of_local).max())
    if name in ("clustered", "water12k_rcut9"):
        return  # atoms left to the per-atom search add their neighbor terms with atomics
    if (name.startswith("alloy") or name == "box_just_over_8rc") and dtype == torch.float64:
        # FP64 cells over the first table size (here: 5^3-cell neighborhoods
        # of the two-cell search) add their neighbor terms with atomics
        return
    # deterministic: a second call gives the same bits
    f2, v2 = ops.force_virial_cells(T(nd, dtype), nlist, cells, mapping, ext_c, ext_t, T(std, dtype), rcut, 0.5,
                                    sel, s.box)
    np.testing.assert_array_equal(f2.cpu().numpy(), f1)
    np.testing.assert_array_equal(v2.cpu().numpy(), v1)


def test_force_virial_cells_long_rows():
    """sum(sel) = 3600: the search goes atom by atom (its candidate tables do
    not fit in LDS for such rows) and the cell-wise force entry point must
    then take every neighbor term itself."""
    s = systems.water192(0)
    sel = [1800, 1800]
    _say("force/virial cells sum(sel)=3600")
    ext_c, ext_t, mapping, nlist, ovf, _, cells = ops.build_nlist(T(s.coord), T(s.atype), s.box, 6.0, sel,
                                                                  cells=True)
    assert (cells[0] == -1).all()
    rng = np.random.default_rng(3)
    std = 0.5 + rng.uniform(size=(2, 3600, 4))
    nd = rng.normal(size=(s.nloc, 3600, 4))
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        f0, v0 = ops.force_virial_recompute(T(nd, dtype), nlist, mapping, ext_c, ext_t, T(std, dtype), 6.0, 0.5, sel)
        f1, v1 = ops.force_virial_cells(T(nd, dtype), nlist, cells, mapping, ext_c, ext_t, T(std, dtype), 6.0, 0.5,
                                        sel, s.box, workspace=_poisoned_cells_workspace(s, 6.0, dtype))
        f0, v0, f1, v1 = (x.cpu().numpy() for x in (f0, v0, f1, v1))
        assert np.abs(f0).max() > 0
        np.testing.assert_allclose(f1, f0, rtol=tol, atol=tol * np.abs(f0).max())
        np.testing.assert_allclose(v1, v0, rtol=tol, atol=tol * np.abs(v0).max())


def test_build_nlist_rejects_foreign_workspace():
    """A neighbor-list workspace prepared for another system (other nloc) or
    never prepared is refused before any list is written."""
    import ctypes

    from dpmi import _lib

    a, b = systems.water192(0), systems.water12k(1)
    _say("build_nlist foreign workspace")
    sel = np.array([46, 92], dtype=np.int32)
    ext_c, ext_t, mapping, nlist, ovf, _, (cslot, ws) = ops.build_nlist(T(b.coord), T(b.atype), b.box, 6.0, sel,
                                                                         cells=True)
    a_ext, a_t, a_map, a_nl, _, _ = ops.build_nlist(T(a.coord), T(a.atype), a.box, 6.0, sel)
    out = torch.full((a.nloc, 138), -7, dtype=torch.int32, device=DEV)
    ovf_a = torch.empty(a.nloc, dtype=torch.int32, device=DEV)
    for wsp in (ws, torch.zeros_like(ws)):
        with pytest.raises(_lib.DpmiError, match="workspace"):
            _lib.call("dpmi_build_nlist_f64", _lib.ptr(out), _lib.ptr(ovf_a), None, None, _lib.ptr(a_ext),
                      _lib.ptr(a_t), _lib.ptr(a_map), a.nloc, a_ext.shape[0], 6.0, _lib.ptr(sel), 2, _lib.ptr(wsp),
                      wsp.numel(), ops._stream())
        assert (out == -7).all()


def test_force_virial_cells_open_system():
    _say("force/virial cells open system")
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 9, (300, 3))
    atype = rng.integers(0, 2, 300).astype(np.int32)
    sel = [30, 30]
    ext_c, ext_t, mapping, nlist, ovf, _, cells = ops.build_nlist(T(pos), T(atype), None, 3.0, sel, cells=True)
    std = 0.5 + rng.uniform(size=(2, 60, 4))
    nd = rng.normal(size=(300, 60, 4))
    f0, v0 = ops.force_virial_recompute(T(nd), nlist, None, ext_c, ext_t, T(std), 3.0, 0.5, sel)
    f1, v1 = ops.force_virial_cells(T(nd), nlist, cells, None, ext_c, ext_t, T(std), 3.0, 0.5, sel, None)
    np.testing.assert_allclose(f1.cpu().numpy(), f0.cpu().numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(v1.cpu().numpy(), v0.cpu().numpy(), rtol=1e-12, atol=1e-12)


def _tab(golden, pt=True):
    if pt:
        g = golden("tabulate_fusion_se_a_pt.json")
        return (np.array(g["table_tensor"]).reshape(8, 132), np.array(g["table_info_tensor"]),
                np.array(g["em_x_tensor"]).reshape(4, 4), np.array(g["em_tensor"]).reshape(4, 4, 4), g)
    g = golden("tabulate_se_a.json")["TestTabulateSeA"]
    return (np.array(g["table"]).reshape(8, 132), np.array(g["info"]), np.array(g["em_x"]).reshape(16, 1),
            np.array(g["em"]).reshape(4, 4, 4), g)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_tabulate_golden_op(golden, dtype, tol):
    """The reference op test (source/tests/pt/test_tabulate_fusion_se_a.py)
    through torch.ops.deepmd, forward and backward."""
    from dpmi import register_ops

    register_ops()
    table, info, em_x, em, g = _tab(golden)
    _say(f"tabulate golden {dtype}")
    em_x_t = T(em_x, dtype).requires_grad_(True)
    em_t = T(em, dtype).requires_grad_(True)
    out = torch.ops.deepmd.tabulate_fusion_se_a(T(table, dtype), torch.tensor(info, dtype=dtype), em_x_t, em_t, 8)[0]
    np.testing.assert_allclose(out.detach().cpu().numpy().ravel(), g["expected_descriptor_tensor"], rtol=tol, atol=tol)
    out.sum().backward()
    np.testing.assert_allclose(em_x_t.grad.cpu().numpy().ravel(), g["expected_dy_dem_x"], rtol=tol, atol=tol)
    np.testing.assert_allclose(em_t.grad.cpu().numpy().ravel(), g["expected_dy_dem"], rtol=tol, atol=tol)


def test_tabulate_cpp_golden(golden):
    table, info, em_x, em, g = _tab(golden, pt=False)
    _say("tabulate C++ golden")
    out = ops.tabulate_forward(T(table), torch.tensor(info), T(em_x), T(em), 8)
    np.testing.assert_allclose(out.cpu().numpy().ravel(), g["expected_xyz_scatter"], atol=1e-5)
    dx, dem = ops.tabulate_grad(T(table), torch.tensor(info), T(em_x), T(em), torch.ones((4, 4, 8), dtype=torch.float64,
                                                                                      device=DEV))
    np.testing.assert_allclose(dx.cpu().numpy().ravel(), g["expected_dy_dem_x"], atol=1e-5)
    np.testing.assert_allclose(dem.cpu().numpy().ravel(), g["expected_dy_dem"], atol=1e-5)


@pytest.mark.parametrize("M,tail", [(100, True), (128, False), (240, True), (8, True)])
def test_tabulate_vs_oracle(oracle, M, tail):
    _say(f"tabulate vs oracle M={M} tail={tail}")
    rng = np.random.default_rng(M)
    nloc, nnei = 301, 92
    info = np.array([0.0, 2.0, 10.0, 0.01, 0.1, -1.0])
    nspline = 200 + 80
    table = rng.normal(size=(nspline, 6 * M)) * np.array([1, 0.5, 0.2, 0.1, 0.05, 0.02] * M)
    em_x = rng.uniform(-0.5, 12.0, size=(nloc * nnei, 1))
    em = rng.normal(size=(nloc, nnei, 4))
    if tail:
        ex = em_x.reshape(nloc, nnei)
        ex[:, 60:] = ex[:, -1:]
        em[:, 60:, 1:] = 0.0
    dy = rng.normal(size=(nloc, 4, M))
    dzx = rng.normal(size=(nloc * nnei, 1))
    dze = rng.normal(size=(nloc, nnei, 4))
    ti = torch.tensor(info)
    out = ops.tabulate_forward(T(table), ti, T(em_x), T(em), M)
    np.testing.assert_allclose(out.cpu().numpy(), oracle.tabulate_fusion_se_a(table, info, em_x, em, M),
                               rtol=1e-12, atol=1e-12)
    dx, dem = ops.tabulate_grad(T(table), ti, T(em_x), T(em), T(dy))
    odx, odem = oracle.tabulate_fusion_se_a_grad(table, info, em_x, em, dy)
    np.testing.assert_allclose(dx.cpu().numpy().reshape(odx.shape), odx, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dem.cpu().numpy(), odem, rtol=1e-12, atol=1e-12)
    gg = ops.tabulate_grad_grad(T(table), ti, T(em_x), T(em), T(dzx), T(dze), M)
    ogg = oracle.tabulate_fusion_se_a_grad_grad(table, info, em_x, em, dzx, dze, M)
    np.testing.assert_allclose(gg.cpu().numpy(), ogg, rtol=1e-12, atol=1e-12)


def test_tabulate_second_order_fd(golden):
    """Double backward through the op vs central differences in dy
    (the reference check, source/tests/pt/tabulate_test_utils.py)."""
    from dpmi import register_ops

    register_ops()
    table, info, em_x, em, g = _tab(golden)
    _say("tabulate second order FD")
    em_x_t = T(em_x).requires_grad_(True)
    em_t = T(em).requires_grad_(True)
    tt = T(table)
    ti = torch.tensor(info)
    out = torch.ops.deepmd.tabulate_fusion_se_a(tt, ti, em_x_t, em_t, 8)[0]
    dy = torch.linspace(-0.5, 0.75, out.numel(), dtype=torch.float64, device=DEV).reshape(out.shape)
    dy.requires_grad_(True)
    g1 = torch.autograd.grad(out, (em_x_t, em_t), grad_outputs=dy, create_graph=True)
    w = [torch.linspace(-0.5, 0.75, x.numel(), dtype=torch.float64, device=DEV).reshape(x.shape) for x in g1]
    proj = sum((a * b).sum() for a, b in zip(g1, w))
    dz = torch.autograd.grad(proj, dy)[0]

    def project(dyv):
        gg = torch.autograd.grad(out, (em_x_t, em_t), grad_outputs=dyv, retain_graph=True)
        return sum((a * b).sum() for a, b in zip(gg, w)).item()

    eps = 1e-6
    flat = dy.detach().reshape(-1)
    for idx in range(0, flat.numel(), 7):
        p = flat.clone()
        m = flat.clone()
        p[idx] += eps
        m[idx] -= eps
        fd = (project(p.reshape(out.shape)) - project(m.reshape(out.shape))) / (2 * eps)
        assert abs(fd - dz.reshape(-1)[idx].item()) < 1e-6


def test_prod_env_mat_op_and_backward(oracle):
    """torch.ops.deepmd.prod_env_mat_a: forward vs oracle (list mapped to
    local indices) and the coordinate gradient vs central differences."""
    from dpmi import register_ops

    register_ops()
    s = systems.water192(0)
    _say("prod_env_mat_a op")
    sel = [46, 92]
    davg = torch.zeros((2, 138, 4), dtype=torch.float64, device=DEV)
    dstd = torch.ones((2, 138, 4), dtype=torch.float64, device=DEV)
    coord = T(s.coord[None]).requires_grad_(True)
    em, ed, rij, nl = torch.ops.deepmd.prod_env_mat_a(coord, T(s.atype[None]), T(s.box[None]), davg, dstd, 6.0, 0.5,
                                                      sel)
    oc, ot, om, onl, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, 6.0, sel)
    oem, _, _ = oracle.env_mat_a(onl, oc, ot, davg.cpu().numpy(), dstd.cpu().numpy(), 6.0, 0.5, sel)
    np.testing.assert_allclose(em.detach().cpu().numpy().ravel(), oem.ravel(), rtol=1e-12, atol=1e-13)
    np.testing.assert_array_equal(nl.cpu().numpy().ravel(), np.where(onl >= 0, om[np.maximum(onl, 0)], -1).ravel())
    w = torch.linspace(-1, 1, em.numel(), dtype=torch.float64, device=DEV).reshape(em.shape)
    (em * w).sum().backward()
    grad = coord.grad.cpu().numpy()[0]
    h = 1e-6
    for i, d in [(0, 0), (5, 1), (77, 2), (191, 0)]:
        cp = s.coord.copy()
        cm = s.coord.copy()
        cp[i, d] += h
        cm[i, d] -= h
        ep = torch.ops.deepmd.prod_env_mat_a(T(cp[None]), T(s.atype[None]), T(s.box[None]), davg, dstd, 6.0, 0.5,
                                             sel)[0]
        emm = torch.ops.deepmd.prod_env_mat_a(T(cm[None]), T(s.atype[None]), T(s.box[None]), davg, dstd, 6.0, 0.5,
                                              sel)[0]
        fd = ((ep - emm) * w).sum().item() / (2 * h)
        assert abs(fd - grad[i, d]) < 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("name,make,rcut,sel", [NLIST_CASES[0], NLIST_CASES[3], NLIST_CASES[-2]],
                         ids=[NLIST_CASES[i][0] for i in (0, 3, -2)])
@pytest.mark.parametrize("counts", [True, False], ids=["numneigh", "padded"])
def test_format_external_list_matches_oracle(oracle, name, make, rcut, sel, counts):
    """An external neighbor list (the LAMMPS / forward_lower input: the
    reference's candidate lists, each row shuffled, plus extra far atoms)
    formatted on the GPU equals format_nlist_cpu of the oracle, bitwise."""
    s = make()
    _say(f"format external list {name} counts={counts}")
    oc, ot, om, onl, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, rcut, sel)
    nn, cand, mem = oracle.candidates(oc, s.nloc, rcut, box=s.box)
    rng = np.random.default_rng(17)
    nall = oc.shape[0]
    extra = 6
    rows = np.full((s.nloc, mem + extra), -1, np.int32)
    counts_v = np.zeros(s.nloc, np.int32)
    for i in range(s.nloc):
        r = np.concatenate([cand[i, :nn[i]], rng.integers(0, nall, extra)])
        r = r[r != i]
        r = np.unique(r)  # distinct candidates
        rng.shuffle(r)
        rows[i, :len(r)] = r
        counts_v[i] = len(r)
    exp, eovf = oracle.format_nlist(oc, ot, s.nloc, counts_v, rows, rows.shape[1], rcut, sel)
    got, ovf, nmap = ops.format_nlist(T(rows), T(oc), T(ot), s.nloc, rcut, sel,
                                      numneigh=T(counts_v) if counts else None, mapping=T(om))
    np.testing.assert_array_equal(got.cpu().numpy(), exp)
    np.testing.assert_array_equal(ovf.cpu().numpy(), eovf)
    np.testing.assert_array_equal(nmap.cpu().numpy(), np.where(exp >= 0, om[np.maximum(exp, 0)], -1))
