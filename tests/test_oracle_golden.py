"""Pin the CPU oracle to the reference's own golden vectors and to
properties the mathematics fixes (finite differences, identities)."""

import numpy as np
import pytest


def _sorted_ghosts(c, t, m, nloc):
    """The reference sorts ghosts before comparing copy_coord output
    (source/lib/tests/test_coord.cc:121-158)."""
    keys = [tuple(np.round(c[i], 8)) + (t[i], m[i]) for i in range(nloc, len(t))]
    order = sorted(range(len(keys)), key=lambda k: keys[k])
    idx = list(range(nloc)) + [nloc + k for k in order]
    return c[idx], t[idx], m[idx]


def test_normalize_coord_golden(oracle, golden):
    g = golden("coord.json")["TestNormCoord"]
    for r in ("r0", "r1", "r2"):
        out = oracle.normalize_coord(np.array(g[r]).reshape(-1, 3), np.array(g["boxt"]))
        np.testing.assert_allclose(out.ravel(), g["posi"], atol=1e-12, rtol=0)


@pytest.mark.parametrize("case", ["TestCopyCoord", "TestCopyCoordMoreCell"])
def test_copy_coord_golden(oracle, golden, case):
    g = golden("coord.json")[case]
    posi = np.array(g["posi"]).reshape(-1, 3)
    nloc = posi.shape[0]
    atype = np.array(g["atype"][:nloc], np.int32)
    c, t, m = oracle.copy_coord(posi, atype, np.array(g["boxt"]), g["rc"])
    exp_c = np.array(g["_expected_posi_cpy"]).reshape(-1, 3)
    exp_t = np.array(g["_expected_atype_cpy"], np.int32)
    exp_m = np.array(g["_expected_mapping"], np.int32)
    assert len(t) == len(exp_t)
    c1, t1, m1 = _sorted_ghosts(c, t, m, nloc)
    c2, t2, m2 = _sorted_ghosts(exp_c, exp_t, exp_m, nloc)
    np.testing.assert_allclose(c1, c2, atol=1e-12, rtol=0)
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(m1, m2)


@pytest.mark.parametrize("case,sel", [("TestFormatNlist", [10, 10]), ("TestFormatNlistShortSel", [2, 2])])
@pytest.mark.parametrize("brute", [True, False])
def test_format_nlist_golden(oracle, golden, case, sel, brute):
    """Exact extended indices, including the ghost layout and -1 padding."""
    g = golden("fmt_nlist.json")[case]
    posi = np.array(g["posi"]).reshape(-1, 3)
    atype = np.array(g["atype"], np.int32)
    box = np.diag([13.0, 13.0, 13.0]).ravel()
    ec, et, em, nlist, ovf = oracle.build_nlist_pbc(posi, atype, box, 6.0, sel, brute=brute)
    np.testing.assert_array_equal(nlist.ravel(), np.array(g["expect_nlist_cpy"]))
    if case == "TestFormatNlistShortSel":
        assert (ovf == 1).all()  # format_nlist_i_cpu returns the overflowing type
    else:
        assert (ovf == -1).all()


def _envmat_case(oracle, golden, case, sel):
    g = golden("env_mat_a.json")[case]
    posi = np.array(g["posi"]).reshape(-1, 3)
    atype = np.array(g["atype"], np.int32)
    box = np.diag([13.0, 13.0, 13.0]).ravel()
    ec, et, emap, nlist, _ = oracle.build_nlist_pbc(posi, atype, box, 6.0, sel)
    nnei = sum(sel)
    avg = np.zeros((2, nnei, 4))
    std = np.ones((2, nnei, 4))
    em, ed, rij = oracle.env_mat_a(nlist, ec, et, avg, std, 6.0, 0.8, sel)
    return g, ec, et, nlist, em, ed, rij, avg, std


@pytest.mark.parametrize("case,sel", [("TestEnvMatA", [10, 10]), ("TestEnvMatAShortSel", [2, 2])])
def test_env_mat_golden(oracle, golden, case, sel):
    g, ec, et, nlist, em, ed, rij, avg, std = _envmat_case(oracle, golden, case, sel)
    np.testing.assert_allclose(em.ravel(), g["expected_env"], atol=1e-5, rtol=0)


def test_env_mat_deriv_finite_difference(oracle, golden):
    """em_deriv is d em / d r_i (the centre), i.e. -d em / d r_ij
    (reference check: test_env_mat_a.cc:372-415)."""
    g, ec, et, nlist, em, ed, rij, avg, std = _envmat_case(oracle, golden, "TestEnvMatA", [10, 10])
    h = 1e-6
    nloc = nlist.shape[0]
    for i in range(nloc):
        for d in range(3):
            cp = ec.copy()
            cm = ec.copy()
            cp[i, d] += h
            cm[i, d] -= h
            # only the centre moves; neighbours (ghost images of other atoms) stay
            emp, _, _ = oracle.env_mat_a(nlist, cp, et, avg, std, 6.0, 0.8, [10, 10])
            emm, _, _ = oracle.env_mat_a(nlist, cm, et, avg, std, 6.0, 0.8, [10, 10])
            fd = (emp[i] - emm[i]) / (2 * h)
            np.testing.assert_allclose(ed[i, :, :, d], fd, atol=1e-6, rtol=0)


def test_env_mat_invariants(oracle):
    """Rotation: comps 1..3 rotate as vectors, comp 0 invariant; the radial
    part equals sw(r)/r with the quintic switch."""
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 4, (12, 3))
    atype = (np.arange(12) % 2).astype(np.int32)
    sel = [8, 8]
    ec, et, m, nlist, _ = oracle.build_nlist_nopbc(pos, atype, 3.5, sel)
    avg = np.zeros((2, 16, 4))
    std = np.ones((2, 16, 4))
    em, ed, rij = oracle.env_mat_a(nlist, ec, et, avg, std, 3.5, 1.0, sel)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    em2, _, _ = oracle.env_mat_a(nlist, ec @ R.T, et, avg, std, 3.5, 1.0, sel)
    np.testing.assert_allclose(em2[..., 0], em[..., 0], atol=1e-13)
    np.testing.assert_allclose(em2[..., 1:], em[..., 1:] @ R.T, atol=1e-13)
    r = np.linalg.norm(rij, axis=-1)
    valid = nlist >= 0
    u = np.clip((r - 1.0) / 2.5, 0, 1)
    sw = u**3 * (-6 * u**2 + 15 * u - 10) + 1
    np.testing.assert_allclose(em[..., 0][valid], (sw / np.where(valid, r, 1))[valid], rtol=1e-12, atol=1e-14)


def _tab_inputs(golden, pt=False):
    if pt:
        g = golden("tabulate_fusion_se_a_pt.json")
        return (np.array(g["table_tensor"]), np.array(g["table_info_tensor"]), np.array(g["em_x_tensor"]),
                np.array(g["em_tensor"]).reshape(4, 4, 4), g, 8)
    g = golden("tabulate_se_a.json")["TestTabulateSeA"]
    return np.array(g["table"]), np.array(g["info"]), np.array(g["em_x"]), np.array(g["em"]).reshape(4, 4, 4), g, 8


def test_tabulate_golden_cpp(oracle, golden):
    table, info, em_x, em, g, M = _tab_inputs(golden)
    out = oracle.tabulate_fusion_se_a(table, info, em_x, em, M)
    np.testing.assert_allclose(out.ravel(), g["expected_xyz_scatter"], atol=1e-5, rtol=0)
    dx, dem = oracle.tabulate_fusion_se_a_grad(table, info, em_x, em, np.ones((4, 4, M)))
    np.testing.assert_allclose(dx.ravel(), g["expected_dy_dem_x"], atol=1e-5, rtol=0)
    np.testing.assert_allclose(dem.ravel(), g["expected_dy_dem"], atol=1e-5, rtol=0)


def test_tabulate_golden_pt(oracle, golden):
    """PT op test arrays at the op test's fp64 tolerance (1e-10)."""
    table, info, em_x, em, g, M = _tab_inputs(golden, pt=True)
    out = oracle.tabulate_fusion_se_a(table, info, em_x, em, M)
    np.testing.assert_allclose(out.ravel(), g["expected_descriptor_tensor"], atol=1e-10, rtol=1e-10)
    dx, dem = oracle.tabulate_fusion_se_a_grad(table, info, em_x, em, np.ones((4, 4, M)))
    np.testing.assert_allclose(dx.ravel(), g["expected_dy_dem_x"], atol=1e-10, rtol=1e-10)
    np.testing.assert_allclose(dem.ravel(), g["expected_dy_dem"], atol=1e-10, rtol=1e-10)


def test_tabulate_grad_finite_difference(oracle, golden):
    """grad = d<dy, out>/d(em_x, em) away from the early-exit tail."""
    table, info, em_x, em, g, M = _tab_inputs(golden, pt=True)
    rng = np.random.default_rng(0)
    dy = rng.normal(size=(4, 4, M))
    dx, dem = oracle.tabulate_fusion_se_a_grad(table, info, em_x, em, dy, is_sorted=False)
    h = 1e-7

    def f(ex, e):
        return np.sum(dy * oracle.tabulate_fusion_se_a(table, info, ex, e, M, is_sorted=False))

    ex = em_x.reshape(4, 4)
    for i in range(4):
        for j in range(4):
            p = ex.copy()
            m = ex.copy()
            p[i, j] += h
            m[i, j] -= h
            fd = (f(p, em) - f(m, em)) / (2 * h)
            assert abs(fd - dx[i, j]) < 1e-6 * max(1, abs(fd))
            for c in range(4):
                p = em.copy()
                m = em.copy()
                p[i, j, c] += h
                m[i, j, c] -= h
                fd = (f(ex, p) - f(ex, m)) / (2 * h)
                assert abs(fd - dem[i, j, c]) < 1e-6 * max(1, abs(fd))


def test_tabulate_grad_grad_is_adjoint(oracle, golden):
    """grad_grad(dz) = d<dz, grad(dy)>/d dy; grad is linear in dy, so a
    central difference in dy is exact up to rounding."""
    table, info, em_x, em, g, M = _tab_inputs(golden, pt=True)
    rng = np.random.default_rng(1)
    dz_x = rng.normal(size=(4, 4))
    dz_e = rng.normal(size=(4, 4, 4))
    gg = oracle.tabulate_fusion_se_a_grad_grad(table, info, em_x, em, dz_x, dz_e, M)
    dy0 = rng.normal(size=(4, 4, M))
    for idx in [(0, 0, 0), (1, 2, 3), (3, 3, 7), (2, 1, 5)]:
        e = np.zeros_like(dy0)
        e[idx] = 1.0
        dxp, demp = oracle.tabulate_fusion_se_a_grad(table, info, em_x, em, dy0 + e)
        dxm, demm = oracle.tabulate_fusion_se_a_grad(table, info, em_x, em, dy0 - e)
        fd = (np.sum(dz_x * (dxp - dxm)) + np.sum(dz_e * (demp - demm))) / 2
        assert abs(fd - gg[idx]) < 1e-10 * max(1, abs(fd))


def _force_case(oracle, golden, name):
    g = golden(f"{name}.json")["TestProdForceA" if name == "prod_force_a" else "TestProdVirialA"]
    posi = np.array(g["posi"]).reshape(-1, 3)
    atype = np.array(g["atype"], np.int32)
    box = np.diag([13.0, 13.0, 13.0]).ravel()
    sel = [5, 5]
    ec, et, emap, nlist, _ = oracle.build_nlist_pbc(posi, atype, box, 6.0, sel)
    avg = np.zeros((2, 10, 4))
    std = np.ones((2, 10, 4))
    em, ed, rij = oracle.env_mat_a(nlist, ec, et, avg, std, 6.0, 0.8, sel)
    nloc = 6
    net_deriv = 10 - np.arange(nloc * 40) * 0.01
    return g, ec, nlist, ed, rij, net_deriv


def test_prod_force_golden(oracle, golden):
    g, ec, nlist, ed, rij, nd = _force_case(oracle, golden, "prod_force_a")
    nall = ec.shape[0]
    # two identical frames, as the reference test does (test_prod_force_a.cc:108-113)
    f = oracle.prod_force_a(np.concatenate([nd, nd]), np.concatenate([ed.ravel(), ed.ravel()]),
                            np.concatenate([nlist, nlist]), nall, nframes=2)
    np.testing.assert_allclose(f.ravel(), g["expected_force"] * 2, atol=1e-5, rtol=0)
    # no external field: forces on all (local + ghost) atoms sum to zero
    np.testing.assert_allclose(f.sum(axis=1), 0, atol=1e-12)


def test_prod_virial_golden(oracle, golden):
    g, ec, nlist, ed, rij, nd = _force_case(oracle, golden, "prod_virial_a")
    nall = ec.shape[0]
    v, av = oracle.prod_virial_a(nd, ed, rij, nlist, nall)
    np.testing.assert_allclose(v, g["expected_virial"], atol=1e-5, rtol=0)
    np.testing.assert_allclose(av.ravel(), g["expected_atom_virial"], atol=1e-5, rtol=0)
    np.testing.assert_allclose(av.sum(axis=0), v, atol=1e-10)


def _grad_case(oracle, golden, name):
    cls = "TestProdForceGradA" if name == "prod_force_grad_a" else "TestProdVirialGradA"
    g = golden(f"{name}.json")[cls]
    posi = np.array(g["posi"]).reshape(-1, 3)
    atype = np.array(g["atype"], np.int32)
    box = np.diag([13.0, 13.0, 13.0]).ravel()
    sel = [5, 5]
    ec, et, _, nlist, _ = oracle.build_nlist_pbc(posi, atype, box, g["rc"], sel)
    em, ed, rij = oracle.env_mat_a(nlist, ec, et, np.zeros((2, 10, 4)), np.ones((2, 10, 4)), g["rc"],
                                   g["rc_smth"], sel)
    return g, ec, nlist, ed, rij


def test_prod_force_grad_golden(oracle, golden):
    g, ec, nlist, ed, rij = _grad_case(oracle, golden, "prod_force_grad_a")
    nloc = 6
    # grad[i] = 10 - 0.1 i, two identical frames (test_prod_force_grad_a.cc:116-124)
    grad = 10 - np.arange(nloc * 3) * 0.1
    gn = oracle.prod_force_grad_a(np.concatenate([grad, grad]), np.concatenate([ed.ravel(), ed.ravel()]),
                                  np.concatenate([nlist, nlist]), nframes=2)
    np.testing.assert_allclose(gn.ravel(), g["expected_grad_net"] * 2, atol=1e-5, rtol=0)


def test_prod_virial_grad_golden(oracle, golden):
    g, ec, nlist, ed, rij = _grad_case(oracle, golden, "prod_virial_grad_a")
    grad = 10 - np.arange(9) * 1.0  # test_prod_virial_grad_a.cc:116-119
    gn = oracle.prod_virial_grad_a(grad, ed, rij, nlist)
    np.testing.assert_allclose(gn.ravel(), g["expected_grad_net"], atol=1e-5, rtol=0)


def test_prod_grads_are_adjoints(oracle):
    """<g, prod_force(nd)> = <prod_force_grad(g), nd> and likewise for the
    virial, on a mapped (local-index) list as the training step uses."""
    from dpmi import systems

    s = systems.water192(3)
    sel = [46, 92]
    ec, et, mp, nlist, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, 6.0, sel)
    em, ed, rij = oracle.env_mat_a(nlist, ec, et, np.zeros((2, 138, 4)), np.ones((2, 138, 4)), 6.0, 0.5, sel)
    nlm = np.where(nlist >= 0, mp[np.maximum(nlist, 0)], -1).astype(np.int32)
    rng = np.random.default_rng(2)
    nd = rng.normal(size=(s.nloc, 4 * 138))
    gf = rng.normal(size=(s.nloc, 3))
    f = oracle.prod_force_a(nd, ed, nlm, s.nloc)[0]
    lhs = np.sum(gf * f)
    rhs = np.sum(oracle.prod_force_grad_a(gf, ed, nlm) * nd)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    gv = rng.normal(size=9)
    v, _ = oracle.prod_virial_a(nd, ed, rij, nlm, s.nloc)
    lhs = np.sum(gv * v)
    rhs = np.sum(oracle.prod_virial_grad_a(gv, ed, rij, nlm) * nd)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_cell_candidates_equal_brute(oracle):
    """The O(N) candidate search finds exactly the brute-force set."""
    from dpmi import systems

    s = systems.water192(0)
    nc = oracle.normalize_coord(s.coord, s.box)
    ec, et, em = oracle.copy_coord(nc, s.atype, s.box, 6.0)
    nn1, c1, m1 = oracle.candidates(ec, s.nloc, 6.0, box=s.box)
    nn2, c2, m2 = oracle.candidates(ec, s.nloc, 6.0, brute=True)
    np.testing.assert_array_equal(nn1, nn2)
    for i in range(s.nloc):
        assert sorted(c1[i, : nn1[i]]) == sorted(c2[i, : nn2[i]])


def test_triclinic_cell_candidates_equal_brute(oracle):
    from dpmi import systems

    s = systems.alloy262k(4, ncell=8)
    nc = oracle.normalize_coord(s.coord, s.box)
    ec, et, em = oracle.copy_coord(nc, s.atype, s.box, 8.0)
    nn1, c1, _ = oracle.candidates(ec, s.nloc, 8.0, box=s.box)
    nn2, c2, _ = oracle.candidates(ec, s.nloc, 8.0, brute=True)
    np.testing.assert_array_equal(nn1, nn2)
    for i in range(0, s.nloc, 7):
        assert sorted(c1[i, : nn1[i]]) == sorted(c2[i, : nn2[i]])
