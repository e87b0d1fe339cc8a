"""The oracle against the reference's own CPU functions, compiled from the
reference sources into oracle/_ref (skipped where that library is absent)."""

import ctypes

import numpy as np
import pytest

from dpmi import systems


@pytest.fixture(scope="module")
def ref(oracle):
    r = oracle.reflib()
    if r is None:
        pytest.skip("reference library oracle/_ref not built")
    return r


def P(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def PI(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _ref_pipeline(ref, coord, atype, box, rcut, rcut_smth, sel):
    nloc = len(atype)
    c = np.ascontiguousarray(coord, np.float64).copy()
    b = np.ascontiguousarray(box, np.float64)
    ref.ref_normalize_coord(P(c), nloc, P(b))
    mem = 1 << 20
    oc = np.zeros((mem, 3))
    ot = np.zeros(mem, np.int32)
    om = np.zeros(mem, np.int32)
    nall = ctypes.c_int(0)
    t = np.ascontiguousarray(atype, np.int32)
    assert ref.ref_copy_coord(P(oc), PI(ot), PI(om), ctypes.byref(nall), P(c), PI(t), nloc, mem, rcut, P(b)) == 0
    n = nall.value
    oc, ot, om = oc[:n].copy(), ot[:n].copy(), om[:n].copy()
    cm = 512
    nn = np.zeros(nloc, np.int32)
    cand = np.zeros((nloc, cm), np.int32)
    mx = ctypes.c_int(0)
    assert ref.ref_build_nlist(PI(nn), PI(cand), cm, ctypes.byref(mx), P(oc), nloc, n, rcut) == 0
    nnei = sum(sel)
    ntypes = len(sel)
    rng = np.random.default_rng(3)
    avg = rng.normal(size=(ntypes, nnei, 4)) * 0.1
    std = 0.5 + rng.uniform(size=(ntypes, nnei, 4))
    em = np.zeros((nloc, nnei, 4))
    ed = np.zeros((nloc, nnei, 4, 3))
    rij = np.zeros((nloc, nnei, 3))
    nlist = np.zeros((nloc, nnei), np.int32)
    s = np.array(sel, np.int32)
    ref.ref_prod_env_mat_a(P(em), P(ed), P(rij), PI(nlist), P(oc), PI(ot), nloc, n, PI(nn), PI(cand), cm,
                           P(avg), P(std), rcut, rcut_smth, PI(s), ntypes)
    return dict(norm=c, ext_coord=oc, ext_type=ot, mapping=om, nn=nn, cand=cand, nlist=nlist,
                em=em, ed=ed, rij=rij, avg=avg, std=std)


CASES = [
    ("water192", lambda: systems.water192(0), 6.0, 0.5, [46, 92]),
    ("alloy_small", lambda: systems.alloy262k(4, ncell=10), 8.0, 2.0, [60, 60, 60, 60]),
    ("water_tight_sel", lambda: systems.water192(2), 6.0, 0.5, [20, 40]),
]


@pytest.mark.parametrize("name,make,rcut,rs,sel", CASES)
def test_oracle_equals_reference_bitwise(oracle, ref, name, make, rcut, rs, sel):
    s = make()
    r = _ref_pipeline(ref, s.coord, s.atype, s.box, rcut, rs, sel)
    nc = oracle.normalize_coord(s.coord, s.box)
    np.testing.assert_array_equal(nc, r["norm"])
    ec, et, em_ = oracle.copy_coord(nc, s.atype, s.box, rcut)
    np.testing.assert_array_equal(ec, r["ext_coord"])
    np.testing.assert_array_equal(et, r["ext_type"])
    np.testing.assert_array_equal(em_, r["mapping"])
    nn, cand, mem = oracle.candidates(ec, s.nloc, rcut, brute=True)
    np.testing.assert_array_equal(nn, r["nn"])
    for i in range(s.nloc):
        np.testing.assert_array_equal(cand[i, : nn[i]], r["cand"][i, : nn[i]])
    # cell-list candidates give the same formatted list
    nn2, cand2, mem2 = oracle.candidates(ec, s.nloc, rcut, box=s.box)
    nlist, ovf = oracle.format_nlist(ec, et, s.nloc, nn2, cand2, mem2, rcut, sel)
    np.testing.assert_array_equal(nlist, r["nlist"])
    em, ed, rij = oracle.env_mat_a(nlist, ec, et, r["avg"], r["std"], rcut, rs, sel)
    np.testing.assert_array_equal(em, r["em"])
    np.testing.assert_array_equal(ed, r["ed"])
    np.testing.assert_array_equal(rij, r["rij"])

    # force / virial on the same inputs
    rng = np.random.default_rng(9)
    nd = rng.normal(size=(s.nloc, sum(sel) * 4))
    nall = ec.shape[0]
    f_ref = np.zeros((1, nall, 3))
    ref.ref_prod_force_a(P(f_ref), P(nd), P(ed), PI(nlist), s.nloc, nall, sum(sel), 1)
    f = oracle.prod_force_a(nd, ed, nlist, nall)
    np.testing.assert_array_equal(f, f_ref)
    v_ref = np.zeros(9)
    av_ref = np.zeros((nall, 9))
    ref.ref_prod_virial_a(P(v_ref), P(av_ref), P(nd), P(ed), P(rij), PI(nlist), s.nloc, nall, sum(sel))
    v, av = oracle.prod_virial_a(nd, ed, rij, nlist, nall)
    # the reference accumulates with omp atomics in thread order
    np.testing.assert_allclose(v, v_ref, rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(av, av_ref, rtol=1e-12, atol=1e-10)

    # their transposes (the training-step ops), same order of operations
    gf = rng.normal(size=(s.nloc, 3))
    gn_ref = np.zeros((s.nloc, sum(sel) * 4))
    ref.ref_prod_force_grad_a(P(gn_ref), P(gf), P(ed), PI(nlist), s.nloc, sum(sel), 1)
    np.testing.assert_array_equal(oracle.prod_force_grad_a(gf, ed, nlist), gn_ref)
    gv = rng.normal(size=9)
    ref.ref_prod_virial_grad_a(P(gn_ref), P(gv), P(ed), P(rij), PI(nlist), s.nloc, sum(sel))
    np.testing.assert_array_equal(oracle.prod_virial_grad_a(gv, ed, rij, nlist), gn_ref)


def _random_table(rng, nspline, M):
    return rng.normal(size=(nspline, 6 * M)) * np.array([1, 0.5, 0.2, 0.1, 0.05, 0.02] * M)


@pytest.mark.parametrize("sorted_tail", [True, False])
def test_tabulate_equals_reference_bitwise(oracle, ref, sorted_tail):
    rng = np.random.default_rng(11)
    nloc, nnei, M = 37, 46, 24
    info = np.array([0.0, 2.0, 10.0, 0.01, 0.1, -1.0])
    nspline = int((2.0 - 0.0) / 0.01) + int((10.0 - 2.0) / 0.1)
    table = _random_table(rng, nspline, M)
    em_x = rng.uniform(-0.5, 12.0, size=(nloc, nnei))
    em = rng.normal(size=(nloc, nnei, 4))
    if sorted_tail:
        # padded tail: identical em_x and zero direction, as after the nlist sort
        em_x[:, 30:] = em_x[:, -1:]
        em[:, 30:, 1:] = 0.0
    out_ref = np.zeros((nloc, 4, M))
    ref.ref_tabulate_fusion_se_a(P(out_ref), P(table), P(info), P(em_x), P(em), nloc, nnei, M)
    np.testing.assert_array_equal(oracle.tabulate_fusion_se_a(table, info, em_x, em, M), out_ref)
    dy = rng.normal(size=(nloc, 4, M))
    dx_ref = np.zeros((nloc, nnei))
    dem_ref = np.zeros((nloc, nnei, 4))
    ref.ref_tabulate_fusion_se_a_grad(P(dx_ref), P(dem_ref), P(table), P(info), P(em_x), P(em), P(dy), nloc, nnei, M)
    dx, dem = oracle.tabulate_fusion_se_a_grad(table, info, em_x, em, dy)
    np.testing.assert_array_equal(dx, dx_ref)
    np.testing.assert_array_equal(dem, dem_ref)
    dzx = rng.normal(size=(nloc, nnei))
    dze = rng.normal(size=(nloc, nnei, 4))
    gg_ref = np.zeros((nloc, 4, M))
    ref.ref_tabulate_fusion_se_a_grad_grad(P(gg_ref), P(table), P(info), P(em_x), P(em), P(dzx), P(dze), nloc, nnei, M)
    np.testing.assert_array_equal(oracle.tabulate_fusion_se_a_grad_grad(table, info, em_x, em, dzx, dze, M), gg_ref)
