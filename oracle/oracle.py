"""ctypes front end of the CPU oracle (test infrastructure only).

Loads ``oracle/_build/liboracle.so`` (the plain-C restatement of the reference
CPU semantics, ``dpmd_oracle.c``) and, when it was built, the reference's own
CPU library compiled from its sources, ``oracle/_ref/libdpref.so``.  Only the
tests, the smoke check and the CPU-baseline leg of the benchmark import this
module; the product package never does.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
REF_PATH = HERE / "_ref" / "libdpref.so"

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int
_f = ctypes.c_float
_l = ctypes.c_long

_lib = None
_ref = None


def build(ref: bool | None = None) -> None:
    """Compile the oracle (and the reference library when its sources exist)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if ref is None:
        ref = Path("/root/reference/source/lib/src").is_dir()
    if ref:
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def _p(a: np.ndarray, kind):
    return a.ctypes.data_as(kind)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build(ref=False)
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_normalize_coord.argtypes = [_dp, _i, _dp]
        L.oracle_cell_grid.argtypes = [_dp, _f, _ip, _ip]
        L.oracle_copy_coord.argtypes = [_dp, _ip, _ip, _l, _dp, _ip, _i, _f, _dp]
        L.oracle_copy_coord.restype = _l
        L.oracle_build_nlist_brute.argtypes = [_ip, _ip, _i, _ip, _dp, _i, _i, _f]
        L.oracle_build_nlist_cell.argtypes = [_ip, _ip, _i, _ip, _dp, _i, _i, _f, _dp]
        L.oracle_format_nlist.argtypes = [_ip, _ip, _dp, _ip, _i, _ip, _ip, _i, _f, _ip, _i]
        L.oracle_env_mat_a.argtypes = [_dp, _dp, _dp, _ip, _dp, _ip, _i, _dp, _dp, _f, _f, _ip, _i]
        L.oracle_tabulate_fusion_se_a.argtypes = [_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _i]
        L.oracle_tabulate_fusion_se_a_grad.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _i, _i, _i]
        L.oracle_tabulate_fusion_se_a_grad_grad.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _i, _i, _i]
        L.oracle_prod_force_a.argtypes = [_dp, _dp, _dp, _ip, _i, _i, _i, _i]
        L.oracle_prod_virial_a.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, _i, _i, _i]
        L.oracle_prod_force_grad_a.argtypes = [_dp, _dp, _dp, _ip, _i, _i, _i]
        L.oracle_prod_virial_grad_a.argtypes = [_dp, _dp, _dp, _dp, _ip, _i, _i]
        L.oracle_num_threads.restype = _i
        _lib = L
    return _lib


def ref_available() -> bool:
    return REF_PATH.exists()


def reflib():
    """The reference's own CPU functions (None when not built)."""
    global _ref
    if _ref is None and REF_PATH.exists():
        R = ctypes.CDLL(str(REF_PATH))
        R.ref_normalize_coord.argtypes = [_dp, _i, _dp]
        R.ref_copy_coord.argtypes = [_dp, _ip, _ip, _ip, _dp, _ip, _i, _i, _f, _dp]
        R.ref_build_nlist.argtypes = [_ip, _ip, _i, _ip, _dp, _i, _i, _f]
        R.ref_prod_env_mat_a.argtypes = [_dp, _dp, _dp, _ip, _dp, _ip, _i, _i, _ip, _ip, _i, _dp, _dp, _f, _f, _ip, _i]
        R.ref_tabulate_fusion_se_a.argtypes = [_dp, _dp, _dp, _dp, _dp, _i, _i, _i]
        R.ref_tabulate_fusion_se_a_f32.argtypes = [_fp, _fp, _fp, _fp, _fp, _i, _i, _i]
        R.ref_tabulate_fusion_se_a_grad.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _i, _i]
        R.ref_tabulate_fusion_se_a_grad_grad.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _i, _i]
        R.ref_prod_force_a.argtypes = [_dp, _dp, _dp, _ip, _i, _i, _i, _i]
        R.ref_prod_virial_a.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, _i, _i, _i]
        _ref = R
    return _ref


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- geometry
def normalize_coord(coord: np.ndarray, box: np.ndarray) -> np.ndarray:
    c = _c(coord, np.float64).copy()
    b = _c(box, np.float64)
    lib().oracle_normalize_coord(_p(c, _dp), c.shape[0], _p(b, _dp))
    return c


def cell_grid(box, rcut: float):
    nc = np.zeros(3, np.int32)
    ng = np.zeros(3, np.int32)
    b = _c(box, np.float64)
    lib().oracle_cell_grid(_p(b, _dp), rcut, _p(nc, _ip), _p(ng, _ip))
    return nc, ng


def copy_coord(coord: np.ndarray, atype: np.ndarray, box: np.ndarray, rcut: float):
    """Ghost construction of the reference; coord must already be normalised."""
    c = _c(coord, np.float64)
    t = _c(atype, np.int32)
    b = _c(box, np.float64)
    nloc = t.shape[0]
    cap = 0
    dummy_c = np.zeros(3, np.float64)
    dummy_i = np.zeros(1, np.int32)
    need = lib().oracle_copy_coord(_p(dummy_c, _dp), _p(dummy_i, _ip), _p(dummy_i, _ip), cap,
                                   _p(c, _dp), _p(t, _ip), nloc, rcut, _p(b, _dp))
    nall = -need if need < 0 else need
    oc = np.zeros((max(nall, 1), 3), np.float64)
    ot = np.zeros(max(nall, 1), np.int32)
    om = np.zeros(max(nall, 1), np.int32)
    got = lib().oracle_copy_coord(_p(oc, _dp), _p(ot, _ip), _p(om, _ip), nall,
                                  _p(c, _dp), _p(t, _ip), nloc, rcut, _p(b, _dp))
    assert got == nall
    return oc[:nall], ot[:nall], om[:nall]


def candidates(ext_coord: np.ndarray, nloc: int, rcut: float, box=None, brute: bool = False):
    """Candidate lists [nloc, mem] (unsorted) + counts, reference semantics."""
    c = _c(ext_coord, np.float64)
    nall = c.shape[0]
    mem = 256
    while True:
        nn = np.zeros(nloc, np.int32)
        cand = np.full((nloc, mem), -1, np.int32)
        mx = np.zeros(1, np.int32)
        if brute or box is None:
            r = lib().oracle_build_nlist_brute(_p(nn, _ip), _p(cand, _ip), mem, _p(mx, _ip),
                                               _p(c, _dp), nloc, nall, rcut)
        else:
            b = _c(box, np.float64)
            r = lib().oracle_build_nlist_cell(_p(nn, _ip), _p(cand, _ip), mem, _p(mx, _ip),
                                              _p(c, _dp), nloc, nall, rcut, _p(b, _dp))
        if r == 0:
            return nn, cand, mem
        mem = int(mx[0]) + 16


def format_nlist(ext_coord, ext_type, nloc, nn, cand, mem, rcut, sel):
    c = _c(ext_coord, np.float64)
    t = _c(ext_type, np.int32)
    sec = np.concatenate([[0], np.cumsum(sel)]).astype(np.int32)
    nnei = int(sec[-1])
    nlist = np.zeros((nloc, nnei), np.int32)
    ovf = np.zeros(nloc, np.int32)
    lib().oracle_format_nlist(_p(nlist, _ip), _p(ovf, _ip), _p(c, _dp), _p(t, _ip), nloc,
                              _p(_c(nn, np.int32), _ip), _p(_c(cand, np.int32), _ip), mem,
                              rcut, _p(sec, _ip), len(sel))
    return nlist, ovf


def build_nlist_pbc(coord, atype, box, rcut, sel, brute: bool = False):
    """normalize -> copy_coord -> candidates -> format, all reference semantics.

    Returns (ext_coord, ext_type, mapping, nlist[nloc, nnei] in extended
    indices, overflow[nloc])."""
    nc = normalize_coord(coord, box)
    ec, et, em = copy_coord(nc, atype, box, rcut)
    nloc = len(atype)
    nn, cand, mem = candidates(ec, nloc, rcut, box=box, brute=brute)
    nlist, ovf = format_nlist(ec, et, nloc, nn, cand, mem, rcut, sel)
    return ec, et, em, nlist, ovf


def build_nlist_nopbc(coord, atype, rcut, sel):
    c = _c(coord, np.float64)
    t = _c(atype, np.int32)
    nloc = len(t)
    nn, cand, mem = candidates(c, nloc, rcut, box=None, brute=True)
    nlist, ovf = format_nlist(c, t, nloc, nn, cand, mem, rcut, sel)
    return c, t, np.arange(nloc, dtype=np.int32), nlist, ovf


# ---------------------------------------------------------------- env-mat
def env_mat_a(nlist, ext_coord, ext_type, avg, std, rcut, rcut_smth, sel):
    nloc, nnei = nlist.shape
    sec = np.concatenate([[0], np.cumsum(sel)]).astype(np.int32)
    em = np.zeros((nloc, nnei, 4), np.float64)
    ed = np.zeros((nloc, nnei, 4, 3), np.float64)
    rij = np.zeros((nloc, nnei, 3), np.float64)
    a = _c(avg, np.float64)
    s = _c(std, np.float64)
    lib().oracle_env_mat_a(_p(em, _dp), _p(ed, _dp), _p(rij, _dp), _p(_c(nlist, np.int32), _ip),
                           _p(_c(ext_coord, np.float64), _dp), _p(_c(ext_type, np.int32), _ip),
                           nloc, _p(a, _dp), _p(s, _dp), rcut, rcut_smth, _p(sec, _ip), len(sel))
    return em, ed, rij


# ---------------------------------------------------------------- tabulate
def tabulate_fusion_se_a(table, info, em_x, em, M, is_sorted=True):
    em = _c(em, np.float64)
    nloc, nnei = em.shape[0], em.shape[1]
    out = np.zeros((nloc, 4, M), np.float64)
    lib().oracle_tabulate_fusion_se_a(_p(out, _dp), _p(_c(table, np.float64), _dp),
                                      _p(_c(info, np.float64), _dp), _p(_c(em_x, np.float64), _dp),
                                      _p(em, _dp), nloc, nnei, M, int(is_sorted))
    return out


def tabulate_fusion_se_a_grad(table, info, em_x, em, dy, is_sorted=True):
    em = _c(em, np.float64)
    dy = _c(dy, np.float64)
    nloc, nnei = em.shape[0], em.shape[1]
    M = dy.shape[-1]
    dx = np.zeros((nloc, nnei), np.float64)
    dem = np.zeros((nloc, nnei, 4), np.float64)
    lib().oracle_tabulate_fusion_se_a_grad(_p(dx, _dp), _p(dem, _dp), _p(_c(table, np.float64), _dp),
                                           _p(_c(info, np.float64), _dp),
                                           _p(_c(em_x, np.float64), _dp), _p(em, _dp), _p(dy, _dp),
                                           nloc, nnei, M, int(is_sorted))
    return dx, dem


def tabulate_fusion_se_a_grad_grad(table, info, em_x, em, dz_dem_x, dz_dem, M, is_sorted=True):
    em = _c(em, np.float64)
    nloc, nnei = em.shape[0], em.shape[1]
    out = np.zeros((nloc, 4, M), np.float64)
    lib().oracle_tabulate_fusion_se_a_grad_grad(_p(out, _dp), _p(_c(table, np.float64), _dp),
                                                _p(_c(info, np.float64), _dp),
                                                _p(_c(em_x, np.float64), _dp), _p(em, _dp),
                                                _p(_c(dz_dem_x, np.float64), _dp),
                                                _p(_c(dz_dem, np.float64), _dp), nloc, nnei, M,
                                                int(is_sorted))
    return out


# ---------------------------------------------------------------- force / virial
def prod_force_a(net_deriv, env_deriv, nlist, nall, nframes=1):
    nlist = _c(nlist, np.int32)
    nnei = nlist.shape[-1]
    nloc = nlist.size // nnei // nframes
    f = np.zeros((nframes, nall, 3), np.float64)
    lib().oracle_prod_force_a(_p(f, _dp), _p(_c(net_deriv, np.float64), _dp),
                              _p(_c(env_deriv, np.float64), _dp), _p(nlist, _ip), nloc, nall,
                              nnei, nframes)
    return f


def prod_force_grad_a(grad, env_deriv, nlist, nframes=1):
    """grad [nframes, nloc, 3] -> grad_net [nframes * nloc, 4 nnei]."""
    nlist = _c(nlist, np.int32)
    nnei = nlist.shape[-1]
    nloc = nlist.size // nnei // nframes
    gn = np.zeros((nframes * nloc, 4 * nnei), np.float64)
    lib().oracle_prod_force_grad_a(_p(gn, _dp), _p(_c(grad, np.float64), _dp), _p(_c(env_deriv, np.float64), _dp),
                                   _p(nlist, _ip), nloc, nnei, nframes)
    return gn


def prod_virial_grad_a(grad, env_deriv, rij, nlist):
    """grad [9] -> grad_net [nloc, 4 nnei] (one frame)."""
    nlist = _c(nlist, np.int32)
    nnei = nlist.shape[-1]
    nloc = nlist.size // nnei
    gn = np.zeros((nloc, 4 * nnei), np.float64)
    lib().oracle_prod_virial_grad_a(_p(gn, _dp), _p(_c(grad, np.float64), _dp), _p(_c(env_deriv, np.float64), _dp),
                                    _p(_c(rij, np.float64), _dp), _p(nlist, _ip), nloc, nnei)
    return gn


def prod_virial_a(net_deriv, env_deriv, rij, nlist, nall):
    nlist = _c(nlist, np.int32)
    nnei = nlist.shape[-1]
    nloc = nlist.size // nnei
    v = np.zeros(9, np.float64)
    av = np.zeros((nall, 9), np.float64)
    lib().oracle_prod_virial_a(_p(v, _dp), _p(av, _dp), _p(_c(net_deriv, np.float64), _dp),
                               _p(_c(env_deriv, np.float64), _dp), _p(_c(rij, np.float64), _dp),
                               _p(nlist, _ip), nloc, nall, nnei)
    return v, av


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def set_threads(n: int) -> None:
    os.environ["OMP_NUM_THREADS"] = str(n)
