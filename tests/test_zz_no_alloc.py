"""dpmi_prod_virial_se_a_noalloc_f64 (the library's own scratch instead of a
caller workspace) against dpmi_prod_virial_se_a_f64 on the same inputs, on
two calls in a row (the second smaller), and its refusal of inputs beyond
the scratch size."""

import numpy as np
import pytest
import torch

from dpmi import _lib, ops, systems

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _lib.load()


def _inputs(oracle, frames):
    """em_deriv, rij, nlist of each frame (the oracle's C++ env-mat) and a
    seeded net_deriv, stacked over frames."""
    eds, rijs, nls, nds = [], [], [], []
    nall = 0
    for k, s in enumerate(frames):
        oc, ot, om, onl, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, 6.0, [46, 92])
        avg, std = np.zeros((2, 138, 4)), np.ones((2, 138, 4))
        _, oed, orij = oracle.env_mat_a(onl, oc, ot, avg, std, 6.0, 0.5, [46, 92])
        eds.append(oed.reshape(-1))
        rijs.append(orij.reshape(-1))
        nls.append(onl.reshape(-1))
        nds.append(np.random.default_rng(k).normal(size=s.nloc * 138 * 4))
        nall = max(nall, oc.shape[0])
    T = lambda xs, dt: torch.as_tensor(np.stack(xs), device=DEV, dtype=dt)
    return (T(nds, torch.float64), T(eds, torch.float64), T(rijs, torch.float64), T(nls, torch.int32),
            frames[0].nloc, nall, len(frames))


def _noalloc(lib, nd, ed, rij, nl, nloc, nall, nf):
    virial = torch.empty((nf, 9), dtype=torch.float64, device=DEV)
    av = torch.empty((nf, nall * 9), dtype=torch.float64, device=DEV)
    rc = lib.dpmi_prod_virial_se_a_noalloc_f64(_lib.ptr(virial), _lib.ptr(av), _lib.ptr(nd), _lib.ptr(ed),
                                               _lib.ptr(rij), _lib.ptr(nl), nloc, nall, 138, nf, ops._stream())
    assert rc == 0, lib.dpmi_last_error()
    return virial, av


def test_noalloc_virial_matches_workspace_entry_point(oracle, lib):
    big = [systems.water12k(1), systems.water12k(2)]
    small = [systems.water192(0)]
    for name, frames in (("2 x 12,288 atoms", big), ("1 x 192 atoms", small)):
        _say(f"prod_virial noalloc {name}")
        nd, ed, rij, nl, nloc, nall, nf = _inputs(oracle, frames)
        v0, av0 = ops.prod_virial(nd, ed, rij, nl, nloc, nall, nf)
        v1, av1 = _noalloc(lib, nd, ed, rij, nl, nloc, nall, nf)
        # same kernels and the same fixed-order reduction: the virial is bitwise equal
        np.testing.assert_array_equal(v1.cpu().numpy(), v0.cpu().numpy())
        # the atomic virial is scattered with atomics (order-dependent last bits)
        a0 = av0.cpu().numpy()
        np.testing.assert_allclose(av1.cpu().numpy(), a0, rtol=1e-12, atol=1e-12 * np.abs(a0).max())
        assert np.abs(v0.cpu().numpy()).max() > 0


def test_noalloc_virial_rejects_inputs_beyond_scratch(lib):
    _say("prod_virial noalloc size limit")
    # 116,509 four-atom groups need 8,388,648 bytes of partials: one group too many
    nloc = 4 * 116509
    rc = lib.dpmi_prod_virial_se_a_noalloc_f64(None, None, None, None, None, None, nloc, nloc, 138, 1,
                                               ops._stream())
    assert rc == 1
    assert b"exceeds the library's scratch" in lib.dpmi_last_error()
    # the largest input that fits is accepted (no pointer is touched with nnei = 0 ... except outputs)
    virial = torch.empty((1, 9), dtype=torch.float64, device=DEV)
    nloc_max = 4 * 116508
    empty = torch.empty(0, dtype=torch.float64, device=DEV)
    nl = torch.empty(0, dtype=torch.int32, device=DEV)
    rc = lib.dpmi_prod_virial_se_a_noalloc_f64(_lib.ptr(virial), None, _lib.ptr(empty), _lib.ptr(empty),
                                               _lib.ptr(empty), _lib.ptr(nl), nloc_max, nloc_max, 0, 1, ops._stream())
    assert rc == 0, lib.dpmi_last_error()
    np.testing.assert_array_equal(virial.cpu().numpy(), np.zeros((1, 9)))
