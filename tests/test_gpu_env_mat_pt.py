"""dpmi.ops.prod_env_mat (HIP kernels behind dpmi_env_mat_pt_*) against the
oracle's restatement of the PyTorch backend's prod_env_mat
(deepmd/pt/model/descriptor/env_mat.py:11-92): forward, coordinate gradient,
the double backward of a force loss, and central finite differences, on
shuffled lists with entries beyond rcut and -1 padding in the middle."""

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


def _frames(oracle, nframes=2, rc_list=7.0, sel=(60, 120), seed=0):
    """Per frame: extended coordinates (padded to a common nall), atype and
    a list built at rc_list > rcut (so it holds entries beyond rcut),
    shuffled within each row, with some slots knocked out to -1."""
    rng = np.random.default_rng(seed)
    exts, lists, types = [], [], []
    for f in range(nframes):
        s = systems.water192(f)
        ec, et, mp, nl, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, rc_list, list(sel))
        nl = nl.copy()
        for i in range(nl.shape[0]):
            nl[i] = nl[i, rng.permutation(nl.shape[1])]
        nl[rng.uniform(size=nl.shape) < 0.05] = -1
        exts.append(ec)
        lists.append(nl)
        types.append(s.atype)
    nall = max(e.shape[0] for e in exts)
    ext = np.zeros((nframes, nall, 3))
    for f, e in enumerate(exts):
        ext[f, : e.shape[0]] = e
        ext[f, e.shape[0]:] = 1000.0 + f  # unreferenced padding atoms
    return ext, np.stack(lists), np.stack(types)


def _stats(ntypes, nnei, ncomp=4, seed=3):
    rng = np.random.default_rng(seed)
    mean = np.zeros((ntypes, nnei, ncomp))
    mean[..., 0] = rng.uniform(0.0, 0.2, size=(ntypes, 1))
    std = np.broadcast_to(0.5 + rng.uniform(size=(ntypes, 1, ncomp)), (ntypes, nnei, ncomp)).copy()
    return mean, std


CASES = [
    ("plain", dict()),
    ("protection", dict(protection=0.1)),
    ("radial_only", dict(radial_only=True)),
    ("exp_switch", dict(use_exp_switch=True)),
]


def _oracle(ext, nl, at, mean, std, kw):
    from oracle.model_ref import prod_env_mat_pt

    return prod_env_mat_pt(torch.as_tensor(ext).reshape(ext.shape[0], -1), torch.as_tensor(nl, dtype=torch.long),
                           torch.as_tensor(at, dtype=torch.long), torch.as_tensor(mean), torch.as_tensor(std), 6.0,
                           0.5, **kw)


def _gpu(ext, nl, at, mean, std, kw, dtype=torch.float64, grad=False):
    c = torch.as_tensor(ext.reshape(ext.shape[0], -1), device=DEV, dtype=dtype)
    if grad:
        c.requires_grad_(True)
    out = ops.prod_env_mat(c, torch.as_tensor(nl, device=DEV), torch.as_tensor(at, device=DEV),
                           torch.as_tensor(mean, device=DEV, dtype=dtype), torch.as_tensor(std, device=DEV, dtype=dtype),
                           6.0, 0.5, **kw)
    return c, out


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_forward_matches_oracle(oracle, name, kw):
    _say(f"prod_env_mat forward {name}")
    ext, nl, at = _frames(oracle)
    ncomp = 1 if kw.get("radial_only") else 4
    mean, std = _stats(2, nl.shape[2], ncomp)
    oe, od, ow = _oracle(ext, nl, at, mean, std, kw)
    _, (e, d, w) = _gpu(ext, nl, at, mean, std, kw)
    assert e.shape == oe.shape and d.shape == od.shape and w.shape == ow.shape
    np.testing.assert_allclose(e.cpu().numpy(), oe.numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(d.cpu().numpy(), od.numpy(), rtol=0, atol=1e-13)
    np.testing.assert_allclose(w.cpu().numpy(), ow.numpy(), rtol=1e-12, atol=1e-14)
    # float32 coordinates: the same kernel in float
    _, (e32, d32, w32) = _gpu(ext, nl, at, mean, std, kw, dtype=torch.float32)
    np.testing.assert_allclose(e32.cpu().numpy(), oe.numpy(), rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(w32.cpu().numpy(), ow.numpy(), rtol=0, atol=2e-6)


def _weights(shape_e, shape_d, shape_w, seed=1):
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape_e), rng.normal(size=shape_d), rng.normal(size=shape_w)


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_coordinate_gradient_matches_oracle(oracle, name, kw):
    _say(f"prod_env_mat gradient {name}")
    ext, nl, at = _frames(oracle)
    ncomp = 1 if kw.get("radial_only") else 4
    mean, std = _stats(2, nl.shape[2], ncomp)
    nf, nloc, nnei = nl.shape
    ge, gd, gw = _weights((nf, nloc, nnei, ncomp), (nf, nloc, nnei, 3), (nf, nloc, nnei, 1))
    from oracle.model_ref import prod_env_mat_pt

    oc = torch.tensor(ext.reshape(nf, -1), requires_grad=True)
    oe, od, ow = prod_env_mat_pt(oc, torch.as_tensor(nl, dtype=torch.long), torch.as_tensor(at, dtype=torch.long),
                                 torch.as_tensor(mean), torch.as_tensor(std), 6.0, 0.5, **kw)
    (og,) = torch.autograd.grad((oe * torch.as_tensor(ge)).sum() + (od * torch.as_tensor(gd)).sum()
                                + (ow * torch.as_tensor(gw)).sum(), oc)
    c, (e, d, w) = _gpu(ext, nl, at, mean, std, kw, grad=True)
    L = (e * torch.as_tensor(ge, device=DEV)).sum() + (d * torch.as_tensor(gd, device=DEV)).sum() + \
        (w * torch.as_tensor(gw, device=DEV)).sum()
    (g,) = torch.autograd.grad(L, c, retain_graph=True)
    assert g.shape == c.shape
    np.testing.assert_allclose(g.cpu().numpy(), og.numpy(), rtol=1e-11, atol=1e-11 * np.abs(og.numpy()).max())
    # env-mat alone (diff and switch unused: their gradients are absent)
    (g1,) = torch.autograd.grad((e * torch.as_tensor(ge, device=DEV)).sum(), c)
    oc2 = torch.tensor(ext.reshape(nf, -1), requires_grad=True)
    oe2, _, _ = prod_env_mat_pt(oc2, torch.as_tensor(nl, dtype=torch.long), torch.as_tensor(at, dtype=torch.long),
                                torch.as_tensor(mean), torch.as_tensor(std), 6.0, 0.5, **kw)
    (og1,) = torch.autograd.grad((oe2 * torch.as_tensor(ge)).sum(), oc2)
    np.testing.assert_allclose(g1.cpu().numpy(), og1.numpy(), rtol=1e-11, atol=1e-11 * np.abs(og1.numpy()).max())


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_double_backward_matches_oracle(oracle, name, kw):
    """L2 = <V, d<G, env>/dx> differentiated w.r.t. G (the path of a force
    loss's parameter gradient) against the oracle's torch double backward."""
    _say(f"prod_env_mat double backward {name}")
    ext, nl, at = _frames(oracle, nframes=1)
    ncomp = 1 if kw.get("radial_only") else 4
    mean, std = _stats(2, nl.shape[2], ncomp)
    nf, nloc, nnei = nl.shape
    ge, gd, gw = _weights((nf, nloc, nnei, ncomp), (nf, nloc, nnei, 3), (nf, nloc, nnei, 1), seed=2)
    V = np.random.default_rng(4).normal(size=(nf, ext.shape[1] * 3))
    from oracle.model_ref import prod_env_mat_pt

    def run(c, G, Gd, Gw, make):
        e, d, w = make(c)
        L = (e * G).sum() + (d * Gd).sum() + (w * Gw).sum()
        (g,) = torch.autograd.grad(L, c, create_graph=True)
        return torch.autograd.grad((g * torch.as_tensor(V, device=c.device)).sum(), (G, Gd, Gw))

    oG, oGd, oGw = (torch.tensor(x, requires_grad=True) for x in (ge, gd, gw))
    oc = torch.tensor(ext.reshape(nf, -1), requires_grad=True)
    want = run(oc, oG, oGd, oGw, lambda c: prod_env_mat_pt(
        c, torch.as_tensor(nl, dtype=torch.long), torch.as_tensor(at, dtype=torch.long), torch.as_tensor(mean),
        torch.as_tensor(std), 6.0, 0.5, **kw))
    G, Gd, Gw = (torch.tensor(x, device=DEV, requires_grad=True) for x in (ge, gd, gw))
    c = torch.tensor(ext.reshape(nf, -1), device=DEV, requires_grad=True)
    got = run(c, G, Gd, Gw, lambda c: ops.prod_env_mat(
        c, torch.as_tensor(nl, device=DEV), torch.as_tensor(at, device=DEV), torch.as_tensor(mean, device=DEV),
        torch.as_tensor(std, device=DEV), 6.0, 0.5, **kw))
    for a, b in zip(got, want):
        np.testing.assert_allclose(a.cpu().numpy(), b.numpy(), rtol=1e-11, atol=1e-11 * np.abs(b.numpy()).max())


def test_gradient_central_difference(oracle):
    _say("prod_env_mat finite differences")
    ext, nl, at = _frames(oracle, nframes=1)
    mean, std = _stats(2, nl.shape[2])
    ge = np.random.default_rng(9).normal(size=nl.shape + (4,))
    G = torch.as_tensor(ge, device=DEV)
    n, a, m, s = (torch.as_tensor(x, device=DEV) for x in (nl, at, mean, std))

    def loss(c):
        return (ops.prod_env_mat(c, n, a, m, s, 6.0, 0.5, protection=0.02)[0] * G).sum().item()

    c = torch.tensor(ext.reshape(1, -1), device=DEV, requires_grad=True)
    (g,) = torch.autograd.grad((ops.prod_env_mat(c, n, a, m, s, 6.0, 0.5, protection=0.02)[0] * G).sum(), c)
    g = g.cpu().numpy().reshape(-1, 3)
    h = 1e-6
    for atom in (0, 7, 150, int(nl[0, 3, 5]), int(nl[0, 77, 100])):
        for k in range(3):
            cp = torch.tensor(ext.reshape(1, -1), device=DEV)
            cm = cp.clone()
            cp[0, 3 * atom + k] += h
            cm[0, 3 * atom + k] -= h
            fd = (loss(cp) - loss(cm)) / (2 * h)
            assert abs(fd - g[atom, k]) <= 1e-6 * max(1.0, abs(fd)), (atom, k, fd, g[atom, k])


def test_invalid_rows_are_nan_and_bad_cutoff_raises():
    _say("prod_env_mat invalid inputs")
    ext = torch.zeros((1, 4, 3), dtype=torch.float64, device=DEV)
    ext[0, 1, 0] = 1.0
    nl = torch.tensor([[[1, 9], [0, -1]]], device=DEV)  # index 9 >= nall
    at = torch.tensor([[0, 5]], device=DEV)  # type 5 >= ntypes
    m = torch.zeros((2, 2, 4), dtype=torch.float64, device=DEV)
    s = torch.ones((2, 2, 4), dtype=torch.float64, device=DEV)
    e, d, w = ops.prod_env_mat(ext.reshape(1, -1), nl, at, m, s, 6.0, 0.5)
    e = e.cpu().numpy()
    assert np.isfinite(e[0, 0, 0]).all() and np.isnan(e[0, 0, 1]).all() and np.isnan(e[0, 1]).all()
    with pytest.raises(ValueError, match="rmin"):
        ops.prod_env_mat(ext.reshape(1, -1), nl, at, m, s, 6.0, 6.0)
