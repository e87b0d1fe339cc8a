"""Pins of the oracle's PyTorch-backend env-mat (oracle.model_ref.prod_env_mat_pt,
a restatement of deepmd/pt/model/descriptor/env_mat.py:11-92): agreement with
the C restatement of env_mat_a_cpu (itself pinned bitwise to the compiled
reference) on formatted lists, the closed forms of padding, beyond-rcut and
switch values, protection, the radial-only form, permutation of the list, and
the derivative against central finite differences."""

import numpy as np
import pytest
import torch

from dpmi import systems


def _water_list(oracle, rc_list=6.0, sel=(46, 92)):
    s = systems.water192(0)
    ec, et, mp, nl, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, rc_list, list(sel))
    return s, ec, et, nl


def _stats(ntypes, nnei, seed=3, ncomp=4):
    rng = np.random.default_rng(seed)
    # se_a statistics are constant over slots: [avg_r, 0, 0, 0] / [std_r, std_a x3]
    mean = np.zeros((ntypes, nnei, ncomp))
    mean[..., 0] = rng.uniform(0.0, 0.2, size=(ntypes, 1))
    std = np.broadcast_to(0.5 + rng.uniform(size=(ntypes, 1, ncomp)), (ntypes, nnei, ncomp)).copy()
    return mean, std


def _pt(ec, nl, at, mean, std, rc=6.0, rs=0.5, **kw):
    from oracle.model_ref import prod_env_mat_pt

    c = torch.as_tensor(ec, dtype=torch.float64).reshape(1, -1)
    n = torch.as_tensor(np.ascontiguousarray(nl), dtype=torch.long).reshape(1, *nl.shape)
    a = torch.as_tensor(at, dtype=torch.long).reshape(1, -1)
    return prod_env_mat_pt(c, n, a, torch.as_tensor(mean), torch.as_tensor(std), rc, rs, **kw)


def test_matches_cpp_env_mat_on_formatted_list(oracle):
    s, ec, et, nl = _water_list(oracle)
    mean, std = _stats(2, 138)
    env, diff, sw = _pt(ec, nl, s.atype, mean, std)
    cem, _, crij = oracle.env_mat_a(nl, ec, et, mean, std, 6.0, 0.5, [46, 92])
    np.testing.assert_allclose(env[0].numpy(), cem.reshape(s.nloc, 138, 4), rtol=1e-13, atol=1e-13)
    # diff is r_ij of the C++ env-mat on listed slots, 0 on padding
    np.testing.assert_allclose(diff[0].numpy(), crij.reshape(s.nloc, 138, 3), rtol=0, atol=1e-14)


def test_padding_and_beyond_rcut_rows_closed_form(oracle):
    s, ec, et, nl = _water_list(oracle, rc_list=7.5, sel=(80, 160))
    mean, std = _stats(2, 240)
    env, diff, sw = _pt(ec, nl, s.atype, mean, std)
    env, diff, sw = env[0].numpy(), diff[0].numpy(), sw[0, :, :, 0].numpy()
    r = np.linalg.norm(ec[np.where(nl >= 0, nl, 0)] - ec[: s.nloc, None, :], axis=-1)
    const = -mean[s.atype] / std[s.atype]
    beyond = (nl >= 0) & (r > 6.0)
    pad = nl < 0
    assert beyond.sum() > 1000 and pad.sum() > 1000
    np.testing.assert_array_equal(env[beyond], const[beyond])
    np.testing.assert_array_equal(env[pad], const[pad])
    assert (sw[beyond] == 0).all() and (sw[pad] == 0).all()
    assert (diff[pad] == 0).all()
    np.testing.assert_allclose(np.linalg.norm(diff[beyond], axis=-1), r[beyond], rtol=1e-15)


def test_switch_values_closed_form():
    from oracle.model_ref import exp_switch, smooth_weight

    r = torch.tensor([0.1, 0.5, 3.25, 6.0, 7.0], dtype=torch.float64)
    np.testing.assert_allclose(smooth_weight(r, 0.5, 6.0).numpy(), [1.0, 1.0, 0.5, 0.0, 0.0], rtol=0, atol=1e-16)
    np.testing.assert_allclose(exp_switch(torch.tensor([0.5], dtype=torch.float64), 0.5, 6.0).numpy(), [np.exp(-1)],
                               rtol=1e-15)
    with pytest.raises(ValueError):
        smooth_weight(r, 6.0, 6.0)


def test_single_pair_with_protection():
    x = np.array([[0.0, 0.0, 0.0], [1.2, -0.7, 2.1]])
    nl = np.array([[1, -1]])
    mean, std = np.zeros((1, 2, 4)), np.ones((1, 2, 4))
    for p in (0.0, 0.3):
        env, diff, sw = _pt(x, nl, np.array([0, 0]), mean, std, protection=p)
        d = x[1] - x[0]
        r = np.linalg.norm(d)
        u = (r - 0.5) / 5.5
        w = u**3 * (-6 * u**2 + 15 * u - 10) + 1
        want = np.concatenate([[w / (r + p)], w * d / (r + p) ** 2])
        np.testing.assert_allclose(env[0, 0, 0].numpy(), want, rtol=1e-14)
        np.testing.assert_array_equal(env[0, 0, 1].numpy(), [0.0, 0.0, 0.0, 0.0])


def test_radial_only_is_first_column(oracle):
    s, ec, et, nl = _water_list(oracle)
    mean, std = _stats(2, 138)
    full, _, _ = _pt(ec, nl, s.atype, mean, std)
    rad, _, _ = _pt(ec, nl, s.atype, mean[..., :1], std[..., :1], radial_only=True)
    np.testing.assert_array_equal(rad[..., 0].numpy(), full[..., 0].numpy())


def test_list_permutation_permutes_rows(oracle):
    s, ec, et, nl = _water_list(oracle, rc_list=7.0, sel=(60, 120))
    mean, std = _stats(2, 180)
    mean[:] = mean[:, :1]
    std[:] = std[:, :1]
    perm = np.random.default_rng(0).permutation(180)
    a, da, wa = _pt(ec, nl, s.atype, mean, std)
    b, db, wb = _pt(ec, nl[:, perm], s.atype, mean, std)
    np.testing.assert_array_equal(b.numpy(), a.numpy()[:, :, perm])
    np.testing.assert_array_equal(db.numpy(), da.numpy()[:, :, perm])


@pytest.mark.parametrize("exp_sw", [False, True])
def test_gradient_matches_finite_difference(oracle, exp_sw):
    from oracle.model_ref import prod_env_mat_pt

    s, ec, et, nl = _water_list(oracle, rc_list=7.0, sel=(60, 120))
    nl = nl[:24]
    mean, std = _stats(2, 180)
    rng = np.random.default_rng(1)
    ge, gd, gs = rng.normal(size=(24, 180, 4)), rng.normal(size=(24, 180, 3)), rng.normal(size=(24, 180, 1))

    def loss(c):
        e, d, w = prod_env_mat_pt(c.reshape(1, -1), torch.as_tensor(nl).reshape(1, 24, 180),
                                  torch.as_tensor(s.atype[:24], dtype=torch.long).reshape(1, 24), torch.as_tensor(mean),
                                  torch.as_tensor(std), 6.0, 0.5, protection=0.05, use_exp_switch=exp_sw)
        return (e[0] * torch.as_tensor(ge)).sum() + (d[0] * torch.as_tensor(gd)).sum() + (w[0] * torch.as_tensor(gs)).sum()

    c = torch.tensor(ec, requires_grad=True)
    (g,) = torch.autograd.grad(loss(c), c)
    h = 1e-6
    for atom in (0, 5, 23, int(nl[3, 7]), int(nl[10, 100])):
        for k in range(3):
            cp = torch.tensor(ec)
            cm = torch.tensor(ec)
            cp[atom, k] += h
            cm[atom, k] -= h
            fd = (loss(cp) - loss(cm)).item() / (2 * h)
            assert abs(fd - g[atom, k].item()) <= 1e-6 * max(1.0, abs(fd)), (atom, k, fd, g[atom, k].item())
