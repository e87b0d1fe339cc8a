"""Table builder (dpmi.table): pinned by closed forms, not by retyping.

* The quintic Hermite step reproduces any degree-5 polynomial exactly, and a
  smooth function to O(h^6) (checked through the oracle's
  tabulate_fusion_se_a, which is itself pinned to the reference's goldens).
* The value / slope / curvature chain of the embedding net equals torch
  autograd (first and second derivative) with and without the timestep.
* A table built with the timestep folded in reproduces the network it
  replaces to 1e-9 on the fine grid; built the reference's way (timestep
  ignored) it reproduces the network without its timestep.
"""

import numpy as np
import pytest
import torch

from dpmi import systems, table


def _eval_table(oracle, tab, info, x):
    """v_k(x) of every channel through the oracle's tabulate (nnei = 1,
    em = [1, 0, 0, 0])."""
    n = x.shape[0]
    M = tab.shape[1] // 6
    em = np.zeros((n, 1, 4))
    em[:, 0, 0] = 1.0
    out = oracle.tabulate_fusion_se_a(tab, info, x.reshape(-1, 1), em, M)
    return np.asarray(out).reshape(n, 4, M)[:, 0, :]


def test_hermite_reproduces_quintic(oracle):
    rng = np.random.default_rng(0)
    c = rng.normal(size=(6, 3))  # three channels
    f = lambda x: sum(c[i] * x[:, None] ** i for i in range(6))
    df = lambda x: sum(i * c[i] * x[:, None] ** (i - 1) for i in range(1, 6))
    d2f = lambda x: sum(i * (i - 1) * c[i] * x[:, None] ** (i - 2) for i in range(2, 6))
    lo, up, ext, s0, s1 = 0.0, 2.0, 5.0, 0.01, 0.1
    xx = np.append(np.append(np.arange(lo, up, s0), np.arange(up, ext * up, s1)), [ext * up])
    nspline = int((up - lo) / s0 + (ext * up - up) / s1)
    tt = np.full((nspline, 1), s1)
    tt[: int((up - lo) / s0)] = s0
    tab = table.hermite_table(f(xx), df(xx), d2f(xx), tt)
    info = np.array([lo, up, ext * up, s0, s1, -1.0])
    x = rng.uniform(lo, ext * up * 0.999, 400)
    got = _eval_table(oracle, tab, info, x)
    ref = f(x)
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_hermite_sixth_order(oracle):
    """sin on two grids: the error falls by ~2^6 when h halves."""
    errs = []
    for h in (0.2, 0.1):
        xx = np.arange(0.0, 4.0 + h / 2, h)
        n = xx.shape[0] - 1
        tab = table.hermite_table(np.sin(xx)[:, None], np.cos(xx)[:, None], -np.sin(xx)[:, None],
                                  np.full((n, 1), h))
        info = np.array([0.0, 4.0, 4.0, h, h, -1.0])
        x = np.linspace(0.0, 3.99, 997)
        errs.append(np.abs(_eval_table(oracle, tab, info, x)[:, 0] - np.sin(x)).max())
    assert errs[1] < errs[0] / 40, errs


@pytest.mark.parametrize("fold", [False, True])
def test_derivative_chain_matches_autograd(fold):
    spec = systems.ModelSpec(ntypes=1, sel=[8], neuron=[5, 10, 20], axis_neuron=4, fitting=[8, 8, 8])
    layers = systems.make_params(spec, seed=11)["embedding"][0]
    x = torch.linspace(-1.0, 3.0, 41, dtype=torch.float64, requires_grad=True)
    y = x[:, None]
    for l in layers:
        w, b, idt = table._layer(l)
        z = torch.tanh(y @ torch.tensor(w) + torch.tensor(b))
        if fold and idt is not None:
            z = z * torch.tensor(idt)
        nin, nout = w.shape
        y = z + (y if nout == nin else torch.cat([y, y], 1) if nout == 2 * nin else 0)
    M = y.shape[1]
    d1 = torch.stack([torch.autograd.grad(y[:, k].sum(), x, create_graph=True)[0] for k in range(M)], 1)
    d2 = torch.stack([torch.autograd.grad(d1[:, k].sum(), x, retain_graph=True)[0] for k in range(M)], 1)
    v, dv, d2v = table.mlp_value_derivs(x.detach().numpy(), layers, fold)
    np.testing.assert_allclose(v, y.detach().numpy(), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(dv, d1.detach().numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(d2v, d2.detach().numpy(), rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("fold", [False, True])
def test_table_reproduces_network(oracle, fold):
    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=3)
    rng = np.random.default_rng(1)
    nnei = sum(spec.sel)
    davg = np.zeros((spec.ntypes, nnei, 4))
    dstd = np.ones((spec.ntypes, nnei, 4))
    tabs = table.build_tables(params["embedding"], davg, dstd, spec.rcut, spec.rcut_smth, 0.9, spec.sel,
                              fold_idt=fold)
    tab, info = tabs[1]
    assert tab.shape == (280, 600)  # the config-1 grid: [0, 2) / 0.01 + [2, 10) / 0.1
    np.testing.assert_array_equal(info, [0, 2, 10, 0.01, 0.1, -1])
    x = rng.uniform(0.0, 2.0, 500)
    ref = table.mlp_value_derivs(x, params["embedding"][1], fold)[0]
    got = _eval_table(oracle, tab, info, x)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9)
    xe = rng.uniform(2.0, 9.99, 200)  # coarse extrapolation grid
    np.testing.assert_allclose(_eval_table(oracle, tab, info, xe),
                               table.mlp_value_derivs(xe, params["embedding"][1], fold)[0], rtol=0, atol=1e-4)
