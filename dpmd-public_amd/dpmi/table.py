"""Table builder for the compressed (tabulated) se_e2_a embedding.

Restates DPTabulate.build for descriptor type "A" (deepmd/utils/tabulate.py:
66-147, 273-367, 468-486; derivative chain of deepmd/utils/tabulate_math.py:
136-251, 353-479): each embedding net G(s) (one per neighbor type,
type_one_side) is sampled on the grid

    x = [lower, lower + s0, ..., upper) ++ [upper, upper + s1, ..., ext*upper) ++ [ext*upper]

with its value, first and second derivative at every node, and each
interval [x_k, x_k+1] of width t becomes the quintic Hermite polynomial that
matches value, slope and curvature at both ends (coefficients a0..a5 per
output channel, stored channel-major as table[k, 6*c + i]).  table_info =
[lower, upper, ext*upper, s0, s1, check_frequency] as
DescrptBlockSeA.enable_compression builds it (se_a.py:733-744).

The reference builder evaluates the layers without their timestep idt
(tabulate_math.py reads only w and b), so for a resnet_dt model the table
is not the network it replaces.  fold_idt=True builds the table of the
network as evaluated (y = tanh(xW + b) * idt + skip); the default follows
the reference.
"""

from __future__ import annotations

import numpy as np


def _layer(l):
    v = l.get("@variables", l)
    w = np.asarray(v["w"], dtype=np.float64)
    b = np.asarray(v["b"], dtype=np.float64).reshape(-1)
    idt = v.get("idt")
    idt = None if idt is None else np.asarray(idt, dtype=np.float64).reshape(-1)
    return w, b, idt


def mlp_value_derivs(x: np.ndarray, layers, fold_idt: bool = False):
    """Embedding net G(x) for scalar inputs x [n] with dG/dx and d2G/dx2,
    each [n, M] (float64 numpy, tanh activation, MLPLayer skip rules): the
    host form of the chain dpmi_tabulate_build_f64 runs on the GPU."""
    y = np.asarray(x, dtype=np.float64).reshape(-1, 1)
    dy = np.ones_like(y)
    d2y = np.zeros_like(y)
    for l in layers:
        w, b, idt = _layer(l)
        nin, nout = w.shape
        z = y @ w + b
        dz = dy @ w
        d2z = d2y @ w
        t = np.tanh(z)
        g1 = 1.0 - t * t
        g2 = -2.0 * t * g1
        scale = idt if (fold_idt and idt is not None) else 1.0
        out = t * scale
        dout = g1 * dz * scale
        d2out = (g2 * dz * dz + g1 * d2z) * scale
        if nout == nin:
            out, dout, d2out = out + y, dout + dy, d2out + d2y
        elif nout == 2 * nin:
            out = out + np.concatenate([y, y], axis=1)
            dout = dout + np.concatenate([dy, dy], axis=1)
            d2out = d2out + np.concatenate([d2y, d2y], axis=1)
        y, dy, d2y = out, dout, d2out
    return y, dy, d2y


def build_on_gpu(xx: np.ndarray, tt: np.ndarray, layers, fold_idt: bool = False, device="cuda"):
    """One net's table on the GPU (dpmi_tabulate_build_f64): the value /
    slope / curvature chain at the grid nodes xx and the quintic Hermite
    coefficients of the intervals of widths tt.  Returns (table [nspline,
    6 M], node values [3, nnodes, M]) as float64 device tensors."""
    import torch

    from . import ops
    from ._lib import ptr

    dev = torch.device(device)
    params, dims = [], []
    for l in layers:
        w, b, idt = _layer(l)
        nin, nout = w.shape
        scale = idt if (fold_idt and idt is not None) else np.ones(nout)
        params += [w.reshape(-1), b, scale]
        dims += [nin, nout]
    p = torch.as_tensor(np.concatenate(params), device=dev)
    tt = np.asarray(tt, dtype=np.float64).reshape(-1)
    xx = np.asarray(xx, dtype=np.float64).reshape(-1)
    if xx.shape[0] < tt.shape[0] + 1:
        raise ValueError("need a grid node at both ends of every interval")
    # the intervals' nodes (a grid from arange may carry one more point)
    x = torch.as_tensor(xx[: tt.shape[0] + 1], device=dev).contiguous()
    t = torch.as_tensor(tt, device=dev).contiguous()
    n, M = x.numel(), dims[-1]
    vdd = torch.empty((3, n, M), dtype=torch.float64, device=dev)
    tab = torch.empty((n - 1, 6 * M), dtype=torch.float64, device=dev)
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32))
    ops._gpu(x)
    ops._launch("dpmi_tabulate_build_f64", ptr(tab), ptr(vdd), ptr(x), ptr(t), n, ptr(p), ptr(d), len(layers),
                ops._stream())
    return tab, vdd


def hermite_table(vv: np.ndarray, dd: np.ndarray, d2: np.ndarray, tt: np.ndarray) -> np.ndarray:
    """Quintic Hermite coefficients of each interval from node values,
    slopes and curvatures ([nspline + 1, M] each) and interval widths
    tt [nspline, 1] -> table [nspline, 6 M] (channel-major)."""
    n = tt.shape[0]
    v0, v1 = vv[:n], vv[1:n + 1]
    d0, d1 = dd[:n], dd[1:n + 1]
    c0, c1 = d2[:n], d2[1:n + 1]
    h = v1 - v0
    a = np.empty((n, vv.shape[1], 6))
    a[..., 0] = v0
    a[..., 1] = d0
    a[..., 2] = 0.5 * c0
    a[..., 3] = (20 * h - (8 * d1 + 12 * d0) * tt - (3 * c0 - c1) * tt * tt) / (2 * tt ** 3)
    a[..., 4] = (-30 * h + (14 * d1 + 16 * d0) * tt + (3 * c0 - 2 * c1) * tt * tt) / (2 * tt ** 4)
    a[..., 5] = (12 * h - 6 * (d1 + d0) * tt + (c1 - c0) * tt * tt) / (2 * tt ** 5)
    return a.reshape(n, -1)


def spline5_switch(x: float, rmin: float, rmax: float) -> float:
    if x < rmin:
        return 1.0
    if x < rmax:
        u = (x - rmin) / (rmax - rmin)
        return u * u * u * (-6 * u * u + 15 * u - 10) + 1
    return 0.0


def env_mat_range(davg, dstd, rcut: float, rcut_smth: float, min_nbor_dist: float):
    """Per neighbor-type lower / upper of the tabulated input s = em[..., 0]
    (tabulate.py:468-486): floor(-davg/dstd), ceil((sw(d)/d - davg)/dstd)."""
    davg = np.asarray(davg, dtype=np.float64)
    dstd = np.asarray(dstd, dtype=np.float64)
    a0 = davg[..., 0].reshape(davg.shape[0], -1)
    s0 = dstd[..., 0].reshape(dstd.shape[0], -1)
    sw = spline5_switch(min_nbor_dist, rcut_smth, rcut)
    lower = np.floor(np.min(-a0 / s0, axis=1))
    upper = np.ceil(np.max(((1.0 / min_nbor_dist) * sw - a0) / s0, axis=1))
    return lower, upper


def build_tables(emb_layers, davg, dstd, rcut: float, rcut_smth: float, min_nbor_dist: float, sel,
                 extrapolate: float = 5.0, stride0: float = 0.01, stride1: float = 0.1,
                 check_frequency: int = -1, fold_idt: bool = False, device=None):
    """Tables of a type_one_side se_e2_a model: a list (one per neighbor type)
    of (table [nspline, 6 M] float64, table_info [6] float64).  device: where
    the tables are built: None for the host (numpy), a device for the
    dpmi_tabulate_build_f64 kernels."""
    lower, upper = env_mat_range(davg, dstd, rcut, rcut_smth, min_nbor_dist)
    live = np.asarray(sel) > 0
    ll = float(np.min(lower[live]))
    uu = float(np.max(upper[live]))
    xx = np.arange(ll, uu, stride0, dtype=np.float64)
    xx = np.append(xx, np.arange(uu, extrapolate * uu, stride1, dtype=np.float64))
    xx = np.append(xx, np.array([extrapolate * uu]))
    nspline = int((uu - ll) / stride0 + (extrapolate * uu - uu) / stride1)
    tt = np.full((nspline, 1), stride1)
    tt[: int((uu - ll) / stride0)] = stride0
    out = []
    for layers in emb_layers:
        if device is not None:
            table = build_on_gpu(xx, tt, layers, fold_idt, device)[0].cpu().numpy()
        else:
            vv, dd, d2 = mlp_value_derivs(xx, layers, fold_idt)
            table = hermite_table(vv, dd, d2, tt)
        info = np.array([int(ll), int(uu), int(uu) * extrapolate, stride0, stride1, check_frequency],
                        dtype=np.float64)
        out.append((table, info))
    return out
