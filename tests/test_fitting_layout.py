"""CPU check of the fused fitting kernel's fragment layout: a numpy emulation
of k_fit's dataflow (the kernel's own fragment indexing, MFMA 16x16x32 lane
layout A lane -> row l%16, K 8*(l//16)+u; B lane -> col l%16, same K; D lane ->
col l%16, rows 4*(l//16)+v) on the packed fragments, against a float64 MLP
with the same bf16-rounded weights and torch autograd for dE/dD."""

import numpy as np
import pytest
import torch

from dpmi import _lib, fitting, systems

L = np.arange(64)
LG, LN = L // 16, L % 16


def mfma(a, b, c):
    """a, b [64, 8] fragments -> c + A @ B in D layout [64, 4]."""
    A = np.zeros((16, 32))
    B = np.zeros((32, 16))
    for u in range(8):
        A[LN, 8 * LG + u] = a[:, u]
        B[8 * LG + u, LN] = b[:, u]
    D = A @ B
    return c + np.stack([D[4 * LG + v, LN] for v in range(4)], axis=1)


def bf16(x):
    return torch.tensor(np.asarray(x, dtype=np.float32)).to(torch.bfloat16).double().numpy()


def emulate(frags, cst, ebias, X, k1):
    """k_fit on one tile: X [64, k1] -> (e [64], dD [64, k1])."""
    nk1 = 2 * ((k1 + 63) // 64)
    nm1 = k1 // 16
    w1f = 0
    w2f = w1f + 16 * nk1 * 64
    w3f, w3b, w2b = w2f + 16 * 8 * 64, w2f + 2 * 16 * 8 * 64, w2f + 3 * 16 * 8 * 64
    w1b = w2f + 4 * 16 * 8 * 64
    F = frags.reshape(-1, 8)  # [lane-fragments, 8]
    kH = 256
    Xp = np.zeros((64, 32 * nk1))
    Xp[:, :k1] = bf16(X)

    def B_of(buf, s, kbase):  # lds_b: atom 16 s + n, K = kbase + 8 g + u
        return np.stack([buf[16 * s + LN, kbase + 8 * LG + u] for u in range(8)], axis=1)

    def to_buf(buf, w, mt, s, vals):  # lds_put (bf16)
        for v in range(4):
            buf[16 * s + LN, 16 * (4 * w + mt) + 4 * LG + v] = bf16(vals[:, v])

    def product(off, nk, buf_or_x, w):
        acc = np.zeros((4, 4, 64, 4))
        for kc in range(nk):
            for mt in range(4):
                a = F[off + ((4 * w + mt) * nk + kc) * 64 + L]
                for s in range(4):
                    acc[mt, s] = mfma(a, B_of(buf_or_x, s, 32 * kc), acc[mt, s])
        return acc

    hbuf = np.zeros((64, kH))
    h = np.zeros((4, 4, 4, 64, 4))  # [w, mt, s, lane, v]
    f = np.zeros((3, 4, 4, 4, 64, 4))
    for w in range(4):
        acc = product(w1f, nk1, Xp, w)
        for mt in range(4):
            j = 16 * (4 * w + mt) + 4 * LG
            for s in range(4):
                for v in range(4):
                    tt = np.tanh(acc[mt, s][:, v] + cst[j + v])
                    h[w, mt, s][:, v] = tt * cst[kH + j + v]
                    f[0, w, mt, s][:, v] = cst[kH + j + v] * (1 - tt * tt)
                to_buf(hbuf, w, mt, s, h[w, mt, s])
    for layer in range(2):
        nbuf = np.zeros((64, kH))
        for w in range(4):
            acc = product(w2f if layer == 0 else w3f, 8, hbuf, w)
            for mt in range(4):
                j = 16 * (4 * w + mt) + 4 * LG
                for s in range(4):
                    for v in range(4):
                        tt = np.tanh(acc[mt, s][:, v] + cst[(2 + 2 * layer) * kH + j + v])
                        idt = cst[(3 + 2 * layer) * kH + j + v]
                        h[w, mt, s][:, v] += tt * idt
                        f[1 + layer, w, mt, s][:, v] = idt * (1 - tt * tt)
                    to_buf(nbuf, w, mt, s, h[w, mt, s])
        hbuf = nbuf
    e = np.zeros(64)
    dz = np.zeros((64, kH))
    for w in range(4):
        for s in range(4):
            part = np.zeros(64)
            for mt in range(4):
                j = 16 * (4 * w + mt) + 4 * LG
                vals = np.zeros((64, 4))
                for v in range(4):
                    part += h[w, mt, s][:, v] * cst[6 * kH + j + v]
                    vals[:, v] = cst[6 * kH + j + v] * f[2, w, mt, s][:, v]
                to_buf(dz, w, mt, s, vals)
            for n in range(16):
                e[16 * s + n] += part[LN == n].sum()
    e = e + cst[7 * kH] + ebias
    dh = np.zeros((4, 4, 4, 64, 4))
    for layer in range(2):
        nbuf = np.zeros((64, kH))
        for w in range(4):
            acc = product(w3b if layer == 0 else w2b, 8, dz, w)
            for mt in range(4):
                j = 16 * (4 * w + mt) + 4 * LG
                for s in range(4):
                    vals = np.zeros((64, 4))
                    for v in range(4):
                        if layer == 0:
                            dh[w, mt, s][:, v] = acc[mt, s][:, v] + cst[6 * kH + j + v]
                            vals[:, v] = dh[w, mt, s][:, v] * f[1, w, mt, s][:, v]
                        else:
                            vals[:, v] = (acc[mt, s][:, v] + dh[w, mt, s][:, v]) * f[0, w, mt, s][:, v]
                    to_buf(nbuf, w, mt, s, vals)
        dz = nbuf
    dD = np.zeros((64, k1))
    mtpw = (nm1 + 3) // 4
    for w in range(4):
        for m in range(mtpw):
            mt = w * mtpw + m
            if mt >= nm1:
                break
            for s in range(4):
                d = np.zeros((64, 4))
                for kc in range(8):
                    d = mfma(F[w1b + (mt * 8 + kc) * 64 + L], B_of(dz, s, 32 * kc), d)
                for v in range(4):
                    dD[16 * s + LN, 16 * mt + 4 * LG + v] = d[:, v]
    return e, dD


def reference(layers, ebias, X):
    """float64 MLP with bf16-rounded weight matrices, autograd dE/dX."""
    x = torch.tensor(X, requires_grad=True)
    y = x
    for k, l in enumerate(layers):
        w, b, idt = fitting._layer(l)
        wt = torch.tensor(bf16(w))
        z = y @ wt + torch.tensor(b)
        if k == len(layers) - 1:
            y = z
            break
        t = torch.tanh(z)
        if idt is not None:
            t = t * torch.tensor(idt)
        y = t + y if w.shape[0] == w.shape[1] else t
    e = y[:, 0] + ebias
    e.sum().backward()
    return e.detach().numpy(), x.grad.numpy()


@pytest.mark.parametrize("k1,h", [(64, 40), (192, 48)])
def test_fitting_emulation(k1, h):
    rng = np.random.default_rng(k1)
    spec = systems.ModelSpec(ntypes=1, sel=[8], rcut=4.0, rcut_smth=1.0, neuron=[4, 8, 16], axis_neuron=4,
                             fitting=[h, h, h])
    params = systems.make_params(spec, seed=3)
    layers = params["fitting"][0]
    # replace layer 1 by a k1-wide input (the kernel is generic in k1)
    layers = [dict(layers[0]), *layers[1:]]
    layers[0]["w"] = rng.normal(size=(k1, h)) / np.sqrt(k1)
    bits, consts, ebias, kk = fitting.pack_fitting([layers], [0.25])
    assert kk == k1
    frags = torch.from_numpy(bits.astype(np.int16)).view(torch.bfloat16).double().numpy()[0]
    X = rng.normal(size=(64, k1)) * 0.5
    e, dD = emulate(frags, consts[0].astype(np.float64), ebias[0], X, k1)
    # the kernel rounds the hidden activations and adjoints to bf16 as GEMM
    # operands; the float64 reference does not
    e_ref, dD_ref = reference(layers, ebias[0], bf16(X))
    # (measured: 1.4e-3 and 2.9e-3 of the largest value; an indexing mistake is O(1))
    np.testing.assert_allclose(e, e_ref, rtol=0, atol=5e-3 * np.abs(e_ref).max())
    assert np.abs(dD - dD_ref).max() <= 6e-3 * np.abs(dD_ref).max()


def test_fitting_frag_count():
    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=1)
    bits, consts, ebias, k1 = fitting.pack_fitting(params["fitting"], params["bias_atom_e"])
    cnt = np.zeros(1, dtype=np.int64)
    _lib.call("dpmi_se_a_fitting_frag_count", _lib.ptr(cnt), k1)
    assert bits.shape[1] * 64 == cnt[0]
    assert consts.shape == (spec.ntypes, fitting.N_CONST)
    assert fitting.supported(params["fitting"])
