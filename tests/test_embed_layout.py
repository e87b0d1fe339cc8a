"""CPU check of the fused-embedding register layout: a numpy emulation of the
kernels' per-lane dataflow (MFMA 16x16x32 with the lane layout verified on
gfx950: A lane -> row l%16, K 8*(l//16)+u; B lane -> col l%16, same K;
D lane -> col l%16, rows 4*(l//16)+v) using the packed fragments, against
the plain MLP / descriptor math in float64."""

import numpy as np
import pytest
import torch

from dpmi import embed, systems

L = np.arange(64)
LG, LR = L // 16, L % 16


def bf16(x):
    return torch.tensor(np.asarray(x, dtype=np.float32)).to(torch.bfloat16).double().numpy()


def mfma(a, b, c=None):
    A = np.zeros((16, 32))
    B = np.zeros((32, 16))
    for u in range(8):
        A[LR, 8 * LG + u] = a[:, u]
        B[8 * LG + u, LR] = b[:, u]
    D = A @ B
    out = np.stack([D[4 * LG + v, LR] for v in range(4)], axis=1)
    return out if c is None else out + c


def phi(g, u):
    return 16 * (u // 4) + 4 * g + (u % 4)


def tanh_scaled(z):
    """The kernels' tanh of x from z = K2 x (weights and biases pre-scaled)."""
    return 1.0 - 2.0 / (1.0 + np.exp2(z))


def forward_tile(s, frag, cst):
    """s: [64] radial input of lane's row (row = l%16)."""
    t1 = np.zeros((64, 8))
    h1 = np.zeros((64, 8))
    for u in range(8):
        q = phi(LG, u)
        t1[:, u] = tanh_scaled(s * cst[q] + cst[32 + q])
        h1[:, u] = t1[:, u] * cst[64 + q]
    # the bias is the MFMA accumulator input
    z2 = [mfma(frag[mt], h1, np.stack([cst[96 + 16 * mt + 4 * LG + v] for v in range(4)], axis=1)) for mt in range(4)]
    t2 = np.zeros((64, 16))
    h2 = np.zeros((64, 16))
    for mt in range(4):
        q = 16 * mt + 4 * LG
        for v in range(4):
            t2[:, mt * 4 + v] = tanh_scaled(z2[mt][:, v])
            h2[:, mt * 4 + v] = t2[:, mt * 4 + v] * cst[160 + q + v] + h1[:, (mt & 1) * 4 + v]
    b2 = [np.concatenate([h2[:, (2 * kc) * 4:(2 * kc) * 4 + 4], h2[:, (2 * kc + 1) * 4:(2 * kc + 1) * 4 + 4]], axis=1)
          for kc in range(2)]
    t3 = np.zeros((64, 32))
    G = np.zeros((64, 32))
    for mt in range(8):
        c0 = np.stack([cst[224 + 16 * mt + 4 * LG + v] for v in range(4)], axis=1)
        z = mfma(frag[5 + 2 * mt], b2[1], mfma(frag[4 + 2 * mt], b2[0], c0))
        q = 16 * mt + 4 * LG
        for v in range(4):
            t3[:, mt * 4 + v] = tanh_scaled(z[:, v])
            G[:, mt * 4 + v] = t3[:, mt * 4 + v] * cst[352 + q + v] + h2[:, (mt & 3) * 4 + v]
    return t1, h1, t2, h2, t3, G


def lanes_to_rows(G):
    """[64, 32] lane layout -> [16 rows, 128 q3]."""
    out = np.zeros((16, 128))
    for mt in range(8):
        for v in range(4):
            out[LR, 16 * mt + 4 * LG + v] = G[:, mt * 4 + v]
    return out


def mlp_ref(s, layers):
    x = s[:, None]
    for l in layers:
        w, b, idt = embed._layer(l)
        y = np.tanh(x @ w + b) * idt
        nin, nout = w.shape
        if nin == nout:
            y = y + x
        elif 2 * nin == nout:
            y = y + np.concatenate([x, x], axis=1)
        x = y
    return x


def effective_layers(layers):
    """The float64 MLP the kernels evaluate: W2, W3 as bf16(K2 W) / K2."""
    out = []
    for l in layers:
        d = dict(l)
        w = np.asarray(l["w"], dtype=np.float64)
        if w.shape[0] > 1:
            d["w"] = torch.tensor(embed.K2 * w).to(torch.bfloat16).double().numpy() / embed.K2
        out.append(d)
    return out


@pytest.fixture(scope="module")
def packed():
    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)
    bits, consts, n1 = embed.pack_embedding(params["embedding"])
    frags = torch.from_numpy(bits.astype(np.int16)).view(torch.bfloat16).float().numpy().astype(np.float64)
    return params, frags, consts.astype(np.float64), n1


def test_forward_tile_matches_mlp(packed):
    params, frags, consts, n1 = packed
    rng = np.random.default_rng(0)
    q1, q2, q3 = embed.internal_maps(n1)
    for t in range(2):
        s = rng.normal(size=16) * 2
        _, _, _, _, _, G = forward_tile(s[LR], frags[t], consts[t])
        got = lanes_to_rows(G)
        ref = mlp_ref(s, effective_layers(params["embedding"][t]))
        # biases / timesteps / w1 are stored in fp32 in the constant table
        np.testing.assert_allclose(got[:, q3], ref, rtol=0, atol=1e-5)
        pad = np.setdiff1d(np.arange(128), q3)
        assert np.all(got[:, pad] == 0)


def test_forward_gr_transpose(packed):
    """k_embed_fwd's G^T R for one 32-row chunk: G of each 16-row tile
    transposed by MFMA against identity fragments, rows in psi order, em^T as
    the A operand -> GR^T (M = c, N = feature), against em^T G in float64
    with G rounded to bf16."""
    params, frags, consts, n1 = packed
    rng = np.random.default_rng(5)
    em = rng.normal(size=(32, 4))
    eye = [np.array([[1.0 if (u >> 2) == j and 4 * LG[l] + (u & 3) == LR[l] else 0.0 for u in range(8)]
                     for l in range(64)]) for j in range(2)]
    gt = np.zeros((2, 4, 64, 8))
    Gfull = np.zeros((32, 128))
    for rt in range(2):
        _, _, _, _, _, G = forward_tile(em[16 * rt + LR, 0], frags[1], consts[1])
        Gb = bf16(G)
        Gfull[16 * rt:16 * rt + 16] = lanes_to_rows(Gb)
        for kc in range(4):
            gb = np.stack([Gb[:, (2 * kc + (u >> 2)) * 4 + (u & 3)] for u in range(8)], axis=1)
            for j in range(2):
                gt[rt, kc][:, 4 * j:4 * j + 4] = mfma(gb, eye[j])
    psi = np.array([[4 * g + u if u < 4 else 16 + 4 * g + u - 4 for u in range(8)] for g in range(4)])
    emb = bf16(em)
    a = np.where((LR < 4)[:, None], emb[psi[LG], np.minimum(LR, 3)[:, None]], 0.0)
    got = np.zeros((4, 128))
    for nt in range(8):
        b = np.concatenate([gt[0, nt >> 1][:, 4 * (nt & 1):4 * (nt & 1) + 4],
                            gt[1, nt >> 1][:, 4 * (nt & 1):4 * (nt & 1) + 4]], axis=1)
        d = mfma(a, b)
        for v in range(4):
            got[v, 16 * nt + LR[LG == 0]] = d[LG == 0][:, v]
    ref = emb.T @ Gfull
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_backward_tile_matches_autograd(packed):
    """dE/ds and dE/d em of one tile for E = sum dGR . (em^T G)."""
    import torch

    params, frags, consts, n1 = packed
    q1, q2, q3 = embed.internal_maps(n1)
    rng = np.random.default_rng(1)
    t = 1
    em = rng.normal(size=(16, 4))
    dgr = np.zeros((4, 128))
    dgr[:, q3] = rng.normal(size=(4, 4 * n1))
    f = frags[t]
    c = consts[t]
    t1, h1, t2, h2, t3, G = forward_tile(em[LR, 0], f, c)
    # the kernel's backward dataflow
    a_dgt = [np.where(((LG == 0)[:, None]) & (np.arange(8)[None, :] < 4),
                      dgr[np.minimum(np.arange(8), 3)[None, :], (16 * mt + LR)[:, None]], 0.0) for mt in range(8)]
    a_dg = [np.where((LR < 4)[:, None], dgr[np.minimum(LR, 3)[:, None], 32 * kc + phi(LG[:, None], np.arange(8)[None, :])],
                     0.0) for kc in range(4)]
    b_em = np.zeros((64, 8))
    b_em[LG == 0, :4] = em[LR[LG == 0]]
    dem = np.zeros((64, 4))
    for kc in range(4):
        bg = np.stack([G[:, (2 * kc + (u >> 2)) * 4 + (u & 3)] for u in range(8)], axis=1)
        dem = mfma(a_dg[kc], bg, dem)
    dgv = np.zeros((64, 32))
    bdz3 = [np.zeros((64, 8)) for _ in range(4)]
    for mt in range(8):
        d = mfma(a_dgt[mt], b_em)
        q = 16 * mt + 4 * LG
        for v in range(4):
            dgv[:, mt * 4 + v] = d[:, v]
            bdz3[mt >> 1][:, (mt & 1) * 4 + v] = d[:, v] * c[576 + q + v] * (1 - t3[:, mt * 4 + v] ** 2)
    dh2 = np.zeros((64, 16))
    bdz2 = [np.zeros((64, 8)) for _ in range(2)]
    for mt in range(4):
        d = np.zeros((64, 4))
        for kc in range(4):
            d = mfma(f[20 + mt * 4 + kc], bdz3[kc], d)
        q = 16 * mt + 4 * LG
        for v in range(4):
            x = d[:, v] + dgv[:, mt * 4 + v] + dgv[:, (mt + 4) * 4 + v]
            dh2[:, mt * 4 + v] = x
            bdz2[mt >> 1][:, (mt & 1) * 4 + v] = x * c[512 + q + v] * (1 - t2[:, mt * 4 + v] ** 2)
    ds = np.zeros(64)
    for mt in range(2):
        d = np.zeros((64, 4))
        for kc in range(2):
            d = mfma(f[36 + mt * 2 + kc], bdz2[kc], d)
        for v in range(4):
            u = mt * 4 + v
            q = phi(LG, u)
            x = d[:, v] + dh2[:, mt * 4 + v] + dh2[:, (mt + 2) * 4 + v]
            ds += x * c[64 + q] * (1 - t1[:, u] ** 2) * c[480 + q]
    ds_tot = np.array([ds[LR == r].sum() for r in range(16)])
    nd = np.zeros((16, 4))
    nd[:, 0] = dem[LG == 0][:, 0][np.argsort(LR[LG == 0])] + ds_tot
    nd[:, 1:] = dem[LG == 0][:, 1:][np.argsort(LR[LG == 0])]
    # autograd reference with the same (bf16-rounded) weights
    emt = torch.tensor(em, requires_grad=True)
    x = emt[:, :1]
    for l in effective_layers(params["embedding"][t]):
        w, b, idt = embed._layer(l)
        wt = torch.tensor(w)
        y = torch.tanh(x @ wt + torch.tensor(b)) * torch.tensor(idt)
        if w.shape[0] == w.shape[1]:
            y = y + x
        elif 2 * w.shape[0] == w.shape[1]:
            y = y + torch.cat([x, x], dim=1)
        x = y
    e = (torch.tensor(dgr[:, q3]) * (emt.t() @ x)).sum()
    e.backward()
    np.testing.assert_allclose(nd, emt.grad.numpy(), rtol=1e-5, atol=1e-5)  # fp32 constants


@pytest.mark.parametrize("n1", [25, 28, 29, 32, 8])
def test_skip_columns_are_empty(n1):
    """With n1 <= 28 the kernels skip the tanh of layer-1 slot column u = 7
    (slots 19, 23, 27, 31) and of the (mt odd, v = 3) rows of layers 2 and 3:
    no feature may live there.  Every map is a bijection onto its slots."""
    q1, q2, q3 = embed.internal_maps(n1)
    assert len(set(q1)) == n1 and len(set(q2)) == 2 * n1 and len(set(q3)) == 4 * n1
    assert q1.max() < 32 and q2.max() < 64 and q3.max() < 128
    skipped1 = {phi(g, 7) for g in range(4)}
    skipped3 = {16 * mt + 4 * g + 3 for mt in (1, 3, 5, 7) for g in range(4)}
    if embed.skip_layout(n1):
        assert skipped1 == {19, 23, 27, 31}
        assert not skipped1 & set(q1.tolist())
        assert not {q for q in skipped3 if q < 64} & set(q2.tolist())
        assert not skipped3 & set(q3.tolist())
    else:
        assert n1 > 28
