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
    r3 = np.zeros((64, 32))
    G = np.zeros((64, 32))
    for mt in range(8):
        c0 = np.stack([cst[224 + 16 * mt + 4 * LG + v] for v in range(4)], axis=1)
        z = mfma(frag[5 + 2 * mt], b2[1], mfma(frag[4 + 2 * mt], b2[0], c0))
        q = 16 * mt + 4 * LG
        for v in range(4):
            r3[:, mt * 4 + v] = 1.0 / (1.0 + np.exp2(z[:, v]))
            # what the kernels contract: idt3 (1 - 2 r3) + [h2, h2] (never formed there)
            G[:, mt * 4 + v] = (1.0 - 2.0 * r3[:, mt * 4 + v]) * cst[352 + q + v] + h2[:, (mt & 3) * 4 + v]
    return t1, h1, t2, h2, r3, G


def frag_of(x, kc):
    """[64, 32] lane values -> the MFMA operand of 32-feature chunk kc (K = phi slot)."""
    return np.stack([x[:, (2 * kc + (u >> 2)) * 4 + (u & 3)] for u in range(8)], axis=1)


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
    """k_embed_fwd's G^T R for one 32-row chunk: r3 and h2 of each 16-row
    tile transposed by MFMA against identity fragments, rows in psi order,
    em^T as the A operand (rows c >= 4 left as they come) ->
    GR^T = idt3 (em^T 1 - 2 em^T r3) + em^T [h2, h2] (M = c, N = feature),
    against em^T G in float64 with r3 and h2 rounded to bf16."""
    params, frags, consts, n1 = packed
    c = consts[1]
    rng = np.random.default_rng(5)
    em = rng.normal(size=(32, 4))
    eye = [np.array([[1.0 if (u >> 2) == j and 4 * LG[l] + (u & 3) == LR[l] else 0.0 for u in range(8)]
                     for l in range(64)]) for j in range(2)]
    rt3 = np.zeros((2, 4, 64, 8))
    ht2 = np.zeros((2, 2, 64, 8))
    Gfull = np.zeros((32, 128))
    for rt in range(2):
        _, _, _, h2, r3, _ = forward_tile(em[16 * rt + LR, 0], frags[1], c)
        r3b, h2b = bf16(r3), bf16(h2)
        Gb = np.zeros((64, 32))
        for mt in range(8):
            q = 16 * mt + 4 * LG
            for v in range(4):
                Gb[:, mt * 4 + v] = (1.0 - 2.0 * r3b[:, mt * 4 + v]) * c[352 + q + v] + h2b[:, (mt & 3) * 4 + v]
        Gfull[16 * rt:16 * rt + 16] = lanes_to_rows(Gb)
        for kc in range(4):
            for j in range(2):
                rt3[rt, kc][:, 4 * j:4 * j + 4] = mfma(frag_of(r3b, kc), eye[j])
        for kc in range(2):
            for j in range(2):
                ht2[rt, kc][:, 4 * j:4 * j + 4] = mfma(frag_of(h2b, kc), eye[j])
    psi = np.array([[4 * g + u if u < 4 else 16 + 4 * g + u - 4 for u in range(8)] for g in range(4)])
    emb = bf16(em)
    a = emb[psi[LG], (LR & 3)[:, None]]
    acc = [mfma(a, np.concatenate([rt3[0, nt >> 1][:, 4 * (nt & 1):4 * (nt & 1) + 4],
                                   rt3[1, nt >> 1][:, 4 * (nt & 1):4 * (nt & 1) + 4]], axis=1)) for nt in range(8)]
    acc_h = [mfma(a, np.concatenate([ht2[0, nt >> 1][:, 4 * (nt & 1):4 * (nt & 1) + 4],
                                     ht2[1, nt >> 1][:, 4 * (nt & 1):4 * (nt & 1) + 4]], axis=1)) for nt in range(4)]
    acc_s = mfma(a, np.ones((64, 8)))
    got = np.zeros((4, 128))
    l0 = LG == 0
    for nt in range(8):
        q = 16 * nt + LR[l0]
        for v in range(4):
            got[v, q] = c[352 + q] * (acc_s[l0][:, v] - 2.0 * acc[nt][l0][:, v]) + acc_h[nt & 3][l0][:, v]
    ref = emb.T @ Gfull
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_backward_tile_matches_autograd(packed):
    """dE/d em of one tile for E = sum dGR . (em^T G) through k_embed_bwd's
    dataflow: the forward recompute with its tangent in s (h1', h2', z3'
    from the same fragments), the A entries written where the stage writes
    them (lane l: features l and l + 64), dem = K_c + A r3 + A h2 and
    tg = A r3 (r3 - 1) z3' + A' h2' = (K2 / 2) sum_q dGR dG/ds."""
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
    # activations unrounded here: this checks the dataflow, not the bf16 error
    t1, h1, t2, h2, r3, G = forward_tile(em[LR, 0], f, c)
    # the stage: every entry of the slot's 1024 written exactly once
    dop = np.full(1024, np.nan)
    at = 32 * ((L >> 2) & 3) + 4 * ((L >> 4) & 1) + (L & 3) + 128 * (L >> 5)
    for cc in range(4):
        xa, xb = dgr[cc, L], dgr[cc, L + 64]
        for off, val in ((0, -2.0 * c[352 + L] * xa), (256, -2.0 * c[352 + 64 + L] * xb), (512, xa + xb),
                         (768, 0.5 * embed.K2 * (xa + xb))):
            assert np.isnan(dop[at + 8 * cc + off]).all()
            dop[at + 8 * cc + off] = val
    assert not np.isnan(dop).any()
    # A operand of chunk j, lane (r, g): entry 16 j + 4 g + (r & 3)
    A = [np.stack([dop[(16 * j + 4 * LG + (LR & 3)) * 8 + u] for u in range(8)], axis=1) for j in range(8)]
    # tangents
    h1t = np.zeros((64, 8))
    for u in range(8):
        h1t[:, u] = (1.0 - t1[:, u] ** 2) * c[480 + phi(LG, u)]
    h2t = np.zeros((64, 16))
    for mt in range(4):
        z2t = mfma(f[mt], h1t)
        q = 16 * mt + 4 * LG
        for v in range(4):
            h2t[:, mt * 4 + v] = (1.0 - t2[:, mt * 4 + v] ** 2) * c[512 + q + v] * z2t[:, v] + h1t[:, (mt & 1) * 4 + v]
    bt = np.zeros((64, 32))
    for mt in range(8):
        z3t = mfma(f[5 + 2 * mt], frag_of(h2t, 1), mfma(f[4 + 2 * mt], frag_of(h2t, 0)))
        for v in range(4):
            rr = r3[:, mt * 4 + v]
            bt[:, mt * 4 + v] = (rr * rr - rr) * z3t[:, v]
    dem = np.zeros((64, 4))
    dem[LG == 0] = (c[352:480][None, :] * dgr).sum(axis=1)  # K_c of slot 0
    tg = np.zeros((64, 4))
    for kc in range(4):
        dem = mfma(A[kc], frag_of(r3, kc), dem)
        tg = mfma(A[kc], frag_of(bt, kc), tg)
    for kc in range(2):
        dem = mfma(A[4 + kc], frag_of(h2, kc), dem)
        tg = mfma(A[6 + kc], frag_of(h2t, kc), tg)
    # lane group 0 holds slot 0: M rows c = 0..3 of each row LR
    l0 = LG == 0
    order = np.argsort(LR[l0])
    ds = (2.0 / embed.K2) * (em[LR[l0]] * tg[l0]).sum(axis=1)
    nd = dem[l0][order].copy()
    nd[:, 0] += ds[order]
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
