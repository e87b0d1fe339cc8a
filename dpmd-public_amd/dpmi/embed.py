"""Host side of the fused se_e2_a embedding kernels (csrc/embed_mfma16.hip).

pack_embedding() lays the embedding-net weights of each neighbor type out
as MFMA A-operand fragments (bf16) in the order the kernels read them, plus a
small table of per-feature constants (fp32), both in the kernels' internal
feature order:

    q1(f) = f (f < 16 or n1 > 28), else 16 + 4 ((f-16) // 3) + (f-16) % 3
                                               layer 1, width n1 <= 32
    q2(f) = q1(f)            (f <  n1)         layer 2, width 2 n1
          = 32 + q1(f - n1)  (f >= n1)
    q3(f) = q2(f)        (f <  2 n1)           layer 3, width 4 n1 = M
          = 64 + q2(f - 2 n1)

(for n1 <= 28 the layer-1 slots 19, 23, 27, 31 stay empty: one B-slot
column of every lane group, whose tanh the kernels skip in all layers)

so that the [x, x] skip connections of MLPLayer (deepmd/pt/model/network/
mlp.py:188-219) pair features on the same lane.  A fragment is 64 lanes x 8
bf16; lane l holds output feature 16*mt + l%16 and the K slots
phi(g, u) = 16*(u//4) + 4*g + u%4 of its 32-wide chunk (g = l//16).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import ops
from ._lib import ptr

N_FRAG = 20  # layer-2 (4) and layer-3 (16) fragments per type
N_CONST = 576
# tanh(x) = 1 - 2 / (2^(K2 x) + 1): the forward weights and biases are stored
# multiplied by K2 so the kernels' MFMA results feed exp2 directly
K2 = 2.0 / np.log(2.0)


def _phi():
    g = np.arange(4)[:, None]
    u = np.arange(8)[None, :]
    return 16 * (u // 4) + 4 * g + (u % 4)  # [4, 8]


def skip_layout(n1: int) -> bool:
    """n1 <= 28: layer-1 slots 19, 23, 27, 31 (B-slot u = 7 of every lane
    group) stay empty, so the kernels skip those tanh columns in all layers."""
    return n1 <= 28


def internal_maps(n1: int):
    f1 = np.arange(n1)
    if skip_layout(n1):
        q1 = np.where(f1 < 16, f1, 16 + 4 * ((f1 - 16) // 3) + (f1 - 16) % 3)
    else:
        q1 = f1
    f2 = np.arange(2 * n1)
    q2 = np.where(f2 < n1, q1[np.minimum(f2, n1 - 1)], 32 + q1[np.clip(f2 - n1, 0, n1 - 1)])
    f3 = np.arange(4 * n1)
    q3 = np.where(f3 < 2 * n1, q2[np.minimum(f3, 2 * n1 - 1)], 64 + q2[np.clip(f3 - 2 * n1, 0, 2 * n1 - 1)])
    return q1, q2, q3


def supported(neuron) -> bool:
    n1 = neuron[0]
    return len(neuron) == 3 and n1 <= 32 and neuron[1] == 2 * n1 and neuron[2] == 4 * n1


def _layer(l):
    v = l.get("@variables", l)
    w = np.asarray(v["w"], dtype=np.float64)
    b = np.asarray(v["b"], dtype=np.float64)
    idt = v.get("idt")
    idt = np.ones_like(b) if idt is None else np.asarray(idt, dtype=np.float64)
    return w, b, idt


# MFMA operand dtype -> entry-point suffix of the 16-bit kernels
SUFFIX = {torch.bfloat16: "bf16", torch.float16: "f16"}
# descriptor output dtype -> desc_kind of dpmi_se_a_descriptor_fwd
DESC_KIND = {torch.float32: 0, torch.bfloat16: 1, torch.float16: 2}


def pack_embedding(emb_layers: list, dtype=torch.bfloat16) -> tuple[np.ndarray, np.ndarray, int]:
    """emb_layers[t] = the three layer dicts of neighbor type t.
    Returns (frags uint16 [ntypes, 20, 64, 8] holding the bits of the 16-bit
    operand dtype (bfloat16 or float16),
    consts float32 [ntypes, 576], n1).  consts: K2 w1 [0:32], K2 b1 [32:64],
    idt1 [64:96], K2 b2 [96:160], idt2 [160:224], K2 b3 [224:352],
    idt3 [352:480], idt1 w1 [480:512], idt2 / K2 [512:576] (the last two
    give the backward's tangents h1' = idt1 w1 (1 - t1^2) and
    h2' = (idt2 / K2) (1 - t2^2) z2' + h1'); the fragments hold dtype(K2 W),
    which the backward's tangent products use as well, so forces are the
    gradient of the evaluated energy."""
    ntypes = len(emb_layers)
    (w1, _, _) = _layer(emb_layers[0][0])
    n1 = w1.shape[1]
    q1, q2, q3 = internal_maps(n1)
    phi = _phi()  # [g, u]
    lane = np.arange(64)
    lg, lr = lane // 16, lane % 16
    frags = np.zeros((ntypes, N_FRAG, 64, 8), dtype=np.float32)
    consts = np.zeros((ntypes, N_CONST), dtype=np.float32)
    for t, layers in enumerate(emb_layers):
        (w1, b1, i1), (w2, b2, i2), (w3, b3, i3) = (_layer(x) for x in layers)
        if w1.shape != (1, n1) or w2.shape != (n1, 2 * n1) or w3.shape != (2 * n1, 4 * n1):
            raise ValueError("fused embedding needs widths (n1, 2 n1, 4 n1) with n1 <= 32")
        W2 = np.zeros((32, 64))
        W2[np.ix_(q1, q2)] = w2
        W3 = np.zeros((64, 128))
        W3[np.ix_(q2, q3)] = w3
        c = consts[t]
        c[q1] = K2 * w1[0]
        c[32 + q1] = K2 * b1
        c[64 + q1] = i1
        c[96 + q2] = K2 * b2
        c[160 + q2] = i2
        c[224 + q3] = K2 * b3
        c[352 + q3] = i3
        c[480 + q1] = i1 * w1[0]
        c[512 + q2] = i2 / K2
        K = phi[lg]  # [64, 8] K slot -> feature offset within a 32-chunk
        f = frags[t]
        for mt in range(4):  # layer 2, A = K2 W2^T (M = q2, K = q1)
            f[mt] = K2 * W2[K, (16 * mt + lr)[:, None]]
        for mt in range(8):  # layer 3, A = K2 W3^T (M = q3, K = q2)
            for kc in range(2):
                f[4 + 2 * mt + kc] = K2 * W3[32 * kc + K, (16 * mt + lr)[:, None]]
    bits = torch.from_numpy(frags).to(dtype).view(torch.int16).numpy().view(np.uint16)
    return bits, consts, n1


class FusedEmbedding:
    """Device-resident packed weights + the four fused launches.  dtype is
    the MFMA operand type: torch.bfloat16 or torch.float16."""

    def __init__(self, emb_layers, sel, axis: int, device, dtype=torch.bfloat16):
        if dtype not in SUFFIX:
            raise ValueError("the fused embedding takes bfloat16 or float16 operands")
        self.dtype = dtype
        self._fwd = "dpmi_se_a_embed_fwd_" + SUFFIX[dtype]
        self._bwd = "dpmi_se_a_embed_bwd_" + SUFFIX[dtype]
        bits, consts, n1 = pack_embedding(emb_layers, dtype)
        self.n1 = n1
        self.M = 4 * n1
        self.axis = int(axis)
        self.sel = np.ascontiguousarray(np.asarray(sel, dtype=np.int32))
        self.ntypes = len(sel)
        self.nnei = int(self.sel.sum())
        self.frags = torch.from_numpy(bits.astype(np.int16)).to(device)
        self.consts = torch.from_numpy(consts).to(device)

    def forward_part(self, em: torch.Tensor, pad=None) -> torch.Tensor:
        """em [nloc, nnei, 4] fp32 -> gr_part [ntypes, nloc, M, 4]: each
        neighbor type's (em^T G)^T (original feature order, the four
        components of a feature contiguous, not yet / nnei).
        pad = (counts int16 [nloc, ntypes], atype int32 [nloc], pad_em fp32
        [ntypes, ntypes, 4]) enables the padding shortcut: the slots after
        each section's listed neighbors all hold pad_em[type(i), t]."""
        nloc = em.shape[0]
        gr_part = torch.empty((self.ntypes, nloc, self.M, 4), dtype=torch.float32, device=em.device)
        counts, atype, pad_em = pad if pad is not None else (None, None, None)
        ops._launch(self._fwd, ptr(gr_part), ptr(em.contiguous()), ptr(self.frags), ptr(self.consts),
                    nloc, ptr(self.sel), self.ntypes, self.n1, ptr(counts), ptr(atype), ptr(pad_em), ops._stream())
        return gr_part

    def forward(self, em: torch.Tensor, desc_dtype=torch.float32, pad=None):
        """em [nloc, nnei, 4] fp32 -> (gr [nloc, 4, M], desc [nloc, M*axis] in
        desc_dtype: float32, or the operand dtype for the fused fitting kernel)."""
        nloc = em.shape[0]
        dev = em.device
        gr_part = self.forward_part(em, pad)
        gr = torch.empty((nloc, 4, self.M), dtype=torch.float32, device=dev)
        desc = torch.empty((nloc, self.M * self.axis), dtype=desc_dtype, device=dev)
        ops._launch("dpmi_se_a_descriptor_fwd", ptr(gr), ptr(desc), ptr(gr_part), nloc, self.ntypes, self.nnei,
                    self.n1, self.axis, DESC_KIND[desc_dtype], ops._stream())
        return gr, desc

    def backward(self, em: torch.Tensor, gr: torch.Tensor, ddesc: torch.Tensor, counts=None) -> torch.Tensor:
        """dE/d desc -> dE/d em [nloc, nnei, 4] fp32 (gr as returned by forward)."""
        nloc = em.shape[0]
        dgr = torch.empty((nloc, 128, 4), dtype=torch.float32, device=em.device)
        ops._launch("dpmi_se_a_descriptor_bwd", ptr(dgr), ptr(ddesc.contiguous()), ptr(gr), nloc, self.nnei, self.n1,
                    self.axis, ops._stream())
        return self.backward_dgr(em, dgr, counts)

    def backward_dgr(self, em: torch.Tensor, dgr: torch.Tensor, counts=None) -> torch.Tensor:
        """dgr [nloc, 128, 4] = dE/d(sum_t gr_part) in the internal order ->
        dE/d em [nloc, nnei, 4] fp32 (recomputing the embedding forward).
        With counts (the forward's padding shortcut) the padding slots after
        each section's listed neighbors get 0."""
        nloc = em.shape[0]
        nd = torch.empty((nloc, self.nnei, 4), dtype=torch.float32, device=em.device)
        ops._launch(self._bwd, ptr(nd), ptr(em.contiguous()), ptr(dgr.contiguous()),
                    ptr(self.frags), ptr(self.consts), nloc, ptr(self.sel), self.ntypes, self.n1, ptr(counts),
                    ops._stream())
        return nd

    def section_counts(self, nlist: torch.Tensor) -> torch.Tensor:
        """Listed neighbors per atom and type section: int16 [nloc, ntypes]."""
        nloc = nlist.shape[0]
        counts = torch.empty((nloc, self.ntypes), dtype=torch.int16, device=nlist.device)
        ops._launch("dpmi_se_a_section_counts", ptr(counts), ptr(nlist.contiguous()), nloc, ptr(self.sel),
                    self.ntypes, ops._stream())
        return counts


# --------------------------------------------------------------------------- FP64 form


def supported_f64(neuron) -> bool:
    return supported(neuron)


def pack_embedding_f64(emb_layers: list):
    """FP64 kernels (csrc/embed_f64.hip): per type W2 [32, 64] and W3 [64, 128]
    in the internal order q1 = f, q2 = q1(f mod n1) + 32 [f >= n1],
    q3 = q2(f mod 2 n1) + 64 [f >= 2 n1] (zero padded), and consts [480]:
    w1 [0:32], b1 [32:64], idt1 [64:96], b2 [96:160], idt2 [160:224],
    b3 [224:352], idt3 [352:480].  Returns (w2, w3, consts, n1), float64."""
    ntypes = len(emb_layers)
    n1 = _layer(emb_layers[0][0])[0].shape[1]
    f1 = np.arange(n1)
    q1 = f1
    f2 = np.arange(2 * n1)
    q2 = np.where(f2 < n1, f2, 32 + f2 - n1)
    f3 = np.arange(4 * n1)
    q3 = np.where(f3 < 2 * n1, q2[np.minimum(f3, 2 * n1 - 1)], 64 + q2[np.clip(f3 - 2 * n1, 0, 2 * n1 - 1)])
    w2 = np.zeros((ntypes, 32, 64))
    w3 = np.zeros((ntypes, 64, 128))
    consts = np.zeros((ntypes, 480))
    for t, layers in enumerate(emb_layers):
        (a1, b1, i1), (a2, b2, i2), (a3, b3, i3) = (_layer(x) for x in layers)
        if a1.shape != (1, n1) or a2.shape != (n1, 2 * n1) or a3.shape != (2 * n1, 4 * n1):
            raise ValueError("fused embedding needs widths (n1, 2 n1, 4 n1) with n1 <= 32")
        w2[t][np.ix_(q1, q2)] = a2
        w3[t][np.ix_(q2, q3)] = a3
        c = consts[t]
        c[q1] = a1[0]
        c[32 + q1] = b1
        c[64 + q1] = i1
        c[96 + q2] = b2
        c[160 + q2] = i2
        c[224 + q3] = b3
        c[352 + q3] = i3
    return w2, w3, consts, n1


class FusedEmbeddingF64:
    """FP64 fused embedding + G^T R and its recompute adjoint (MFMA f64)."""

    def __init__(self, emb_layers, sel, device):
        w2, w3, consts, n1 = pack_embedding_f64(emb_layers)
        self.n1 = n1
        self.M = 4 * n1
        self.sel = np.ascontiguousarray(np.asarray(sel, dtype=np.int32))
        self.ntypes = len(sel)
        self.nnei = int(self.sel.sum())
        self.w2 = torch.from_numpy(w2).to(device)
        self.w3 = torch.from_numpy(w3).to(device)
        self.consts = torch.from_numpy(consts).to(device)

    def forward(self, em: torch.Tensor, pad=None) -> torch.Tensor:
        """em [nloc, nnei, 4] float64 -> GR [nloc, 4, M] = sum_t em_t^T G_t / nnei.
        pad = (counts, atype, pad_em float64): the padding shortcut."""
        nloc = em.shape[0]
        em = em.contiguous()
        part = torch.empty((self.ntypes, nloc, 4, self.M), dtype=torch.float64, device=em.device)
        counts, atype, pad_em = pad if pad is not None else (None, None, None)
        ops._launch("dpmi_se_a_embed_fwd_f64", ptr(part), ptr(em), ptr(self.w2), ptr(self.w3), ptr(self.consts),
                    nloc, ptr(self.sel), self.ntypes, self.n1, ptr(counts), ptr(atype), ptr(pad_em), ops._stream())
        return part.sum(dim=0) / self.nnei

    def backward(self, em: torch.Tensor, dgr: torch.Tensor, counts=None) -> torch.Tensor:
        """dE/dGR [nloc, 4, M] -> dE/d em [nloc, nnei, 4] float64 (padding slots
        0 with counts)."""
        nloc = em.shape[0]
        dgr_t = (dgr / self.nnei).contiguous()
        nd = torch.empty((nloc, self.nnei, 4), dtype=torch.float64, device=em.device)
        ops._launch("dpmi_se_a_embed_bwd_f64", ptr(nd), ptr(em.contiguous()), ptr(dgr_t), ptr(self.w2),
                    ptr(self.w3), ptr(self.consts), nloc, ptr(self.sel), self.ntypes, self.n1, ptr(counts),
                    ops._stream())
        return nd

    def forward_part(self, em: torch.Tensor, pad=None) -> torch.Tensor:
        """em [nloc, nnei, 4] float64 -> per-type partials [ntypes, nloc, 4, M]
        of sum_rows em_t^T G_t (GR = their sum / nnei), as the FP64 fitting
        kernel takes them."""
        nloc = em.shape[0]
        em = em.contiguous()
        part = torch.empty((self.ntypes, nloc, 4, self.M), dtype=torch.float64, device=em.device)
        counts, atype, pad_em = pad if pad is not None else (None, None, None)
        ops._launch("dpmi_se_a_embed_fwd_f64", ptr(part), ptr(em), ptr(self.w2), ptr(self.w3), ptr(self.consts),
                    nloc, ptr(self.sel), self.ntypes, self.n1, ptr(counts), ptr(atype), ptr(pad_em), ops._stream())
        return part

    def backward_part(self, em: torch.Tensor, dpart: torch.Tensor, counts=None) -> torch.Tensor:
        """dE/d(sum of the partials) [nloc, 4, M] float64 (what the FP64
        fitting kernel writes) -> dE/d em [nloc, nnei, 4]."""
        nloc = em.shape[0]
        nd = torch.empty((nloc, self.nnei, 4), dtype=torch.float64, device=em.device)
        ops._launch("dpmi_se_a_embed_bwd_f64", ptr(nd), ptr(em.contiguous()), ptr(dpart.contiguous()), ptr(self.w2),
                    ptr(self.w3), ptr(self.consts), nloc, ptr(self.sel), self.ntypes, self.n1, ptr(counts),
                    ops._stream())
        return nd
