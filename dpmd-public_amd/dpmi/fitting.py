"""Host side of the fused energy fitting kernel (csrc/fitting_mfma16.hip).

pack_fitting() lays each type's fitting net (three hidden layers of width
H <= 256, the second and third with the MLPLayer skip connection, a linear
H -> 1 output; deepmd/dpmodel/utils/network.py:1031-1090) out as MFMA A-operand
fragments (bf16) in the order the kernel reads them, padded with zeros to the
kernel's hidden width 256.  A fragment is 64 lanes x 8 bf16; lane l = 16 g + m
holds row 16 mt + m of the A matrix and its K slots 32 kc + 8 g + u:

    w1f  A = W1^T   (M = hidden feature, K = descriptor feature)   16 x nk1
    w2f  A = W2^T,  w3f  A = W3^T                                  16 x 8 each
    w3b  A = W3,    w2b  A = W2      (M = input, K = output)       16 x 8 each
    w1b  A = W1     (M = descriptor feature, K = hidden feature)   k1/16 x 8

consts per type (fp32, 7*256 + 8): b1 idt1 b2 idt2 b3 idt3 w4 and b4 (idt = 1
where a layer has no timestep), and ebias (fp64) = bias_atom_e + out_bias.
"""

from __future__ import annotations

import numpy as np
import torch

from . import ops
from ._lib import ptr

KH = 256
KKB = 64
N_CONST = 7 * KH + 8


def _layer(l):
    v = l.get("@variables", l)
    w = np.asarray(v["w"], dtype=np.float64)
    b = np.asarray(v["b"], dtype=np.float64).reshape(-1)
    idt = v.get("idt")
    idt = None if idt is None else np.asarray(idt, dtype=np.float64).reshape(-1)
    return w, b, idt


def supported(fit_layers: list) -> bool:
    """Three hidden layers of one width H <= 256 (in -> H without skip, then
    H -> H with skip) and a linear H -> 1 output layer, for every type."""
    for layers in fit_layers:
        if len(layers) != 4:
            return False
        ws = [_layer(l)[0] for l in layers]
        k1, h = ws[0].shape
        if not (h <= KH and k1 % 16 == 0 and k1 != h and 2 * k1 != h):
            return False
        if ws[1].shape != (h, h) or ws[2].shape != (h, h) or ws[3].shape != (h, 1):
            return False
    return True


def _frags(A: np.ndarray, nm: int, nk: int) -> np.ndarray:
    """A [rows, cols] -> fragments [nm, nk, 64, 8] of the zero-padded
    [16 nm, 32 nk] matrix; lane 16 g + m, slot u = A[16 mt + m, 32 kc + 8 g + u]."""
    P = np.zeros((16 * nm, 32 * nk), dtype=np.float64)
    r, c = min(A.shape[0], P.shape[0]), min(A.shape[1], P.shape[1])
    P[:r, :c] = A[:r, :c]
    return P.reshape(nm, 16, nk, 4, 8).transpose(0, 2, 3, 1, 4).reshape(nm, nk, 64, 8)


def pack_fitting(fit_layers: list, bias_atom_e, out_bias=None, dtype=torch.bfloat16):
    """Returns (frags uint16 [ntypes, nfrag, 64, 8] holding the bits of the
    16-bit operand dtype (bfloat16 or float16), consts float32
    [ntypes, N_CONST], ebias float64 [ntypes], k1)."""
    if not supported(fit_layers):
        raise ValueError("fused fitting needs hidden widths (H, H, H), H <= 256, and a linear H -> 1 output")
    ntypes = len(fit_layers)
    k1 = _layer(fit_layers[0][0])[0].shape[0]
    nk1 = 2 * ((k1 + KKB - 1) // KKB)
    nm1 = k1 // 16
    blocks, consts = [], np.zeros((ntypes, N_CONST), dtype=np.float32)
    for t, layers in enumerate(fit_layers):
        (w1, b1, i1), (w2, b2, i2), (w3, b3, i3), (w4, b4, _) = (_layer(x) for x in layers)
        if w1.shape[0] != k1:
            raise ValueError("every type's fitting net must take the same descriptor width")
        h = w1.shape[1]
        parts = [
            _frags(w1.T, 16, nk1),
            _frags(w2.T, 16, 8),
            _frags(w3.T, 16, 8),
            _frags(w3, 16, 8),
            _frags(w2, 16, 8),
            _frags(w1, nm1, 8),
        ]
        blocks.append(np.concatenate([p.reshape(-1, 64, 8) for p in parts]))
        c = consts[t]
        for k, v in enumerate((b1, i1, b2, i2, b3, i3, w4[:, 0])):
            c[k * KH:k * KH + h] = 1.0 if v is None else v
        c[7 * KH] = b4[0] if b4.size else 0.0
    frags = np.stack(blocks).astype(np.float32)
    bits = torch.from_numpy(frags).to(dtype).view(torch.int16).numpy().view(np.uint16)
    ob = np.zeros(ntypes) if out_bias is None else np.asarray(out_bias, dtype=np.float64).reshape(-1)
    ebias = np.asarray(bias_atom_e, dtype=np.float64).reshape(-1) + ob
    return bits, consts, ebias, k1


class FusedFitting:
    """Device-resident packed fitting nets + the fused launch.  dtype is the
    MFMA operand type: torch.bfloat16 or torch.float16."""

    def __init__(self, fit_layers, bias_atom_e, out_bias, device, dtype=torch.bfloat16):
        sfx = {torch.bfloat16: "bf16", torch.float16: "f16"}.get(dtype)
        if sfx is None:
            raise ValueError("the fused fitting takes bfloat16 or float16 operands")
        self.dtype = dtype
        self._fit = "dpmi_se_a_fitting_" + sfx
        self._fit_gr = "dpmi_se_a_fitting_gr_" + sfx
        bits, consts, ebias, k1 = pack_fitting(fit_layers, bias_atom_e, out_bias, dtype)
        self.k1 = k1
        self.ntypes = len(fit_layers)
        self.frags = torch.from_numpy(bits.astype(np.int16)).to(device)
        self.consts = torch.from_numpy(consts).to(device)
        self.ebias = torch.from_numpy(ebias).to(device)

    def sort_by_type(self, atype: torch.Tensor):
        """(perm int32 [nloc] atoms ordered by type, type_start int32 [ntypes+1])."""
        a = atype.to(torch.int64)
        perm = torch.argsort(a, stable=True).to(torch.int32)
        counts = torch.zeros(self.ntypes, dtype=torch.int64, device=a.device)
        counts.scatter_add_(0, a, torch.ones_like(a))
        start = torch.zeros(self.ntypes + 1, dtype=torch.int32, device=a.device)
        start[1:] = torch.cumsum(counts, 0).to(torch.int32)
        return perm, start

    def __call__(self, desc: torch.Tensor, atype: torch.Tensor, order=None):
        """desc [nloc, k1] in the operand dtype -> (atom_energy float64 [nloc],
        d E / d desc fp32 [nloc, k1])."""
        nloc = desc.shape[0]
        if desc.shape[1] != self.k1:
            raise ValueError(f"descriptor width {desc.shape[1]} != fitting input {self.k1}")
        if desc.dtype != self.dtype:
            raise TypeError(f"the fused fitting kernel takes the descriptor in {self.dtype}")
        perm, start = self.sort_by_type(atype) if order is None else order
        desc = desc.contiguous()
        ae = torch.empty(nloc, dtype=torch.float64, device=desc.device)
        dd = torch.empty((nloc, self.k1), dtype=torch.float32, device=desc.device)
        ops._launch(self._fit, ptr(ae), ptr(dd), ptr(desc), ptr(perm), ptr(start), ptr(self.frags),
                    ptr(self.consts), ptr(self.ebias), nloc, self.ntypes, self.k1, ops._stream())
        return ae, dd

    def supports_gr(self, n1: int, axis: int) -> bool:
        """Whether the fused descriptor form applies: axis 16, 16 <= 4 n1 <= 128,
        at most 4 types (GR staged in fp32 up to 4 n1 = 104, in fp16 above)."""
        return axis == 16 and 4 <= n1 <= 32 and self.k1 == 64 * n1 and self.ntypes <= 4

    def from_gr(self, gr_part: torch.Tensor, atype: torch.Tensor, n1: int, nnei: int, order=None):
        """Fused descriptor form: gr_part [ntypes, nloc, 4 n1, 4] fp32 (the
        embedding forward's partials) -> (atom_energy float64 [nloc],
        dE/d(sum_t gr_part) fp32 [nloc, 128, 4] in the embedding kernels'
        internal order)."""
        nloc = gr_part.shape[1]
        if gr_part.dtype != torch.float32 or gr_part.shape[2:] != (4 * n1, 4):
            raise ValueError("gr_part must be float32 [ntypes, nloc, 4 n1, 4]")
        perm, start = self.sort_by_type(atype) if order is None else order
        gr_part = gr_part.contiguous()
        ae = torch.empty(nloc, dtype=torch.float64, device=gr_part.device)
        dgr = torch.empty((nloc, 128, 4), dtype=torch.float32, device=gr_part.device)
        ops._launch(self._fit_gr, ptr(ae), ptr(dgr), ptr(gr_part), ptr(perm), ptr(start),
                    ptr(self.frags), ptr(self.consts), ptr(self.ebias), nloc, self.ntypes, n1, nnei, ops._stream())
        return ae, dgr


# --------------------------------------------------------------------------- FP64 form (csrc/fitting_f64.hip)


def _frags64(A: np.ndarray, nm: int, nk: int) -> np.ndarray:
    """A [rows, cols] -> MFMA f64 16x16x4 A fragments [nm, nk, 64] of the
    zero-padded [16 nm, 4 nk] matrix: lane l = 16 kk + r holds
    A[16 mt + r, 4 ks + kk]."""
    P = np.zeros((16 * nm, 4 * nk), dtype=np.float64)
    r, c = min(A.shape[0], P.shape[0]), min(A.shape[1], P.shape[1])
    P[:r, :c] = A[:r, :c]
    return P.reshape(nm, 16, nk, 4).transpose(0, 2, 3, 1).reshape(nm, nk, 64)


def supported_f64(fit_layers: list, M: int, axis: int) -> bool:
    """The FP64 kernel's shapes: the fused-fitting structure above with the
    descriptor D = GR^T GR[:, :16] of a G^T R width 16 <= M <= 128."""
    if axis != 16 or not (16 <= M <= 128) or not supported(fit_layers):
        return False
    return all(_layer(layers[0])[0].shape[0] == 16 * M for layers in fit_layers)


def pack_fitting_f64(fit_layers: list, bias_atom_e, out_bias, M: int):
    """(frags float64 [ntypes, per-type doubles], consts float64
    [ntypes, 7*256 + 1], ebias float64 [ntypes], H) in the order the FP64
    kernel reads them, per type:

        w1f  A = W1^T  [16 tiles, 4 M K steps]   w2f, w3f  A = W2^T, W3^T  [16, nkh]
        w3b, w2b  A = W3, W2  [16, nkh]          w1b  A = W1  [M tiles, nkh]

    (nkh = ceil(H / 4)); consts b1 idt1 b2 idt2 b3 idt3 w4 [256 each], b4."""
    if not supported_f64(fit_layers, M, 16):
        raise ValueError("FP64 fused fitting needs axis 16, 16 <= M <= 128 and hidden widths (H, H, H), H <= 256")
    ntypes = len(fit_layers)
    h = _layer(fit_layers[0][0])[0].shape[1]
    nkh = (h + 3) // 4
    blocks, consts = [], np.zeros((ntypes, 7 * KH + 1), dtype=np.float64)
    for t, layers in enumerate(fit_layers):
        (w1, b1, i1), (w2, b2, i2), (w3, b3, i3), (w4, b4, _) = (_layer(x) for x in layers)
        if w1.shape[1] != h:
            raise ValueError("every type's fitting net must have the same hidden width")
        parts = [_frags64(w1.T, 16, 4 * M), _frags64(w2.T, 16, nkh), _frags64(w3.T, 16, nkh),
                 _frags64(w3, 16, nkh), _frags64(w2, 16, nkh), _frags64(w1, M, nkh)]
        blocks.append(np.concatenate([x.reshape(-1) for x in parts]))
        c = consts[t]
        for k, v in enumerate((b1, i1, b2, i2, b3, i3, w4[:, 0])):
            c[k * KH:k * KH + h] = 1.0 if v is None else v
        c[7 * KH] = b4[0] if b4.size else 0.0
    ob = np.zeros(ntypes) if out_bias is None else np.asarray(out_bias, dtype=np.float64).reshape(-1)
    ebias = np.asarray(bias_atom_e, dtype=np.float64).reshape(-1) + ob
    return np.stack(blocks), consts, ebias, h


class FusedFitting64:
    """Device-resident FP64 fitting nets + the fused descriptor / fitting /
    adjoint launch (dpmi_se_a_fitting_gr_f64)."""

    def __init__(self, fit_layers, bias_atom_e, out_bias, M: int, device):
        frags, consts, ebias, h = pack_fitting_f64(fit_layers, bias_atom_e, out_bias, M)
        import ctypes

        n = ctypes.c_int64(0)
        ops.call("dpmi_se_a_fitting_f64_frag_doubles", ctypes.byref(n), M, h)
        if frags.shape[1] != n.value:
            raise RuntimeError(f"packed {frags.shape[1]} doubles per type, the kernel reads {n.value}")
        self.M, self.H, self.ntypes = M, h, len(fit_layers)
        self.frags = torch.from_numpy(frags).to(device)
        self.consts = torch.from_numpy(consts).to(device)
        self.ebias = torch.from_numpy(ebias).to(device)

    sort_by_type = FusedFitting.sort_by_type

    def from_parts(self, gr_part: torch.Tensor, atype: torch.Tensor, nnei: int, order=None):
        """gr_part [nparts, nloc, 4, M] float64 (GR = sum over parts / nnei)
        -> (atom_energy float64 [nloc], dE/d(sum of parts) float64
        [nloc, 4, M])."""
        if gr_part.dtype != torch.float64 or gr_part.dim() != 4 or tuple(gr_part.shape[2:]) != (4, self.M):
            raise ValueError(f"gr_part must be float64 [nparts, nloc, 4, {self.M}]")
        nparts, nloc = gr_part.shape[:2]
        perm, start = self.sort_by_type(atype) if order is None else order
        gr_part = gr_part.contiguous()
        ae = torch.empty(nloc, dtype=torch.float64, device=gr_part.device)
        dgr = torch.empty((nloc, 4, self.M), dtype=torch.float64, device=gr_part.device)
        ops._launch("dpmi_se_a_fitting_gr_f64", ptr(ae), ptr(dgr), ptr(gr_part), nparts, ptr(perm), ptr(start),
                    ptr(self.frags), ptr(self.consts), ptr(self.ebias), nloc, self.ntypes, self.M, self.H, nnei,
                    ops._stream())
        return ae, dgr


def energy_sum(atom_energy: torch.Tensor) -> torch.Tensor:
    """Compensated (Neumaier), deterministic FP64 sum of atomic energies on
    the GPU (dpmi_energy_sum_f64): a 0-d float64 tensor."""
    ae = atom_energy.to(torch.float64).contiguous()
    out = torch.empty((), dtype=torch.float64, device=ae.device)
    ops._launch("dpmi_energy_sum_f64", ptr(out), ptr(ae), ae.numel(), ops._stream())
    return out
