"""se_e2_a energy model on the GPU: energy, forces and virial per step.

Mirrors the composition of the reference deepmd/pt model
(deepmd/pt/model/descriptor/se_a.py:750-866, network/mlp.py:188-219,
task/fitting.py:734-900, model/transform_output.py:65-206) with the data
path moved onto libdpmi:

  coordinates --nlist kernels--> formatted list (bit-exact C++ semantics)
              --env-mat kernel--> em [nloc, nnei, 4]
              --embedding + G^T R, D, fitting--> atomic energies, E (FP64 sum)
  dE/d em     --recompute adjoint kernel--> forces (ghosts folded), virial

Forces are the hand-written adjoint of the env-mat (the derivative is
recomputed, never stored); the network part of dE/d em comes from the
embedding/fitting adjoint.  Precision forms:

* ``fp64``  — everything in double ("plain FP64").
* ``fp32``  — the same path in float (pair differences in double), energy
  reduction in fp64.
* ``bf16`` / ``fp16`` — embedding, descriptor and fitting GEMM operands in
  bf16/fp16 with fp32 accumulation, env-mat and its adjoint in fp32, energy
  and virial reductions in fp64.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, embed, fitting, ops

_ACT = torch.tanh


class _LowPrecisionMM(torch.autograd.Function):
    """y = x @ w with bf16/fp16 operands, fp32 accumulation and an fp32
    result; the backward products use the same operand precision."""

    @staticmethod
    def forward(ctx, x, w, gemm_dtype):
        xl, wl = x.to(gemm_dtype), w.to(gemm_dtype)
        ctx.save_for_backward(xl, wl)
        ctx.out = x.dtype
        return torch.mm(xl, wl, out_dtype=x.dtype)

    @staticmethod
    def backward(ctx, gy):
        xl, wl = ctx.saved_tensors
        gl = gy.to(xl.dtype)
        gx = torch.mm(gl, wl.t().contiguous(), out_dtype=ctx.out) if ctx.needs_input_grad[0] else None
        gw = torch.mm(xl.t().contiguous(), gl, out_dtype=ctx.out) if ctx.needs_input_grad[1] else None
        return gx, gw, None


class _LowPrecisionBMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, gemm_dtype):
        al, bl = a.to(gemm_dtype).contiguous(), b.to(gemm_dtype).contiguous()
        ctx.save_for_backward(al, bl)
        ctx.out = a.dtype
        return torch.bmm(al, bl, out_dtype=a.dtype)

    @staticmethod
    def backward(ctx, gy):
        al, bl = ctx.saved_tensors
        gl = gy.to(al.dtype).contiguous()
        ga = torch.bmm(gl, bl.transpose(1, 2).contiguous(), out_dtype=ctx.out) if ctx.needs_input_grad[0] else None
        gb = torch.bmm(al.transpose(1, 2).contiguous(), gl, out_dtype=ctx.out) if ctx.needs_input_grad[1] else None
        return ga, gb, None


_SPLIT_ROWS = 1 << 15


def _split_k_linear(x, w, b):
    """x @ w + b as a batched product over row blocks with the weight
    broadcast, so the weight gradient is a batched product summed over the
    blocks (split K) and stays differentiable for the double backward.  The
    bias is broadcast by ops.RowBroadcast, whose gradient is the fixed-order
    HIP row sum (no library multi-block reduction: those clear a semaphore
    with a memory-set node, which a captured training step must not hold)."""
    rows = x.shape[0]
    nb = min(256, rows // 1024)
    pad = (-rows) % nb
    if pad:
        x = torch.cat([x, x.new_zeros((pad, x.shape[1]))])
    y = torch.bmm(x.view(nb, -1, x.shape[1]), w.unsqueeze(0).expand(nb, -1, -1)).reshape(-1, w.shape[1])
    if pad:
        y = y[:rows]
    return y + ops.RowBroadcast.apply(b, rows)


class Net:
    """Stack of reference MLP layers (MLPLayer.forward, mlp.py:188-219): y =
    tanh(x W + b) [* idt] [+ x or + (x, x)]; the last layer of a fitting net is
    linear without skip."""

    def __init__(self, layers, device, dtype, last_linear: bool):
        self.w, self.b, self.idt, self.act = [], [], [], []
        for k, l in enumerate(layers):
            v = l.get("@variables", l)
            self.w.append(torch.as_tensor(np.asarray(v["w"], dtype=np.float64), device=device, dtype=dtype))
            self.b.append(torch.as_tensor(np.asarray(v["b"], dtype=np.float64), device=device, dtype=dtype))
            idt = v.get("idt")
            self.idt.append(None if idt is None else torch.as_tensor(np.asarray(idt, dtype=np.float64),
                                                                     device=device, dtype=dtype))
            self.act.append(not (last_linear and k == len(layers) - 1))

    def __call__(self, x, gemm_dtype=None):
        for w, b, idt, act in zip(self.w, self.b, self.idt, self.act):
            rows_bc = False
            if gemm_dtype is None and w.requires_grad and x.dim() == 2 and x.shape[0] >= _SPLIT_ROWS:
                # training on neighbor rows: dL/dW would be one GEMM with a tiny
                # output and K = all rows (a single workgroup); split K
                y = _split_k_linear(x, w, b)
                rows_bc = True
            elif gemm_dtype is None or w.shape[1] < 16:
                # fp64 path, or the 240 -> 1 output layer: a per-atom dot product
                # (the low-precision N=1 GEMM path of the BLAS is pathologically slow)
                y = torch.addmm(b, x, w) if x.dim() == 2 else x @ w + b
            else:
                # low-precision operands, fp32 accumulation and fp32 result
                y = _LowPrecisionMM.apply(x, w, gemm_dtype) + b
            if act:
                y = _ACT(y)
            if idt is not None:
                y = y * (ops.RowBroadcast.apply(idt, y.shape[0]) if rows_bc else idt)
            if act:
                nin, nout = w.shape
                if nin == nout:
                    y = y + x
                elif 2 * nin == nout:
                    y = y + torch.cat([x, x], dim=-1)
            x = y
        return x


class SeAEnergyModel:
    # neighbor rows per block of the FP64/FP32 network evaluation
    ROWS_PER_BLOCK = 1 << 22

    def __init__(self, rcut, rcut_smth, sel, axis, davg, dstd, emb, fit, bias_atom_e, out_bias=None,
                 device="cuda", precision: str = "fp64", tables=None):
        if precision not in ("fp64", "fp32", "bf16", "fp16"):
            raise ValueError(f"precision must be fp64, fp32, bf16 or fp16, not {precision}")
        self.device = torch.device(device)
        self.precision = precision
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        self.gemm_dtype = {"fp64": None, "fp32": None, "bf16": torch.bfloat16, "fp16": torch.float16}[precision]
        self.rcut = float(rcut)
        self.rcut_smth = float(rcut_smth)
        self.sel = [int(s) for s in sel]
        self.sec = np.concatenate([[0], np.cumsum(self.sel)]).astype(int)
        self._sel_i32 = np.ascontiguousarray(np.asarray(self.sel, dtype=np.int32))
        self.nnei = int(sum(self.sel))
        self.ntypes = len(self.sel)
        self.axis = int(axis)
        dt = self.dtype
        self.davg = torch.as_tensor(np.asarray(davg, dtype=np.float64), device=self.device, dtype=dt)
        self.dstd = torch.as_tensor(np.asarray(dstd, dtype=np.float64), device=self.device, dtype=dt)
        self.emb = [Net(layers, self.device, dt, last_linear=False) for layers in emb]
        neuron = [w.shape[1] for w in self.emb[0].w]
        # compressed form: per neighbor type (table, table_info) of the
        # tabulated embedding (dpmi.table.build_tables), fp64 or fp32
        self.tables = None
        if tables is not None:
            if precision == "bf16" or precision == "fp16":
                raise ValueError("the tabulated embedding runs in fp64 or fp32")
            self.tables = [(torch.as_tensor(np.asarray(t), device=self.device, dtype=dt),
                            torch.as_tensor(np.asarray(i), dtype=dt)) for t, i in tables]
        # FP64 form: fused FP64-MFMA embedding kernels for E+F+V (the torch
        # layers above stay for training, which differentiates the forces)
        self.fused64 = None
        if precision == "fp64" and tables is None and embed.supported_f64(neuron):
            self.fused64 = embed.FusedEmbeddingF64(emb, self.sel, self.device)
        # bf16 / fp16 forms: fused MFMA embedding kernels (else the torch layers above)
        self.fused = None
        if precision in ("bf16", "fp16") and embed.supported(neuron) and self.axis <= 16:
            self.fused = embed.FusedEmbedding(emb, self.sel, self.axis, self.device, self.gemm_dtype)
        self.fit = [Net(layers, self.device, dt, last_linear=True) for layers in fit]
        self.M = self.emb[0].w[-1].shape[1]
        self.bias_atom_e = torch.as_tensor(np.asarray(bias_atom_e, dtype=np.float64).reshape(-1),
                                           device=self.device)
        ob = np.zeros(self.ntypes) if out_bias is None else np.asarray(out_bias, dtype=np.float64).reshape(-1)
        self.out_bias = torch.as_tensor(ob, device=self.device)
        # bf16 / fp16 forms: fused MFMA fitting kernel (else the torch layers above)
        self.fused_fit = None
        if self.fused is not None and fitting.supported(fit):
            self.fused_fit = fitting.FusedFitting(fit, bias_atom_e, ob, self.device, self.gemm_dtype)
        # the descriptor fused into the fitting kernel (axis 16, M <= 128)
        self.fused_gr = self.fused_fit is not None and self.fused_fit.supports_gr(self.fused.n1, self.axis)
        # FP64 form (network or table embedding): the descriptor, the fitting
        # nets and their adjoint as one FP64-MFMA kernel (else torch, _fit_gr64)
        self.fused_fit64 = None
        if precision == "fp64" and fitting.supported_f64(fit, self.M, self.axis):
            self.fused_fit64 = fitting.FusedFitting64(fit, bias_atom_e, ob, self.M, self.device)
        self._order_cache = None
        self._host_order_cache = None
        self._types_checked = None
        self._status_host = None
        self._status_event = None
        # padding shortcut of the fused embedding kernels (each davg / dstd
        # section constant, so all padding slots of a section share one em row)
        self.pad_em = self._padding_rows() if (self.fused is not None or self.fused64 is not None) else None

    def named_parameters(self):
        """[((kind, net, layer, name), tensor)] of the trainable tensors, made
        leaves that require grad: w, b and idt of every embedding layer (not
        with tables: the compressed embedding is fixed, as in the reference)
        and every fitting layer."""
        groups = [("fitting", self.fit)]
        if self.tables is None:
            groups.insert(0, ("embedding", self.emb))
        out = []
        for kind, nets in groups:
            for n, net in enumerate(nets):
                for k in range(len(net.w)):
                    for name in ("w", "b", "idt"):
                        t = getattr(net, name)[k]
                        if t is not None:
                            out.append(((kind, n, k, name), t.requires_grad_(True)))
        return out

    def parameters(self):
        return [t for _, t in self.named_parameters()]

    # ----------------------------------------------------------------- constructors
    @classmethod
    def from_params(cls, spec, params, device="cuda", precision="fp64", tables=None):
        return cls(spec.rcut, spec.rcut_smth, spec.sel, spec.axis_neuron, params["davg"], params["dstd"],
                   params["embedding"], params["fitting"], params["bias_atom_e"], device=device,
                   precision=precision, tables=tables)

    @classmethod
    def from_serialized(cls, m: dict, device="cuda", precision="fp64"):
        """From the 'model' block of the reference serialization
        (deepmd/pt/model/descriptor/se_a.py:392-443 layout)."""
        d = m["descriptor"]
        f = m["fitting"]
        emb = [net["layers"] for net in d["embeddings"]["networks"]]
        fit = [net["layers"] for net in f["nets"]["networks"]]
        return cls(d["rcut"], d["rcut_smth"], d["sel"], d["axis_neuron"], d["@variables"]["davg"],
                   d["@variables"]["dstd"], emb, fit, f["@variables"]["bias_atom_e"],
                   m["@variables"]["out_bias"], device=device, precision=precision)

    # ----------------------------------------------------------------- pieces
    def neighbor_list(self, coord, atype, box):
        """(ext_coord, ext_type, mapping, nlist, overflow, None, cells); cells
        = (candidate slots, workspace) for the cell-wise force kernel.  With a
        periodic box the pipeline runs without host waits (the extended
        arrays are sized by the capacity bound); its status is read by
        check_status."""
        out = ops.build_nlist(coord, atype, box, self.rcut, self.sel, cells=True, wait=box is None)
        if box is not None and not torch.cuda.is_current_stream_capturing():
            st = ops.nlist_status(out[6][1])
            if self._status_host is None:
                self._status_host = torch.zeros(1, dtype=torch.int32).pin_memory()
            self._status_host.copy_(st, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            self._status_event = ev
        return out

    def check_status(self) -> None:
        """Raise if the last neighbor list built without host waits reported
        an error (its forces are NaN then).  compute() calls it for the
        previous step; call it after the last step (waits for that step's
        list).  A step replayed from a captured graph is not tracked: check
        its forces for NaN."""
        ev = self._status_event
        if ev is None or torch.cuda.is_current_stream_capturing():
            return
        ev.synchronize()
        self._status_event = None
        code = int(self._status_host[0])
        if code:
            what = []
            if code & 1:
                what.append("an atom has more than 1024 candidates within rcut")
            if code & 2:
                what.append("the neighbor-list workspace belongs to another system")
            if code & 4:
                what.append("more ghost atoms than the capacity bound")
            raise RuntimeError(f"neighbor list failed (status {code}: {'; '.join(what)}); its forces are NaN")

    def forces(self, net_deriv, nlist, cells, mapping, ext_c, ext_t, box):
        return ops.force_virial_cells(net_deriv, nlist, cells, mapping, ext_c, ext_t, self.dstd, self.rcut,
                                      self.rcut_smth, self.sel, box)

    def env_mat(self, nlist, ext_c, ext_t):
        em, _, _ = ops.env_mat_a(nlist, ext_c, ext_t, self.davg, self.dstd, self.rcut, self.rcut_smth, self.sel,
                                 with_deriv=False, with_rij=False, dtype=self.dtype)
        return em

    def descriptor(self, em):
        """G^T R / nnei and D = (G^T R)^T (G^T R)[:, :axis] (se_a.py:802-858)."""
        nloc = em.shape[0]
        gr = None
        for t in range(self.ntypes):
            rr = em[:, self.sec[t]:self.sec[t + 1], :]
            if self.tables is not None:
                # tabulate_fusion_se_a in place of embedding + G^T R (se_a.py:833-839)
                table, info = self.tables[t]
                c = ops.TabulateFusionSeAOp.apply(table, info, rr[:, :, :1].reshape(-1, 1), rr.contiguous(),
                                                  self.M)
                gr = c if gr is None else gr + c
                continue
            g = self.emb[t](rr[:, :, :1].reshape(-1, 1), self.gemm_dtype).reshape(nloc, -1, self.M)
            if self.gemm_dtype is None:
                c = torch.bmm(rr.transpose(1, 2), g)
            else:
                c = _LowPrecisionBMM.apply(rr.transpose(1, 2), g, self.gemm_dtype)
            gr = c if gr is None else gr + c
        gr = gr / self.nnei
        d = torch.bmm(gr.transpose(1, 2), gr[:, :, : self.axis])
        return d.reshape(nloc, -1)

    def atom_energy(self, desc, atype, order=None):
        """Per-atom energies of desc [nloc, k1] with each type's fitting net.
        order = (perm, bounds) of _host_type_order: one gather into
        type-sorted rows, contiguous slices per type and one scatter back,
        without the per-type host synchronisation of nonzero()."""
        if order is not None:
            perm, bounds = order
            ds = desc.index_select(0, perm)
            ys = []
            for t in range(self.ntypes):
                if bounds[t + 1] > bounds[t]:
                    y = self.fit[t](ds[bounds[t]:bounds[t + 1]], self.gemm_dtype)[:, 0].to(torch.float64)
                    ys.append(y + self.bias_atom_e[t] + self.out_bias[t])
            es = torch.cat(ys) if ys else ds.new_zeros(0, dtype=torch.float64)
            return torch.zeros(desc.shape[0], dtype=torch.float64, device=desc.device).index_add(0, perm, es)
        e = torch.empty(desc.shape[0], dtype=torch.float64, device=desc.device)
        for t in range(self.ntypes):
            idx = (atype == t).nonzero().squeeze(1)
            if idx.numel() == 0:
                continue
            y = self.fit[t](desc.index_select(0, idx), self.gemm_dtype)[:, 0].to(torch.float64)
            e.index_copy_(0, idx, y + self.bias_atom_e[t] + self.out_bias[t])
        return e

    # ----------------------------------------------------------------- E + F + V
    def check_types(self, atype: torch.Tensor) -> None:
        """Raise ValueError unless every atom type lies in [0, ntypes).  The
        check reads the range back once per atom-type tensor (cached while the
        same, unmodified tensor comes back).  The reference treats type < 0 as
        a virtual atom (make_atom_mask, deepmd/pt/model/atomic_model/
        base_atomic_model.py:278-298) and indexes its per-type nets with every
        other type; this model path takes real atoms only."""
        c = self._types_checked
        if c is not None and c[0] is atype and c[1] == atype._version:
            return
        if atype.numel():
            lo, hi = (int(x) for x in torch.aminmax(atype.reshape(-1)))
            if lo < 0 or hi >= self.ntypes:
                raise ValueError(f"atom types must lie in [0, {self.ntypes}), got types {lo}..{hi}; virtual atoms "
                                 "(type < 0) are not supported by the model path: leave them out of the frame")
        self._types_checked = (atype, atype._version)

    def compute(self, coord: torch.Tensor, atype: torch.Tensor, box, need_atom_energy: bool = False):
        """One E+F+V evaluation of one frame.  coord [nloc, 3] float64 on the
        GPU, atype [nloc] int with every type in [0, ntypes); box: 9 values or
        None.  The neighbor list is rebuilt every call."""
        self.check_types(atype)
        self.check_status()
        atype_in = atype
        atype = atype.to(torch.int32)
        ext_c, ext_t, mapping, nlist, _ovf, _, cells = self.neighbor_list(coord, atype, box)
        if self.fused is not None:
            return self._compute_fused(atype, ext_c, ext_t, mapping, nlist, cells, box, need_atom_energy,
                                       order=self._type_order(atype_in, atype))
        if self.fused64 is not None:
            return self._compute_fused64(atype, ext_c, ext_t, mapping, nlist, cells, box, need_atom_energy,
                                         atype_in=atype_in)
        if self.tables is not None:
            return self._compute_tab(atype, ext_c, ext_t, mapping, nlist, cells, box, need_atom_energy,
                                     atype_in=atype_in)
        em = self.env_mat(nlist, ext_c, ext_t)
        # atoms in blocks: an atom's energy and dE/d em depend on its own em
        # rows only, so the network's working set stays bounded (about
        # ROWS_PER_BLOCK neighbor rows of every layer width)
        nloc = em.shape[0]
        blk = max(256, self.ROWS_PER_BLOCK // max(self.nnei, 1))
        net_deriv = torch.empty_like(em)
        ae = torch.empty(nloc, dtype=torch.float64, device=em.device)
        # one block: the cached type order of the atom-type tensor serves it
        order = self._host_type_order(atype_in, atype) if nloc <= blk else None
        if order is not None and order[1][-1] != nloc:  # types outside [0, ntypes)
            order = None
        for a in range(0, nloc, blk):
            b = min(nloc, a + blk)
            em_b = em[a:b].detach().requires_grad_(True)
            with torch.enable_grad():
                ae_b = self.atom_energy(self.descriptor(em_b), atype[a:b], order)
                (net_deriv[a:b],) = torch.autograd.grad(ae_b.sum(), em_b)
            ae[a:b] = ae_b.detach()
        energy = fitting.energy_sum(ae)
        force, virial = self.forces(net_deriv, nlist, cells, mapping, ext_c, ext_t, box)
        out = {"energy": energy, "force": force, "virial": virial}
        if need_atom_energy:
            out["atom_energy"] = ae
        return out

    def compute_lower(self, ext_coord: torch.Tensor, ext_type: torch.Tensor, cand: torch.Tensor, nloc: int,
                      numneigh=None, mapping=None):
        """E+F+V from extended coordinates and an external neighbor list, as
        forward_lower (deepmd/pt/model/model/make_model.py:334-415) and the
        LAMMPS path (source/api_cc/src/DeepPotPT.cc:180-330) take them: the
        list is formatted on the GPU (format_nlist_cpu semantics).  Returns
        energy, atom_energy [nloc], force [nall, 3] on every extended atom
        (ghost forces left for the caller to fold or communicate; folded onto
        local atoms when mapping is given) and virial [9]."""
        self.check_types(ext_type)
        ext_c = ext_coord.to(torch.float64).reshape(-1, 3).contiguous()
        ext_t = ext_type.to(torch.int32).reshape(-1).contiguous()
        atype = ext_t[:nloc]
        nlist, _ovf, _ = ops.format_nlist(cand, ext_c, ext_t, nloc, self.rcut, self.sel, numneigh=numneigh)
        em = self.env_mat(nlist, ext_c, ext_t)
        if self.fused is not None:
            nd, ae = self._fused_chain(em, atype, pad=self._pad_args(nlist, atype))
        elif self.fused64 is not None:
            pad = self._pad_args(nlist, atype)
            ae, dpart = self._fit_parts64(self.fused64.forward_part(em, pad), atype)
            nd = self.fused64.backward_part(em, dpart, pad[0] if pad is not None else None)
        else:
            em = em.detach().requires_grad_(True)
            with torch.enable_grad():
                ae = self.atom_energy(self.descriptor(em), atype)
                (nd,) = torch.autograd.grad(ae.sum(), em)
        force, virial = ops.force_virial_recompute(nd, nlist, mapping, ext_c, ext_t, self.dstd, self.rcut,
                                                   self.rcut_smth, self.sel)
        ae = ae.detach()
        return {"energy": fitting.energy_sum(ae), "atom_energy": ae, "force": force, "virial": virial}

    def _fit_parts64(self, parts, atype, atype_in=None):
        """(atom energies [nloc], dE/d(sum of parts) [nloc, 4, M]) of the
        per-type G^T R partials parts [nparts, nloc, 4, M] (GR = their sum /
        nnei): the FP64 fused descriptor / fitting kernel, or the torch nets
        of _fit_gr64 where its shapes are not covered."""
        if self.fused_fit64 is not None and parts.dtype == torch.float64:
            order = self._type_order(atype_in if atype_in is not None else atype, atype)
            return self.fused_fit64.from_parts(parts, atype, self.nnei, order=order)
        ae, dgr = self._fit_gr64(parts.sum(dim=0) / self.nnei, atype, atype_in)
        return ae, dgr / self.nnei

    def _compute_fused64(self, atype, ext_c, ext_t, mapping, nlist, cells, box, need_atom_energy, atype_in=None):
        """FP64 plain form on hand-written kernels end to end: fused FP64
        embedding + G^T R partials, the fused FP64 descriptor / fitting /
        adjoint kernel (dE/d partials), the embedding's recompute adjoint,
        the env-mat adjoint for forces and virial, and the compensated
        energy sum."""
        em = self.env_mat(nlist, ext_c, ext_t)
        pad = self._pad_args(nlist, atype)
        part = self.fused64.forward_part(em, pad)
        ae, dpart = self._fit_parts64(part, atype, atype_in)
        nd = self.fused64.backward_part(em, dpart, pad[0] if pad is not None else None)
        force, virial = self.forces(nd, nlist, cells, mapping, ext_c, ext_t, box)
        out = {"energy": fitting.energy_sum(ae), "force": force, "virial": virial}
        if need_atom_energy:
            out["atom_energy"] = ae
        return out

    def _compute_tab(self, atype, ext_c, ext_t, mapping, nlist, cells, box, need_atom_energy, atype_in=None):
        """Compressed model (tabulated embedding) with direct launches of the
        table kernels: tabulate_fusion_se_a per neighbor type (se_a.py:833-839)
        into the per-type G^T R partials, the descriptor and fitting nets with
        their adjoint (_fit_parts64: the FP64 fused kernel, or torch for an
        fp32 table), then tabulate_fusion_se_a_grad per type on
        dy = dE/d(em^T G) = dE/d(sum of partials), whose em_x part adds to
        column 0 of dE/d em (em_x is em[..., 0]); env-mat adjoint for forces
        and virial."""
        em = self.env_mat(nlist, ext_c, ext_t)
        nloc = em.shape[0]
        parts = []
        gr_part = torch.empty((self.ntypes, nloc, 4, self.M), dtype=em.dtype, device=em.device)
        for t in range(self.ntypes):
            rr = em[:, self.sec[t]:self.sec[t + 1], :].contiguous()
            ex = rr[:, :, 0].reshape(-1, 1).contiguous()
            table, info = self.tables[t]
            gr_part[t] = ops.tabulate_forward(table, info, ex, rr, self.M)
            parts.append((table, info, ex, rr))
        ae, dy = self._fit_parts64(gr_part, atype, atype_in)
        dy = dy.to(em.dtype).contiguous()
        nd = torch.empty_like(em)
        for t, (table, info, ex, rr) in enumerate(parts):
            dx, dem = ops.tabulate_grad(table, info, ex, rr, dy)
            dem[:, :, 0] += dx.reshape(dem.shape[0], -1)
            nd[:, self.sec[t]:self.sec[t + 1], :] = dem
        force, virial = self.forces(nd, nlist, cells, mapping, ext_c, ext_t, box)
        out = {"energy": fitting.energy_sum(ae), "force": force, "virial": virial}
        if need_atom_energy:
            out["atom_energy"] = ae
        return out

    def _host_type_order(self, atype_in, atype):
        """(perm int64 [nloc] atoms ordered by type, bounds: host list of each
        type's first sorted index + nloc), reused while the caller passes the
        same, unmodified atype tensor (identity and version counter)."""
        c = self._host_order_cache
        if atype_in is not None and c is not None and c[0] is atype_in and c[1] == atype_in._version:
            return c[2]
        a = atype.to(torch.int64)
        perm = torch.argsort(a, stable=True)
        counts = torch.bincount(a, minlength=self.ntypes)[: self.ntypes].cpu().tolist()
        bounds = [0]
        for n in counts:
            bounds.append(bounds[-1] + int(n))
        order = (perm, bounds)
        if atype_in is not None:
            self._host_order_cache = (atype_in, atype_in._version, order)
        return order

    def _fit_gr64(self, gr, atype, atype_in=None):
        """Descriptor D = GR^T GR< and the per-type fitting nets (in the dtype
        of gr: FP64, or FP32 for an fp32 compressed model) with
        their autograd adjoint, on G^T R [nloc, 4, M] gathered once into
        type-sorted order: each type's net runs on a contiguous slice (blocks
        of up to 2^17 atoms), and the atomic energies and dE/dGR go back to
        the original order once.  Returns (atom_energy [nloc], dgr)."""
        perm, bounds = self._host_type_order(atype_in, atype)
        gs = gr.index_select(0, perm)
        dgs = torch.empty_like(gs)
        aes = torch.empty(gs.shape[0], dtype=torch.float64, device=gr.device)
        blk = max(1024, self.ROWS_PER_BLOCK // 32)
        for t in range(self.ntypes):
            for a in range(bounds[t], bounds[t + 1], blk):
                b = min(bounds[t + 1], a + blk)
                g = gs[a:b].detach().requires_grad_(True)
                with torch.enable_grad():
                    d = torch.bmm(g.transpose(1, 2), g[:, :, : self.axis]).reshape(b - a, -1)
                    y = self.fit[t](d, self.gemm_dtype)[:, 0].to(torch.float64)
                    (dgs[a:b],) = torch.autograd.grad(y.sum(), g)
                aes[a:b] = y.detach() + self.bias_atom_e[t] + self.out_bias[t]
        ae = torch.empty_like(aes).index_copy_(0, perm, aes)
        dgr = torch.empty_like(gr).index_copy_(0, perm, dgs)
        return ae, dgr

    def _padding_rows(self):
        """pad_em [ntypes (center), ntypes (neighbor), 4] on the device, in
        the env-mat's dtype: the em row of every padding slot of a section,
        (0 - davg) / dstd as the env-mat computes it (env_mat_a_cpu leaves
        an empty slot's row at 0, source/lib/src/env_mat.cc:149-160, and the
        normalization of prod_env_mat.cc:85-88 applies to every slot); None
        when davg or dstd varies within a section."""
        npdt = np.float32 if self.dtype == torch.float32 else np.float64
        av = self.davg.detach().cpu().numpy().astype(npdt).reshape(self.ntypes, self.nnei, 4)
        sd = self.dstd.detach().cpu().numpy().astype(npdt).reshape(self.ntypes, self.nnei, 4)
        pad = np.zeros((self.ntypes, self.ntypes, 4), dtype=npdt)
        for ti in range(self.ntypes):
            for t in range(self.ntypes):
                a, d = av[ti, self.sec[t]:self.sec[t + 1]], sd[ti, self.sec[t]:self.sec[t + 1]]
                if len(a) == 0:
                    continue
                if not (np.all(a == a[0]) and np.all(d == d[0])):
                    return None
                pad[ti, t] = (npdt(0.0) - a[0]) / d[0]
        return torch.from_numpy(pad).to(self.device)

    def _pad_args(self, nlist, atype):
        """(counts, atype, pad_em) of the padding shortcut for this list, or None."""
        if self.pad_em is None:
            return None
        counts = torch.empty((nlist.shape[0], self.ntypes), dtype=torch.int16, device=nlist.device)
        ops._launch("dpmi_se_a_section_counts", _lib.ptr(counts), _lib.ptr(nlist.contiguous()), nlist.shape[0],
                    _lib.ptr(self._sel_i32), self.ntypes, ops._stream())
        return counts, atype.to(torch.int32).contiguous(), self.pad_em

    def _type_order(self, atype_in, atype):
        """(perm, type_start) of the fused fitting kernel, reused while the
        caller passes the same, unmodified atype tensor (the cache holds the
        tensor itself and its version counter, so neither a new tensor nor an
        in-place change can reuse a stale order)."""
        ff = self.fused_fit if self.fused_fit is not None else self.fused_fit64
        if ff is None:
            return None
        c = self._order_cache
        if c is not None and c[0] is atype_in and c[1] == atype_in._version:
            return c[2]
        order = ff.sort_by_type(atype)
        self._order_cache = (atype_in, atype_in._version, order)
        return order

    def _fused_chain(self, em, atype, use_gr=None, order=None, pad=None):
        """bf16 / fp16 embedding -> descriptor -> fitting -> adjoints: (dE/d em, atom
        energies).  With the descriptor fused into the fitting kernel
        (fused_gr) the G^T R partials go straight to it and dE/dGR comes back;
        otherwise the descriptor kernels run between (and torch runs the
        fitting net when its shape has no fused kernel)."""
        use_gr = self.fused_gr if use_gr is None else use_gr
        counts = pad[0] if pad is not None else None
        if use_gr:
            gr_part = self.fused.forward_part(em, pad)
            ae, dgr = self.fused_fit.from_gr(gr_part, atype, self.fused.n1, self.fused.nnei, order=order)
            return self.fused.backward_dgr(em, dgr, counts), ae
        gr, desc = self.fused.forward(em, self.gemm_dtype if self.fused_fit is not None else torch.float32, pad)
        if self.fused_fit is not None:
            ae, ddesc = self.fused_fit(desc, atype, order=order)
        else:
            desc.requires_grad_(True)
            with torch.enable_grad():
                ae = self.atom_energy(desc, atype)
                (ddesc,) = torch.autograd.grad(ae.sum(), desc)
        return self.fused.backward(em, gr, ddesc, counts), ae

    def _compute_fused(self, atype, ext_c, ext_t, mapping, nlist, cells, box, need_atom_energy, order=None):
        """bf16 / fp16 forms with hand-written adjoints end to end: fused embedding +
        G^T R forward, fused fitting forward + adjoint (or hipBLASLt GEMMs with
        autograd through the fitting net when its shape is not supported),
        descriptor and embedding adjoints recomputing the forward, env-mat
        adjoint for forces and virial."""
        em = self.env_mat(nlist, ext_c, ext_t)
        nd, ae = self._fused_chain(em, atype, order=order, pad=self._pad_args(nlist, atype))
        energy = fitting.energy_sum(ae.detach())
        force, virial = self.forces(nd, nlist, cells, mapping, ext_c, ext_t, box)
        out = {"energy": energy, "force": force, "virial": virial}
        if need_atom_energy:
            out["atom_energy"] = ae.detach()
        return out


class CapturedStep:
    """One E+F+V step of a model captured in a HIP graph (torch.cuda.CUDAGraph):
    the neighbor list (periodic box: the form without host waits), env-mat,
    networks, forces and virial replay as one graph launch.  coord and atype
    are the captured inputs: copy new coordinates into ``self.coord`` (or pass
    them to ``replay``) and replay; the box and the atom count are fixed.
    ``self.out`` holds the outputs of the last replay.  The model's status
    check does not see replays: a failed list shows as NaN forces."""

    def __init__(self, model: SeAEnergyModel, coord: torch.Tensor, atype: torch.Tensor, box, warmup: int = 2):
        self.model = model
        self.coord = coord.detach().clone()
        self.atype = atype.detach().clone()
        self.box = box
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(max(1, warmup)):
                model.compute(self.coord, self.atype, box)
            model.check_status()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = model.compute(self.coord, self.atype, box)

    def replay(self, coord: torch.Tensor | None = None) -> dict:
        if coord is not None:
            self.coord.copy_(coord)
        self.graph.replay()
        return self.out
