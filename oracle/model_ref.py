"""CPU model oracle for se_e2_a energy models (test infrastructure only).

A float64 torch-CPU restatement of the reference deepmd/pt model path:

* env-mat: deepmd/pt/model/descriptor/env_mat.py:11-92 and
  deepmd/pt/utils/preprocess.py:9-19 (smooth weight);
* embedding/fitting layers: deepmd/pt/model/network/mlp.py:188-219
  (linear, tanh, timestep, resnet with x or [x, x]);
* descriptor: deepmd/pt/model/descriptor/se_a.py:750-866 (type_one_side,
  G^T R / nnei, D = (G^T R)^T (G^T R)[:, :axis]); compressed: the C
  oracle's tabulate_fusion_se_a in place of the embedding + G^T R per type;
* fitting: deepmd/pt/model/task/fitting.py:734-900 (own-type net + bias_atom_e)
  and the out_bias of base_atomic_model.py:571-592;
* energy/force/virial: deepmd/pt/model/model/transform_output.py:65-206,
  317-383 (autograd force, F (x) x virial, ghost -> local fold);
* training loss: deepmd/pt/loss/ener.py:243-271,302-328 (E + F mse) and its
  parameter gradient by torch double backward; compressed: the C
  tabulate_fusion_se_a_grad_grad as the backward of the tabulate grad.

The neighbor list comes from the C oracle (reference C++ semantics).  Pinned
by source/tests/infer/deeppot-testcase.yaml (reference tolerance: 12 decimal
places, source/tests/infer/test_models.py:30-32) and by finite differences.
"""

from __future__ import annotations

import numpy as np
import torch

from . import oracle


def _arr(x):
    return None if x is None else torch.as_tensor(np.asarray(x, dtype=np.float64))


class Layer:
    def __init__(self, d: dict, activation: bool = True):
        d = d.get("@variables", d)  # serialized layers keep arrays under @variables
        self.w = _arr(d["w"])
        self.b = _arr(d["b"])
        self.idt = _arr(d.get("idt"))
        self.act = activation

    def __call__(self, x):
        y = x @ self.w + self.b
        if self.act:
            y = torch.tanh(y)
        if self.idt is not None:
            y = y * self.idt
        nin, nout = self.w.shape
        if self.act:  # hidden layers carry resnet=True
            if nin == nout:
                y = y + x
            elif 2 * nin == nout:
                y = y + torch.cat([x, x], dim=-1)
        return y


class _Tabulate(torch.autograd.Function):
    """Compressed embedding of one neighbor-type section on the CPU: the C
    oracle's tabulate_fusion_se_a and its grad (the op and its backward,
    tabulate_multi_device.cc:702-830; called as in se_a.py:826-839)."""

    @staticmethod
    def forward(ctx, em_x, em, table, info, M):
        ctx.save_for_backward(em_x, em)
        ctx.table, ctx.info = table, info
        out = oracle.tabulate_fusion_se_a(table, info, em_x.detach().numpy(), em.detach().numpy(), M)
        return torch.as_tensor(out)

    @staticmethod
    def backward(ctx, dy):
        em_x, em = ctx.saved_tensors
        dx, dem = _TabulateGrad.apply(em_x, em, dy, ctx.table, ctx.info)
        return dx, dem, None, None, None


class _TabulateGrad(torch.autograd.Function):
    """(dy_dem_x, dy_dem) of the compressed embedding; its own backward is
    tabulate_fusion_se_a_grad_grad w.r.t. dy only, as the reference's
    TabulateFusionSeAGradOp::backward (tabulate_multi_device.cc:669-700)."""

    @staticmethod
    def forward(ctx, em_x, em, dy, table, info):
        ctx.save_for_backward(em_x, em)
        ctx.table, ctx.info, ctx.M = table, info, dy.shape[-1]
        dx, dem = oracle.tabulate_fusion_se_a_grad(table, info, em_x.detach().numpy(), em.detach().numpy(),
                                                   dy.detach().numpy())
        return torch.as_tensor(dx).reshape(em_x.shape), torch.as_tensor(dem)

    @staticmethod
    def backward(ctx, g_dx, g_dem):
        em_x, em = ctx.saved_tensors
        g_dx = torch.zeros_like(em_x) if g_dx is None else g_dx
        g_dem = torch.zeros_like(em) if g_dem is None else g_dem
        dz = oracle.tabulate_fusion_se_a_grad_grad(ctx.table, ctx.info, em_x.detach().numpy(),
                                                   em.detach().numpy(),
                                                   g_dx.detach().numpy().reshape(em.shape[0], em.shape[1]),
                                                   g_dem.detach().numpy(), ctx.M)
        return None, None, torch.as_tensor(dz), None, None


def smooth_weight(distance, rmin, rmax):
    """compute_smooth_weight (deepmd/pt/utils/preprocess.py:9-19)."""
    if rmin >= rmax:
        raise ValueError("rmin should be less than rmax.")
    distance = torch.clamp(distance, min=rmin, max=rmax)
    uu = (distance - rmin) / (rmax - rmin)
    uu2 = uu * uu
    return uu2 * uu * (-6 * uu2 + 15 * uu - 10) + 1


def exp_switch(distance, rmin, rmax):
    """compute_exp_sw (deepmd/pt/utils/preprocess.py:22-34), C = 20."""
    if rmin >= rmax:
        raise ValueError("rmin should be less than rmax.")
    distance = torch.clamp(distance, min=0.0, max=rmax)
    a = 20 / rmin
    return torch.exp(-torch.exp(a * (distance - rmin)))


def prod_env_mat_pt(extended_coord, nlist, atype, mean, stddev, rcut, rcut_smth, radial_only=False,
                    protection=0.0, use_exp_switch=False):
    """prod_env_mat / _make_env_mat of the PyTorch backend
    (deepmd/pt/model/descriptor/env_mat.py:11-92), step for step, on CPU
    tensors (differentiable w.r.t. extended_coord through torch autograd):
    padding indices point at an extra coordinate coord[-1] + rcut, padded
    rows get length + 1 and weight 0, and the result is normalised by the
    centre type's mean / stddev.  Returns (env_mat, diff, switch)."""
    bsz, natoms, nnei = nlist.shape
    coord = extended_coord.view(bsz, -1, 3)
    nall = coord.shape[1]
    mask = nlist >= 0
    nl = torch.where(mask, nlist, nall)
    coord_l = coord[:, :natoms].view(bsz, -1, 1, 3)
    index = nl.view(bsz, -1).unsqueeze(-1).expand(-1, -1, 3)
    coord_pad = torch.concat([coord, coord[:, -1:, :] + rcut], dim=1)
    coord_r = torch.gather(coord_pad, 1, index).view(bsz, natoms, nnei, 3)
    diff = coord_r - coord_l
    length = torch.linalg.norm(diff, dim=-1, keepdim=True)
    length = length + ~mask.unsqueeze(-1)
    t0 = 1 / (length + protection)
    t1 = diff / (length + protection) ** 2
    sw = exp_switch(length, rcut_smth, rcut) if use_exp_switch else smooth_weight(length, rcut_smth, rcut)
    weight = sw * mask.unsqueeze(-1)
    env = t0 * weight if radial_only else torch.cat([t0, t1], dim=-1) * weight
    env = (env - mean[atype]) / stddev[atype]
    return env, diff * mask.unsqueeze(-1), weight


def loss_prefactors(start_pref_e, start_pref_f, limit_pref_e=None, limit_pref_f=None, learning_rate=None,
                    starter_learning_rate=None):
    """(pref_e, pref_f) of EnergyStdLoss.forward (deepmd/pt/loss/ener.py:
    229-231): coef = learning_rate / starter_learning_rate and
    pref = limit + (start - limit) * coef.  Without a learning rate, or
    without limits, the start prefactors."""
    if learning_rate is None:
        return float(start_pref_e), float(start_pref_f)
    coef = float(learning_rate) / float(starter_learning_rate)
    le = start_pref_e if limit_pref_e is None else limit_pref_e
    lf = start_pref_f if limit_pref_f is None else limit_pref_f
    return le + (start_pref_e - le) * coef, lf + (start_pref_f - lf) * coef


class Model:
    """se_e2_a + energy fitting, from a dict in the reference serialization
    layout (the 'model' block of deeppot_sea.yaml) or from dpmi.systems
    parameters (see from_params).  tables (optional): per neighbor type
    (table, table_info) of the compressed embedding."""

    def __init__(self, rcut, rcut_smth, sel, axis, davg, dstd, emb, fit, bias_atom_e, out_bias=None, tables=None):
        self.tables = tables
        self.rcut = float(rcut)
        self.rcut_smth = float(rcut_smth)
        self.sel = list(sel)
        self.nnei = int(sum(sel))
        self.sec = np.concatenate([[0], np.cumsum(sel)])
        self.axis = int(axis)
        self.davg = _arr(davg)
        self.dstd = _arr(dstd)
        self.emb = [[Layer(l) for l in net] for net in emb]
        self.fit = [[Layer(l, activation=(k < len(net) - 1)) for k, l in enumerate(net)] for net in fit]
        self.bias_atom_e = _arr(bias_atom_e).reshape(-1)
        nt = len(sel)
        self.out_bias = _arr(out_bias).reshape(nt) if out_bias is not None else torch.zeros(nt, dtype=torch.float64)

    @classmethod
    def from_serialized(cls, m: dict) -> "Model":
        d = m["descriptor"]
        f = m["fitting"]
        emb = [net["layers"] for net in d["embeddings"]["networks"]]
        fit = [net["layers"] for net in f["nets"]["networks"]]
        return cls(d["rcut"], d["rcut_smth"], d["sel"], d["axis_neuron"], d["@variables"]["davg"],
                   d["@variables"]["dstd"], emb, fit, f["@variables"]["bias_atom_e"],
                   m["@variables"]["out_bias"])

    @classmethod
    def from_params(cls, spec, params: dict, tables=None) -> "Model":
        return cls(spec.rcut, spec.rcut_smth, spec.sel, spec.axis_neuron, params["davg"], params["dstd"],
                   params["embedding"], params["fitting"], params["bias_atom_e"], tables=tables)

    # ------------------------------------------------------------------ pieces
    def env_mat(self, ext, nlist, atype):
        nloc, nnei = nlist.shape
        nl = torch.as_tensor(nlist, dtype=torch.long)
        mask = nl >= 0
        idx = torch.where(mask, nl, torch.zeros_like(nl))
        diff = ext[idx] - ext[:nloc].unsqueeze(1)
        length = torch.linalg.norm(diff, dim=-1, keepdim=True)
        length = length + (~mask).unsqueeze(-1)
        t0 = 1.0 / length
        t1 = diff / length**2
        u = (torch.clamp(length, self.rcut_smth, self.rcut) - self.rcut_smth) / (self.rcut - self.rcut_smth)
        sw = u * u * u * (-6 * u * u + 15 * u - 10) + 1
        sw = sw * mask.unsqueeze(-1)
        em = torch.cat([t0, t1], dim=-1) * sw
        at = torch.as_tensor(atype[:nloc], dtype=torch.long)
        return (em - self.davg[at]) / self.dstd[at]

    def descriptor(self, em):
        nloc = em.shape[0]
        M = self.emb[0][-1].w.shape[1]
        gr = torch.zeros((nloc, 4, M), dtype=torch.float64)
        for t in range(len(self.sel)):
            rr = em[:, self.sec[t]:self.sec[t + 1], :]
            if self.tables is not None:
                table, info = self.tables[t]
                gr = gr + _Tabulate.apply(rr[:, :, :1].reshape(-1, 1), rr.contiguous(), table, info, M)
                continue
            g = rr[:, :, :1]
            for layer in self.emb[t]:
                g = layer(g)
            gr = gr + rr.transpose(1, 2) @ g
        gr = gr / self.nnei
        d = gr.transpose(1, 2) @ gr[:, :, : self.axis]
        return d.reshape(nloc, -1)

    def atom_energy(self, desc, atype):
        at = torch.as_tensor(atype, dtype=torch.long)
        e = torch.zeros(desc.shape[0], dtype=torch.float64)
        for t, net in enumerate(self.fit):
            y = desc
            for layer in net:
                y = layer(y)
            y = y[:, 0] + self.bias_atom_e[t]
            e = torch.where(at == t, y, e)
        return e + self.out_bias[at]

    # ------------------------------------------------------------------ training
    def named_parameters(self):
        """[((kind, net, layer, name), tensor)] for w, b, idt of every
        embedding layer (none with tables: the compressed embedding is fixed)
        and every fitting layer, made leaves that require grad."""
        groups = [("fitting", self.fit)]
        if self.tables is None:
            groups.insert(0, ("embedding", self.emb))
        out = []
        for kind, nets in groups:
            for n, net in enumerate(nets):
                for k, l in enumerate(net):
                    for name in ("w", "b", "idt"):
                        t = getattr(l, name)
                        if t is not None:
                            t = t.detach().clone().requires_grad_(True)
                            setattr(l, name, t)
                            out.append(((kind, n, k, name), t))
        return out

    def parameters(self):
        return [t for _, t in self.named_parameters()]

    def energy_force(self, coord, atype, box, create_graph=False):
        """E (fp64 scalar tensor) and F [nloc, 3] = -dE/dx with ghost forces
        folded to their owners (transform_output.py:152-206,317-383)."""
        coord = np.asarray(coord, dtype=np.float64).reshape(-1, 3)
        atype = np.asarray(atype, dtype=np.int32)
        if box is None:
            ec, et, mp, nlist, _ = oracle.build_nlist_nopbc(coord, atype, self.rcut, self.sel)
        else:
            ec, et, mp, nlist, _ = oracle.build_nlist_pbc(coord, atype, box, self.rcut, self.sel)
        ext = torch.tensor(ec, requires_grad=True)
        energy = self.atom_energy(self.descriptor(self.env_mat(ext, nlist, et)), atype).sum()
        (g,) = torch.autograd.grad(energy, ext, create_graph=create_graph)
        f = torch.zeros((len(atype), 3), dtype=torch.float64).index_add(0, torch.as_tensor(mp, dtype=torch.long), -g)
        return energy, f

    def train_loss(self, frames, energy_label, force_label, pref_e=0.02, pref_f=1000.0, params=None,
                   limit_pref_e=None, limit_pref_f=None, learning_rate=None, starter_learning_rate=None):
        """EnergyStdLoss (deepmd/pt/loss/ener.py:243-271,302-328; mse, 1/N
        energy normalisation) over frames [(coord, atype, box)], and with
        params (from parameters()) its gradient by a double backward.
        pref_e / pref_f are the start prefactors; with a learning rate they
        move to limit_pref_* as ener.py:229-231 moves them."""
        pref_e, pref_f = loss_prefactors(pref_e, pref_f, limit_pref_e, limit_pref_f, learning_rate,
                                         starter_learning_rate)
        es, fs = [], []
        for c, t, b in frames:
            e, f = self.energy_force(c, t, b, create_graph=params is not None)
            es.append(e)
            fs.append(f)
        energy = torch.stack(es)
        force = torch.stack(fs)
        natoms = force.shape[1]
        l2_e = torch.mean(torch.square(energy - torch.as_tensor(np.asarray(energy_label, dtype=np.float64))))
        l2_f = torch.mean(torch.square((torch.as_tensor(np.asarray(force_label, dtype=np.float64)) - force).reshape(-1)))
        loss = (1.0 / natoms) * (pref_e * l2_e) + pref_f * l2_f
        if params is None:
            return float(loss.item()), None
        grads = torch.autograd.grad(loss, params)
        return float(loss.item()), [g.detach().numpy() for g in grads]

    # ------------------------------------------------------------------ E, F, V
    def evaluate(self, coord, atype, box):
        coord = np.asarray(coord, dtype=np.float64).reshape(-1, 3)
        atype = np.asarray(atype, dtype=np.int32)
        nloc = len(atype)
        if box is None:
            ec, et, mp, nlist, _ = oracle.build_nlist_nopbc(coord, atype, self.rcut, self.sel)
        else:
            ec, et, mp, nlist, _ = oracle.build_nlist_pbc(coord, atype, box, self.rcut, self.sel)
        ext = torch.tensor(ec, requires_grad=True)
        em = self.env_mat(ext, nlist, et)
        desc = self.descriptor(em)
        ae = self.atom_energy(desc, atype)
        energy = ae.sum()
        (g,) = torch.autograd.grad(energy, ext)
        fext = -g.detach().numpy()
        virial = np.einsum("ik,ij->kj", fext, ec).reshape(9)
        force = np.zeros((nloc, 3))
        np.add.at(force, mp, fext)
        return {
            "energy": float(energy.item()),
            "atom_energy": ae.detach().numpy(),
            "force": force,
            "virial": virial,
            "descriptor": desc.detach().numpy(),
            "nlist": nlist,
            "ext_coord": ec,
            "mapping": mp,
        }
