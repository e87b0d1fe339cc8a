"""Training step of the se_e2_a energy model on the GPU (E + F loss).

Follows the reference step (deepmd/pt/train/training.py:1393-1460): model
forward with the force graph kept (create_graph, make_model.py:410), the
energy/force loss of EnergyStdLoss (deepmd/pt/loss/ener.py:197-330, "mse",
legacy 1/N energy normalisation), loss.backward() — a double backward
through the force — and an Adam step.

Data path per batch of frames:

  coordinates --nlist kernels--> formatted list + list mapped to local atoms
              --env-mat kernel--> em, d em / d r_ij (stored: frames are small)
  em          --embedding (or tabulate_fusion_se_a) + fitting--> E per frame
  dE/d em     (autograd, graph kept) --ProdForceOp--> F per local atom
  loss.backward(): F's backward is the ProdForceSeAGrad kernel; on the
  tabulated path the tabulate grad's backward is tabulate_fusion_se_a_grad_grad.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from . import ops


@dataclass
class Batch:
    """Frames of one configuration size, prepared for the model (all on the
    GPU): em [nf * nloc, nnei, 4], em_deriv [nf, nloc * nnei * 12], nlist
    mapped to local atoms [nf, nloc * nnei], atype [nf * nloc]."""

    nframes: int
    nloc: int
    em: torch.Tensor
    em_deriv: torch.Tensor
    nlist: torch.Tensor
    atype: torch.Tensor


def prepare(model, coords, atypes, boxes) -> Batch:
    """Neighbor list (bit-exact C++ semantics), list mapped to local atoms
    and env-mat with its derivative, for each frame."""
    ems, eds, nls, ats = [], [], [], []
    nloc = None
    for c, t, b in zip(coords, atypes, boxes):
        c = torch.as_tensor(c, dtype=torch.float64, device=model.device).reshape(-1, 3)
        t = torch.as_tensor(t, device=model.device).to(torch.int32)
        if nloc is None:
            nloc = c.shape[0]
        elif c.shape[0] != nloc:
            raise ValueError("all frames of a batch need the same number of atoms")
        ext_c, ext_t, _map, nlist, _ovf, nl_map = ops.build_nlist(c, t, b, model.rcut, model.sel, mapped=True)
        em, ed, _ = ops.env_mat_a(nlist, ext_c, ext_t, model.davg, model.dstd, model.rcut, model.rcut_smth,
                                  model.sel, with_deriv=True, with_rij=False, dtype=model.dtype)
        ems.append(em)
        eds.append(ed.reshape(-1))
        nls.append(nl_map.reshape(-1))
        ats.append(t)
    nf = len(ems)
    return Batch(nf, nloc, torch.cat(ems), torch.stack(eds), torch.stack(nls), torch.cat(ats))


def energy_force(model, batch: Batch, create_graph: bool):
    """E [nf] (fp64) and F [nf, nloc, 3] of every frame; with create_graph
    the force stays differentiable w.r.t. the model parameters."""
    em = batch.em.detach().requires_grad_(True)
    with torch.enable_grad():
        desc = model.descriptor(em)
        ae = model.atom_energy(desc, batch.atype)
        energy = ae.reshape(batch.nframes, batch.nloc).sum(dim=1)
        (nd,) = torch.autograd.grad(energy.sum(), em, create_graph=create_graph)
        nd = nd.reshape(batch.nframes, -1)
        force = ops.ProdForceOp.apply(nd, batch.em_deriv, batch.nlist, batch.nloc, batch.nloc, batch.nframes)
    return energy, force.reshape(batch.nframes, batch.nloc, 3)


class EnergyStdLoss:
    """loss = pref_e / N * mean_f (E - E*)^2 + pref_f * mean (F - F*)^2
    (EnergyStdLoss.forward, deepmd/pt/loss/ener.py:243-271,302-328 with
    loss_func "mse", no Huber, intensive_ener_virial False).

    The prefactors follow the learning rate as in the reference
    (ener.py:229-231): pref = limit + (start - limit) * lr / starter_lr.
    limit_pref_* default to start_pref_*, which keeps them constant."""

    def __init__(self, start_pref_e: float = 0.02, start_pref_f: float = 1000.0, limit_pref_e: float | None = None,
                 limit_pref_f: float | None = None, starter_learning_rate: float = 1.0):
        self.start_pref_e = float(start_pref_e)
        self.start_pref_f = float(start_pref_f)
        self.limit_pref_e = self.start_pref_e if limit_pref_e is None else float(limit_pref_e)
        self.limit_pref_f = self.start_pref_f if limit_pref_f is None else float(limit_pref_f)
        self.starter_learning_rate = float(starter_learning_rate)

    def prefactors(self, learning_rate: float | None = None):
        """(pref_e, pref_f) at learning_rate (the starter rate when None)."""
        lr = self.starter_learning_rate if learning_rate is None else float(learning_rate)
        coef = lr / self.starter_learning_rate
        pref_e = self.limit_pref_e + (self.start_pref_e - self.limit_pref_e) * coef
        pref_f = self.limit_pref_f + (self.start_pref_f - self.limit_pref_f) * coef
        return pref_e, pref_f

    def __call__(self, energy, force, energy_label, force_label, natoms: int, learning_rate: float | None = None):
        pref_e, pref_f = self.prefactors(learning_rate)
        l2_e = torch.mean(torch.square(energy - energy_label))
        diff_f = (force_label - force).reshape(-1)
        l2_f = torch.mean(torch.square(diff_f))
        loss = (1.0 / natoms) * (pref_e * l2_e) + (pref_f * l2_f).to(torch.float64)
        return loss, {"rmse_e": l2_e.detach().sqrt() / natoms, "rmse_f": l2_f.detach().sqrt()}


class Trainer:
    """Adam on the model's trainable tensors (embedding and fitting nets;
    the fitting nets only with a tabulated embedding).  The loss prefactors
    move with the optimizer's current learning rate (lr of its first param
    group) from start_pref_* at lr to limit_pref_*, as the reference's."""

    def __init__(self, model, lr: float = 1e-3, pref_e: float = 0.02, pref_f: float = 1000.0,
                 limit_pref_e: float | None = None, limit_pref_f: float | None = None):
        self.model = model
        self.params = model.parameters()
        self.loss = EnergyStdLoss(pref_e, pref_f, limit_pref_e, limit_pref_f, starter_learning_rate=lr)
        self.opt = torch.optim.Adam(self.params, lr=lr)

    def loss_and_grad(self, batch: Batch, energy_label, force_label):
        for p in self.params:
            p.grad = None
        energy, force = energy_force(self.model, batch, create_graph=True)
        lr = self.opt.param_groups[0]["lr"]
        loss, more = self.loss(energy, force, energy_label, force_label, batch.nloc, learning_rate=lr)
        loss.backward()
        return loss.detach(), more

    def step(self, batch: Batch, energy_label, force_label):
        loss, more = self.loss_and_grad(batch, energy_label, force_label)
        self.opt.step()
        return loss, more
