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

import numpy as np
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
    # lists built without host waits: int32 [nframes] status words (0 = ok)
    status: torch.Tensor | None = None


def prepare(model, coords, atypes, boxes, wait: bool = True) -> Batch:
    """Neighbor list (bit-exact C++ semantics), list mapped to local atoms
    and env-mat with its derivative, for each frame.  wait=False builds the
    periodic lists without host waits (capacity-bounded ghosts; a failed list
    is reported by ops.nlist_status and poisons its rows instead)."""
    ems, eds, nls, ats, sts = [], [], [], [], []
    nloc = None
    for c, t, b in zip(coords, atypes, boxes):
        c = torch.as_tensor(c, dtype=torch.float64, device=model.device).reshape(-1, 3)
        t = torch.as_tensor(t, device=model.device).to(torch.int32)
        if nloc is None:
            nloc = c.shape[0]
        elif c.shape[0] != nloc:
            raise ValueError("all frames of a batch need the same number of atoms")
        nowait = not wait and b is not None
        ext_c, ext_t, _map, nlist, _ovf, nl_map, *ws = ops.build_nlist(c, t, b, model.rcut, model.sel, mapped=True,
                                                                       cells=nowait, wait=not nowait)
        if nowait:
            sts.append(ops.nlist_status(ws[0][1]).reshape(1))
        em, ed, _ = ops.env_mat_a(nlist, ext_c, ext_t, model.davg, model.dstd, model.rcut, model.rcut_smth,
                                  model.sel, with_deriv=True, with_rij=False, dtype=model.dtype)
        ems.append(em)
        eds.append(ed.reshape(-1))
        nls.append(nl_map.reshape(-1))
        ats.append(t)
    nf = len(ems)
    status = torch.cat(sts) if sts and len(sts) == nf else None
    return Batch(nf, nloc, torch.cat(ems), torch.stack(eds), torch.stack(nls), torch.cat(ats), status)


def energy_force(model, batch: Batch, create_graph: bool, order=None):
    """E [nf] (fp64) and F [nf, nloc, 3] of every frame; with create_graph
    the force stays differentiable w.r.t. the model parameters.  order =
    (perm, bounds) of the batch's types runs the fitting nets on contiguous
    type slices without the per-type host reads of nonzero()."""
    em = batch.em.detach().requires_grad_(True)
    with torch.enable_grad():
        desc = model.descriptor(em)
        ae = model.atom_energy(desc, batch.atype, order)
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

    def __call__(self, energy, force, energy_label, force_label, natoms: int, learning_rate: float | None = None,
                 prefs=None):
        """prefs = (pref_e, pref_f) as device scalars (the learning-rate tick's
        output) replaces the host-side prefactors of learning_rate."""
        pref_e, pref_f = self.prefactors(learning_rate) if prefs is None else prefs
        l2_e = torch.mean(torch.square(energy - energy_label))
        diff_f = (force_label - force).reshape(-1)
        l2_f = torch.mean(torch.square(diff_f))
        loss = (1.0 / natoms) * (pref_e * l2_e) + (pref_f * l2_f).to(torch.float64)
        return loss, {"rmse_e": l2_e.detach().sqrt() / natoms, "rmse_f": l2_f.detach().sqrt()}


class LearningRateExp:
    """The reference trainer's exponential-decay schedule (BaseLR.__init__
    and LearningRateExp, deepmd/dpmodel/utils/learning_rate.py:30-166,
    219-397): the constructor derives stop_lr, the warmup, decay_steps and
    decay_rate as the reference does; the rate at step k is evaluated on the
    GPU by dpmi_lr_tick from ``sched()``, under the LambdaLR of
    training.py:1138-1150."""

    def __init__(self, start_lr: float, num_steps: int, stop_lr: float | None = None,
                 stop_lr_ratio: float | None = None, decay_steps: int = 5000, decay_rate: float | None = None,
                 warmup_steps: int = 0, warmup_ratio: float | None = None, warmup_start_factor: float = 0.0,
                 smooth: bool = False):
        if not (start_lr > 0 and np.isfinite(start_lr)):
            raise ValueError(f"start_lr ({start_lr}) must be positive and finite.")
        if (stop_lr is None) == (stop_lr_ratio is None):
            raise ValueError("give exactly one of stop_lr and stop_lr_ratio")
        self.stop_lr = start_lr * stop_lr_ratio if stop_lr_ratio is not None else float(stop_lr)
        if warmup_steps != 0 and warmup_ratio is not None:
            raise ValueError("warmup_steps and warmup_ratio are mutually exclusive")
        self.warmup_steps = int(warmup_ratio * num_steps) if warmup_ratio is not None else int(warmup_steps)
        if num_steps < 0 or self.warmup_steps < 0:
            raise ValueError("num_steps and warmup_steps must be non-negative")
        if num_steps > 0 and self.warmup_steps >= num_steps:
            raise ValueError("warmup_steps must be smaller than num_steps")
        if num_steps == 0 and self.warmup_steps != 0:
            raise ValueError("warmup_steps must be 0 when num_steps is 0")
        self.warmup_start_lr = warmup_start_factor * start_lr
        self.start_lr = float(start_lr)
        self.num_steps = int(num_steps)
        decay_total = self.num_steps - self.warmup_steps
        if decay_steps <= 0:
            raise ValueError(f"decay_steps ({decay_steps}) must be positive.")
        self.decay_steps = int(decay_steps)
        if decay_rate is None and self.decay_steps >= decay_total:
            self.decay_steps = 100 if decay_total // 10 > 100 else decay_total // 100 + 1
        if decay_rate is not None:
            self.decay_rate = float(decay_rate)
        elif decay_total == 0:
            self.decay_rate = 1.0
        else:
            clamped = max(self.stop_lr, 1e-10)
            self.decay_rate = np.exp(np.log(clamped / self.start_lr) / (decay_total / self.decay_steps)).item()
        self.smooth = bool(smooth)

    @classmethod
    def constant(cls, lr: float) -> "LearningRateExp":
        return cls(lr, num_steps=0, stop_lr=lr, decay_rate=1.0)

    def sched(self, loss: "EnergyStdLoss") -> np.ndarray:
        """The 11 values dpmi_lr_tick reads (schedule, then the prefactors)."""
        return np.array([self.start_lr, self.decay_rate, self.stop_lr, self.decay_steps, self.warmup_steps,
                         self.warmup_start_lr, float(self.smooth), loss.start_pref_e, loss.limit_pref_e,
                         loss.start_pref_f, loss.limit_pref_f], dtype=np.float64)


class DeviceAdam:
    """Adam of the reference trainer (torch.optim.Adam(fused=True),
    training.py:1028-1030) as one dpmi_adam_* launch over a flat buffer: the
    parameters are re-pointed to views of ``flat``, so the kernel updates
    them in place.  The learning-rate tick (``tick``) writes [lr, pref_e,
    pref_f, t] into ``hyper`` on the device at the start of a step; nothing
    here reads back to the host, so a step can be captured in a HIP graph."""

    def __init__(self, params, schedule: LearningRateExp, loss: "EnergyStdLoss", betas=(0.9, 0.999),
                 eps: float = 1e-8, weight_decay: float = 0.0):
        if not params:
            raise ValueError("no parameters")
        dtype, dev = params[0].dtype, params[0].device
        if any(p.dtype != dtype or p.device != dev for p in params):
            raise ValueError("all parameters need one dtype and device")
        n = sum(p.numel() for p in params)
        self.flat = torch.empty(n, dtype=dtype, device=dev)
        off = 0
        for p in params:
            view = self.flat[off:off + p.numel()].view_as(p)
            view.copy_(p.detach())
            p.data = view
            off += p.numel()
        self.params = list(params)
        self.exp_avg = torch.zeros_like(self.flat)
        self.exp_avg_sq = torch.zeros_like(self.flat)
        self.hyper = torch.zeros(4, dtype=torch.float64, device=dev)
        self.step_count = torch.zeros(1, dtype=torch.int64, device=dev)
        self.sched = schedule.sched(loss)
        self.betas, self.eps, self.weight_decay = tuple(betas), float(eps), float(weight_decay)

    def tick(self) -> None:
        ops.lr_tick(self.hyper, self.step_count, self.sched)

    def step(self, grads) -> None:
        g = torch.cat([x.reshape(-1) for x in grads])
        ops.adam_step(self.flat, g, self.exp_avg, self.exp_avg_sq, self.hyper, self.betas, self.eps,
                      self.weight_decay)


class Trainer:
    """Adam on the model's trainable tensors (embedding and fitting nets;
    the fitting nets only with a tabulated embedding), on the device
    (DeviceAdam).  The loss prefactors move with the learning rate from
    start_pref_* at the starter rate to limit_pref_*, as the reference's;
    the rate follows ``schedule`` (constant ``lr`` when None)."""

    def __init__(self, model, lr: float = 1e-3, pref_e: float = 0.02, pref_f: float = 1000.0,
                 limit_pref_e: float | None = None, limit_pref_f: float | None = None,
                 schedule: LearningRateExp | None = None, betas=(0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 0.0):
        self.model = model
        self.params = model.parameters()
        self.schedule = schedule if schedule is not None else LearningRateExp.constant(lr)
        self.loss = EnergyStdLoss(pref_e, pref_f, limit_pref_e, limit_pref_f,
                                  starter_learning_rate=self.schedule.start_lr)
        self.opt = DeviceAdam(self.params, self.schedule, self.loss, betas, eps, weight_decay)

    def _type_order(self, batch: Batch):
        """(perm, bounds) of the batch's atom types: one host read of the
        type counts, after which the fitting nets run on contiguous slices."""
        return self.model._host_type_order(None, batch.atype)

    def _loss_grads(self, batch: Batch, energy_label, force_label, order):
        energy, force = energy_force(self.model, batch, create_graph=True, order=order)
        h = self.opt.hyper
        loss, more = self.loss(energy, force, energy_label, force_label, batch.nloc, prefs=(h[1], h[2]))
        grads = torch.autograd.grad(loss, self.params)
        return loss.detach(), more, grads

    def loss_and_grad(self, batch: Batch, energy_label, force_label):
        """Loss and dL/dparam (left in p.grad) at the current step's rate,
        without updating the parameters or the step count."""
        self.opt.tick()
        loss, more, grads = self._loss_grads(batch, energy_label, force_label, self._type_order(batch))
        self.opt.step_count.sub_(1)
        for p, g in zip(self.params, grads):
            p.grad = g
        return loss, more

    def step(self, batch: Batch, energy_label, force_label):
        self.opt.tick()
        loss, more, grads = self._loss_grads(batch, energy_label, force_label, self._type_order(batch))
        self.opt.step(grads)
        return loss, more


class CapturedTrainStep:
    """A whole training step captured once as a HIP graph and replayed: the
    lists and env-mat of the batch (the box form without host waits), the
    forward with the force graph, the loss, the double backward and the Adam
    launch, all with the rate and prefactors of the device-side tick.  The
    frames' coordinates live in ``coords`` (copy new ones in before a replay);
    the boxes, atom types and frame sizes are fixed by the capture, the
    labels live in ``energy_label`` / ``force_label``.  ``replay`` returns the
    step's loss tensor (read it only when needed: reading it waits)."""

    def __init__(self, trainer: Trainer, coords, atypes, boxes, energy_label, force_label, warmup: int = 2):
        if any(b is None for b in boxes):
            raise ValueError("the captured step needs a periodic box for every frame")
        self.tr = trainer
        m = trainer.model
        self.coords = [torch.as_tensor(c, dtype=torch.float64, device=m.device).reshape(-1, 3).clone() for c in coords]
        self.atypes = [torch.as_tensor(t, device=m.device).to(torch.int32).clone() for t in atypes]
        self.boxes = [np.asarray(b, dtype=np.float64).copy() for b in boxes]
        self.energy_label = torch.as_tensor(energy_label, device=m.device).clone()
        self.force_label = torch.as_tensor(force_label, device=m.device).clone()
        probe = prepare(m, self.coords, self.atypes, self.boxes, wait=False)
        self.order = trainer.model._host_type_order(probe.atype, probe.atype)
        # warm-up steps on a side stream (they train: the step count and the
        # parameters move on as for any other step), then the capture
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(max(1, warmup)):
                self._step()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.loss = self._step()

    def _step(self):
        tr = self.tr
        batch = prepare(tr.model, self.coords, self.atypes, self.boxes, wait=False)
        self.status = batch.status
        tr.opt.tick()
        loss, _, grads = tr._loss_grads(batch, self.energy_label, self.force_label, self.order)
        tr.opt.step(grads)
        return loss

    def replay(self):
        self.graph.replay()
        return self.loss

    def check(self) -> None:
        """Raise if a list of the last replay failed (waits for the step)."""
        st = self.status.cpu().tolist()
        if any(st):
            raise RuntimeError(f"neighbor list status per frame {st}: 1 = more than 1024 candidates, "
                               "2 = workspace rejected, 4 = ghost capacity exceeded")
