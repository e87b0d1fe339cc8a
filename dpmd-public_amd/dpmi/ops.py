"""Kernel wrappers and the ``torch.ops.deepmd`` op surface.

Registers, in namespace ``deepmd`` (the namespace deepmd/pt calls through):

* ``tabulate_fusion_se_a(table, table_info, em_x, em, last_layer_size) -> Tensor[]``
  — same schema, dtype dispatch (double/float by ``table.dtype``) and autograd
  structure as source/op/pt/tabulate_multi_device.cc:605-830,1498-1552:
  forward -> grad (w.r.t. em_x, em) -> grad_grad (w.r.t. dy only).
* ``prod_env_mat_a``, ``prod_force_se_a``, ``prod_virial_se_a`` — the C++/TF
  ops of source/op/tf/prod_env_mat_multi_device.cc:11-28,
  prod_force_multi_device.cc:6-14, prod_virial_multi_device.cc:5-15, which
  deepmd/pt lacks; argument meaning follows the TF ops (natoms[0] = nloc,
  natoms[1] = nall), with coordinates per frame.

Every op runs the HIP kernels of libdpmi; tensors must live on the GPU.
There is no CPU path: a CPU tensor raises, exactly as the reference raises
for a GPU tensor when its library was built without GPU support.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr

# --------------------------------------------------------------------------- helpers

_TIMED: dict | None = None


class KernelTimer:
    """Records HIP events on the launching stream around the named libdpmi
    entry points while active (used by bench.py for live per-kernel time)."""

    def __init__(self, *names):
        self.names = names
        self.events = {n: [] for n in names}

    def __enter__(self):
        global _TIMED
        _TIMED = self.events
        return self

    def __exit__(self, *exc):
        global _TIMED
        _TIMED = None

    def summary(self) -> dict:
        torch.cuda.synchronize()
        out = {}
        for n, evs in self.events.items():
            ms = [a.elapsed_time(b) for a, b in evs]
            out[n] = {"launches": len(ms), "total_ms": float(sum(ms)),
                      "avg_ms": float(sum(ms) / len(ms)) if ms else 0.0}
        return out


def _launch(name, *args):
    if _TIMED is not None and name in _TIMED:
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        call(name, *args)
        b.record()
        _TIMED[name].append((a, b))
    else:
        call(name, *args)



def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _gpu(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise RuntimeError(
                "dpmi implements the GPU path only: the input tensor is on the CPU "
                "(move it to the MI355X device)"
            )


def _c(t, dtype=None):
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def _sfx(dtype):
    if dtype == torch.float64:
        return "f64"
    if dtype == torch.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {dtype}: only float64 and float32")


def _sel_host(sel):
    return np.ascontiguousarray(np.asarray(sel, dtype=np.int32))


def _box_host(box):
    if box is None:
        return None
    if isinstance(box, torch.Tensor):
        box = box.detach().to("cpu", torch.float64).numpy()
    return np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(9))


# --------------------------------------------------------------------------- neighbor list


def nlist_workspace_bytes(box, nloc: int, rcut: float) -> int:
    out = ctypes.c_int64(0)
    b = _box_host(box)
    call("dpmi_nlist_workspace_bytes", ctypes.byref(out), ptr(b) if b is not None else None, nloc, rcut)
    return int(out.value)


def build_nlist(coord: torch.Tensor, atype: torch.Tensor, box, rcut: float, sel, mapped: bool = False,
                mem_nall: int | None = None, cells: bool = False, wait: bool = True):
    """Periodic (box given) or open neighbor list of one frame.

    coord [nloc, 3] float64, atype [nloc] int.  Returns (ext_coord [nall, 3],
    ext_type [nall] int32, mapping [nall] int32, nlist [nloc, nnei] int32 in
    extended indices, overflow [nloc] int32, nlist_mapped or None).  With
    cells=True a seventh item (cand_slot [nloc, nnei] int16, workspace) is
    returned for force_virial_cells.

    wait=False (periodic boxes): the pipeline without host waits; the
    extended arrays then have dpmi_copy_coord_capacity rows, of which the
    first nall hold atoms (nall stays on the device), and errors are read
    later through nlist_status(workspace).
    """
    _gpu(coord, atype)
    dev = coord.device
    coord = _c(coord, torch.float64).view(-1, 3)
    atype = _c(atype, torch.int32).view(-1)
    nloc = coord.shape[0]
    b = _box_host(box)
    ws_bytes = nlist_workspace_bytes(b, nloc, rcut)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    if not wait and b is not None:
        capv = ctypes.c_int64(0)
        call("dpmi_copy_coord_capacity", ctypes.byref(capv), ptr(b), nloc, rcut)
        cap = int(capv.value)
        ext_c = torch.empty((cap, 3), dtype=torch.float64, device=dev)
        ext_t = torch.empty(cap, dtype=torch.int32, device=dev)
        mapping = torch.empty(cap, dtype=torch.int32, device=dev)
        _launch("dpmi_copy_coord_nowait_f64", ptr(ext_c), ptr(ext_t), ptr(mapping), cap, ptr(coord), ptr(atype), nloc,
                ptr(b), rcut, ptr(ws), ws_bytes, _stream())
        s = _sel_host(sel)
        nnei = int(s.sum())
        nlist = torch.empty((nloc, nnei), dtype=torch.int32, device=dev)
        overflow = torch.empty(nloc, dtype=torch.int32, device=dev)
        nl_map = torch.empty((nloc, nnei), dtype=torch.int32, device=dev) if mapped else None
        cslot = torch.empty((nloc, nnei), dtype=torch.int16, device=dev) if cells else None
        _launch("dpmi_build_nlist_nowait_f64", ptr(nlist), ptr(overflow), ptr(nl_map), ptr(cslot), ptr(ext_c),
                ptr(ext_t), ptr(mapping), nloc, rcut, ptr(s), len(s), ptr(ws), ws_bytes, _stream())
        if cells:
            return ext_c, ext_t, mapping, nlist, overflow, nl_map, (cslot, ws)
        return ext_c, ext_t, mapping, nlist, overflow, nl_map
    cap = mem_nall if mem_nall is not None else (nloc * 4 + 4096 if b is not None else nloc)
    nall = ctypes.c_int64(0)
    for _ in range(8):
        ext_c = torch.empty((max(cap, 1), 3), dtype=torch.float64, device=dev)
        ext_t = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        mapping = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        try:
            _launch("dpmi_copy_coord_f64", ptr(ext_c), ptr(ext_t), ptr(mapping), ctypes.byref(nall), cap,
                 ptr(coord), ptr(atype), nloc, ptr(b) if b is not None else None, rcut, ptr(ws), ws_bytes,
                 _stream())
            break
        except _lib.DpmiCapacityError:
            cap = int(nall.value)
    n = int(nall.value)
    ext_c, ext_t, mapping = ext_c[:n], ext_t[:n], mapping[:n]
    s = _sel_host(sel)
    nnei = int(s.sum())
    nlist = torch.empty((nloc, nnei), dtype=torch.int32, device=dev)
    overflow = torch.empty(nloc, dtype=torch.int32, device=dev)
    nl_map = torch.empty((nloc, nnei), dtype=torch.int32, device=dev) if mapped else None
    cslot = torch.empty((nloc, nnei), dtype=torch.int16, device=dev) if cells else None
    _launch("dpmi_build_nlist_f64", ptr(nlist), ptr(overflow), ptr(nl_map), ptr(cslot), ptr(ext_c), ptr(ext_t),
         ptr(mapping), nloc, n, rcut, ptr(s), len(s), ptr(ws), ws_bytes, _stream())
    if cells:
        return ext_c, ext_t, mapping, nlist, overflow, nl_map, (cslot, ws)
    return ext_c, ext_t, mapping, nlist, overflow, nl_map


def nlist_status(workspace) -> torch.Tensor:
    """Status word of the neighbor-list pipeline that used this workspace,
    as an int32 device tensor (0: no error; bit 1: an atom with more than
    1024 candidates, 2: a workspace of another system, 4: more ghosts than
    the extended arrays hold).  Queued on the stream, no wait."""
    out = torch.empty(1, dtype=torch.int32, device=workspace.device)
    _launch("dpmi_nlist_status", ptr(out), ptr(workspace), workspace.numel(), _stream())
    return out


def format_nlist(cand, ext_coord, ext_type, nloc: int, rcut: float, sel, numneigh=None, mapping=None):
    """format_nlist_cpu of an external list (the LAMMPS / forward_lower input):
    cand [nloc, ncand] extended indices (-1 padded, or numneigh [nloc] valid
    entries per row) -> (nlist [nloc, nnei], overflow [nloc], nlist mapped to
    local atoms or None)."""
    _gpu(cand, ext_coord, ext_type)
    dev = ext_coord.device
    cand = _c(cand, torch.int32).view(nloc, -1)
    s = _sel_host(sel)
    nnei = int(s.sum())
    nlist = torch.empty((nloc, nnei), dtype=torch.int32, device=dev)
    overflow = torch.empty(nloc, dtype=torch.int32, device=dev)
    nl_map = torch.empty((nloc, nnei), dtype=torch.int32, device=dev) if mapping is not None else None
    status = torch.empty(1, dtype=torch.int32, device=dev)
    nn = _c(numneigh, torch.int32) if numneigh is not None else None
    _launch("dpmi_format_nlist_f64", ptr(nlist), ptr(overflow), ptr(nl_map), ptr(cand), ptr(nn), cand.shape[1],
            ptr(_c(ext_coord, torch.float64)), ptr(_c(ext_type, torch.int32)),
            ptr(_c(mapping, torch.int32)) if mapping is not None else None, nloc, ext_coord.shape[0], rcut,
            ptr(s), len(s), ptr(status), 4, _stream())
    return nlist, overflow, nl_map


# --------------------------------------------------------------------------- env-mat


def env_mat_a(nlist, ext_coord, ext_type, avg, std, rcut, rcut_smth, sel, with_deriv=True, with_rij=True,
              dtype=None):
    """prod_env_mat_a on a formatted list of one frame -> (em, em_deriv, rij).
    dtype=float32 with float64 coordinates selects the mixed kernel (pair
    differences in double, the rest in float)."""
    _gpu(nlist, ext_coord, ext_type, avg, std)
    dt = ext_coord.dtype if dtype is None else dtype
    sfx = _sfx(dt)
    if dt == torch.float32 and ext_coord.dtype == torch.float64:
        sfx = "f32c64"
    nloc, nnei = nlist.shape
    dev = ext_coord.device
    em = torch.empty((nloc, nnei, 4), dtype=dt, device=dev)
    ed = torch.empty((nloc, nnei, 4, 3), dtype=dt, device=dev) if with_deriv else None
    rij = torch.empty((nloc, nnei, 3), dtype=dt, device=dev) if with_rij else None
    s = _sel_host(sel)
    _launch(f"dpmi_prod_env_mat_a_{sfx}", ptr(em), ptr(ed), ptr(rij), ptr(_c(nlist, torch.int32)), ptr(_c(ext_coord)),
         ptr(_c(ext_type, torch.int32)), nloc, ptr(_c(avg, dt)), ptr(_c(std, dt)), rcut, rcut_smth, ptr(s), len(s),
         _stream())
    return em, ed, rij


# --------------------------------------------------------------------------- env-mat, PyTorch-backend form


def _pt_dims(coord, nlist):
    nf, nloc, nnei = nlist.shape
    nall = coord.numel() // (3 * nf) if nf > 0 else 0
    return nf, nloc, nnei, nall


def _pt_tail(coord, nlist, cfg):
    """Size, cutoff and stream arguments shared by the dpmi_env_mat_pt_* entry points."""
    nf, nloc, nnei, nall = _pt_dims(coord, nlist)
    rcut, rcut_smth, protection, exp_sw, ncomp, ntypes = cfg
    return (nf, nloc, nall, nnei, ntypes, ncomp, rcut, rcut_smth, protection, exp_sw, _stream())


class _EnvMatPtGrad(torch.autograd.Function):
    """Coordinate gradient of the PyTorch-backend env-mat; differentiable
    w.r.t. the incoming gradients (dpmi_env_mat_pt_grad_grad_*), which is
    what the parameter gradient of a force loss needs.  The second
    derivative w.r.t. the coordinates is not formed."""

    @staticmethod
    def forward(ctx, g_env, g_diff, g_sw, coord, nlist, atype, std, cfg):
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(coord, nlist, atype, std)
        ctx.cfg = cfg
        ctx.has = (g_env is not None, g_diff is not None, g_sw is not None)
        ctx.shapes = tuple(None if g is None else g.shape for g in (g_env, g_diff, g_sw))
        nf, nloc, nnei, nall = _pt_dims(coord, nlist)
        gc = torch.empty((nf, nall, 3), dtype=coord.dtype, device=coord.device)
        g_env, g_diff, g_sw = (None if g is None else _c(g, coord.dtype) for g in (g_env, g_diff, g_sw))
        _launch(f"dpmi_env_mat_pt_grad_{_sfx(coord.dtype)}", ptr(gc), ptr(g_env), ptr(g_diff), ptr(g_sw),
                ptr(coord), ptr(nlist), ptr(atype), ptr(std), *_pt_tail(coord, nlist, cfg))
        return gc

    @staticmethod
    def backward(ctx, v):
        if v is None:
            return (None,) * 8
        coord, nlist, atype, std = ctx.saved_tensors
        cfg = ctx.cfg
        nf, nloc, nnei, nall = _pt_dims(coord, nlist)
        ncomp = cfg[4]
        need = ctx.needs_input_grad
        ge = torch.empty((nf, nloc, nnei, ncomp), dtype=coord.dtype, device=coord.device) if need[0] else None
        gd = torch.empty((nf, nloc, nnei, 3), dtype=coord.dtype, device=coord.device) if need[1] else None
        gs = torch.empty((nf, nloc, nnei, 1), dtype=coord.dtype, device=coord.device) if need[2] else None
        _launch(f"dpmi_env_mat_pt_grad_grad_{_sfx(coord.dtype)}", ptr(ge), ptr(gd), ptr(gs),
                ptr(_c(v, coord.dtype).reshape(nf, nall, 3)), ptr(coord), ptr(nlist), ptr(atype), ptr(std),
                *_pt_tail(coord, nlist, cfg))
        shp = ctx.shapes
        return (ge.reshape(shp[0]) if ge is not None else None, gd.reshape(shp[1]) if gd is not None else None,
                gs.reshape(shp[2]) if gs is not None else None, None, None, None, None, None)


class _EnvMatPt(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coord, nlist, atype, mean, std, cfg):
        ctx.set_materialize_grads(False)
        nf, nloc, nnei, nall = _pt_dims(coord, nlist)
        ncomp = cfg[4]
        dt, dev = coord.dtype, coord.device
        env = torch.empty((nf, nloc, nnei, ncomp), dtype=dt, device=dev)
        diff = torch.empty((nf, nloc, nnei, 3), dtype=dt, device=dev)
        sw = torch.empty((nf, nloc, nnei, 1), dtype=dt, device=dev)
        _launch(f"dpmi_env_mat_pt_{_sfx(dt)}", ptr(env), ptr(diff), ptr(sw), ptr(coord), ptr(nlist), ptr(atype),
                ptr(mean), ptr(std), *_pt_tail(coord, nlist, cfg))
        ctx.save_for_backward(coord, nlist, atype, std)
        ctx.cfg = cfg
        return env, diff, sw

    @staticmethod
    def backward(ctx, g_env, g_diff, g_sw):
        coord, nlist, atype, std = ctx.saved_tensors
        if g_env is None and g_diff is None and g_sw is None:
            return (None,) * 6
        gc = _EnvMatPtGrad.apply(g_env, g_diff, g_sw, coord, nlist, atype, std, ctx.cfg)
        return gc, None, None, None, None, None


def prod_env_mat(extended_coord, nlist, atype, mean, stddev, rcut: float, rcut_smth: float, radial_only: bool = False,
                 protection: float = 0.0, use_exp_switch: bool = False):
    """Drop-in for deepmd.pt.model.descriptor.env_mat.prod_env_mat
    (env_mat.py:51-92), which DescrptBlockSeA.forward calls
    (se_a.py:781-790): (env_mat [nf, nloc, nnei, 4 or 1], diff
    [nf, nloc, nnei, 3], switch [nf, nloc, nnei, 1]) from extended_coord
    [nf, nall * 3] (or [nf, nall, 3]), nlist [nf, nloc, nnei] (any order,
    -1 padding, entries beyond rcut allowed), atype [nf, nloc] and
    mean / stddev [ntypes, nnei, 4 or 1].  Differentiable w.r.t.
    extended_coord (HIP kernels for the gradient and for its own backward
    w.r.t. the incoming gradients, so an E + F loss trains through it);
    float64 or float32."""
    _gpu(extended_coord, nlist, atype, mean, stddev)
    dt = extended_coord.dtype
    _sfx(dt)
    if nlist.dim() != 3:
        raise ValueError("nlist must be [nframes, nloc, nnei]")
    nf, nloc, nnei = nlist.shape
    ncomp = 1 if radial_only else 4
    if mean.shape[-1] != ncomp or stddev.shape[-1] != ncomp:
        raise ValueError(f"mean / stddev must end in {ncomp} components")
    ntypes = mean.shape[0]
    if mean.numel() != ntypes * nnei * ncomp or stddev.numel() != mean.numel():
        raise ValueError("mean / stddev must be [ntypes, nnei, ncomp]")
    if atype.numel() != nf * nloc:
        raise ValueError("atype must be [nframes, nloc]")
    if extended_coord.numel() % (3 * max(nf, 1)) != 0:
        raise ValueError("extended_coord must be [nframes, nall * 3]")
    if float(rcut_smth) >= float(rcut):
        raise ValueError("rmin should be less than rmax.")
    coord = extended_coord.reshape(nf, -1, 3)
    if not coord.is_contiguous():
        coord = coord.contiguous()
    cfg = (float(rcut), float(rcut_smth), float(protection), int(bool(use_exp_switch)), ncomp, int(ntypes))
    env, diff, sw = _EnvMatPt.apply(coord, _c(nlist, torch.int32), _c(atype, torch.int32).reshape(nf, nloc),
                                    _c(mean, dt), _c(stddev, dt), cfg)
    return env, diff, sw


def force_virial_recompute(net_deriv, nlist, mapping, ext_coord, ext_type, std, rcut, rcut_smth, sel,
                           want_virial=True):
    """Adjoint of the env-mat: forces (-dE/dx) from dE/d em, with the env-mat
    derivative recomputed.  mapping given -> forces folded onto local atoms."""
    _gpu(net_deriv, nlist, ext_coord, ext_type, std)
    sfx = _sfx(net_deriv.dtype)
    nloc = nlist.shape[0]
    nall = ext_coord.shape[0]
    dev = ext_coord.device
    force = torch.empty((nloc if mapping is not None else nall, 3), dtype=torch.float64, device=dev)
    virial = torch.empty(9, dtype=torch.float64, device=dev) if want_virial else None
    wsb = ctypes.c_int64(0)
    call("dpmi_prod_virial_workspace_bytes", ctypes.byref(wsb), nloc, 1)
    ws = torch.empty(int(wsb.value), dtype=torch.uint8, device=dev)
    s = _sel_host(sel)
    _launch(f"dpmi_force_virial_recompute_{sfx}", ptr(force), ptr(virial), ptr(_c(net_deriv)),
         ptr(_c(nlist, torch.int32)), ptr(_c(mapping, torch.int32)) if mapping is not None else None,
         ptr(_c(ext_coord, torch.float64)), ptr(_c(ext_type, torch.int32)), nloc, nall,
         ptr(_c(std, net_deriv.dtype)), rcut, rcut_smth, ptr(s), len(s), ptr(ws), int(wsb.value), _stream())
    return force, virial


def force_virial_cells(net_deriv, nlist, cells, mapping, ext_coord, ext_type, std, rcut, rcut_smth, sel, box,
                       want_virial=True, workspace=None):
    """force_virial_recompute for a list from build_nlist(..., cells=True)
    (cells = its (cand_slot, workspace)): no memory atomics for the neighbor
    forces, deterministic.  Forces folded onto local atoms.  workspace
    (optional): a uint8 device tensor of at least
    dpmi_force_virial_cells_workspace_bytes, whose contents are ignored."""
    _gpu(net_deriv, nlist, ext_coord, ext_type, std)
    sfx = _sfx(net_deriv.dtype)
    cslot, nlws = cells
    nloc = nlist.shape[0]
    nall = ext_coord.shape[0]
    dev = ext_coord.device
    b = _box_host(box)
    force = torch.empty((nloc, 3), dtype=torch.float64, device=dev)
    virial = torch.empty(9, dtype=torch.float64, device=dev) if want_virial else None
    wsb = ctypes.c_int64(0)
    call("dpmi_force_virial_cells_workspace_bytes", ctypes.byref(wsb), ptr(b) if b is not None else None, nloc,
         rcut, int(net_deriv.dtype == torch.float64))
    if workspace is None:
        ws = torch.empty(int(wsb.value), dtype=torch.uint8, device=dev)
    else:
        if workspace.numel() < int(wsb.value) or workspace.dtype != torch.uint8:
            raise ValueError(f"workspace needs {int(wsb.value)} bytes (uint8)")
        ws = workspace
    s = _sel_host(sel)
    _launch(f"dpmi_force_virial_cells_{sfx}", ptr(force), ptr(virial), ptr(_c(net_deriv)),
         ptr(_c(nlist, torch.int32)), ptr(cslot), ptr(_c(mapping, torch.int32)) if mapping is not None else None,
         ptr(_c(ext_coord, torch.float64)), ptr(_c(ext_type, torch.int32)), nloc, nall,
         ptr(_c(std, net_deriv.dtype)), rcut, rcut_smth, ptr(s), len(s), ptr(b) if b is not None else None,
         ptr(nlws), ptr(ws), int(wsb.value), _stream())
    return force, virial


def prod_force(net_deriv, in_deriv, nlist, nloc: int, nall: int, nframes: int = 1):
    _gpu(net_deriv, in_deriv, nlist)
    dt = net_deriv.dtype
    nnei = nlist.numel() // (nloc * nframes) if nloc * nframes > 0 else 0
    force = torch.empty((nframes, nall * 3), dtype=dt, device=net_deriv.device)
    _launch(f"dpmi_prod_force_se_a_{_sfx(dt)}", ptr(force), ptr(_c(net_deriv)), ptr(_c(in_deriv, dt)),
         ptr(_c(nlist, torch.int32)), nloc, nall, nnei, nframes, _stream())
    return force


def prod_virial(net_deriv, in_deriv, rij, nlist, nloc: int, nall: int, nframes: int = 1, atomic: bool = True):
    _gpu(net_deriv, in_deriv, rij, nlist)
    dt = net_deriv.dtype
    dev = net_deriv.device
    nnei = nlist.numel() // (nloc * nframes) if nloc * nframes > 0 else 0
    virial = torch.empty((nframes, 9), dtype=dt, device=dev)
    atom_virial = torch.empty((nframes, nall * 9), dtype=dt, device=dev) if atomic else None
    wsb = ctypes.c_int64(0)
    call("dpmi_prod_virial_workspace_bytes", ctypes.byref(wsb), nloc, nframes)
    ws = torch.empty(int(wsb.value), dtype=torch.uint8, device=dev)
    _launch(f"dpmi_prod_virial_se_a_{_sfx(dt)}", ptr(virial), ptr(atom_virial), ptr(_c(net_deriv)),
         ptr(_c(in_deriv, dt)), ptr(_c(rij, dt)), ptr(_c(nlist, torch.int32)), nloc, nall, nnei, nframes, ptr(ws),
         int(wsb.value), _stream())
    return virial, atom_virial


def prod_force_grad(grad, in_deriv, nlist, nloc: int, nall: int, nframes: int = 1):
    """grad [nframes, nall * 3] -> grad_net [nframes, nloc * 4 nnei]
    (ProdForceSeAGrad; nall = nloc is the reference op)."""
    _gpu(grad, in_deriv, nlist)
    dt = in_deriv.dtype
    nnei = nlist.numel() // (nloc * nframes) if nloc * nframes > 0 else 0
    gn = torch.empty((nframes, nloc * nnei * 4), dtype=dt, device=grad.device)
    _launch(f"dpmi_prod_force_grad_a_{_sfx(dt)}", ptr(gn), ptr(_c(grad, dt)), ptr(_c(in_deriv)),
            ptr(_c(nlist, torch.int32)), nloc, nall, nnei, nframes, _stream())
    return gn


def prod_virial_grad(grad, in_deriv, rij, nlist, nloc: int, nframes: int = 1):
    """grad [nframes, 9] -> grad_net [nframes, nloc * 4 nnei] (ProdVirialSeAGrad)."""
    _gpu(grad, in_deriv, rij, nlist)
    dt = in_deriv.dtype
    nnei = nlist.numel() // (nloc * nframes) if nloc * nframes > 0 else 0
    gn = torch.empty((nframes, nloc * nnei * 4), dtype=dt, device=grad.device)
    _launch(f"dpmi_prod_virial_grad_a_{_sfx(dt)}", ptr(gn), ptr(_c(grad, dt)), ptr(_c(in_deriv)),
            ptr(_c(rij, dt)), ptr(_c(nlist, torch.int32)), nloc, nnei, nframes, _stream())
    return gn


class ProdForceOp(torch.autograd.Function):
    """force = prod_force(net_deriv) with its gradient w.r.t. net_deriv from
    the ProdForceSeAGrad kernel (the reference's registered gradient,
    deepmd/tf/op/_prod_force_se_a_grad.py:16-28).  The backward is itself
    linear in the incoming gradient, so a second differentiation needs none
    of its own."""

    @staticmethod
    def forward(ctx, net_deriv, in_deriv, nlist, nloc, nall, nframes):
        ctx.save_for_backward(in_deriv, nlist)
        ctx.dims = (nloc, nall, nframes)
        return prod_force(net_deriv, in_deriv, nlist, nloc, nall, nframes)

    @staticmethod
    def backward(ctx, g):
        in_deriv, nlist = ctx.saved_tensors
        nloc, nall, nframes = ctx.dims
        gn = prod_force_grad(g.contiguous(), in_deriv, nlist, nloc, nall, nframes)
        return gn, None, None, None, None, None


class ProdVirialOp(torch.autograd.Function):
    """(virial, atom_virial) = prod_virial(net_deriv); the gradient w.r.t.
    net_deriv takes the virial's gradient only, as the reference's registered
    gradient does (deepmd/tf/op/_prod_virial_se_a_grad.py:16-30)."""

    @staticmethod
    def forward(ctx, net_deriv, in_deriv, rij, nlist, nloc, nall, nframes):
        ctx.save_for_backward(in_deriv, rij, nlist)
        ctx.dims = (nloc, nframes)
        return prod_virial(net_deriv, in_deriv, rij, nlist, nloc, nall, nframes)

    @staticmethod
    def backward(ctx, g, g_atom):
        in_deriv, rij, nlist = ctx.saved_tensors
        nloc, nframes = ctx.dims
        if g is None:
            return None, None, None, None, None, None, None
        gn = prod_virial_grad(g.contiguous(), in_deriv, rij, nlist, nloc, nframes)
        return gn, None, None, None, None, None, None


# --------------------------------------------------------------------------- tabulate


def _info_host(table_info, dtype):
    # The reference reads table_info on the host (source/op/pt/tabulate_multi_device.cc,
    # tabulate.cu:999-1001); the caller passes a CPU tensor.
    arr = table_info.detach().to("cpu").numpy().astype(np.float64 if dtype == torch.float64 else np.float32)
    return np.ascontiguousarray(arr.reshape(-1))


def _tab_dims(em_x, em):
    if em_x.dim() != 2:
        raise ValueError("Dim of input should be 2")
    if em.dim() != 3:
        raise ValueError("Dim of input should be 3")
    return em.shape[0], em.shape[1]


def tabulate_forward(table, table_info, em_x, em, last_layer_size: int, is_sorted: bool = True):
    if table.dim() != 2:
        raise ValueError("Dim of table should be 2")
    _gpu(table, em_x, em)
    dt = table.dtype
    nloc, nnei = _tab_dims(em_x, em)
    out = torch.empty((nloc, 4, last_layer_size), dtype=dt, device=table.device)
    info = _info_host(table_info, dt)
    _launch(f"dpmi_tabulate_fusion_se_a_{_sfx(dt)}", ptr(out), ptr(_c(table)), ptr(info), ptr(_c(em_x, dt)),
         ptr(_c(em, dt)), nloc, nnei, last_layer_size, int(is_sorted), _stream())
    return out


def tabulate_grad(table, table_info, em_x, em, dy, is_sorted: bool = True):
    if dy.dim() != 3:
        raise ValueError("Dim of dy_tensor should be 3")
    _gpu(table, em_x, em, dy)
    dt = table.dtype
    nloc, nnei = _tab_dims(em_x, em)
    M = dy.shape[2]
    dx = torch.empty_like(em_x, dtype=dt)
    dem = torch.empty_like(em, dtype=dt)
    info = _info_host(table_info, dt)
    _launch(f"dpmi_tabulate_fusion_se_a_grad_{_sfx(dt)}", ptr(dx), ptr(dem), ptr(_c(table)), ptr(info),
         ptr(_c(em_x, dt)), ptr(_c(em, dt)), ptr(_c(dy, dt)), nloc, nnei, M, int(is_sorted), _stream())
    return dx, dem


def tabulate_grad_grad(table, table_info, em_x, em, dz_dy_dem_x, dz_dy_dem, M: int, is_sorted: bool = True):
    if dz_dy_dem_x.dim() != 2:
        raise ValueError("Dim of dz_dy_dem_x should be 2")
    if dz_dy_dem.dim() != 3:
        raise ValueError("Dim of dz_dy_dem should be 3")
    _gpu(table, em_x, em, dz_dy_dem_x, dz_dy_dem)
    dt = table.dtype
    nloc, nnei = _tab_dims(em_x, em)
    out = torch.empty((nloc, 4, M), dtype=dt, device=table.device)
    info = _info_host(table_info, dt)
    _launch(f"dpmi_tabulate_fusion_se_a_grad_grad_{_sfx(dt)}", ptr(out), ptr(_c(table)), ptr(info), ptr(_c(em_x, dt)),
         ptr(_c(em, dt)), ptr(_c(dz_dy_dem_x, dt)), ptr(_c(dz_dy_dem, dt)), nloc, nnei, M, int(is_sorted), _stream())
    return out


class TabulateFusionSeAGradOp(torch.autograd.Function):
    """(dy_dem_x, dy_dem) = grad(dy); differentiable w.r.t. dy only
    (TabulateFusionSeAGradOp, tabulate_multi_device.cc:605-700)."""

    @staticmethod
    def forward(ctx, table, table_info, em_x, em, dy, descriptor):
        dx, dem = tabulate_grad(table, table_info, em_x, em, dy)
        ctx.save_for_backward(table, table_info, em_x, em, descriptor)
        return dx, dem

    @staticmethod
    def backward(ctx, g_dx, g_dem):
        table, table_info, em_x, em, descriptor = ctx.saved_tensors
        g_dx = g_dx.contiguous() if g_dx is not None else torch.zeros_like(em_x)
        g_dem = g_dem.contiguous() if g_dem is not None else torch.zeros_like(em)
        dz_dy = tabulate_grad_grad(table, table_info, em_x, em, g_dx, g_dem, descriptor.shape[2], True)
        return None, None, None, None, dz_dy, None


class TabulateFusionSeAOp(torch.autograd.Function):
    """descriptor = tabulate(em_x, em) (TabulateFusionSeAOp,
    tabulate_multi_device.cc:702-830)."""

    @staticmethod
    def forward(ctx, table, table_info, em_x, em, last_layer_size):
        out = tabulate_forward(table, table_info, em_x, em, last_layer_size)
        ctx.save_for_backward(table, table_info, em_x, em, out)
        return out

    @staticmethod
    def backward(ctx, dy):
        table, table_info, em_x, em, out = ctx.saved_tensors
        dx, dem = TabulateFusionSeAGradOp.apply(table, table_info, em_x, em, dy.contiguous(), out)
        return None, None, dx, dem, None


def tabulate_fusion_se_a(table, table_info, em_x, em, last_layer_size: int):
    return [TabulateFusionSeAOp.apply(table, table_info, em_x, em, last_layer_size)]


# --------------------------------------------------------------------------- prod_env_mat_a op


class ProdEnvMatAOp(torch.autograd.Function):
    """Env-mat of every frame with the neighbor list built on the GPU; the
    backward is prod_force with the returned derivative (forces on local
    atoms, the list being mapped to local indices as ProdEnvMatAOp returns it)."""

    @staticmethod
    def forward(ctx, coord, atype, box, davg, dstd, rcut, rcut_smth, sel):
        nf = coord.shape[0]
        coord = coord.reshape(nf, -1, 3)
        atype = atype.reshape(nf, -1)
        nloc = coord.shape[1]
        ems, eds, rijs, nls = [], [], [], []
        for f in range(nf):
            b = None if box is None else box.reshape(nf, 9)[f]
            ext_c, ext_t, mapping, nlist, _ovf, nl_map = build_nlist(coord[f].detach(), atype[f], b, rcut, sel,
                                                                      mapped=True)
            em, ed, rij = env_mat_a(nlist, ext_c.to(coord.dtype), ext_t, davg, dstd, rcut, rcut_smth, sel)
            ems.append(em.reshape(nloc, -1))
            eds.append(ed.reshape(nloc, -1))
            rijs.append(rij.reshape(nloc, -1))
            nls.append(nl_map)
        em = torch.stack(ems).reshape(nf, -1)
        ed = torch.stack(eds).reshape(nf, -1)
        rij = torch.stack(rijs).reshape(nf, -1)
        nl = torch.stack(nls).reshape(nf, -1)
        ctx.save_for_backward(ed, nl)
        ctx.dims = (nf, nloc)
        ctx.mark_non_differentiable(ed, rij, nl)
        return em, ed, rij, nl

    @staticmethod
    def backward(ctx, g_em, g_ed, g_rij, g_nl):
        ed, nl = ctx.saved_tensors
        nf, nloc = ctx.dims
        force = prod_force(g_em.contiguous(), ed, nl, nloc, nloc, nf)
        return -force.reshape(nf, nloc, 3), None, None, None, None, None, None, None


def prod_env_mat_a(coord, atype, box, davg, dstd, rcut: float, rcut_smth: float, sel):
    return list(ProdEnvMatAOp.apply(coord, atype, box, davg, dstd, float(rcut), float(rcut_smth), list(sel)))


def prod_force_se_a(net_deriv, in_deriv, nlist, natoms, n_a_sel: int, n_r_sel: int):
    nf = net_deriv.shape[0]
    na = natoms.detach().to("cpu").tolist()
    return ProdForceOp.apply(net_deriv, in_deriv, nlist, int(na[0]), int(na[1]), nf)


def prod_virial_se_a(net_deriv, in_deriv, rij, nlist, natoms, n_a_sel: int, n_r_sel: int):
    nf = net_deriv.shape[0]
    na = natoms.detach().to("cpu").tolist()
    v, av = ProdVirialOp.apply(net_deriv, in_deriv, rij, nlist, int(na[0]), int(na[1]), nf)
    return [v, av]


# --------------------------------------------------------------------------- training step

def lr_tick(hyper: torch.Tensor, step: torch.Tensor, sched) -> None:
    """Start of a training step on the device: hyper [4] float64 = [lr_k,
    pref_e, pref_f, k + 1] for k = step (int64 [1]), then step = k + 1
    (dpmi_lr_tick; sched is the 11-value schedule of the C header)."""
    _gpu(hyper, step)
    if hyper.dtype != torch.float64 or hyper.numel() < 4 or not hyper.is_contiguous():
        raise ValueError("hyper must be a contiguous float64 tensor of 4 values")
    if step.dtype != torch.int64 or step.numel() != 1:
        raise ValueError("step must be one int64 value")
    s = np.ascontiguousarray(np.asarray(sched, dtype=np.float64))
    if s.shape != (11,):
        raise ValueError("the schedule has 11 values")
    call("dpmi_lr_tick", ptr(hyper), ptr(step), ptr(s), _stream())


def adam_step(param, grad, exp_avg, exp_avg_sq, hyper, betas=(0.9, 0.999), eps: float = 1e-8,
              weight_decay: float = 0.0) -> None:
    """In-place Adam over flat contiguous buffers of one dtype (dpmi_adam_*),
    lr and t read from hyper on the device."""
    _gpu(param, grad, exp_avg, exp_avg_sq, hyper)
    ts = (param, grad, exp_avg, exp_avg_sq)
    if any(t.dtype != param.dtype or not t.is_contiguous() or t.numel() != param.numel() for t in ts):
        raise ValueError("param, grad, exp_avg and exp_avg_sq must be contiguous, of one dtype and size")
    _launch(f"dpmi_adam_{_sfx(param.dtype)}", ptr(param), ptr(grad), ptr(exp_avg), ptr(exp_avg_sq), param.numel(),
            ptr(hyper), float(betas[0]), float(betas[1]), float(eps), float(weight_decay), _stream())


def row_sum(x: torch.Tensor) -> torch.Tensor:
    """Column sums of a 2-D tensor [R, F] in a fixed order (dpmi_row_sum_*)."""
    _gpu(x)
    if x.dim() != 2:
        raise ValueError("row_sum takes a 2-D tensor")
    x = x.contiguous()
    R, F = x.shape
    out = torch.empty(F, dtype=x.dtype, device=x.device)
    part = torch.empty(256 * F, dtype=x.dtype, device=x.device)
    _launch(f"dpmi_row_sum_{_sfx(x.dtype)}", ptr(out), ptr(x), ptr(part), R, F, _stream())
    return out


class RowBroadcast(torch.autograd.Function):
    """v [F] -> [R, F] with every row v; the gradient is the fixed-order
    row sum (RowSum), whose gradient is this broadcast again, so the pair is
    differentiable to any order."""

    @staticmethod
    def forward(ctx, v, rows: int):
        ctx.rows = rows
        return v.unsqueeze(0).expand(rows, v.shape[0])

    @staticmethod
    def backward(ctx, g):
        return RowSum.apply(g), None


class RowSum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        ctx.rows = x.shape[0]
        return row_sum(x)

    @staticmethod
    def backward(ctx, g):
        return RowBroadcast.apply(g, ctx.rows)


# --------------------------------------------------------------------------- registration

_REGISTERED = False
_LIBS = []


def register() -> None:
    """Define the deepmd:: ops (once per process)."""
    global _REGISTERED
    if _REGISTERED:
        return
    existing = getattr(torch.ops, "deepmd", None)
    if existing is not None and hasattr(existing, "tabulate_fusion_se_a"):
        try:
            existing.tabulate_fusion_se_a.default  # noqa: B018
            raise RuntimeError("torch.ops.deepmd.tabulate_fusion_se_a is already defined by another library")
        except AttributeError:
            pass
    lib = torch.library.Library("deepmd", "FRAGMENT")
    lib.define("tabulate_fusion_se_a(Tensor table, Tensor table_info, Tensor em_x, Tensor em, "
               "int last_layer_size) -> Tensor[]")
    lib.impl("tabulate_fusion_se_a", tabulate_fusion_se_a, "CompositeImplicitAutograd")
    lib.define("prod_env_mat_a(Tensor coord, Tensor atype, Tensor? box, Tensor davg, Tensor dstd, "
               "float rcut, float rcut_smth, int[] sel) -> Tensor[]")
    lib.impl("prod_env_mat_a", prod_env_mat_a, "CompositeImplicitAutograd")
    lib.define("prod_force_se_a(Tensor net_deriv, Tensor in_deriv, Tensor nlist, Tensor natoms, "
               "int n_a_sel, int n_r_sel) -> Tensor")
    lib.impl("prod_force_se_a", prod_force_se_a, "CompositeImplicitAutograd")
    lib.define("prod_virial_se_a(Tensor net_deriv, Tensor in_deriv, Tensor rij, Tensor nlist, Tensor natoms, "
               "int n_a_sel, int n_r_sel) -> Tensor[]")
    lib.impl("prod_virial_se_a", prod_virial_se_a, "CompositeImplicitAutograd")
    _LIBS.append(lib)
    _REGISTERED = True
