"""Neighbor-list strategy with the interface of the reference's pluggable
builders (NeighborList.build, deepmd/dpmodel/utils/neighbor_list.py:54-105;
the default all-pairs builder is DefaultNeighborList,
deepmd/dpmodel/utils/default_neighbor_list.py:25-66).

GpuNeighborList().build(coord, atype, box, rcut, sel) returns the extended
representation the lower interface takes, built by the cell-list kernels
(normalize + periodic ghosts in the copy_coord layout, then the formatted
list).  It is injected where the reference takes a strategy:
``model(coord, atype, box, neighbor_list=GpuNeighborList())``
(deepmd/dpmodel/model/make_model.py:99-117).

Differences from DefaultNeighborList, none of which changes the energy:

* ghosts follow copy_coord (neighbor_list.cc:745-845: the ghost cells of a
  cell grid, locals first) instead of whole periodic images, so nall is
  smaller and differs between frames; frames are padded to the largest nall
  with atoms of type -1 (virtual: make_atom_mask treats type < 0 as absent,
  no list refers to them), coordinate 0 and mapping 0;
* the list is already formatted: per type section of sel, nearest first
  (float r^2, then index), -1 padding.  The lower interface's
  _format_nlist keeps it (its width is sum(sel)) and nlist_distinguish_types
  reproduces the same sections.
"""

from __future__ import annotations

import torch

from . import ops


class GpuNeighborList:
    """Stateless strategy: rcut and sel come with each call, as in the
    reference's NeighborList."""

    def build(self, coord, atype, box, rcut: float, sel, return_mode: str = "extended"):
        """(extended_coord [nf, nall, 3], extended_atype [nf, nall],
        nlist [nf, nloc, sum(sel)], mapping [nf, nall]) of every frame.
        coord [nf, nloc, 3] or [nf, nloc * 3]; atype [nf, nloc]; box
        [nf, 9], [nf, 3, 3] or None (open system).  Index tensors are int64,
        as torch builds them in the reference."""
        if return_mode != "extended":
            raise NotImplementedError("GpuNeighborList only supports the extended-coordinate contract.")
        ops._gpu(coord, atype)
        nf, nloc = atype.shape[:2]
        c = coord.reshape(nf, nloc, 3)
        b = None if box is None else box.reshape(nf, 9)
        out = []
        for f in range(nf):
            ext_c, ext_t, mapping, nlist, _ovf, _ = ops.build_nlist(
                c[f].detach().to(torch.float64), atype[f], None if b is None else b[f], rcut, sel)
            out.append((ext_c, ext_t, mapping, nlist))
        nall = max([o[0].shape[0] for o in out], default=nloc)
        dev = coord.device
        ext_coord = torch.zeros((nf, nall, 3), dtype=coord.dtype, device=dev)
        ext_atype = torch.full((nf, nall), -1, dtype=atype.dtype, device=dev)
        mapping = torch.zeros((nf, nall), dtype=torch.int64, device=dev)
        nlist = torch.empty((nf, nloc, sum(int(s) for s in sel)), dtype=torch.int64, device=dev)
        for f, (ec, et, mp, nl) in enumerate(out):
            n = ec.shape[0]
            ext_coord[f, :n] = ec.to(coord.dtype)
            ext_atype[f, :n] = et.to(atype.dtype)
            mapping[f, :n] = mp.to(torch.int64)
            nlist[f] = nl.to(torch.int64)
        return ext_coord, ext_atype, nlist, mapping
