// Cell grid shared by the neighbor search (nlist.hip) and the cell-wise
// force kernels (env_mat.hip): grid parameters, the neighbor-list workspace
// layout, and the candidate list of a cell's neighborhood.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dpmi {
namespace {

constexpr int kMaxGridNoPbc = 64;
// candidate table of one cell's neighborhood kept in LDS: first pass (all
// cells) and second pass (the cells listed as too big for the first), and
// the number of neighbor cells ((2 nsearch + 1)^3 <= 125)
constexpr int kCandCap = 1536;
constexpr int kCandCapBig = 3072;
constexpr int kMaxNb = 128;
// most workgroups sharing one cell in the cell-wise search (small boxes)
constexpr int kMaxCellSplit = 8;

// workspace status words: misc[0] nall, misc[1] error bits (1: more
// candidates than a row can hold, 2: the workspace was not set up by
// dpmi_copy_coord_f64 for this nloc / nall or is smaller than its layout),
// misc[2] per-atom redo count, misc[3] big-cell count

struct GridParams {
  double box[9];
  double rec[9];
  double lo[3];  // open systems: grid origin
  double cs[3];  // open systems: cell width
  int ncell[3];
  int ngcell[3];
  int ext[3];
  int nsearch;
  int pbc;
  int nloc;
  long long lay_nlc;   // cell counts the workspace layout was sized for
  long long lay_next;
  int magic;           // kGridMagic once dpmi_copy_coord_f64 has written the header
};

constexpr int kGridMagic = 0x494d5044;  // "DPMI"

struct WsLayout {
  int64_t params, misc, bbox, cell_of, cell_count, cell_start, cell_cursor, cell_unordered, cell_atoms, ghost_off,
      redo, bigcells, pos_of, total;
};

__host__ __device__ inline int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// Header (grid parameters, status words, bounding box) at fixed offsets, then
// the per-atom and per-cell arrays.  Host and device derive the same offsets
// from (nloc, nlc, next), so kernels need only the workspace base pointer.
__host__ __device__ inline WsLayout ws_layout(int64_t nloc, int64_t nlc, int64_t next) {
  WsLayout w;
  int64_t na = nloc > 0 ? nloc : 1;
  int64_t o = 0;
  w.params = o;
  o = align256(o + (int64_t)sizeof(GridParams));
  w.misc = o;
  o = align256(o + 4 * 16);
  w.bbox = o;
  o = align256(o + 8 * 8);
  w.redo = o;  // offset independent of the cell counts (the host reads it)
  o = align256(o + 4 * na);
  w.cell_of = o;
  o = align256(o + 4 * na);
  w.cell_count = o;
  o = align256(o + 4 * nlc);
  w.cell_start = o;
  o = align256(o + 4 * (nlc + 1));
  w.cell_cursor = o;
  o = align256(o + 4 * nlc);
  w.cell_unordered = o;
  o = align256(o + 4 * na);
  w.cell_atoms = o;
  o = align256(o + 4 * na);
  w.ghost_off = o;
  o = align256(o + 4 * (next + 1));
  w.bigcells = o;  // cells whose neighborhood exceeds the first search pass's table
  o = align256(o + 4 * nlc);
  w.pos_of = o;  // position of each local atom in cell_atoms (the cell-sorted order)
  o = align256(o + 4 * na);
  w.total = o;
  return w;
}

// compute_pbc_shift (neighbor_list.cc:727-743)
__device__ inline int pbc_shift(int idx, int n) {
  int s = 0;
  if (idx < 0) {
    s = 1;
    while (idx + s * n < 0) s++;
  } else if (idx >= n) {
    s = -1;
    while (idx + s * n >= n) s--;
  }
  return s;
}

// Neighborhood of local cell cc: for each of the (2 ns + 1)^3 neighbor cells
// (x-major offset order) the first atom (cell_atoms position for a local
// cell, extended index for a ghost cell), whether it is local, and the
// exclusive prefix of atom counts.  Block-cooperative (two barriers); returns
// the total candidate count T = s_pref[nnb].  A candidate's position
// g = s_pref[t] + k (k-th atom of neighbor cell t) is the candidate slot used
// by the search and by the cell-wise force kernels alike.
__device__ inline int nb_table(const GridParams& p, const int* __restrict__ cell_start,
                               const int* __restrict__ ghost_off, int cc, int* s_base, int* s_local, int* s_pref) {
  const int tid = threadIdx.x;
  const int ns = p.nsearch, side = 2 * ns + 1, nnb = side * side * side;
  const int c2 = cc % p.ncell[2], c1 = (cc / p.ncell[2]) % p.ncell[1], c0 = cc / (p.ncell[1] * p.ncell[2]);
  if (tid < nnb) {
    const int ii0 = c0 + tid / (side * side) - ns, ii1 = c1 + (tid / side) % side - ns, ii2 = c2 + tid % side - ns;
    int base = 0, cnt = 0, local = 1;
    if (p.pbc) {
      if (!(ii0 < -p.ngcell[0] || ii0 >= p.ncell[0] + p.ngcell[0] || ii1 < -p.ngcell[1] ||
            ii1 >= p.ncell[1] + p.ngcell[1] || ii2 < -p.ngcell[2] || ii2 >= p.ncell[2] + p.ngcell[2])) {
        local = ii0 >= 0 && ii0 < p.ncell[0] && ii1 >= 0 && ii1 < p.ncell[1] && ii2 >= 0 && ii2 < p.ncell[2];
        const int j0 = ii0 + pbc_shift(ii0, p.ncell[0]) * p.ncell[0];
        const int j1 = ii1 + pbc_shift(ii1, p.ncell[1]) * p.ncell[1];
        const int j2 = ii2 + pbc_shift(ii2, p.ncell[2]) * p.ncell[2];
        const int q = (j0 * p.ncell[1] + j1) * p.ncell[2] + j2;
        base = cell_start[q];
        cnt = cell_start[q + 1] - base;
        if (!local) {
          const int e = ((ii0 + p.ngcell[0]) * p.ext[1] + (ii1 + p.ngcell[1])) * p.ext[2] + (ii2 + p.ngcell[2]);
          base = p.nloc + ghost_off[e];
        }
      }
    } else if (!(ii0 < 0 || ii0 >= p.ncell[0] || ii1 < 0 || ii1 >= p.ncell[1] || ii2 < 0 || ii2 >= p.ncell[2])) {
      const int q = (ii0 * p.ncell[1] + ii1) * p.ncell[2] + ii2;
      base = cell_start[q];
      cnt = cell_start[q + 1] - base;
    }
    s_base[tid] = base;
    s_local[tid] = local;
    s_pref[tid + 1] = cnt;
  }
  if (tid == 0) s_pref[0] = 0;
  __syncthreads();
  if (tid == 0)
    for (int t = 0; t < nnb; ++t) s_pref[t + 1] += s_pref[t];
  __syncthreads();
  return s_pref[nnb];
}

// Neighbor cell t (last t with s_pref[t] <= g) and extended index of
// candidate slot g of the table above.
__device__ inline int nb_candidate(int g, int nnb, const int* s_base, const int* s_local, const int* s_pref,
                                   const int* __restrict__ cell_atoms, int* t_out) {
  int lo = 0, hi = nnb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_pref[mid] <= g) lo = mid; else hi = mid;
  }
  const int k = g - s_pref[lo];
  *t_out = lo;
  return s_local[lo] ? cell_atoms[s_base[lo] + k] : s_base[lo] + k;
}

}  // namespace
}  // namespace dpmi
