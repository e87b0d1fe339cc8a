// Periodic cell-list neighbor search for se_e2_a on gfx950.
//
// Pipeline (one stream, no host round trip except reading nall once):
//   1. normalize + bin   one thread per local atom: wrap into the box exactly
//                        as normalize_coord_cpu does, assign the build_clist
//                        cell, count atoms per cell (atomics).
//   2. scan              one workgroup: exclusive prefix sums of the cell
//                        counts and of the ghost-cell counts in the
//                        (x, y, z) loop order of copy_coord; gives nall.
//   3. order             stable per-cell order by atom index (rank = number
//                        of cell members with a smaller index), so the ghost
//                        layout equals copy_coord's push_back order.
//   4. ghosts            one wavefront per extended cell writes its images.
//   5. search + format   one wavefront per local atom scans the 27 (or 125)
//                        neighbouring cells, filters with the double test of
//                        build_nlist_cpu and the float test of
//                        format_nlist_i_cpu, compacts candidates into LDS with
//                        ballot/mbcnt prefix sums, ranks them by the 64-bit key
//                        (type, float r^2 bits, extended index) and writes the
//                        type sections of the formatted row.
//
// Bit-exactness: every floating-point expression follows the reference
// evaluation order and this file is compiled with -ffp-contract=off, so no
// multiply-add is fused; divisions and sqrt are correctly rounded on gfx950.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <cmath>
#include <cstring>
#include <vector>

#include "cells.h"
#include "dpmi_internal.h"

namespace dpmi {
namespace {

constexpr int kMaxTypes = 64;



struct SelParams {
  int ntypes;
  int nnei;
  int sel[kMaxTypes];
  int sec[kMaxTypes + 1];
};

// sel and the section starts, copied to LDS once per workgroup: indexed by a
// per-lane type, the kernel-argument arrays would be read from memory by
// every lane (a round trip per key)
struct SelLds {
  int sel[kMaxTypes];
  int sec[kMaxTypes + 1];
};

__device__ inline void stage_sel(const SelParams& sp, SelLds* sl) {
  if (threadIdx.x == 0)
    for (int t = 0; t < sp.ntypes; ++t) {
      sl->sel[t] = sp.sel[t];
      sl->sec[t] = sp.sec[t];
    }
}

// ---------------------------------------------------------------- host geometry
// Same formulas as the reference region code (source/lib/src/region.cc:35-122,
// SimulationRegion_Impl.h:344-470); x86-64 host code has no FMA, so these
// scalars are bit-identical to the reference's.
double box_volume(const double* b) {
  double v = b[0] * (b[4] * b[8] - b[7] * b[5]) - b[1] * (b[3] * b[8] - b[6] * b[5]) +
             b[2] * (b[3] * b[7] - b[6] * b[4]);
  return std::fabs(v);
}

void rec_box(const double* b, double* rec) {
  double vi = 1. / box_volume(b);
  rec[0] = (b[4] * b[8] - b[7] * b[5]) * vi;
  rec[4] = (b[0] * b[8] - b[6] * b[2]) * vi;
  rec[8] = (b[0] * b[4] - b[3] * b[1]) * vi;
  rec[1] = (-b[3] * b[8] + b[6] * b[5]) * vi;
  rec[2] = (b[3] * b[7] - b[6] * b[4]) * vi;
  rec[3] = (-b[1] * b[8] + b[7] * b[2]) * vi;
  rec[5] = (-b[0] * b[7] + b[6] * b[1]) * vi;
  rec[6] = (b[1] * b[5] - b[4] * b[2]) * vi;
  rec[7] = (-b[0] * b[5] + b[3] * b[2]) * vi;
}

void face_distance(const double* b, double* dd) {
  double vol = box_volume(b);
  auto cp = [](const double* a, const double* c, double* o) {
    o[0] = a[1] * c[2] - a[2] * c[1];
    o[1] = a[2] * c[0] - a[0] * c[2];
    o[2] = a[0] * c[1] - a[1] * c[0];
  };
  double t[3];
  cp(b + 3, b + 6, t);
  dd[0] = vol * (1. / std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
  cp(b + 6, b + 0, t);
  dd[1] = vol * (1. / std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
  cp(b + 0, b + 3, t);
  dd[2] = vol * (1. / std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
}

// Cell grid of copy_coord (neighbor_list.cc:756-770): ncell = int(face / rc),
// ngcell = int(rc / cell_size) + 1, rc being the float cutoff as a double.
GridParams pbc_grid(const double* box, float rcut, int nloc) {
  GridParams g;
  std::memset(&g, 0, sizeof(g));
  std::memcpy(g.box, box, sizeof(double) * 9);
  rec_box(box, g.rec);
  double to_face[3], rc = (double)rcut;
  face_distance(box, to_face);
  g.nsearch = 1;
  for (int d = 0; d < 3; ++d) {
    g.ncell[d] = (int)(to_face[d] / rc);
    if (g.ncell[d] == 0) g.ncell[d] = 1;
    double cell_size = to_face[d] / g.ncell[d];
    g.ngcell[d] = (int)(rc / cell_size) + 1;
    g.ext[d] = g.ncell[d] + 2 * g.ngcell[d];
    // a cell only just wider than rc leaves no margin for rounding at a cell
    // boundary: search two cells out there (never needed for C0-C4).
    if (cell_size - rc < 1e-6 * to_face[d]) g.nsearch = 2;
  }
  g.pbc = 1;
  g.nloc = nloc;
  return g;
}

bool grid_sizes(const double* box_host, int64_t nloc, float rcut, int64_t* nlc, int64_t* next) {
  if (box_host) {
    GridParams g = pbc_grid(box_host, rcut, (int)nloc);
    *nlc = (int64_t)g.ncell[0] * g.ncell[1] * g.ncell[2];
    *next = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
  } else {
    *nlc = (int64_t)kMaxGridNoPbc * kMaxGridNoPbc * kMaxGridNoPbc;
    *next = *nlc;
  }
  return *nlc > 0 && *next < (int64_t(1) << 31);
}

// ---------------------------------------------------------------- device helpers
__device__ inline void to_inter(double* ri, const double* rec, const double* rp) {
  ri[0] = rp[0] * rec[0] + rp[1] * rec[1] + rp[2] * rec[2];
  ri[1] = rp[0] * rec[3] + rp[1] * rec[4] + rp[2] * rec[5];
  ri[2] = rp[0] * rec[6] + rp[1] * rec[7] + rp[2] * rec[8];
}

__device__ inline void to_phys(double* rp, const double* b, const double* ri) {
  rp[0] = ri[0] * b[0] + ri[1] * b[3] + ri[2] * b[6];
  rp[1] = ri[0] * b[1] + ri[1] * b[4] + ri[2] * b[7];
  rp[2] = ri[0] * b[2] + ri[1] * b[5] + ri[2] * b[8];
}


__global__ void k_fill_int(int* p, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_write_params(GridParams p, GridParams* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = p;
}

// Sets error bit 2 of misc[1] unless the workspace header was written by
// dpmi_copy_coord_f64 for this nloc and nall and its layout fits in
// ws_bytes; the search kernels then leave every output untouched.
__global__ void k_ws_validate(char* ws, int nloc, int nall, long long ws_bytes) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const GridParams& p = *reinterpret_cast<const GridParams*>(ws);
  int* misc = reinterpret_cast<int*>(ws + ws_layout(0, 0, 0).misc);
  bool ok = p.magic == kGridMagic && p.nloc == nloc && (nall < 0 || misc[0] == nall) && misc[0] >= nloc &&
            misc[0] < (1 << 24) && p.lay_nlc > 0 && p.lay_next > 0 && p.lay_nlc < (1ll << 31) &&
            p.lay_next < (1ll << 31);
  if (ok) ok = ws_layout(p.nloc, p.lay_nlc, p.lay_next).total <= ws_bytes;
  if (ok) ok = misc[4] == 0;  // the ghost copy ran out of room: nall indexes past the arrays
  if (!ok) atomicOr(&misc[1], 2);
}

// After a rejected search (status bit 2) the outputs were never written:
// fill them with -1 (no neighbors), so that every later kernel reads valid
// indices; its forces are then NaN (dpmi_force_virial_cells_*).
__global__ void k_poison_rows(int* __restrict__ nlist, int* __restrict__ nlist_mapped, int16_t* __restrict__ cand_slot,
                              int* __restrict__ overflow, int64_t nslots, int nloc, const char* __restrict__ ws) {
  const int* misc = reinterpret_cast<const int*>(ws + ws_layout(0, 0, 0).misc);
  if ((misc[1] & 2) == 0) return;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nslots; k += (int64_t)gridDim.x * blockDim.x) {
    nlist[k] = -1;
    if (nlist_mapped) nlist_mapped[k] = -1;
    if (cand_slot) cand_slot[k] = -1;
    if (overflow && k < nloc) overflow[k] = -1;
  }
}

// Bounding box of an open system (one workgroup).
__global__ void __launch_bounds__(1024) k_bbox(const double* c, int n, double* out) {
  __shared__ double red[6][1024 / kWave];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    for (int d = 0; d < 3; ++d) {
      double v = c[3 * i + d];
      lo[d] = fmin(lo[d], v);
      hi[d] = fmax(hi[d], v);
    }
  for (int off = 32; off > 0; off >>= 1)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, kWave));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, kWave));
    }
  int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  if (lane_id() == 0)
    for (int d = 0; d < 3; ++d) {
      red[d][w] = lo[d];
      red[3 + d][w] = hi[d];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) {
      double a = red[d][0], b = red[3 + d][0];
      for (int k = 1; k < nw; ++k) {
        a = fmin(a, red[d][k]);
        b = fmax(b, red[3 + d][k]);
      }
      out[d] = a;
      out[3 + d] = b;
    }
  }
}

// Step 1: normalize_coord_cpu (coord.cc:13-28) + build_clist cell index
// (neighbor_list.cc:61-111 with nat_stt = ext_stt = 0, grid = ncell).
__global__ void k_normalize_bin(const double* __restrict__ in_c, const int* __restrict__ in_t,
                                double* __restrict__ out_c, int* __restrict__ out_t, int* __restrict__ mapping,
                                int* __restrict__ cell_of, int* __restrict__ cell_count, GridParams p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.nloc) return;
  double x[3] = {in_c[3 * i + 0], in_c[3 * i + 1], in_c[3 * i + 2]};
  int idx[3];
  if (p.pbc) {
    double ri[3];
    to_inter(ri, p.rec, x);
    for (int d = 0; d < 3; ++d) {
      ri[d] = fmod(ri[d], 1.);
      if (ri[d] < 0.) ri[d] += 1.;
    }
    to_phys(x, p.box, ri);
    double inter[3];
    to_inter(inter, p.rec, x);
    for (int d = 0; d < 3; ++d) {
      double cell_size = 1. / p.ncell[d];
      idx[d] = (int)((inter[d] - 0.) / cell_size);
      if (inter[d] - 0. < 0.) idx[d]--;
      if (idx[d] < 0) idx[d] = 0;
      else if (idx[d] >= p.ncell[d]) idx[d] = p.ncell[d] - 1;
    }
  } else {
    for (int d = 0; d < 3; ++d) {
      idx[d] = (int)floor((x[d] - p.lo[d]) / p.cs[d]);
      if (idx[d] < 0) idx[d] = 0;
      else if (idx[d] >= p.ncell[d]) idx[d] = p.ncell[d] - 1;
    }
  }
  int c = (idx[0] * p.ncell[1] + idx[1]) * p.ncell[2] + idx[2];
  out_c[3 * i + 0] = x[0];
  out_c[3 * i + 1] = x[1];
  out_c[3 * i + 2] = x[2];
  out_t[i] = in_t[i];
  mapping[i] = i;
  cell_of[i] = c;
  atomicAdd(&cell_count[c], 1);
}

// Block-wide exclusive scan helper (blockDim = 1024 = 16 waves).
__device__ int block_exclusive_scan(int v, int* total, int* smem) {
  int lane = lane_id(), w = threadIdx.x / kWave;
  int incl = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    int t = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += t;
  }
  if (lane == kWave - 1) smem[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int k = 0; k < (int)(blockDim.x / kWave); ++k) {
      int t = smem[k];
      smem[k] = run;
      run += t;
    }
    smem[16] = run;
  }
  __syncthreads();
  int excl = smem[w] + incl - v;
  *total = smem[16];
  __syncthreads();
  return excl;
}

// Step 2: prefix sums of local cell counts (cell_start) and of ghost counts in
// the ii[0], ii[1], ii[2] loop order of copy_coord (ghost_off); misc[0] = nall.
__global__ void __launch_bounds__(1024) k_scan(const int* __restrict__ cell_count, int* __restrict__ cell_start,
                                               int* __restrict__ ghost_off, int* __restrict__ misc, GridParams p) {
  __shared__ int smem[32];
  int nlc = p.ncell[0] * p.ncell[1] * p.ncell[2];
  int carry = 0;
  for (int base = 0; base < nlc; base += blockDim.x) {
    int c = base + threadIdx.x;
    int v = c < nlc ? cell_count[c] : 0;
    int tot;
    int ex = block_exclusive_scan(v, &tot, smem);
    if (c < nlc) cell_start[c] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) cell_start[nlc] = carry;
  int next = p.ext[0] * p.ext[1] * p.ext[2];
  int gcarry = 0;
  for (int base = 0; base < next; base += blockDim.x) {
    int e = base + threadIdx.x;
    int v = 0;
    if (e < next && p.pbc) {
      int e2 = e % p.ext[2], e1 = (e / p.ext[2]) % p.ext[1], e0 = e / (p.ext[1] * p.ext[2]);
      int ii[3] = {e0 - p.ngcell[0], e1 - p.ngcell[1], e2 - p.ngcell[2]};
      bool local = true;
      int jj[3];
      for (int d = 0; d < 3; ++d) {
        if (ii[d] < 0 || ii[d] >= p.ncell[d]) local = false;
        jj[d] = ii[d] + pbc_shift(ii[d], p.ncell[d]) * p.ncell[d];
      }
      if (!local) v = cell_count[(jj[0] * p.ncell[1] + jj[1]) * p.ncell[2] + jj[2]];
    }
    int tot;
    int ex = block_exclusive_scan(v, &tot, smem);
    if (e < next) ghost_off[e] = gcarry + ex;
    gcarry += tot;
  }
  if (threadIdx.x == 0) {
    ghost_off[next] = gcarry;
    misc[0] = p.nloc + gcarry;
  }
}

// Step 3a: unordered bucket fill.
__global__ void k_cell_fill(const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                            int* __restrict__ cursor, int* __restrict__ unordered, int nloc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nloc) return;
  int c = cell_of[i];
  int slot = atomicAdd(&cursor[c], 1);
  unordered[cell_start[c] + slot] = i;
}

// Step 3b: stable order inside each cell (index order, as build_clist pushes).
// pos_of is the inverse of cell_atoms (position of each atom in it).
__global__ void k_cell_order(const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                             const int* __restrict__ unordered, int* __restrict__ cell_atoms,
                             int* __restrict__ pos_of, int nloc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nloc) return;
  int c = cell_of[i];
  int b = cell_start[c], e = cell_start[c + 1];
  int rank = 0;
  for (int k = b; k < e; ++k) rank += unordered[k] < i;
  cell_atoms[b + rank] = i;
  pos_of[i] = b + rank;
}

// Step 4: ghost images, one wavefront per extended cell (copy_coord,
// neighbor_list.cc:796-824: coordinate minus inter2Phys(shift)).
__global__ void __launch_bounds__(64) k_ghosts(double* __restrict__ out_c, int* __restrict__ out_t,
                                               int* __restrict__ mapping, const int* __restrict__ cell_start,
                                               const int* __restrict__ cell_atoms, const int* __restrict__ ghost_off,
                                               GridParams p, int64_t mem_nall, int* __restrict__ misc) {
  int e = blockIdx.x;
  int e2 = e % p.ext[2], e1 = (e / p.ext[2]) % p.ext[1], e0 = e / (p.ext[1] * p.ext[2]);
  int ii[3] = {e0 - p.ngcell[0], e1 - p.ngcell[1], e2 - p.ngcell[2]};
  bool local = true;
  int jj[3];
  double sd[3];
  for (int d = 0; d < 3; ++d) {
    if (ii[d] < 0 || ii[d] >= p.ncell[d]) local = false;
    int s = pbc_shift(ii[d], p.ncell[d]);
    sd[d] = (double)s;
    jj[d] = ii[d] + s * p.ncell[d];
  }
  if (local) return;
  double sv[3];
  to_phys(sv, p.box, sd);
  int c = (jj[0] * p.ncell[1] + jj[1]) * p.ncell[2] + jj[2];
  int b = cell_start[c], cnt = cell_start[c + 1] - b;
  int64_t base = (int64_t)p.nloc + ghost_off[e];
  if (base + cnt > mem_nall) {  // more ghosts than the output holds (the host sized it for fewer)
    if (threadIdx.x == 0) atomicOr(&misc[4], 4);
    cnt = (int)max((int64_t)0, mem_nall - base);
  }
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
    int a = cell_atoms[b + k];
    int64_t q = base + k;
    out_c[3 * q + 0] = out_c[3 * a + 0] - sv[0];
    out_c[3 * q + 1] = out_c[3 * a + 1] - sv[1];
    out_c[3 * q + 2] = out_c[3 * a + 2] - sv[2];
    out_t[q] = out_t[a];
    mapping[q] = a;
  }
}

// Step 5: candidates + format, one wavefront per local atom.
// key = type << 56 | float-bits(r^2) << 24 | extended index (nall < 2^24).
template <int CAP>
__global__ void __launch_bounds__(64) k_search_format(
    int* __restrict__ nlist, int* __restrict__ overflow, int* __restrict__ nlist_mapped,
    const double* __restrict__ ext_c, const int* __restrict__ ext_t, const int* __restrict__ mapping,
    char* __restrict__ ws, const int* __restrict__ atoms, int16_t* __restrict__ cand_slot, SelParams sp,
    float rcut, const int* __restrict__ count_dev) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ SelLds s_sl;
  stage_sel(sp, &s_sl);
  __syncthreads();
  uint64_t* keys = reinterpret_cast<uint64_t*>(smem_raw);
  int* row = reinterpret_cast<int*>(keys + CAP);
  const GridParams p = *reinterpret_cast<const GridParams*>(ws);
  const WsLayout L = ws_layout(p.nloc, p.lay_nlc, p.lay_next);
  const int* __restrict__ cell_of = reinterpret_cast<const int*>(ws + L.cell_of);
  const int* __restrict__ cell_start = reinterpret_cast<const int*>(ws + L.cell_start);
  const int* __restrict__ cell_atoms = reinterpret_cast<const int*>(ws + L.cell_atoms);
  const int* __restrict__ ghost_off = reinterpret_cast<const int*>(ws + L.ghost_off);
  int* misc = reinterpret_cast<int*>(ws + L.misc);
  if (misc[1] & 2) return;  // workspace rejected by k_ws_validate
  // one atom per workgroup, or (count_dev) a fixed grid striding over the
  // count the cell-wise pass left in device memory
  const int nwork = count_dev ? *count_dev : (int)gridDim.x;
  for (int wk = blockIdx.x; wk < nwork; wk += gridDim.x) {
  const int i = atoms ? atoms[wk] : wk;
  const int lane = lane_id();
  const double rcut2d = (double)(rcut * rcut);
  const float rcut2f = rcut * rcut;
  const double xi0 = ext_c[3 * i + 0], xi1 = ext_c[3 * i + 1], xi2 = ext_c[3 * i + 2];
  const float fi0 = (float)xi0, fi1 = (float)xi1, fi2 = (float)xi2;
  int c = cell_of[i];
  int c2 = c % p.ncell[2], c1 = (c / p.ncell[2]) % p.ncell[1], c0 = c / (p.ncell[1] * p.ncell[2]);
  int n = 0;  // wave-uniform candidate count
  const int ns = p.nsearch;
  for (int d0 = -ns; d0 <= ns; ++d0) {
    for (int d1 = -ns; d1 <= ns; ++d1) {
      for (int d2 = -ns; d2 <= ns; ++d2) {
        int ii0 = c0 + d0, ii1 = c1 + d1, ii2 = c2 + d2;
        bool local;
        int b, cnt;
        int64_t gbase = 0;
        if (p.pbc) {
          if (ii0 < -p.ngcell[0] || ii0 >= p.ncell[0] + p.ngcell[0] || ii1 < -p.ngcell[1] ||
              ii1 >= p.ncell[1] + p.ngcell[1] || ii2 < -p.ngcell[2] || ii2 >= p.ncell[2] + p.ngcell[2])
            continue;
          local = ii0 >= 0 && ii0 < p.ncell[0] && ii1 >= 0 && ii1 < p.ncell[1] && ii2 >= 0 && ii2 < p.ncell[2];
          int j0 = ii0 + pbc_shift(ii0, p.ncell[0]) * p.ncell[0];
          int j1 = ii1 + pbc_shift(ii1, p.ncell[1]) * p.ncell[1];
          int j2 = ii2 + pbc_shift(ii2, p.ncell[2]) * p.ncell[2];
          int cc = (j0 * p.ncell[1] + j1) * p.ncell[2] + j2;
          b = cell_start[cc];
          cnt = cell_start[cc + 1] - b;
          if (!local) {
            int e = ((ii0 + p.ngcell[0]) * p.ext[1] + (ii1 + p.ngcell[1])) * p.ext[2] + (ii2 + p.ngcell[2]);
            gbase = (int64_t)p.nloc + ghost_off[e];
          }
        } else {
          if (ii0 < 0 || ii0 >= p.ncell[0] || ii1 < 0 || ii1 >= p.ncell[1] || ii2 < 0 || ii2 >= p.ncell[2])
            continue;
          local = true;
          int cc = (ii0 * p.ncell[1] + ii1) * p.ncell[2] + ii2;
          b = cell_start[cc];
          cnt = cell_start[cc + 1] - b;
        }
        for (int k0 = 0; k0 < cnt; k0 += kWave) {
          int k = k0 + lane;
          bool pass = false;
          uint64_t key = 0;
          if (k < cnt) {
            int j = local ? cell_atoms[b + k] : (int)(gbase + k);
            if (j != i) {
              double xj0 = ext_c[3 * j + 0], xj1 = ext_c[3 * j + 1], xj2 = ext_c[3 * j + 2];
              double e0 = xi0 - xj0, e1 = xi1 - xj1, e2 = xi2 - xj2;
              double r2 = e0 * e0 + e1 * e1 + e2 * e2;
              if (r2 < rcut2d) {
                int tj = ext_t[j];
                float f0 = (float)xj0 - fi0, f1 = (float)xj1 - fi1, f2 = (float)xj2 - fi2;
                float rr2 = f0 * f0 + f1 * f1 + f2 * f2;
                // types outside [0, ntypes) are not neighbors of any section
                if (tj >= 0 && tj < sp.ntypes && rr2 <= rcut2f) {
                  pass = true;
                  key = ((uint64_t)tj << 56) | ((uint64_t)__float_as_uint(rr2) << 24) | (uint64_t)j;
                }
              }
            }
          }
          uint64_t m = __ballot(pass);
          int pos = n + mask_prefix(m);
          if (pass && pos < CAP) keys[pos] = key;
          n += __popcll(m);
        }
      }
    }
  }
  if (n > CAP) {
    if (lane == 0) atomicOr(&misc[1], 1);
    n = CAP;
  }
  for (int s = lane; s < sp.nnei; s += kWave) row[s] = -1;
  __syncthreads();
  int ovf = -1;
  for (int k = lane; k < n; k += kWave) {
    uint64_t kk = keys[k];
    int t = (int)(kk >> 56);
    uint64_t tbase = (uint64_t)t << 56;
    uint64_t rel = kk - tbase;
    int rank = 0;
    for (int m = 0; m < n; ++m) rank += (keys[m] - tbase) < rel;
    const int sel_t = s_sl.sel[t], sec_t = s_sl.sec[t];
    
    if (rank < sel_t)
      row[sec_t + rank] = (int)(kk & 0xFFFFFFull);
    else
      ovf = max(ovf, t);
  }
  __syncthreads();
  int64_t ro = (int64_t)i * sp.nnei;
  for (int s = lane; s < sp.nnei; s += kWave) {
    int j = row[s];
    nlist[ro + s] = j;
    if (nlist_mapped) nlist_mapped[ro + s] = j >= 0 ? mapping[j] : -1;
    if (cand_slot) cand_slot[ro + s] = -1;  // not in a cell's candidate table
  }
  for (int off = 32; off > 0; off >>= 1) ovf = max(ovf, __shfl_xor(ovf, off, kWave));
  if (overflow && lane == 0) overflow[i] = ovf;
  __syncthreads();
  }
}

// Step 5, cell-wise: one workgroup per local cell (grid-stride over cells).
// The candidate atoms of the cell's (2 nsearch + 1)^3 neighborhood are
// gathered into LDS once and every atom of the cell tests them from there
// (one wavefront per atom), instead of each atom re-walking the cell lists.
// Same candidate tests, keys and ranking as k_search_format; an atom with
// more than kKeyCap candidates, or a cell whose neighborhood exceeds
// kCandCap atoms, is appended to the redo list for k_search_format.
constexpr int kKeyCap = 512;  // plus 8 padding slots per wavefront
// three or four types: each type's 64-bit keys in a segment of its own
constexpr int kSegTypes = 4;
constexpr int kSeg4 = kKeyCap / kSegTypes;
constexpr int kSeg4Cap = kSeg4 - 8;  // room for the padding in each segment
constexpr size_t kMaxCellsLds = 150 * 1024;  // dynamic LDS the cell-wise search may take
constexpr int kBadType = 127;                // packed type code of a candidate that never passes

// Bound on |x| of any extended coordinate: the box grown by the ghost cells
// (periodic), or the grid of an open system.
__device__ inline double coord_bound(const GridParams& p) {
  double xm = 0.;
  for (int d = 0; d < 3; ++d) {
    double x = 0.;
    if (p.pbc) {
      for (int i = 0; i < 3; ++i) x += (1.0 + (double)p.ngcell[i] / p.ncell[i]) * fabs(p.box[3 * i + d]);
    } else {
      x = fmax(fabs(p.lo[d]), fabs(p.lo[d] + p.ncell[d] * p.cs[d]));
    }
    xm = fmax(xm, x);
  }
  return xm * 1.001 + 1e-30;
}

// Width of the band below rc^2 in which the float r^2 of float-rounded
// coordinates cannot decide the double test |xi - xj|^2 < rc^2.  Each
// float coordinate is within ulp(X)/2 of the double one and the float
// difference adds at most ulp(rc)/2, so every component differs by at most
// dl = ulp(X) + ulp(rc)/2; then |rr2 - r2| <= 2 sqrt(3) rc dl + 3 dl^2 plus
// the rounding of the float squares and sums (<= 3 ulp(rc^2)).  Twice that.
__device__ inline double screen_margin(double xmax, float rcut) {
  const double ux = ldexp(1.0, ilogb(xmax) - 23), urc = ldexp(1.0, ilogb((double)rcut) - 23);
  const double url = ldexp(1.0, ilogb((double)rcut * rcut) - 23);
  const double dl = ux + 0.5 * urc;
  return 2.0 * (3.5 * (double)rcut * dl + 3.0 * dl * dl + 3.0 * url);
}

// wavefronts per workgroup of the cell-wise search: eight share the candidate
// table of the first pass (two workgroups per CU), four the larger table of
// the second
constexpr int search_waves(bool big) { return big ? 4 : 8; }

template <int CAP, bool kBigPass, bool kSeg>
__global__ void __launch_bounds__(64 * search_waves(kBigPass)) k_search_cells(
    int* __restrict__ nlist, int* __restrict__ overflow, int* __restrict__ nlist_mapped,
    int16_t* __restrict__ cand_slot, const double* __restrict__ ext_c, const int* __restrict__ ext_t,
    const int* __restrict__ mapping, char* __restrict__ ws, SelParams sp, float rcut) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ SelLds s_sl;
  stage_sel(sp, &s_sl);
  __syncthreads();
  // candidate g: float coordinates and (type code << 24 | extended index)
  float4* cand = reinterpret_cast<float4*>(smem_raw);
  uint64_t* keys_all = reinterpret_cast<uint64_t*>(cand + CAP);
  constexpr int kW = search_waves(kBigPass);
  int* rows_all = reinterpret_cast<int*>(keys_all + kW * (kKeyCap + 8));
  int16_t* kslot_all = reinterpret_cast<int16_t*>(rows_all + kW * sp.nnei);
  int16_t* rslot_all = kslot_all + kW * (kKeyCap + 8);
  __shared__ int s_base[kMaxNb], s_local[kMaxNb], s_pref[kMaxNb + 1];

  const GridParams p = *reinterpret_cast<const GridParams*>(ws);
  const WsLayout L = ws_layout(p.nloc, p.lay_nlc, p.lay_next);
  const int* __restrict__ cell_start = reinterpret_cast<const int*>(ws + L.cell_start);
  const int* __restrict__ cell_atoms = reinterpret_cast<const int*>(ws + L.cell_atoms);
  const int* __restrict__ ghost_off = reinterpret_cast<const int*>(ws + L.ghost_off);
  int* misc = reinterpret_cast<int*>(ws + L.misc);
  int* redo = reinterpret_cast<int*>(ws + L.redo);
  int* bigcells = reinterpret_cast<int*>(ws + L.bigcells);
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const double rcut2d = (double)(rcut * rcut);
  const float rcut2f = rcut * rcut;
  const double rsure = rcut2d - screen_margin(coord_bound(p), rcut);
  const int ns = p.nsearch, side = 2 * ns + 1, nnb = side * side * side;
  const int ncells = kBigPass ? misc[3] : p.ncell[0] * p.ncell[1] * p.ncell[2];
  uint64_t* keys = keys_all + w * (kKeyCap + 8);
  // Two types: the key region holds instead 32-bit keys (type << 31 | float
  // bits of r^2), which order the candidates like the 64-bit keys except for
  // exact r^2 ties, followed by the candidates' extended indices.
  const bool short_keys = sp.ntypes <= 2;
  // (the host picks kSeg for three or four types)
  constexpr bool seg_keys = kSeg;
  uint32_t* k32 = reinterpret_cast<uint32_t*>(keys);
  int* kj = reinterpret_cast<int*>(k32 + (kKeyCap + 8));
  int* row = rows_all + w * sp.nnei;
  int16_t* kslot = kslot_all + w * (kKeyCap + 8);
  int16_t* rslot = rslot_all + w * sp.nnei;

  if (misc[1] & 2) return;  // workspace rejected by k_ws_validate
  // Fewer cells than workgroups (small boxes): nsplit workgroups share each
  // cell, each building the same candidate table and taking every
  // nsplit-th group of kW of its atoms; large boxes keep one per cell.
  const int nsplit = ncells > 0 ? max(1, min(kMaxCellSplit, (int)gridDim.x / ncells)) : 1;
  for (int wi = blockIdx.x; wi < ncells * nsplit; wi += gridDim.x) {
    const int ci = wi % ncells, part = wi / ncells;
    const int cc = kBigPass ? bigcells[ci] : ci;
    const int a0 = cell_start[cc], na = cell_start[cc + 1] - a0;
    if (na <= kW * part) continue;  // no atom for this part (uniform in the workgroup)
    const int T = nb_table(p, cell_start, ghost_off, cc, s_base, s_local, s_pref);
    if (T > CAP) {
      if (part == 0) {
        if (!kBigPass) {
          if (tid == 0) bigcells[atomicAdd(&misc[3], 1)] = cc;  // second pass, larger table
        } else {
          for (int k = tid; k < na; k += blockDim.x) redo[atomicAdd(&misc[2], 1)] = cell_atoms[a0 + k];
        }
      }
      __syncthreads();
      continue;
    }
    for (int g = tid; g < T; g += blockDim.x) {
      int t;
      const int j = nb_candidate(g, nnb, s_base, s_local, s_pref, cell_atoms, &t);
      const int tj = ext_t[j];
      const int code = (tj >= 0 && tj < sp.ntypes) ? tj : kBadType;
      cand[g] = make_float4((float)ext_c[3 * j + 0], (float)ext_c[3 * j + 1], (float)ext_c[3 * j + 2],
                            __int_as_float((code << 24) | j));
    }
    __syncthreads();
    // the centre cell's members are the table's run at s_pref[nnb / 2], in
    // cell order: an atom's index and float coordinates come from LDS (its
    // double coordinates only where the float test cannot decide)
    const int ctr = s_pref[nnb / 2];
    for (int ai = kW * part + w; ai < na; ai += kW * nsplit) {
      const float4 me = cand[ctr + ai];
      const int i = __float_as_int(me.w) & 0xFFFFFF;
      const float fi0 = me.x, fi1 = me.y, fi2 = me.z;
      int n = 0, n0 = 0;
      int ns[kSegTypes] = {0, 0, 0, 0};  // keys per type segment (seg_keys)
      for (int g0 = 0; g0 < T; g0 += kWave) {
        const int g = g0 + lane;
        bool pass = false;
        uint64_t key = 0;
        uint32_t key32 = 0;
        int jpass = 0;
        if (g < T) {
          const float4 c = cand[g];
          const int packed = __float_as_int(c.w);
          const int j = packed & 0xFFFFFF, tj = packed >> 24;
          // the float test of format_nlist_i_cpu (exact: the same float
          // coordinates), then the double test of build_nlist_cpu, which the
          // float r^2 decides outside a narrow band below rc^2
          const float f0 = c.x - fi0, f1 = c.y - fi1, f2 = c.z - fi2;
          const float rr2 = f0 * f0 + f1 * f1 + f2 * f2;
          if (j != i && tj != kBadType && rr2 <= rcut2f) {
            bool in = (double)rr2 < rsure;
            if (!in) {
              const double e0 = ext_c[3 * i + 0] - ext_c[3 * j + 0], e1 = ext_c[3 * i + 1] - ext_c[3 * j + 1],
                           e2 = ext_c[3 * i + 2] - ext_c[3 * j + 2];
              in = e0 * e0 + e1 * e1 + e2 * e2 < rcut2d;
            }
            if (in) {
              pass = true;
              key = ((uint64_t)tj << 56) | ((uint64_t)__float_as_uint(rr2) << 24) | (uint64_t)j;
              key32 = ((uint32_t)tj << 31) | __float_as_uint(rr2);
              jpass = j;
            }
          }
        }
        const uint64_t m = __ballot(pass);
        if (seg_keys) {
          const int tp = (int)(key >> 56);
          int r = 0;
#pragma unroll
          for (int t = 0; t < kSegTypes; ++t) {
            const uint64_t mt = __ballot(pass && tp == t);
            if (tp == t) r = ns[t] + mask_prefix(mt);
            ns[t] += __popcll(mt);
          }
          if (pass && r < kSeg4Cap) {
            keys[kSeg4 * tp + r] = key;
            kslot[kSeg4 * tp + r] = (int16_t)g;
          }
        }
        const int pos = n + mask_prefix(m);
        if (!seg_keys && pass && pos < kKeyCap) {
          if (short_keys) {
            k32[pos] = key32;
            kj[pos] = jpass;
          } else {
            keys[pos] = key;
          }
          kslot[pos] = (int16_t)g;
        }
        n += __popcll(m);
        n0 += __popcll(__ballot(pass && (key >> 56) == 0));
      }
      int nsmax = 0;
#pragma unroll
      for (int t = 0; t < kSegTypes; ++t) nsmax = max(nsmax, ns[t]);
      // wave-uniform: leave this atom to the per-atom kernel
      if (seg_keys ? nsmax > kSeg4Cap : n > kKeyCap) {
        if (lane == 0) redo[atomicAdd(&misc[2], 1)] = i;
        continue;
      }
      for (int s = lane; s < sp.nnei; s += kWave) {
        row[s] = -1;
        rslot[s] = -1;
      }
      // rank of each key within its type: two keys per lane per pass over the
      // key list, read 8 at a time (independent LDS reads)
      const bool seg_fast = seg_keys && nsmax <= kWave;
      if (seg_keys) {
        // each segment padded to a multiple of 8 (all of them to the longest
        // when every type fits one lane per key)
#pragma unroll
        for (int t = 0; t < kSegTypes; ++t) {
          const int e = seg_fast ? (nsmax + 7) & ~7 : (ns[t] + 7) & ~7;
          if (t < sp.ntypes)
            for (int k = ns[t] + lane; k < e; k += kWave) keys[kSeg4 * t + k] = ~0ull;
        }
      } else if (n & 7) {  // pad to a multiple of 8 with keys that rank after everything
        if (lane < 8 - (n & 7)) {
          if (short_keys) {
            k32[n + lane] = ~0u;
            kj[n + lane] = 0xFFFFFF;
          } else {
            keys[n + lane] = ~0ull;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const int n8 = (n + 7) & ~7;
      int ovf = -1;
      // Two types, at most 128 keys: rank by the 32-bit keys, counting the
      // keys <= each key (itself included; the type bit puts type-0 keys
      // first, so a type-1 key subtracts their count n0).  Without ties these
      // counts are 1..n_t within each type, which their sum certifies; with
      // a tie it exceeds sum n_t (n_t + 1) / 2 and the 64-bit ranking below
      // decides (index order among equal r^2).
      bool ranked = false;
      if (short_keys && n <= 2 * kWave) {
        const int ka = lane, kb = kWave + lane;
        const uint32_t a32 = ka < n ? k32[ka] : ~0u, b32 = kb < n ? k32[kb] : ~0u;
        int ra = 0, rb = 0;
        for (int q = 0; q < n8; q += 8) {
          uint32_t kq[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) kq[u] = k32[q + u];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            ra += kq[u] <= a32;
            rb += kq[u] <= b32;
          }
        }
        ra -= (a32 >> 31) ? n0 : 0;
        rb -= (b32 >> 31) ? n0 : 0;
        int sum = (ka < n ? ra : 0) + (kb < n ? rb : 0);
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, kWave);
        const int n1 = n - n0;
        if (sum == n0 * (n0 + 1) / 2 + n1 * (n1 + 1) / 2) {
          ranked = true;
          if (ka < n) {
            const int t = (int)(a32 >> 31), r = ra - 1;
            const int sel_t = s_sl.sel[t], sec_t = s_sl.sec[t];
            
            if (r < sel_t) {
              row[sec_t + r] = kj[ka];
              rslot[sec_t + r] = kslot[ka];
            } else {
              ovf = max(ovf, t);
            }
          }
          if (kb < n) {
            const int t = (int)(b32 >> 31), r = rb - 1;
            const int sel_t = s_sl.sel[t], sec_t = s_sl.sec[t];
            
            if (r < sel_t) {
              row[sec_t + r] = kj[kb];
              rslot[sec_t + r] = kslot[kb];
            } else {
              ovf = max(ovf, t);
            }
          }
        }
      }
      // the 64-bit key of entry q (rebuilt from the short layout for two types)
      auto key_at = [&](int q, auto shortk) -> uint64_t {
        if constexpr (decltype(shortk)::value) {
          const uint32_t k = k32[q];
          return k == ~0u ? ~0ull
                          : ((uint64_t)(k >> 31) << 56) | ((uint64_t)(k & 0x7FFFFFFFu) << 24) | (uint64_t)kj[q];
        } else {
          return keys[q];
        }
      };
      // keys base .. base + cnt - 1, the list padded to a multiple of 8
      auto rank64 = [&](auto shortk, int base, int cnt) {
      const int cnt8 = (cnt + 7) & ~7;
      for (int k0 = 0; k0 < cnt; k0 += 2 * kWave) {
        const int ka = k0 + lane, kb = k0 + kWave + lane;
        const uint64_t kka = ka < cnt ? key_at(base + ka, shortk) : ~0ull,
                       kkb = kb < cnt ? key_at(base + kb, shortk) : ~0ull;
        const uint64_t ta = kka & (0xFFull << 56), tb = kkb & (0xFFull << 56);
        const uint64_t rela = kka - ta, relb = kkb - tb;
        int ra = 0, rb = 0;
        for (int q = 0; q < cnt8; q += 8) {
          uint64_t kq[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) kq[u] = key_at(base + q + u, shortk);
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            ra += (kq[u] - ta) < rela;
            rb += (kq[u] - tb) < relb;
          }
        }
        if (ka < cnt) {
          const int t = (int)(kka >> 56);
          const int sel_t = s_sl.sel[t], sec_t = s_sl.sec[t];
          if (ra < sel_t) {
            row[sec_t + ra] = (int)(kka & 0xFFFFFFull);
            rslot[sec_t + ra] = kslot[base + ka];
          } else {
            ovf = max(ovf, t);
          }
        }
        if (kb < cnt) {
          const int t = (int)(kkb >> 56);
          const int sel_t = s_sl.sel[t], sec_t = s_sl.sec[t];
          if (rb < sel_t) {
            row[sec_t + rb] = (int)(kkb & 0xFFFFFFull);
            rslot[sec_t + rb] = kslot[base + kb];
          } else {
            ovf = max(ovf, t);
          }
        }
      }
      };
      // three or four types: each key against its own segment only; lane l
      // ranks key l of every segment when each fits one lane per key (the
      // rank is the number of smaller keys: keys of one type differ in j)
      if (seg_fast) {
        ranked = true;
        const int q8 = (nsmax + 7) & ~7;
        uint64_t kt[kSegTypes];
        int rk[kSegTypes];
#pragma unroll
        for (int t = 0; t < kSegTypes; ++t) {
          kt[t] = t < sp.ntypes && lane < ns[t] ? keys[kSeg4 * t + lane] : ~0ull;
          rk[t] = 0;
        }
        for (int q = 0; q < q8; q += 8) {
#pragma unroll
          for (int t = 0; t < kSegTypes; ++t) {
            if (t >= sp.ntypes) break;  // uniform
            uint64_t kq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) kq[u] = keys[kSeg4 * t + q + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) rk[t] += kq[u] < kt[t];
          }
        }
#pragma unroll
        for (int t = 0; t < kSegTypes; ++t) {
          if (t < sp.ntypes && lane < ns[t]) {
            const int sel_t = s_sl.sel[t], sec_t = s_sl.sec[t];
            if (rk[t] < sel_t) {
              row[sec_t + rk[t]] = (int)(kt[t] & 0xFFFFFFull);
              rslot[sec_t + rk[t]] = kslot[kSeg4 * t + lane];
            } else {
              ovf = max(ovf, t);
            }
          }
        }
      }
      if (!ranked) {
        if (short_keys) {
          rank64(std::true_type{}, 0, n);
        } else if (seg_keys) {
#pragma unroll
          for (int t = 0; t < kSegTypes; ++t)
            if (t < sp.ntypes) rank64(std::false_type{}, kSeg4 * t, ns[t]);
        } else {
          rank64(std::false_type{}, 0, n);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const int64_t ro = (int64_t)i * sp.nnei;
      for (int s = lane; s < sp.nnei; s += kWave) {
        const int j = row[s];
        nlist[ro + s] = j;
        if (nlist_mapped) nlist_mapped[ro + s] = j >= 0 ? mapping[j] : -1;
        if (cand_slot) cand_slot[ro + s] = rslot[s];
      }
      for (int off = 32; off > 0; off >>= 1) ovf = max(ovf, __shfl_xor(ovf, off, kWave));
      if (overflow && lane == 0) overflow[i] = ovf;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();  // the next cell overwrites the candidate arrays
  }
}

// format_nlist_cpu (source/lib/src/fmt_nlist.cc:145-177) of an external
// list: row i of cand (width ncand, extended indices) holds numneigh[i]
// candidates (or, without counts, entries until the row ends, -1 skipped).
// One wavefront per local atom keeps the candidates of type >= 0 (and <
// ntypes) with float r^2 <= rc^2 (format_nlist_i_cpu, fmt_nlist.cc:98-143)
// and ranks them by the same (type, float r^2, index) key as the search.
template <int CAP>
__global__ void __launch_bounds__(64) k_format_external(
    int* __restrict__ nlist, int* __restrict__ overflow, int* __restrict__ nlist_mapped, int* __restrict__ status,
    const int* __restrict__ cand, const int* __restrict__ numneigh, int64_t ncand, const double* __restrict__ ext_c,
    const int* __restrict__ ext_t, const int* __restrict__ mapping, int nall, SelParams sp, float rcut) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ SelLds s_sl;
  stage_sel(sp, &s_sl);
  __syncthreads();
  uint64_t* keys = reinterpret_cast<uint64_t*>(smem_raw);
  int* row = reinterpret_cast<int*>(keys + CAP);
  const int i = blockIdx.x;
  const int lane = lane_id();
  const float rcut2f = rcut * rcut;
  const float fi0 = (float)ext_c[3 * i + 0], fi1 = (float)ext_c[3 * i + 1], fi2 = (float)ext_c[3 * i + 2];
  const int64_t nn = numneigh ? (int64_t)numneigh[i] : ncand;
  const int64_t cnt = nn < ncand ? nn : ncand;
  const int* ci = cand + (int64_t)i * ncand;
  int n = 0;  // wave-uniform count of kept neighbors
  for (int64_t k0 = 0; k0 < cnt; k0 += kWave) {
    const int64_t k = k0 + lane;
    bool pass = false;
    uint64_t key = 0;
    if (k < cnt) {
      const int j = ci[k];
      if (j >= 0 && j < nall) {
        const int tj = ext_t[j];
        if (tj >= 0 && tj < sp.ntypes) {
          const float f0 = (float)ext_c[3 * j + 0] - fi0, f1 = (float)ext_c[3 * j + 1] - fi1,
                      f2 = (float)ext_c[3 * j + 2] - fi2;
          const float rr2 = f0 * f0 + f1 * f1 + f2 * f2;
          if (rr2 <= rcut2f) {
            pass = true;
            key = ((uint64_t)tj << 56) | ((uint64_t)__float_as_uint(rr2) << 24) | (uint64_t)j;
          }
        }
      }
    }
    const uint64_t m = __ballot(pass);
    const int pos = n + mask_prefix(m);
    if (pass && pos < CAP) keys[pos] = key;
    n += __popcll(m);
  }
  if (n > CAP) {
    if (lane == 0) atomicOr(status, 1);
    n = CAP;
  }
  for (int s = lane; s < sp.nnei; s += kWave) row[s] = -1;
  __syncthreads();
  int ovf = -1;
  for (int k = lane; k < n; k += kWave) {
    const uint64_t kk = keys[k];
    const int t = (int)(kk >> 56);
    const uint64_t tbase = (uint64_t)t << 56;
    const uint64_t rel = kk - tbase;
    int rank = 0;
    for (int m2 = 0; m2 < n; ++m2) rank += (keys[m2] - tbase) < rel;
    const int sel_t = s_sl.sel[t], sec_t = s_sl.sec[t];
    
    if (rank < sel_t)
      row[sec_t + rank] = (int)(kk & 0xFFFFFFull);
    else
      ovf = max(ovf, t);
  }
  __syncthreads();
  const int64_t ro = (int64_t)i * sp.nnei;
  for (int s = lane; s < sp.nnei; s += kWave) {
    const int j = row[s];
    nlist[ro + s] = j;
    if (nlist_mapped) nlist_mapped[ro + s] = j >= 0 ? mapping[j] : -1;
  }
  for (int off = 32; off > 0; off >>= 1) ovf = max(ovf, __shfl_xor(ovf, off, kWave));
  if (overflow && lane == 0) overflow[i] = ovf;
}

}  // namespace

// Dynamic LDS of the cell-wise search with a table of cap candidates.
size_t search_cells_lds(int cap, int nnei, int waves) {
  return (size_t)cap * sizeof(float4) +
         waves * ((size_t)(kKeyCap + 8) * (sizeof(uint64_t) + sizeof(int16_t)) + (sizeof(int) + sizeof(int16_t)) * nnei);
}

bool search_uses_cells(int nnei) {
  return search_cells_lds(kCandCapBig, nnei, search_waves(true)) <= kMaxCellsLds &&
         search_cells_lds(kCandCap, nnei, search_waves(false)) <= kMaxCellsLds;
}

int cell_grid_info(const double* box_host, int64_t nloc, float rcut, int64_t* nlc, int* nsearch) {
  int64_t next;
  DPMI_CHECK_ARG(rcut > 0.f, "rcut must be positive");
  DPMI_CHECK_ARG(grid_sizes(box_host, nloc, rcut, nlc, &next), "degenerate box");
  *nsearch = box_host ? pbc_grid(box_host, rcut, (int)nloc).nsearch : 1;
  return DPMI_OK;
}

int nlist_workspace(int64_t* bytes, const double* box_host, int64_t nloc, float rcut) {
  int64_t nlc, next;
  DPMI_CHECK_ARG(nloc >= 0, "nloc must be >= 0");
  DPMI_CHECK_ARG(rcut > 0.f, "rcut must be positive");
  DPMI_CHECK_ARG(grid_sizes(box_host, nloc, rcut, &nlc, &next), "degenerate box");
  *bytes = ws_layout(nloc, nlc, next).total;
  return DPMI_OK;
}

}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_nlist_workspace_bytes(int64_t* bytes, const double* box_host, int64_t nloc, float rcut) {
  DPMI_CHECK_ARG(bytes != nullptr, "bytes is NULL");
  return nlist_workspace(bytes, box_host, nloc, rcut);
}

namespace dpmi {

// Upper bound on nall for a periodic box: every local cell has at most
// prod_d ceil((ncell_d + 2 ngcell_d) / ncell_d) images among the extended
// cells (itself included), so nall <= nloc times that product.
int64_t copy_coord_capacity(const double* box_host, int64_t nloc, float rcut) {
  const GridParams g = pbc_grid(box_host, rcut, (int)nloc);
  int64_t m = 1;
  for (int d = 0; d < 3; ++d) m *= (g.ncell[d] + 2 * g.ngcell[d] + g.ncell[d] - 1) / g.ncell[d];
  return std::max<int64_t>(nloc * m, 1);
}

static int copy_coord_impl(double* out_coord, int* out_type, int* mapping, int64_t* nall, int64_t mem_nall,
                           const double* in_coord, const int* in_type, int64_t nloc, const double* box_host,
                           float rcut, void* workspace, int64_t workspace_bytes, void* stream, bool wait) {
  DPMI_CHECK_ARG(wait || box_host != nullptr, "the form without waits needs a periodic box");
  DPMI_CHECK_ARG(!wait || nall != nullptr, "nall is NULL");
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 24), "nloc out of range [0, 2^24)");
  DPMI_CHECK_ARG(mem_nall >= nloc, "mem_nall must be >= nloc");
  if (!wait) {
    const int64_t cap = copy_coord_capacity(box_host, nloc, rcut);
    DPMI_CHECK_ARG(mem_nall >= cap, "mem_nall below dpmi_copy_coord_capacity");
    DPMI_CHECK_ARG(cap < (int64_t(1) << 24), "the capacity bound reaches 2^24 extended atoms, beyond the 64-bit "
                                             "neighbor key: use dpmi_copy_coord_f64");
  }
  int64_t need;
  int rc = nlist_workspace(&need, box_host, nloc, rcut);
  if (rc) return rc;
  DPMI_CHECK_ARG(workspace != nullptr && workspace_bytes >= need, "workspace too small");
  hipStream_t st = as_stream(stream);
  int64_t nlc, next;
  grid_sizes(box_host, nloc, rcut, &nlc, &next);
  WsLayout L = ws_layout(nloc, nlc, next);
  char* ws = static_cast<char*>(workspace);
  int* cell_of = reinterpret_cast<int*>(ws + L.cell_of);
  int* cell_count = reinterpret_cast<int*>(ws + L.cell_count);
  int* cell_start = reinterpret_cast<int*>(ws + L.cell_start);
  int* cursor = reinterpret_cast<int*>(ws + L.cell_cursor);
  int* unordered = reinterpret_cast<int*>(ws + L.cell_unordered);
  int* cell_atoms = reinterpret_cast<int*>(ws + L.cell_atoms);
  int* pos_of = reinterpret_cast<int*>(ws + L.pos_of);
  int* ghost_off = reinterpret_cast<int*>(ws + L.ghost_off);
  int* misc = reinterpret_cast<int*>(ws + L.misc);
  GridParams* gp = reinterpret_cast<GridParams*>(ws + L.params);
  double* bbox = reinterpret_cast<double*>(ws + L.bbox);

  GridParams g;
  if (box_host) {
    g = pbc_grid(box_host, rcut, (int)nloc);
  } else {
    std::memset(&g, 0, sizeof(g));
    g.pbc = 0;
    g.nloc = (int)nloc;
    g.nsearch = 1;
    double h[6] = {0, 0, 0, 0, 0, 0};
    if (nloc > 0) {
      hipLaunchKernelGGL(k_bbox, dim3(1), dim3(1024), 0, st, in_coord, (int)nloc, bbox);
      DPMI_LAUNCH_CHECK();
      DPMI_HIP(hipMemcpyAsync(h, bbox, sizeof(h), hipMemcpyDeviceToHost, st));
      DPMI_HIP(hipStreamSynchronize(st));
    }
    for (int d = 0; d < 3; ++d) {
      double extent = h[3 + d] - h[d];
      int nc = (int)(extent / (double)rcut);
      nc = std::max(1, std::min(nc, kMaxGridNoPbc));
      g.ncell[d] = nc;
      g.ext[d] = nc;
      g.lo[d] = h[d];
      g.cs[d] = std::max(extent / nc, (double)rcut) * (1.0 + 1e-12) + 1e-300;
    }
  }
  g.lay_nlc = nlc;
  g.lay_next = next;
  g.magic = kGridMagic;
  const int64_t nlc_use = (int64_t)g.ncell[0] * g.ncell[1] * g.ncell[2];
  const int64_t next_use = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
  hipLaunchKernelGGL(k_write_params, dim3(1), dim3(1), 0, st, g, gp);
  DPMI_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(nlc_use, 256)), dim3(256), 0, st, cell_count, nlc_use, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(nlc_use, 256)), dim3(256), 0, st, cursor, nlc_use, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(16), 0, st, misc, (int64_t)16, 0);
  DPMI_LAUNCH_CHECK();
  if (nloc > 0) {
    hipLaunchKernelGGL(k_normalize_bin, dim3(ceil_div(nloc, 256)), dim3(256), 0, st, in_coord, in_type, out_coord,
                       out_type, mapping, cell_of, cell_count, g);
    DPMI_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, cell_count, cell_start, ghost_off, misc, g);
  DPMI_LAUNCH_CHECK();
  if (nloc > 0) {
    hipLaunchKernelGGL(k_cell_fill, dim3(ceil_div(nloc, 256)), dim3(256), 0, st, cell_of, cell_start, cursor,
                       unordered, (int)nloc);
    hipLaunchKernelGGL(k_cell_order, dim3(ceil_div(nloc, 256)), dim3(256), 0, st, cell_of, cell_start, unordered,
                       cell_atoms, pos_of, (int)nloc);
    DPMI_LAUNCH_CHECK();
  }
  if (!wait) {
    // nall stays in the workspace (status word 0); the capacity bound makes
    // every ghost fit, and k_ghosts flags it if one would not
    hipLaunchKernelGGL(k_ghosts, dim3(next_use), dim3(64), 0, st, out_coord, out_type, mapping, cell_start,
                       cell_atoms, ghost_off, g, mem_nall, misc);
    DPMI_LAUNCH_CHECK();
    return DPMI_OK;
  }
  int h_nall = 0;
  DPMI_HIP(hipMemcpyAsync(&h_nall, misc, sizeof(int), hipMemcpyDeviceToHost, st));
  DPMI_HIP(hipStreamSynchronize(st));
  *nall = h_nall;
  if (h_nall > mem_nall) {
    set_error("nall exceeds mem_nall");
    return DPMI_ERR_CAPACITY;
  }
  DPMI_CHECK_ARG(h_nall < (1 << 24), "nall >= 2^24 is not supported by the 64-bit neighbor key");
  if (g.pbc && next_use > 0 && h_nall > nloc) {
    hipLaunchKernelGGL(k_ghosts, dim3(next_use), dim3(64), 0, st, out_coord, out_type, mapping, cell_start,
                       cell_atoms, ghost_off, g, mem_nall, misc);
    DPMI_LAUNCH_CHECK();
  }
  return DPMI_OK;
}

}  // namespace dpmi

extern "C" int dpmi_copy_coord_f64(double* out_coord, int* out_type, int* mapping, int64_t* nall, int64_t mem_nall,
                                   const double* in_coord, const int* in_type, int64_t nloc,
                                   const double* box_host, float rcut, void* workspace, int64_t workspace_bytes,
                                   void* stream) {
  return copy_coord_impl(out_coord, out_type, mapping, nall, mem_nall, in_coord, in_type, nloc, box_host, rcut,
                         workspace, workspace_bytes, stream, true);
}

extern "C" int dpmi_copy_coord_capacity(int64_t* capacity, const double* box_host, int64_t nloc, float rcut) {
  DPMI_CHECK_ARG(capacity != nullptr && box_host != nullptr && nloc >= 0 && rcut > 0.f, "bad arguments");
  int64_t nlc, next;
  DPMI_CHECK_ARG(grid_sizes(box_host, nloc, rcut, &nlc, &next), "degenerate box");
  *capacity = copy_coord_capacity(box_host, nloc, rcut);
  return DPMI_OK;
}

extern "C" int dpmi_copy_coord_nowait_f64(double* out_coord, int* out_type, int* mapping, int64_t mem_nall,
                                          const double* in_coord, const int* in_type, int64_t nloc,
                                          const double* box_host, float rcut, void* workspace,
                                          int64_t workspace_bytes, void* stream) {
  return copy_coord_impl(out_coord, out_type, mapping, nullptr, mem_nall, in_coord, in_type, nloc, box_host, rcut,
                         workspace, workspace_bytes, stream, false);
}

namespace dpmi {

// wait: read the redo count and the status back (the form that returns
// errors); otherwise the redo pass runs on a fixed grid that reads the count
// from the workspace, nall comes from the workspace (nall < 0) and errors
// stay in the status words for dpmi_nlist_status.
static int build_nlist_impl(int* nlist, int* overflow, int* nlist_mapped, int16_t* cand_slot, const double* ext_coord,
                            const int* ext_type, const int* mapping, int64_t nloc, int64_t nall, float rcut,
                            const int* sel_host, int ntypes, const void* workspace, int64_t workspace_bytes,
                            void* stream, bool wait) {
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= kMaxTypes, "ntypes out of range [1, 64]");
  DPMI_CHECK_ARG(sel_host != nullptr, "sel is NULL");
  DPMI_CHECK_ARG(!wait || (nall >= nloc && nall < (int64_t(1) << 24)), "nall out of range");
  DPMI_CHECK_ARG(nlist_mapped == nullptr || mapping != nullptr, "nlist_mapped needs mapping");
  SelParams sp;
  std::memset(&sp, 0, sizeof(sp));
  sp.ntypes = ntypes;
  sp.sec[0] = 0;
  for (int t = 0; t < ntypes; ++t) {
    DPMI_CHECK_ARG(sel_host[t] >= 0, "sel must be >= 0");
    sp.sel[t] = sel_host[t];
    sp.sec[t + 1] = sp.sec[t] + sel_host[t];
  }
  sp.nnei = sp.sec[ntypes];
  DPMI_CHECK_ARG(sp.nnei <= 4096, "sum(sel) must be <= 4096");
  if (nloc == 0) return DPMI_OK;
  DPMI_CHECK_ARG(workspace != nullptr && workspace_bytes >= (int64_t)(512 + 64), "workspace too small");
  hipStream_t st = as_stream(stream);
  char* ws = const_cast<char*>(static_cast<const char*>(workspace));
  int* misc = reinterpret_cast<int*>(ws + ws_layout(0, 0, 0).misc);
  // status words misc[1..3] (capacity error, redo count, big-cell count)
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, misc + 1, (int64_t)3, 0);
  hipLaunchKernelGGL(k_ws_validate, dim3(1), dim3(64), 0, st, ws, (int)nloc, wait ? (int)nall : -1,
                     (long long)workspace_bytes);
  const size_t shm_cells = search_cells_lds(kCandCap, sp.nnei, search_waves(false)),
               shm_big = search_cells_lds(kCandCapBig, sp.nnei, search_waves(true));
  if (search_uses_cells(sp.nnei)) {
    // first pass over every cell; cells whose neighborhood exceeds its table
    // are listed for the second pass (larger table, fewer workgroups per CU);
    // three or four types keep each type's keys in a segment of its own
    auto passes = [&](auto first, auto second) -> int {
      if (shm_cells > 65536)
        DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(first), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)shm_cells));
      DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(second), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)shm_big));
      hipLaunchKernelGGL(first, dim3((unsigned)std::min<int64_t>(nloc, 2048)), dim3(64 * search_waves(false)),
                         shm_cells, st, nlist, overflow, nlist_mapped, cand_slot, ext_coord, ext_type, mapping, ws, sp,
                         rcut);
      hipLaunchKernelGGL(second, dim3((unsigned)std::min<int64_t>(nloc, 2048)), dim3(64 * search_waves(true)),
                         shm_big, st, nlist, overflow, nlist_mapped, cand_slot, ext_coord, ext_type, mapping, ws, sp,
                         rcut);
      return DPMI_OK;
    };
    const int rc = sp.ntypes > 2 && sp.ntypes <= kSegTypes
                       ? passes(k_search_cells<kCandCap, false, true>, k_search_cells<kCandCapBig, true, true>)
                       : passes(k_search_cells<kCandCap, false, false>, k_search_cells<kCandCapBig, true, false>);
    if (rc) return rc;
  } else {
    // very long rows (sum(sel)): every atom through the per-atom kernel
    size_t shm = 1024 * sizeof(uint64_t) + sizeof(int) * sp.nnei;
    hipLaunchKernelGGL(k_search_format<1024>, dim3((unsigned)nloc), dim3(64), shm, st, nlist, overflow, nlist_mapped,
                       ext_coord, ext_type, mapping, ws, nullptr, cand_slot, sp, rcut, nullptr);
  }
  DPMI_LAUNCH_CHECK();
  if (!wait) {
    if (search_uses_cells(sp.nnei)) {
      const int* redo = reinterpret_cast<const int*>(ws + ws_layout(nloc, 0, 0).redo);
      size_t shm = 1024 * sizeof(uint64_t) + sizeof(int) * sp.nnei;
      hipLaunchKernelGGL(k_search_format<1024>, dim3((unsigned)std::min<int64_t>(nloc, 2048)), dim3(64), shm, st,
                         nlist, overflow, nlist_mapped, ext_coord, ext_type, mapping, ws, redo, cand_slot, sp, rcut,
                         misc + 2);
      DPMI_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_poison_rows, dim3(256), dim3(256), 0, st, nlist, nlist_mapped, cand_slot, overflow,
                       nloc * (int64_t)sp.nnei, (int)nloc, ws);
    DPMI_LAUNCH_CHECK();
    return DPMI_OK;
  }
  int stat[2] = {0, 0};
  DPMI_HIP(hipMemcpyAsync(stat, misc + 1, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  DPMI_HIP(hipStreamSynchronize(st));
  if (stat[1] > 0) {
    // atoms left by the cell kernel: one wavefront each, up to 1024 candidates
    const int* redo = reinterpret_cast<const int*>(ws + ws_layout(nloc, 0, 0).redo);
    size_t shm = 1024 * sizeof(uint64_t) + sizeof(int) * sp.nnei;
    hipLaunchKernelGGL(k_search_format<1024>, dim3(stat[1]), dim3(64), shm, st, nlist, overflow, nlist_mapped,
                       ext_coord, ext_type, mapping, ws, redo, cand_slot, sp, rcut, nullptr);
    DPMI_LAUNCH_CHECK();
    DPMI_HIP(hipMemcpyAsync(stat, misc + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    DPMI_HIP(hipStreamSynchronize(st));
  }
  int err = stat[0];
  if (err & 2) {
    set_error("the workspace was not prepared by dpmi_copy_coord_f64 for this nloc and nall, or is smaller than "
              "dpmi_nlist_workspace_bytes gave for that system");
    return DPMI_ERR_ARG;
  }
  if (err) {
    set_error("more than 1024 candidates within rcut for some atom");
    return DPMI_ERR_CAPACITY;
  }
  return DPMI_OK;
}

// status word reported by dpmi_nlist_status: the search's error bits and
// those of the ghost copy (bit 4: more ghosts than the output holds)
__global__ void k_status(int* __restrict__ out, const char* __restrict__ ws) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int* misc = reinterpret_cast<const int*>(ws + ws_layout(0, 0, 0).misc);
    out[0] = misc[1] | misc[4];
  }
}

}  // namespace dpmi

extern "C" int dpmi_build_nlist_f64(int* nlist, int* overflow, int* nlist_mapped, int16_t* cand_slot,
                                    const double* ext_coord, const int* ext_type, const int* mapping, int64_t nloc,
                                    int64_t nall, float rcut, const int* sel_host, int ntypes, const void* workspace,
                                    int64_t workspace_bytes, void* stream) {
  return build_nlist_impl(nlist, overflow, nlist_mapped, cand_slot, ext_coord, ext_type, mapping, nloc, nall, rcut,
                          sel_host, ntypes, workspace, workspace_bytes, stream, true);
}

extern "C" int dpmi_build_nlist_nowait_f64(int* nlist, int* overflow, int* nlist_mapped, int16_t* cand_slot,
                                           const double* ext_coord, const int* ext_type, const int* mapping,
                                           int64_t nloc, float rcut, const int* sel_host, int ntypes,
                                           const void* workspace, int64_t workspace_bytes, void* stream) {
  return build_nlist_impl(nlist, overflow, nlist_mapped, cand_slot, ext_coord, ext_type, mapping, nloc, -1, rcut,
                          sel_host, ntypes, workspace, workspace_bytes, stream, false);
}

extern "C" int dpmi_nlist_status(int* status, const void* workspace, int64_t workspace_bytes, void* stream) {
  DPMI_CHECK_ARG(status != nullptr && workspace != nullptr && workspace_bytes >= (int64_t)(512 + 64),
                 "bad arguments");
  hipLaunchKernelGGL(k_status, dim3(1), dim3(64), 0, as_stream(stream), status, static_cast<const char*>(workspace));
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_format_nlist_f64(int* nlist, int* overflow, int* nlist_mapped, const int* cand,
                                     const int* numneigh, int64_t ncand, const double* ext_coord, const int* ext_type,
                                     const int* mapping, int64_t nloc, int64_t nall, float rcut, const int* sel_host,
                                     int ntypes, void* workspace, int64_t workspace_bytes, void* stream) {
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= kMaxTypes, "ntypes out of range [1, 64]");
  DPMI_CHECK_ARG(sel_host != nullptr, "sel is NULL");
  DPMI_CHECK_ARG(nloc >= 0 && nall >= nloc && nall < (int64_t(1) << 24), "nloc/nall out of range");
  DPMI_CHECK_ARG(ncand >= 0, "ncand must be >= 0");
  DPMI_CHECK_ARG(nlist_mapped == nullptr || mapping != nullptr, "nlist_mapped needs mapping");
  DPMI_CHECK_ARG(workspace != nullptr && workspace_bytes >= (int64_t)sizeof(int), "workspace too small (4 bytes)");
  SelParams sp;
  std::memset(&sp, 0, sizeof(sp));
  sp.ntypes = ntypes;
  for (int t = 0; t < ntypes; ++t) {
    DPMI_CHECK_ARG(sel_host[t] >= 0, "sel must be >= 0");
    sp.sel[t] = sel_host[t];
    sp.sec[t + 1] = sp.sec[t] + sel_host[t];
  }
  sp.nnei = sp.sec[ntypes];
  DPMI_CHECK_ARG(sp.nnei <= 4096, "sum(sel) must be <= 4096");
  if (nloc == 0) return DPMI_OK;
  hipStream_t st = as_stream(stream);
  int* status = static_cast<int*>(workspace);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, status, (int64_t)1, 0);
  const size_t shm = 1024 * sizeof(uint64_t) + sizeof(int) * sp.nnei;
  hipLaunchKernelGGL(k_format_external<1024>, dim3((unsigned)nloc), dim3(64), shm, st, nlist, overflow, nlist_mapped,
                     status, cand, numneigh, ncand, ext_coord, ext_type, mapping, (int)nall, sp, rcut);
  DPMI_LAUNCH_CHECK();
  int h = 0;
  DPMI_HIP(hipMemcpyAsync(&h, status, sizeof(int), hipMemcpyDeviceToHost, st));
  DPMI_HIP(hipStreamSynchronize(st));
  if (h) {
    set_error("more than 1024 neighbors within rcut for some atom");
    return DPMI_ERR_CAPACITY;
  }
  return DPMI_OK;
}
