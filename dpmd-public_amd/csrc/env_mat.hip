// Environment matrix, force and virial kernels for se_e2_a on gfx950.
//
// Mapping: one 64-wide wavefront per centre atom, four atoms per 256-thread
// workgroup.  A lane owns neighbor slots lane, lane+64, ...; the slot's four
// em values (and twelve derivatives) are contiguous, so each store
// instruction of a wave writes one contiguous span of the [nloc, nnei, 4]
// (or [.., 12]) row: the stores are fully coalesced.  Centre-atom sums are
// wavefront reductions; only neighbor forces use global atomics.  The virial
// is reduced deterministically: per-workgroup partials in LDS, then one
// compensated (Neumaier) sum per component.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <type_traits>

#include <cstring>

#include "cells.h"
#include "dpmi_internal.h"

namespace dpmi {
namespace {

constexpr int kMaxTypes = 64;
constexpr int kAtomsPerBlock = 4;

struct SecParams {
  int ntypes;
  int nnei;
  int sec[kMaxTypes + 1];
};

inline int make_sec(SecParams* sp, const int* sel_host, int ntypes) {
  DPMI_CHECK_ARG(sel_host != nullptr, "sel is NULL");
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= kMaxTypes, "ntypes out of range [1, 64]");
  std::memset(sp, 0, sizeof(*sp));
  sp->ntypes = ntypes;
  for (int t = 0; t < ntypes; ++t) {
    DPMI_CHECK_ARG(sel_host[t] >= 0, "sel must be >= 0");
    sp->sec[t + 1] = sp->sec[t] + sel_host[t];
  }
  sp->nnei = sp->sec[ntypes];
  return DPMI_OK;
}

// Type section of slot s: the number of section starts at or below s (the
// loop index is wave-uniform, so the section bounds are scalar loads; a
// per-lane index into the kernel arguments would be a memory round trip).
__device__ inline int slot_type(const SecParams& sp, int s) {
  int t = 0;
  for (int k = 1; k < sp.ntypes; ++k) t += (int)(s >= sp.sec[k]);
  return t;
}

// lim[t] = end of section t, written from uniform indices
__device__ inline void section_ends(const SecParams& sp, int* lim) {
  if (lane_id() == 0)
    for (int t = 0; t < sp.ntypes; ++t) lim[t] = sp.sec[t + 1];
}

// spline5_switch, source/lib/include/switcher.h:60-85 (rmin, rmax are float).
template <typename FP>
__device__ inline void spline5(FP& vv, FP& dd, FP xx, float rmin, float rmax) {
  if (xx < rmin) {
    dd = (FP)0.;
    vv = (FP)1.;
  } else if (xx < rmax) {
    FP uu = (xx - rmin) / (FP)(rmax - rmin);
    FP du = (FP)1. / (FP)(rmax - rmin);
    vv = uu * uu * uu * ((FP)-6. * uu * uu + (FP)15. * uu - (FP)10.) + (FP)1.;
    dd = ((FP)3. * uu * uu * ((FP)-6. * uu * uu + (FP)15. * uu - (FP)10.) +
          uu * uu * uu * ((FP)-12. * uu + (FP)15.)) *
         du;
  } else {
    dd = (FP)0.;
    vv = (FP)0.;
  }
}

// env_mat_a_cpu for one neighbor (source/lib/src/env_mat.cc:159-219): the four
// values and their derivative w.r.t. the centre atom, before normalisation.
template <typename FP>
__device__ inline void env_one(const FP rr[3], float rmin, float rmax, FP v[4], FP d[12]) {
  FP nr2 = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
  FP inr = (FP)1. / sqrt(nr2);
  FP nr = nr2 * inr;
  FP inr2 = inr * inr;
  FP inr4 = inr2 * inr2;
  FP inr3 = inr4 * nr;
  FP sw, dsw;
  spline5(sw, dsw, nr, rmin, rmax);
  v[0] = (FP)1. / nr;
  v[1] = rr[0] / nr2;
  v[2] = rr[1] / nr2;
  v[3] = rr[2] / nr2;
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) d[dd] = rr[dd] * inr3 * sw - v[0] * dsw * rr[dd] * inr;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) {
      FP t = (FP)2. * rr[c] * rr[dd] * inr4;
      if (c == dd) t = t - inr2;
      d[3 + 3 * c + dd] = t * sw - v[1 + c] * dsw * rr[dd] * inr;
    }
  v[0] *= sw;
  v[1] *= sw;
  v[2] *= sw;
  v[3] *= sw;
}

// The same in fp32 with hardware reciprocals instead of IEEE divisions (the
// recompute adjoint of the fp32 forms: a few ulp, far below their bf16 / fp32
// operand rounding): 1/r from v_rsq_f32, 1/(rmax - rmin) precomputed.
__device__ inline void env_one_fast(const float rr[3], float rmin, float inv_range, float v[4], float d[12]) {
  const float nr2 = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
  const float inr = __builtin_amdgcn_rsqf(nr2);
  const float nr = nr2 * inr;
  const float inr2 = inr * inr;
  const float inr4 = inr2 * inr2;
  const float inr3 = inr4 * nr;
  float sw, dsw;
  const float uu = (nr - rmin) * inv_range;
  if (uu <= 0.f) {
    sw = 1.f;
    dsw = 0.f;
  } else if (uu < 1.f) {
    sw = uu * uu * uu * (-6.f * uu * uu + 15.f * uu - 10.f) + 1.f;
    dsw = (3.f * uu * uu * (-6.f * uu * uu + 15.f * uu - 10.f) + uu * uu * uu * (-12.f * uu + 15.f)) * inv_range;
  } else {
    sw = 0.f;
    dsw = 0.f;
  }
  v[0] = inr;
  v[1] = rr[0] * inr2;
  v[2] = rr[1] * inr2;
  v[3] = rr[2] * inr2;
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) d[dd] = rr[dd] * inr3 * sw - v[0] * dsw * rr[dd] * inr;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) {
      float t = 2.f * rr[c] * rr[dd] * inr4;
      if (c == dd) t = t - inr2;
      d[3 + 3 * c + dd] = t * sw - v[1 + c] * dsw * rr[dd] * inr;
    }
  v[0] *= sw;
  v[1] *= sw;
  v[2] *= sw;
  v[3] *= sw;
}

// First -1 in each type section of the wave's row: slots at or after it are
// empty (env_mat_a_cpu breaks at the first negative entry of a section).
// lim is the wavefront's own LDS row of kMaxTypes ints.
__device__ inline void section_limits_wave(const int* __restrict__ nl, const SecParams& sp, int* lim) {
  const int lane = lane_id();
  section_ends(sp, lim);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int s = lane; s < sp.nnei; s += kWave)
    if (nl[s] < 0) atomicMin(&lim[slot_type(sp, s)], s);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// FP: arithmetic and output type; CT: coordinate type (the difference
// x_j - x_i is formed in CT, then rounded to FP).
template <typename FP, typename CT>
__global__ void __launch_bounds__(256)
    k_env_mat_a(FP* __restrict__ em, FP* __restrict__ em_deriv, FP* __restrict__ rij, const int* __restrict__ nlist,
                const CT* __restrict__ ext_c, const int* __restrict__ ext_t, int nloc, const FP* __restrict__ avg,
                const FP* __restrict__ std_, float rmin, float rmax, SecParams sp) {
  __shared__ int lim_all[kAtomsPerBlock][kMaxTypes];
  const int w = wave_id(), lane = lane_id();
  const int i = blockIdx.x * kAtomsPerBlock + w;
  const bool active = i < nloc;
  const int nnei = sp.nnei;
  const int ii = active ? i : 0;
  const int* nl = nlist + (int64_t)ii * nnei;
  int ti = ext_t[ii];
  if (ti >= sp.ntypes) ti = -1;  // no statistics for such a type: treated as virtual
  const CT xi0 = ext_c[3 * ii + 0], xi1 = ext_c[3 * ii + 1], xi2 = ext_c[3 * ii + 2];
  const int nem = nnei * 4;
  const float inv_range = 1.f / (rmax - rmin);
  // one neighbor slot: coordinates loaded, limits known
  auto slot = [&](int s, int j, bool valid, const CT xj[3]) {
    FP rr[3] = {(FP)0., (FP)0., (FP)0.};
    FP v[4] = {(FP)0., (FP)0., (FP)0., (FP)0.};
    FP d[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) d[k] = (FP)0.;
    if (valid) {
      rr[0] = (FP)(xj[0] - xi0);
      rr[1] = (FP)(xj[1] - xi1);
      rr[2] = (FP)(xj[2] - xi2);
      // fp32: the reciprocal-based core the fp32 force recompute uses too
      if constexpr (sizeof(FP) == 4)
        env_one_fast(rr, rmin, inv_range, v, d);
      else
        env_one(rr, rmin, rmax, v, d);
    }
    int64_t o = (int64_t)i * nem + 4 * s;
    if (ti >= 0) {
      const FP* av = avg + (int64_t)ti * nem + 4 * s;
      const FP* sd = std_ + (int64_t)ti * nem + 4 * s;
      if constexpr (sizeof(FP) == 4) {
        // fp32: times the hardware reciprocal of std (1 ulp) instead of four
        // IEEE divisions (and twelve more for the derivative)
        FP inv[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) inv[c] = __builtin_amdgcn_rcpf(sd[c]);
#pragma unroll
        for (int c = 0; c < 4; ++c) em[o + c] = (v[c] - av[c]) * inv[c];
        if (em_deriv) {
#pragma unroll
          for (int k = 0; k < 12; ++k) em_deriv[3 * o + k] = d[k] * inv[k / 3];
        }
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) em[o + c] = (v[c] - av[c]) / sd[c];
        if (em_deriv) {
#pragma unroll
          for (int k = 0; k < 12; ++k) em_deriv[3 * o + k] = d[k] / sd[k / 3];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) em[o + c] = (FP)0.;
      if (em_deriv) {
#pragma unroll
        for (int k = 0; k < 12; ++k) em_deriv[3 * o + k] = (FP)0.;
      }
    }
    if (rij) {
      int64_t r = (int64_t)i * nnei * 3 + 3 * s;
      rij[r + 0] = rr[0];
      rij[r + 1] = rr[1];
      rij[r + 2] = rr[2];
    }
  };
  if (nnei <= kWave * 4) {
    // the row in registers, every coordinate load issued before any arithmetic
    int jv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) jv[k] = lane + kWave * k < nnei ? nl[lane + kWave * k] : 0;
    // the limits row is the wavefront's own: wave-local ordering suffices
    section_ends(sp, lim_all[w]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int s = lane + kWave * k;
      if (s < nnei && jv[k] < 0) atomicMin(&lim_all[w][slot_type(sp, s)], s);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (!active) return;
    bool valid[4];
    CT xj[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int s = lane + kWave * k;
      valid[k] = s < nnei && s < lim_all[w][slot_type(sp, s)];
      const int j = valid[k] ? jv[k] : i;
#pragma unroll
      for (int c = 0; c < 3; ++c) xj[k][c] = ext_c[3 * j + c];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (lane + kWave * k < nnei) slot(lane + kWave * k, jv[k], valid[k], xj[k]);
  } else {
    section_limits_wave(nl, sp, lim_all[w]);
    if (!active) return;
    for (int s = lane; s < nnei; s += kWave) {
      const int j = nl[s];
      const bool valid = s < lim_all[w][slot_type(sp, s)];
      const CT xj[3] = {valid ? ext_c[3 * j + 0] : xi0, valid ? ext_c[3 * j + 1] : xi1,
                        valid ? ext_c[3 * j + 2] : xi2};
      slot(s, j, valid, xj);
    }
  }
}

template <typename FP>
__global__ void k_fill(FP* p, int64_t n, FP v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// prod_force_a (source/lib/src/prod_force.cc:23-76): F_i -= sum_a nd*ed,
// F_j += sum_{a in slot j} nd*ed.
template <typename FP>
__global__ void __launch_bounds__(256)
    k_prod_force(FP* __restrict__ force, const FP* __restrict__ nd, const FP* __restrict__ ed,
                 const int* __restrict__ nlist, int nloc, int nall, int nnei, int nframes) {
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int64_t gi = (int64_t)blockIdx.x * kAtomsPerBlock + w;
  if (gi >= (int64_t)nloc * nframes) return;
  const int64_t f = gi / nloc, li = gi % nloc;
  const int64_t nd_off = gi * nnei * 4;
  FP ci[3] = {(FP)0., (FP)0., (FP)0.};
  FP* fr = force + f * nall * 3;
  for (int s = lane; s < nnei; s += kWave) {
    FP g[3] = {(FP)0., (FP)0., (FP)0.};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      FP n = nd[nd_off + 4 * s + c];
      const FP* e = ed + (nd_off + 4 * s + c) * 3;
      g[0] += n * e[0];
      g[1] += n * e[1];
      g[2] += n * e[2];
    }
    ci[0] += g[0];
    ci[1] += g[1];
    ci[2] += g[2];
    int j = nlist[gi * nnei + s];
    if (j >= 0) {
      atomicAdd(&fr[3 * (int64_t)j + 0], g[0]);
      atomicAdd(&fr[3 * (int64_t)j + 1], g[1]);
      atomicAdd(&fr[3 * (int64_t)j + 2], g[2]);
    }
  }
  ci[0] = wave_sum(ci[0]);
  ci[1] = wave_sum(ci[1]);
  ci[2] = wave_sum(ci[2]);
  if (lane == 0) {
    atomicAdd(&fr[3 * li + 0], -ci[0]);
    atomicAdd(&fr[3 * li + 1], -ci[1]);
    atomicAdd(&fr[3 * li + 2], -ci[2]);
  }
}

// Block partial of nine virial components -> partial[block*9 + k] (double).
template <typename FP>
__device__ inline void block_virial_partial(double v[9], double* __restrict__ partial) {
  __shared__ double red[kAtomsPerBlock][9];
  const int w = threadIdx.x / kWave, lane = lane_id();
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 9; ++k) red[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 9) {
    double s = 0.;
    for (int q = 0; q < kAtomsPerBlock; ++q) s += red[q][threadIdx.x];
    partial[(int64_t)blockIdx.x * 9 + threadIdx.x] = s;
  }
}

// prod_virial_a (source/lib/src/prod_virial.cc:22-70):
// V[d0][d1] += nd * rij[d1] * ed[d0] over valid neighbor slots.
template <typename FP>
__global__ void __launch_bounds__(256)
    k_prod_virial(double* __restrict__ partial, FP* __restrict__ atom_virial, const FP* __restrict__ nd,
                  const FP* __restrict__ ed, const FP* __restrict__ rij, const int* __restrict__ nlist, int nloc,
                  int nall, int nnei, int blocks_per_frame) {
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int f = blockIdx.x / blocks_per_frame;
  const int64_t li = (int64_t)(blockIdx.x % blocks_per_frame) * kAtomsPerBlock + w;
  double v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = 0.;
  if (li < nloc) {
    const int64_t gi = (int64_t)f * nloc + li;
    for (int s = lane; s < nnei; s += kWave) {
      int j = nlist[gi * nnei + s];
      if (j < 0) continue;
      const FP* r = rij + (gi * nnei + s) * 3;
      FP g[3] = {(FP)0., (FP)0., (FP)0.};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        FP n = nd[gi * nnei * 4 + 4 * s + c];
        const FP* e = ed + (gi * nnei * 4 + 4 * s + c) * 3;
        g[0] += n * e[0];
        g[1] += n * e[1];
        g[2] += n * e[2];
      }
#pragma unroll
      for (int d0 = 0; d0 < 3; ++d0)
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1) {
          FP tmp = g[d0] * r[d1];
          v[d0 * 3 + d1] += (double)tmp;
          if (atom_virial) atomicAdd(&atom_virial[((int64_t)f * nall + j) * 9 + d0 * 3 + d1], tmp);
        }
    }
  }
  block_virial_partial<FP>(v, partial);
}

// prod_force_grad_a (source/lib/src/prod_force_grad.cc:22-79), the transpose
// of prod_force: gn[i, 4s+c] = -<g_i, ed[i,4s+c]> + <g_j, ed[i,4s+c]> for a
// valid neighbor j of slot s.  grad holds nall rows per frame; an index
// j >= nall reads row j % nloc (with nall = nloc: the reference's rule).
// One lane per (atom, slot): 12 contiguous ed values in, 4 gn values out.
template <typename FP>
__global__ void __launch_bounds__(256)
    k_prod_force_grad(FP* __restrict__ gn, const FP* __restrict__ grad, const FP* __restrict__ ed,
                      const int* __restrict__ nlist, int64_t nslots, int nloc, int nall, int nnei) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const int64_t gi = t / nnei;  // frame * nloc + atom
  const int64_t f = gi / nloc;
  const FP* e = ed + t * 12;
  const FP* g = grad + (f * nall + (gi - f * nloc)) * 3;
  FP o[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    FP a = (FP)0.;
    a -= g[0] * e[3 * c + 0];
    a -= g[1] * e[3 * c + 1];
    a -= g[2] * e[3 * c + 2];
    o[c] = a;
  }
  int j = nlist[t];
  if (j >= nall) j = j % nloc;
  if (j >= 0) {
    const FP* gj = grad + (f * nall + j) * 3;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      o[c] += gj[0] * e[3 * c + 0];
      o[c] += gj[1] * e[3 * c + 1];
      o[c] += gj[2] * e[3 * c + 2];
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) gn[t * 4 + c] = o[c];
}

// prod_virial_grad_a (source/lib/src/prod_virial_grad.cc:20-62):
// gn[i, 4s+c] = sum_{d0,d1} g[d0,d1] rij[i,s,d1] ed[i,4s+c,d0] for a valid
// neighbor, 0 for an empty slot; g is the frame's [9] virial gradient.
template <typename FP>
__global__ void __launch_bounds__(256)
    k_prod_virial_grad(FP* __restrict__ gn, const FP* __restrict__ grad, const FP* __restrict__ ed,
                       const FP* __restrict__ rij, const int* __restrict__ nlist, int64_t nslots, int nloc,
                       int nnei) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const int64_t f = t / ((int64_t)nloc * nnei);
  FP o[4] = {(FP)0., (FP)0., (FP)0., (FP)0.};
  if (nlist[t] >= 0) {
    const FP* e = ed + t * 12;
    const FP* r = rij + t * 3;
    const FP* g = grad + f * 9;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int d0 = 0; d0 < 3; ++d0)
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1) o[c] += g[d0 * 3 + d1] * r[d1] * e[3 * c + d0];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) gn[t * 4 + c] = o[c];
}

// Compensated, deterministic sum of per-block partials: one workgroup per
// frame, 9 x 64 lanes; lane q of component k sums blocks q, q+64, ...
// (Neumaier), then a fixed-order tree combines the 64 (sum, carry) pairs.
template <typename FP>
__global__ void __launch_bounds__(576)
    k_virial_reduce(FP* __restrict__ out, const double* __restrict__ partial, int blocks_per_frame) {
  const int k = threadIdx.x / kWave, lane = lane_id();
  const int f = blockIdx.x;
  const double* p = partial + (int64_t)f * blocks_per_frame * 9;
  double s = 0., c = 0.;
  // eight loads in flight, summed in block order (b = lane, lane + 64, ...)
  constexpr int kB = 8;
  for (int b0 = lane; b0 < blocks_per_frame; b0 += kB * kWave) {
    double xs[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int b = b0 + u * kWave;
      xs[u] = b < blocks_per_frame ? p[(int64_t)b * 9 + k] : 0.;
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      if (b0 + u * kWave >= blocks_per_frame) break;
      const double x = xs[u], t = s + x;
      c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
      s = t;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    double so = __shfl_xor(s, off, kWave), co = __shfl_xor(c, off, kWave);
    double lo_s = lane & off ? so : s, hi_s = lane & off ? s : so;
    double t = lo_s + hi_s;
    double e = (fabs(lo_s) >= fabs(hi_s)) ? ((lo_s - t) + hi_s) : ((hi_s - t) + lo_s);
    c = (lane & off ? co + c : c + co) + e;
    s = t;
  }
  if (lane == 0) out[(int64_t)f * 9 + k] = (FP)(s + c);
}

// Fused backward of the env-mat for the model path: em_deriv is recomputed
// from coordinates instead of being stored, forces on extended atoms are
// folded onto their local owners through `mapping` (the ghost -> local
// scatter of communicate_extended_output), and the virial partials are
// formed in the same pass.  net_deriv is dE/d em (normalised em).
template <typename FP, typename ACC>
__global__ void __launch_bounds__(256)
    k_force_virial_recompute(ACC* __restrict__ force, double* __restrict__ partial, const FP* __restrict__ nd,
                             const int* __restrict__ nlist, const int* __restrict__ mapping,
                             const double* __restrict__ ext_c, const int* __restrict__ ext_t, int nloc,
                             const FP* __restrict__ std_, float rmin, float rmax, SecParams sp) {
  __shared__ int lim_all[kAtomsPerBlock][kMaxTypes];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int nnei = sp.nnei;
  double v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = 0.;
  // persistent grid: each wavefront walks atoms; one virial partial per workgroup
  for (int i = blockIdx.x * kAtomsPerBlock + w; i < nloc; i += gridDim.x * kAtomsPerBlock) {
    const int* nl = nlist + (int64_t)i * nnei;
    section_limits_wave(nl, sp, lim_all[w]);
    const int ti = ext_t[i] < sp.ntypes ? ext_t[i] : -1;
    const double xi0 = ext_c[3 * i + 0], xi1 = ext_c[3 * i + 1], xi2 = ext_c[3 * i + 2];
    ACC ci[3] = {0, 0, 0};
    for (int s = lane; s < nnei && ti >= 0; s += kWave) {
      int t = slot_type(sp, s);
      if (s >= lim_all[w][t]) continue;
      int j = nl[s];
      double r64[3] = {ext_c[3 * j + 0] - xi0, ext_c[3 * j + 1] - xi1, ext_c[3 * j + 2] - xi2};
      FP rr[3] = {(FP)r64[0], (FP)r64[1], (FP)r64[2]};
      FP val[4], d[12];
      env_one(rr, rmin, rmax, val, d);
      const FP* sd = std_ + ((int64_t)ti * nnei + s) * 4;
      const FP* n = nd + ((int64_t)i * nnei + s) * 4;
      FP g[3] = {(FP)0., (FP)0., (FP)0.};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        FP nc = n[c] / sd[c];
        g[0] += nc * d[3 * c + 0];
        g[1] += nc * d[3 * c + 1];
        g[2] += nc * d[3 * c + 2];
      }
      ci[0] += g[0];
      ci[1] += g[1];
      ci[2] += g[2];
      int o = mapping ? mapping[j] : j;
      atomicAdd(&force[3 * (int64_t)o + 0], (ACC)g[0]);
      atomicAdd(&force[3 * (int64_t)o + 1], (ACC)g[1]);
      atomicAdd(&force[3 * (int64_t)o + 2], (ACC)g[2]);
#pragma unroll
      for (int d0 = 0; d0 < 3; ++d0)
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1) v[d0 * 3 + d1] += (double)g[d0] * r64[d1];
    }
    ci[0] = wave_sum(ci[0]);
    ci[1] = wave_sum(ci[1]);
    ci[2] = wave_sum(ci[2]);
    if (lane == 0) {
      atomicAdd(&force[3 * (int64_t)i + 0], -ci[0]);
      atomicAdd(&force[3 * (int64_t)i + 1], -ci[1]);
      atomicAdd(&force[3 * (int64_t)i + 2], -ci[2]);
    }
  }
  block_virial_partial<FP>(v, partial);
}

// Cell-wise form of the recompute adjoint, without memory atomics for the
// neighbor forces.  Pass 1: one workgroup per local cell (grid-stride); the
// cell's neighbor terms g(i, s) accumulate in LDS at the neighbor's position in
// the cell's candidate table (cand_slot from the neighbor search), then each
// table entry is written once to partial[neighbor-cell index][position of the
// target in the cell-sorted order] (whole runs per neighbor cell).  The
// centre terms and the virial are as in k_force_virial_recompute.  Atoms or
// cells the search left to its per-atom path (cand_slot < 0, or a table
// over the table capacity) add their neighbor terms with atomics instead.
// The table sums do not depend on timing: in fp32 one table of 64-bit
// fixed-point sums (2^-32 resolution, |term| < 2^31) takes integer LDS atomics,
// which are exact and so order-free; in fp64 each wavefront accumulates into
// its own copy (a wave's lanes never share a slot, so plain adds in program
// order suffice) and the copies are summed in a fixed order.  The first pass
// takes every cell with tables up to kCandCap; cells the search moved to its
// second pass (bigger tables) are left to a second force pass (kBigPass, fp32:
// 3072 x 24 B of LDS) or, in FP64 (4 x 1536 x 24 B already), processed with
// neighbor atomics.
template <typename FP, int CAP, bool kBigPass, int kChunks>
__global__ void __launch_bounds__(256)
    k_force_cells(double* __restrict__ force, FP* __restrict__ partial, int* __restrict__ cell_ok,
                  double* __restrict__ vpartial, const FP* __restrict__ nd, const int* __restrict__ nlist,
                  const int16_t* __restrict__ cand_slot, const int* __restrict__ mapping,
                  const double* __restrict__ ext_c, const int* __restrict__ ext_t, const FP* __restrict__ std_,
                  const char* __restrict__ nlws, float rmin, float rmax, SecParams sp, int nloc) {
  __shared__ int s_base[kMaxNb], s_local[kMaxNb], s_pref[kMaxNb + 1];
  __shared__ int lim_all[kAtomsPerBlock][kMaxTypes];
  extern __shared__ unsigned char smem_raw[];
  constexpr int kCap = CAP;
  constexpr bool kFixed = sizeof(FP) == 4;   // fixed-point table (fp32) or per-wave copies (fp64)
  constexpr bool kLeaveBig = !kBigPass && kFixed;  // a second pass takes the big cells
  constexpr float kFix = 4294967296.0f;      // 2^32
  constexpr double kUnfix = 1.0 / 4294967296.0;
  FP* acc = reinterpret_cast<FP*>(smem_raw);  // fp64: [4 wavefronts][kCap][3]
  unsigned long long* accq = reinterpret_cast<unsigned long long*>(smem_raw);  // fp32: [kCap][3]
  constexpr int kAccWords = kFixed ? 3 * kCap : kAtomsPerBlock * 3 * kCap;
  const GridParams p = *reinterpret_cast<const GridParams*>(nlws);
  const WsLayout L = ws_layout(p.nloc, p.lay_nlc, p.lay_next);
  const int* __restrict__ cell_start = reinterpret_cast<const int*>(nlws + L.cell_start);
  const int* __restrict__ cell_atoms = reinterpret_cast<const int*>(nlws + L.cell_atoms);
  const int* __restrict__ ghost_off = reinterpret_cast<const int*>(nlws + L.ghost_off);
  const int* __restrict__ misc = reinterpret_cast<const int*>(nlws + L.misc);
  const int* __restrict__ bigcells = reinterpret_cast<const int*>(nlws + L.bigcells);
  const int* __restrict__ pos_of = reinterpret_cast<const int*>(nlws + L.pos_of);
  const int tid = threadIdx.x, w = wave_id(), lane = lane_id();
  const int side = 2 * p.nsearch + 1, nnb = side * side * side;
  // a neighbor-list workspace of another system: no cell is processed (the
  // gather then writes NaN forces)
  const bool ws_ok = p.magic == kGridMagic && p.nloc == nloc;
  const int ncells = !ws_ok ? 0 : kBigPass ? misc[3] : p.ncell[0] * p.ncell[1] * p.ncell[2];
  const int nnei = sp.nnei;
  const float inv_range = 1.f / (rmax - rmin);
  double v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = 0.;
  for (int ci = blockIdx.x; ci < ncells; ci += gridDim.x) {
    const int cc = kBigPass ? bigcells[ci] : ci;
    const int a0 = cell_start[cc], na = cell_start[cc + 1] - a0;
    if (na == 0) {  // uniform in the workgroup
      if (tid == 0) cell_ok[cc] = 0;
      continue;
    }
    const int T = nb_table(p, cell_start, ghost_off, cc, s_base, s_local, s_pref);
    const bool ok = T <= kCap;
    if (kLeaveBig && !ok) continue;  // uniform; the second pass writes cell_ok
    if (tid == 0) cell_ok[cc] = ok ? 1 : 0;
    if (ok) {
      if constexpr (kFixed)
        for (int k = tid; k < kAccWords; k += blockDim.x) accq[k] = 0ull;
      else
        for (int k = tid; k < kAccWords; k += blockDim.x) acc[k] = (FP)0.;
    }
    __syncthreads();
    FP* accw = acc + w * 3 * kCap;
    // neighbor term of slot s of atom i (neighbor j, its coordinates, dE/d em
    // row, std row and table slot already loaded); centre sum in ci
    auto term = [&](double ci[3], double xi0, double xi1, double xi2, int j, const double xj[3], const FP n[4],
                    const FP sd[4], int cs) {
      const double r64[3] = {xj[0] - xi0, xj[1] - xi1, xj[2] - xi2};
      const FP rr[3] = {(FP)r64[0], (FP)r64[1], (FP)r64[2]};
      FP val[4], d[12];
      if constexpr (kFixed)
        env_one_fast(rr, rmin, inv_range, val, d);
      else
        env_one(rr, rmin, rmax, val, d);
      FP g[3] = {(FP)0., (FP)0., (FP)0.};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        FP nc;
        if constexpr (kFixed)
          nc = n[c] * __builtin_amdgcn_rcpf(sd[c]);
        else
          nc = n[c] / sd[c];
        g[0] += nc * d[3 * c + 0];
        g[1] += nc * d[3 * c + 1];
        g[2] += nc * d[3 * c + 2];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) ci[k] += (double)g[k];
#pragma unroll
      for (int d0 = 0; d0 < 3; ++d0)
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1) v[d0 * 3 + d1] += (double)g[d0] * r64[d1];
      if (cs >= 0 && cs < T) {
        if constexpr (kFixed) {
#pragma unroll
          for (int k = 0; k < 3; ++k)
            atomicAdd(&accq[3 * cs + k], (unsigned long long)__float2ll_rn((float)g[k] * kFix));
        } else {
          accw[3 * cs + 0] += g[0];
          accw[3 * cs + 1] += g[1];
          accw[3 * cs + 2] += g[2];
        }
      } else {
        const int o = mapping ? mapping[j] : j;
        atomicAdd(&force[3 * (int64_t)o + 0], (double)g[0]);
        atomicAdd(&force[3 * (int64_t)o + 1], (double)g[1]);
        atomicAdd(&force[3 * (int64_t)o + 2], (double)g[2]);
      }
    };
    auto centre = [&](int i, double ci[3]) {
#pragma unroll
      for (int k = 0; k < 3; ++k) ci[k] = wave_sum(ci[k]);
      if (lane == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) atomicAdd(&force[3 * (int64_t)i + k], -ci[k]);
    };
    if (nnei <= kWave * kChunks) {
      // The atom's whole row in registers.  In fp32 the loads that need
      // nothing but the atom (list, dE/d em, table slots, its coordinates)
      // are issued one atom ahead, so a wavefront waits for one round trip
      // per atom (the neighbor coordinates) instead of two.
      struct Row {
        int i, ti;
        double xi[3];
        int jv[kChunks], csv[kChunks];
        FP nv[kChunks][4];
      };
      auto load_row = [&](int ai, Row& R) {
        const int i = ai < na ? cell_atoms[a0 + ai] : cell_atoms[a0];
        R.i = i;
        const int tr = ext_t[i];
        R.ti = ai < na && tr < sp.ntypes ? tr : -1;
#pragma unroll
        for (int c = 0; c < 3; ++c) R.xi[c] = ext_c[3 * i + c];
        const int* nl = nlist + (int64_t)i * nnei;
#pragma unroll
        for (int k = 0; k < kChunks; ++k) {
          const int s = lane + kWave * k;
          const bool in = s < nnei && R.ti >= 0;
          R.jv[k] = s < nnei ? nl[s] : 0;
          const FP* n = nd + ((int64_t)i * nnei + (in ? s : 0)) * 4;
#pragma unroll
          for (int c = 0; c < 4; ++c) R.nv[k][c] = in ? n[c] : (FP)0.;
          R.csv[k] = in && ok ? (int)cand_slot[(int64_t)i * nnei + s] : -1;
        }
      };
      // (fp64 holds twice the registers per row: it loads at use)
      constexpr bool kAhead = kFixed;
      Row nx;
      if constexpr (kAhead) load_row(w, nx);
      for (int ai = w; ai < na; ai += kAtomsPerBlock) {
        Row R;
        if constexpr (kAhead) {
          R = nx;
          if (ai + kAtomsPerBlock < na) load_row(ai + kAtomsPerBlock, nx);
        } else {
          load_row(ai, R);
        }
        const int i = R.i, ti = R.ti;
        // section limits (first -1 of each type section) from the registers
        section_ends(sp, lim_all[w]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < kChunks; ++k) {
          const int s = lane + kWave * k;
          if (s < nnei && R.jv[k] < 0) atomicMin(&lim_all[w][slot_type(sp, s)], s);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        bool act[kChunks];
        double xj[kChunks][3];
        FP sv[kChunks][4];
#pragma unroll
        for (int k = 0; k < kChunks; ++k) {
          const int s = lane + kWave * k;
          act[k] = ti >= 0 && s < nnei && s < lim_all[w][slot_type(sp, s)];
          const int j = act[k] ? R.jv[k] : i;
#pragma unroll
          for (int c = 0; c < 3; ++c) xj[k][c] = ext_c[3 * j + c];
          // the std rows (ntypes x nnei x 4, cache resident) at use
          const FP* sd = std_ + ((int64_t)(ti < 0 ? 0 : ti) * nnei + (s < nnei ? s : 0)) * 4;
#pragma unroll
          for (int c = 0; c < 4; ++c) sv[k][c] = act[k] ? sd[c] : (FP)1.;
        }
        double ci[3] = {0., 0., 0.};
#pragma unroll
        for (int k = 0; k < kChunks; ++k)
          if (act[k]) term(ci, R.xi[0], R.xi[1], R.xi[2], R.jv[k], xj[k], R.nv[k], sv[k], R.csv[k]);
        centre(i, ci);
      }
    } else {
      for (int ai = w; ai < na; ai += kAtomsPerBlock) {
        const int i = cell_atoms[a0 + ai];
        const int* nl = nlist + (int64_t)i * nnei;
        const int ti = ext_t[i] < sp.ntypes ? ext_t[i] : -1;
        const double xi0 = ext_c[3 * i + 0], xi1 = ext_c[3 * i + 1], xi2 = ext_c[3 * i + 2];
        double ci[3] = {0., 0., 0.};
        section_limits_wave(nl, sp, lim_all[w]);
        for (int s = lane; s < nnei && ti >= 0; s += kWave) {
          const int t = slot_type(sp, s);
          if (s >= lim_all[w][t]) continue;
          const int j = nl[s];
          const double xj[3] = {ext_c[3 * j + 0], ext_c[3 * j + 1], ext_c[3 * j + 2]};
          const FP* sd = std_ + ((int64_t)ti * nnei + s) * 4;
          const FP* n = nd + ((int64_t)i * nnei + s) * 4;
          const FP nn[4] = {n[0], n[1], n[2], n[3]}, ss[4] = {sd[0], sd[1], sd[2], sd[3]};
          term(ci, xi0, xi1, xi2, j, xj, nn, ss, ok ? (int)cand_slot[(int64_t)i * nnei + s] : -1);
        }
        centre(i, ci);
      }
    }
    __syncthreads();
    if (ok)
      for (int g = tid; g < T; g += blockDim.x) {
        int t;
        const int j = nb_candidate(g, nnb, s_base, s_local, s_pref, cell_atoms, &t);
        const int o = mapping ? mapping[j] : j;
        // [offset][cell-sorted position][3]: the candidates of one neighbor
        // cell (local, or a ghost image in the same member order) are
        // consecutive positions, so each run of the table is one contiguous write
        FP* dst = partial + ((int64_t)t * nloc + pos_of[o]) * 3;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          FP x;
          if constexpr (kFixed) {
            x = (FP)((double)(long long)accq[3 * g + k] * kUnfix);
          } else {
            x = acc[3 * g + k];
#pragma unroll
            for (int ww = 1; ww < kAtomsPerBlock; ++ww) x += acc[ww * 3 * kCap + 3 * g + k];
          }
          dst[k] = x;
        }
      }
    __syncthreads();
  }
  block_virial_partial<FP>(v, vpartial);
}

// Pass 2: thread per position of the cell-sorted atom order (consecutive
// threads read consecutive partials); the neighbor-cell index t of a cell's
// table holding this atom's image identifies that cell (c - offset(t),
// wrapped), so the atom sums its partial[t][pos] over t in a fixed order.
template <typename FP>
__global__ void __launch_bounds__(256)
    k_force_gather(double* __restrict__ force, const FP* __restrict__ partial, const int* __restrict__ cell_ok,
                   const char* __restrict__ nlws, int nloc) {
  const GridParams p = *reinterpret_cast<const GridParams*>(nlws);
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= nloc) return;
  // a workspace of another system, or a neighbor list whose search reported
  // an error (status words of dpmi_nlist_status): NaN forces
  const int* misc = reinterpret_cast<const int*>(nlws + ws_layout(0, 0, 0).misc);
  if (p.magic != kGridMagic || p.nloc != nloc || (misc[1] | misc[4]) != 0) {
    const double nan = __builtin_nan("");
    force[3 * (int64_t)pos + 0] = nan;
    force[3 * (int64_t)pos + 1] = nan;
    force[3 * (int64_t)pos + 2] = nan;
    return;
  }
  const WsLayout L = ws_layout(p.nloc, p.lay_nlc, p.lay_next);
  const int* __restrict__ cell_of = reinterpret_cast<const int*>(nlws + L.cell_of);
  const int o = reinterpret_cast<const int*>(nlws + L.cell_atoms)[pos];
  const int ns = p.nsearch, side = 2 * ns + 1, nnb = side * side * side;
  const int c = cell_of[o];
  const int c2 = c % p.ncell[2], c1 = (c / p.ncell[2]) % p.ncell[1], c0 = c / (p.ncell[1] * p.ncell[2]);
  double f[3] = {0., 0., 0.};
  // the neighbor cell that holds this atom in its table at offset t, or -1
  // when no table holds it there
  auto holder = [&](int t) -> int {
    const int o0 = t / (side * side) - ns, o1 = (t / side) % side - ns, o2 = t % side - ns;
    int q0 = c0 - o0, q1 = c1 - o1, q2 = c2 - o2;
    if (p.pbc) {
      q0 += pbc_shift(q0, p.ncell[0]) * p.ncell[0];
      q1 += pbc_shift(q1, p.ncell[1]) * p.ncell[1];
      q2 += pbc_shift(q2, p.ncell[2]) * p.ncell[2];
      const int i0 = q0 + o0, i1 = q1 + o1, i2 = q2 + o2;
      if (i0 < -p.ngcell[0] || i0 >= p.ncell[0] + p.ngcell[0] || i1 < -p.ngcell[1] ||
          i1 >= p.ncell[1] + p.ngcell[1] || i2 < -p.ngcell[2] || i2 >= p.ncell[2] + p.ngcell[2])
        return -1;
    } else if (q0 < 0 || q0 >= p.ncell[0] || q1 < 0 || q1 >= p.ncell[1] || q2 < 0 || q2 >= p.ncell[2]) {
      return -1;
    }
    return (q0 * p.ncell[1] + q1) * p.ncell[2] + q2;
  };
  if (nnb == 27) {
    // every flag and partial load issued before the sums (a partial slot
    // that no table wrote holds stale data and is selected away, not added);
    // the sums run in the same fixed order t = 0 .. 26
    int okv[27];
    FP x[27][3];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int q = holder(t);
      okv[t] = q >= 0 ? cell_ok[q] : 0;
      const FP* src = partial + ((int64_t)t * nloc + pos) * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) x[t][k] = src[k];
    }
#pragma unroll
    for (int t = 0; t < 27; ++t)
      if (okv[t])
#pragma unroll
        for (int k = 0; k < 3; ++k) f[k] += (double)x[t][k];
    force[3 * (int64_t)o + 0] += f[0];
    force[3 * (int64_t)o + 1] += f[1];
    force[3 * (int64_t)o + 2] += f[2];
    return;
  }
  for (int t = 0; t < nnb; ++t) {
    const int o0 = t / (side * side) - ns, o1 = (t / side) % side - ns, o2 = t % side - ns;
    int q0 = c0 - o0, q1 = c1 - o1, q2 = c2 - o2;
    if (p.pbc) {
      q0 += pbc_shift(q0, p.ncell[0]) * p.ncell[0];
      q1 += pbc_shift(q1, p.ncell[1]) * p.ncell[1];
      q2 += pbc_shift(q2, p.ncell[2]) * p.ncell[2];
      // cell q's table holds offset t only when q + offset(t) lies in the
      // ghost range (nb_table); with nsearch > ngcell some offsets fall
      // outside it and their partial slots are never written
      const int i0 = q0 + o0, i1 = q1 + o1, i2 = q2 + o2;
      if (i0 < -p.ngcell[0] || i0 >= p.ncell[0] + p.ngcell[0] || i1 < -p.ngcell[1] ||
          i1 >= p.ncell[1] + p.ngcell[1] || i2 < -p.ngcell[2] || i2 >= p.ncell[2] + p.ngcell[2])
        continue;
    } else if (q0 < 0 || q0 >= p.ncell[0] || q1 < 0 || q1 >= p.ncell[1] || q2 < 0 || q2 >= p.ncell[2]) {
      continue;
    }
    if (!cell_ok[(q0 * p.ncell[1] + q1) * p.ncell[2] + q2]) continue;
    const FP* src = partial + ((int64_t)t * nloc + pos) * 3;
    f[0] += (double)src[0];
    f[1] += (double)src[1];
    f[2] += (double)src[2];
  }
  force[3 * (int64_t)o + 0] += f[0];
  force[3 * (int64_t)o + 1] += f[1];
  force[3 * (int64_t)o + 2] += f[2];
}

template <typename FP, typename CT>
int env_mat_impl(FP* em, FP* em_deriv, FP* rij, const int* nlist, const CT* ext_coord, const int* ext_type,
                 int64_t nloc, const FP* avg, const FP* std_, float rcut, float rcut_smth, const int* sel_host,
                 int ntypes, void* stream) {
  SecParams sp;
  int rc = make_sec(&sp, sel_host, ntypes);
  if (rc) return rc;
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  DPMI_CHECK_ARG(em != nullptr, "em is NULL");
  if (nloc == 0) return DPMI_OK;
  hipLaunchKernelGGL((k_env_mat_a<FP, CT>), dim3(ceil_div(nloc, kAtomsPerBlock)), dim3(256), 0, as_stream(stream), em,
                     em_deriv, rij, nlist, ext_coord, ext_type, (int)nloc, avg, std_, rcut_smth, rcut, sp);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename FP>
int prod_force_impl(FP* force, const FP* nd, const FP* ed, const int* nlist, int64_t nloc, int64_t nall, int64_t nnei,
                    int64_t nframes, void* stream) {
  DPMI_CHECK_ARG(nloc >= 0 && nall >= nloc && nnei >= 0 && nframes >= 0, "bad sizes");
  hipStream_t st = as_stream(stream);
  int64_t nf = nframes * nall * 3;
  if (nf > 0) hipLaunchKernelGGL(k_fill<FP>, dim3(ceil_div(nf, 256)), dim3(256), 0, st, force, nf, (FP)0.);
  int64_t natoms = nloc * nframes;
  if (natoms > 0)
    hipLaunchKernelGGL(k_prod_force<FP>, dim3(ceil_div(natoms, kAtomsPerBlock)), dim3(256), 0, st, force, nd, ed,
                       nlist, (int)nloc, (int)nall, (int)nnei, (int)nframes);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

int64_t virial_ws(int64_t nloc, int64_t nframes) {
  return std::max<int64_t>(1, nframes) * std::max<int64_t>(1, ceil_div(nloc, kAtomsPerBlock)) * 9 * 8;
}

template <typename FP>
int prod_virial_impl(FP* virial, FP* atom_virial, const FP* nd, const FP* ed, const FP* rij, const int* nlist,
                     int64_t nloc, int64_t nall, int64_t nnei, int64_t nframes, void* ws, int64_t ws_bytes,
                     void* stream) {
  DPMI_CHECK_ARG(nloc >= 0 && nall >= nloc && nnei >= 0 && nframes >= 0, "bad sizes");
  DPMI_CHECK_ARG(ws != nullptr && ws_bytes >= virial_ws(nloc, nframes), "workspace too small");
  hipStream_t st = as_stream(stream);
  if (nframes == 0) return DPMI_OK;
  if (atom_virial) {
    int64_t na = nframes * nall * 9;
    if (na > 0) hipLaunchKernelGGL(k_fill<FP>, dim3(ceil_div(na, 256)), dim3(256), 0, st, atom_virial, na, (FP)0.);
  }
  int bpf = (int)std::max<int64_t>(1, ceil_div(nloc, kAtomsPerBlock));
  double* partial = static_cast<double*>(ws);
  hipLaunchKernelGGL(k_prod_virial<FP>, dim3(bpf * nframes), dim3(256), 0, st, partial, atom_virial, nd, ed, rij,
                     nlist, (int)nloc, (int)nall, (int)nnei, bpf);
  hipLaunchKernelGGL(k_virial_reduce<FP>, dim3(nframes), dim3(576), 0, st, virial, partial, bpf);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename FP>
int prod_force_grad_impl(FP* gn, const FP* grad, const FP* ed, const int* nlist, int64_t nloc, int64_t nall,
                         int64_t nnei, int64_t nframes, void* stream) {
  DPMI_CHECK_ARG(nloc >= 0 && nall >= nloc && nall < (int64_t(1) << 31) && nnei >= 0 && nframes >= 0,
                 "bad sizes");
  const int64_t n = nframes * nloc * nnei;
  if (n == 0) return DPMI_OK;
  DPMI_CHECK_ARG(gn && grad && ed && nlist, "NULL argument");
  hipLaunchKernelGGL(k_prod_force_grad<FP>, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), gn, grad, ed,
                     nlist, n, (int)nloc, (int)nall, (int)nnei);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename FP>
int prod_virial_grad_impl(FP* gn, const FP* grad, const FP* ed, const FP* rij, const int* nlist, int64_t nloc,
                          int64_t nnei, int64_t nframes, void* stream) {
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31) && nnei >= 0 && nframes >= 0, "bad sizes");
  const int64_t n = nframes * nloc * nnei;
  if (n == 0) return DPMI_OK;
  DPMI_CHECK_ARG(gn && grad && ed && rij && nlist, "NULL argument");
  hipLaunchKernelGGL(k_prod_virial_grad<FP>, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), gn, grad, ed,
                     rij, nlist, n, (int)nloc, (int)nnei);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_prod_force_grad_a_f64(double* grad_net, const double* grad, const double* in_deriv,
                                          const int* nlist, int64_t nloc, int64_t nall, int64_t nnei,
                                          int64_t nframes, void* stream) {
  return prod_force_grad_impl<double>(grad_net, grad, in_deriv, nlist, nloc, nall, nnei, nframes, stream);
}

extern "C" int dpmi_prod_force_grad_a_f32(float* grad_net, const float* grad, const float* in_deriv,
                                          const int* nlist, int64_t nloc, int64_t nall, int64_t nnei,
                                          int64_t nframes, void* stream) {
  return prod_force_grad_impl<float>(grad_net, grad, in_deriv, nlist, nloc, nall, nnei, nframes, stream);
}

extern "C" int dpmi_prod_virial_grad_a_f64(double* grad_net, const double* grad, const double* in_deriv,
                                           const double* rij, const int* nlist, int64_t nloc, int64_t nnei,
                                           int64_t nframes, void* stream) {
  return prod_virial_grad_impl<double>(grad_net, grad, in_deriv, rij, nlist, nloc, nnei, nframes, stream);
}

extern "C" int dpmi_prod_virial_grad_a_f32(float* grad_net, const float* grad, const float* in_deriv,
                                           const float* rij, const int* nlist, int64_t nloc, int64_t nnei,
                                           int64_t nframes, void* stream) {
  return prod_virial_grad_impl<float>(grad_net, grad, in_deriv, rij, nlist, nloc, nnei, nframes, stream);
}

extern "C" int dpmi_prod_env_mat_a_f64(double* em, double* em_deriv, double* rij, const int* nlist,
                                       const double* ext_coord, const int* ext_type, int64_t nloc, const double* avg,
                                       const double* std_, float rcut, float rcut_smth, const int* sel_host,
                                       int ntypes, void* stream) {
  return env_mat_impl<double, double>(em, em_deriv, rij, nlist, ext_coord, ext_type, nloc, avg, std_, rcut,
                                      rcut_smth, sel_host, ntypes, stream);
}

extern "C" int dpmi_prod_env_mat_a_f32(float* em, float* em_deriv, float* rij, const int* nlist,
                                       const float* ext_coord, const int* ext_type, int64_t nloc, const float* avg,
                                       const float* std_, float rcut, float rcut_smth, const int* sel_host,
                                       int ntypes, void* stream) {
  return env_mat_impl<float, float>(em, em_deriv, rij, nlist, ext_coord, ext_type, nloc, avg, std_, rcut,
                                    rcut_smth, sel_host, ntypes, stream);
}

extern "C" int dpmi_prod_env_mat_a_f32c64(float* em, float* em_deriv, float* rij, const int* nlist,
                                          const double* ext_coord, const int* ext_type, int64_t nloc,
                                          const float* avg, const float* std_, float rcut, float rcut_smth,
                                          const int* sel_host, int ntypes, void* stream) {
  return env_mat_impl<float, double>(em, em_deriv, rij, nlist, ext_coord, ext_type, nloc, avg, std_, rcut,
                                     rcut_smth, sel_host, ntypes, stream);
}

extern "C" int dpmi_prod_force_se_a_f64(double* force, const double* net_deriv, const double* in_deriv,
                                        const int* nlist, int64_t nloc, int64_t nall, int64_t nnei, int64_t nframes,
                                        void* stream) {
  return prod_force_impl<double>(force, net_deriv, in_deriv, nlist, nloc, nall, nnei, nframes, stream);
}

extern "C" int dpmi_prod_force_se_a_f32(float* force, const float* net_deriv, const float* in_deriv,
                                        const int* nlist, int64_t nloc, int64_t nall, int64_t nnei, int64_t nframes,
                                        void* stream) {
  return prod_force_impl<float>(force, net_deriv, in_deriv, nlist, nloc, nall, nnei, nframes, stream);
}

extern "C" int dpmi_prod_virial_workspace_bytes(int64_t* bytes, int64_t nloc, int64_t nframes) {
  DPMI_CHECK_ARG(bytes != nullptr, "bytes is NULL");
  *bytes = virial_ws(nloc, nframes);
  return DPMI_OK;
}

extern "C" int dpmi_prod_virial_se_a_f64(double* virial, double* atom_virial, const double* net_deriv,
                                         const double* in_deriv, const double* rij, const int* nlist, int64_t nloc,
                                         int64_t nall, int64_t nnei, int64_t nframes, void* workspace,
                                         int64_t workspace_bytes, void* stream) {
  return prod_virial_impl<double>(virial, atom_virial, net_deriv, in_deriv, rij, nlist, nloc, nall, nnei, nframes,
                                  workspace, workspace_bytes, stream);
}

extern "C" int dpmi_prod_virial_se_a_f32(float* virial, float* atom_virial, const float* net_deriv,
                                         const float* in_deriv, const float* rij, const int* nlist, int64_t nloc,
                                         int64_t nall, int64_t nnei, int64_t nframes, void* workspace,
                                         int64_t workspace_bytes, void* stream) {
  return prod_virial_impl<float>(virial, atom_virial, net_deriv, in_deriv, rij, nlist, nloc, nall, nnei, nframes,
                                 workspace, workspace_bytes, stream);
}

// Scratch of dpmi_prod_virial_se_a_noalloc_f64: the per-workgroup virial
// partials (9 doubles per 4 local atoms per frame) that the workspace of
// dpmi_prod_virial_se_a_f64 holds, in device memory the library owns.  Every
// GPU the library runs on keeps its own copy until the process ends; every
// call on one GPU shares it, so calls that overlap on one GPU race.  Each
// workgroup writes all of its partials before the reduction reads them, so
// nothing needs clearing between calls.
constexpr int64_t kScratchBytes = int64_t(8) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (the host-side name is only a stand-in for the device array).
int scratch_address(void** out) {
  int dev = 0;
  DPMI_HIP(hipGetDevice(&dev));
  DPMI_CHECK_ARG(dev >= 0 && dev < kMaxScratchDevices, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    DPMI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DPMI_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DPMI_OK;
}

}  // namespace

extern "C" int dpmi_prod_virial_se_a_noalloc_f64(double* virial, double* atom_virial, const double* net_deriv,
                                                 const double* in_deriv, const double* rij, const int* nlist,
                                                 int64_t nloc, int64_t nall, int64_t nnei, int64_t nframes,
                                                 void* stream) {
  DPMI_CHECK_ARG(nloc >= 0 && nall >= nloc && nnei >= 0 && nframes >= 0, "bad sizes");
  const int64_t need = virial_ws(nloc, nframes);
  if (need > kScratchBytes) {
    set_error("dpmi_prod_virial_se_a_noalloc_f64: nframes * ceil(nloc / 4) * 72 bytes = " + std::to_string(need) +
              " exceeds the library's scratch of " + std::to_string(kScratchBytes) +
              " bytes (at most 116,508 four-atom groups over all frames); use dpmi_prod_virial_se_a_f64 with a "
              "caller workspace");
    return DPMI_ERR_ARG;
  }
  void* ws = nullptr;
  int rc = scratch_address(&ws);
  if (rc) return rc;
  return prod_virial_impl<double>(virial, atom_virial, net_deriv, in_deriv, rij, nlist, nloc, nall, nnei, nframes,
                                  ws, kScratchBytes, stream);
}

constexpr int64_t kForceBlocks = 2048;

template <typename FP>
static int force_virial_recompute_impl(double* force, double* virial, const FP* net_deriv, const int* nlist,
                                       const int* mapping, const double* ext_coord, const int* ext_type,
                                       int64_t nloc, int64_t nall, const FP* std_, float rcut, float rcut_smth,
                                       const int* sel_host, int ntypes, void* workspace, int64_t workspace_bytes,
                                       void* stream) {
  SecParams sp;
  int rc = make_sec(&sp, sel_host, ntypes);
  if (rc) return rc;
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31) && nall >= nloc, "nloc/nall out of range");
  DPMI_CHECK_ARG(workspace != nullptr && workspace_bytes >= virial_ws(nloc, 1), "workspace too small");
  hipStream_t st = as_stream(stream);
  const int64_t nout = (mapping ? nloc : nall) * 3;
  if (nout > 0) hipLaunchKernelGGL(k_fill<double>, dim3(ceil_div(nout, 256)), dim3(256), 0, st, force, nout, 0.);
  // persistent grid (a few workgroups per CU): few virial partials to reduce
  int bpf = (int)std::min<int64_t>(kForceBlocks, std::max<int64_t>(1, ceil_div(nloc, kAtomsPerBlock)));
  double* partial = static_cast<double*>(workspace);
  hipLaunchKernelGGL((k_force_virial_recompute<FP, double>), dim3(bpf), dim3(256), 0, st, force, partial, net_deriv,
                     nlist, mapping, ext_coord, ext_type, (int)nloc, std_, rcut_smth, rcut, sp);
  if (virial) hipLaunchKernelGGL(k_virial_reduce<double>, dim3(1), dim3(576), 0, st, virial, partial, bpf);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_force_virial_recompute_f64(double* force, double* virial, const double* net_deriv,
                                               const int* nlist, const int* mapping, const double* ext_coord,
                                               const int* ext_type, int64_t nloc, int64_t nall, const double* std_,
                                               float rcut, float rcut_smth, const int* sel_host, int ntypes,
                                               void* workspace, int64_t workspace_bytes, void* stream) {
  return force_virial_recompute_impl<double>(force, virial, net_deriv, nlist, mapping, ext_coord, ext_type, nloc,
                                             nall, std_, rcut, rcut_smth, sel_host, ntypes, workspace,
                                             workspace_bytes, stream);
}

extern "C" int dpmi_force_virial_recompute_f32(double* force, double* virial, const float* net_deriv,
                                               const int* nlist, const int* mapping, const double* ext_coord,
                                               const int* ext_type, int64_t nloc, int64_t nall, const float* std_,
                                               float rcut, float rcut_smth, const int* sel_host, int ntypes,
                                               void* workspace, int64_t workspace_bytes, void* stream) {
  return force_virial_recompute_impl<float>(force, virial, net_deriv, nlist, mapping, ext_coord, ext_type, nloc,
                                            nall, std_, rcut, rcut_smth, sel_host, ntypes, workspace,
                                            workspace_bytes, stream);
}

namespace dpmi {
namespace {

constexpr int kBigBlocks = 512;  // workgroups of the second (big-cell) force pass

int64_t force_cells_ws(const double* box_host, int64_t nloc, float rcut, int64_t elem, int64_t* nnb_out) {
  int64_t nlc;
  int ns;
  if (cell_grid_info(box_host, nloc, rcut, &nlc, &ns) != DPMI_OK) return -1;
  const int64_t side = 2 * ns + 1, nnb = side * side * side;
  if (nnb_out) *nnb_out = nnb;
  const int64_t na = std::max<int64_t>(nloc, 1);
  return align256(na * nnb * 3 * elem) + align256(4 * nlc) + align256((kForceBlocks + kBigBlocks) * 9 * 8);
}

template <typename FP>
int force_cells_impl(double* force, double* virial, const FP* net_deriv, const int* nlist, const int16_t* cand_slot,
                     const int* mapping, const double* ext_coord, const int* ext_type, int64_t nloc, int64_t nall,
                     const FP* std_, float rcut, float rcut_smth, const int* sel_host, int ntypes,
                     const double* box_host, const void* nlist_workspace, void* workspace, int64_t workspace_bytes,
                     void* stream) {
  SecParams sp;
  int rc = make_sec(&sp, sel_host, ntypes);
  if (rc) return rc;
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31) && nall >= nloc, "nloc/nall out of range");
  DPMI_CHECK_ARG(mapping != nullptr || nall == nloc, "ghost atoms need mapping");
  DPMI_CHECK_ARG(cand_slot != nullptr && nlist_workspace != nullptr, "cand_slot and the neighbor-list workspace are required");
  int64_t nnb = 0;
  const int64_t need = force_cells_ws(box_host, nloc, rcut, sizeof(FP), &nnb);
  DPMI_CHECK_ARG(need > 0 && workspace != nullptr && workspace_bytes >= need,
                 "workspace too small (dpmi_force_virial_cells_workspace_bytes)");
  hipStream_t st = as_stream(stream);
  if (nloc > 0) hipLaunchKernelGGL(k_fill<double>, dim3(ceil_div(nloc * 3, 256)), dim3(256), 0, st, force, nloc * 3, 0.);
  char* ws = static_cast<char*>(workspace);
  FP* partial = reinterpret_cast<FP*>(ws);
  int* cell_ok = reinterpret_cast<int*>(ws + align256(std::max<int64_t>(nloc, 1) * nnb * 3 * sizeof(FP)));
  int64_t nlc;
  int ns;
  cell_grid_info(box_host, nloc, rcut, &nlc, &ns);
  double* vpartial = reinterpret_cast<double*>(reinterpret_cast<char*>(cell_ok) + align256(4 * nlc));
  if (!search_uses_cells(sp.nnei)) {
    // the search went atom by atom (rows too long for its candidate tables):
    // every cand_slot is -1 and no big-cell list exists, so the per-atom
    // adjoint takes every neighbor term
    const int bpa = (int)std::min<int64_t>(kForceBlocks, std::max<int64_t>(1, ceil_div(nloc, kAtomsPerBlock)));
    if (nloc > 0)
      hipLaunchKernelGGL((k_force_virial_recompute<FP, double>), dim3(bpa), dim3(256), 0, st, force, vpartial,
                         net_deriv, nlist, mapping, ext_coord, ext_type, (int)nloc, std_, rcut_smth, rcut, sp);
    else
      hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(64), 0, st, vpartial, (int64_t)9, 0.);
    if (virial)
      hipLaunchKernelGGL(k_virial_reduce<double>, dim3(1), dim3(576), 0, st, virial, vpartial, nloc > 0 ? bpa : 1);
    DPMI_LAUNCH_CHECK();
    return DPMI_OK;
  }
  const int bpf = (int)std::min<int64_t>(kForceBlocks, std::max<int64_t>(1, nloc));
  // LDS table: fp32 one fixed-point copy (8 B per component), fp64 one copy per wavefront
  const size_t per_slot = sizeof(FP) == 4 ? 3 * sizeof(unsigned long long) : kAtomsPerBlock * 3 * sizeof(FP);
  const size_t shm = (size_t)kCandCap * per_slot;
  const size_t shm_big = (size_t)kCandCapBig * per_slot;
  if (nloc > 0) {
    // a row of up to 3 (4) chunks of 64 neighbors is loaded whole per atom
    auto launch = [&](auto chunks) -> int {
      constexpr int kCh = decltype(chunks)::value;
      if (shm > 65536)
        DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_force_cells<FP, kCandCap, false, kCh>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      hipLaunchKernelGGL((k_force_cells<FP, kCandCap, false, kCh>), dim3(bpf), dim3(256), shm, st, force, partial,
                         cell_ok, vpartial, net_deriv, nlist, cand_slot, mapping, ext_coord, ext_type, std_,
                         static_cast<const char*>(nlist_workspace), rcut_smth, rcut, sp, (int)nloc);
      if constexpr (sizeof(FP) == 4) {
        // the search's big cells (second table size); its virial partials follow the first pass's
        DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_force_cells<FP, kCandCapBig, true, kCh>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm_big));
        hipLaunchKernelGGL((k_force_cells<FP, kCandCapBig, true, kCh>), dim3(kBigBlocks), dim3(256), shm_big, st,
                           force, partial, cell_ok, vpartial + 9 * bpf, net_deriv, nlist, cand_slot, mapping,
                           ext_coord, ext_type, std_, static_cast<const char*>(nlist_workspace), rcut_smth, rcut, sp,
                           (int)nloc);
      }
      return DPMI_OK;
    };
    rc = sp.nnei <= 3 * kWave ? launch(std::integral_constant<int, 3>{}) : launch(std::integral_constant<int, 4>{});
    if (rc) return rc;
    hipLaunchKernelGGL((k_force_gather<FP>), dim3(ceil_div(nloc, 256)), dim3(256), 0, st, force, partial, cell_ok,
                       static_cast<const char*>(nlist_workspace), (int)nloc);
  } else {
    hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(64), 0, st, vpartial, (int64_t)9, 0.);
  }
  const int nparts = nloc > 0 ? bpf + (sizeof(FP) == 4 ? kBigBlocks : 0) : 1;
  if (virial) hipLaunchKernelGGL(k_virial_reduce<double>, dim3(1), dim3(576), 0, st, virial, vpartial, nparts);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

}  // namespace
}  // namespace dpmi

extern "C" int dpmi_force_virial_cells_workspace_bytes(int64_t* bytes, const double* box_host, int64_t nloc,
                                                       float rcut, int fp64) {
  DPMI_CHECK_ARG(bytes != nullptr && nloc >= 0, "bad arguments");
  const int64_t b = force_cells_ws(box_host, nloc, rcut, fp64 ? 8 : 4, nullptr);
  DPMI_CHECK_ARG(b > 0, "degenerate box");
  *bytes = b;
  return DPMI_OK;
}

extern "C" int dpmi_force_virial_cells_f64(double* force, double* virial, const double* net_deriv, const int* nlist,
                                           const int16_t* cand_slot, const int* mapping, const double* ext_coord,
                                           const int* ext_type, int64_t nloc, int64_t nall, const double* std_,
                                           float rcut, float rcut_smth, const int* sel_host, int ntypes,
                                           const double* box_host, const void* nlist_workspace, void* workspace,
                                           int64_t workspace_bytes, void* stream) {
  return force_cells_impl<double>(force, virial, net_deriv, nlist, cand_slot, mapping, ext_coord, ext_type, nloc,
                                  nall, std_, rcut, rcut_smth, sel_host, ntypes, box_host, nlist_workspace,
                                  workspace, workspace_bytes, stream);
}

extern "C" int dpmi_force_virial_cells_f32(double* force, double* virial, const float* net_deriv, const int* nlist,
                                           const int16_t* cand_slot, const int* mapping, const double* ext_coord,
                                           const int* ext_type, int64_t nloc, int64_t nall, const float* std_,
                                           float rcut, float rcut_smth, const int* sel_host, int ntypes,
                                           const double* box_host, const void* nlist_workspace, void* workspace,
                                           int64_t workspace_bytes, void* stream) {
  return force_cells_impl<float>(force, virial, net_deriv, nlist, cand_slot, mapping, ext_coord, ext_type, nloc,
                                 nall, std_, rcut, rcut_smth, sel_host, ntypes, box_host, nlist_workspace, workspace,
                                 workspace_bytes, stream);
}
