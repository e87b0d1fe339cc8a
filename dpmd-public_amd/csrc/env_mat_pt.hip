// Environment matrix with the semantics of the PyTorch backend's prod_env_mat
// (deepmd/pt/model/descriptor/env_mat.py:11-92, smooth weights of
// deepmd/pt/utils/preprocess.py:9-34), on gfx950, and the two derivative
// products its autograd needs: the coordinate gradient (vector-Jacobian
// product) and, for the double backward of an energy + force loss, the
// product of the env-mat Jacobian with a coordinate-space vector.
//
// Unlike the C++ env-mat (env_mat.hip), the list here is whatever the caller
// passes: entries in any order, -1 padding anywhere, neighbors beyond rcut
// (their smooth weight is 0, so their rows hold -mean / std and only diff
// carries them).  Per slot, with d = x_j - x_i, r = |d|, q = r + protection
// and w = sw(r):
//   env = ([1 / q, d / q^2] * w - mean[t_i]) / std[t_i]   (radial: [1 / q] * w)
//   diff = d,  switch = w                                  (padding: 0, 0)
//
// Mapping: one 64-wide wavefront per (frame, local atom) row, four rows per
// 256-thread workgroup, a lane per neighbor slot; the slot's env values are
// contiguous, so every store instruction of a wave covers one contiguous
// span of the row.  The coordinate gradient sums the centre's terms over the
// wavefront (one atomic add per component) and scatters each neighbor's term
// with an atomic add, as the reference's autograd scatter does; its
// summation order, and so its last bits, is therefore not fixed.
#include <hip/hip_runtime.h>

#include <cmath>

#include "dpmi_internal.h"

namespace dpmi {
namespace {

constexpr int kRowsPerBlock = 4;
constexpr int64_t kMaxBlocks = 16384;

template <typename FP>
struct PtEnv {
  FP rmin, rmax, range, protection;
  int exp_switch;
  int ncomp;  // 4, or 1 for the radial-only env-mat
  int nnei, nloc, nall, ntypes;
};

// Smooth weight w(r) and dw/dr: compute_smooth_weight (clamp to [rmin,
// rmax], u = (r - rmin) / (rmax - rmin), w = u^3 (-6 u^2 + 15 u - 10) + 1) or
// compute_exp_sw (clamp to [0, rmax], w = exp(-exp(a (r - rmin))),
// a = 20 / rmin).  The clamp makes dw/dr zero outside the interval.
template <typename FP>
__device__ inline void pt_weight(const PtEnv<FP>& p, FP r, FP* w, FP* dw) {
  if (p.exp_switch) {
    const FP a = (FP)20. / p.rmin;
    const FP rr = r < (FP)0. ? (FP)0. : (r > p.rmax ? p.rmax : r);
    const FP e = exp(a * (rr - p.rmin));
    *w = exp(-e);
    *dw = (r >= (FP)0. && r <= p.rmax) ? -(*w) * e * a : (FP)0.;
  } else {
    const FP rr = r < p.rmin ? p.rmin : (r > p.rmax ? p.rmax : r);
    const FP u = (rr - p.rmin) / p.range;
    const FP u2 = u * u;
    *w = u2 * u * ((FP)-6. * u2 + (FP)15. * u - (FP)10.) + (FP)1.;
    *dw = (r >= p.rmin && r <= p.rmax)
              ? ((FP)3. * u2 * ((FP)-6. * u2 + (FP)15. * u - (FP)10.) + u2 * u * ((FP)-12. * u + (FP)15.)) / p.range
              : (FP)0.;
  }
}

__device__ inline int wave_in_block() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

template <typename FP>
__global__ void __launch_bounds__(256)
    k_env_pt_fwd(FP* __restrict__ env, FP* __restrict__ diff, FP* __restrict__ sw, const FP* __restrict__ coord,
                 const int* __restrict__ nlist, const int* __restrict__ atype, const FP* __restrict__ mean,
                 const FP* __restrict__ stdv, PtEnv<FP> p, int64_t rows) {
  const int lane = lane_id();
  const FP nan = (FP)__builtin_nan("");
  for (int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + wave_in_block(); row < rows;
       row += (int64_t)gridDim.x * kRowsPerBlock) {
    const int64_t f = row / p.nloc, i = row % p.nloc;
    const FP* cf = coord + f * (int64_t)p.nall * 3;
    const FP xi0 = cf[3 * i + 0], xi1 = cf[3 * i + 1], xi2 = cf[3 * i + 2];
    const int ti = atype[row];
    const bool tok = ti >= 0 && ti < p.ntypes;
    for (int s = lane; s < p.nnei; s += kWave) {
      const int64_t o = row * p.nnei + s;
      const int j = nlist[o];
      const bool real = j >= 0 && j < p.nall;
      FP d[3] = {(FP)0., (FP)0., (FP)0.}, e[4] = {(FP)0., (FP)0., (FP)0., (FP)0.}, w = (FP)0.;
      if (real) {
        d[0] = cf[3 * (int64_t)j + 0] - xi0;
        d[1] = cf[3 * (int64_t)j + 1] - xi1;
        d[2] = cf[3 * (int64_t)j + 2] - xi2;
        const FP r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        FP dw;
        pt_weight(p, r, &w, &dw);
        const FP q = r + p.protection;
        e[0] = (FP)1. / q * w;
        const FP q2 = q * q;
        e[1] = d[0] / q2 * w;
        e[2] = d[1] / q2 * w;
        e[3] = d[2] / q2 * w;
      }
      // an index beyond the extended atoms, or a centre type outside
      // [0, ntypes), has no defined env-mat: NaN marks it
      const bool bad = !tok || j >= p.nall;
      const int64_t so = ((int64_t)(tok ? ti : 0) * p.nnei + s) * p.ncomp;
      FP* eo = env + o * p.ncomp;
      if (p.ncomp == 4) {
#pragma unroll
        for (int c = 0; c < 4; ++c) eo[c] = bad ? nan : (e[c] - mean[so + c]) / stdv[so + c];
      } else {
        eo[0] = bad ? nan : (e[0] - mean[so]) / stdv[so];
      }
      if (diff) {
        diff[o * 3 + 0] = bad ? nan : d[0];
        diff[o * 3 + 1] = bad ? nan : d[1];
        diff[o * 3 + 2] = bad ? nan : d[2];
      }
      if (sw) sw[o] = bad ? nan : w;
    }
  }
}

// g_coord[f, j] += g(slot), g_coord[f, i] -= g(slot), with
// g = dL/dd = sum_c a_c dE_c/dd + g_diff + g_sw dw/dr d/r, a_c = g_env_c / std_c:
//   dE_0/dd = (w'/q - w/q^2) d/r
//   dE_k/dd = w/q^2 e_k + d_k (w'/q^2 - 2 w/q^3) d/r
// Padding slots carry no gradient (their rows are constants).
template <typename FP>
__global__ void __launch_bounds__(256)
    k_env_pt_vjp(FP* __restrict__ gcoord, const FP* __restrict__ g_env, const FP* __restrict__ g_diff,
                 const FP* __restrict__ g_sw, const FP* __restrict__ coord, const int* __restrict__ nlist,
                 const int* __restrict__ atype, const FP* __restrict__ stdv, PtEnv<FP> p, int64_t rows) {
  const int lane = lane_id();
  for (int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + wave_in_block(); row < rows;
       row += (int64_t)gridDim.x * kRowsPerBlock) {
    const int64_t f = row / p.nloc, i = row % p.nloc;
    const FP* cf = coord + f * (int64_t)p.nall * 3;
    FP* gf = gcoord + f * (int64_t)p.nall * 3;
    const FP xi0 = cf[3 * i + 0], xi1 = cf[3 * i + 1], xi2 = cf[3 * i + 2];
    const int ti = atype[row];
    const bool tok = ti >= 0 && ti < p.ntypes;
    FP ci[3] = {(FP)0., (FP)0., (FP)0.};
    for (int s = lane; s < p.nnei; s += kWave) {
      const int64_t o = row * p.nnei + s;
      const int j = nlist[o];
      if (j < 0) continue;
      if (j >= p.nall || !tok) {
        ci[0] = (FP)__builtin_nan("");
        continue;
      }
      const FP d[3] = {cf[3 * (int64_t)j + 0] - xi0, cf[3 * (int64_t)j + 1] - xi1, cf[3 * (int64_t)j + 2] - xi2};
      const FP r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      FP w, dw;
      pt_weight(p, r, &w, &dw);
      const FP q = r + p.protection, q2 = q * q;
      const int64_t so = ((int64_t)ti * p.nnei + s) * p.ncomp;
      FP radial = (FP)0., g[3] = {(FP)0., (FP)0., (FP)0.};
      if (g_env) {
        const FP* ge = g_env + o * p.ncomp;
        const FP a0 = ge[0] / stdv[so];
        radial += a0 * (dw / q - w / q2);
        if (p.ncomp == 4) {
          const FP a[3] = {ge[1] / stdv[so + 1], ge[2] / stdv[so + 2], ge[3] / stdv[so + 3]};
          radial += (a[0] * d[0] + a[1] * d[1] + a[2] * d[2]) * (dw / q2 - (FP)2. * w / (q2 * q));
#pragma unroll
          for (int k = 0; k < 3; ++k) g[k] += w / q2 * a[k];
        }
      }
      if (g_sw) radial += g_sw[o] * dw;
      if (g_diff) {
#pragma unroll
        for (int k = 0; k < 3; ++k) g[k] += g_diff[o * 3 + k];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        g[k] += radial * d[k] / r;
        atomicAdd(&gf[3 * (int64_t)j + k], g[k]);
        ci[k] -= g[k];
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) ci[k] = wave_sum(ci[k]);
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 3; ++k) atomicAdd(&gf[3 * i + k], ci[k]);
  }
}

// Derivative of the coordinate gradient above with respect to (g_env,
// g_diff, g_sw), applied to v = dL/d g_coord: with u = v_j - v_i,
//   dL/dg_env_c = (dE_c/dd . u) / std_c,  dL/dg_diff = u,
//   dL/dg_sw = w' (d . u) / r;  0 on padding slots.
template <typename FP>
__global__ void __launch_bounds__(256)
    k_env_pt_jvp(FP* __restrict__ gg_env, FP* __restrict__ gg_diff, FP* __restrict__ gg_sw,
                 const FP* __restrict__ v, const FP* __restrict__ coord, const int* __restrict__ nlist,
                 const int* __restrict__ atype, const FP* __restrict__ stdv, PtEnv<FP> p, int64_t rows) {
  const int lane = lane_id();
  const FP nan = (FP)__builtin_nan("");
  for (int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + wave_in_block(); row < rows;
       row += (int64_t)gridDim.x * kRowsPerBlock) {
    const int64_t f = row / p.nloc, i = row % p.nloc;
    const FP* cf = coord + f * (int64_t)p.nall * 3;
    const FP* vf = v + f * (int64_t)p.nall * 3;
    const FP xi0 = cf[3 * i + 0], xi1 = cf[3 * i + 1], xi2 = cf[3 * i + 2];
    const FP vi0 = vf[3 * i + 0], vi1 = vf[3 * i + 1], vi2 = vf[3 * i + 2];
    const int ti = atype[row];
    const bool tok = ti >= 0 && ti < p.ntypes;
    for (int s = lane; s < p.nnei; s += kWave) {
      const int64_t o = row * p.nnei + s;
      const int j = nlist[o];
      FP ge[4] = {(FP)0., (FP)0., (FP)0., (FP)0.}, gd[3] = {(FP)0., (FP)0., (FP)0.}, gs = (FP)0.;
      if (j >= 0 && (j >= p.nall || !tok)) {
        ge[0] = ge[1] = ge[2] = ge[3] = gd[0] = gd[1] = gd[2] = gs = nan;
      } else if (j >= 0) {
        const FP d[3] = {cf[3 * (int64_t)j + 0] - xi0, cf[3 * (int64_t)j + 1] - xi1, cf[3 * (int64_t)j + 2] - xi2};
        const FP u[3] = {vf[3 * (int64_t)j + 0] - vi0, vf[3 * (int64_t)j + 1] - vi1, vf[3 * (int64_t)j + 2] - vi2};
        const FP r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        FP w, dw;
        pt_weight(p, r, &w, &dw);
        const FP q = r + p.protection, q2 = q * q;
        const FP du = (d[0] * u[0] + d[1] * u[1] + d[2] * u[2]) / r;
        const int64_t so = ((int64_t)ti * p.nnei + s) * p.ncomp;
        ge[0] = (dw / q - w / q2) * du / stdv[so];
        if (p.ncomp == 4) {
          const FP b = (dw / q2 - (FP)2. * w / (q2 * q)) * du;
#pragma unroll
          for (int k = 0; k < 3; ++k) ge[k + 1] = (w / q2 * u[k] + d[k] * b) / stdv[so + k + 1];
        }
        gd[0] = u[0];
        gd[1] = u[1];
        gd[2] = u[2];
        gs = dw * du;
      }
      if (gg_env) {
        FP* go = gg_env + o * p.ncomp;
        for (int c = 0; c < p.ncomp; ++c) go[c] = ge[c];
      }
      if (gg_diff) {
        gg_diff[o * 3 + 0] = gd[0];
        gg_diff[o * 3 + 1] = gd[1];
        gg_diff[o * 3 + 2] = gd[2];
      }
      if (gg_sw) gg_sw[o] = gs;
    }
  }
}

template <typename FP>
__global__ void k_fill_pt(FP* p, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    p[k] = (FP)0.;
}

template <typename FP>
int make_params(PtEnv<FP>* p, int64_t nframes, int64_t nloc, int64_t nall, int64_t nnei, int ntypes, int ncomp,
                double rcut, double rcut_smth, double protection, int use_exp_switch) {
  DPMI_CHECK_ARG(nframes >= 0 && nloc >= 0 && nall >= nloc && nnei >= 0, "bad sizes (need nall >= nloc >= 0)");
  DPMI_CHECK_ARG(nloc < (int64_t(1) << 31) && nall < (int64_t(1) << 31) && nnei < (int64_t(1) << 20),
                 "sizes out of range");
  DPMI_CHECK_ARG(ntypes > 0, "ntypes must be positive");
  DPMI_CHECK_ARG(ncomp == 4 || ncomp == 1, "ncomp must be 4 (full) or 1 (radial only)");
  // compute_smooth_weight / compute_exp_sw raise for rmin >= rmax
  DPMI_CHECK_ARG(rcut_smth < rcut, "rmin should be less than rmax (rcut_smth < rcut)");
  DPMI_CHECK_ARG(!use_exp_switch || rcut_smth > 0., "the exponential switch needs rcut_smth > 0");
  p->rmin = (FP)rcut_smth;
  p->rmax = (FP)rcut;
  p->range = (FP)(rcut - rcut_smth);
  p->protection = (FP)protection;
  p->exp_switch = use_exp_switch ? 1 : 0;
  p->ncomp = ncomp;
  p->nnei = (int)nnei;
  p->nloc = (int)nloc;
  p->nall = (int)nall;
  p->ntypes = ntypes;
  return DPMI_OK;
}

inline unsigned grid_for(int64_t rows) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, ceil_div(rows, kRowsPerBlock)));
}

template <typename FP>
int env_pt_impl(FP* env, FP* diff, FP* sw, const FP* coord, const int* nlist, const int* atype, const FP* mean,
                const FP* stdv, int64_t nframes, int64_t nloc, int64_t nall, int64_t nnei, int ntypes, int ncomp,
                double rcut, double rcut_smth, double protection, int use_exp_switch, void* stream) {
  PtEnv<FP> p;
  int rc = make_params(&p, nframes, nloc, nall, nnei, ntypes, ncomp, rcut, rcut_smth, protection, use_exp_switch);
  if (rc) return rc;
  const int64_t rows = nframes * nloc;
  if (rows == 0 || nnei == 0) return DPMI_OK;
  DPMI_CHECK_ARG(env && coord && nlist && atype && mean && stdv, "NULL input or env_mat");
  hipLaunchKernelGGL(k_env_pt_fwd<FP>, dim3(grid_for(rows)), dim3(256), 0, as_stream(stream), env, diff, sw, coord,
                     nlist, atype, mean, stdv, p, rows);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename FP>
int env_pt_grad_impl(FP* gcoord, const FP* g_env, const FP* g_diff, const FP* g_sw, const FP* coord,
                     const int* nlist, const int* atype, const FP* stdv, int64_t nframes, int64_t nloc, int64_t nall,
                     int64_t nnei, int ntypes, int ncomp, double rcut, double rcut_smth, double protection,
                     int use_exp_switch, void* stream) {
  PtEnv<FP> p;
  int rc = make_params(&p, nframes, nloc, nall, nnei, ntypes, ncomp, rcut, rcut_smth, protection, use_exp_switch);
  if (rc) return rc;
  DPMI_CHECK_ARG(gcoord != nullptr || nframes * nall == 0, "grad_coord is NULL");
  hipStream_t st = as_stream(stream);
  const int64_t n = nframes * nall * 3;
  if (n > 0)
    hipLaunchKernelGGL(k_fill_pt<FP>, dim3((unsigned)std::min<int64_t>(4096, ceil_div(n, 256))), dim3(256), 0, st,
                       gcoord, n);
  const int64_t rows = nframes * nloc;
  if (rows > 0 && nnei > 0) {
    DPMI_CHECK_ARG(coord && nlist && atype && stdv, "NULL input");
    hipLaunchKernelGGL(k_env_pt_vjp<FP>, dim3(grid_for(rows)), dim3(256), 0, st, gcoord, g_env, g_diff, g_sw, coord,
                       nlist, atype, stdv, p, rows);
  }
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename FP>
int env_pt_grad_grad_impl(FP* gg_env, FP* gg_diff, FP* gg_sw, const FP* v, const FP* coord, const int* nlist,
                          const int* atype, const FP* stdv, int64_t nframes, int64_t nloc, int64_t nall, int64_t nnei,
                          int ntypes, int ncomp, double rcut, double rcut_smth, double protection, int use_exp_switch,
                          void* stream) {
  PtEnv<FP> p;
  int rc = make_params(&p, nframes, nloc, nall, nnei, ntypes, ncomp, rcut, rcut_smth, protection, use_exp_switch);
  if (rc) return rc;
  const int64_t rows = nframes * nloc;
  if (rows == 0 || nnei == 0) return DPMI_OK;
  DPMI_CHECK_ARG(v && coord && nlist && atype && stdv, "NULL input");
  hipLaunchKernelGGL(k_env_pt_jvp<FP>, dim3(grid_for(rows)), dim3(256), 0, as_stream(stream), gg_env, gg_diff, gg_sw,
                     v, coord, nlist, atype, stdv, p, rows);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

#define DPMI_ENV_PT(SFX, FP)                                                                                         \
  extern "C" int dpmi_env_mat_pt_##SFX(FP* env_mat, FP* diff, FP* sw, const FP* ext_coord, const int* nlist,       \
                                       const int* atype, const FP* mean, const FP* stddev, int64_t nframes,         \
                                       int64_t nloc, int64_t nall, int64_t nnei, int ntypes, int ncomp,             \
                                       double rcut, double rcut_smth, double protection, int use_exp_switch,       \
                                       void* stream) {                                                              \
    return env_pt_impl<FP>(env_mat, diff, sw, ext_coord, nlist, atype, mean, stddev, nframes, nloc, nall, nnei,     \
                           ntypes, ncomp, rcut, rcut_smth, protection, use_exp_switch, stream);                     \
  }                                                                                                                 \
  extern "C" int dpmi_env_mat_pt_grad_##SFX(FP* grad_coord, const FP* g_env_mat, const FP* g_diff, const FP* g_sw,  \
                                            const FP* ext_coord, const int* nlist, const int* atype,                \
                                            const FP* stddev, int64_t nframes, int64_t nloc, int64_t nall,           \
                                            int64_t nnei, int ntypes, int ncomp, double rcut, double rcut_smth,      \
                                            double protection, int use_exp_switch, void* stream) {                  \
    return env_pt_grad_impl<FP>(grad_coord, g_env_mat, g_diff, g_sw, ext_coord, nlist, atype, stddev, nframes,      \
                                nloc, nall, nnei, ntypes, ncomp, rcut, rcut_smth, protection, use_exp_switch,       \
                                stream);                                                                            \
  }                                                                                                                 \
  extern "C" int dpmi_env_mat_pt_grad_grad_##SFX(FP* gg_env_mat, FP* gg_diff, FP* gg_sw, const FP* v,               \
                                                 const FP* ext_coord, const int* nlist, const int* atype,           \
                                                 const FP* stddev, int64_t nframes, int64_t nloc, int64_t nall,      \
                                                 int64_t nnei, int ntypes, int ncomp, double rcut,                  \
                                                 double rcut_smth, double protection, int use_exp_switch,           \
                                                 void* stream) {                                                    \
    return env_pt_grad_grad_impl<FP>(gg_env_mat, gg_diff, gg_sw, v, ext_coord, nlist, atype, stddev, nframes,       \
                                     nloc, nall, nnei, ntypes, ncomp, rcut, rcut_smth, protection, use_exp_switch,  \
                                     stream);                                                                       \
  }

DPMI_ENV_PT(f64, double)
DPMI_ENV_PT(f32, float)
