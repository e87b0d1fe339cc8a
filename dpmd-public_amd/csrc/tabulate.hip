// Tabulated (fifth-order polynomial) se_a embedding fused with G^T R on gfx950.
//
// Semantics: tabulate_fusion_se_a_cpu / _grad_cpu / _grad_grad_cpu
// (source/lib/src/tabulate.cc:19-383), including locate_xx's two-stride
// table and the is_sorted early exit: the first neighbor j whose em_x equals
// the atom's last em_x and whose direction (em[1..3]) is zero ends the loop,
// its term weighted by (nnei - j).
//
// Mapping (one 64-wide wavefront per centre atom, 4 atoms per workgroup):
//  - forward and grad_grad accumulate out[c, k] over neighbors: lanes own
//    channels k (k = lane + 64 q), so no cross-lane reduction is needed; for a
//    given neighbor the wave reads one contiguous table row (6*M values).
//  - grad reduces over channels: a wave processes 16 neighbors at once, four
//    lanes per neighbor splitting the channels, so each output needs only a
//    two-step shuffle reduction.  dy[i] is staged in LDS.
#include <hip/hip_runtime.h>

#include "dpmi_internal.h"

namespace dpmi {
namespace {

constexpr int kAtoms = 4;

template <typename FP>
struct TabInfo {
  FP lower, upper, mx, s0, s1;
};

// locate_xx (tabulate.cc:19-42)
template <typename FP>
__device__ inline void locate_xx(const TabInfo<FP>& t, FP& xx, int& idx) {
  if (xx < t.lower) {
    idx = 0;
    xx = (FP)0.;
  } else if (xx < t.upper) {
    idx = (int)((xx - t.lower) / t.s0);
    xx -= (idx * t.s0 + t.lower);
  } else if (xx < t.mx) {
    int first = (int)((t.upper - t.lower) / t.s0);
    idx = first + (int)((xx - t.upper) / t.s1);
    xx -= ((idx - first) * t.s1 + t.upper);
  } else {
    idx = (int)((t.upper - t.lower) / t.s0) + (int)((t.mx - t.upper) / t.s1) - 1;
    xx = (FP)0.;
  }
}

// Index of the early-exit neighbor (nnei when there is none).
template <typename FP>
__device__ inline int exit_index(const FP* __restrict__ ex, const FP* __restrict__ e, int nnei, int is_sorted) {
  if (!is_sorted) return nnei;
  const FP ago = ex[nnei - 1];
  int first = nnei;
  for (int j0 = 0; j0 < nnei; j0 += kWave) {
    int j = j0 + lane_id();
    bool hit = false;
    if (j < nnei) hit = ex[j] == ago && e[4 * j + 1] == (FP)0. && e[4 * j + 2] == (FP)0. && e[4 * j + 3] == (FP)0.;
    uint64_t m = __ballot(hit);
    if (m) {
      first = j0 + __ffsll((unsigned long long)m) - 1;
      break;
    }
  }
  return first;
}

template <typename FP, int Q>
__global__ void __launch_bounds__(256)
    k_tab_fwd(FP* __restrict__ out, const FP* __restrict__ table, TabInfo<FP> info, const FP* __restrict__ em_x,
              const FP* __restrict__ em, int nloc, int nnei, int M, int is_sorted) {
  const int w = wave_id(), lane = lane_id();
  const int i = blockIdx.x * kAtoms + w;
  if (i >= nloc) return;
  const FP* ex = em_x + (int64_t)i * nnei;
  const FP* e = em + (int64_t)i * nnei * 4;
  const int jx = exit_index(ex, e, nnei, is_sorted);
  const int jend = jx < nnei ? jx + 1 : nnei;
  FP acc[Q][4];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[q][c] = (FP)0.;
  for (int j = 0; j < jend; ++j) {
    FP ll[4] = {e[4 * j + 0], e[4 * j + 1], e[4 * j + 2], e[4 * j + 3]};
    FP xx = ex[j];
    int idx;
    locate_xx(info, xx, idx);
    const FP wgt = j == jx ? (FP)(nnei - j) : (FP)1.;
    const FP* row = table + (int64_t)idx * M * 6;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      int k = lane + kWave * q;
      if (k < M) {
        const FP* a = row + 6 * k;
        FP var = a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * xx) * xx) * xx) * xx) * xx;
        FP wv = wgt * var;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[q][c] += wv * ll[c];
      }
    }
  }
  FP* o = out + (int64_t)i * 4 * M;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    int k = lane + kWave * q;
    if (k < M)
#pragma unroll
      for (int c = 0; c < 4; ++c) o[c * M + k] = acc[q][c];
  }
}

template <typename FP, int Q>
__global__ void __launch_bounds__(256)
    k_tab_grad_grad(FP* __restrict__ dz_dy, const FP* __restrict__ table, TabInfo<FP> info,
                    const FP* __restrict__ em_x, const FP* __restrict__ em, const FP* __restrict__ dz_dem_x,
                    const FP* __restrict__ dz_dem, int nloc, int nnei, int M, int is_sorted) {
  const int w = wave_id(), lane = lane_id();
  const int i = blockIdx.x * kAtoms + w;
  if (i >= nloc) return;
  const FP* ex = em_x + (int64_t)i * nnei;
  const FP* e = em + (int64_t)i * nnei * 4;
  const FP* hx = dz_dem_x + (int64_t)i * nnei;
  const FP* h = dz_dem + (int64_t)i * nnei * 4;
  const int jx = exit_index(ex, e, nnei, is_sorted);
  const int jend = jx < nnei ? jx + 1 : nnei;
  FP acc[Q][4];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[q][c] = (FP)0.;
  for (int j = 0; j < jend; ++j) {
    FP ll[4] = {e[4 * j + 0], e[4 * j + 1], e[4 * j + 2], e[4 * j + 3]};
    FP hh[4] = {h[4 * j + 0], h[4 * j + 1], h[4 * j + 2], h[4 * j + 3]};
    FP xx = ex[j];
    const FP dzx = hx[j];
    int idx;
    locate_xx(info, xx, idx);
    const FP wgt = j == jx ? (FP)(nnei - j) : (FP)1.;
    const FP* row = table + (int64_t)idx * M * 6;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      int k = lane + kWave * q;
      if (k < M) {
        const FP* a = row + 6 * k;
        FP var = a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * xx) * xx) * xx) * xx) * xx;
        FP vg = a[1] + ((FP)2. * a[2] + ((FP)3. * a[3] + ((FP)4. * a[4] + (FP)5. * a[5] * xx) * xx) * xx) * xx;
        FP gd = dzx * vg;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[q][c] += wgt * (var * hh[c] + gd * ll[c]);
      }
    }
  }
  FP* o = dz_dy + (int64_t)i * 4 * M;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    int k = lane + kWave * q;
    if (k < M)
#pragma unroll
      for (int c = 0; c < 4; ++c) o[c * M + k] = acc[q][c];
  }
}

// grad: lane = 4*jg + kg; neighbor j = jb + jg, channels k = kg, kg+4, ...
template <typename FP>
__global__ void __launch_bounds__(256)
    k_tab_grad(FP* __restrict__ dy_dem_x, FP* __restrict__ dy_dem, const FP* __restrict__ table, TabInfo<FP> info,
               const FP* __restrict__ em_x, const FP* __restrict__ em, const FP* __restrict__ dy, int nloc, int nnei,
               int M, int is_sorted) {
  extern __shared__ unsigned char smem_raw[];
  FP* dys = reinterpret_cast<FP*>(smem_raw);  // [kAtoms][4*M]
  const int w = threadIdx.x / kWave, lane = lane_id();  // (wave_id() measured slower here)
  const int i = blockIdx.x * kAtoms + w;
  FP* dyw = dys + (int64_t)w * 4 * M;
  if (i < nloc) {
    const FP* src = dy + (int64_t)i * 4 * M;
    for (int k = lane; k < 4 * M; k += kWave) dyw[k] = src[k];
  }
  __syncthreads();
  if (i >= nloc) return;
  const FP* ex = em_x + (int64_t)i * nnei;
  const FP* e = em + (int64_t)i * nnei * 4;
  const int jx = exit_index(ex, e, nnei, is_sorted);
  const int jg = lane >> 2, kg = lane & 3;
  for (int jb = 0; jb < nnei; jb += 16) {
    const int j = jb + jg;
    FP gsum = (FP)0., d0 = (FP)0., d1 = (FP)0., d2 = (FP)0., d3 = (FP)0.;
    const bool live = j < nnei && j <= jx;
    if (live) {
      FP ll[4] = {e[4 * j + 0], e[4 * j + 1], e[4 * j + 2], e[4 * j + 3]};
      FP xx = ex[j];
      int idx;
      locate_xx(info, xx, idx);
      const FP* row = table + (int64_t)idx * M * 6;
      for (int k = kg; k < M; k += 4) {
        const FP* a = row + 6 * k;
        FP r0 = dyw[k], r1 = dyw[M + k], r2 = dyw[2 * M + k], r3 = dyw[3 * M + k];
        FP res = a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * xx) * xx) * xx) * xx) * xx;
        FP g = (a[1] + ((FP)2. * a[2] + ((FP)3. * a[3] + ((FP)4. * a[4] + (FP)5. * a[5] * xx) * xx) * xx) * xx);
        gsum += g * (ll[0] * r0 + ll[1] * r1 + ll[2] * r2 + ll[3] * r3);
        d0 += res * r0;
        d1 += res * r1;
        d2 += res * r2;
        d3 += res * r3;
      }
    }
#pragma unroll
    for (int off = 1; off < 4; off <<= 1) {
      gsum += __shfl_xor(gsum, off, kWave);
      d0 += __shfl_xor(d0, off, kWave);
      d1 += __shfl_xor(d1, off, kWave);
      d2 += __shfl_xor(d2, off, kWave);
      d3 += __shfl_xor(d3, off, kWave);
    }
    if (kg == 0 && j < nnei) {
      const FP wgt = j == jx ? (FP)(nnei - j) : (FP)1.;
      const int64_t o = (int64_t)i * nnei + j;
      if (live) {
        dy_dem_x[o] = gsum * wgt;
        dy_dem[4 * o + 0] = d0 * wgt;
        dy_dem[4 * o + 1] = d1 * wgt;
        dy_dem[4 * o + 2] = d2 * wgt;
        dy_dem[4 * o + 3] = d3 * wgt;
      } else {
        dy_dem_x[o] = (FP)0.;
        dy_dem[4 * o + 0] = (FP)0.;
        dy_dem[4 * o + 1] = (FP)0.;
        dy_dem[4 * o + 2] = (FP)0.;
        dy_dem[4 * o + 3] = (FP)0.;
      }
    }
  }
}

template <typename FP>
int get_info(TabInfo<FP>* t, const FP* info_host) {
  DPMI_CHECK_ARG(info_host != nullptr, "table_info is NULL (it must be host memory)");
  t->lower = info_host[0];
  t->upper = info_host[1];
  t->mx = info_host[2];
  t->s0 = info_host[3];
  t->s1 = info_host[4];
  return DPMI_OK;
}

template <typename FP>
int check_sizes(int64_t nloc, int64_t nnei, int64_t M) {
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  DPMI_CHECK_ARG(nnei > 0 && nnei <= 65536, "nnei out of range");
  DPMI_CHECK_ARG(M > 0 && M <= 1024,
                 "In the process of model compression, the size of the last layer of embedding net must be less "
                 "than 1024!");
  return DPMI_OK;
}

#define DPMI_Q_DISPATCH(KERNEL, FP, ...)                                         \
  do {                                                                           \
    int q_ = (int)ceil_div(M, kWave);                                            \
    if (q_ <= 1) hipLaunchKernelGGL((KERNEL<FP, 1>), __VA_ARGS__);               \
    else if (q_ <= 2) hipLaunchKernelGGL((KERNEL<FP, 2>), __VA_ARGS__);          \
    else if (q_ <= 4) hipLaunchKernelGGL((KERNEL<FP, 4>), __VA_ARGS__);          \
    else if (q_ <= 8) hipLaunchKernelGGL((KERNEL<FP, 8>), __VA_ARGS__);          \
    else hipLaunchKernelGGL((KERNEL<FP, 16>), __VA_ARGS__);                      \
  } while (0)

template <typename FP>
int tab_fwd(FP* out, const FP* table, const FP* info_host, const FP* em_x, const FP* em, int64_t nloc, int64_t nnei,
            int64_t M, int is_sorted, void* stream) {
  int rc = check_sizes<FP>(nloc, nnei, M);
  if (rc) return rc;
  TabInfo<FP> t;
  if ((rc = get_info(&t, info_host))) return rc;
  if (nloc == 0) return DPMI_OK;
  DPMI_Q_DISPATCH(k_tab_fwd, FP, dim3(ceil_div(nloc, kAtoms)), dim3(256), 0, as_stream(stream), out, table, t, em_x,
                  em, (int)nloc, (int)nnei, (int)M, is_sorted);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename FP>
int tab_grad(FP* dx, FP* dem, const FP* table, const FP* info_host, const FP* em_x, const FP* em, const FP* dy,
             int64_t nloc, int64_t nnei, int64_t M, int is_sorted, void* stream) {
  int rc = check_sizes<FP>(nloc, nnei, M);
  if (rc) return rc;
  TabInfo<FP> t;
  if ((rc = get_info(&t, info_host))) return rc;
  if (nloc == 0) return DPMI_OK;
  size_t shm = sizeof(FP) * kAtoms * 4 * M;
  if (shm > 65536) {  // gfx950 has 160 KB of LDS per CU; opt in above 64 KB
    DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tab_grad<FP>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  }
  hipLaunchKernelGGL(k_tab_grad<FP>, dim3(ceil_div(nloc, kAtoms)), dim3(256), shm, as_stream(stream), dx, dem, table,
                     t, em_x, em, dy, (int)nloc, (int)nnei, (int)M, is_sorted);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename FP>
int tab_grad_grad(FP* dz_dy, const FP* table, const FP* info_host, const FP* em_x, const FP* em, const FP* dzx,
                  const FP* dze, int64_t nloc, int64_t nnei, int64_t M, int is_sorted, void* stream) {
  int rc = check_sizes<FP>(nloc, nnei, M);
  if (rc) return rc;
  TabInfo<FP> t;
  if ((rc = get_info(&t, info_host))) return rc;
  if (nloc == 0) return DPMI_OK;
  DPMI_Q_DISPATCH(k_tab_grad_grad, FP, dim3(ceil_div(nloc, kAtoms)), dim3(256), 0, as_stream(stream), dz_dy, table, t,
                  em_x, em, dzx, dze, (int)nloc, (int)nnei, (int)M, is_sorted);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_tabulate_fusion_se_a_f64(double* out, const double* table, const double* table_info_host,
                                             const double* em_x, const double* em, int64_t nloc, int64_t nnei,
                                             int64_t last_layer_size, int is_sorted, void* stream) {
  return tab_fwd<double>(out, table, table_info_host, em_x, em, nloc, nnei, last_layer_size, is_sorted, stream);
}
extern "C" int dpmi_tabulate_fusion_se_a_f32(float* out, const float* table, const float* table_info_host,
                                             const float* em_x, const float* em, int64_t nloc, int64_t nnei,
                                             int64_t last_layer_size, int is_sorted, void* stream) {
  return tab_fwd<float>(out, table, table_info_host, em_x, em, nloc, nnei, last_layer_size, is_sorted, stream);
}
extern "C" int dpmi_tabulate_fusion_se_a_grad_f64(double* dy_dem_x, double* dy_dem, const double* table,
                                                  const double* table_info_host, const double* em_x, const double* em,
                                                  const double* dy, int64_t nloc, int64_t nnei,
                                                  int64_t last_layer_size, int is_sorted, void* stream) {
  return tab_grad<double>(dy_dem_x, dy_dem, table, table_info_host, em_x, em, dy, nloc, nnei, last_layer_size,
                          is_sorted, stream);
}
extern "C" int dpmi_tabulate_fusion_se_a_grad_f32(float* dy_dem_x, float* dy_dem, const float* table,
                                                  const float* table_info_host, const float* em_x, const float* em,
                                                  const float* dy, int64_t nloc, int64_t nnei,
                                                  int64_t last_layer_size, int is_sorted, void* stream) {
  return tab_grad<float>(dy_dem_x, dy_dem, table, table_info_host, em_x, em, dy, nloc, nnei, last_layer_size,
                         is_sorted, stream);
}
extern "C" int dpmi_tabulate_fusion_se_a_grad_grad_f64(double* dz_dy, const double* table,
                                                       const double* table_info_host, const double* em_x,
                                                       const double* em, const double* dz_dy_dem_x,
                                                       const double* dz_dy_dem, int64_t nloc, int64_t nnei,
                                                       int64_t last_layer_size, int is_sorted, void* stream) {
  return tab_grad_grad<double>(dz_dy, table, table_info_host, em_x, em, dz_dy_dem_x, dz_dy_dem, nloc, nnei,
                               last_layer_size, is_sorted, stream);
}
extern "C" int dpmi_tabulate_fusion_se_a_grad_grad_f32(float* dz_dy, const float* table, const float* table_info_host,
                                                       const float* em_x, const float* em, const float* dz_dy_dem_x,
                                                       const float* dz_dy_dem, int64_t nloc, int64_t nnei,
                                                       int64_t last_layer_size, int is_sorted, void* stream) {
  return tab_grad_grad<float>(dz_dy, table, table_info_host, em_x, em, dz_dy_dem_x, dz_dy_dem, nloc, nnei,
                              last_layer_size, is_sorted, stream);
}
