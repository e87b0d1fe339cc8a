// Table builder of the compressed (tabulated) embedding on gfx950, FP64.
//
// Reference semantics: DPTabulate.build for se_e2_a (deepmd/utils/
// tabulate.py:67-147, 273-367) with the value / slope / curvature chain of
// the embedding net (deepmd/utils/tabulate_math.py:136-251, 353-470): each
// grid node x is pushed through the net (y = tanh(x W + b) [* idt] [+ x or
// + (x, x)] per layer) together with dy/dx and d2y/dx2, and every interval
// [x_k, x_k+1] of width t becomes the quintic that matches value, slope and
// curvature at both ends:
//   a0 = v0, a1 = d0, a2 = c0 / 2,
//   a3 = (20 h - (8 d1 + 12 d0) t - (3 c0 - c1) t^2) / (2 t^3),
//   a4 = (-30 h + (14 d1 + 16 d0) t + (3 c0 - 2 c1) t^2) / (2 t^4),
//   a5 = (12 h - 6 (d1 + d0) t + (c1 - c0) t^2) / (2 t^5),   h = v1 - v0,
// stored channel-major, table[k][6 c + i] (the layout tabulate_fusion_se_a
// reads).
//
// Mapping: one workgroup per grid node (grid-stride), a thread per output
// feature of the current layer; the layer's input (value, slope, curvature)
// sits in LDS, double-buffered across layers.  Then one thread per
// (interval, channel) forms the six coefficients.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "dpmi_internal.h"
#include "tanh64.h"

namespace dpmi {
namespace {

constexpr int kMaxLayers = 8;
constexpr int kMaxWidth = 1024;

struct NetDims {
  int nlayers;
  int nin[kMaxLayers], nout[kMaxLayers];
  long long off[kMaxLayers];  // start of layer l in params: w [nin][nout], b [nout], idt [nout]
};

__global__ void __launch_bounds__(256)
    k_net_chain(double* __restrict__ vdd, const double* __restrict__ xx, int64_t nnodes,
                const double* __restrict__ params, NetDims d, int M) {
  extern __shared__ double sm[];
  double* buf[2][3] = {{sm, sm + kMaxWidth, sm + 2 * kMaxWidth},
                       {sm + 3 * kMaxWidth, sm + 4 * kMaxWidth, sm + 5 * kMaxWidth}};
  for (int64_t node = blockIdx.x; node < nnodes; node += gridDim.x) {
    if (threadIdx.x == 0) {
      buf[0][0][0] = xx[node];
      buf[0][1][0] = 1.;
      buf[0][2][0] = 0.;
    }
    __syncthreads();
    int cur = 0;
    for (int l = 0; l < d.nlayers; ++l) {
      const int nin = d.nin[l], nout = d.nout[l];
      const double* w = params + d.off[l];
      const double* b = w + (int64_t)nin * nout;
      const double* idt = b + nout;
      const double* y = buf[cur][0];
      const double* dy = buf[cur][1];
      const double* d2y = buf[cur][2];
      double* yo = buf[cur ^ 1][0];
      double* dyo = buf[cur ^ 1][1];
      double* d2yo = buf[cur ^ 1][2];
      for (int j = threadIdx.x; j < nout; j += blockDim.x) {
        double z = b[j], dz = 0., d2z = 0.;
        for (int i = 0; i < nin; ++i) {
          const double wij = w[(int64_t)i * nout + j];
          z = fma(y[i], wij, z);
          dz = fma(dy[i], wij, dz);
          d2z = fma(d2y[i], wij, d2z);
        }
        const double t = tanh64(z), g1 = 1. - t * t, g2 = -2. * t * g1, s = idt[j];
        double o = t * s, od = g1 * dz * s, od2 = (g2 * dz * dz + g1 * d2z) * s;
        if (nout == nin) {
          o += y[j];
          od += dy[j];
          od2 += d2y[j];
        } else if (nout == 2 * nin) {
          const int k = j < nin ? j : j - nin;
          o += y[k];
          od += dy[k];
          od2 += d2y[k];
        }
        yo[j] = o;
        dyo[j] = od;
        d2yo[j] = od2;
      }
      __syncthreads();
      cur ^= 1;
    }
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
      vdd[node * M + j] = buf[cur][0][j];
      vdd[(nnodes + node) * M + j] = buf[cur][1][j];
      vdd[(2 * nnodes + node) * M + j] = buf[cur][2][j];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256)
    k_hermite(double* __restrict__ table, const double* __restrict__ vdd, const double* __restrict__ tt,
              int64_t nspline, int64_t nnodes, int M) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nspline * M) return;
  const int64_t k = q / M;
  const int c = (int)(q - k * M);
  const double* v = vdd;
  const double* dv = vdd + nnodes * M;
  const double* cv = vdd + 2 * nnodes * M;
  const double v0 = v[k * M + c], v1 = v[(k + 1) * M + c];
  const double d0 = dv[k * M + c], d1 = dv[(k + 1) * M + c];
  const double c0 = cv[k * M + c], c1 = cv[(k + 1) * M + c];
  const double t = tt[k], h = v1 - v0, t2 = t * t, t3 = t2 * t;
  double* o = table + k * 6 * M + 6 * c;
  o[0] = v0;
  o[1] = d0;
  o[2] = 0.5 * c0;
  o[3] = (20. * h - (8. * d1 + 12. * d0) * t - (3. * c0 - c1) * t2) / (2. * t3);
  o[4] = (-30. * h + (14. * d1 + 16. * d0) * t + (3. * c0 - 2. * c1) * t2) / (2. * t3 * t);
  o[5] = (12. * h - 6. * (d1 + d0) * t + (c1 - c0) * t2) / (2. * t3 * t2);
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_tabulate_build_f64(double* table, double* node_vdd, const double* xx, const double* tt,
                                       int64_t nnodes, const double* params, const int* dims_host, int nlayers,
                                       void* stream) {
  DPMI_CHECK_ARG(nlayers > 0 && nlayers <= kMaxLayers, "nlayers out of range [1, 8]");
  DPMI_CHECK_ARG(dims_host != nullptr, "dims is NULL");
  DPMI_CHECK_ARG(nnodes >= 2 && nnodes < (int64_t(1) << 31), "need at least two grid nodes");
  NetDims d;
  std::memset(&d, 0, sizeof(d));
  d.nlayers = nlayers;
  long long off = 0;
  int width = 1;
  for (int l = 0; l < nlayers; ++l) {
    d.nin[l] = dims_host[2 * l];
    d.nout[l] = dims_host[2 * l + 1];
    DPMI_CHECK_ARG(d.nin[l] == width, "layer input width must match the previous layer's output (first: 1)");
    DPMI_CHECK_ARG(d.nout[l] > 0 && d.nout[l] <= kMaxWidth, "layer width out of range [1, 1024]");
    d.off[l] = off;
    off += (long long)d.nin[l] * d.nout[l] + 2LL * d.nout[l];
    width = d.nout[l];
  }
  DPMI_CHECK_ARG(table && node_vdd && xx && tt && params, "NULL argument");
  const int M = width;
  hipStream_t st = as_stream(stream);
  const size_t shm = 6 * kMaxWidth * sizeof(double);
  DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_net_chain), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shm));
  hipLaunchKernelGGL(k_net_chain, dim3((unsigned)std::min<int64_t>(nnodes, 4096)), dim3(256), shm, st, node_vdd, xx,
                     nnodes, params, d, M);
  const int64_t nq = (nnodes - 1) * M;
  hipLaunchKernelGGL(k_hermite, dim3((unsigned)ceil_div(nq, 256)), dim3(256), 0, st, table, node_vdd, tt, nnodes - 1,
                     nnodes, M);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}
