// Optimizer side of the training step on the GPU: the learning-rate schedule
// tick (with the loss prefactors that follow the rate) and Adam over one flat
// parameter buffer.  Both read the step count and the rate from device memory,
// so a whole training step (lists, env-mat, forward, double backward, Adam)
// can be captured once and replayed as a HIP graph.
//
// Reference semantics:
//  * learning rate: LambdaLR over LearningRateExp (deepmd/pt/train/
//    training.py:1138-1150; deepmd/dpmodel/utils/learning_rate.py:167-217,
//    367-397): value(k) = start_lr * r^floor(k / s) (r^(k / s) when smooth),
//    clipped below at stop_lr, after a linear warmup from warmup_start_lr;
//    the optimizer's rate is base_lr * (value(k) / base_lr).
//  * loss prefactors: pref = limit + (start - limit) * lr / start_lr
//    (deepmd/pt/loss/ener.py:229-231).
//  * Adam: torch.optim.Adam(fused=True) as built at training.py:1028-1030
//    (betas, eps, weight_decay; no amsgrad, no maximize), stepped once per
//    training step (training.py:1451-1452): with t = steps taken so far,
//      g += wd p;  m = b1 m + (1 - b1) g;  v = b2 v + (1 - b2) g g;
//      p -= (lr / (1 - b1^t)) m / (sqrt(v) / sqrt(1 - b2^t) + eps).
// The unit is built without floating-point contraction, so each update is
// the rounded sequence written above.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "dpmi_internal.h"

namespace dpmi {
namespace {

// host layout of the schedule argument of dpmi_lr_tick
enum {
  kStartLr, kDecayRate, kStopLr, kDecaySteps, kWarmupSteps, kWarmupStartLr, kSmooth,
  kStartPrefE, kLimitPrefE, kStartPrefF, kLimitPrefF, kSchedLen
};

struct Schedule {
  double start_lr, decay_rate, stop_lr, warmup_start_lr;
  long long decay_steps, warmup_steps;
  int smooth;
  double start_pref_e, limit_pref_e, start_pref_f, limit_pref_f;
};

__device__ double decay_value(const Schedule& s, double step) {
  const double e = s.smooth ? step / (double)s.decay_steps : floor(step / (double)s.decay_steps);
  const double lr = s.start_lr * pow(s.decay_rate, e);
  return lr < s.stop_lr ? s.stop_lr : lr;
}

// hyper = [lr, pref_e, pref_f, t] of the step that starts; step counts the
// steps begun so far (the schedule index k before, t = k + 1 after)
__global__ void k_lr_tick(double* __restrict__ hyper, long long* __restrict__ step, Schedule s) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const long long k = *step;
  const double kd = (double)k;
  double value;
  if (s.warmup_steps == 0) {
    value = decay_value(s, kd);
  } else {
    const double warm = s.warmup_start_lr + (s.start_lr - s.warmup_start_lr) * (kd / (double)s.warmup_steps);
    const double dk = kd - (double)s.warmup_steps;
    const double decay = decay_value(s, dk > 0.0 ? dk : 0.0);
    value = k < s.warmup_steps ? warm : decay;
  }
  const double lr = s.start_lr * (value / s.start_lr);
  const double coef = lr / s.start_lr;
  hyper[0] = lr;
  hyper[1] = s.limit_pref_e + (s.start_pref_e - s.limit_pref_e) * coef;
  hyper[2] = s.limit_pref_f + (s.start_pref_f - s.limit_pref_f) * coef;
  hyper[3] = (double)(k + 1);
  *step = k + 1;
}

// Adam over n elements of one flat buffer; the step scalars come from hyper
// (written by k_lr_tick earlier on the stream).  Elementwise, HBM bound:
// 4 reads and 3 writes per element.
template <typename T>
__global__ void __launch_bounds__(256)
    k_adam(T* __restrict__ param, const T* __restrict__ grad, T* __restrict__ exp_avg, T* __restrict__ exp_avg_sq,
           long long n, const double* __restrict__ hyper, double beta1, double beta2, double eps, double wd) {
  const double lr = hyper[0], t = hyper[3];
  const T step_size = (T)(lr / (1.0 - pow(beta1, t)));
  const T bc2_sqrt = (T)sqrt(1.0 - pow(beta2, t));
  const T b1 = (T)beta1, b2 = (T)beta2, one_b1 = (T)(1.0 - beta1), one_b2 = (T)(1.0 - beta2);
  const T e = (T)eps, w = (T)wd;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    T p = param[i], g = grad[i];
    if (wd != 0.0) g = g + p * w;
    const T m = b1 * exp_avg[i] + one_b1 * g;
    const T v = b2 * exp_avg_sq[i] + one_b2 * g * g;
    const T denom = sqrt(v) / bc2_sqrt + e;
    p = p - step_size * m / denom;
    exp_avg[i] = m;
    exp_avg_sq[i] = v;
    param[i] = p;
  }
}

template <typename T>
int adam_impl(T* param, const T* grad, T* exp_avg, T* exp_avg_sq, int64_t n, const double* hyper, double beta1,
              double beta2, double eps, double weight_decay, void* stream) {
  DPMI_CHECK_ARG(n >= 0, "n must be >= 0");
  if (n == 0) return DPMI_OK;
  DPMI_CHECK_ARG(param && grad && exp_avg && exp_avg_sq && hyper, "NULL argument");
  DPMI_CHECK_ARG(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0, "betas must lie in [0, 1)");
  DPMI_CHECK_ARG(eps >= 0.0 && weight_decay >= 0.0, "eps and weight_decay must be >= 0");
  const int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 4096);
  hipLaunchKernelGGL(k_adam<T>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), param, grad, exp_avg,
                     exp_avg_sq, (long long)n, hyper, beta1, beta2, eps, weight_decay);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

// Column sums of x [R, F] in a fixed order, in two passes: block b of pass 1
// sums rows [b chunk, (b + 1) chunk) into part [b, F] (lanes of one column
// over interleaved rows, combined in lane order), pass 2 sums the parts of
// each column in a fixed pairwise tree.  The reduction the training graph needs for the gradients of
// broadcast biases and timesteps, without the multi-block library reduction
// (whose semaphore memset must not sit in a captured graph).
constexpr int kRowSumBlocks = 256;

template <typename T>
__global__ void __launch_bounds__(256)
    k_rowsum_part(T* __restrict__ part, const T* __restrict__ x, long long R, int F, long long chunk) {
  __shared__ T s[256];
  const int W = F < 256 ? F : 256, nj = 256 / W;
  const int t = threadIdx.x, j = t / W, c = blockIdx.y * W + t % W;
  T acc = 0;
  if (j < nj && c < F) {
    const long long r0 = (long long)blockIdx.x * chunk, r1 = r0 + chunk < R ? r0 + chunk : R;
    for (long long r = r0 + j; r < r1; r += nj) acc += x[r * F + c];
  }
  s[t] = acc;
  __syncthreads();
  if (t < W && blockIdx.y * W + t < F) {
    T v = 0;
    for (int q = 0; q < nj; ++q) v += s[q * W + t];
    part[(long long)blockIdx.x * F + blockIdx.y * W + t] = v;
  }
}

// one workgroup per column: part b in lane b, then a fixed pairwise tree
template <typename T>
__global__ void __launch_bounds__(kRowSumBlocks) k_rowsum_final(T* __restrict__ out, const T* __restrict__ part, int nb,
                                                               int F) {
  __shared__ T s[kRowSumBlocks];
  const int c = blockIdx.x, t = threadIdx.x;
  s[t] = t < nb ? part[(long long)t * F + c] : T(0);
  __syncthreads();
#pragma unroll
  for (int h = kRowSumBlocks / 2; h > 0; h >>= 1) {
    if (t < h) s[t] += s[t + h];
    __syncthreads();
  }
  if (t == 0) out[c] = s[0];
}

template <typename T>
int row_sum_impl(T* out, const T* x, T* part, int64_t R, int64_t F, void* stream) {
  DPMI_CHECK_ARG(R >= 0 && F >= 1 && F < (int64_t(1) << 30), "need R >= 0 and 1 <= F < 2^30");
  DPMI_CHECK_ARG(out && part && (x || R == 0), "NULL argument");
  hipStream_t st = as_stream(stream);
  const int nb = (int)std::min<int64_t>(kRowSumBlocks, std::max<int64_t>(1, ceil_div(R, 256)));
  const int64_t chunk = std::max<int64_t>(1, ceil_div(R, nb));
  const int W = F < 256 ? (int)F : 256;
  hipLaunchKernelGGL(k_rowsum_part<T>, dim3((unsigned)nb, (unsigned)ceil_div(F, W)), dim3(256), 0, st, part, x,
                     (long long)R, (int)F, (long long)chunk);
  hipLaunchKernelGGL(k_rowsum_final<T>, dim3((unsigned)F), dim3(kRowSumBlocks), 0, st, out, part, nb, (int)F);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_row_sum_f64(double* out, const double* x, double* part, int64_t R, int64_t F, void* stream) {
  return row_sum_impl<double>(out, x, part, R, F, stream);
}

extern "C" int dpmi_row_sum_f32(float* out, const float* x, float* part, int64_t R, int64_t F, void* stream) {
  return row_sum_impl<float>(out, x, part, R, F, stream);
}

extern "C" int dpmi_lr_tick(double* hyper, int64_t* step, const double* sched_host, void* stream) {
  DPMI_CHECK_ARG(hyper && step && sched_host, "NULL argument");
  const double* a = sched_host;
  DPMI_CHECK_ARG(a[kStartLr] > 0.0, "start_lr must be > 0");
  DPMI_CHECK_ARG(a[kDecaySteps] >= 1.0 && a[kWarmupSteps] >= 0.0, "decay_steps must be >= 1, warmup_steps >= 0");
  Schedule s;
  s.start_lr = a[kStartLr];
  s.decay_rate = a[kDecayRate];
  s.stop_lr = a[kStopLr];
  s.decay_steps = (long long)a[kDecaySteps];
  s.warmup_steps = (long long)a[kWarmupSteps];
  s.warmup_start_lr = a[kWarmupStartLr];
  s.smooth = a[kSmooth] != 0.0;
  s.start_pref_e = a[kStartPrefE];
  s.limit_pref_e = a[kLimitPrefE];
  s.start_pref_f = a[kStartPrefF];
  s.limit_pref_f = a[kLimitPrefF];
  hipLaunchKernelGGL(k_lr_tick, dim3(1), dim3(64), 0, as_stream(stream), hyper, reinterpret_cast<long long*>(step), s);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_adam_f64(double* param, const double* grad, double* exp_avg, double* exp_avg_sq, int64_t n,
                             const double* hyper, double beta1, double beta2, double eps, double weight_decay,
                             void* stream) {
  return adam_impl<double>(param, grad, exp_avg, exp_avg_sq, n, hyper, beta1, beta2, eps, weight_decay, stream);
}

extern "C" int dpmi_adam_f32(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                             const double* hyper, double beta1, double beta2, double eps, double weight_decay,
                             void* stream) {
  return adam_impl<float>(param, grad, exp_avg, exp_avg_sq, n, hyper, beta1, beta2, eps, weight_decay, stream);
}
