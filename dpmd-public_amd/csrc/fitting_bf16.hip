// Fused energy fitting net on gfx950 MFMA (bf16 operands, fp32 accumulation):
// atomic energies and dE/dD of every local atom in one pass.
//
// Reference semantics: GeneralFitting._forward_common / EnergyFittingNet
// (deepmd/pt/model/task/fitting.py:734-900, ener.py) with FittingNet layers
// (deepmd/dpmodel/utils/network.py:1031-1090, MLPLayer.forward of
// deepmd/pt/model/network/mlp.py:188-219), one network per atom type:
//   h1 = tanh(D W1 + b1) [* idt1],                      (K1 -> H)
//   h_l = tanh(h_{l-1} W_l + b_l) [* idt_l] + h_{l-1},   (H -> H, l = 2, 3)
//   e  = h3 . w4 + b4 + bias_atom_e[t],                  (H -> 1, linear)
// and the adjoint dE/dD for E = sum_i e_i, which the descriptor adjoint
// turns into dE/d em.
//
// MI355X mapping.  A workgroup owns 64 atoms of one type (the list is sorted
// by type); its four wavefronts split every hidden layer by output feature
// (wavefront w owns features 64w .. 64w+63 of the width padded to 256), so a
// feature's tanh factor, skip value and adjoint stay in the registers of one
// lane for the whole pass.  Layers are computed transposed, y^T = W^T x^T,
// with v_mfma_f32_16x16x32_bf16: the weight fragment is the A operand (M =
// output feature), the layer input is the B operand (K = input feature,
// N = atom), read from an LDS buffer [atom][feature] that every layer's
// output is written back to.  The D matrix streams through LDS in 64-wide K
// blocks (bf16, as the descriptor kernel writes it), prefetched one block ahead.  The
// backward pass runs the transposed products (dh_{l-1} = dz_l W_l^T + skip)
// through the same LDS buffer and ends with dD = dz1 W1^T, whose B operands
// (dz1 for all 64 atoms) stay in registers while W1's fragments stream from
// L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "dpmi_internal.h"

namespace dpmi {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kH = 256;         // padded hidden width (4 wavefronts x 64 features)
constexpr int kTile = 64;       // atoms per workgroup
constexpr int kKB = 64;         // layer-1 input features staged per step
constexpr int kXS = kKB + 8;    // LDS row stride (bf16) of the staged D block
constexpr int kHS = kH + 8;     // LDS row stride (bf16) of the hidden-layer buffer
constexpr int kMaxTypes = 16;
constexpr int kCst = 7 * kH + 8;  // b1 idt1 b2 idt2 b3 idt3 w4 [kH each], b4

// Fragment offsets (in 64-lane fragments of bf16x8) inside one type's block.
struct FitParams {
  int ntypes;
  int k1;    // width of D (layer-1 input)
  int nk1;   // 32-wide K chunks of layer 1 (k1 padded to a multiple of 64, / 32)
  int nm1;   // 16-row M tiles of the dD product (k1 / 16)
  int64_t type_frags;  // fragments per type
  int64_t w1f, w2f, w3f, w3b, w2b, w1b;
};

__device__ inline float fast_tanh(float x) {
  float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + e);
}

__device__ inline f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// B fragment of 16 atoms (rows 16s..16s+15 of an LDS [atom][stride] buffer),
// K = features kbase + 8g .. kbase + 8g + 7.
__device__ inline bf16x8 lds_b(const __bf16* buf, int stride, int s, int n, int kbase, int g) {
  return *reinterpret_cast<const bf16x8*>(buf + (16 * s + n) * stride + kbase + 8 * g);
}

// Writes a layer output (D layout: lane -> atom n of subtile s, features
// 4g + v of M tile mt) to the LDS buffer as bf16.
__device__ inline void lds_put(__bf16* buf, int w, int mt, int s, int n, int g, const float* v4) {
  bf16x4 q;
  q[0] = (__bf16)v4[0];
  q[1] = (__bf16)v4[1];
  q[2] = (__bf16)v4[2];
  q[3] = (__bf16)v4[3];
  *reinterpret_cast<bf16x4*>(buf + (16 * s + n) * kHS + 16 * (4 * w + mt) + 4 * g) = q;
}

// One hidden H -> H layer product for this wavefront's four M tiles and the
// 64 atoms: acc[mt][s] = sum_k W(frag)[feature][k] * in[atom][k].
__device__ inline void hidden_product(f32x4 (&acc)[4][4], const bf16x8* __restrict__ frag, const __bf16* s_h,
                                      int w, int lane, int n, int g) {
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[mt][s] = f32x4{0.f, 0.f, 0.f, 0.f};
  bf16x8 a[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) a[mt] = frag[((4 * w + mt) * 8 + 0) * 64 + lane];
#pragma unroll 1
  for (int kc = 0; kc < 8; ++kc) {
    bf16x8 an[4];
    const int kn = kc + 1 < 8 ? kc + 1 : kc;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) an[mt] = frag[((4 * w + mt) * 8 + kn) * 64 + lane];
    bf16x8 b[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) b[s] = lds_b(s_h, kHS, s, n, 32 * kc, g);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[mt][s] = mfma(a[mt], b[s], acc[mt][s]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) a[mt] = an[mt];
  }
}

__global__ void __launch_bounds__(256)
    k_fit(double* __restrict__ atom_energy, float* __restrict__ ddesc, const __bf16* __restrict__ desc,
          const int* __restrict__ perm, const int* __restrict__ type_start, const bf16x8* __restrict__ frag_all,
          const float* __restrict__ consts, const double* __restrict__ ebias, FitParams p) {
  __shared__ __attribute__((aligned(16))) __bf16 s_x[2][kTile * kXS];
  __shared__ __attribute__((aligned(16))) __bf16 s_h[kTile * kHS];
  __shared__ float s_e[4][kTile];
  __shared__ int s_atom[kTile];

  // tile -> (type, first sorted index); every thread of the workgroup
  // computes the same answer, so the early exit is uniform
  int t = -1, start = 0, cnt = 0, before = 0;
  for (int tt = 0; tt < p.ntypes; ++tt) {
    const int lo = type_start[tt], hi = type_start[tt + 1];
    const int nt = (hi - lo + kTile - 1) / kTile;
    if (t < 0 && (int)blockIdx.x < before + nt) {
      t = tt;
      start = lo + ((int)blockIdx.x - before) * kTile;
      cnt = min(kTile, hi - start);
    }
    before += nt;
  }
  if (t < 0) return;

  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, n = lane & 15, g = lane >> 4;
  if (tid < kTile) s_atom[tid] = tid < cnt ? perm[start + tid] : -1;
  const bf16x8* frag = frag_all + (int64_t)t * p.type_frags;
  const float* cst = consts + (int64_t)t * kCst;
  __syncthreads();

  // ---- layer 1: D block loads (thread -> atom tid/4, features 16*(tid%4) .. +15 of the block, bf16)
  const int xa = s_atom[tid >> 2], xq = tid & 3;
  const __bf16* xrow = desc + (int64_t)(xa < 0 ? 0 : xa) * p.k1;
  const int nkb = p.nk1 / 2;
  auto load_block = [&](int kb, bf16x8 (&r)[2]) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int f = kb * kKB + 16 * xq + 8 * j;
      // streamed once: non-temporal, so it does not evict the weight fragments from L2
      r[j] = (xa >= 0 && f < p.k1) ? __builtin_nontemporal_load(reinterpret_cast<const bf16x8*>(xrow + f))
                                   : bf16x8{};
    }
  };
  auto store_block = [&](int buf, const bf16x8 (&r)[2]) {
    __bf16* dst = s_x[buf] + (tid >> 2) * kXS + 16 * xq;
    *reinterpret_cast<bf16x8*>(dst) = r[0];
    *reinterpret_cast<bf16x8*>(dst + 8) = r[1];
  };

  f32x4 acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[mt][s] = f32x4{0.f, 0.f, 0.f, 0.f};
  {
    bf16x8 r[2];
    load_block(0, r);
    store_block(0, r);
  }
  __syncthreads();
  const bf16x8* w1f = frag + p.w1f;
  // weight fragments of the next K block are loaded one block ahead
  bf16x8 a_cur[2][4];
#pragma unroll
  for (int kc = 0; kc < 2; ++kc)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) a_cur[kc][mt] = w1f[((4 * w + mt) * p.nk1 + kc) * 64 + lane];
#pragma unroll 1
  for (int kb = 0; kb < nkb; ++kb) {
    bf16x8 r[2];
    bf16x8 a_nxt[2][4];
    const bool more = kb + 1 < nkb;
    const int kn = more ? kb + 1 : kb;
    if (more) load_block(kb + 1, r);
#pragma unroll
    for (int kc = 0; kc < 2; ++kc)
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) a_nxt[kc][mt] = w1f[((4 * w + mt) * p.nk1 + 2 * kn + kc) * 64 + lane];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
      bf16x8 b[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) b[s] = lds_b(s_x[kb & 1], kXS, s, n, 32 * kc, g);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[mt][s] = mfma(a_cur[kc][mt], b[s], acc[mt][s]);
    }
    if (more) store_block((kb + 1) & 1, r);
    __syncthreads();
#pragma unroll
    for (int kc = 0; kc < 2; ++kc)
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) a_cur[kc][mt] = a_nxt[kc][mt];
  }

  // ---- layer-1 epilogue: h (fp32, skip of layer 2), tanh factors f1 (fp32)
  float h[4][4][4], f1[4][4][4], f2[4][4][4], f3[4][4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const int j0 = 16 * (4 * w + mt) + 4 * g;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float tt = fast_tanh(acc[mt][s][v] + cst[0 * kH + j0 + v]);
        const float idt = cst[1 * kH + j0 + v];
        h[mt][s][v] = tt * idt;
        f1[mt][s][v] = idt * (1.f - tt * tt);
      }
      lds_put(s_h, w, mt, s, n, g, h[mt][s]);
    }
  }
  __syncthreads();

  // ---- layers 2 and 3 (H -> H with skip)
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    hidden_product(acc, frag + (layer == 0 ? p.w2f : p.w3f), s_h, w, lane, n, g);
    __syncthreads();  // every wavefront has read the layer input
    const float* bl = cst + (2 + 2 * layer) * kH;
    const float* il = cst + (3 + 2 * layer) * kH;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const int j0 = 16 * (4 * w + mt) + 4 * g;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float tt = fast_tanh(acc[mt][s][v] + bl[j0 + v]);
          const float idt = il[j0 + v];
          h[mt][s][v] += tt * idt;
          (layer == 0 ? f2 : f3)[mt][s][v] = idt * (1.f - tt * tt);
        }
        if (layer == 0) lds_put(s_h, w, mt, s, n, g, h[mt][s]);
      }
    }
    if (layer == 0) __syncthreads();
  }

  // ---- output layer and atomic energy; dz3 = w4 * f3 into the LDS buffer
  const float* w4 = cst + 6 * kH;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    float e = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const int j0 = 16 * (4 * w + mt) + 4 * g;
      float dz[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        e += h[mt][s][v] * w4[j0 + v];
        dz[v] = w4[j0 + v] * f3[mt][s][v];
      }
      lds_put(s_h, w, mt, s, n, g, dz);
    }
    e += __shfl_xor(e, 16, 64);
    e += __shfl_xor(e, 32, 64);
    if (g == 0) s_e[w][16 * s + n] = e;
  }
  __syncthreads();
  if (tid < kTile && tid < cnt) {
    const float e = s_e[0][tid] + s_e[1][tid] + s_e[2][tid] + s_e[3][tid] + cst[7 * kH];
    atom_energy[s_atom[tid]] = (double)e + ebias[t];
  }

  // ---- backward through layers 3 and 2: dh_{l-1} = dz_l W_l^T + dh_l, dz = dh * f
  // dh3 = w4 (dE/de = 1); the skip of layer 3 carries it to dh2 unchanged
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    hidden_product(acc, frag + (layer == 0 ? p.w3b : p.w2b), s_h, w, lane, n, g);
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const int j0 = 16 * (4 * w + mt) + 4 * g;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        float dz[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float dh;
          if (layer == 0) {
            dh = acc[mt][s][v] + w4[j0 + v];
            h[mt][s][v] = dh;  // h now holds dh2
            dz[v] = dh * f2[mt][s][v];
          } else {
            dh = acc[mt][s][v] + h[mt][s][v];
            dz[v] = dh * f1[mt][s][v];
          }
        }
        lds_put(s_h, w, mt, s, n, g, dz);
      }
    }
    __syncthreads();
  }

  // ---- dD = dz1 W1^T: B operands (dz1 of the 64 atoms, K = 256) in registers
  bf16x8 bz[8][4];
#pragma unroll
  for (int kc = 0; kc < 8; ++kc)
#pragma unroll
    for (int s = 0; s < 4; ++s) bz[kc][s] = lds_b(s_h, kHS, s, n, 32 * kc, g);
  const bf16x8* w1b = frag + p.w1b;
  int atom_s[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) atom_s[s] = s_atom[16 * s + n];
  const int mt_per_wave = (p.nm1 + 3) / 4;
  const int mt0 = w * mt_per_wave, mt1 = min(p.nm1, mt0 + mt_per_wave);
  // the next M tile's eight weight fragments are loaded one tile ahead
  bf16x8 wa[8];
  if (mt0 < mt1)
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) wa[kc] = w1b[((int64_t)mt0 * 8 + kc) * 64 + lane];
#pragma unroll 1
  for (int mt = mt0; mt < mt1; ++mt) {
    bf16x8 wn[8];
    const int mn = mt + 1 < mt1 ? mt + 1 : mt;
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) wn[kc] = w1b[((int64_t)mn * 8 + kc) * 64 + lane];
    f32x4 d[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) d[s] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 8; ++kc)
#pragma unroll
      for (int s = 0; s < 4; ++s) d[s] = mfma(wa[kc], bz[kc][s], d[s]);
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) wa[kc] = wn[kc];
    const int f = 16 * mt + 4 * g;
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (atom_s[s] >= 0)
        __builtin_nontemporal_store(d[s], reinterpret_cast<f32x4*>(ddesc + (int64_t)atom_s[s] * p.k1 + f));
  }
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_se_a_fitting_bf16(double* atom_energy, float* ddesc, const void* desc, const int* perm,
                                      const int* type_start, const void* frags, const float* consts,
                                      const double* ebias, int64_t nloc, int ntypes, int k1, void* stream) {
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= kMaxTypes, "ntypes out of range [1, 16]");
  DPMI_CHECK_ARG(k1 > 0 && k1 % 16 == 0, "descriptor width must be a positive multiple of 16");
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  if (nloc == 0) return DPMI_OK;
  FitParams p;
  std::memset(&p, 0, sizeof(p));
  p.ntypes = ntypes;
  p.k1 = k1;
  p.nk1 = (int)(2 * ceil_div(k1, kKB));
  p.nm1 = k1 / 16;
  // fragment blocks per type: W1^T (16 x nk1), W2^T, W3^T (16 x 8), W3, W2 (16 x 8), W1 (nm1 x 8)
  p.w1f = 0;
  p.w2f = p.w1f + (int64_t)16 * p.nk1 * 64;
  p.w3f = p.w2f + 16 * 8 * 64;
  p.w3b = p.w3f + 16 * 8 * 64;
  p.w2b = p.w3b + 16 * 8 * 64;
  p.w1b = p.w2b + 16 * 8 * 64;
  p.type_frags = p.w1b + (int64_t)p.nm1 * 8 * 64;
  // one workgroup per 64-atom tile of each type (upper bound; extra ones exit)
  const int64_t grid = ceil_div(nloc, kTile) + ntypes;
  hipLaunchKernelGGL(k_fit, dim3((unsigned)grid), dim3(256), 0, as_stream(stream), atom_energy, ddesc,
                     static_cast<const __bf16*>(desc), perm,
                     type_start, static_cast<const bf16x8*>(frags), consts, ebias, p);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_se_a_fitting_frag_count(int64_t* count, int k1) {
  DPMI_CHECK_ARG(count != nullptr && k1 > 0 && k1 % 16 == 0, "bad arguments");
  const int64_t nk1 = 2 * ceil_div(k1, kKB);
  *count = (16 * nk1 + 4 * 16 * 8 + (int64_t)(k1 / 16) * 8) * 64;
  return DPMI_OK;
}
