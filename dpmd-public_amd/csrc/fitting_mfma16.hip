// Fused energy fitting net on gfx950 MFMA (bf16 or fp16 operands, fp32
// accumulation): atomic energies and dE/dD of every local atom in one pass.
// The kernel is a template on the 16-bit operand type H (__bf16 or
// _Float16); both MFMA forms have the same lane layout.
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
// MI355X mapping.  A workgroup owns 32 atoms of one type (the list is sorted
// by type); its four wavefronts split every hidden layer by output feature:
// wavefront w owns features 64 w .. 64 w + 63 of the width padded to 256, so
// a feature's tanh factor, skip value and adjoint stay in the registers of
// one lane for the whole pass.  The per-lane state (four feature tiles x two
// 16-atom subtiles) fits the 256 registers of two waves per SIMD, and the
// workgroup's LDS (71 KB) lets two workgroups share a CU, so that one's
// barrier-separated phases (staging, epilogues) overlap the other's products.  Layers are computed transposed, y^T = W^T x^T,
// with v_mfma_f32_16x16x32_bf16 / _f16: the weight fragment is the A operand (M =
// output feature), the layer input is the B operand (K = input feature,
// N = atom), read from an LDS buffer [atom][feature] that every layer's
// output is written back to.  The D matrix streams through LDS in 64-wide K
// blocks (H, as the descriptor kernel writes it), prefetched one block ahead.  The
// backward pass runs the transposed products (dh_{l-1} = dz_l W_l^T + skip)
// through the same LDS buffer and ends with dD = dz1 W1^T, whose B operands
// (dz1 for the 32 atoms) stay in registers while W1's fragments stream from
// L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "dpmi_internal.h"
#include "embed_layout.h"

namespace dpmi {
namespace {

template <typename H>
using h8 = H __attribute__((ext_vector_type(8)));
template <typename H>
using h4 = H __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kH = 256;         // padded hidden width (4 feature groups x 64 features)
constexpr int kTile = 32;       // atoms per workgroup
constexpr int kThreads = 256;   // four wavefronts, one per feature group
constexpr int kSub = 2;         // 16-atom subtiles per wavefront (all of the tile)
constexpr int kKB = 64;         // layer-1 input features staged per step
constexpr int kXS = kKB + 8;    // LDS row stride (16-bit elements) of the staged D block
constexpr int kHS = kH + 8;     // LDS row stride (16-bit elements) of the hidden-layer buffer
constexpr int kMaxTypes = 16;
constexpr int kCst = 7 * kH + 8;  // b1 idt1 b2 idt2 b3 idt3 w4 [kH each], b4
constexpr int kMaxGR = 104;       // widest G^T R (M) the fused descriptor form stages in fp32
constexpr int kGRS = 4 * kMaxGR + 4;  // LDS floats per atom of the staged GR ([m][c], padded)
// 104 < M <= 128: GR staged in fp16 (fp32 would take 66 KB of LDS per
// workgroup, and two workgroups per CU would no longer fit)
constexpr int kMaxGR16 = 128;
constexpr int kGRS16 = 4 * kMaxGR16 + 8;  // LDS halves per atom of the fp16-staged GR
constexpr int kMaxTypesGR = 4;        // neighbor types whose partials the fused form sums

// Fragment offsets (in 64-lane fragments of 8 16-bit values) inside one type's block.
struct FitParams {
  int ntypes;
  int nloc;
  int M;       // fused descriptor form: G^T R width (4 n1), axis 16
  int n1;
  float inv_nnei;
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

__device__ inline f32x4 mfma(const h8<__bf16>& a, const h8<__bf16>& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline f32x4 mfma(const h8<_Float16>& a, const h8<_Float16>& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Sum over the four 16-lane groups, the same value in every lane: the gfx950
// row swaps (lanes l <-> l ^ 16, then l <-> l ^ 32) stay in the VALU.
__device__ inline float sum_lane_groups(float x) {
  const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  const float y = __uint_as_float(a[0]) + __uint_as_float(a[1]);
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(y), __float_as_uint(y), false, false);
  return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// B fragment of 16 atoms (rows 16s..16s+15 of an LDS [atom][stride] buffer),
// K = features kbase + 8g .. kbase + 8g + 7.
template <typename H>
__device__ inline h8<H> lds_b(const H* buf, int stride, int s, int n, int kbase, int g) {
  return *reinterpret_cast<const h8<H>*>(buf + (16 * s + n) * stride + kbase + 8 * g);
}

// Writes a layer output (D layout: lane -> atom n of subtile s, features
// 4g + v of M tile mt of feature group fg) to the LDS buffer as H.
template <typename H>
__device__ inline void lds_put(H* buf, int fg, int mt, int s, int n, int g, const float* v4) {
  h4<H> q;
  q[0] = (H)v4[0];
  q[1] = (H)v4[1];
  q[2] = (H)v4[2];
  q[3] = (H)v4[3];
  *reinterpret_cast<h4<H>*>(buf + (16 * s + n) * kHS + 16 * (4 * fg + mt) + 4 * g) = q;
}

// One hidden H -> H layer product for this wavefront's four M tiles (feature
// group fg) and its two subtiles (s0, s0 + 1): acc[mt][s] = sum_k
// W(frag)[feature][k] * in[atom][k].
template <typename H>
__device__ inline void hidden_product(f32x4 (&acc)[4][kSub], const h8<H>* __restrict__ frag, const H* s_h,
                                      int fg, int s0, int lane, int n, int g) {
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int s = 0; s < kSub; ++s) acc[mt][s] = f32x4{0.f, 0.f, 0.f, 0.f};
  // fragments two K chunks ahead, in a ring of three register sets (a loop
  // body of three chunks, so no set is copied into another)
  auto load = [&](int kc, h8<H>(&a)[4]) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) a[mt] = frag[((4 * fg + mt) * 8 + kc) * 64 + lane];
  };
  auto step = [&](int kc, const h8<H>(&cur)[4], h8<H>(&nxt2)[4]) {
    if (kc + 2 < 8) load(kc + 2, nxt2);
    h8<H> b[kSub];
#pragma unroll
    for (int s = 0; s < kSub; ++s) b[s] = lds_b(s_h, kHS, s0 + s, n, 32 * kc, g);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int s = 0; s < kSub; ++s) acc[mt][s] = mfma(cur[mt], b[s], acc[mt][s]);
  };
  h8<H> a0[4], a1[4], a2[4];
  load(0, a0);
  load(1, a1);
#pragma unroll 1
  for (int kc = 0; kc < 8; kc += 3) {
    step(kc, a0, a2);
    if (kc + 1 < 8) step(kc + 1, a1, a0);
    if (kc + 2 < 8) step(kc + 2, a2, a1);
  }
}

// With kGR > 0 the descriptor is formed inside the kernel (the fused
// descriptor form, axis 16): the workgroup stages GR = sum_t gr_part / nnei
// of its 32 atoms in LDS ([atom][m][c]; fp32 for kGR = 1, M <= 104, fp16 for
// kGR = 2, M <= 128, products and sums in fp32 either way), builds each 64-wide block of
// D[m][a] = sum_c GR[c][m] GR[c][a] (se_a.py:851-858) into the layer-1
// staging buffer, and ends with dE/dGR instead of dE/dD:
//   dGR[c][m] = sum_a dD[m][a] GR[c][a] + [m < 16] sum_m' dD[m'][m] GR[c][m'],
// written / nnei in the embedding kernels' internal order as dgr
// [nloc, 128, 4] (component fastest: one 16-byte store per atom and feature).
template <typename H, int kGR>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_fit(double* __restrict__ atom_energy, float* __restrict__ dout, const H* __restrict__ desc,
          const float* __restrict__ gr_part, const int* __restrict__ perm, const int* __restrict__ type_start,
          const h8<H>* __restrict__ frag_all, const float* __restrict__ consts, const double* __restrict__ ebias,
          FitParams p) {
  // the layer-1 D blocks (four buffers) and, after layer 1, the hidden-layer
  // buffer share one region; the adjoint's dE/dGR stages reuse it at the end
  constexpr int kXH = 4 * kTile * kXS > kTile * kHS ? 4 * kTile * kXS : kTile * kHS;
  static_assert(kTile * kHS * 2 >= 2 * kTile * 16 * 16, "the dE/dGR stages must fit the shared region");
  __shared__ __attribute__((aligned(16))) H s_xh[kXH];
  auto s_x = [&](int b) { return s_xh + b * (kTile * kXS); };
  H* s_h = s_xh;
  __shared__ float s_e[4][kTile];
  __shared__ int s_atom[kTile];
  constexpr int kGRFloats = kGR == 1 ? kTile * kGRS : kGR == 2 ? kTile * kGRS16 / 2 : 4;
  __shared__ __attribute__((aligned(16))) float s_gr[kGRFloats];

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

  // wavefront w: feature group fg = w (features 64 fg .. 64 fg + 63), subtiles s0, s0 + 1
  const int tid = threadIdx.x, w = wave_id(), lane = tid & 63, n = lane & 15, g = lane >> 4;
  const int fg = w & 3, s0 = 0;
  if (tid < kTile) s_atom[tid] = tid < cnt ? perm[start + tid] : -1;
  const h8<H>* frag = frag_all + (int64_t)t * p.type_frags;
  const float* cst = consts + (int64_t)t * kCst;
  __syncthreads();

  // ---- layer 1: D block loads (thread -> atom tid/8, features 8*(tid%8) .. +7 of the block, H)
  const int xl = tid >> 3, xa = s_atom[xl], xq = tid & 7;
  const H* xrow = desc + (int64_t)(xa < 0 ? 0 : xa) * p.k1;
  const int nkb = p.nk1 / 2;
  auto load_block = [&](int kb, h8<H>& r) {
    const int f = kb * kKB + 8 * xq;
    // streamed once: non-temporal, so it does not evict the weight fragments from L2
    r = (xa >= 0 && f < p.k1) ? __builtin_nontemporal_load(reinterpret_cast<const h8<H>*>(xrow + f)) : h8<H>{};
  };
  auto store_block = [&](int buf, const h8<H>& r) {
    *reinterpret_cast<h8<H>*>(s_x(buf) + xl * kXS + 8 * xq) = r;
  };
  // GR[atom][m][0..3] of the staged tile, as fp32 or fp16 words
  auto gr_ld = [&](int atom, int m) -> float4 {
    if constexpr (kGR == 2) {
      const h4<_Float16> q = *reinterpret_cast<const h4<_Float16>*>(
          reinterpret_cast<const _Float16*>(s_gr) + atom * kGRS16 + 4 * m);
      return make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
    } else {
      return reinterpret_cast<const float4*>(s_gr)[atom * (kGRS / 4) + m];
    }
  };
  auto gr_st = [&](int atom, int m, const float4& x) {
    if constexpr (kGR == 2) {
      h4<_Float16> q;
      q[0] = (_Float16)x.x;
      q[1] = (_Float16)x.y;
      q[2] = (_Float16)x.z;
      q[3] = (_Float16)x.w;
      *reinterpret_cast<h4<_Float16>*>(reinterpret_cast<_Float16*>(s_gr) + atom * kGRS16 + 4 * m) = q;
    } else {
      reinterpret_cast<float4*>(s_gr)[atom * (kGRS / 4) + m] = x;
    }
  };
  // D block kb of atom xl, features (m = 4 kb + 2 mh + ml, a = 4 aq .. 4 aq + 3)
  // with aq = xq % 4, mh = xq / 4: the thread's GR[.][a] (ga) stay in
  // registers through layer 1
  const int aq = xq & 3, mh = xq >> 2;
  float4 ga[4];
  auto gen_block = [&](int kb, h8<H>& r) {
#pragma unroll
    for (int ml = 0; ml < 2; ++ml) {
      const float4 gm = gr_ld(xl, 4 * kb + 2 * mh + ml);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        r[4 * ml + j] = (H)(gm.x * ga[j].x + gm.y * ga[j].y + gm.z * ga[j].z + gm.w * ga[j].w);
    }
  };
  auto store_gen = [&](int buf, const h8<H>& r) {
    // r[4 ml + j] -> feature 16 (2 mh + ml) + 4 aq + j
    H* dst = s_x(buf) + xl * kXS + 32 * mh + 4 * aq;
#pragma unroll
    for (int ml = 0; ml < 2; ++ml) {
      h4<H> q;
#pragma unroll
      for (int j = 0; j < 4; ++j) q[j] = r[4 * ml + j];
      *reinterpret_cast<h4<H>*>(dst + 16 * ml) = q;
    }
  };

  f32x4 acc[4][kSub];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int s = 0; s < kSub; ++s) acc[mt][s] = f32x4{0.f, 0.f, 0.f, 0.f};
  if constexpr (kGR > 0) {
    // all of the tile's GR at once: work item (atom la, feature m), kBatch
    // items per thread in flight; the four components of every type are read
    // as one float4 (gr_part [t][i][m][c]: consecutive lanes read consecutive
    // 16-byte words) and written as one float4 GR[.][m] (consecutive lanes
    // write consecutive LDS addresses: no bank conflicts) -> s_gr[atom][m][c]
    constexpr int kBatch = 4;
    const int nit = kTile * p.M;
    const float inv_m = 1.f / (float)p.M;
#pragma unroll 1
    for (int k0 = tid; k0 < nit; k0 += kThreads * kBatch) {
      f32x4 x[kBatch][kMaxTypesGR];
#pragma unroll
      for (int b = 0; b < kBatch; ++b) {
        const int k = k0 + kThreads * b;
        const int la = k < nit ? __float2int_rz(((float)k + 0.5f) * inv_m) : 0;
        const int ai = k < nit ? s_atom[la] : -1;
        const int m = k - la * p.M;
#pragma unroll
        for (int tt = 0; tt < kMaxTypesGR; ++tt) {
          const f32x4* src = reinterpret_cast<const f32x4*>(gr_part) + ((int64_t)tt * p.nloc + (ai < 0 ? 0 : ai)) * p.M + m;
          x[b][tt] = (ai >= 0 && tt < p.ntypes) ? __builtin_nontemporal_load(src) : f32x4{0.f, 0.f, 0.f, 0.f};
        }
      }
#pragma unroll
      for (int b = 0; b < kBatch; ++b) {
        const int k = k0 + kThreads * b;
        if (k >= nit) break;
        const int la = __float2int_rz(((float)k + 0.5f) * inv_m), m = k - la * p.M;
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          v[c] = x[b][0][c];
#pragma unroll
          for (int tt = 1; tt < kMaxTypesGR; ++tt) v[c] += x[b][tt][c];
        }
        gr_st(la, m, make_float4(v[0] * p.inv_nnei, v[1] * p.inv_nnei, v[2] * p.inv_nnei, v[3] * p.inv_nnei));
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) ga[j] = gr_ld(xl, 4 * aq + j);
    h8<H> r;
    gen_block(0, r);
    store_gen(0, r);
    if (nkb > 1) {
      gen_block(1, r);
      store_gen(1, r);
    }
  } else {
    h8<H> r;
    load_block(0, r);
    store_block(0, r);
    if (nkb > 1) {
      load_block(1, r);
      store_block(1, r);
    }
  }
  __syncthreads();
  const h8<H>* w1f = frag + p.w1f;
  // weight fragments are loaded two K blocks ahead (an L2 round trip is
  // longer than one block's products)
  auto load_w1 = [&](int kb, h8<H> (&a)[2][4]) {
#pragma unroll
    for (int kc = 0; kc < 2; ++kc)
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) a[kc][mt] = w1f[((4 * fg + mt) * p.nk1 + 2 * kb + kc) * 64 + lane];
  };
  // one K block: products with the fragments cur, block kb + 2's fragments
  // into nxt2 (a ring of three register sets over a loop body of three
  // blocks, so no set is copied into another).  The D blocks are in a ring
  // of four LDS buffers, block kb + 2 written while block kb is read, so
  // the workgroup meets at a barrier after every second block: a buffer is
  // rewritten two blocks after it was read, with a barrier in between.
  auto l1_step = [&](int kb, const h8<H>(&cur)[2][4], h8<H>(&nxt2)[2][4]) {
    h8<H> r;
    const bool more = kb + 2 < nkb;
    // fused form: D block kb + 2 is built from the staged GR
    if constexpr (kGR > 0) {
      if (more) gen_block(kb + 2, r);
    } else {
      if (more) load_block(kb + 2, r);
    }
    load_w1(kb + 2 < nkb ? kb + 2 : nkb - 1, nxt2);
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
      h8<H> b[kSub];
#pragma unroll
      for (int s = 0; s < kSub; ++s) b[s] = lds_b(s_x(kb & 3), kXS, s0 + s, n, 32 * kc, g);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int s = 0; s < kSub; ++s) acc[mt][s] = mfma(cur[kc][mt], b[s], acc[mt][s]);
    }
    if constexpr (kGR > 0) {
      if (more) store_gen((kb + 2) & 3, r);
    } else {
      if (more) store_block((kb + 2) & 3, r);
    }
    if (kb & 1) __syncthreads();  // uniform
  };
  h8<H> a0[2][4], a1[2][4], a2[2][4];
  load_w1(0, a0);
  load_w1(nkb > 1 ? 1 : 0, a1);
#pragma unroll 1
  for (int kb = 0; kb < nkb; kb += 3) {
    l1_step(kb, a0, a2);
    if (kb + 1 < nkb) l1_step(kb + 1, a1, a0);
    if (kb + 2 < nkb) l1_step(kb + 2, a2, a1);
  }
  __syncthreads();  // every wavefront is done with the D blocks
  // ---- layer-1 epilogue: h (fp32, skip of layer 2), tanh factors f1 (fp32)
  float h[4][kSub][4], f1[4][kSub][4], f2[4][kSub][4], f3[4][kSub][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const int j0 = 16 * (4 * fg + mt) + 4 * g;
#pragma unroll
    for (int s = 0; s < kSub; ++s) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float tt = fast_tanh(acc[mt][s][v] + cst[0 * kH + j0 + v]);
        const float idt = cst[1 * kH + j0 + v];
        h[mt][s][v] = tt * idt;
        f1[mt][s][v] = idt * (1.f - tt * tt);
      }
      lds_put(s_h, fg, mt, s0 + s, n, g, h[mt][s]);
    }
  }
  __syncthreads();

  // ---- layers 2 and 3 (H -> H with skip)
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    hidden_product(acc, frag + (layer == 0 ? p.w2f : p.w3f), s_h, fg, s0, lane, n, g);
    __syncthreads();  // every wavefront has read the layer input
    const float* bl = cst + (2 + 2 * layer) * kH;
    const float* il = cst + (3 + 2 * layer) * kH;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const int j0 = 16 * (4 * fg + mt) + 4 * g;
#pragma unroll
      for (int s = 0; s < kSub; ++s) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float tt = fast_tanh(acc[mt][s][v] + bl[j0 + v]);
          const float idt = il[j0 + v];
          h[mt][s][v] += tt * idt;
          (layer == 0 ? f2 : f3)[mt][s][v] = idt * (1.f - tt * tt);
        }
        if (layer == 0) lds_put(s_h, fg, mt, s0 + s, n, g, h[mt][s]);
      }
    }
    if (layer == 0) __syncthreads();
  }

  // ---- output layer and atomic energy; dz3 = w4 * f3 into the LDS buffer
  const float* w4 = cst + 6 * kH;
#pragma unroll
  for (int s = 0; s < kSub; ++s) {
    float e = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const int j0 = 16 * (4 * fg + mt) + 4 * g;
      float dz[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        e += h[mt][s][v] * w4[j0 + v];
        dz[v] = w4[j0 + v] * f3[mt][s][v];
      }
      lds_put(s_h, fg, mt, s0 + s, n, g, dz);
    }
    e += __shfl_xor(e, 16, 64);
    e += __shfl_xor(e, 32, 64);
    if (g == 0) s_e[fg][16 * (s0 + s) + n] = e;
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
    hidden_product(acc, frag + (layer == 0 ? p.w3b : p.w2b), s_h, fg, s0, lane, n, g);
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const int j0 = 16 * (4 * fg + mt) + 4 * g;
#pragma unroll
      for (int s = 0; s < kSub; ++s) {
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
        lds_put(s_h, fg, mt, s0 + s, n, g, dz);
      }
    }
    __syncthreads();
  }

  // ---- dD = dz1 W1^T: B operands (dz1 of the wavefront's 32 atoms, K = 256)
  // in registers; the feature group picks the M tiles
  h8<H> bz[8][kSub];
#pragma unroll
  for (int kc = 0; kc < 8; ++kc)
#pragma unroll
    for (int s = 0; s < kSub; ++s) bz[kc][s] = lds_b(s_h, kHS, s0 + s, n, 32 * kc, g);
  const h8<H>* w1b = frag + p.w1b;
  int atom_s[kSub];
#pragma unroll
  for (int s = 0; s < kSub; ++s) atom_s[s] = s_atom[16 * (s0 + s) + n];
  // lane group g < kSub stores the first term of subtile s0 + g
  const int atom_g = g < kSub ? s_atom[16 * (s0 + g) + n] : -1;
  const int mt_per_wave = (p.nm1 + 3) / 4;
  const int mt0 = fg * mt_per_wave, mt1 = min(p.nm1, mt0 + mt_per_wave);
  // fused form: GR[.][a] of the lane's atoms (a = 4 g + v) in registers,
  // acc2 = this wavefront's part of sum_m dD[m][a] GR[.][m]; the shared
  // region is reused (after the barrier) for the first term of m < 16 and
  // for the reduction over the feature groups
  float4 gra[kSub][4], acc2[kSub][4];
  float4* t1buf = reinterpret_cast<float4*>(s_xh);
  float4* red = reinterpret_cast<float4*>(s_xh) + kTile * 16;
  if constexpr (kGR > 0) {
#pragma unroll
    for (int s = 0; s < kSub; ++s)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        gra[s][v] = gr_ld(16 * (s0 + s) + n, 4 * g + v);
        acc2[s][v] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    __syncthreads();  // every wavefront has read dz1 from s_h
  }
  // the M tiles' eight weight fragments are loaded two tiles ahead (one
  // tile's products are shorter than an L2 round trip)
  h8<H> wa[8], wb[8];
#pragma unroll
  for (int kc = 0; kc < 8; ++kc) {
    wa[kc] = mt0 < mt1 ? w1b[((int64_t)mt0 * 8 + kc) * 64 + lane] : h8<H>{};
    wb[kc] = mt0 + 1 < mt1 ? w1b[((int64_t)(mt0 + 1) * 8 + kc) * 64 + lane] : h8<H>{};
  }
#pragma unroll 1
  for (int mt = mt0; mt < mt1; ++mt) {
    h8<H> wn[8];
    const int mn = mt + 2 < mt1 ? mt + 2 : mt;
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) wn[kc] = w1b[((int64_t)mn * 8 + kc) * 64 + lane];
    f32x4 d[kSub];
#pragma unroll
    for (int s = 0; s < kSub; ++s) d[s] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 8; ++kc)
#pragma unroll
      for (int s = 0; s < kSub; ++s) d[s] = mfma(wa[kc], bz[kc][s], d[s]);
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) {
      wa[kc] = wb[kc];
      wb[kc] = wn[kc];
    }
    if constexpr (kGR > 0) {
      // M tile mt is m = mt, its rows a = 4 g + v (axis 16)
      const int q3 = mt >= 16 ? q3_of(mt, p.n1) : 0;
      float4 p1[kSub];
#pragma unroll
      for (int s = 0; s < kSub; ++s) {
        const float4 gm = gr_ld(16 * (s0 + s) + n, mt);
        p1[s] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float x = d[s][v];
          p1[s].x += x * gra[s][v].x;
          p1[s].y += x * gra[s][v].y;
          p1[s].z += x * gra[s][v].z;
          p1[s].w += x * gra[s][v].w;
          acc2[s][v].x += x * gm.x;
          acc2[s][v].y += x * gm.y;
          acc2[s][v].z += x * gm.z;
          acc2[s][v].w += x * gm.w;
        }
        // first term: sum over a = the four lane groups
        p1[s].x = sum_lane_groups(p1[s].x);
        p1[s].y = sum_lane_groups(p1[s].y);
        p1[s].z = sum_lane_groups(p1[s].z);
        p1[s].w = sum_lane_groups(p1[s].w);
      }
      // lane group g < kSub holds atom 16 (s0 + g) + n: one 16-byte
      // (four-component) store per lane
      const float4 x = g == 0 ? p1[0] : p1[1];
      if (atom_g >= 0) {
        if (mt >= 16)
          *reinterpret_cast<float4*>(dout + ((int64_t)atom_g * 128 + q3) * 4) =
              make_float4(x.x * p.inv_nnei, x.y * p.inv_nnei, x.z * p.inv_nnei, x.w * p.inv_nnei);
        else
          t1buf[(16 * (s0 + g) + n) * 16 + mt] = x;
      }
    } else {
      const int f = 16 * mt + 4 * g;
#pragma unroll
      for (int s = 0; s < kSub; ++s)
        if (atom_s[s] >= 0)
          __builtin_nontemporal_store(d[s], reinterpret_cast<f32x4*>(dout + (int64_t)atom_s[s] * p.k1 + f));
    }
  }
  if constexpr (kGR > 0) {
    // second term summed over the four feature groups in a fixed order
#pragma unroll 1
    for (int ph = 0; ph < 4; ++ph) {
      if (fg == ph)
#pragma unroll
        for (int s = 0; s < kSub; ++s)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            float4* r4 = red + (16 * (s0 + s) + n) * 16 + 4 * g + v;
            float4 x = acc2[s][v];
            if (ph > 0) {
              const float4 y = *r4;
              x.x += y.x;
              x.y += y.y;
              x.z += y.z;
              x.w += y.w;
            }
            *r4 = x;
          }
      __syncthreads();
    }
    // dGR for m < 16 (both terms) and zeros in the padding slots
    for (int k = tid; k < kTile * 128; k += kThreads) {
      const int la = k >> 7, q = k & 127, ai = s_atom[la];
      if (ai < 0) continue;
      const int m = orig_of_q3(q, p.n1);
      if (m >= 16) continue;
      float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
      if (m >= 0) {
        const float4 y = red[la * 16 + m], z = t1buf[la * 16 + m];
        x = make_float4((y.x + z.x) * p.inv_nnei, (y.y + z.y) * p.inv_nnei, (y.z + z.z) * p.inv_nnei,
                        (y.w + z.w) * p.inv_nnei);
      }
      *reinterpret_cast<float4*>(dout + ((int64_t)ai * 128 + q) * 4) = x;
    }
  }
}

// Fragment blocks per type: W1^T (16 x nk1), W2^T, W3^T (16 x 8), W3, W2 (16 x 8), W1 (nm1 x 8).
void set_offsets(FitParams& p) {
  p.nk1 = (int)(2 * ceil_div(p.k1, kKB));
  p.nm1 = p.k1 / 16;
  p.w1f = 0;
  p.w2f = p.w1f + (int64_t)16 * p.nk1 * 64;
  p.w3f = p.w2f + 16 * 8 * 64;
  p.w3b = p.w3f + 16 * 8 * 64;
  p.w2b = p.w3b + 16 * 8 * 64;
  p.w1b = p.w2b + 16 * 8 * 64;
  p.type_frags = p.w1b + (int64_t)p.nm1 * 8 * 64;
}

template <typename H>
int fitting(double* atom_energy, float* ddesc, const void* desc, const int* perm, const int* type_start,
            const void* frags, const float* consts, const double* ebias, int64_t nloc, int ntypes, int k1,
            void* stream) {
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= kMaxTypes, "ntypes out of range [1, 16]");
  DPMI_CHECK_ARG(k1 > 0 && k1 % 16 == 0, "descriptor width must be a positive multiple of 16");
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  if (nloc == 0) return DPMI_OK;
  FitParams p;
  std::memset(&p, 0, sizeof(p));
  p.ntypes = ntypes;
  p.k1 = k1;
  set_offsets(p);
  // one workgroup per 64-atom tile of each type (upper bound; extra ones exit)
  const int64_t grid = ceil_div(nloc, kTile) + ntypes;
  hipLaunchKernelGGL((k_fit<H, 0>), dim3((unsigned)grid), dim3(kThreads), 0, as_stream(stream), atom_energy, ddesc,
                     static_cast<const H*>(desc), nullptr, perm, type_start, static_cast<const h8<H>*>(frags),
                     consts, ebias, p);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename H>
int fitting_gr(double* atom_energy, float* dgr, const float* gr_part, const int* perm, const int* type_start,
               const void* frags, const float* consts, const double* ebias, int64_t nloc, int ntypes, int n1,
               int nnei, void* stream) {
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= kMaxTypesGR, "fused descriptor form: ntypes out of range [1, 4]");
  DPMI_CHECK_ARG(n1 >= 4 && 4 * n1 <= kMaxGR16, "fused descriptor form needs 16 <= 4 n1 <= 128");
  DPMI_CHECK_ARG(nnei > 0, "nnei must be positive");
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  if (nloc == 0) return DPMI_OK;
  FitParams p;
  std::memset(&p, 0, sizeof(p));
  p.ntypes = ntypes;
  p.nloc = (int)nloc;
  p.M = 4 * n1;
  p.n1 = n1;
  p.inv_nnei = 1.f / (float)nnei;
  p.k1 = 16 * p.M;
  set_offsets(p);
  const int64_t grid = ceil_div(nloc, kTile) + ntypes;
  if (p.M <= kMaxGR)
    hipLaunchKernelGGL((k_fit<H, 1>), dim3((unsigned)grid), dim3(kThreads), 0, as_stream(stream), atom_energy, dgr,
                       nullptr, gr_part, perm, type_start, static_cast<const h8<H>*>(frags), consts, ebias, p);
  else
    hipLaunchKernelGGL((k_fit<H, 2>), dim3((unsigned)grid), dim3(kThreads), 0, as_stream(stream), atom_energy, dgr,
                       nullptr, gr_part, perm, type_start, static_cast<const h8<H>*>(frags), consts, ebias, p);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_se_a_fitting_bf16(double* atom_energy, float* ddesc, const void* desc, const int* perm,
                                      const int* type_start, const void* frags, const float* consts,
                                      const double* ebias, int64_t nloc, int ntypes, int k1, void* stream) {
  return fitting<__bf16>(atom_energy, ddesc, desc, perm, type_start, frags, consts, ebias, nloc, ntypes, k1, stream);
}

extern "C" int dpmi_se_a_fitting_f16(double* atom_energy, float* ddesc, const void* desc, const int* perm,
                                     const int* type_start, const void* frags, const float* consts,
                                     const double* ebias, int64_t nloc, int ntypes, int k1, void* stream) {
  return fitting<_Float16>(atom_energy, ddesc, desc, perm, type_start, frags, consts, ebias, nloc, ntypes, k1,
                           stream);
}

extern "C" int dpmi_se_a_fitting_gr_bf16(double* atom_energy, float* dgr, const float* gr_part, const int* perm,
                                         const int* type_start, const void* frags, const float* consts,
                                         const double* ebias, int64_t nloc, int ntypes, int n1, int nnei,
                                         void* stream) {
  return fitting_gr<__bf16>(atom_energy, dgr, gr_part, perm, type_start, frags, consts, ebias, nloc, ntypes, n1,
                            nnei, stream);
}

extern "C" int dpmi_se_a_fitting_gr_f16(double* atom_energy, float* dgr, const float* gr_part, const int* perm,
                                        const int* type_start, const void* frags, const float* consts,
                                        const double* ebias, int64_t nloc, int ntypes, int n1, int nnei,
                                        void* stream) {
  return fitting_gr<_Float16>(atom_energy, dgr, gr_part, perm, type_start, frags, consts, ebias, nloc, ntypes, n1,
                              nnei, stream);
}

extern "C" int dpmi_se_a_fitting_frag_count(int64_t* count, int k1) {
  DPMI_CHECK_ARG(count != nullptr && k1 > 0 && k1 % 16 == 0, "bad arguments");
  const int64_t nk1 = 2 * ceil_div(k1, kKB);
  *count = (16 * nk1 + 4 * 16 * 8 + (int64_t)(k1 / 16) * 8) * 64;
  return DPMI_OK;
}
