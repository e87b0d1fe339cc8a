// Fused se_e2_a embedding net + G^T R on gfx950 MFMA (bf16 or fp16 operands,
// fp32 accumulation), its recompute-based adjoint, and the (G^T R)(G^T R<)^T
// descriptor with its adjoint.  The kernels are templates on the 16-bit
// operand type H (__bf16 or _Float16); both MFMA forms have the same lane
// layout.
//
// Reference semantics: DescrptBlockSeA.forward (deepmd/pt/model/descriptor/
// se_a.py:802-858) with the embedding layers of MLPLayer.forward
// (deepmd/pt/model/network/mlp.py:188-219), type_one_side: for neighbor type t
// the rows are the sel_t slots of every atom, s = em[..., 0],
//   h1 = tanh(s w1 + b1) * idt1,           (1 -> n1)
//   h2 = tanh(h1 W2 + b2) * idt2 + [h1,h1], (n1 -> 2 n1)
//   G  = tanh(h2 W3 + b3) * idt3 + [h2,h2], (2 n1 -> 4 n1 = M)
//   GR = sum_t em_t^T G_t / nnei,   D = GR^T GR[:, :axis].
//
// MI355X mapping.  One 64-lane wavefront owns one atom's section of one
// neighbor type (four atoms of the same type per workgroup); rows are
// processed 16 at a time.  With v_mfma_f32_16x16x32_bf16 / _f16 (A: lane -> M row
// lane%16, K 8*(lane/16)+u; B: lane -> N col lane%16, same K; D: lane -> col
// lane%16, rows 4*(lane/16)+v) each layer is computed transposed,
// y^T = W^T x^T, so the neighbor row is the N (lane) index everywhere and the
// D fragment of one layer is directly the B fragment of the next: the K order
// inside a 32-wide chunk is the permutation phi(g,u) = 16*(u/4) + 4g + u%4,
// baked into the pre-arranged weight fragments.  Features use an internal
// order in which the [x, x] skip connections land on the same lane and
// register: q1 = f (f < n1), q2 = q1(f mod n1) + 32*[f >= n1],
// q3 = q2(f mod 2n1) + 64*[f >= 2n1]; padded features are exact zeros.
// G itself is never formed: with r3 = 1 / (1 + exp(2 x3)), tanh = 1 - 2 r3 and
// G = idt3 (1 - 2 r3) + [h2, h2], so G^T R = idt3 (em^T 1) - 2 idt3 (em^T r3)
// + em^T [h2, h2]; layer 3 is computed transposed, so r3 comes out with the
// rows in the K position of the em^T r3 product, h2 is transposed per
// 16-row tile through LDS (ds_read_b64_tr_b16).  The backward pass
// recomputes the forward per tile together with its tangent in the scalar
// input s (forward-mode: the same fragments, the same register layout) and
// contracts both with dGR, so it needs no transposes and no adjoint chain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "dpmi_internal.h"
#include "embed_layout.h"

namespace dpmi {
namespace {

template <typename H>
using h8 = H __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxTypes = 16;
constexpr int kFragFwd = 20;   // 4 (layer 2) + 16 (layer 3), per neighbor type
constexpr int kConsts = 576;   // k w1, k b1, idt1 [32]; k b2, idt2 [64]; k b3, idt3 [128]; idt1 w1 [32];
                               // idt2 / k [64]   (k = 2 log2 e)
constexpr int kQ3 = 128;       // internal width of G
constexpr int kAtoms = 4;      // atoms (wavefronts) per workgroup
constexpr int kH2P = 80;       // row pitch (16-bit elements) of a tile's h2 in LDS: conflict-free transposed reads

struct EmbedParams {
  int ntypes;
  int nnei;
  int n1;
  int sec[kMaxTypes + 1];
  // optional padding shortcut: counts[i * ntypes + t] = listed neighbors of
  // atom i in section t (the rows after them are padding slots whose em row
  // is pad_em[(type(i) * ntypes + t) * 4 ..], the same in every such slot)
  const int16_t* counts;
  const int* atype;
  const float* pad_em;
};

// tanh(x) from z = 2 log2(e) x: 1 - 2 / (2^z + 1), exact limits at +-inf.
// The forward weights, biases and layer-1 constants are stored pre-scaled by
// 2 log2(e), so the MFMA (with the bias as its accumulator input) yields z.
__device__ inline float tanh_scaled(float z) {
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z));
}

__device__ inline f32x4 mfma(const h8<__bf16>& a, const h8<__bf16>& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline f32x4 mfma(const h8<_Float16>& a, const h8<_Float16>& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

template <typename H>
using h4 = H __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
__device__ inline f32x4 mfma16(const h4<__bf16>& a, const h4<__bf16>& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s16x4, a), __builtin_bit_cast(s16x4, b), c, 0,
                                                    0, 0);
}
__device__ inline f32x4 mfma16(const h4<_Float16>& a, const h4<_Float16>& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// Layer-1 feature owned by lane group g at B-slot u (also the K order phi).
__device__ inline int phi(int g, int u) { return 16 * (u >> 2) + 4 * g + (u & 3); }

// The forward chain of one 16-row tile.  Layers 1 and 2 keep the row as the
// lane index (N); layer 3 is computed transposed, z3^T = h2^T (k W3): h2's
// registers serve as the A operand (M = row, the same K slots) and the
// layer-3 weight fragment as the B operand (N = feature, the same lane
// contents), so r3 comes out with the feature as the lane index and lane
// group g holding rows 4 g .. 4 g + 3 - the K = 16 operand of the em^T r3
// product, with no transpose.  Layer 3 is returned as r3 = 1 / (1 + 2^z3)
// only: with tanh = 1 - 2 r3 the embedding output is G = idt3 (1 - 2 r3) +
// [h2, h2], and the kernel contracts r3 and h2 separately (G^T R = idt3
// (sum_rows em) - 2 idt3 (em^T r3) + em^T [h2, h2]), so G is never formed.
template <typename H>
struct TileFwd {
  h8<H> b2[2];   // h2 in H: features 32 kc + phi(g, u) of row lane % 16
  float r3t[32];  // slot mt * 4 + v: feature 16 mt + lane % 16 of row 4 g + v
};

template <typename H, bool kSkip>
__device__ inline void forward_tile(float s, const h8<H>* __restrict__ frag, const float* __restrict__ cst,
                                    const f32x4* __restrict__ b3x4, int g, int lane, TileFwd<H>& f) {
  h8<H> b1;
  float h1[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    if (kSkip && u == 7) {  // empty slot in every lane group
      h1[u] = 0.f;
      b1[u] = (H)0.f;
      continue;
    }
    const int q = phi(g, u);
    h1[u] = tanh_scaled(s * cst[q] + cst[32 + q]) * cst[64 + q];
    b1[u] = (H)h1[u];
  }
  f32x4 z2[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
    z2[mt] = mfma(frag[mt * 64 + lane], b1, *reinterpret_cast<const f32x4*>(cst + 96 + 16 * mt + 4 * g));
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const int q = 16 * mt + 4 * g;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (kSkip && (mt & 1) && v == 3) {
        f.b2[mt >> 1][(mt & 1) * 4 + v] = (H)0.f;
        continue;
      }
      f.b2[mt >> 1][(mt & 1) * 4 + v] = (H)(tanh_scaled(z2[mt][v]) * cst[160 + q + v] + h1[(mt & 1) * 4 + v]);
    }
  }
  const int n = lane & 15;
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    // fragments read per pair of output tiles, not all ahead of the products
    // (sixteen fragments in flight would hold 64 registers)
    if ((mt & 1) == 0) asm volatile("" ::: "memory");
    f32x4 z = mfma(f.b2[0], frag[(4 + 2 * mt) * 64 + lane], b3x4[16 * mt + n]);
    z = mfma(f.b2[1], frag[(5 + 2 * mt) * 64 + lane], z);
#pragma unroll
    for (int v = 0; v < 4; ++v) f.r3t[mt * 4 + v] = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z[v]));
  }
}

// Row stream.  A wavefront takes a contiguous range of atoms for one
// neighbor type t and walks their rows of section t in atom order: the
// listed rows with counts (the padding shortcut), all sel rows without.  It
// reads up to 64 atoms at a time (a window): the window's row counts sit one
// per lane with their inclusive prefix sum, so the start and end of any atom
// of the window are readlanes away.  The stream is cut into chunks of at most
// 32 rows holding rows of at most four consecutive atoms; atom k of the
// window owns slot k % 4, and the em^T products use M row 4 slot + c, so the
// four 16-lane groups of the accumulators hold four atoms (instead of one
// atom in rows c < 4 and twelve unused rows).  Tiles are full except at the
// end of a range or where a chunk closes on its fourth atom, so the tanh
// chain runs on listed rows instead of whole 16-row tiles per atom.
constexpr int kSlots = 4;

struct Chunk {
  int s0;          // stream position where atom k0 starts
  int e0, e1, e2;  // where atoms k0, k0 + 1, k0 + 2 end
  int p1;          // end of the chunk (its rows are p0 .. p1 - 1)
  int nfin;        // atoms k0 .. k0 + nfin - 1 end inside the chunk
};

// inclusive prefix of the window's row counts at atom k (the total past it)
__device__ inline int pre_at(int pre, int k, int n) { return __builtin_amdgcn_readlane(pre, min(k, n - 1)); }

// the chunk that starts at stream position p0 with atom k0 < n unfinished
// and may hold rows of atoms k0 .. klast (klast <= k0 + 3)
__device__ inline Chunk make_chunk(int pre, int k0, int p0, int n, int rows, int klast) {
  Chunk c;
  klast = min(klast, n - 1);
  c.s0 = k0 > 0 ? pre_at(pre, k0 - 1, n) : 0;
  c.e0 = pre_at(pre, k0, n);
  c.e1 = pre_at(pre, k0 + 1, n);
  c.e2 = pre_at(pre, k0 + 2, n);
  const int e3 = pre_at(pre, k0 + 3, n);
  c.p1 = min(p0 + rows, pre_at(pre, klast, n));
  c.nfin = (int)(c.e0 <= c.p1) + (int)(k0 + 1 <= klast && c.e1 <= c.p1) + (int)(k0 + 2 <= klast && c.e2 <= c.p1) +
           (int)(k0 + 3 <= klast && e3 <= c.p1);
  return c;
}

// the same with at most two atoms per chunk (the backward's 16-row tiles)
__device__ inline Chunk make_chunk2(int pre, int k0, int p0, int n, int rows) {
  Chunk c;
  c.s0 = k0 > 0 ? pre_at(pre, k0 - 1, n) : 0;
  c.e0 = pre_at(pre, k0, n);
  c.e1 = pre_at(pre, k0 + 1, n);
  c.e2 = c.e1;
  c.p1 = min(p0 + rows, c.e1);
  c.nfin = (int)(c.e0 <= c.p1) + (int)(k0 + 1 < n && c.e1 <= c.p1);
  return c;
}

// the atom (window index) of stream position pos inside chunk c, and its row
__device__ inline int chunk_atom(const Chunk& c, int k0, int pos, int* row) {
  const int kk = (int)(pos >= c.e0) + (int)(pos >= c.e1) + (int)(pos >= c.e2);
  const int start = kk == 0 ? c.s0 : kk == 1 ? c.e0 : kk == 2 ? c.e1 : c.e2;
  *row = pos - start;
  return k0 + kk;
}

// A workgroup's atoms [A0, A0 + nwg) (at most one per thread) split among
// its kW wavefronts by rows, not by atoms: with R rows in all, atom i (E_i
// rows before it) goes to wavefront min(kW - 1, kW E_i / R), so every
// wavefront gets a contiguous range of about R / kW rows and none waits long
// for the last one at the end.  Wavefront k takes atoms s_wstart[k] ..
// s_wstart[k + 1] - 1 (relative to A0).  Contains two workgroup barriers.
template <int kW>
__device__ inline void split_rows(const EmbedParams& p, int t, int sel, int A0, int nwg, int w, int lane,
                                  int* s_wsum, int* s_wstart) {
  auto count = [&](int i) {
    return i < nwg ? (p.counts != nullptr ? (int)p.counts[(int64_t)(A0 + i) * p.ntypes + t] : sel) : 0;
  };
  const int ia = threadIdx.x;
  const int cnt_i = count(ia);
  int inc = cnt_i;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(inc, off, 64);
    if (lane >= off) inc += y;
  }
  if (lane == 63) s_wsum[w] = inc;
  if (threadIdx.x <= kW) s_wstart[threadIdx.x] = nwg;
  __syncthreads();
  int base = 0, R = 0;
  for (int ww = 0; ww < kW; ++ww) {
    const int x = s_wsum[ww];
    base += ww < w ? x : 0;
    R += x;
  }
  const int ex = base + inc - cnt_i;  // E_i
  auto wave_of = [&](int e) { return R > 0 ? min(kW - 1, (int)(((int64_t)kW * e) / R)) : 0; };
  // atom i starts the wavefronts (wave_of(E_{i-1}), wave_of(E_i)]
  int e_prev = __shfl_up(ex, 1, 64);
  if (lane == 0) e_prev = ia > 0 ? base - count(ia - 1) : -1;
  if (ia < nwg) {
    const int k1 = wave_of(ex), k0 = ia > 0 ? wave_of(e_prev) : -1;
    for (int k = k0 + 1; k <= k1; ++k) s_wstart[k] = ia;
  }
  __syncthreads();
}

// Forward: GR partial [t][i][m][c] (original feature order m < 4 n1) of
// the atoms of each wavefront's range, over 16-row tiles of the row stream
// holding rows of up to four atoms.  A workgroup loads its type's weight
// fragments once; its four wavefronts stream independently (no block
// barrier after set-up).  G^T R contracts over neighbor rows: r3 comes out
// of the transposed layer-3 product with the feature as the lane index and
// lane group g holding rows 4g .. 4g + 3, the K = 16 operand of em^T r3; h2
// comes out of layer 2 with the row as the lane index, so the tile's h2 is
// stored to LDS as 16-bit [row][feature] and read back with the transposing
// ds_read_b64_tr_b16 into the same layout for em^T h2.  The A
// operand of these products and of em^T 1 holds em[row][c] in M row
// 4 slot(row) + c and zero in the other slots' rows.
template <typename H, bool kSkip>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
    k_embed_fwd(float* __restrict__ gr_part, const float* __restrict__ em, const h8<H>* __restrict__ frag_all,
                const float* __restrict__ consts, int nloc, int per_wave, EmbedParams p) {
  __shared__ h8<H> s_frag[kFragFwd * 64];
  __shared__ __attribute__((aligned(16))) float s_cst[kConsts];
  __shared__ __attribute__((aligned(16))) H s_em[kAtoms][16 * 16];  // em in H, [row][4 slot + c]
  // h2 of the tile in H, [row][feature] (pitch kH2P): read back transposed
  __shared__ __attribute__((aligned(16))) H s_h2[kAtoms][16 * kH2P];
  // layer-3 bias (k b3) of feature 16 mt + n in all four rows: the
  // accumulator input of the transposed product, one 16-byte read
  __shared__ f32x4 s_b3x4[8 * 16];
  const int t = blockIdx.y;
  const int w = wave_id(), lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const h8<H>* frag_t = frag_all + (int64_t)t * kFragFwd * 64;
  for (int k = threadIdx.x; k < kFragFwd * 64; k += blockDim.x) s_frag[k] = frag_t[k];
  for (int k = threadIdx.x; k < kConsts; k += blockDim.x) s_cst[k] = consts[(int64_t)t * kConsts + k];
  for (int k = threadIdx.x; k < 8 * 16; k += blockDim.x) {
    const float b = consts[(int64_t)t * kConsts + 224 + k];
    s_b3x4[k] = f32x4{b, b, b, b};
  }
  const float* cst = s_cst;
  const int sel = p.sec[t + 1] - p.sec[t];
  __syncthreads();
  // Padding shortcut: an atom's slots after its last listed neighbor all
  // hold the same em row (its center type's -davg/dstd); they are not
  // evaluated, their G^T R share is n_pad em_pad^T G_pad with G_pad
  // evaluated once per center type here (with the H roundings the tile path
  // applies).
  extern __shared__ float s_gpad[];  // [ntypes][kQ3] (dynamic: sized by the launch)
  if (p.counts != nullptr) {
    for (int ti = w; ti < p.ntypes; ti += kAtoms) {
      TileFwd<H> f;
      forward_tile<H, kSkip>(p.pad_em[(ti * p.ntypes + t) * 4], s_frag, cst, s_b3x4, g, lane, f);
      // every row holds the padding row: row 0 (lane group 0, v = 0) gives
      // r3 of feature 16 mt + r, row 0 (lane r = 0) h2 of 16 mt + 4 g + v
      if (g == 0)
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
          const int q = 16 * mt + r;
          s_gpad[ti * kQ3 + q] = cst[352 + q] * (1.f - 2.f * (float)(H)f.r3t[mt * 4]);
        }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (r == 0)
#pragma unroll
        for (int mt = 0; mt < 8; ++mt)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            s_gpad[ti * kQ3 + 16 * mt + 4 * g + v] += (float)f.b2[(mt & 3) >> 1][((mt & 3) & 1) * 4 + v];
    }
    __syncthreads();
  }
  H* sem = s_em[w];
  H* sh2 = s_h2[w];
  const float* em_t = em + (int64_t)p.sec[t] * 4;
  const int a_lo = min((blockIdx.x * kAtoms + w) * per_wave, nloc), a_hi = min(a_lo + per_wave, nloc);
  for (int wa = a_lo; wa < a_hi; wa += 64) {
    const int n = min(64, a_hi - wa);
    int cnt = 0;
    if (lane < n) cnt = p.counts != nullptr ? (int)p.counts[(int64_t)(wa + lane) * p.ntypes + t] : sel;
    int pre = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(pre, off, 64);
      if (lane >= off) pre += y;
    }
    // em row of this lane in the tile of chunk c (read one tile ahead)
    auto load_row = [&](const Chunk& c, int k0, int p0) {
      const int pos = p0 + r;
      int row;
      const int k = chunk_atom(c, k0, pos, &row);
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (pos < c.p1) v = *reinterpret_cast<const float4*>(em_t + ((int64_t)(wa + k) * p.nnei + row) * 4);
      return v;
    };
    // D(M = 4 slot + c, N = internal feature 16 nt + r): acc = em^T r3,
    // acc_h = em^T h2, acc_s = sum over rows of em (the same in every column)
    f32x4 acc[8], acc_h[4], acc_s = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc_h[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    // Finished atoms keep their accumulators until three have ended (or the
    // window is done), and are then written in one pass with one lane group
    // each; until then the tiles hold rows of atoms up to kf + 3 only, kf the
    // first atom not written yet.
    int k0 = 0, p0 = 0, kf = 0;
    Chunk c = make_chunk(pre, 0, 0, n, 16, 3);
    float4 ev_next = load_row(c, 0, 0);
    while (k0 < n) {
      const int k0n = k0 + c.nfin, p0n = c.p1;
      const bool flush = k0n >= n || k0n - kf >= 3;
      const int kfn = flush ? k0n : kf;
      const float4 ev = ev_next;
      Chunk cn = c;
      if (k0n < n) {
        cn = make_chunk(pre, k0n, p0n, n, 16, kfn + 3);
        ev_next = load_row(cn, k0n, p0n);
      }
      if (c.p1 > p0) {
        // this lane's row: its slot (-1 past the tile's rows)
        const int pos = p0 + r;
        int row;
        const int sl = pos < c.p1 ? (chunk_atom(c, k0, pos, &row) & (kSlots - 1)) : -1;
        // em staged as [row][4 slot + c]: lane group g writes the slot-g quarter of its row
        {
          h4<H> e4;
          e4[0] = (H)(sl == g ? ev.x : 0.f);
          e4[1] = (H)(sl == g ? ev.y : 0.f);
          e4[2] = (H)(sl == g ? ev.z : 0.f);
          e4[3] = (H)(sl == g ? ev.w : 0.f);
          *reinterpret_cast<h4<H>*>(sem + 16 * r + 4 * g) = e4;
        }
        // weights and constants stay in LDS (re-read per tile), not in registers
        asm volatile("" ::: "memory");
        TileFwd<H> f;
        forward_tile<H, kSkip>(ev.x, s_frag, cst, s_b3x4, g, lane, f);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // em^T as A (M = 4 slot + c, K = row 4 g + u)
        h4<H> a;
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = sem[16 * (4 * g + u) + r];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
          h4<H> b;
#pragma unroll
          for (int v = 0; v < 4; ++v) b[v] = (H)f.r3t[mt * 4 + v];
          acc[mt] = mfma16(a, b, acc[mt]);
        }
        // h2 (lane = row r, features 32 kc + 16 hh + 4 g + 0..3 in b2[kc]
        // slots 4 hh ..) into [row][feature], read back transposed: in lane
        // group G, lane i addresses row 4 G + i / 4, features 16 j + 4 (i % 4)
        // .. + 3, and ds_read_b64_tr_b16 hands lane i rows 4 G .. 4 G + 3 of
        // feature 16 j + i, the K = 16 operand of em^T h2
#pragma unroll
        for (int kc = 0; kc < 2; ++kc)
#pragma unroll
          for (int hh = 0; hh < 2; ++hh) {
            h4<H> x;
#pragma unroll
            for (int v = 0; v < 4; ++v) x[v] = f.b2[kc][4 * hh + v];
            *reinterpret_cast<h4<H>*>(sh2 + r * kH2P + 32 * kc + 16 * hh + 4 * g) = x;
          }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          typedef short s16x4_lds __attribute__((ext_vector_type(4)));
          auto* src = (__attribute__((address_space(3))) s16x4_lds*)(sh2 + (4 * g + (r >> 2)) * kH2P + 16 * j + 4 * (r & 3));
          const h4<H> b = __builtin_bit_cast(h4<H>, __builtin_amdgcn_ds_read_tr16_b64_v4i16(src));
          acc_h[j] = mfma16(a, b, acc_h[j]);
        }
        h4<H> ones;
#pragma unroll
        for (int u = 0; u < 4; ++u) ones[u] = (H)1.f;
        acc_s = mfma16(a, ones, acc_s);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      if (flush) {
        // the finished atoms kf .. k0n - 1: lane group g holds the atom of
        // slot g, which is kf + kk; its GR_t^T rows in the original feature order
        const int kk = (g - kf) & (kSlots - 1);
        const bool fin = kf + kk < k0n;
        const int cnt_a = __shfl(cnt, min(kf + kk, 63), 64);
        if (fin) {
          const int ai = wa + kf + kk;
          // (n1 passes through an opaque copy so the index arithmetic below is
          // not hoisted out of the tile loop, where registers are scarce)
          int n1 = p.n1;
          asm volatile("" : "+s"(n1));
          const int M = 4 * n1;
          float4* out = reinterpret_cast<float4*>(gr_part) + ((int64_t)t * nloc + ai) * M;
          // padding rows not evaluated: n_pad em_pad^T G_pad
          float padc[4] = {0.f, 0.f, 0.f, 0.f};
          int tp = 0;
          const int n_pad = sel - cnt_a;
          if (p.counts != nullptr && n_pad > 0) {
            const int ti = p.atype[ai];
            if (ti >= 0 && ti < p.ntypes) {
              tp = ti;
#pragma unroll
              for (int v = 0; v < 4; ++v) padc[v] = (float)n_pad * (float)(H)p.pad_em[(ti * p.ntypes + t) * 4 + v];
            }
          }
#pragma unroll
          for (int nt = 0; nt < 8; ++nt) {
            const int q = 16 * nt + r;
            const int m = orig_of_q3(q, n1);
            const float idt3 = cst[352 + q];
            const float gp = n_pad > 0 ? s_gpad[tp * kQ3 + q] : 0.f;
            float o[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) o[v] = idt3 * (acc_s[v] - 2.f * acc[nt][v]) + acc_h[nt & 3][v] + padc[v] * gp;
            if (m >= 0) out[m] = make_float4(o[0], o[1], o[2], o[3]);
          }
        }
        // the finished slots start the next atom from zero
#pragma unroll
        for (int v = 0; v < 4; ++v) {
#pragma unroll
          for (int nt = 0; nt < 8; ++nt) acc[nt][v] = fin ? 0.f : acc[nt][v];
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) acc_h[nt][v] = fin ? 0.f : acc_h[nt][v];
          acc_s[v] = fin ? 0.f : acc_s[v];
        }
      }
      k0 = k0n;
      p0 = p0n;
      kf = kfn;
      c = cn;
    }
  }
}

// GR = sum_t GR_t / nnei (partials [t][i][m][c]; GR [i][c][m] in the
// original order, saved for the backward) and
// D[m][a] = sum_c GR[c][m] GR[c][a] (se_a.py:851-858).
template <typename DT>
__global__ void __launch_bounds__(256)
    k_descrpt_fwd(float* __restrict__ gr, DT* __restrict__ desc, const float* __restrict__ gr_part, int nloc,
                  int ntypes, float inv_nnei, int n1, int axis) {
  __shared__ float s[kAtoms][4 * kQ3];
  const int w = wave_id(), lane = threadIdx.x & 63;
  const int i = blockIdx.x * kAtoms + w;
  if (i >= nloc) return;
  const int M = 4 * n1;
  for (int k = lane; k < 4 * M; k += 64) {  // k = 4 m + c
    float v = 0.f;
    for (int t = 0; t < ntypes; ++t) v += gr_part[((int64_t)t * nloc + i) * 4 * M + k];
    v *= inv_nnei;
    const int ck = (k & 3) * M + (k >> 2);
    s[w][ck] = v;
    gr[(int64_t)i * 4 * M + ck] = v;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int k = lane; k < M * axis; k += 64) {
    const int m = k / axis, a = k - m * axis;
    float d = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) d += s[w][c * M + m] * s[w][c * M + a];
    desc[(int64_t)i * M * axis + k] = (DT)d;
  }
}

// dGR[c][q3(k)] = (sum_a dD[k][a] GR[c][a] + [k < axis] sum_m dD[m][k] GR[c][m]) / nnei
// (GR in the original order, dGR in the internal order of the embedding kernels)
__global__ void __launch_bounds__(256)
    k_descrpt_bwd(float* __restrict__ dgr, const float* __restrict__ ddesc, const float* __restrict__ gr, int nloc,
                  float inv_nnei, int n1, int axis) {
  __shared__ float s[kAtoms][4 * kQ3];
  __shared__ float sd[kAtoms][128 * 16];
  const int w = wave_id(), lane = threadIdx.x & 63;
  const int i = blockIdx.x * kAtoms + w;
  if (i >= nloc) return;
  const int M = 4 * n1;
  for (int k = lane; k < 4 * M; k += 64) s[w][k] = gr[(int64_t)i * 4 * M + k];
  for (int k = lane; k < M * axis; k += 64) sd[w][k] = ddesc[(int64_t)i * M * axis + k];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int q = lane; q < kQ3; q += 64) {
    const int k = orig_of_q3(q, n1);
    float out[4] = {0.f, 0.f, 0.f, 0.f};
    if (k >= 0) {
      for (int a = 0; a < axis; ++a) {
        float dd = sd[w][k * axis + a];
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] += dd * s[w][c * M + a];
      }
      if (k < axis) {
        for (int m = 0; m < M; ++m) {
          float dd = sd[w][m * axis + k];
#pragma unroll
          for (int c = 0; c < 4; ++c) out[c] += dd * s[w][c * M + m];
        }
      }
    }
    *reinterpret_cast<float4*>(dgr + ((int64_t)i * kQ3 + q) * 4) =
        make_float4(out[0] * inv_nnei, out[1] * inv_nnei, out[2] * inv_nnei, out[3] * inv_nnei);
  }
}

// Same contraction for axis <= 16, one wavefront per atom, without cross-lane
// reductions in the loops: the atom's dD row is staged in LDS (row stride
// axis + 1, conflict-free column reads), the first term runs with one lane
// per m over a, the second with lane (m0, a) over m = m0 + j (64 / AXIS),
// and only its final sum crosses lanes.
template <int AXIS>
__global__ void __launch_bounds__(256)
    k_descrpt_bwd_lanes(float* __restrict__ dgr, const float* __restrict__ ddesc, const float* __restrict__ gr,
                        int nloc, float inv_nnei, int n1) {
  constexpr int kGroups = 64 / AXIS;
  constexpr int kS = AXIS + 1;
  __shared__ float s_gr[kAtoms][4 * kQ3];
  __shared__ float s_out[kAtoms][4 * kQ3];
  __shared__ __attribute__((aligned(16))) float s_ga[kAtoms][AXIS * 4];
  __shared__ float s_dd[kAtoms][kQ3 * kS];
  const int w = wave_id(), lane = threadIdx.x & 63;
  const int i = blockIdx.x * kAtoms + w;
  if (i >= nloc) return;
  const int M = 4 * n1;
  float* sg = s_gr[w];
  float* so = s_out[w];
  float* sga = s_ga[w];
  float* sdd = s_dd[w];
  const float* gri = gr + (int64_t)i * 4 * M;
  const float4* dd4 = reinterpret_cast<const float4*>(ddesc + (int64_t)i * M * AXIS);
  const int n4 = M * AXIS / 4;
#pragma unroll 4
  for (int k = lane; k < n4; k += 64) {
    const float4 x = dd4[k];
    const int m = (4 * k) / AXIS, a = (4 * k) % AXIS;
    sdd[m * kS + a + 0] = x.x;
    sdd[m * kS + a + 1] = x.y;
    sdd[m * kS + a + 2] = x.z;
    sdd[m * kS + a + 3] = x.w;
  }
  // GR rows staged at the fixed stride kQ3 (constant LDS offsets per c)
  for (int k = lane; k < 4 * kQ3; k += 64) so[k] = 0.f;
#pragma unroll
  for (int c = 0; c < 4; ++c)
    for (int m = lane; m < M; m += 64) sg[c * kQ3 + m] = gri[c * M + m];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < AXIS * 4) {
    const int a = lane >> 2, c = lane & 3;
    sga[4 * a + c] = sg[c * kQ3 + a];
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // first term: so[c][q3(m)] = sum_a dD[m][a] GR[c][a]
  for (int m = lane; m < M; m += 64) {
    float t[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < AXIS; ++a) {
      const float x = sdd[m * kS + a];
      const float4 g4 = *reinterpret_cast<const float4*>(sga + 4 * a);
      t[0] += x * g4.x;
      t[1] += x * g4.y;
      t[2] += x * g4.z;
      t[3] += x * g4.w;
    }
    const int qm = q3_of(m, n1);
#pragma unroll
    for (int c = 0; c < 4; ++c) so[c * kQ3 + qm] = t[c];
  }
  // second term: sum_m dD[m][a] GR[c][m] for a < axis
  const int a = lane % AXIS, m0 = lane / AXIS;
  float t2[4] = {0.f, 0.f, 0.f, 0.f};
  for (int m = m0; m < M; m += kGroups) {
    const float x = sdd[m * kS + a];
#pragma unroll
    for (int c = 0; c < 4; ++c) t2[c] += x * sg[c * kQ3 + m];
  }
#pragma unroll
  for (int off = AXIS; off < 64; off <<= 1)
#pragma unroll
    for (int c = 0; c < 4; ++c) t2[c] += __shfl_xor(t2[c], off, 64);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (m0 == 0) {
    const int qa = q3_of(a, n1);
#pragma unroll
    for (int c = 0; c < 4; ++c) so[c * kQ3 + qa] += t2[c];
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  float* out = dgr + (int64_t)i * 4 * kQ3;  // [q3][c]
  for (int k = lane; k < 4 * kQ3; k += 64) out[k] = so[(k & 3) * kQ3 + (k >> 2)] * inv_nnei;
}

// Backward: net_deriv[i][sec_t + row][c] = dE/d em, recomputing the forward
// over the same row stream as the forward, with 16-row tiles that hold rows
// of at most two atoms (slot = window index % 2).  The embedding net's input
// is the scalar s = em[row][0], so dE/ds is taken in forward mode: the tile
// carries the tangents dh1/ds, dh2/ds, dz3/ds beside the recompute (same
// weight fragments, same lane layout), and with dE/dG[row][q] =
// sum_c em[row][c] dGR[c][q] (G^T R = em^T G),
//   dE/ds = sum_c em[row][c] T[c][row],  T[c] = sum_q dGR[c][q] dG[q]/ds,
//   dE/d em[row][c] = sum_q G[q] dGR[c][q] (+ dE/ds for c = 0).
// Both contractions run on MFMA with A (M = 4 slot + c, K = feature) built
// from the atom's dGR when its first row comes up:
//   G  = idt3 (1 - 2 r3) + [h2, h2]:  K_c + (-2 idt3 dGR) r3 + (dGR[q2] + dGR[q2 + 64]) h2,
//   dG/ds = idt3 4 r3 (1 - r3) dz3/ds + [h2', h2']:  with z3' = k dz3/ds (the
//   fragments hold k W3) the same A (-2 idt3 dGR) against r3 (r3 - 1) z3' and
//   (k / 2) (dGR[q2] + dGR[q2 + 64]) against h2' give (k / 2) T,
// so the adjoint chain (dG = em dGR, dz3, dh2, dz2, dh1) is never formed.
// A workgroup loads its type's forward fragments into LDS once.
constexpr int kWavesBwd = 16;  // wavefronts per backward workgroup: four per SIMD
constexpr int kDopBwd = 8 * 16;  // A entries per slot: -2 idt3 dGR (4 chunks), dGR[q2] + dGR[q2 + 64] (2), times k / 2 (2)
constexpr float kK2 = 2.8853900817779268f;  // k = 2 log2(e)

template <typename H, bool kSkip>
__global__ void __launch_bounds__(64 * kWavesBwd) __attribute__((amdgpu_waves_per_eu(4, 4)))
    k_embed_bwd(float* __restrict__ net_deriv, const float* __restrict__ em, const float* __restrict__ dgr,
                const h8<H>* __restrict__ frag_all, const float* __restrict__ consts, int nloc, int per_wg,
                EmbedParams p) {
  constexpr int kS = 2;  // atoms per tile
  __shared__ h8<H> s_frag[kFragFwd * 64];
  __shared__ h8<H> s_dop[kWavesBwd][kS][kDopBwd];
  __shared__ __attribute__((aligned(16))) float s_kin[kWavesBwd][kS][4];
  __shared__ __attribute__((aligned(16))) float s_cst[kConsts];
  const int t = blockIdx.y;
  const int w = wave_id(), lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const h8<H>* frag_t = frag_all + (int64_t)t * kFragFwd * 64;
  for (int k = threadIdx.x; k < kFragFwd * 64; k += blockDim.x) s_frag[k] = frag_t[k];
  for (int k = threadIdx.x; k < kConsts; k += blockDim.x) s_cst[k] = consts[(int64_t)t * kConsts + k];
  const float* cst = s_cst;
  const int sel = p.sec[t + 1] - p.sec[t];
  const float* em_t = em + (int64_t)p.sec[t] * 4;
  float* nd_t = net_deriv + (int64_t)p.sec[t] * 4;
  // this workgroup's atoms, split among its wavefronts by rows
  __shared__ int s_wsum[kWavesBwd], s_wstart[kWavesBwd + 1];
  const int A0 = min((int)blockIdx.x * per_wg, nloc), nwg = min(per_wg, nloc - A0);
  split_rows<kWavesBwd>(p, t, sel, A0, nwg, w, lane, s_wsum, s_wstart);

  // fp16 operands: everything below is linear in dGR, so each atom's dGR is
  // staged times 2^-e (2^e the power of two at or below its largest
  // magnitude) and dE/d em multiplied by 2^e at the end, both exact; the A
  // operands then sit in fp16's normal range instead of underflowing
  float unscale[kS] = {1.f, 1.f};
  // Stage atom ai (listed rows cnt) into slot s: the padding slots after its
  // listed rows get dE/d em = 0 (no force term reads them); lane l holds
  // dGR[.][q3 = l] and dGR[.][q3 = l + 64] and writes the A entries of those
  // features (entry 16 chunk + 4 gg + c, element u: feature 32 chunk +
  // phi(gg, u)); K_c = sum_q3 idt3 dGR[c] goes to s_kin.  Rows of the other
  // slot (and of a slot not staged yet) only reach M rows that are not written.
  auto stage = [&](int ai, int s, int cnt) {
    if (p.counts != nullptr)
      for (int row = cnt + lane; row < sel; row += 64)
        *reinterpret_cast<float4*>(nd_t + ((int64_t)ai * p.nnei + row) * 4) = make_float4(0.f, 0.f, 0.f, 0.f);
    if (cnt == 0) return;
    // (the lane index passes through an opaque copy: hoisted address
    // arithmetic here made the allocator spill inside the loops)
    const float4* dg = reinterpret_cast<const float4*>(dgr + (int64_t)ai * 4 * kQ3);
    int l = lane;
    asm volatile("" : "+v"(l));
    float4 x0 = dg[l], x1 = dg[l + 64];
    if constexpr (std::is_same<H, _Float16>::value) {
      float mx = fmaxf(fmaxf(fmaxf(fabsf(x0.x), fabsf(x0.y)), fmaxf(fabsf(x0.z), fabsf(x0.w))),
                       fmaxf(fmaxf(fabsf(x1.x), fabsf(x1.y)), fmaxf(fabsf(x1.z), fabsf(x1.w))));
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
      // exponent field of the maximum (0 for zero and subnormals), kept in
      // [27, 227] so that both powers of two below are normal floats
      int ef = (int)((__float_as_uint(mx) >> 23) & 255u);
      ef = __builtin_amdgcn_readfirstlane(min(max(ef, 27), 227));
      const float sc = __uint_as_float((uint32_t)(254 - ef) << 23);  // 2^-e
      unscale[s] = __uint_as_float((uint32_t)ef << 23);              // 2^e
      x0 = make_float4(x0.x * sc, x0.y * sc, x0.z * sc, x0.w * sc);
      x1 = make_float4(x1.x * sc, x1.y * sc, x1.z * sc, x1.w * sc);
    }
    const float i0 = cst[352 + l], i1 = cst[352 + 64 + l];
    const float xa[4] = {x0.x, x0.y, x0.z, x0.w}, xb[4] = {x1.x, x1.y, x1.z, x1.w};
    float kp[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) kp[c] = fmaf(i0, xa[c], i1 * xb[c]);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
#pragma unroll
      for (int c = 0; c < 4; ++c) kp[c] += __shfl_xor(kp[c], off, 64);
    // the previous tiles' reads of this slot are complete
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    H* dop = reinterpret_cast<H*>(s_dop[w][s]);
    const int at = 32 * ((l >> 2) & 3) + 4 * ((l >> 4) & 1) + (l & 3) + 128 * (l >> 5);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float a2 = xa[c] + xb[c];
      dop[at + 8 * c] = (H)(-2.f * i0 * xa[c]);
      dop[at + 8 * c + 256] = (H)(-2.f * i1 * xb[c]);
      dop[at + 8 * c + 512] = (H)a2;
      dop[at + 8 * c + 768] = (H)(0.5f * kK2 * a2);
    }
    if (lane == 0) *reinterpret_cast<float4*>(s_kin[w][s]) = make_float4(kp[0], kp[1], kp[2], kp[3]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  const int a_lo = A0 + s_wstart[w], a_hi = A0 + s_wstart[w + 1];
  for (int wa = a_lo; wa < a_hi; wa += 64) {
    const int n = min(64, a_hi - wa);
    int cnt = 0;
    if (lane < n) cnt = p.counts != nullptr ? (int)p.counts[(int64_t)(wa + lane) * p.ntypes + t] : sel;
    int pre = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(pre, off, 64);
      if (lane >= off) pre += y;
    }
    // em row of this lane in the tile of chunk c (read one tile ahead)
    auto load_row = [&](const Chunk& c, int k0, int p0) {
      const int pos = p0 + r;
      int row;
      const int k = chunk_atom(c, k0, pos, &row);
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (pos < c.p1) v = *reinterpret_cast<const float4*>(em_t + ((int64_t)(wa + k) * p.nnei + row) * 4);
      return v;
    };
    int k0 = 0, p0 = 0, staged = 0;
    Chunk c = make_chunk2(pre, 0, 0, n, 16);
    float4 ev_next = load_row(c, 0, 0);
    while (k0 < n) {
      const int k0n = k0 + c.nfin, p0n = c.p1;
      // stage the atoms whose rows start in this tile (and the atoms without
      // rows before them)
      const int last = max(c.p1 > c.e0 ? k0 + 1 : k0, k0 + c.nfin - 1);
      for (; staged <= last; ++staged) stage(wa + staged, staged & 1, __builtin_amdgcn_readlane(cnt, staged));
      const float4 ev = ev_next;
      Chunk cn = c;
      if (k0n < n) {
        cn = make_chunk2(pre, k0n, p0n, n, 16);
        ev_next = load_row(cn, k0n, p0n);
      }
      if (c.p1 > p0) {
        // this lane's row: atom k (slot sl, -1 past the tile's rows), row index
        const int pos = p0 + r;
        int row;
        const int k = chunk_atom(c, k0, pos, &row);
        const int sl = pos < c.p1 ? (k & 1) : -1;
        // weights and constants are re-read from LDS every tile instead of
        // being held in registers across the loop
        asm volatile("" ::: "memory");
        // layer 1 and its tangent: h1' = idt1 w1 (1 - t1^2)
        h8<H> b1, b1t;
        float h1[8], h1t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (kSkip && u == 7) {  // empty slot in every lane group
            h1[u] = h1t[u] = 0.f;
            b1[u] = b1t[u] = (H)0.f;
            continue;
          }
          const int q = phi(g, u);
          const float tt = tanh_scaled(ev.x * cst[q] + cst[32 + q]);
          h1[u] = tt * cst[64 + q];
          h1t[u] = fmaf(-tt, tt, 1.f) * cst[480 + q];
          b1[u] = (H)h1[u];
          b1t[u] = (H)h1t[u];
        }
        // layer 2: z2 (k scaled, bias in the accumulator) and z2' = k W2^T h1';
        // h2' = (idt2 / k) (1 - t2^2) z2' + h1'
        h8<H> b2[2], b2t[2];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          const h8<H> fr = s_frag[mt * 64 + lane];
          const f32x4 z = mfma(fr, b1, *reinterpret_cast<const f32x4*>(cst + 96 + 16 * mt + 4 * g));
          const f32x4 zt = mfma(fr, b1t, f32x4{0.f, 0.f, 0.f, 0.f});
          const int q = 16 * mt + 4 * g;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int u = (mt & 1) * 4 + v;
            if (kSkip && (mt & 1) && v == 3) {
              b2[mt >> 1][u] = b2t[mt >> 1][u] = (H)0.f;
              continue;
            }
            const float tt = tanh_scaled(z[v]);
            b2[mt >> 1][u] = (H)(tt * cst[160 + q + v] + h1[u]);
            b2t[mt >> 1][u] = (H)fmaf(fmaf(-tt, tt, 1.f) * cst[512 + q + v], zt[v], h1t[u]);
          }
        }
        // layer 3 by 32-feature chunks, each contracted at once: dem += A r3,
        // tg += A r3 (r3 - 1) z3' (M = 4 slot + c, N = rows; lane groups 2, 3
        // hold rows that are not written)
        const h8<H>* dop = s_dop[w][(r >> 2) & 1] + 4 * g + (r & 3);
        f32x4 dem = f32x4{0.f, 0.f, 0.f, 0.f}, tg = f32x4{0.f, 0.f, 0.f, 0.f};
        if (g < kS) dem = *reinterpret_cast<const f32x4*>(s_kin[w][g]);
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
          // fragments read per chunk, not all ahead of the products
          asm volatile("" ::: "memory");
          h8<H> br, bt;
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const int mt = 2 * kc + j;
            const h8<H> f0 = s_frag[(4 + 2 * mt) * 64 + lane], f1 = s_frag[(5 + 2 * mt) * 64 + lane];
            f32x4 z = mfma(f0, b2[0], *reinterpret_cast<const f32x4*>(cst + 224 + 16 * mt + 4 * g));
            z = mfma(f1, b2[1], z);
            f32x4 zt = mfma(f0, b2t[0], f32x4{0.f, 0.f, 0.f, 0.f});
            zt = mfma(f1, b2t[1], zt);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              if (kSkip && j == 1 && v == 3) {
                br[4 * j + v] = bt[4 * j + v] = (H)0.f;
                continue;
              }
              const float r3 = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z[v]));
              br[4 * j + v] = (H)r3;
              bt[4 * j + v] = (H)(fmaf(r3, r3, -r3) * zt[v]);
            }
          }
          const h8<H> a = dop[16 * kc];
          dem = mfma(a, br, dem);
          tg = mfma(a, bt, tg);
        }
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
          dem = mfma(dop[16 * (4 + kc)], b2[kc], dem);
          tg = mfma(dop[16 * (6 + kc)], b2t[kc], tg);
        }
        if (g == sl) {
          const float us = sl == 0 ? unscale[0] : unscale[1];
          const float ds = (2.f / kK2) * (ev.x * tg[0] + ev.y * tg[1] + ev.z * tg[2] + ev.w * tg[3]);
          *reinterpret_cast<float4*>(nd_t + ((int64_t)(wa + k) * p.nnei + row) * 4) =
              make_float4((dem[0] + ds) * us, dem[1] * us, dem[2] * us, dem[3] * us);
        }
      }
      k0 = k0n;
      p0 = p0n;
      c = cn;
    }
  }
}

// counts[i * ntypes + t] = leading entries >= 0 of section t of row i (the
// slots env_mat_a_cpu treats as neighbors: it stops at the first -1).
__global__ void __launch_bounds__(256)
    k_section_counts(int16_t* __restrict__ counts, const int* __restrict__ nlist, int nloc, EmbedParams p) {
  const int w = wave_id(), lane = threadIdx.x & 63;
  const int i = blockIdx.x * kAtoms + w;
  if (i >= nloc) return;
  const int* nl = nlist + (int64_t)i * p.nnei;
  for (int t = 0; t < p.ntypes; ++t) {
    int first = p.sec[t + 1];
    for (int s = p.sec[t] + lane; s < p.sec[t + 1]; s += 64)
      if (nl[s] < 0) first = min(first, s);
    for (int off = 32; off > 0; off >>= 1) first = min(first, __shfl_xor(first, off, 64));
    if (lane == 0) counts[(int64_t)i * p.ntypes + t] = (int16_t)(first - p.sec[t]);
  }
}

int make_params(EmbedParams* p, const int* sel_host, int ntypes) {
  DPMI_CHECK_ARG(sel_host != nullptr, "sel is NULL");
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= kMaxTypes, "ntypes out of range [1, 16]");
  std::memset(p, 0, sizeof(*p));
  p->ntypes = ntypes;
  for (int t = 0; t < ntypes; ++t) {
    DPMI_CHECK_ARG(sel_host[t] >= 0, "sel must be >= 0");
    p->sec[t + 1] = p->sec[t] + sel_host[t];
  }
  p->nnei = p->sec[ntypes];
  return DPMI_OK;
}

template <typename H>
int embed_fwd(float* gr_part, const float* em, const void* frags, const float* consts, int64_t nloc,
              const int* sel_host, int ntypes, int n1, const int16_t* counts, const int* atype, const float* pad_em,
              void* stream) {
  EmbedParams p;
  int rc = make_params(&p, sel_host, ntypes);
  if (rc) return rc;
  DPMI_CHECK_ARG(counts == nullptr || (atype != nullptr && pad_em != nullptr),
                 "the padding shortcut needs atype and pad_em with counts");
  p.counts = counts;
  p.atype = atype;
  p.pad_em = pad_em;
  DPMI_CHECK_ARG(n1 > 0 && n1 <= 32, "n1 out of range [1, 32]");
  p.n1 = n1;
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  if (nloc == 0) return DPMI_OK;
  // up to 1,024 workgroups per type (256 CUs four times over); each
  // wavefront streams a contiguous range of per_wave atoms (four workgroups
  // per CU: one whose wavefronts end early is soon followed by another)
  dim3 grid((unsigned)std::min<int64_t>(ceil_div(nloc, kAtoms), 1024), (unsigned)ntypes);
  const int per_wave = (int)ceil_div(nloc, (int64_t)grid.x * kAtoms);
  // G of the padding row per center type (dynamic LDS, ntypes x 128 floats)
  const size_t shm_gpad = (size_t)ntypes * kQ3 * sizeof(float);
  if (skip_layout(n1))
    hipLaunchKernelGGL((k_embed_fwd<H, true>), grid, dim3(256), shm_gpad, as_stream(stream), gr_part, em,
                       static_cast<const h8<H>*>(frags), consts, (int)nloc, per_wave, p);
  else
    hipLaunchKernelGGL((k_embed_fwd<H, false>), grid, dim3(256), shm_gpad, as_stream(stream), gr_part, em,
                       static_cast<const h8<H>*>(frags), consts, (int)nloc, per_wave, p);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

template <typename H>
int embed_bwd(float* net_deriv, const float* em, const float* dgr, const void* frags, const float* consts,
              int64_t nloc, const int* sel_host, int ntypes, int n1, const int16_t* counts, void* stream) {
  EmbedParams p;
  int rc = make_params(&p, sel_host, ntypes);
  if (rc) return rc;
  p.counts = counts;
  DPMI_CHECK_ARG(n1 > 0 && n1 <= 32, "n1 out of range [1, 32]");
  p.n1 = n1;
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  if (nloc == 0) return DPMI_OK;
  // one workgroup of sixteen wavefronts per CU (87 KB of LDS); a workgroup
  // takes per_wg <= 1,024 consecutive atoms (about 256 workgroups per type)
  // and splits them among its wavefronts by rows
  const int per_wg = (int)std::min<int64_t>(64 * kWavesBwd, std::max<int64_t>(kWavesBwd, ceil_div(nloc, 256)));
  dim3 grid((unsigned)ceil_div(nloc, (int64_t)per_wg), (unsigned)ntypes);
  if (skip_layout(n1))
    hipLaunchKernelGGL((k_embed_bwd<H, true>), grid, dim3(64 * kWavesBwd), 0, as_stream(stream), net_deriv, em,
                       dgr, static_cast<const h8<H>*>(frags), consts, (int)nloc, per_wg, p);
  else
    hipLaunchKernelGGL((k_embed_bwd<H, false>), grid, dim3(64 * kWavesBwd), 0, as_stream(stream), net_deriv, em,
                       dgr, static_cast<const h8<H>*>(frags), consts, (int)nloc, per_wg, p);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_se_a_embed_fwd_bf16(float* gr_part, const float* em, const void* frags, const float* consts,
                                        int64_t nloc, const int* sel_host, int ntypes, int n1, const int16_t* counts,
                                        const int* atype, const float* pad_em, void* stream) {
  return embed_fwd<__bf16>(gr_part, em, frags, consts, nloc, sel_host, ntypes, n1, counts, atype, pad_em, stream);
}

extern "C" int dpmi_se_a_embed_fwd_f16(float* gr_part, const float* em, const void* frags, const float* consts,
                                       int64_t nloc, const int* sel_host, int ntypes, int n1, const int16_t* counts,
                                       const int* atype, const float* pad_em, void* stream) {
  return embed_fwd<_Float16>(gr_part, em, frags, consts, nloc, sel_host, ntypes, n1, counts, atype, pad_em, stream);
}

extern "C" int dpmi_se_a_section_counts(int16_t* counts, const int* nlist, int64_t nloc, const int* sel_host,
                                        int ntypes, void* stream) {
  EmbedParams p;
  int rc = make_params(&p, sel_host, ntypes);
  if (rc) return rc;
  DPMI_CHECK_ARG(p.nnei < 32768, "sum(sel) must be < 32768");
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  if (nloc == 0) return DPMI_OK;
  hipLaunchKernelGGL(k_section_counts, dim3((unsigned)ceil_div(nloc, kAtoms)), dim3(256), 0, as_stream(stream), counts,
                     nlist, (int)nloc, p);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_se_a_descriptor_fwd(float* gr, void* desc, const float* gr_part, int64_t nloc, int ntypes,
                                        int nnei, int n1, int axis, int desc_kind, void* stream) {
  DPMI_CHECK_ARG(n1 > 0 && n1 <= 32 && axis > 0 && axis <= 4 * n1 && axis <= 16, "unsupported widths");
  DPMI_CHECK_ARG(desc_kind >= 0 && desc_kind <= 2, "desc_kind must be 0 (fp32), 1 (bf16) or 2 (fp16)");
  if (nloc == 0) return DPMI_OK;
  if (desc_kind == 1)
    hipLaunchKernelGGL(k_descrpt_fwd<__bf16>, dim3(ceil_div(nloc, kAtoms)), dim3(256), 0, as_stream(stream), gr,
                       static_cast<__bf16*>(desc), gr_part, (int)nloc, ntypes, 1.f / (float)nnei, n1, axis);
  else if (desc_kind == 2)
    hipLaunchKernelGGL(k_descrpt_fwd<_Float16>, dim3(ceil_div(nloc, kAtoms)), dim3(256), 0, as_stream(stream), gr,
                       static_cast<_Float16*>(desc), gr_part, (int)nloc, ntypes, 1.f / (float)nnei, n1, axis);
  else
    hipLaunchKernelGGL(k_descrpt_fwd<float>, dim3(ceil_div(nloc, kAtoms)), dim3(256), 0, as_stream(stream), gr,
                       static_cast<float*>(desc), gr_part, (int)nloc, ntypes, 1.f / (float)nnei, n1, axis);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_se_a_descriptor_bwd(float* dgr, const float* ddesc, const float* gr, int64_t nloc, int nnei,
                                        int n1, int axis, void* stream) {
  DPMI_CHECK_ARG(n1 > 0 && n1 <= 32 && axis > 0 && axis <= 4 * n1 && axis <= 16, "unsupported widths");
  if (nloc == 0) return DPMI_OK;
  const dim3 grid((unsigned)ceil_div(nloc, kAtoms));
  const float inv = 1.f / (float)nnei;
  if (axis == 16)
    hipLaunchKernelGGL(k_descrpt_bwd_lanes<16>, grid, dim3(256), 0, as_stream(stream), dgr, ddesc, gr, (int)nloc, inv, n1);
  else if (axis == 8)
    hipLaunchKernelGGL(k_descrpt_bwd_lanes<8>, grid, dim3(256), 0, as_stream(stream), dgr, ddesc, gr, (int)nloc, inv, n1);
  else if (axis == 4)
    hipLaunchKernelGGL(k_descrpt_bwd_lanes<4>, grid, dim3(256), 0, as_stream(stream), dgr, ddesc, gr, (int)nloc, inv, n1);
  else
    hipLaunchKernelGGL(k_descrpt_bwd, grid, dim3(256), 0, as_stream(stream), dgr, ddesc, gr, (int)nloc, inv, n1, axis);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_se_a_embed_bwd_bf16(float* net_deriv, const float* em, const float* dgr, const void* frags,
                                        const float* consts, int64_t nloc, const int* sel_host, int ntypes, int n1,
                                        const int16_t* counts, void* stream) {
  return embed_bwd<__bf16>(net_deriv, em, dgr, frags, consts, nloc, sel_host, ntypes, n1, counts, stream);
}

extern "C" int dpmi_se_a_embed_bwd_f16(float* net_deriv, const float* em, const float* dgr, const void* frags,
                                       const float* consts, int64_t nloc, const int* sel_host, int ntypes, int n1,
                                       const int16_t* counts, void* stream) {
  return embed_bwd<_Float16>(net_deriv, em, dgr, frags, consts, nloc, sel_host, ntypes, n1, counts, stream);
}
