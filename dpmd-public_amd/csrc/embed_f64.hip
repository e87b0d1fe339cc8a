// Fused se_e2_a embedding net + G^T R in FP64 on gfx950 (v_mfma_f64_16x16x4f64),
// and its recompute-based adjoint: the plain FP64 form of the same kernels as
// embed_bf16.hip.
//
// Reference semantics: DescrptBlockSeA.forward (deepmd/pt/model/descriptor/
// se_a.py:802-858) with MLPLayer.forward (deepmd/pt/model/network/mlp.py:
// 188-219), type_one_side:
//   h1 = tanh(s w1 + b1) * idt1,  h2 = tanh(h1 W2 + b2) * idt2 + [h1, h1],
//   G  = tanh(h2 W3 + b3) * idt3 + [h2, h2],   GR_t = em_t^T G_t.
//
// MFMA f64 16x16x4 lane layout on gfx950 (verified on the device by a probe):
// A lane l -> (M row l%16, K l/16), B lane l -> (K l/16, N col l%16),
// D lane l, slot v -> (M row 4v + l/16, N col l%16).  Every layer is computed
// transposed, y^T = W^T x^T (M = output feature, N = neighbor row, K = input
// feature), so a lane always holds one neighbor row (r = l%16) and, in D slot
// v of feature tile mt, feature 16 mt + 4 v + g (g = l/16).  That value is
// exactly the B operand of K-chunk (mt, v) of the next layer: the layers chain
// in registers with the weights read as plain matrices from LDS.  Internal
// feature order: q1 = f (f < n1), q2 = q1(f mod n1) + 32 [f >= n1],
// q3 = q2(f mod 2 n1) + 64 [f >= 2 n1]; the [x, x] skip halves share lane and
// slot.  K-chunks made only of padding (4 consecutive slots >= n1) are skipped.
// GR = sum_rows em^T G contracts over rows, so each G tile is transposed by
// four MFMAs against 0/1 selector fragments (exact) and reduced with em^T.
// tanh is evaluated as E / (E + 2) with E = expm1(2|x|) from a Cody-Waite
// reduction and a degree-13 Taylor polynomial (relative error ~1e-16).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "dpmi_internal.h"
#include "tanh64.h"

namespace dpmi {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxTypes64 = 16;
constexpr int kP2 = 65;     // LDS row pitch (doubles) of W2 [32][64]
constexpr int kP3 = 129;    // LDS row pitch (doubles) of W3 [64][128]
constexpr int kC64 = 480;   // w1[32] b1[32] idt1[32] b2[64] idt2[64] b3[128] idt3[128]
constexpr int kWaves64 = 8;  // wavefronts per workgroup: two per SIMD, so FP64 MFMA and VALU of different waves overlap
constexpr size_t kLdsW = (size_t)(32 * kP2 + 64 * kP3 + kC64) * sizeof(double);
// forward: G of the padding row of each center type after the weights
constexpr size_t kLdsFwd = kLdsW + (size_t)kMaxTypes64 * 128 * sizeof(double);

struct Embed64Params {
  int ntypes;
  int nnei;
  int n1;
  int nblk_total;
  int sec[kMaxTypes64 + 1];
  int blk[kMaxTypes64 + 1];  // workgroup ranges per neighbor type
  // optional padding shortcut (as in the bf16 kernels): counts[i * ntypes + t]
  // listed neighbors, atype[i], pad_em[(type * ntypes + t) * 4 + c]
  const int16_t* counts;
  const int* atype;
  const double* pad_em;
};

__device__ inline f64x4 mfma64(double a, double b, const f64x4& c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ inline int orig_of_q3_64(int q, int n1) {
  const int base = q >= 64 ? 2 * n1 : 0, q2 = q & 63;
  const int f2 = q2 >= 32 ? (q2 - 32 < n1 ? n1 + (q2 - 32) : -1) : (q2 < n1 ? q2 : -1);
  return f2 < 0 ? -1 : base + f2;
}

// Chunks are skipped at compile time: NL = n1 rounded up to a multiple of 4
// (24 for n1 <= 24, else 28 or 32); a 4-slot chunk starting at layer-1 slot
// s1 holds a feature iff s1 < NL (the rest of a live chunk is zero padding).
template <int NL>
__device__ constexpr bool live_chunk(int s1) { return s1 < NL; }

// Weights of type t into LDS: W2 [q1][q2] (pitch kP2), W3 [q2][q3] (pitch
// kP3), constants.
__device__ inline void load_weights(double* sw, const double* __restrict__ w2, const double* __restrict__ w3,
                                    const double* __restrict__ consts, int t) {
  double* sW2 = sw;
  double* sW3 = sW2 + 32 * kP2;
  double* sC = sW3 + 64 * kP3;
  const double* w2t = w2 + (int64_t)t * 32 * 64;
  const double* w3t = w3 + (int64_t)t * 64 * 128;
  for (int k = threadIdx.x; k < 32 * 64; k += blockDim.x) sW2[(k >> 6) * kP2 + (k & 63)] = w2t[k];
  for (int k = threadIdx.x; k < 64 * 128; k += blockDim.x) sW3[(k >> 7) * kP3 + (k & 127)] = w3t[k];
  for (int k = threadIdx.x; k < kC64; k += blockDim.x) sC[k] = consts[(int64_t)t * kC64 + k];
}

__device__ inline int type_of_block(const Embed64Params& p, int b) {
  int t = 0;
  while (t + 1 < p.ntypes && b >= p.blk[t + 1]) ++t;
  return t;
}

// Layers 1 and 2 of one 16-row tile: h1 (8 slots: chunk (mt1, v) = 4 mt1 + v),
// h2 (16 slots: chunk (mt2, v) = 4 mt2 + v), tanh values t1, t2 (for the
// adjoint; may be null).
template <int NL, bool kKeep>
__device__ inline void layers12(double s, const double* __restrict__ sW2, const double* __restrict__ sC, int g,
                                int r, double (&h1)[8], double (&h2)[16], double (&t1)[8], double (&t2)[16]) {
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int s1 = 16 * (c >> 2) + 4 * (c & 3);
    if (!live_chunk<NL>(s1)) {
      h1[c] = 0.;
      if (kKeep) t1[c] = 0.;
      continue;
    }
    const int q = s1 + g;
    const double t = tanh64(fma(s, sC[q], sC[32 + q]));
    if (kKeep) t1[c] = t;
    h1[c] = t * sC[64 + q];
  }
  auto z2 = [&](int mt) {
    asm volatile("" ::: "memory");
    f64x4 d;
#pragma unroll
    for (int v = 0; v < 4; ++v) d[v] = sC[96 + 16 * mt + 4 * v + g];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int s1 = 16 * (c >> 2) + 4 * (c & 3);
      if (!live_chunk<NL>(s1)) continue;
      d = mfma64(sW2[(s1 + g) * kP2 + 16 * mt + r], h1[c], d);  // A: W2[q1 = s1 + k][q2 = 16 mt + m]
    }
    return d;
  };
  f64x4 d_next = z2(0);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const f64x4 d = d_next;
    if (mt < 3) d_next = z2(mt + 1);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = 4 * mt + v, s1 = 16 * (mt & 1) + 4 * v;
      if (!live_chunk<NL>(s1)) {
        h2[c] = 0.;
        if (kKeep) t2[c] = 0.;
        continue;
      }
      const double t = tanh64(d[v]);
      if (kKeep) t2[c] = t;
      h2[c] = t * sC[160 + 16 * mt + 4 * v + g] + h1[4 * (mt & 1) + v];
    }
  }
}

// Layer-3 feature tile mt: the MFMA chain (z3 of the tile) ...
template <int NL>
__device__ inline f64x4 layer3_z(int mt, const double (&h2)[16], const double* __restrict__ sW3,
                                 const double* __restrict__ sC, int g, int r) {
  // keep the weight reads of this tile here: hoisting all tiles' LDS reads
  // ahead would need hundreds of registers
  asm volatile("" ::: "memory");
  f64x4 d;
#pragma unroll
  for (int v = 0; v < 4; ++v) d[v] = sC[224 + 16 * mt + 4 * v + g];
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const int s1 = 16 * ((c >> 2) & 1) + 4 * (c & 3);
    if (!live_chunk<NL>(s1)) continue;
    d = mfma64(sW3[(16 * (c >> 2) + 4 * (c & 3) + g) * kP3 + 16 * mt + r], h2[c], d);
  }
  return d;
}

// ... and its activation: G (4 slots) and the tanh values.  The loops issue
// the next tile's MFMA chain before this, so the matrix core and the VALU
// work concurrently.
template <int NL>
__device__ inline void layer3_act(int mt, const f64x4& d, const double (&h2)[16], const double* __restrict__ sC,
                                  int g, double (&G)[4], double (&t3)[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int s1 = 16 * (mt & 1) + 4 * v;
    if (!live_chunk<NL>(s1)) {
      G[v] = 0.;
      t3[v] = 0.;
      continue;
    }
    const double t = tanh64(d[v]);
    t3[v] = t;
    G[v] = t * sC[352 + 16 * mt + 4 * v + g] + h2[4 * (mt & 3) + v];
  }
}

// Forward: gr_part[t][i][c][m] = sum over type-t neighbors of em[c] G[m]
// (original feature order).  Workgroups are split over neighbor types in
// proportion to sel; each wavefront walks atoms, 16 rows per tile.
template <int NL>
__global__ void __launch_bounds__(64 * kWaves64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_embed64_fwd(double* __restrict__ gr_part, const double* __restrict__ em, const double* __restrict__ w2,
                  const double* __restrict__ w3, const double* __restrict__ consts, int nloc, Embed64Params p) {
  extern __shared__ double smem64[];
  const int t = type_of_block(p, blockIdx.x);
  load_weights(smem64, w2, w3, consts, t);
  __syncthreads();
  const double* sW2 = smem64;
  const double* sW3 = sW2 + 32 * kP2;
  const double* sC = sW3 + 64 * kP3;
  const int w = wave_id(), lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const int n1 = p.n1, M = 4 * n1;
  const int sel = p.sec[t + 1] - p.sec[t], ntile = (sel + 15) / 16;
  const int nb = p.blk[t + 1] - p.blk[t], b0 = blockIdx.x - p.blk[t];
  // padding shortcut: G of each center type's padding row, once per workgroup
  double* s_gpad = const_cast<double*>(sC) + kC64;  // [ntypes][128]
  if (p.counts != nullptr) {
    for (int ti = w; ti < p.ntypes; ti += kWaves64) {
      double h1[8], h2[16], t1[8], t2[16];
      layers12<NL, false>(p.pad_em[(ti * p.ntypes + t) * 4], sW2, sC, g, r, h1, h2, t1, t2);
#pragma unroll
      for (int mt = 0; mt < 8; ++mt) {
        double G[4], t3[4];
        layer3_act<NL>(mt, layer3_z<NL>(mt, h2, sW3, sC, g, r), h2, sC, g, G, t3);
        if (r == 0)
#pragma unroll
          for (int v = 0; v < 4; ++v) s_gpad[ti * 128 + 16 * mt + 4 * v + g] = G[v];
      }
    }
    __syncthreads();
  }
  for (int i = b0 * kWaves64 + w; i < nloc; i += nb * kWaves64) {
    const double* emi = em + ((int64_t)i * p.nnei + p.sec[t]) * 4;
    // tiles holding listed neighbors (all of the section without counts)
    const int nti = p.counts != nullptr ? (p.counts[(int64_t)i * p.ntypes + t] + 15) >> 4 : ntile;  // wave-uniform
    // GR[c = g][16 mt + r]: slot 0 of each tile's GR^T product (the other
    // slots are zero rows c >= 4), accumulated here so that no full
    // accumulator stays live across tiles
    double acc[8];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) acc[mt] = 0.;
    for (int tl = 0; tl < nti; ++tl) {
      const int row = 16 * tl + r;
      const double s = row < sel ? emi[4 * row] : 0.;
      // em^T as A of GR^T: (M = c = r < 4, K = row 16 tl + 4 kc + g)
      double a_em[4];
#pragma unroll
      for (int kc = 0; kc < 4; ++kc) {
        const int rr = 16 * tl + 4 * kc + g;
        a_em[kc] = (r < 4 && rr < sel) ? emi[4 * rr + (r & 3)] : 0.;
      }
      double h1[8], h2[16], t1[8], t2[16];
      layers12<NL, false>(s, sW2, sC, g, r, h1, h2, t1, t2);
      f64x4 z_next = layer3_z<NL>(0, h2, sW3, sC, g, r);
#pragma unroll
      for (int mt = 0; mt < 8; ++mt) {
        const f64x4 z = z_next;
        if (mt < 7) z_next = layer3_z<NL>(mt + 1, h2, sW3, sC, g, r);
        double G[4], t3[4];
        layer3_act<NL>(mt, z, h2, sC, g, G, t3);
        // transpose: T[m = row][n] = G[row][16 mt + n]; T slot v' holds rows 4 v' + g, column r
        f64x4 T = f64x4{0., 0., 0., 0.};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (!live_chunk<NL>(16 * (mt & 1) + 4 * v)) continue;
          T = mfma64(G[v], r == 4 * v + g ? 1. : 0., T);
        }
        // GR^T[c][16 mt + n] += sum_rows em[row][c] G[row][16 mt + n]
        {
          f64x4 grt = f64x4{0., 0., 0., 0.};
#pragma unroll
          for (int kc = 0; kc < 4; ++kc) grt = mfma64(a_em[kc], T[kc], grt);
          acc[mt] += grt[0];
        }
      }
    }
    // padding slots not evaluated: n_pad em_pad[c] G_pad
    const int n_pad = sel - min(sel, 16 * nti);
    if (n_pad > 0) {
      const int ti = p.atype[i];
      if (ti >= 0 && ti < p.ntypes) {
        const double e = (double)n_pad * p.pad_em[(ti * p.ntypes + t) * 4 + g];
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) acc[mt] += e * s_gpad[ti * 128 + 16 * mt + r];
      }
    }
    // acc[mt] slot 0 of lane group g = c holds GR[c][q3 = 16 mt + r]
    double* out = gr_part + ((int64_t)t * nloc + i) * 4 * M;
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
      const int m = orig_of_q3_64(16 * mt + r, n1);
      if (m >= 0) out[g * M + m] = acc[mt];
    }
  }
}

// Backward: net_deriv[i][sec_t + row][c] = dE/d em from dgr = dE/d GR_t
// ([nloc, 4, M], original order; the same for every type), recomputing the
// forward per tile.  dG = em dGR, direct term dem = G . dGR, then the
// adjoint chain dh2 = W3 dz3 + skip, dh1 = W2 dz2 + skip, ds = dz1 . w1.
template <int NL>
__global__ void __launch_bounds__(64 * kWaves64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_embed64_bwd(double* __restrict__ net_deriv, const double* __restrict__ em, const double* __restrict__ dgr,
                  const double* __restrict__ w2, const double* __restrict__ w3, const double* __restrict__ consts,
                  int nloc, Embed64Params p) {
  extern __shared__ double smem64[];
  const int t = type_of_block(p, blockIdx.x);
  load_weights(smem64, w2, w3, consts, t);
  __syncthreads();
  const double* sW2 = smem64;
  const double* sW3 = sW2 + 32 * kP2;
  const double* sC = sW3 + 64 * kP3;
  const int w = wave_id(), lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  double* sdg = const_cast<double*>(sC) + kC64 + w * 128 * 4;  // [q3][c]
  const int n1 = p.n1, M = 4 * n1;
  const int sel = p.sec[t + 1] - p.sec[t], ntile = (sel + 15) / 16;
  const int nb = p.blk[t + 1] - p.blk[t], b0 = blockIdx.x - p.blk[t];
  for (int i = b0 * kWaves64 + w; i < nloc; i += nb * kWaves64) {
    const double* emi = em + ((int64_t)i * p.nnei + p.sec[t]) * 4;
    for (int k = lane; k < 4 * 128; k += 64) {
      const int q = k >> 2, c = k & 3, m = orig_of_q3_64(q, n1);
      sdg[k] = m >= 0 ? dgr[((int64_t)i * 4 + c) * M + m] : 0.;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double* ndi = net_deriv + ((int64_t)i * p.nnei + p.sec[t]) * 4;
    // padding shortcut: tiles past the listed neighbors are not evaluated and
    // every padding slot gets dE/d em = 0 (no force term reads it)
    const int nv = p.counts != nullptr ? (int)p.counts[(int64_t)i * p.ntypes + t] : sel;
    const int nti = (nv + 15) >> 4;
    for (int row = 16 * nti + r; row < sel; row += 16) ndi[4 * row + g] = 0.;
    for (int tl = 0; tl < nti; ++tl) {
      const int row = 16 * tl + r;
      const bool in = row < sel;
      const double s = in ? emi[4 * row] : 0.;
      const double em_g = in ? emi[4 * row + g] : 0.;  // B of dG: (K = c = g, N = row)
      double h1[8], h2[16], t1[8], t2[16];
      layers12<NL, true>(s, sW2, sC, g, r, h1, h2, t1, t2);
      f64x4 dem = f64x4{0., 0., 0., 0.};
      f64x4 dh2[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) dh2[mt] = f64x4{0., 0., 0., 0.};
      f64x4 z_next = layer3_z<NL>(0, h2, sW3, sC, g, r);
#pragma unroll
      for (int mt = 0; mt < 8; ++mt) {
        const f64x4 z = z_next;
        if (mt < 7) z_next = layer3_z<NL>(mt + 1, h2, sW3, sC, g, r);
        double G[4], t3[4];
        layer3_act<NL>(mt, z, h2, sC, g, G, t3);
        // direct term dem[c][row] += sum_q dGR[c][q] G[row][q]: A (M = c = r < 4, K = q3 chunk)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (!live_chunk<NL>(16 * (mt & 1) + 4 * v)) continue;
          const double a = r < 4 ? sdg[(16 * mt + 4 * v + g) * 4 + (r & 3)] : 0.;
          dem = mfma64(a, G[v], dem);
        }
        // dG[row][q] = sum_c em[row][c] dGR[c][q]: A (M = q = 16 mt + r, K = c = g)
        const f64x4 dG = mfma64(sdg[(16 * mt + r) * 4 + g], em_g, f64x4{0., 0., 0., 0.});
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (!live_chunk<NL>(16 * (mt & 1) + 4 * v)) continue;
          dh2[mt & 3][v] += dG[v];  // skip: G = ... + [h2, h2] carries dG to dh2[q3 mod 64]
          const double dz3 = dG[v] * sC[352 + 16 * mt + 4 * v + g] * fma(-t3[v], t3[v], 1.);
          asm volatile("" ::: "memory");
          // dh2[q2] += sum_q3 W3[q2][q3] dz3[q3]: A (M = q2 = 16 mt2 + r, K = q3 = 16 mt + 4 v + g)
#pragma unroll
          for (int mt2 = 0; mt2 < 4; ++mt2) {
            if (!live_chunk<NL>(16 * (mt2 & 1))) continue;
            dh2[mt2] = mfma64(sW3[(16 * mt2 + r) * kP3 + 16 * mt + 4 * v + g], dz3, dh2[mt2]);
          }
        }
      }
      // dh2 += skip; dz2 = dh2 idt2 (1 - t2^2); dh1[q1] = sum_q2 W2[q1][q2] dz2[q2]
      f64x4 dh1[2] = {f64x4{0., 0., 0., 0.}, f64x4{0., 0., 0., 0.}};
      double dh2v[16];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = 4 * mt + v;
          dh2v[c] = dh2[mt][v];
          if (!live_chunk<NL>(16 * (mt & 1) + 4 * v)) continue;
          const double dz2 = dh2v[c] * sC[160 + 16 * mt + 4 * v + g] * fma(-t2[c], t2[c], 1.);
          asm volatile("" ::: "memory");
#pragma unroll
          for (int mt1 = 0; mt1 < 2; ++mt1) {
            if (!live_chunk<NL>(16 * mt1)) continue;
            dh1[mt1] = mfma64(sW2[(16 * mt1 + r) * kP2 + 16 * mt + 4 * v + g], dz2, dh1[mt1]);
          }
        }
      // dh1 += dh2[q1] + dh2[q1 + 32];  ds = sum_q1 dh1 idt1 (1 - t1^2) w1
      double ds = 0.;
#pragma unroll
      for (int mt1 = 0; mt1 < 2; ++mt1)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = 4 * mt1 + v, s1 = 16 * mt1 + 4 * v;
          if (!live_chunk<NL>(s1)) continue;
          const double x = dh1[mt1][v] + dh2v[c] + dh2v[c + 8];
          ds += x * sC[64 + s1 + g] * fma(-t1[c], t1[c], 1.) * sC[s1 + g];
        }
      ds += __shfl_xor(ds, 16, 64);
      ds += __shfl_xor(ds, 32, 64);
      // dem slot 0 of lane group g = c holds dE/d em[row r][c] (direct part)
      if (in) ndi[4 * row + g] = row < nv ? dem[0] + (g == 0 ? ds : 0.) : 0.;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

int make_params64(Embed64Params* p, const int* sel_host, int ntypes, int n1, int64_t nloc) {
  DPMI_CHECK_ARG(sel_host != nullptr, "sel is NULL");
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= kMaxTypes64, "ntypes out of range [1, 16]");
  DPMI_CHECK_ARG(n1 > 0 && n1 <= 32, "n1 out of range [1, 32]");
  std::memset(p, 0, sizeof(*p));
  p->ntypes = ntypes;
  p->n1 = n1;
  for (int t = 0; t < ntypes; ++t) {
    DPMI_CHECK_ARG(sel_host[t] >= 0, "sel must be >= 0");
    p->sec[t + 1] = p->sec[t] + sel_host[t];
  }
  p->nnei = p->sec[ntypes];
  // One workgroup per CU at a time (the weights fill most of the LDS): two
  // full rounds over the CUs, apportioned to the types in proportion to their
  // rows by largest remainder so that the total is exact (a rounded-up share
  // per type would leave one workgroup running alone in a third round).
  static int ncu = 0;
  if (ncu == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      ncu = v;
    else
      ncu = 256;
  }
  const int64_t cap = ceil_div(std::max<int64_t>(nloc, 1), kWaves64);  // workgroups with an atom for each wave
  const int64_t total = std::max<int64_t>(ntypes, std::min<int64_t>(2 * ncu, cap * ntypes));
  int64_t share[kMaxTypes64], rem[kMaxTypes64], used = 0;
  for (int t = 0; t < ntypes; ++t) {
    const int64_t num = p->nnei > 0 ? total * sel_host[t] : total;
    const int64_t den = p->nnei > 0 ? p->nnei : ntypes;
    share[t] = num / den;
    rem[t] = num % den;
    used += share[t];
  }
  for (int64_t left = total - used; left > 0; --left) {  // the largest remainders get one more
    int best = 0;
    for (int t = 1; t < ntypes; ++t)
      if (rem[t] > rem[best]) best = t;
    ++share[best];
    rem[best] = -1;
  }
  p->blk[0] = 0;
  for (int t = 0; t < ntypes; ++t) {
    const int64_t b = std::max<int64_t>(1, std::min<int64_t>(share[t], cap));
    p->blk[t + 1] = p->blk[t] + (int)b;
  }
  p->nblk_total = p->blk[ntypes];
  return DPMI_OK;
}

size_t lds_bwd() { return kLdsW + (size_t)kWaves64 * 128 * 4 * sizeof(double); }

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_se_a_embed_fwd_f64(double* gr_part, const double* em, const double* w2, const double* w3,
                                       const double* consts, int64_t nloc, const int* sel_host, int ntypes, int n1,
                                       const int16_t* counts, const int* atype, const double* pad_em,
                                       void* stream) {
  Embed64Params p;
  int rc = make_params64(&p, sel_host, ntypes, n1, nloc);
  if (rc) return rc;
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  DPMI_CHECK_ARG(counts == nullptr || (atype != nullptr && pad_em != nullptr),
                 "the padding shortcut needs atype and pad_em with counts");
  p.counts = counts;
  p.atype = atype;
  p.pad_em = pad_em;
  if (nloc == 0) return DPMI_OK;
  auto launch = [&](auto kern) -> int {
    DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kLdsFwd));
    hipLaunchKernelGGL(kern, dim3(p.nblk_total), dim3(64 * kWaves64), kLdsFwd, as_stream(stream), gr_part, em, w2,
                       w3, consts, (int)nloc, p);
    return DPMI_OK;
  };
  rc = n1 <= 24 ? launch(k_embed64_fwd<24>) : n1 <= 28 ? launch(k_embed64_fwd<28>) : launch(k_embed64_fwd<32>);
  if (rc) return rc;
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_se_a_embed_bwd_f64(double* net_deriv, const double* em, const double* dgr, const double* w2,
                                       const double* w3, const double* consts, int64_t nloc, const int* sel_host,
                                       int ntypes, int n1, const int16_t* counts, void* stream) {
  Embed64Params p;
  int rc = make_params64(&p, sel_host, ntypes, n1, nloc);
  if (rc) return rc;
  p.counts = counts;
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  if (nloc == 0) return DPMI_OK;
  const size_t shm = lds_bwd();
  auto launch = [&](auto kern) -> int {
    DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)shm));
    hipLaunchKernelGGL(kern, dim3(p.nblk_total), dim3(64 * kWaves64), shm, as_stream(stream), net_deriv, em, dgr, w2,
                       w3, consts, (int)nloc, p);
    return DPMI_OK;
  };
  rc = n1 <= 24 ? launch(k_embed64_bwd<24>) : n1 <= 28 ? launch(k_embed64_bwd<28>) : launch(k_embed64_bwd<32>);
  if (rc) return rc;
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}
