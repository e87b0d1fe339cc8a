// FP64 descriptor + energy fitting net and their adjoint on gfx950
// (v_mfma_f64_16x16x4f64), and the compensated energy reduction: the plain
// FP64 form of the fused descriptor/fitting kernel of fitting_mfma16.hip.
//
// Reference semantics: the descriptor of DescrptBlockSeA.forward
// (deepmd/pt/model/descriptor/se_a.py:851-858; GR = sum_t G_t^T R_t / nnei,
// D[m][a] = sum_c GR[c][m] GR[c][a] for a < axis = 16, D row-major m * 16 + a),
// GeneralFitting._forward_common / EnergyFittingNet
// (deepmd/pt/model/task/fitting.py:734-900) with the FittingNet layers of
// deepmd/dpmodel/utils/network.py:1031-1090, own-type net per atom:
//   h1 = tanh(D W1 + b1) idt1,  h_l = tanh(h_{l-1} W_l + b_l) idt_l + h_{l-1} (l = 2, 3),
//   e  = h3 . w4 + b4 + bias_atom_e[t] + out_bias[t],
// the adjoint dE/dGR for E = sum_i e_i, and the energy sum of
// transform_output.py:165,181 in FP64, compensated.
//
// MI355X mapping.  A workgroup of eight wavefronts owns 32 atoms of one type
// (the atoms are sorted by type); wavefront w owns hidden features
// 32 w .. 32 w + 31 (two 16-feature M tiles of the width padded to 256), so a
// feature's tanh factors, skip value and adjoint stay in one lane's registers.
// Layers are computed transposed, y^T = W^T x^T: MFMA f64 16x16x4 with the
// weight fragment as A (lane l -> M row l%16, K l/16, streamed from L2) and the
// layer input as B (lane l -> K l/16, N col l%16 = atom).  Layer 1 never
// reads D: the tile's GR is staged in LDS ([atom][m][c]) and every lane forms
// its B value D[n][16 m + 4 q + g] = sum_c GR[n][c][m] GR[n][c][4 q + g] with
// four FMAs, the GR[n][.][4 q + g] (a < 16) held in registers.  The hidden
// layers read their input from an LDS [atom][feature] buffer that aliases the
// GR stage; the backward runs the transposed products through the same
// buffer and ends with dD = dz1 W1^T, one m (16 rows a) per M tile, which is
// contracted at once: dGR[c][m] gets sum_a dD[m][a] GR[c][a] (summed over the
// four lane groups) and dGR[c][a] (a < 16) gets sum_m dD[m][a] GR[c][m]
// (per-wavefront partials summed over the eight wavefronts in a fixed order).
// Every sum has a fixed order: the results do not depend on timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "dpmi_internal.h"
#include "tanh64.h"

namespace dpmi {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kT = 32;         // atoms per workgroup (two 16-atom N tiles)
constexpr int kWv = 8;         // wavefronts per workgroup
constexpr int kThreads = 64 * kWv;
constexpr int kHP = 256;       // padded hidden width (16 M tiles)
constexpr int kHS = kHP + 4;   // LDS row pitch (doubles) of the hidden buffer
constexpr int kMaxM = 128;     // widest G^T R
constexpr int kCst64 = 7 * kHP + 1;  // b1 idt1 b2 idt2 b3 idt3 w4 [256 each], b4
constexpr int kMaxParts = 16;
constexpr int kHB = 6;         // K steps per prefetched batch of the hidden-layer products
constexpr int kDB = 6;         // K steps per prefetched batch of the dD product

struct Fit64Params {
  int ntypes;
  int nloc;
  int M;       // G^T R width; D has 16 M columns (axis 16)
  int nparts;  // partial G^T R arrays summed into GR
  int nmt;     // live hidden M tiles, ceil(H / 16)
  int nkh;     // K steps of the hidden layers, ceil(H / 4)
  double inv_nnei;
  int64_t type_doubles;  // doubles per type in the fragment array
  int64_t w1f, w2f, w3f, w3b, w2b, w1b;
};

__device__ inline f64x4 mfma64(double a, double b, const f64x4& c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

inline size_t fit64_lds_doubles(int M) {
  return (size_t)std::max(kT * (4 * M + 2), kT * kHS) + (size_t)kT * 64 + (size_t)kWv * kT;
}

// acc[i][s] = sum over K steps of W(frag tile mt0 + i)[feature][k] * in[atom][k]
// for the wavefront's two M tiles (i) and the two atom tiles (s); the input
// comes from the LDS [atom][feature] buffer.  Tiles at or past nmt are skipped.
__device__ inline void hidden64(f64x4 (&acc)[2][2], const double* __restrict__ wf, const double* s_h, int mt0,
                                int nmt, int nkh, int lane, int n, int g) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int s = 0; s < 2; ++s) acc[i][s] = f64x4{0., 0., 0., 0.};
  const bool live1 = mt0 + 1 < nmt;
  const double* w0 = wf + (int64_t)mt0 * nkh * 64 + lane;
  const double* w1 = w0 + (int64_t)nkh * 64;
  // software pipeline: the fragments of the next kHB K steps load while this
  // batch's products run (one batch is shorter than an L2 round trip)
  double an[kHB][2];
  auto load = [&](int kb, double (&a)[kHB][2]) {
#pragma unroll
    for (int u = 0; u < kHB; ++u) {
      const int ks = kb + u;
      a[u][0] = ks < nkh ? w0[(int64_t)ks * 64] : 0.;
      a[u][1] = live1 && ks < nkh ? w1[(int64_t)ks * 64] : 0.;
    }
  };
  load(0, an);
#pragma unroll 1
  for (int kb = 0; kb < nkh; kb += kHB) {
    double a[kHB][2];
#pragma unroll
    for (int u = 0; u < kHB; ++u) {
      a[u][0] = an[u][0];
      a[u][1] = an[u][1];
    }
    if (kb + kHB < nkh) load(kb + kHB, an);
#pragma unroll
    for (int u = 0; u < kHB; ++u) {
      const int ks = kb + u;
      if (ks >= nkh) break;
      const double b0 = s_h[n * kHS + 4 * ks + g];
      const double b1 = s_h[(16 + n) * kHS + 4 * ks + g];
      acc[0][0] = mfma64(a[u][0], b0, acc[0][0]);
      acc[0][1] = mfma64(a[u][0], b1, acc[0][1]);
      if (live1) {
        acc[1][0] = mfma64(a[u][1], b0, acc[1][0]);
        acc[1][1] = mfma64(a[u][1], b1, acc[1][1]);
      }
    }
  }
}

__global__ void __launch_bounds__(kThreads)
    k_fit64(double* __restrict__ atom_energy, double* __restrict__ dgr, const double* __restrict__ gr_part,
            const int* __restrict__ perm, const int* __restrict__ type_start, const double* __restrict__ frag_all,
            const double* __restrict__ consts, const double* __restrict__ ebias, Fit64Params p) {
  extern __shared__ double smem[];
  const int M = p.M, gs = 4 * M + 2;
  double* s_buf = smem;  // GR stage [atom][m][c] (pitch gs), then the hidden buffer [atom][kHS]
  double* s_t1 = smem + std::max(kT * gs, kT * kHS);  // term 1 of dGR for m < 16: [atom][c][16]
  double* s_e = s_t1 + kT * 64;                        // energy partials [wave][atom]
  __shared__ int s_atom[kT];

  // tile -> (type, first sorted index); the same answer in every thread
  int t = -1, start = 0, cnt = 0, before = 0;
  for (int tt = 0; tt < p.ntypes; ++tt) {
    const int lo = type_start[tt], hi = type_start[tt + 1];
    const int nt = (hi - lo + kT - 1) / kT;
    if (t < 0 && (int)blockIdx.x < before + nt) {
      t = tt;
      start = lo + ((int)blockIdx.x - before) * kT;
      cnt = min(kT, hi - start);
    }
    before += nt;
  }
  if (t < 0) return;

  const int tid = threadIdx.x, w = wave_id(), lane = tid & 63, n = lane & 15, g = lane >> 4;
  if (tid < kT) s_atom[tid] = tid < cnt ? perm[start + tid] : -1;
  const double* frag = frag_all + (int64_t)t * p.type_doubles;
  const double* cst = consts + (int64_t)t * kCst64;
  __syncthreads();

  // ---- GR = inv_nnei * sum over partials, staged as [atom][m][c]: thread
  // per (atom, m), its four components read as coalesced scalars (lanes over
  // consecutive m) and written as four consecutive doubles
  const int per_atom = 4 * M;
  for (int k = tid; k < kT * M; k += kThreads) {
    const int la = k / M, m = k - la * M;
    const int ai = s_atom[la];
    double v[4] = {0., 0., 0., 0.};
    if (ai >= 0)
      for (int q = 0; q < p.nparts; ++q) {
        const double* src = gr_part + ((int64_t)q * p.nloc + ai) * per_atom + m;
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] += src[c * M];
      }
    double* dst = s_buf + la * gs + 4 * m;
#pragma unroll
    for (int c = 0; c < 4; ++c) dst[c] = v[c] * p.inv_nnei;
  }
  __syncthreads();
  const int mt0 = 2 * w;
  const bool live0 = mt0 < p.nmt, live1 = mt0 + 1 < p.nmt;
  // ---- layer 1: K = 16 M, k = 16 m + 4 q + kk (kk = K row of the lane = g)
  f64x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int s = 0; s < 2; ++s) acc[i][s] = f64x4{0., 0., 0., 0.};
  if (live0) {
    // GR[atom][c][a] of this lane's atoms for a = 4 q + g (q < 4): the second
    // factor of every D value it forms
    double ga[2][4][4];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) ga[s][q][c] = s_buf[(16 * s + n) * gs + 4 * (4 * q + g) + c];
    const int nk1 = 4 * M;
    const double* w0 = frag + p.w1f + (int64_t)mt0 * nk1 * 64 + lane;
    const double* w1 = w0 + (int64_t)nk1 * 64;
    // W1 fragments two m-steps ahead (one step's products are shorter than
    // an L2 round trip)
    auto load_m = [&](int m, double (&a)[2][4]) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[0][q] = m < M ? w0[(int64_t)(4 * m + q) * 64] : 0.;
        a[1][q] = live1 && m < M ? w1[(int64_t)(4 * m + q) * 64] : 0.;
      }
    };
    double an[2][4], ann[2][4];
    load_m(0, an);
    load_m(1, ann);
#pragma unroll 1
    for (int m = 0; m < M; ++m) {
      double a[2][4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[0][q] = an[0][q];
        a[1][q] = an[1][q];
        an[0][q] = ann[0][q];
        an[1][q] = ann[1][q];
      }
      if (m + 2 < M) load_m(m + 2, ann);
      double gm[2][4];
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int c = 0; c < 4; ++c) gm[s][c] = s_buf[(16 * s + n) * gs + 4 * m + c];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double b[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
          b[s] = fma(gm[s][3], ga[s][q][3], fma(gm[s][2], ga[s][q][2], fma(gm[s][1], ga[s][q][1], gm[s][0] * ga[s][q][0])));
        acc[0][0] = mfma64(a[0][q], b[0], acc[0][0]);
        acc[0][1] = mfma64(a[0][q], b[1], acc[0][1]);
        if (live1) {
          acc[1][0] = mfma64(a[1][q], b[0], acc[1][0]);
          acc[1][1] = mfma64(a[1][q], b[1], acc[1][1]);
        }
      }
    }
  }
  // layer-1 epilogue: D slot v of tile mt holds feature 16 mt + 4 v + g of atom n
  double h[2][2][4], f1[2][2][4], f2[2][2][4], dz[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = 16 * (mt0 + i) + 4 * v + g;
        const double tt = tanh64(acc[i][s][v] + cst[j]);
        const double idt = cst[kHP + j];
        h[i][s][v] = tt * idt;
        f1[i][s][v] = idt * (1. - tt * tt);
      }
  __syncthreads();  // every wavefront is done with the GR stage
  auto put = [&](const double (&x)[2][2][4]) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (i == 0 ? !live0 : !live1) continue;
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int v = 0; v < 4; ++v) s_buf[(16 * s + n) * kHS + 16 * (mt0 + i) + 4 * v + g] = x[i][s][v];
    }
  };
  put(h);
  __syncthreads();

  // ---- layers 2 and 3 (H -> H with the skip connection)
#pragma unroll 1
  for (int layer = 0; layer < 2; ++layer) {
    if (live0) hidden64(acc, frag + (layer == 0 ? p.w2f : p.w3f), s_buf, mt0, p.nmt, p.nkh, lane, n, g);
    __syncthreads();  // every wavefront has read the layer input
    const double* bl = cst + (2 + 2 * layer) * kHP;
    const double* il = cst + (3 + 2 * layer) * kHP;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = 16 * (mt0 + i) + 4 * v + g;
          const double tt = tanh64(acc[i][s][v] + bl[j]);
          const double idt = il[j];
          h[i][s][v] += tt * idt;
          if (layer == 0)
            f2[i][s][v] = idt * (1. - tt * tt);
          else  // dz3 = dE/dh3 f3 with dE/dh3 = w4
            dz[i][s][v] = cst[6 * kHP + j] * idt * (1. - tt * tt);
        }
    if (layer == 0) {
      put(h);
      __syncthreads();
    }
  }

  // ---- atomic energies (fixed-order sums) and dz3 into the buffer
  const double* w4 = cst + 6 * kHP;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    double e = 0.;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) e = fma(h[i][s][v], w4[16 * (mt0 + i) + 4 * v + g], e);
    e += __shfl_xor(e, 16, 64);
    e += __shfl_xor(e, 32, 64);
    if (g == 0) s_e[w * kT + 16 * s + n] = e;
  }
  put(dz);
  __syncthreads();
  if (tid < cnt) {
    double e = 0.;
    for (int ww = 0; ww < kWv; ++ww) e += s_e[ww * kT + tid];
    atom_energy[s_atom[tid]] = e + cst[7 * kHP] + ebias[t];
  }

  // ---- backward: dh2 = W3 dz3 + w4 (dE/dh3 = w4 and the skip), dz2 = dh2 f2,
  // dh1 = W2 dz2 + dh2, dz1 = dh1 f1
#pragma unroll 1
  for (int layer = 0; layer < 2; ++layer) {
    if (live0) hidden64(acc, frag + (layer == 0 ? p.w3b : p.w2b), s_buf, mt0, p.nmt, p.nkh, lane, n, g);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (layer == 0) {
            const int j = 16 * (mt0 + i) + 4 * v + g;
            h[i][s][v] = acc[i][s][v] + w4[j];  // h now holds dh2
            dz[i][s][v] = h[i][s][v] * f2[i][s][v];
          } else {
            dz[i][s][v] = (acc[i][s][v] + h[i][s][v]) * f1[i][s][v];
          }
        }
    put(dz);
    __syncthreads();
  }

  // ---- dD = W1 dz1 one m at a time (M tile md: rows a = 4 v + g), contracted into dGR
  double t2[2][4][4];  // [s][v][c]: sum_m dD[m][4 v + g] GR[c][m]
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int c = 0; c < 4; ++c) t2[s][v][c] = 0.;
  int ai_s[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) ai_s[s] = s_atom[16 * s + n];
  // GR[c][a] for a = 4 v + g again (the stage was overwritten), from the partials
  auto gr_at = [&](int ai, int c, int m) {
    double v = 0.;
    if (ai >= 0)
      for (int q = 0; q < p.nparts; ++q) v += gr_part[(((int64_t)q * p.nloc + ai) * 4 + c) * M + m];
    return v * p.inv_nnei;
  };
  double ga[2][4][4];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int c = 0; c < 4; ++c) ga[s][v][c] = gr_at(ai_s[s], c, 4 * v + g);
  const double* w1b = frag + p.w1b;
  // software pipeline over (m, K batch): the fragments of the next batch
  // (or of the first batch of this wavefront's next m) and the GR column of
  // the next m load while the current products run
  auto load_d = [&](int md, int kb, double (&a)[kDB]) {
    const double* wrow = w1b + (int64_t)md * p.nkh * 64 + lane;
#pragma unroll
    for (int u = 0; u < kDB; ++u) a[u] = md < M && kb + u < p.nkh ? wrow[(int64_t)(kb + u) * 64] : 0.;
  };
  auto load_gm = [&](int md, double (&gm)[2][4]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c) gm[s][c] = md < M ? gr_at(ai_s[s], c, md) : 0.;
  };
  double an[kDB], gmn[2][4];
  load_d(w, 0, an);
  load_gm(w, gmn);
#pragma unroll 1
  for (int md = w; md < M; md += kWv) {
    f64x4 d[2] = {f64x4{0., 0., 0., 0.}, f64x4{0., 0., 0., 0.}};
    double gms[2][4];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c) gms[s][c] = gmn[s][c];
    load_gm(md + kWv, gmn);
#pragma unroll 1
    for (int kb = 0; kb < p.nkh; kb += kDB) {
      double a[kDB];
#pragma unroll
      for (int u = 0; u < kDB; ++u) a[u] = an[u];
      if (kb + kDB < p.nkh)
        load_d(md, kb + kDB, an);
      else
        load_d(md + kWv, 0, an);
#pragma unroll
      for (int u = 0; u < kDB; ++u) {
        const int ks = kb + u;
        if (ks >= p.nkh) break;
        d[0] = mfma64(a[u], s_buf[n * kHS + 4 * ks + g], d[0]);
        d[1] = mfma64(a[u], s_buf[(16 + n) * kHS + 4 * ks + g], d[1]);
      }
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int ai = ai_s[s];
      double gm[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) gm[c] = gms[s][c];
      double t1[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double x = 0.;
#pragma unroll
        for (int v = 0; v < 4; ++v) x = fma(d[s][v], ga[s][v][c], x);
        x += __shfl_xor(x, 16, 64);
        x += __shfl_xor(x, 32, 64);
        t1[c] = x;
#pragma unroll
        for (int v = 0; v < 4; ++v) t2[s][v][c] = fma(d[s][v], gm[c], t2[s][v][c]);
      }
      if (g == 0 && ai >= 0) {
        if (md < 16) {
#pragma unroll
          for (int c = 0; c < 4; ++c) s_t1[(16 * s + n) * 64 + c * 16 + md] = t1[c];
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) dgr[((int64_t)ai * 4 + c) * M + md] = t1[c] * p.inv_nnei;
        }
      }
    }
  }
  __syncthreads();  // dz1 and the term-1 stage are complete
  // term 2: per-wavefront partials [wave][atom][a][2 components], two components at a time
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc)
          s_buf[((w * kT + 16 * s + n) * 16 + 4 * v + g) * 2 + cc] = t2[s][v][2 * ch + cc];
    __syncthreads();
    for (int k = tid; k < kT * 32; k += kThreads) {
      const int la = k >> 5, a = (k >> 1) & 15, cc = k & 1, c = 2 * ch + cc;
      const int ai = s_atom[la];
      if (ai < 0) continue;
      double x = s_t1[la * 64 + c * 16 + a];
      for (int ww = 0; ww < kWv; ++ww) x += s_buf[((ww * kT + la) * 16 + a) * 2 + cc];
      dgr[((int64_t)ai * 4 + c) * M + a] = x * p.inv_nnei;
    }
    __syncthreads();
  }
}

// E = sum of atom energies: Neumaier sums over a fixed stride, then a
// fixed-order tree of (sum, carry) pairs; one workgroup, deterministic.
__global__ void __launch_bounds__(1024) k_energy_sum(double* __restrict__ out, const double* __restrict__ ae,
                                                     int64_t n) {
  __shared__ double ss[1024], cs[1024];
  double s = 0., c = 0.;
  // eight independent loads in flight per thread, then their sums in order
  for (int64_t k0 = threadIdx.x; k0 < n; k0 += 8 * 1024) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = k0 + 1024 * u < n ? ae[k0 + 1024 * u] : 0.;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double t = s + x[u];
      c += fabs(s) >= fabs(x[u]) ? (s - t) + x[u] : (x[u] - t) + s;
      s = t;
    }
  }
  ss[threadIdx.x] = s;
  cs[threadIdx.x] = c;
  __syncthreads();
  for (int half = 512; half > 0; half >>= 1) {
    if ((int)threadIdx.x < half) {
      const double a = ss[threadIdx.x], b = ss[threadIdx.x + half], t = a + b;
      const double e = fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a;
      ss[threadIdx.x] = t;
      cs[threadIdx.x] = cs[threadIdx.x] + cs[threadIdx.x + half] + e;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = ss[0] + cs[0];
}

}  // namespace
}  // namespace dpmi

using namespace dpmi;

extern "C" int dpmi_se_a_fitting_gr_f64(double* atom_energy, double* dgr, const double* gr_part, int nparts,
                                        const int* perm, const int* type_start, const double* frags,
                                        const double* consts, const double* ebias, int64_t nloc, int ntypes, int M,
                                        int H, int nnei, void* stream) {
  DPMI_CHECK_ARG(nloc >= 0 && nloc < (int64_t(1) << 31), "nloc out of range");
  DPMI_CHECK_ARG(ntypes > 0 && ntypes <= 64, "ntypes out of range [1, 64]");
  DPMI_CHECK_ARG(nparts > 0 && nparts <= kMaxParts, "nparts out of range [1, 16]");
  DPMI_CHECK_ARG(M >= 16 && M <= kMaxM, "G^T R width M must be in [16, 128] (axis 16)");
  DPMI_CHECK_ARG(H > 0 && H <= kHP, "hidden width H must be in [1, 256]");
  DPMI_CHECK_ARG(nnei > 0, "nnei must be positive");
  if (nloc == 0) return DPMI_OK;
  DPMI_CHECK_ARG(atom_energy && dgr && gr_part && perm && type_start && frags && consts && ebias, "NULL argument");
  Fit64Params p;
  std::memset(&p, 0, sizeof(p));
  p.ntypes = ntypes;
  p.nloc = (int)nloc;
  p.M = M;
  p.nparts = nparts;
  p.nmt = (H + 15) / 16;
  p.nkh = (H + 3) / 4;
  p.inv_nnei = 1.0 / (double)nnei;
  const int64_t f1 = 16LL * 4 * M * 64, fh = 16LL * p.nkh * 64, fb1 = (int64_t)M * p.nkh * 64;
  p.w1f = 0;
  p.w2f = f1;
  p.w3f = p.w2f + fh;
  p.w3b = p.w3f + fh;
  p.w2b = p.w3b + fh;
  p.w1b = p.w2b + fh;
  p.type_doubles = p.w1b + fb1;
  // tiles: sum over types of ceil(count / 32) <= nloc / 32 + ntypes
  const int64_t nblk = nloc / kT + ntypes;
  const size_t shm = fit64_lds_doubles(M) * sizeof(double);
  DPMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_fit64), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shm));
  hipLaunchKernelGGL(k_fit64, dim3((unsigned)nblk), dim3(kThreads), shm, as_stream(stream), atom_energy, dgr, gr_part,
                     perm, type_start, frags, consts, ebias, p);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}

extern "C" int dpmi_se_a_fitting_f64_frag_doubles(int64_t* count, int M, int H) {
  DPMI_CHECK_ARG(count != nullptr && M >= 16 && M <= kMaxM && H > 0 && H <= kHP, "bad arguments");
  const int64_t nkh = (H + 3) / 4;
  *count = 16LL * 4 * M * 64 + 4 * 16LL * nkh * 64 + (int64_t)M * nkh * 64;
  return DPMI_OK;
}

extern "C" int dpmi_energy_sum_f64(double* energy, const double* atom_energy, int64_t n, void* stream) {
  DPMI_CHECK_ARG(energy != nullptr && n >= 0 && (n == 0 || atom_energy != nullptr), "bad arguments");
  hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(1024), 0, as_stream(stream), energy, atom_energy, n);
  DPMI_LAUNCH_CHECK();
  return DPMI_OK;
}
