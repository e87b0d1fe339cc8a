// Internal feature order of the fused bf16 embedding kernels, shared with
// the fitting kernel that reads their G^T R partials and writes their dGR.
// Layer-1 features f < n1 live in slots q1 < 32; layer 2 (2 n1 features) in
// q2 = q1(f mod n1) + 32 [f >= n1]; G (4 n1 features) in
// q3 = q2(f mod 2 n1) + 64 [f >= 2 n1], so the [x, x] skip connections land
// on the same lane and register.  Slots without a feature are padding.
#pragma once

#include <hip/hip_runtime.h>

namespace dpmi {

// Layer-1 slots.  With n1 <= 28 the slots 19, 23, 27, 31 stay empty: they are
// B-slot u = 7 of every lane group (phi(g, 7) = 19 + 4 g), and through the
// skip connections the (mt odd, v = 3) accumulator rows of layers 2 and 3, so
// the kernels skip those tanh evaluations outright.  Wider layers use slots
// 0 .. n1 - 1.
__host__ __device__ inline bool skip_layout(int n1) { return n1 <= 28; }

__host__ __device__ inline int q1_of(int f, int n1) {
  return (!skip_layout(n1) || f < 16) ? f : 16 + 4 * ((f - 16) / 3) + (f - 16) % 3;
}

// original layer-1 feature in slot q, or -1 for an empty slot
__host__ __device__ inline int f1_of(int q, int n1) {
  int f = q;
  if (skip_layout(n1) && q >= 16) f = (q & 3) == 3 ? -1 : q - ((q - 16) >> 2);
  return f < n1 ? f : -1;
}

// original feature (0 .. 4 n1 - 1) of internal G index q, or -1 for padding
__host__ __device__ inline int orig_of_q3(int q, int n1) {
  const int base = q >= 64 ? 2 * n1 : 0, q2 = q & 63;
  const int f1 = f1_of(q2 & 31, n1);
  return f1 < 0 ? -1 : base + (q2 >= 32 ? n1 : 0) + f1;
}

// internal G index of original feature f (n1 = width of layer 1)
__host__ __device__ inline int q3_of(int f, int n1) {
  int n2 = 2 * n1;
  int base = f >= n2 ? 64 : 0;
  int f2 = f >= n2 ? f - n2 : f;
  return base + (f2 >= n1 ? 32 + q1_of(f2 - n1, n1) : q1_of(f2, n1));
}

}  // namespace dpmi
