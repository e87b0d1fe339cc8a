// Internal helpers shared by the HIP translation units of libdpmi.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "dpmi.h"

namespace dpmi {

// Thread-local last error text, read through dpmi_last_error().
void set_error(const std::string& msg);

// Cell grid of the neighbor search for this box (nlist.hip): number of local
// cells the workspace is sized for and the search radius in cells.
int cell_grid_info(const double* box_host, int64_t nloc, float rcut, int64_t* nlc, int* nsearch);

// Whether dpmi_build_nlist_f64 searches cell by cell for rows of nnei slots
// (candidate tables and cand_slot), or atom by atom (every cand_slot -1, no
// big-cell list) when the rows are too long for its LDS budget.
bool search_uses_cells(int nnei);

#define DPMI_CHECK_ARG(cond, msg)       \
  do {                                  \
    if (!(cond)) {                      \
      ::dpmi::set_error(msg);           \
      return DPMI_ERR_ARG;              \
    }                                   \
  } while (0)

#define DPMI_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::dpmi::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
      return DPMI_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

// Check the launch that was just queued.
#define DPMI_LAUNCH_CHECK() DPMI_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ----------------------------------------------------------- wave helpers
// MI355X (gfx950) runs 64-wide wavefronts.
constexpr int kWave = 64;

__device__ inline int lane_id() { return __lane_id(); }

// Index of the wavefront in its workgroup, as a wave-uniform (scalar) value:
// the compiler cannot tell that threadIdx.x / 64 is the same in every lane,
// and everything derived from it (atom indices, per-atom loads, addresses)
// would otherwise live in vector registers.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Lanes below this one whose bit is set in mask.
__device__ inline int mask_prefix(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0));
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

}  // namespace dpmi
