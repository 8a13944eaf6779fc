// Device and host utilities shared by the plate and bed kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/tensordem_amd.h"

namespace tdm {

constexpr int kThreads = 256;
constexpr int kCh = 16;  // ledger partial channels per block
constexpr double kEps64 = 2.220446049250313e-16;  // torch.finfo(float64).eps
constexpr float kEps32 = 1.1920928955078125e-07f;

__host__ __device__ inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// FP32 force path: hardware reciprocal and square root (at most 1 ulp), far
// cheaper than the correctly rounded expansions; break decisions within the
// FP32 margin are re-done in FP64, so these never decide a fracture.
__device__ inline float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float sqrt32(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float rsq32(float x) { return __builtin_amdgcn_rsqf(x); }

// Deterministic wave-wide sum of a double: a fixed DPP shift/broadcast tree
// (row_shr 1,2,4,8 inside each 16-lane row, then row_bcast 15 and 31), both
// 32-bit halves moved with the same control.  The total is read from lane 63
// and returned in every lane.
template <int CTRL, int ROWMASK>
__device__ inline double dpp_move(double v)
{
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWMASK, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWMASK, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ inline double wave_sum(double v)
{
  v += dpp_move<0x111, 0xF>(v);
  v += dpp_move<0x112, 0xF>(v);
  v += dpp_move<0x114, 0xF>(v);
  v += dpp_move<0x118, 0xF>(v);
  v += dpp_move<0x142, 0xA>(v);
  v += dpp_move<0x143, 0xC>(v);
  int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Deterministic block sum of kCh channels; result valid in thread 0.
template <int NCH>
__device__ void block_sum(double (&ch)[NCH], double (*lds)[kCh])
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
    // most channels are zero in most waves (breaks, loads, supports, tool):
    // a wave-uniform skip keeps their DPP trees off the critical path
    double s = 0.0;
    if (__ballot(ch[k] != 0.0)) s = wave_sum(ch[k]);
    if (lane == 0) lds[wave][k] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NCH; ++k) ch[k] = ((lds[0][k] + lds[1][k]) + lds[2][k]) + lds[3][k];
  }
}

template <typename R> struct Prec;
template <> struct Prec<double> { static constexpr bool fp64 = true; };
template <> struct Prec<float> { static constexpr bool fp64 = false; };

// grid-stride fill (used instead of hipMemsetAsync so clears stay graph-safe)
template <typename T>
__global__ void fill_kernel(T *p, size_t n, T v)
{
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}

// error reporting for the C ABI (one message per host thread)
std::string &last_error();
int fail(int code, const std::string &msg);

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return tdm::fail(TDM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

}  // namespace tdm

// captured step sequence (tdm_graph in the C ABI)
struct tdm_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};
