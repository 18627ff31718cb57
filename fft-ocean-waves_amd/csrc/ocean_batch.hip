// ocean_batch.hip -- batched frames of many oceans / cascades of uneven N in one
// launch per pass.  No reference counterpart: the reference simulates a single
// N=256 ocean (main.cpp:17); this is the multi-cascade extension.
//
// A batch member is one ocean cascade (its own N, parameters, h0, intermediate
// and outputs).  A frame is two launches over device tables of work items:
//   k_batch_cols: item (member, strip)   -> cols_lds_strip<N, C>, 768 threads,
//                 C = 4096/N columns (N/2 at N = 64) or, per member, half of it
//   k_batch_rows: item (member, band set) -> rows_nrm_band<N, 8>,        256 threads
//                 (bands of the member's R rows)
//                 (k_batch_rows_render: the same with each member's render export;
//                 members of N = 4096 walk their bands in k_batch_rows_wide, 512 threads)
// Both block shapes are the same for every N (C = 4096/N columns per column
// block, 256/T band groups per row block), so members of different N share a
// launch; the kernels dispatch on the member's N.  Blocks are independent (no
// inter-block waits).  Items are ordered member-major, and the XCD remap keeps
// neighbouring items, which share Z lines and halo rows, on one XCD.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "ocean_batch.h"
#include "ocean_device.hpp"

namespace ocn {

namespace {

// LDS of one block: the largest need over the batched sizes N = 64..2048
constexpr int cols_lds_need()
{
    int m = 0;
    for (int N = kBatchMinN; N <= kBatchMaxN; N *= 2) {
        const int need = 3 * batch_cols_per_block(N) * strip_len(N);
        m = m > need ? m : need;
    }
    return m;
}
constexpr int rows_lds_need()
{
    int m = 0;
    for (int N = kBatchMinN; N <= kBatchMaxN; N *= 2) {
        const int need = kBatchRowsThreads / (N / batch_rows_radix(N)) * lds_len(N);
        m = m > need ? m : need;
    }
    return m;
}
constexpr int kColsLds = cols_lds_need();  // 3 * C * strip_len(N)
constexpr int kRowsLds = rows_lds_need();  // (256 / T) * lds_len(N), T = N/8 (N = 4096: the wide launch)

template <int N>
__device__ __forceinline__ void batch_cols_item(const BatchMember& m, int strip, double t_rev, float2* lds)
{
    // m.cols is C or (N <= 1024) C/2 (batch_cols_valid, checked on the host)
    constexpr int C = batch_cols_per_block(N);
    using Sync = typename std::conditional<(N / 16 <= 64), WaveSync, BlockSync>::type;
    if constexpr (N <= 1024 && C >= 2) {
        if (m.cols != C) {
            cols_lds_strip<N, C / 2, Sync, kBatchColsThreads>(m.h0, m.z1, m.z2, m.tw16, m.sc, t_rev, strip, lds);
            return;
        }
    }
    cols_lds_strip<N, C, Sync, kBatchColsThreads>(m.h0, m.z1, m.z2, m.tw16, m.sc, t_rev, strip, lds);
}

template <int N, bool RENDER>
__device__ __forceinline__ void batch_rows_item(const BatchMember& m, int first_band, float2* lds)
{
    constexpr int E = batch_rows_radix(N), T = N / E;
    static_assert(T <= kBatchRowsThreads && kBatchRowsThreads % T == 0, "whole groups per row block");
    // m.band: even, dividing N, and with T > 64 (block barriers) a multiple
    // of the groups per block in bands, so that every group walks one
    // (batch_band_valid, checked on the host)
    const int R = m.band;
    using Sync = typename std::conditional<(T <= 64), WaveSync, BlockSync>::type;
    // a group of T >= 64 threads is made of whole waves: its index is
    // wave-uniform, and reading it from the first lane keeps the band scalar
    const int gl = (int)threadIdx.x / T, j = (int)threadIdx.x % T;
    const int g  = T >= 64 ? __builtin_amdgcn_readfirstlane(gl) : gl;
    const int band = first_band + g;
    if (band >= N / R) return;  // whole group (wave-aligned when GPB > 1)
    // Z2 rows paired too: the walk stays within 256 registers a lane, so two
    // 256-thread blocks still share a CU (the render variant reads them alone
    // to keep its extra outputs within the same budget).  Every band walks
    // downwards: the upward walk of odd bands (ALT) made the 192-cascade frame
    // 0.7% slower in this kernel, which holds every N in one code object.
    rows_nrm_band<N, E, Sync, RENDER, !RENDER>(m.z1, m.z2, m.height, m.dx, m.dz, m.normals, m.jac,
                                               E == 8 ? m.tw8 : m.tw16, lds + g * lds_len(N), j, band, R, m.lambda,
                                               m.inv2d, m.disp, m.rnorm, m.dscale);
}

template <bool RENDER>
__device__ __forceinline__ void batch_rows(const BatchMember* __restrict__ members, const int2* __restrict__ items,
                                           float2* lds)
{
    const int2        it = items[xcd_logical(blockIdx.x, gridDim.x)];
    const BatchMember m  = members[it.x];
    switch (m.sc.N) {
    case 64: batch_rows_item<64, RENDER>(m, it.y, lds); break;
    case 128: batch_rows_item<128, RENDER>(m, it.y, lds); break;
    case 256: batch_rows_item<256, RENDER>(m, it.y, lds); break;
    case 512: batch_rows_item<512, RENDER>(m, it.y, lds); break;
    case 1024: batch_rows_item<1024, RENDER>(m, it.y, lds); break;
    case 2048: batch_rows_item<2048, RENDER>(m, it.y, lds); break;
    default: break;
    }
}

}  // namespace

__global__ __launch_bounds__(kBatchColsThreads) void k_batch_cols(const BatchMember* __restrict__ members,
                                                                  const int2* __restrict__ items, double t_rev)
{
    __shared__ float2 lds[kColsLds];
    const int2        it = items[xcd_logical(blockIdx.x, gridDim.x)];
    const BatchMember m  = members[it.x];
    switch (m.sc.N) {
    case 64: batch_cols_item<64>(m, it.y, t_rev, lds); break;
    case 128: batch_cols_item<128>(m, it.y, t_rev, lds); break;
    case 256: batch_cols_item<256>(m, it.y, t_rev, lds); break;
    case 512: batch_cols_item<512>(m, it.y, t_rev, lds); break;
    case 1024: batch_cols_item<1024>(m, it.y, t_rev, lds); break;
    case 2048: batch_cols_item<2048>(m, it.y, t_rev, lds); break;
    case 4096: batch_cols_item<4096>(m, it.y, t_rev, lds); break;
    default: break;
    }
}

__global__ __launch_bounds__(kBatchRowsThreads) void k_batch_rows(const BatchMember* __restrict__ members,
                                                                  const int2* __restrict__ items)
{
    __shared__ float2 lds[kRowsLds];
    batch_rows<false>(members, items, lds);
}

// The same row pass with the render export of every member (disp, rnorm).
__global__ __launch_bounds__(kBatchRowsThreads) void k_batch_rows_render(const BatchMember* __restrict__ members,
                                                                         const int2* __restrict__ items)
{
    __shared__ float2 lds[kRowsLds];
    batch_rows<true>(members, items, lds);
}

// Row bands of the N = 4096 members: one band of 32 rows per 512-thread block
// (the radix-8 group of rows_nrm_band), Z2 rows read alone as in the
// single-ocean row kernel at this N.  Two waves per SIMD.
template <bool RENDER>
__device__ __forceinline__ void batch_wide_rows(const BatchMember* __restrict__ members,
                                                const int2* __restrict__ items, float2* lds)
{
    constexpr int N = kBatchWideN, E = 8;
    static_assert(N / E == kBatchWideRowsThreads, "one radix-8 group per block");
    const int2        it = items[xcd_logical(blockIdx.x, gridDim.x)];
    const BatchMember m  = members[it.x];
    rows_nrm_band<N, E, BlockSync, RENDER, false>(m.z1, m.z2, m.height, m.dx, m.dz, m.normals, m.jac, m.tw8, lds,
                                                  (int)threadIdx.x, it.y, m.band, m.lambda, m.inv2d, m.disp, m.rnorm,
                                                  m.dscale);
}

__global__ __launch_bounds__(kBatchWideRowsThreads, 2) void k_batch_rows_wide(
    const BatchMember* __restrict__ members, const int2* __restrict__ items)
{
    __shared__ float2 lds[lds_len(kBatchWideN)];
    batch_wide_rows<false>(members, items, lds);
}

__global__ __launch_bounds__(kBatchWideRowsThreads, 2) void k_batch_rows_wide_render(
    const BatchMember* __restrict__ members, const int2* __restrict__ items)
{
    __shared__ float2 lds[lds_len(kBatchWideN)];
    batch_wide_rows<true>(members, items, lds);
}

hipError_t launch_batch_frame(const BatchMember* members, const int2* cols_items, int n_cols, const int2* rows_items,
                              int n_rows, const int2* wide_items, int n_wide, double t_rev, bool render,
                              hipStream_t s)
{
    if (n_cols < 1 || n_rows + n_wide < 1 || n_rows < 0 || n_wide < 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_cols, dim3(n_cols), dim3(kBatchColsThreads), 0, s, members, cols_items, t_rev);
    if (n_rows > 0) {
        if (render)
            hipLaunchKernelGGL(k_batch_rows_render, dim3(n_rows), dim3(kBatchRowsThreads), 0, s, members, rows_items);
        else
            hipLaunchKernelGGL(k_batch_rows, dim3(n_rows), dim3(kBatchRowsThreads), 0, s, members, rows_items);
    }
    if (n_wide > 0) {
        if (render)
            hipLaunchKernelGGL(k_batch_rows_wide_render, dim3(n_wide), dim3(kBatchWideRowsThreads), 0, s, members,
                               wide_items);
        else
            hipLaunchKernelGGL(k_batch_rows_wide, dim3(n_wide), dim3(kBatchWideRowsThreads), 0, s, members,
                               wide_items);
    }
    return hipGetLastError();
}

const void* batch_cols_kernel() { return reinterpret_cast<const void*>(&k_batch_cols); }
const void* batch_rows_kernel(bool render)
{
    return render ? reinterpret_cast<const void*>(&k_batch_rows_render) : reinterpret_cast<const void*>(&k_batch_rows);
}
const void* batch_wide_rows_kernel(bool render)
{
    return render ? reinterpret_cast<const void*>(&k_batch_rows_wide_render)
                  : reinterpret_cast<const void*>(&k_batch_rows_wide);
}

}  // namespace ocn
