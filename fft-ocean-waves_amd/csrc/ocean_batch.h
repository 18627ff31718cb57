// ocean_batch.h -- internal interface of the batched kernels (ocean_batch.hip)
// for the C ABI (ocean_capi.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ocean_kernels.h"

namespace ocn {

// Block shapes shared by every N in a batch.
constexpr int kBatchColsThreads = 768;   // 3 * C * (N/16) with C = kBatchColsWidth / N
constexpr int kBatchColsWidth   = 4096;  // C * N: columns per column block times N (N = 64: C = N/2)
constexpr int kBatchRowsThreads = 256;   // (256 / T) groups of T = N/8 threads
constexpr int kBatchMinN        = 64;
constexpr int kBatchMaxN        = 4096;
// Members of this N walk their row bands in a launch of their own: a radix-8
// row group of N = 4096 is 512 threads, wider than the shared row block.
constexpr int kBatchWideN           = 4096;
constexpr int kBatchWideRowsThreads = kBatchWideN / 8;

// Row-pass radix and default band height per N: radix 8 keeps the fused row
// kernel at two waves per SIMD; short bands at small N keep a row block busy.
// The band height of each member can be changed (ocn_batch_set_schedule,
// ocn_batch_autotune) among the heights batch_band_valid accepts.
__host__ __device__ constexpr int batch_rows_radix(int) { return 8; }
__host__ __device__ constexpr int batch_rows_band(int N) { return N <= 64 ? 2 : N <= 512 ? 8 : N <= 1024 ? 16 : 32; }

// Device-side descriptor of one batch member (one ocean cascade, one frame).
struct BatchMember {
    const float4* h0;
    float2*       z1;
    float2*       z2;
    const float2* tw8;
    const float2* tw16;
    float*        height;
    float*        dx;
    float*        dz;
    float4*       normals;
    float*        jac;
    SpecConst     sc;
    float         lambda;
    float         inv2d;
    float4*       disp;    // render export (k_batch_rows_render only)
    float4*       rnorm;
    float         dscale;  // displacement_scale
    int           band;    // rows per band of the row pass
    int           cols;    // columns per strip of the column pass
};

// Work items: int2 (member, strip) for the column pass, (member, first band)
// for the row pass (each row block covers 256 / T consecutive bands).
__host__ __device__ constexpr int batch_cols_per_block(int N)
{
    return kBatchColsWidth / N < N / 2 ? kBatchColsWidth / N : N / 2;
}
// Column strip widths a member of size N may use: the default C = 4096/N (N/2
// at N = 64), whose 3C transform groups fill the 768-thread block, or half of
// it up to N = 1024 (the transforms then take half the block, every thread
// still folds and stores).  N = 2048 / 4096 transforms synchronise the whole
// block, so they keep the default.
inline bool batch_cols_valid(int N, int C)
{
    const int C0 = batch_cols_per_block(N);
    return C == C0 || (N <= 1024 && C0 >= 2 && C == C0 / 2);
}
// strips of cols_lds_strip: the self-mirror columns 0 and N/2 share one slot
inline int batch_cols_items_per_member(int N, int C) { return (N / 2) / C; }
// groups per row block: one band per block in the wide launch
inline int batch_rows_groups_per_block(int N)
{
    return N == kBatchWideN ? 1 : kBatchRowsThreads / (N / batch_rows_radix(N));
}
inline int batch_rows_items_per_member(int N, int band)
{
    const int gpb = batch_rows_groups_per_block(N), bands = N / band;
    return (bands + gpb - 1) / gpb;
}
// A band height R of the row pass at N: even, at most 32 (the walk is
// sequential within a band), dividing N, and where groups are wider than a
// wave (block barriers) a whole number of bands for every row block.
inline bool batch_band_valid(int N, int R)
{
    if (R < 2 || R > 32 || (R & (R - 1)) || N % R) return false;
    const int T = N / batch_rows_radix(N);
    return T <= 64 || (N / R) % batch_rows_groups_per_block(N) == 0;
}

// render: the row pass also writes each member's render export (disp, rnorm).
// wide_items: (member, band) of the N = 4096 members (n_wide may be 0).
hipError_t  launch_batch_frame(const BatchMember* members, const int2* cols_items, int n_cols, const int2* rows_items,
                               int n_rows, const int2* wide_items, int n_wide, double t_rev, bool render,
                               hipStream_t s);
const void* batch_cols_kernel();
const void* batch_rows_kernel(bool render);
const void* batch_wide_rows_kernel(bool render);

}  // namespace ocn
