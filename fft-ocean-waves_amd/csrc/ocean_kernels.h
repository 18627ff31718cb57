// ocean_kernels.h -- internal interface between the C ABI (ocean_capi.cpp) and
// the gfx950 kernels (ocean_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ocn {

constexpr int kMaxFrames        = 32;  // frames per launch of the fused path
constexpr int kNumColsVariants  = 21;  // column-pass (kernel, radix, columns/block) schedules
constexpr int kNumRowsVariants  = 9;   // row-pass schedules (radix, fused band, parallel bands)

// Per-ocean spectrum constants, FP32 for the kernels plus the FP64 pieces of
// the phase.  Filled by the host from ocn_params.
struct SpecConst {
    int      N;
    float    kscale;      // 2 pi / L
    float    Lp2;         // (V^2 / g)^2
    float    l2;          // suppression^2
    float    A;           // Phillips amplitude
    float    wx, wy;      // unit wind direction
    float    g;           // gravity
    uint32_t key0, key1;  // Philox key (seed)
    uint32_t ocean_id, cascade_id;
    int      conj_mode;   // 0: reference literal (identity), 1: textbook conj
    double   k2_d;        // (2 pi / L)^2
    double   g2_d;        // g^2
};

// Frames of one launch: times and the element strides between frames of the
// intermediate and output buffers.
struct FrameBatch {
    double t_rev[kMaxFrames];  // t / (2 pi) per frame
    size_t z1_stride;          // float2 elements
    size_t z2_stride;          // float2 elements
    size_t out_stride;         // float elements of height/dx/dz/jacobian (normals: 4x)
    int    nframes;            // frames in this launch (<= kMaxFrames)
};

struct FrameArgs {
    const float4* h0;
    const float4* h0t;        // h0 transposed ([x][y]) for the N >= 2048 column strips, else null
    float2*       z1;         // nframes x N x N
    float2*       z2;         // nframes x N/2 x N
    const float2* tw[3];      // Stockham twiddles for radix 4, 8, 16
    SpecConst     sc;
    FrameBatch    fb;
    int           nframes;
    int           cols_variant;
    int           rows_variant;
    float*        height;
    float*        dx;
    float*        dz;
    float4*       normals;    // may be null
    float*        jacobian;   // may be null
    float4*       disp;       // render export, may be null
    float4*       rnorm;      // render export, may be null
    int           normal_mode;
    float         lambda;
    float         dscale;     // render height scale
    float         inv2d;      // N / (2 L)
};

// Static intermediates of ocn_update_frames_noalloc: one file-scope device
// array per GPU (ocean_kernels.hip), 48 MiB = Z1 + Z2 of one N=2048 frame,
// four N=1024 frames, sixteen N=512 frames.
constexpr size_t kStaticScratchBytes = (size_t)48 << 20;
constexpr int    kMaxScratchDevices  = 64;
// Looks up the array's address on `device` (which must be current) once and
// keeps it; a setup call (it may load the code object on that device).
hipError_t static_scratch_resolve(int device);
// The address static_scratch_resolve kept for `device`, or null.
void*      static_scratch_ptr(int device);

// h0 of the spectrum, and when h0t is given its transpose h0t[x * N + y] = h0[y * N + x]
hipError_t launch_h0(int N, float4* h0, float4* h0t, const uint8_t* noise, const SpecConst& sc, hipStream_t s);
hipError_t launch_spectrum_t(int N, const float4* h0, float2* hy, float2* hx, float2* hz, const SpecConst& sc,
                             double t_rev, hipStream_t s);
// stages: bit 0 = spectrum + column transforms, bit 1 = row transforms +
// inversion, bit 2 = normal map + Jacobian (7 = whole frame).
hipError_t launch_frame(int N, const FrameArgs& a, int stages, hipStream_t s);
hipError_t launch_ifft2(int N, const float2* src, float2* tmp, float* out, const float2* const* tw, hipStream_t s);
hipError_t launch_normals(int N, const float* height, const float* dx, const float* dz, float4* normals,
                          float* jacobian, int mode, float lambda, float inv2d, hipStream_t s);
bool       schedule_valid(int N, int cols_variant, int rows_variant);
void       cols_variant_shape(int v, int* E, int* C);
int        rows_variant_radix(int v);
int        rows_variant_band(int v);  // 0: separate normal-map kernel; R: fused bands of R rows
bool       rows_variant_parallel(int v);  // fused bands whose rows are transformed at once

}  // namespace ocn
