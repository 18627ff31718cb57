// ocean_kernels.hip -- gfx950 kernels of the Tessendorf ocean path.
//
// Reference passes (diharaw/fft-ocean-waves) and their kernels here:
//   tilde_h0_k         (main.cpp:553-583, tilde_h0_k_cs.glsl)   -> k_tilde_h0_k, k_tilde_h0_k_noise
//   tilde_h0_t         (main.cpp:587-608, tilde_h0_t_cs.glsl)   -> k_tilde_h0_t            (pass API)
//   butterfly_fft      (main.cpp:612-683, butterfly_cs.glsl,
//                       inversion_cs.glsl)                      -> k_c2c_rows + k_c2c_cols_inv (pass API)
//   generate_normal_map(main.cpp:687-707, normal_map_cs.glsl)   -> k_normals
//   update() frame     (main.cpp:240-244)                       -> k_cols + k_rows + k_normals
//
// Fused frame data flow (per texel, FP32):
//   h0 (float4: h0k, h0mk) --k_cols_fused--> Z1 (N x N float2), Z2 (N/2 x N float2)
//                          --k_rows--> height, dx, dz (float)
//                          --k_normals-----> normals (float4), Jacobian (float)
// k_cols evaluates the spectrum at texel t and at its mirror t' = (-t mod N),
// folds each field F into its Hermitian part A = (F(t) + conj F(t'))/2 (whose
// inverse DFT is Re of F's inverse DFT, which is all inversion_cs.glsl keeps),
// and runs the column (ty) transforms for the half plane tx in [0, N/2].  It
// stores the packed rows Z1 = G_h + i G_dx (full width via G(-tx) = conj G(tx))
// and Z2 = G_dz(2p) + i G_dz(2p + 1).  k_rows runs the row (tx)
// transforms: Re/Im of each result are two real outputs at once.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <type_traits>

#include "fft_core.hpp"
#include "ocean_device.hpp"
#include "ocean_kernels.h"

namespace ocn {

// ------------------------------------------------------------------ A3 tilde_h0_k

// tilde_h0_k_cs.glsl:74-94 with Philox uniforms (counter = (texel, ocean,
// cascade, 0), key = seed).  Output: float4 (h0k.re, h0k.im, h0mk.re, h0mk.im).
__global__ __launch_bounds__(256) void k_tilde_h0_k(float4* __restrict__ h0, SpecConst sc)
{
    const int tx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ty = blockIdx.y;
    if (tx >= sc.N) return;
    const uint32_t idx = (uint32_t)ty * (uint32_t)sc.N + (uint32_t)tx;
    const uint4 r  = philox4x32_10(make_uint4(idx, sc.ocean_id, sc.cascade_id, 0u), make_uint2(sc.key0, sc.key1));
    const float4 xi = box_muller(u01(r.x), u01(r.y), u01(r.z), u01(r.w));
    const float s  = phillips_amp(sc, tx, ty);
    h0[idx] = make_float4(s * xi.x, s * xi.y, s * xi.z, s * xi.w);
}

// h0t[x * N + y] = h0[y * N + x] (N a multiple of 32): 32 x 32 tiles through
// LDS, so both the reads and the writes are whole rows of a tile.  The column
// strips at N >= 2048 read h0 down their columns from this copy.
__global__ __launch_bounds__(256) void k_transpose_h0(const float4* __restrict__ h0, float4* __restrict__ h0t, int N)
{
    __shared__ float4 tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int x = threadIdx.x % 32, y = threadIdx.x / 32;  // 8 tile rows per pass
    for (int r = y; r < 32; r += 8) tile[r][x] = h0[(size_t)(by + r) * N + bx + x];
    __syncthreads();
    for (int r = y; r < 32; r += 8) h0t[(size_t)(bx + r) * N + by + x] = tile[x][r];
}

// Same with the reference's four 256x256 8-bit noise images sampled NEAREST +
// CLAMP_TO_EDGE at gid/N (tilde_h0_k_cs.glsl:53-58, main.cpp:1102-1116).
__global__ __launch_bounds__(256) void k_tilde_h0_k_noise(float4* __restrict__ h0, const uint8_t* __restrict__ noise,
                                                          SpecConst sc)
{
    const int tx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ty = blockIdx.y;
    if (tx >= sc.N) return;
    int ix = (tx * 256) / sc.N, iy = (ty * 256) / sc.N;
    ix = ix > 255 ? 255 : ix;
    iy = iy > 255 ? 255 : iy;
    const int o = iy * 256 + ix;
    const float4 xi = box_muller(noise[o] / 255.0f, noise[65536 + o] / 255.0f, noise[131072 + o] / 255.0f,
                                 noise[196608 + o] / 255.0f);
    const float s = phillips_amp(sc, tx, ty);
    h0[(size_t)ty * sc.N + tx] = make_float4(s * xi.x, s * xi.y, s * xi.z, s * xi.w);
}

// Pass API: the three spectrum textures of tilde_h0_t (dy, dx, dz), float2 each.
__global__ __launch_bounds__(256) void k_tilde_h0_t(const float4* __restrict__ h0, float2* __restrict__ hy,
                                                    float2* __restrict__ hx, float2* __restrict__ hz, SpecConst sc,
                                                    double t_rev)
{
    const int tx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ty = blockIdx.y;
    if (tx >= sc.N) return;
    const size_t i  = (size_t)ty * sc.N + tx;
    const int    nx = tx - sc.N / 2, ny = ty - sc.N / 2;
    float kinv, s, c;
    sincos_turns(phase_turns(sc, nx * nx + ny * ny, t_rev, kinv), s, c);
    const float2 h = evolve_amp(sc, h0[i], s, c);
    hy[i] = h;
    hx[i] = choppy((float)nx * kinv, h);
    hz[i] = choppy((float)ny * kinv, h);
}

// ------------------------------------------------------------------ fused frame, pass 1

// Column pass.  1-D grid of (strip groups x frames x G) blocks.  Block = C
// columns x T threads; thread (c, j) transforms column tx = strip*C + c of the
// half plane (tx <= N/2), rows ty = j + r*T; consecutive threads read
// consecutive columns, so a wave's loads cover 64/C rows of C adjacent texels.
// The three transforms (height, Dx, Dz) share the column's LDS buffer.
// Block order (after the XCD remap): group-major, then frame, then strip, so a
// group's h0 columns (the same for every frame) and the Z lines its G strips
// fill together are served by one XCD's L2.
template <int N, int E, int C>
__global__ __launch_bounds__(C*(N / E)) void k_cols(const float4* __restrict__ h0, float2* __restrict__ z1,
                                                    float2* __restrict__ z2, const float2* __restrict__ tw,
                                                    SpecConst sc, FrameBatch fb)
{
    constexpr int T = N / E;
    constexpr int G = strip_group<N, C>();
    __shared__ float2 lds[C * lds_len(N)];
    const int L     = xcd_logical(blockIdx.x, gridDim.x);
    const int rest  = L / G;
    const int f     = rest % fb.nframes;
    const int strip = (rest / fb.nframes) * G + L % G;
    if (strip > (N / 2) / C) return;  // padding of the last group (whole block)
    const double t_rev = fb.t_rev[f];
    z1 += (size_t)f * fb.z1_stride;
    z2 += (size_t)f * fb.z2_stride;
    const int  c     = threadIdx.x % C;
    const int  j     = threadIdx.x / C;
    const int  tx    = strip * C + c;
    const bool valid = tx <= N / 2;
    const int  txc   = valid ? tx : N / 2;
    const int  mxc   = (N - txc) & (N - 1);
    const int  nx = txc - N / 2, qx = mxc - N / 2;

    float2 ay[E], ax[E], az[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int ty = j + r * T;
        const int my = (N - ty) & (N - 1);
        const float4 p = h0[(size_t)ty * N + txc];
        const float4 q = h0[(size_t)my * N + mxc];
        pair_fold(sc, p, q, nx, ty - N / 2, qx, my - N / 2, t_rev, ay[r], ax[r], az[r]);
    }

    float2* buf = lds + c * lds_len(N);
    fft_inverse<N, E>(ay, j, buf, tw, BlockSync{});
    fft_inverse<N, E>(ax, j, buf, tw, BlockSync{});
    fft_inverse<N, E>(az, j, buf, tw, BlockSync{});
    // G_dz back through LDS in natural order: Z2 pairs adjacent rows 2p, 2p+1,
    // which sit in different threads.
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; ++r) buf[lds_pad(j + r * T)] = az[r];
    __syncthreads();

    if (!valid) return;
    const bool mirror = mxc != tx;
    // two-row tiles (zidx): threads j and j+1 of a column write the two rows
    // of a tile, so the C columns of a strip fill runs of 2C values
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int y = j + r * T;
        z1[zidx<N>(y, tx)] = make_float2(ay[r].x - ax[r].y, ay[r].y + ax[r].x);
        if (mirror) z1[zidx<N>(y, zpos<N>(mxc))] = make_float2(ay[r].x + ax[r].y, ax[r].x - ay[r].y);
    }
#pragma unroll
    for (int r = 0; r < E / 2; ++r) {
        const int    p  = j + r * T;
        const float2 lo = buf[lds_pad(2 * p)], hi = buf[lds_pad(2 * p + 1)];  // rows 2p and 2p+1
        z2[zidx<N>(p, tx)] = make_float2(lo.x - hi.y, lo.y + hi.x);
        if (mirror) z2[zidx<N>(p, zpos<N>(mxc))] = make_float2(lo.x + hi.y, hi.x - lo.y);
    }
}

// Column pass, LDS-staged variant (requires T = N/16 <= 64): the work and
// block order of k_cols, with each strip done by cols_lds_strip (every thread
// folds the spectrum into LDS, each transform runs inside one wave, every
// thread stores).  Block = 3C*T threads.
template <int N, int C, int W = 1>
__global__ __launch_bounds__(W * 3 * C * (N / 16)) void k_cols_lds(const float4* __restrict__ h0,
                                                                   float2* __restrict__ z1, float2* __restrict__ z2,
                                                                   const float2* __restrict__ tw, SpecConst sc,
                                                                   FrameBatch fb)
{
    // W = 2: twice the threads, the extra ones folding and storing only
    constexpr int T = N / 16;
    constexpr int G = strip_group<N, C>();
    static_assert(T <= 64 && 64 % T == 0, "each transform group must lie inside one wave");
    __shared__ float2 lds[3 * C * strip_len(N)];
    const int L     = xcd_logical(blockIdx.x, gridDim.x);
    const int rest  = L / G;
    const int f     = rest % fb.nframes;
    const int strip = (rest / fb.nframes) * G + L % G;
    if (strip >= lds_strips<N, C>()) return;  // padding of the last group (whole block)
    cols_lds_strip<N, C, WaveSync, W * 3 * C * T>(h0, z1 + (size_t)f * fb.z1_stride, z2 + (size_t)f * fb.z2_stride,
                                                  tw, sc, fb.t_rev[f], strip, lds);
}

// k_cols_lds over pairs of frames: block (strip, frames 2i and 2i+1) holds
// both frames of its C columns in LDS, loading each texel pair's h0 and
// evaluating its dispersion once for both (cols_lds_strip_f2); an odd last
// frame takes a block of its own.  Blocks of one strip group and consecutive
// frame pairs are consecutive on one XCD, as in k_cols_lds.  Block = 6C*T
// threads.
template <int N, int C>
__global__ __launch_bounds__(6 * C * (N / 16)) void k_cols_lds_f2(const float4* __restrict__ h0,
                                                                  float2* __restrict__ z1, float2* __restrict__ z2,
                                                                  const float2* __restrict__ tw, SpecConst sc,
                                                                  FrameBatch fb)
{
    constexpr int T = N / 16, NB = 6 * C * T;
    constexpr int G = strip_group<N, C>();
    static_assert(T <= 64 && 64 % T == 0, "each transform group must lie inside one wave");
    __shared__ float2 lds[6 * C * strip_len(N)];
    const int L     = xcd_logical(blockIdx.x, gridDim.x);
    const int pairs = (fb.nframes + 1) / 2;
    const int rest  = L / G;
    const int f     = 2 * (rest % pairs);
    const int strip = (rest / pairs) * G + L % G;
    if (strip >= lds_strips<N, C>()) return;  // padding of the last group (whole block)
    float2* z1a = z1 + (size_t)f * fb.z1_stride;
    float2* z2a = z2 + (size_t)f * fb.z2_stride;
    if (f + 1 < fb.nframes)
        cols_lds_strip_f2<N, C, WaveSync, NB>(h0, z1a, z2a, z1a + fb.z1_stride, z2a + fb.z2_stride, tw, sc,
                                              fb.t_rev[f], fb.t_rev[f + 1], strip, lds);
    else
        cols_lds_strip<N, C, WaveSync, NB>(h0, z1a, z2a, tw, sc, fb.t_rev[f], strip, lds);
}

// k_cols_lds for N >= 2048, whose N/16-thread transform groups span several
// waves: the same strip (cols_lds_strip) with block barriers between the
// transform stages, reading the transposed h0 (h0t), where a strip's columns
// are contiguous runs instead of one 16-byte value per 32 or 64 KB row
// (sequences of frames: N=2048 70 -> 64 us, N=4096 363 -> 350 us per frame).
// Block = 3C*T threads.
template <int N, int C>
__global__ __launch_bounds__(3 * C * (N / 16)) void k_cols_ldsb(const float4* __restrict__ h0t,
                                                                float2* __restrict__ z1, float2* __restrict__ z2,
                                                                const float2* __restrict__ tw, SpecConst sc,
                                                                FrameBatch fb)
{
    constexpr int G = strip_group<N, C>();
    static_assert(N / 16 > 64 && (N / 16) % 64 == 0, "transform groups of whole waves");
    __shared__ float2 lds[3 * C * strip_len(N)];
    const int L     = xcd_logical(blockIdx.x, gridDim.x);
    const int rest  = L / G;
    const int f     = rest % fb.nframes;
    const int strip = (rest / fb.nframes) * G + L % G;
    if (strip >= lds_strips<N, C>()) return;  // padding of the last group (whole block)
    cols_lds_strip<N, C, BlockSync, 3 * C * (N / 16), true>(h0t, z1 + (size_t)f * fb.z1_stride,
                                                            z2 + (size_t)f * fb.z2_stride, tw, sc, fb.t_rev[f], strip,
                                                            lds);
}

// Two-pass variant of k_cols_ldsb (cols_lds2b_strip, N >= 2048): 2C*T threads,
// half the LDS per column, the transposed h0.  Strips of C columns plus the
// column N/2 (the last strip), in the block order of k_cols.
template <int N, int C>
__global__ __launch_bounds__(2 * C * (N / 16)) void k_cols_lds2b(const float4* __restrict__ h0t,
                                                                 float2* __restrict__ z1, float2* __restrict__ z2,
                                                                 const float2* __restrict__ tw16,
                                                                 const float2* __restrict__ tw8, SpecConst sc,
                                                                 FrameBatch fb)
{
    constexpr int G = strip_group<N, C>();
    __shared__ float2 lds[2 * C * lds_len(N)];
    const int L     = xcd_logical(blockIdx.x, gridDim.x);
    const int rest  = L / G;
    const int f     = rest % fb.nframes;
    const int strip = (rest / fb.nframes) * G + L % G;
    if (strip > (N / 2) / C) return;  // padding of the last group (whole block)
    cols_lds2b_strip<N, C>(h0t, z1 + (size_t)f * fb.z1_stride, z2 + (size_t)f * fb.z2_stride, tw16, tw8, sc,
                           fb.t_rev[f], strip, lds);
}

// Two-pass variant of k_cols_lds (cols_lds2_strip): 2C*T threads and half the
// LDS, same block order.
template <int N, int C>
__global__ __launch_bounds__(2 * C * (N / 16)) void k_cols_lds2(const float4* __restrict__ h0,
                                                                float2* __restrict__ z1, float2* __restrict__ z2,
                                                                const float2* __restrict__ tw, SpecConst sc,
                                                                FrameBatch fb)
{
    constexpr int G = strip_group<N, C>();
    __shared__ float2 lds[2 * C * lds_len(N)];
    const int L     = xcd_logical(blockIdx.x, gridDim.x);
    const int rest  = L / G;
    const int f     = rest % fb.nframes;
    const int strip = (rest / fb.nframes) * G + L % G;
    if (strip > (N / 2) / C) return;  // padding of the last group (whole block)
    cols_lds2_strip<N, C>(h0, z1 + (size_t)f * fb.z1_stride, z2 + (size_t)f * fb.z2_stride, tw, sc, fb.t_rev[f],
                          strip, lds);
}

// ------------------------------------------------------------------ fused frame, pass 2

// Row pass.  Grid = (N/2, frames).  Block p handles rows 2p and 2p+1: group 0
// transforms Z1 row 2p, group 1 Z1 row 2p+1, group 2 Z2 row p (whose real and
// imaginary parts are Dz of rows 2p and 2p+1), then applies the inversion step
// (-1)^(x+y) / N^2 (inversion_cs.glsl:35-41).
template <int N, int E>
__global__ __launch_bounds__(3 * (N / E)) void k_rows(const float2* __restrict__ z1, const float2* __restrict__ z2,
                                                      float* __restrict__ height, float* __restrict__ dx,
                                                      float* __restrict__ dz, const float2* __restrict__ tw,
                                                      FrameBatch fb)
{
    constexpr int T = N / E;
    __shared__ float2 lds[3 * lds_len(N)];
    const int f = blockIdx.y;
    z1 += (size_t)f * fb.z1_stride;
    z2 += (size_t)f * fb.z2_stride;
    const size_t o = (size_t)f * fb.out_stride;
    height += o;
    dx += o;
    dz += o;
    const int g = threadIdx.x / T;
    const int j = threadIdx.x % T;
    const int p = blockIdx.x;
    // Z1 rows 2p, 2p+1 form one two-row tile (zidx); Z2 row p is read alone
    const float2* src = g == 2 ? z2 : z1;
    const int     y   = g == 2 ? p : 2 * p + g;
    float2 v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = src[zidx<N>(y, zpos_elem<N, E>(j, r))];
    fft_inverse<N, E>(v, j, lds + g * lds_len(N), tw, BlockSync{});
    constexpr float scale = 1.0f / ((float)N * (float)N);
    const size_t row0 = (size_t)(2 * p) * N, row1 = row0 + N;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int   x = j + r * T;
        const float s = (x & 1) ? -scale : scale;  // sign of row 2p; row 2p+1 has the opposite sign
        if (g == 0) {
            height[row0 + x] = s * v[r].x;
            dx[row0 + x]     = s * v[r].y;
        } else if (g == 1) {
            height[row1 + x] = -s * v[r].x;
            dx[row1 + x]     = -s * v[r].y;
        } else {
            dz[row0 + x] = s * v[r].x;
            dz[row1 + x] = -s * v[r].y;
        }
    }
}

// ------------------------------------------------------------------ fused row pass + normals + Jacobian

// Row pass fused with the normal map (nearest taps) and the Jacobian: each
// group of T = N/E threads walks one band of R rows (rows_nrm_band).  Groups
// lie in one wave when T <= 64 (wave-level ordering), otherwise a block is one
// group.  1-D grid of ceil(bands*frames / groups per block) blocks; after the
// XCD remap, neighbouring bands (which share halo rows) run on the same XCD.
// At least two waves per SIMD: the band walk holds about 260 registers a lane
// when left unbounded, and one wave per SIMD cannot hide its load latency (the
// bound trades that for a few spilled registers; N=1024 frames 18.7 -> 17.6 us).
template <int N, int E, int R, bool RENDER>
__global__ __launch_bounds__(N / E >= 64 ? N / E : 64, 2) void k_rows_nrm(
    const float2* __restrict__ z1, const float2* __restrict__ z2, float* __restrict__ height,
    float* __restrict__ dx, float* __restrict__ dz, float4* __restrict__ normals, float* __restrict__ jac,
    const float2* __restrict__ tw, FrameBatch fb, float lambda, float inv2d, float4* __restrict__ disp,
    float4* __restrict__ rnorm, float dscale)
{
    constexpr int T   = N / E;
    constexpr int GPB = T >= 64 ? 1 : 64 / T;  // groups per block
    constexpr int NB  = N / R;                  // bands per frame
    using Sync = typename std::conditional<(T <= 64), WaveSync, BlockSync>::type;
    __shared__ float2 lds[GPB * lds_len(N)];
    // With one group per block (block = T threads) g is 0: stated at compile
    // time so that the frame and its buffer offsets stay scalar.
    const int g    = GPB == 1 ? 0 : (int)threadIdx.x / T;
    const int j    = GPB == 1 ? (int)threadIdx.x : (int)threadIdx.x % T;
    const int L    = xcd_logical(blockIdx.x, gridDim.x) * GPB + g;
    const int band = L % NB;
    const int f    = L / NB;
    if (f >= fb.nframes) return;  // whole group (a wave part when T <= 64): padding of the last block
    const size_t fo = (size_t)f * fb.out_stride;
    // Z2 rows paired up to N=1024 (N=512 frames -4%, N=1024 -1%), and at N=2048
    // without the render export (247 registers with the upward walk, N=2048
    // frames -0.4%; the render variant would spill).  Odd bands walk upwards
    // from N=1024 (halo tiles shared in time with the neighbours)
    static_assert(R % 2 == 0 && N % R == 0, "bands of an even number of rows");
    constexpr bool PAIR2 = N <= 1024 || (N <= 2048 && !RENDER);
    rows_nrm_band<N, E, Sync, RENDER, PAIR2, (N >= 1024)>(
        z1 + (size_t)f * fb.z1_stride, z2 + (size_t)f * fb.z2_stride, height + fo, dx + fo, dz + fo,
        normals ? normals + fo : nullptr, jac ? jac + fo : nullptr, tw, lds + g * lds_len(N), j, band, R, lambda,
        inv2d, disp ? disp + fo : nullptr, rnorm ? rnorm + fo : nullptr, dscale);
}

// ------------------------------------------------------------------ fused row pass, parallel bands

// Row pass fused with the normal map (nearest taps) and the Jacobian for
// launches of few frames: where the band walk of k_rows_nrm runs its rows one
// after another (long bands, few groups), here one block owns a band of R
// output rows [y0, y0 + R) and transforms all the rows its stencils need at
// once: G1 = R + 2 Z1 rows (y0 - 1 .. y0 + R) and G2 = R/2 + 2 Z2 rows
// (y0/2 - 1 .. y0/2 + R/2), one T-thread group (radix E) each.  The results
// stay in the groups' LDS buffers in natural order; after one block barrier
// every thread takes texels of the band, applies (-1)^(x+y) / N^2
// (inversion_cs.glsl:35-41) to each value it reads and computes the Sobel
// normal (normal_map_cs.glsl:37-53) and the Jacobian of the render
// displacement (grid_tes.glsl:60-62) with the band walk's arithmetic.  1-D
// grid of (N/R) x frames blocks; after the XCD remap neighbouring bands,
// which read the same halo tiles, run on one XCD.
template <int N, int E, int R>
constexpr int par_groups()
{
    return (R + 2) + (R / 2 + 2);
}

template <int N, int E, int R, int W = 1>
__global__ __launch_bounds__((W * par_groups<N, E, R>() * (N / E))) void k_rows_nrm_par(
    const float2* __restrict__ z1, const float2* __restrict__ z2, float* __restrict__ height,
    float* __restrict__ dx, float* __restrict__ dz, float4* __restrict__ normals, float* __restrict__ jac,
    const float2* __restrict__ tw, FrameBatch fb, float lambda, float inv2d)
{
    // W = 2: twice the threads, the extra ones (after the G transform groups)
    // taking part in the stencils only
    constexpr int T = N / E, G1 = R + 2, G = par_groups<N, E, R>(), LEN = lds_len(N), NT = W * G * T;
    static_assert(R % 2 == 0 && N % R == 0, "bands of an even number of rows");
    static_assert(T % 64 == 0 || 64 % T == 0, "transform groups of whole waves or within one wave");
    static_assert(W == 1 || T <= 64, "extra threads only beside wave-local transforms");
    using Sync = typename std::conditional<(T <= 64), WaveSync, BlockSync>::type;
    constexpr float scale = 1.0f / ((float)N * (float)N);
    __shared__ float2 lds[G * LEN];
    const int L    = xcd_logical(blockIdx.x, gridDim.x);
    const int band = L % (N / R);
    const int f    = L / (N / R);
    const int y0   = band * R;
    z1 += (size_t)f * fb.z1_stride;
    z2 += (size_t)f * fb.z2_stride;
    const size_t fo = (size_t)f * fb.out_stride;
    height += fo;
    dx += fo;
    dz += fo;
    if (normals) normals += fo;
    if (jac) jac += fo;

    if (W == 1 || (int)threadIdx.x < G * T) {
        const int g = threadIdx.x / T, j = threadIdx.x % T;
        // group g < G1: Z1 row y0 - 1 + g; else Z2 row y0/2 - 1 + (g - G1) (periodic)
        const bool    is1 = g < G1;
        const float2* src = is1 ? z1 : z2;
        const int     row = is1 ? (y0 - 1 + g) & (N - 1) : ((y0 >> 1) - 1 + (g - G1)) & (N / 2 - 1);
        float2        v[E];
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = src[zidx<N>(row, zpos_elem<N, E>(j, r))];
        float2* buf = lds + g * LEN;
        fft_inverse<N, E>(v, j, buf, tw, Sync{});
        Sync{}();
#pragma unroll
        for (int r = 0; r < E; ++r) buf[lds_pad(j + r * T)] = v[r];
    }
    __syncthreads();

    // Stencils: each thread walks whole columns x of the band, keeping the
    // signed values of x-1, x, x+1 for all band and halo rows in registers
    // (each LDS value is read once per column).  The sign (-1)^(x+y) / N^2 of
    // texel (x, y): rows y0 - 1 + g alternate starting from that of
    // (x, y0 - 1) (y0 is even); columns x +- 1 have the opposite sign of x.
    constexpr int G2 = G - G1;
    for (int x = threadIdx.x; x < N; x += NT) {
        const int   pl = lds_pad((x - 1) & (N - 1)), pm = lds_pad(x), pr = lds_pad((x + 1) & (N - 1));
        const float s0 = ((x + 1) & 1) ? -scale : scale;  // (x, y0 - 1)
        float2 hl[G1], hm[G1], hr[G1];  // (height, Dx) of rows y0 - 1 .. y0 + R at x - 1, x, x + 1
#pragma unroll
        for (int g = 0; g < G1; ++g) {
            const float  sm = (g & 1) ? -s0 : s0;
            const float2 l = lds[g * LEN + pl], m = lds[g * LEN + pm], r = lds[g * LEN + pr];
            hl[g] = make_float2(-sm * l.x, -sm * l.y);
            hm[g] = make_float2(sm * m.x, sm * m.y);
            hr[g] = make_float2(-sm * r.x, -sm * r.y);
        }
        // Dz of rows y0 - 1 .. y0 + R: Z2 row y/2 (buffer G1 + (y+2)/2 - y0/2), part y & 1
        float zl[G1], zm[G1], zr[G1];
#pragma unroll
        for (int q = 0; q < G2; ++q) {
            const float2 l = lds[(G1 + q) * LEN + pl], m = lds[(G1 + q) * LEN + pm], r = lds[(G1 + q) * LEN + pr];
            // Z2 row y0/2 - 1 + q holds rows y0 - 2 + 2q (re) and y0 - 1 + 2q (im), i.e. g = 2q - 1, 2q
#pragma unroll
            for (int part = 0; part < 2; ++part) {
                const int g = 2 * q - 1 + part;
                if (g < 0 || g >= G1) continue;
                const float sm = (g & 1) ? -s0 : s0;
                zl[g] = -sm * (part ? l.y : l.x);
                zm[g] = sm * (part ? m.y : m.x);
                zr[g] = -sm * (part ? r.y : r.x);
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int    g = k + 1;  // grid row y0 + k
            const size_t o = (size_t)(y0 + k) * N + x;
            if (normals) {
                // v_s = h(y-1) + 2 h(y) + h(y+1), d_s = h(y-1) - h(y+1) at x-1, x, x+1;
                // nx = v_s(x-1) - v_s(x+1), nz = d_s(x-1) + 2 d_s(x) + d_s(x+1)
                const float vl  = hl[g - 1].x + 2.0f * hl[g].x + hl[g + 1].x;
                const float vr  = hr[g - 1].x + 2.0f * hr[g].x + hr[g + 1].x;
                const float nx  = vl - vr;
                const float nz  = (hl[g - 1].x - hl[g + 1].x) + 2.0f * (hm[g - 1].x - hm[g + 1].x) +
                                  (hr[g - 1].x - hr[g + 1].x);
                const float inv = rsqrtf(fmaf(nx, nx, fmaf(nz, nz, 1.0f)));
                st_out(&normals[o], make_float4(nx * inv, inv, nz * inv, 1.0f));
            }
            if (jac) {
                const float dxdx = (hr[g].y - hl[g].y) * inv2d;
                const float dzdx = (zr[g] - zl[g]) * inv2d;
                const float dxdz = (hm[g + 1].y - hm[g - 1].y) * inv2d;
                const float dzdz = (zm[g + 1] - zm[g - 1]) * inv2d;
                st_out(&jac[o], (1.0f - lambda * dxdx) * (1.0f - lambda * dzdz) - lambda * lambda * dxdz * dzdx);
            }
            st_out(&height[o], hm[g].x);
            st_out(&dx[o], hm[g].y);
            st_out(&dz[o], zm[g]);
        }
    }
}

// ------------------------------------------------------------------ A7 normals (+ Jacobian)

// normal_map_cs.glsl:24-54 (Sobel on the height, ny = 1, normalize, w = 1),
// and optionally the Jacobian J = (1 - l dDx/dx)(1 - l dDz/dz) - l^2 dDx/dz dDz/dx
// of the render displacement (-l Dx, -l Dz) (grid_tes.glsl:60-62), by periodic
// central differences with spacing L/N.
//
// Tiled: a block owns a 64 x 16 texel tile and first stages height/dx/dz for
// the tile plus a 2-texel periodic halo in LDS with coalesced row loads (the
// bilinear taps reach two texels down/left, the others one), then every tap
// comes from LDS.  1-D grid of (N/64) x (N/16) x frames blocks; frame f reads
// and writes at offset f*stride.  After the XCD remap, tiles of neighbouring
// rows run on the same XCD, so the halo rows are L2 hits.
constexpr int kNrmW = 64, kNrmH = 16, kNrmHalo = 2;
constexpr int kNrmLW = kNrmW + 2 * kNrmHalo, kNrmLH = kNrmH + 2 * kNrmHalo;

// texture() of the staged tile under NEAREST + REPEAT (s[y][x], halo-offset
// coordinates) or LINEAR + REPEAT at a texel corner: 2x2 box over (x-1..x, y-1..y).
__device__ __forceinline__ float stap(const float (*s)[kNrmLW], int y, int x, int mode)
{
    return mode ? 0.25f * (s[y - 1][x - 1] + s[y - 1][x] + s[y][x - 1] + s[y][x]) : s[y][x];
}

__global__ __launch_bounds__(256) void k_normals(const float* __restrict__ height, const float* __restrict__ dx,
                                                 const float* __restrict__ dz, float4* __restrict__ normals,
                                                 float* __restrict__ jac, int N, int mode, float lambda,
                                                 float inv2d, size_t stride, float4* __restrict__ disp,
                                                 float4* __restrict__ rnorm, float dscale)
{
    __shared__ float sh[kNrmLH][kNrmLW], sx[kNrmLH][kNrmLW], sz[kNrmLH][kNrmLW];
    const int tiles_x = N / kNrmW, tiles_y = N / kNrmH;
    const int L  = xcd_logical(blockIdx.x, gridDim.x);
    const int bx = L % tiles_x;
    const int by = (L / tiles_x) % tiles_y;
    const int f  = L / (tiles_x * tiles_y);
    const size_t fo = (size_t)f * stride;
    height += fo;
    dx += fo;
    dz += fo;
    if (normals) normals += fo;
    if (jac) jac += fo;
    if (disp) disp += fo;
    if (rnorm) rnorm += fo;
    const bool need_h = normals || disp || rnorm, need_xz = jac || disp || rnorm;
    const int x0 = bx * kNrmW, y0 = by * kNrmH;
    for (int k = threadIdx.x; k < kNrmLH * kNrmLW; k += 256) {
        const int ly = k / kNrmLW, lx = k % kNrmLW;
        const size_t gi = (size_t)((y0 + ly - kNrmHalo) & (N - 1)) * N + ((x0 + lx - kNrmHalo) & (N - 1));
        if (need_h) sh[ly][lx] = height[gi];
        if (need_xz) {
            sx[ly][lx] = dx[gi];
            sz[ly][lx] = dz[gi];
        }
    }
    __syncthreads();
    const int cx = (threadIdx.x & 63) + kNrmHalo;
#pragma unroll
    for (int k = 0; k < kNrmH / 4; ++k) {
        const int ly = (threadIdx.x >> 6) + 4 * k;
        const int cy = ly + kNrmHalo;
        const size_t i = (size_t)(y0 + ly) * N + x0 + (threadIdx.x & 63);
        if (normals) {
            // z0..z7 at (x-1,y-1) (x,y-1) (x+1,y-1) (x-1,y) (x+1,y) (x-1,y+1) (x,y+1) (x+1,y+1)
            const float z0 = stap(sh, cy - 1, cx - 1, mode), z1 = stap(sh, cy - 1, cx, mode);
            const float z2 = stap(sh, cy - 1, cx + 1, mode), z3 = stap(sh, cy, cx - 1, mode);
            const float z4 = stap(sh, cy, cx + 1, mode), z5 = stap(sh, cy + 1, cx - 1, mode);
            const float z6 = stap(sh, cy + 1, cx, mode), z7 = stap(sh, cy + 1, cx + 1, mode);
            const float nz  = z0 + 2.0f * z1 + z2 - z5 - 2.0f * z6 - z7;
            const float nx  = z0 + 2.0f * z3 + z5 - z2 - 2.0f * z4 - z7;
            const float inv = rsqrtf(fmaf(nx, nx, fmaf(nz, nz, 1.0f)));
            normals[i] = make_float4(nx * inv, inv, nz * inv, 1.0f);
        }
        if (need_xz) {
            const float dxdx = (sx[cy][cx + 1] - sx[cy][cx - 1]) * inv2d;
            const float dxdz = (sx[cy + 1][cx] - sx[cy - 1][cx]) * inv2d;
            const float dzdx = (sz[cy][cx + 1] - sz[cy][cx - 1]) * inv2d;
            const float dzdz = (sz[cy + 1][cx] - sz[cy - 1][cx]) * inv2d;
            const float J    = (1.0f - lambda * dxdx) * (1.0f - lambda * dzdz) - lambda * lambda * dxdz * dzdx;
            if (jac) jac[i] = J;
            if (disp) disp[i] = make_float4(-lambda * sx[cy][cx], dscale * sh[cy][cx], -lambda * sz[cy][cx], J);
            if (rnorm) {
                // n = dP/dv x dP/du of P = (u - l Dx, s h, v - l Dz); its y component is J
                const float dhdx = (sh[cy][cx + 1] - sh[cy][cx - 1]) * inv2d;
                const float dhdz = (sh[cy + 1][cx] - sh[cy - 1][cx]) * inv2d;
                const float nx   = -dscale * (lambda * dhdz * dzdx + (1.0f - lambda * dzdz) * dhdx);
                const float nz   = -dscale * (lambda * dxdz * dhdx + dhdz * (1.0f - lambda * dxdx));
                const float inv  = rsqrtf(fmaf(nx, nx, fmaf(J, J, nz * nz)));
                rnorm[i] = make_float4(nx * inv, J * inv, nz * inv, 1.0f);
            }
        }
    }
}

// ------------------------------------------------------------------ A5+A6 pass API (general complex input)

// Row transforms of a general complex N x N field (the horizontal butterflies).
template <int N, int E>
__global__ __launch_bounds__(N / E) void k_c2c_rows(const float2* __restrict__ src, float2* __restrict__ dst,
                                                    const float2* __restrict__ tw)
{
    constexpr int T = N / E;
    __shared__ float2 lds[lds_len(N)];
    const int j = threadIdx.x;
    const size_t row = (size_t)blockIdx.x * N;
    float2 v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = src[row + j + r * T];
    fft_inverse<N, E>(v, j, lds, tw, BlockSync{});
#pragma unroll
    for (int r = 0; r < E; ++r) dst[row + j + r * T] = v[r];
}

// Column transforms + inversion: out = (-1)^(x+y) Re(.) / N^2.
template <int N, int E, int C>
__global__ __launch_bounds__(C*(N / E)) void k_c2c_cols_inv(const float2* __restrict__ src, float* __restrict__ out,
                                                            const float2* __restrict__ tw)
{
    constexpr int T = N / E;
    __shared__ float2 lds[C * lds_len(N)];
    const int c = threadIdx.x % C;
    const int j = threadIdx.x / C;
    const int x = blockIdx.x * C + c;
    float2 v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = src[(size_t)(j + r * T) * N + x];
    fft_inverse<N, E>(v, j, lds + c * lds_len(N), tw, BlockSync{});
    constexpr float scale = 1.0f / ((float)N * (float)N);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int y = j + r * T;
        out[(size_t)y * N + x] = (((x + y) & 1) ? -scale : scale) * v[r].x;
    }
}

// ------------------------------------------------------------------ launchers

// Column-pass schedules (radix E, columns per block C); the autotuner picks one
// per N.  Row pass: radix E, three transforms per block.
// kind 0: k_cols (register-resident, block-synchronised transforms);
// kind 1: k_cols_lds (LDS-staged, one wave per transform, radix 16 only);
// kind 2: k_cols_lds2 (the same in two passes through half the LDS);
// kind 3: k_cols_lds with twice the threads (folds and stores split finer);
// kind 4: k_cols_ldsb (LDS-staged, block-synchronised transforms, N >= 2048);
// kind 5: k_cols_lds2b (the same in two passes through half the LDS, N >= 2048);
// kind 6: k_cols_lds_f2 (k_cols_lds over pairs of frames).
struct ColsVariant {
    int E, C, kind;
};
constexpr ColsVariant kColsVariants[kNumColsVariants] = {{16, 1, 0}, {16, 2, 0}, {16, 4, 0}, {16, 8, 0}, {8, 2, 0},
                                                         {8, 4, 0},  {8, 8, 0},  {16, 1, 1}, {16, 2, 1}, {16, 4, 1},
                                                         {16, 2, 2}, {16, 4, 2}, {16, 8, 2}, {16, 1, 3}, {16, 2, 3},
                                                         {16, 1, 4}, {16, 2, 4}, {16, 2, 5}, {16, 4, 5},
                                                         {16, 4, 6}, {16, 8, 6}};
// Row-pass schedules: band 0 = k_rows (radix E) followed by k_normals;
// band R > 0 = k_rows_nrm with bands of R rows (radix E), normals fused;
// par = k_rows_nrm_par (bands of R rows whose rows are transformed at once).
struct RowsVariant {
    int  E, band;
    bool par;
    int  W;  // parallel bands: threads per transform-group thread
};
constexpr RowsVariant kRowsVariants[kNumRowsVariants] = {{16, 0, false, 1}, {8, 0, false, 1},  {8, 4, false, 1},
                                                         {8, 8, false, 1},  {8, 16, false, 1}, {8, 32, false, 1},
                                                         {16, 4, true, 1},  {16, 2, true, 1},  {16, 2, true, 2}};

template <int N, int E, int C, int K>
constexpr bool cols_fits()
{
    if constexpr (K == 0)
        return N >= 2 * E && C * (N / E) <= 1024 && C * lds_len(N) * 8 <= 160 * 1024 && C <= N / 2;
    else if constexpr (K == 1 || K == 3 || K == 6)
        return E == 16 && N / 16 <= 64 && (K == 1 ? 1 : 2) * 3 * C * (N / 16) <= 1024 &&
               (K == 6 ? 2 : 1) * 3 * C * strip_len(N) * 8 <= 160 * 1024 && C <= N / 2;
    else if constexpr (K == 4)
        return E == 16 && N / 16 > 64 && 3 * C * (N / 16) <= 1024 && 3 * C * strip_len(N) * 8 <= 160 * 1024;
    else if constexpr (K == 5)
        return E == 16 && N / 16 > 64 && 2 * C * (N / 16) <= 1024 && 2 * C * lds_len(N) * 8 <= 160 * 1024;
    else
        return E == 16 && N >= 64 && N / 16 <= 64 && 2 * C * (N / 16) <= 1024 && 2 * C * lds_len(N) * 8 <= 160 * 1024 &&
               C <= N / 2 && (C * N) % (2 * C * (N / 16)) == 0;
}
template <int N, int E, int B, bool PAR = false, int W = 1>
constexpr bool rows_fits()
{
    if constexpr (PAR) {
        constexpr int T = N / E, G = par_groups<N, E, B>();
        return N >= 2 * E && N % B == 0 && (T % 64 == 0 || 64 % T == 0) && (W == 1 || T <= 64) &&
               W * G * T <= 1024 && G * lds_len(N) * 8 <= 160 * 1024;
    } else if constexpr (B == 0) {
        return N >= 2 * E && 3 * (N / E) <= 1024;
    } else {
        return N >= 2 * E && N % B == 0 && N / E <= 1024;
    }
}

constexpr int radix_slot(int E) { return E == 16 ? 2 : E == 8 ? 1 : 0; }

template <int N, int V>
static hipError_t launch_cols_v(const FrameArgs& a, hipStream_t s)
{
    constexpr int E = kColsVariants[V].E, C = kColsVariants[V].C, K = kColsVariants[V].kind;
    if constexpr (!cols_fits<N, E, C, K>()) {
        return hipErrorInvalidValue;
    } else if constexpr (K == 1 || K == 3) {
        constexpr int W      = K == 3 ? 2 : 1;
        constexpr int G      = strip_group<N, C>();
        constexpr int strips = lds_strips<N, C>();
        constexpr int groups = (strips + G - 1) / G;
        hipLaunchKernelGGL((k_cols_lds<N, C, W>), dim3(groups * G * a.nframes), dim3(W * 3 * C * (N / 16)), 0, s,
                           a.h0, a.z1, a.z2, a.tw[radix_slot(16)], a.sc, a.fb);
        return hipGetLastError();
    } else if constexpr (K == 6) {
        constexpr int G      = strip_group<N, C>();
        constexpr int strips = lds_strips<N, C>();
        constexpr int groups = (strips + G - 1) / G;
        hipLaunchKernelGGL((k_cols_lds_f2<N, C>), dim3(groups * G * ((a.nframes + 1) / 2)), dim3(6 * C * (N / 16)), 0,
                           s, a.h0, a.z1, a.z2, a.tw[radix_slot(16)], a.sc, a.fb);
        return hipGetLastError();
    } else if constexpr (K == 4) {
        constexpr int G      = strip_group<N, C>();
        constexpr int strips = lds_strips<N, C>();
        constexpr int groups = (strips + G - 1) / G;
        if (!a.h0t) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_cols_ldsb<N, C>), dim3(groups * G * a.nframes), dim3(3 * C * (N / 16)), 0, s, a.h0t,
                           a.z1, a.z2, a.tw[radix_slot(16)], a.sc, a.fb);
        return hipGetLastError();
    } else if constexpr (K == 5) {
        constexpr int G      = strip_group<N, C>();
        constexpr int strips = (N / 2) / C + 1;
        constexpr int groups = (strips + G - 1) / G;
        if (!a.h0t) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_cols_lds2b<N, C>), dim3(groups * G * a.nframes), dim3(2 * C * (N / 16)), 0, s, a.h0t,
                           a.z1, a.z2, a.tw[radix_slot(16)], a.tw[radix_slot(8)], a.sc, a.fb);
        return hipGetLastError();
    } else if constexpr (K == 2) {
        constexpr int G      = strip_group<N, C>();
        constexpr int strips = (N / 2) / C + 1;
        constexpr int groups = (strips + G - 1) / G;
        hipLaunchKernelGGL((k_cols_lds2<N, C>), dim3(groups * G * a.nframes), dim3(2 * C * (N / 16)), 0, s, a.h0,
                           a.z1, a.z2, a.tw[radix_slot(16)], a.sc, a.fb);
        return hipGetLastError();
    } else {
        constexpr int G       = strip_group<N, C>();
        constexpr int strips  = (N / 2) / C + 1;  // half plane tx in [0, N/2]
        constexpr int groups  = (strips + G - 1) / G;
        hipLaunchKernelGGL((k_cols<N, E, C>), dim3(groups * G * a.nframes), dim3(C * (N / E)), 0, s, a.h0, a.z1,
                           a.z2, a.tw[radix_slot(E)], a.sc, a.fb);
        return hipGetLastError();
    }
}

template <int N, int V>
static hipError_t launch_rows_v(const FrameArgs& a, hipStream_t s)
{
    constexpr int  E = kRowsVariants[V].E, B = kRowsVariants[V].band, W = kRowsVariants[V].W;
    constexpr bool P = kRowsVariants[V].par;
    if constexpr (!rows_fits<N, E, B, P, W>()) {
        return hipErrorInvalidValue;
    } else if constexpr (P) {
        constexpr int T = N / E;
        hipLaunchKernelGGL((k_rows_nrm_par<N, E, B, W>), dim3((N / B) * a.nframes),
                           dim3(W * par_groups<N, E, B>() * T), 0, s, a.z1, a.z2, a.height, a.dx, a.dz, a.normals,
                           a.jacobian, a.tw[radix_slot(E)], a.fb, a.lambda, a.inv2d);
        return hipGetLastError();
    } else if constexpr (B == 0) {
        hipLaunchKernelGGL((k_rows<N, E>), dim3(N / 2, a.nframes), dim3(3 * (N / E)), 0, s, a.z1, a.z2, a.height,
                           a.dx, a.dz, a.tw[radix_slot(E)], a.fb);
        return hipGetLastError();
    } else {
        constexpr int T = N / E, GPB = T >= 64 ? 1 : 64 / T;
        const int groups = (N / B) * a.nframes;
        if (a.disp || a.rnorm)
            hipLaunchKernelGGL((k_rows_nrm<N, E, B, true>), dim3((groups + GPB - 1) / GPB), dim3(GPB * T), 0, s,
                               a.z1, a.z2, a.height, a.dx, a.dz, a.normals, a.jacobian, a.tw[radix_slot(E)], a.fb,
                               a.lambda, a.inv2d, a.disp, a.rnorm, a.dscale);
        else
            hipLaunchKernelGGL((k_rows_nrm<N, E, B, false>), dim3((groups + GPB - 1) / GPB), dim3(GPB * T), 0, s,
                               a.z1, a.z2, a.height, a.dx, a.dz, a.normals, a.jacobian, a.tw[radix_slot(E)], a.fb,
                               a.lambda, a.inv2d, nullptr, nullptr, 0.0f);
        return hipGetLastError();
    }
}

template <int N>
static hipError_t launch_frame_n(const FrameArgs& a, int stages, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if (stages & 1) {
        switch (a.cols_variant) {
        case 0: e = launch_cols_v<N, 0>(a, s); break;
        case 1: e = launch_cols_v<N, 1>(a, s); break;
        case 2: e = launch_cols_v<N, 2>(a, s); break;
        case 3: e = launch_cols_v<N, 3>(a, s); break;
        case 4: e = launch_cols_v<N, 4>(a, s); break;
        case 5: e = launch_cols_v<N, 5>(a, s); break;
        case 6: e = launch_cols_v<N, 6>(a, s); break;
        case 7: e = launch_cols_v<N, 7>(a, s); break;
        case 8: e = launch_cols_v<N, 8>(a, s); break;
        case 9: e = launch_cols_v<N, 9>(a, s); break;
        case 10: e = launch_cols_v<N, 10>(a, s); break;
        case 11: e = launch_cols_v<N, 11>(a, s); break;
        case 12: e = launch_cols_v<N, 12>(a, s); break;
        case 13: e = launch_cols_v<N, 13>(a, s); break;
        case 14: e = launch_cols_v<N, 14>(a, s); break;
        case 15: e = launch_cols_v<N, 15>(a, s); break;
        case 16: e = launch_cols_v<N, 16>(a, s); break;
        case 17: e = launch_cols_v<N, 17>(a, s); break;
        case 18: e = launch_cols_v<N, 18>(a, s); break;
        case 19: e = launch_cols_v<N, 19>(a, s); break;
        case 20: e = launch_cols_v<N, 20>(a, s); break;
        default: e = hipErrorInvalidValue;
        }
        if (e != hipSuccess) return e;
    }
    // The fused row variants cover the normal map and Jacobian (nearest taps
    // only); with bilinear normals the unfused radix-16 row pass is used.  The
    // parallel-band variants write no render export: with one requested the
    // unfused radix-8 pass and the normal-map kernel write it.
    int rv = a.rows_variant;
    if (rv >= 0 && rv < kNumRowsVariants && kRowsVariants[rv].band > 0 && a.normal_mode != 0) rv = 0;
    if (rv >= 0 && rv < kNumRowsVariants && kRowsVariants[rv].par && (a.disp || a.rnorm)) rv = 1;
    const bool fused = rv >= 0 && rv < kNumRowsVariants && kRowsVariants[rv].band > 0;
    if (stages & 2) {
        switch (rv) {
        case 0: e = launch_rows_v<N, 0>(a, s); break;
        case 1: e = launch_rows_v<N, 1>(a, s); break;
        case 2: e = launch_rows_v<N, 2>(a, s); break;
        case 3: e = launch_rows_v<N, 3>(a, s); break;
        case 4: e = launch_rows_v<N, 4>(a, s); break;
        case 5: e = launch_rows_v<N, 5>(a, s); break;
        case 6: e = launch_rows_v<N, 6>(a, s); break;
        case 7: e = launch_rows_v<N, 7>(a, s); break;
        case 8: e = launch_rows_v<N, 8>(a, s); break;
        default: e = hipErrorInvalidValue;
        }
        if (e != hipSuccess) return e;
    }
    if ((stages & 4) && !fused && (a.normals || a.jacobian || a.disp || a.rnorm)) {
        hipLaunchKernelGGL(k_normals, dim3((N / kNrmW) * (N / kNrmH) * a.nframes), dim3(256), 0, s, a.height, a.dx, a.dz,
                           a.normals, a.jacobian, N, a.normal_mode, a.lambda, a.inv2d, a.fb.out_stride, a.disp,
                           a.rnorm, a.dscale);
        e = hipGetLastError();
    }
    return e;
}

template <int N>
static bool variant_ok_n(int cols_v, int rows_v)
{
    constexpr bool c[kNumColsVariants] = {
#define OCN_FITS(v) cols_fits<N, kColsVariants[v].E, kColsVariants[v].C, kColsVariants[v].kind>()
        OCN_FITS(0), OCN_FITS(1), OCN_FITS(2), OCN_FITS(3), OCN_FITS(4),
        OCN_FITS(5), OCN_FITS(6), OCN_FITS(7), OCN_FITS(8), OCN_FITS(9),
        OCN_FITS(10), OCN_FITS(11), OCN_FITS(12), OCN_FITS(13), OCN_FITS(14), OCN_FITS(15), OCN_FITS(16),
        OCN_FITS(17), OCN_FITS(18), OCN_FITS(19), OCN_FITS(20)};
#undef OCN_FITS
#define OCN_RFITS(v) \
    rows_fits<N, kRowsVariants[v].E, kRowsVariants[v].band, kRowsVariants[v].par, kRowsVariants[v].W>()
    constexpr bool r[kNumRowsVariants] = {OCN_RFITS(0), OCN_RFITS(1), OCN_RFITS(2), OCN_RFITS(3), OCN_RFITS(4),
                                          OCN_RFITS(5), OCN_RFITS(6), OCN_RFITS(7), OCN_RFITS(8)};
#undef OCN_RFITS
    return cols_v >= 0 && cols_v < kNumColsVariants && rows_v >= 0 && rows_v < kNumRowsVariants && c[cols_v] &&
           r[rows_v];
}

template <int N>
static hipError_t launch_ifft2_n(const float2* src, float2* tmp, float* out, const float2* const* tw, hipStream_t s)
{
    constexpr int E = 16, T = N / E, C = (512 / T) < 16 ? (512 / T) : 16;
    hipLaunchKernelGGL((k_c2c_rows<N, E>), dim3(N), dim3(T), 0, s, src, tmp, tw[radix_slot(E)]);
    hipLaunchKernelGGL((k_c2c_cols_inv<N, E, C>), dim3(N / C), dim3(C * T), 0, s, tmp, out, tw[radix_slot(E)]);
    return hipGetLastError();
}

#define OCN_DISPATCH_N(N, CALL, BAD)                \
    switch (N) {                                    \
    case 64: return CALL(64);                       \
    case 128: return CALL(128);                     \
    case 256: return CALL(256);                     \
    case 512: return CALL(512);                     \
    case 1024: return CALL(1024);                   \
    case 2048: return CALL(2048);                   \
    case 4096: return CALL(4096);                   \
    default: return BAD;                            \
    }

hipError_t launch_frame(int N, const FrameArgs& a, int stages, hipStream_t s)
{
    if (a.nframes < 1 || a.nframes > kMaxFrames) return hipErrorInvalidValue;
#define CALL(n) launch_frame_n<n>(a, stages, s)
    OCN_DISPATCH_N(N, CALL, hipErrorInvalidValue)
#undef CALL
}

bool schedule_valid(int N, int cols_variant, int rows_variant)
{
#define CALL(n) variant_ok_n<n>(cols_variant, rows_variant)
    OCN_DISPATCH_N(N, CALL, false)
#undef CALL
}

void cols_variant_shape(int v, int* E, int* C)
{
    *E = kColsVariants[v].E;
    *C = kColsVariants[v].C;
}

int rows_variant_radix(int v) { return kRowsVariants[v].E; }

int rows_variant_band(int v) { return kRowsVariants[v].band; }

bool rows_variant_parallel(int v) { return kRowsVariants[v].par; }

hipError_t launch_ifft2(int N, const float2* src, float2* tmp, float* out, const float2* const* tw, hipStream_t s)
{
#define CALL(n) launch_ifft2_n<n>(src, tmp, out, tw, s)
    OCN_DISPATCH_N(N, CALL, hipErrorInvalidValue)
#undef CALL
}

// ------------------------------------------------------------------ static scratch

// Z1 and Z2 of ocn_update_frames_noalloc, for callers that cannot allocate
// device memory.  Each GPU the library runs on holds its own copy of this
// array until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code the name is a host-side stand-in, not a device
// address): static_scratch_resolve looks up the device address once per
// device with hipGetSymbolAddress, and the frame launches take that pointer.
// Nothing in it needs clearing: the column pass writes every element of Z1
// and Z2 before the row pass of the same launch reads them.
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
std::mutex g_scratch_mu;
void*      g_scratch_addr[kMaxScratchDevices] = {};
}  // namespace

hipError_t static_scratch_resolve(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_addr[device]) return hipSuccess;
    void*      p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = p;
    return hipSuccess;
}

void* static_scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_addr[device];
}

hipError_t launch_h0(int N, float4* h0, float4* h0t, const uint8_t* noise, const SpecConst& sc, hipStream_t s)
{
    dim3 grid((N + 255) / 256, N);
    if (noise) hipLaunchKernelGGL(k_tilde_h0_k_noise, grid, dim3(256), 0, s, h0, noise, sc);
    else hipLaunchKernelGGL(k_tilde_h0_k, grid, dim3(256), 0, s, h0, sc);
    if (h0t) hipLaunchKernelGGL(k_transpose_h0, dim3(N / 32, N / 32), dim3(256), 0, s, h0, h0t, N);
    return hipGetLastError();
}

hipError_t launch_spectrum_t(int N, const float4* h0, float2* hy, float2* hx, float2* hz, const SpecConst& sc,
                             double t_rev, hipStream_t s)
{
    hipLaunchKernelGGL(k_tilde_h0_t, dim3((N + 255) / 256, N), dim3(256), 0, s, h0, hy, hx, hz, sc, t_rev);
    return hipGetLastError();
}

hipError_t launch_normals(int N, const float* height, const float* dx, const float* dz, float4* normals,
                          float* jacobian, int mode, float lambda, float inv2d, hipStream_t s)
{
    hipLaunchKernelGGL(k_normals, dim3((N / kNrmW) * (N / kNrmH)), dim3(256), 0, s, height, dx, dz, normals, jacobian, N, mode,
                       lambda, inv2d, (size_t)0, (float4*)nullptr, (float4*)nullptr, 0.0f);
    return hipGetLastError();
}

}  // namespace ocn
