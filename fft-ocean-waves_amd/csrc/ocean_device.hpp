// ocean_device.hpp -- device functions shared by the single-ocean kernels
// (ocean_kernels.hip) and the batched kernels (ocean_batch.hip): noise and the
// Phillips spectrum, the dispersion phase, the Hermitian fold, the XCD-aware
// block order, the LDS-staged column transform of one strip and the fused
// row + normal-map + Jacobian walk of one band.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "fft_core.hpp"
#include "ocean_kernels.h"

namespace ocn {

// ------------------------------------------------------------------ output stores

// The frame's outputs are written once and never read back by these kernels:
// non-temporal stores keep them from displacing the intermediates (still to be
// read by other blocks of the row pass) in L2 and the Infinity Cache.
__device__ __forceinline__ void st_out(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_out(float4* p, float4 v)
{
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    const f32x4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<f32x4*>(p));
}

// ------------------------------------------------------------------ noise

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k.x += 0x9E3779B9u; k.y += 0xBB67AE85u; }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}

// (m + 1/2) * 2^-23 from the top 23 bits: exact in FP32.
__device__ __forceinline__ float u01(uint32_t x) { return ((float)(x >> 9) + 0.5f) * (1.0f / 8388608.0f); }

// gauss_rnd (tilde_h0_k_cs.glsl:51-68): clamp to [0.001, 1], two Box-Muller pairs.
__device__ __forceinline__ float4 box_muller(float n0, float n1, float n2, float n3)
{
    n0 = fminf(fmaxf(n0, 0.001f), 1.0f);
    n1 = fminf(fmaxf(n1, 0.001f), 1.0f);
    n2 = fminf(fmaxf(n2, 0.001f), 1.0f);
    n3 = fminf(fmaxf(n3, 0.001f), 1.0f);
    float s0, c0, s1, c1;
    sincospif(2.0f * n0, &s0, &c0);
    sincospif(2.0f * n2, &s1, &c1);
    const float r0 = sqrtf(-2.0f * logf(n1));
    const float r1 = sqrtf(-2.0f * logf(n3));
    return make_float4(r0 * c0, r0 * s0, r1 * c1, r1 * s1);
}

// sqrt(P(k))/sqrt(2) before the clamp (tilde_h0_k_cs.glsl:35-45, 76-88).
// k_hat = k / |k|_clamped: equal to normalize(k) for k != 0, and 0 at DC where
// normalize(0) is undefined.  pow(d, 2.0) is d*d.
__device__ __forceinline__ float phillips_amp_raw(const SpecConst& sc, int tx, int ty)
{
    const float kx = sc.kscale * (float)(tx - sc.N / 2);
    const float ky = sc.kscale * (float)(ty - sc.N / 2);
    float kmag = sqrtf(kx * kx + ky * ky);
    if (kmag < 0.00001f) kmag = 0.00001f;
    const float k2  = kmag * kmag;
    const float sup = expf(-k2 * sc.l2);
    const float d   = (kx * sc.wx + ky * sc.wy) / kmag;
    const float P   = sc.A * expf(-1.0f / (k2 * sc.Lp2)) * (d * d) * sup / (k2 * k2);
    return sqrtf(P) * 0.70710678118654752f;
}

// clamp(sqrt(P(k))/sqrt(2), -4000, 4000) (tilde_h0_k_cs.glsl:87-88); P >= 0.
__device__ __forceinline__ float phillips_amp(const SpecConst& sc, int tx, int ty)
{
    return fminf(phillips_amp_raw(sc, tx, ty), 4000.0f);
}

// ------------------------------------------------------------------ A4 time evolution

// Dispersion phase w t / (2 pi) with w = sqrt(g |k|_clamped) (tilde_h0_t_cs.glsl:77-79),
// as a fraction of a turn in [-1/2, 1/2].  w t reaches hundreds of radians,
// where FP32 rounding of w alone would cost ~1e-5 rad, so w is refined in FP64:
// an FP32 estimate w0 plus one Newton step on w^4 = g^2 |k|^2 (|k|^2 is exact
// from the integer wave numbers), then the product with t / (2 pi) is reduced
// to one turn before the FP32 sincos.  m = nx^2 + ny^2.  kinv receives
// kscale / |k|_clamped, the factor of the choppy spectra (-i k/|k| = -i n kinv).
// Two reciprocal square roots give the FP32 parts: 1/sqrt(m) is kinv itself,
// and s = 1/sqrt(g |k|) gives w0 = g |k| s and the Newton step's 1/w0^3 = s^3.
__device__ __forceinline__ double omega_fp64(const SpecConst& sc, int m, float& kinv)
{
    const float mf = (float)m;
    const float r  = __builtin_amdgcn_rsqf(mf);  // +inf at m = 0
    const float kr = sc.kscale * (mf * r);       // |k|; NaN at m = 0
    const bool  lo = !(kr >= 0.00001f);          // |k| below the 1e-5 clamp (or DC)
    const float kf = lo ? 0.00001f : kr;
    kinv = lo ? sc.kscale * 100000.0f : r;
    const float  gk = sc.g * kf;
    const float  s  = __builtin_amdgcn_rsqf(gk);
    const float  w0 = gk * s;
    const double Q  = fmax(sc.k2_d * (double)m, 1e-10) * sc.g2_d;  // g^2 |k|_clamped^2
    const double w  = (double)w0;
    const double w2 = w * w;
    return w - (w2 * w2 - Q) * (double)(0.25f * s * s * s);
}

// w t / (2 pi) reduced to [-1/2, 1/2] turns, w from omega_fp64.
__device__ __forceinline__ float turns(double w, double t_rev)
{
    const double rev = w * t_rev;
    return (float)(rev - rint(rev));
}

__device__ __forceinline__ float phase_turns(const SpecConst& sc, int m, double t_rev, float& kinv)
{
    return turns(omega_fp64(sc, m, kinv), t_rev);
}

// sin and cos of 2 pi fr for fr in [-1/2, 1/2]: reduction by quarter turns
// q = rint(4 fr) (the remainder r = fr - q/4 is exact), Taylor polynomials of
// sin and cos on |2 pi r| <= pi/4 to degree 9 and 10 (truncation < 2e-9),
// then rotation by q quarter turns.
__device__ __forceinline__ void sincos_turns(float fr, float& s, float& c)
{
    const float q  = rintf(4.0f * fr);
    const float th = (fr - 0.25f * q) * 6.28318530717958647692f;
    const float t2 = th * th;
    const float sp = th * fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, 2.75573192239858907e-6f, -1.98412698412698413e-4f),
                                                 8.33333333333333333e-3f), -1.66666666666666667e-1f), 1.0f);
    const float cp = fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, -2.75573192239858907e-7f, 2.48015873015873016e-5f),
                                                       -1.38888888888888889e-3f), 4.16666666666666667e-2f), -0.5f),
                          1.0f);
    // quarter turn qi: (s, c) = (sp, cp), (cp, -sp), (-sp, -cp), (-cp, sp),
    // as a select and two sign-bit flips (no divergent branches)
    const int   qi = (int)q & 3;
    const float a  = (qi & 1) ? cp : sp;
    const float b  = (qi & 1) ? sp : cp;
    s = __int_as_float(__float_as_int(a) ^ ((qi & 2) << 30));
    c = __int_as_float(__float_as_int(b) ^ (((qi + 1) & 2) << 30));
}

// a e^{i th} + C(b) e^{-i th} for h = (a, b) = (h0k, h0mk) (tilde_h0_t_cs.glsl:94-110).
__device__ __forceinline__ float2 evolve_amp(const SpecConst& sc, float4 h, float s, float c)
{
    const float bi = sc.conj_mode ? -h.w : h.w;
    return make_float2(fmaf(h.x, c, -h.y * s) + fmaf(h.z, c, bi * s), fmaf(h.x, s, h.y * c) + fmaf(bi, c, -h.z * s));
}

// -i m * H  (tilde_h0_t_cs.glsl:113-126)
__device__ __forceinline__ float2 choppy(float m, float2 h) { return make_float2(m * h.y, -m * h.x); }

// Hermitian parts (F(t) + conj F(t'))/2 of the three spectra at texel t and its
// mirror t' = -t mod N.  Both share |k| (n' = -n except on the Nyquist row or
// column, where n' = n = -N/2), hence one phase and one sincos for the pair.
// p = h0 at t, q = h0 at t'; (nx, ny) and (qx, qy) are their centred indices.
// pair_fold_at takes the phase fr (turns) and the choppy factor kinv of the
// pair, so that frames of one texel pair share the dispersion evaluation.
// pair_fold_sc takes its sin and cos, for pairs that also share the phase.
__device__ __forceinline__ void pair_fold_sc(const SpecConst& sc, float4 p, float4 q, int nx, int ny, int qx, int qy,
                                             float kinv, float s, float c, float2& ay, float2& ax, float2& az)
{
    const float2 H = evolve_amp(sc, p, s, c);
    const float2 G = evolve_amp(sc, q, s, c);
    const float mx = (float)nx * kinv, mz = (float)ny * kinv;
    const float gx = (float)qx * kinv, gz = (float)qy * kinv;
    ay = make_float2(0.5f * (H.x + G.x), 0.5f * (H.y - G.y));
    ax = make_float2(0.5f * (mx * H.y + gx * G.y), 0.5f * (gx * G.x - mx * H.x));
    az = make_float2(0.5f * (mz * H.y + gz * G.y), 0.5f * (gz * G.x - mz * H.x));
}

__device__ __forceinline__ void pair_fold_at(const SpecConst& sc, float4 p, float4 q, int nx, int ny, int qx, int qy,
                                             float kinv, float fr, float2& ay, float2& ax, float2& az)
{
    float s, c;
    sincos_turns(fr, s, c);
    const float2 H = evolve_amp(sc, p, s, c);
    const float2 G = evolve_amp(sc, q, s, c);
    const float mx = (float)nx * kinv, mz = (float)ny * kinv;
    const float gx = (float)qx * kinv, gz = (float)qy * kinv;
    ay = make_float2(0.5f * (H.x + G.x), 0.5f * (H.y - G.y));
    ax = make_float2(0.5f * (mx * H.y + gx * G.y), 0.5f * (gx * G.x - mx * H.x));
    az = make_float2(0.5f * (mz * H.y + gz * G.y), 0.5f * (gz * G.x - mz * H.x));
}

__device__ __forceinline__ void pair_fold(const SpecConst& sc, float4 p, float4 q, int nx, int ny, int qx, int qy,
                                          double t_rev, float2& ay, float2& ax, float2& az)
{
    float kinv, s, c;
    sincos_turns(phase_turns(sc, nx * nx + ny * ny, t_rev, kinv), s, c);
    const float2 H = evolve_amp(sc, p, s, c);
    const float2 G = evolve_amp(sc, q, s, c);
    const float mx = (float)nx * kinv, mz = (float)ny * kinv;
    const float gx = (float)qx * kinv, gz = (float)qy * kinv;
    ay = make_float2(0.5f * (H.x + G.x), 0.5f * (H.y - G.y));
    // -i m H = (m H.y, -m H.x);  conj(-i g G) = (g G.y, g G.x)
    ax = make_float2(0.5f * (mx * H.y + gx * G.y), 0.5f * (gx * G.x - mx * H.x));
    az = make_float2(0.5f * (mz * H.y + gz * G.y), 0.5f * (gz * G.x - mz * H.x));
}

// ------------------------------------------------------------------ intermediate column order

// Position of column x in a row of the column-pass intermediates (Z1, Z2):
// columns 0..N/2 in place, columns N/2+1..N-1 reversed after them (column
// N - u at N/2 + u).  A strip of half-plane columns tx and its mirror columns
// N - tx then occupy two aligned runs of equal length, so neighbouring strips
// complete whole cache lines instead of splitting the mirror run across lines
// written by different strips.
template <int N>
__device__ __forceinline__ int zpos(int x)
{
    return x <= N / 2 ? x : 3 * N / 2 - x;
}

// zpos of element r of a row transform held as x = j + r*T (T = N/E, j < T),
// with the half the element falls in known at compile time.
template <int N, int E>
__device__ __forceinline__ int zpos_elem(int j, int r)
{
    constexpr int T = N / E;
    if (2 * r < E) return j + r * T;
    if (2 * r > E) return 3 * N / 2 - (j + r * T);
    return j == 0 ? N / 2 : N - j;
}

// Z1 and Z2 are stored in tiles of two rows: element (row y, position p) sits
// at float2 index (y/2 * N + p) * 2 + y%2, the two rows of a tile interleaved.
// A column strip of C columns then writes runs of 2C values (one 16-byte store
// per column and tile) instead of runs of C per row, which halves the write
// requests of the column pass, and the row pass reads both rows of a tile with
// one 16-byte load per position.  Tiles start 16-byte aligned when the buffer
// does.
template <int N>
__device__ __forceinline__ size_t zidx(int y, int p)
{
    return ((size_t)(y >> 1) * N + p) * 2 + (y & 1);
}
template <int N>
__device__ __forceinline__ float4* ztile(float2* z, int m)
{
    return reinterpret_cast<float4*>(z) + (size_t)m * N;
}
template <int N>
__device__ __forceinline__ const float4* ztile(const float2* z, int m)
{
    return reinterpret_cast<const float4*>(z) + (size_t)m * N;
}

// ------------------------------------------------------------------ XCD-aware block order

// MI355X spreads the workgroups of a launch round-robin over its 8 XCDs
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  This maps the
// hardware block index to a logical index such that XCD x processes the
// contiguous logical range [x*q + min(x, r), ...) in order (q = n / 8,
// r = n % 8): consecutive logical blocks share an L2.  It is a bijection on
// [0, n) for any n.
__device__ __forceinline__ int xcd_logical(int b, int n)
{
    constexpr int X = 8;
    const int q = n / X, r = n % X, x = b % X;
    return x * q + (x < r ? x : r) + b / X;
}

// Column strips per block group (consecutive blocks on one XCD): G strips of
// C columns cover 32 columns of every Z tile row from N = 256 on (four whole
// 128-byte lines), 16 below.  Wider groups complete more lines in one L2
// before they are written back: at N = 1024 the column pass writes 14.5
// instead of 16.7 B/texel and the frame drops from 17.3 to 16.6 us; groups
// of 64 or 128 columns were slower again, and at N = 128 the wider group
// cost 1%.
template <int N, int C>
__host__ __device__ constexpr int strip_group()
{
    constexpr int W = N >= 256 ? 32 : 16;
    return C >= W ? 1 : W / C;
}


// ------------------------------------------------------------------ column pass: one strip

// Strips of cols_lds_strip per frame: the N/2 + 1 half-plane columns
// 0 .. N/2 in strips of C, with the two self-mirror columns 0 and N/2 sharing
// one transform (below), i.e. N/2 column slots.
template <int N, int C>
__host__ __device__ constexpr int lds_strips()
{
    return (N / 2) / C;
}

// Stride (float2) of the per-(field, column) LDS buffers of cols_lds_strip:
// lds_len(N) plus a few elements.  The fold writes of 16 lanes (4 columns by
// 4 consecutive rows, then their partner rows N - ty) and the store reads
// then fall on distinct banks more often: with a stride of 1 mod 32 the 32
// banks of a 16-lane b64 access saw up to 4 lanes per bank in the fold
// writes.  A bank model of the strip's accesses gives 1536 instead of 2979
// fold-write cycles per block at N = 1024 and C = 4 (7424 instead of 8960
// over folds, transforms and store reads with one fold per texel pair), and
// no width C in 1..8 more.
__host__ __device__ constexpr int strip_len(int N)
{
    return N == 1024 ? 1091 : N == 512 ? 547 : N == 256 ? 275 : N == 128 ? 152 : N == 64 ? 76 : lds_len(N);
}

// Phases of cols_lds_strip (below), shared with the two-frame strip.
// Strip 0, after the folds: column slot 0 += i * (the Nyquist column's fold),
// its N rows spread over the whole block (strip is the same for every thread
// of the block).  Begins with a block barrier.
template <int N, int C, int NB, bool HT, int CL = C>
__device__ __forceinline__ void strip_nyquist(const float4* __restrict__ h0, const SpecConst& sc, double t_rev,
                                              float2* lds)
{
    auto at = [&](int x, int y) -> float4 { return HT ? h0[(size_t)x * N + y] : h0[(size_t)y * N + x]; };
    constexpr int LEN = strip_len(N);
    __syncthreads();
    for (int ty = threadIdx.x; ty < N; ty += NB) {
        const int my = (N - ty) & (N - 1);
        float2 by, bx, bz;
        pair_fold(sc, at(N / 2, ty), at(N / 2, my), 0, ty - N / 2, 0, my - N / 2, t_rev, by, bx, bz);
        float2& ay = lds[0 * CL * LEN + lds_pad(ty)];
        float2& ax = lds[1 * CL * LEN + lds_pad(ty)];
        float2& az = lds[2 * CL * LEN + lds_pad(ty)];
        ay = make_float2(ay.x - by.y, ay.y + by.x);
        ax = make_float2(ax.x - bx.y, ax.y + bx.x);
        az = make_float2(az.x - bz.y, az.y + bz.x);
    }
}

// The 3C column transforms in place, one per T-thread group of the first NT
// threads (Sync: WaveSync when a group lies in one wave, else BlockSync).
template <int N, int C, class Sync, int NB>
__device__ __forceinline__ void strip_transform(const float2* __restrict__ tw, float2* lds)
{
    constexpr int E = 16, T = N / E, NT = 3 * C * T, LEN = strip_len(N);
    const bool active = NB == NT || (int)threadIdx.x < NT;  // takes part in the transforms
    if (active) {
        const int g = threadIdx.x / T, j = threadIdx.x % T;
        float2* buf = lds + g * LEN;
        float2  v[E];
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = buf[lds_pad(j + r * T)];
        fft_inverse<N, E>(v, j, buf, tw, Sync{});
        Sync{}();
#pragma unroll
        for (int r = 0; r < E; ++r) buf[lds_pad(j + r * T)] = v[r];
    }
}

// Z1 = G_h + i G_dx and Z2 = G_dz(2p) + i G_dz(2p+1) of the strip's columns
// and their mirrors from the transformed LDS buffers, in two-row tiles:
// thread k stores Z1 tile m = k / C (rows 2m, 2m+1) and, for m < N/4, Z2
// tile m (Z2 rows 2m, 2m+1, i.e. G_dz rows 4m..4m+3) of column c = k % C.
template <int N, int C, int NB, int CL = C>
__device__ __forceinline__ void strip_store(float2* __restrict__ z1, float2* __restrict__ z2, int strip,
                                            const float2* lds)
{
    constexpr int LEN = strip_len(N);
    constexpr int ITZ = (C * N / 2 + NB - 1) / NB;
#pragma unroll
    for (int it = 0; it < ITZ; ++it) {
        const int k = threadIdx.x + it * NB;
        if (k < C * N / 2) {
            const int    c   = k % C, m = k / C;
            const int    tx  = strip * C + c;
            const float2 y0  = lds[(0 * CL + c) * LEN + lds_pad(2 * m)];
            const float2 x0  = lds[(1 * CL + c) * LEN + lds_pad(2 * m)];
            const float2 y1  = lds[(0 * CL + c) * LEN + lds_pad(2 * m + 1)];
            const float2 x1  = lds[(1 * CL + c) * LEN + lds_pad(2 * m + 1)];
            float4*      t1  = ztile<N>(z1, m);
            const bool   z2t = m < N / 4;  // Z2 row q pairs G_dz rows 2q and 2q+1
            float2       a{}, b{}, d{}, e{};
            if (z2t) {
                a = lds[(2 * CL + c) * LEN + lds_pad(4 * m)];
                b = lds[(2 * CL + c) * LEN + lds_pad(4 * m + 1)];
                d = lds[(2 * CL + c) * LEN + lds_pad(4 * m + 2)];
                e = lds[(2 * CL + c) * LEN + lds_pad(4 * m + 3)];
            }
            float4* t2 = ztile<N>(z2, m);
            if (tx == 0) {
                // packed self-mirror columns: real parts are column 0's (real)
                // transforms, imaginary parts column N/2's
                t1[0]     = make_float4(y0.x, x0.x, y1.x, x1.x);
                t1[N / 2] = make_float4(y0.y, x0.y, y1.y, x1.y);
                if (z2t) {
                    t2[0]     = make_float4(a.x, b.x, d.x, e.x);
                    t2[N / 2] = make_float4(a.y, b.y, d.y, e.y);
                }
            } else {
                const int mxc = N - tx;
                t1[tx]        = make_float4(y0.x - x0.y, y0.y + x0.x, y1.x - x1.y, y1.y + x1.x);
                t1[zpos<N>(mxc)] = make_float4(y0.x + x0.y, x0.x - y0.y, y1.x + x1.y, x1.x - y1.y);
                if (z2t) {
                    t2[tx]        = make_float4(a.x - b.y, a.y + b.x, d.x - e.y, d.y + e.x);
                    t2[zpos<N>(mxc)] = make_float4(a.x + b.y, b.x - a.y, d.x + e.y, e.x - d.y);
                }
            }
        }
    }
}

// LDS-staged column transforms of one strip: the C half-plane columns
// tx = strip*C + c, block of 3*C*T threads (T = N/16), frame time t_rev;
// z1/z2 point at the frame's intermediate, lds holds 3*C*strip_len(N).
// Phase A: every thread evaluates Hermitian parts (consecutive threads take
// consecutive columns, then rows ty with their partner rows N - ty, whose
// texel pairs share |k| and one phase) into per-(field, column) LDS buffers.
// Phase B: each of the 3C transforms runs inside one T-thread group (Sync:
// WaveSync when a group lies in one wave, else BlockSync) and leaves its
// result in natural order in the same buffer.  Phase C: every thread stores
// Z1 = G_h + i G_dx (columns tx and N-tx) and Z2 = G_dz(2p) + i G_dz(2p+1),
// both in two-row tiles (zidx).
// The self-mirror columns 0 and N/2 (their own mirrors) have Hermitian folded
// spectra, A(-ty) = conj A(ty), so their column transforms are real: column
// slot 0 of strip 0 transforms A_0 + i A_{N/2} and its result holds column 0 in
// the real parts and column N/2 in the imaginary parts.  That leaves N/2
// column slots, a whole number of strips (at N = 1024 and C = 2, exactly one
// block per CU for a single frame instead of 257 blocks on 256 CUs).
//
// HT: h0 is given transposed (h0[x * N + y] holds texel (x, y)), so the strip
// reads its columns as contiguous runs instead of one 16-byte value per row.
template <int N, int C, class Sync, int NB = 3 * C * (N / 16), bool HT = false>
__device__ __forceinline__ void cols_lds_strip(const float4* __restrict__ h0, float2* __restrict__ z1,
                                               float2* __restrict__ z2, const float2* __restrict__ tw,
                                               const SpecConst& sc, double t_rev, int strip, float2* lds)
{
    auto at = [&](int x, int y) -> float4 { return HT ? h0[(size_t)x * N + y] : h0[(size_t)y * N + x]; };
    // NB = threads of the block: all of them fold (phase A) and store (phase
    // C); the 3C transforms take the first NT.  A block wider than NT (a
    // single frame, whose few blocks leave the SIMDs short of waves, or the
    // batch's block shared by every N) splits the folds and stores finer.
    constexpr int E = 16, T = N / E, NT = 3 * C * T, LEN = strip_len(N);
    constexpr int ITERS = (C * (N / 2 + 1) + NB - 1) / NB;  // row-pair tasks per thread
    static_assert(NB >= NT, "block too small for the strip");
    static_assert(NB == NT || !std::is_same<Sync, BlockSync>::value, "block barriers need every thread active");
    static_assert((N / 2) % C == 0, "strips of C columns tile the half plane");
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int k = threadIdx.x + it * NB;
        if (k < C * (N / 2 + 1)) {
            // rows ra and rb = -ra mod N of column c (ny -> -ny): their texel
            // pairs share |k| and so one dispersion phase; rows 0 and N/2 are
            // their own partners and are folded once
            const int c   = k % C, ra = k / C;
            const int rb  = (N - ra) & (N - 1);
            const int tx  = strip * C + c;  // < N/2
            const int mxc = (N - tx) & (N - 1);
            const int nx = tx - N / 2, qx = mxc - N / 2;
            float     kinv, sn, cs;
            sincos_turns(turns(omega_fp64(sc, nx * nx + (ra - N / 2) * (ra - N / 2), kinv), t_rev), sn, cs);
            float2 ay, ax, az;
            pair_fold_sc(sc, at(tx, ra), at(mxc, rb), nx, ra - N / 2, qx, rb - N / 2, kinv, sn, cs, ay, ax, az);
            lds[(0 * C + c) * LEN + lds_pad(ra)] = ay;
            lds[(1 * C + c) * LEN + lds_pad(ra)] = ax;
            lds[(2 * C + c) * LEN + lds_pad(ra)] = az;
            if (rb != ra) {
                pair_fold_sc(sc, at(tx, rb), at(mxc, ra), nx, rb - N / 2, qx, ra - N / 2, kinv, sn, cs, ay, ax, az);
                lds[(0 * C + c) * LEN + lds_pad(rb)] = ay;
                lds[(1 * C + c) * LEN + lds_pad(rb)] = ax;
                lds[(2 * C + c) * LEN + lds_pad(rb)] = az;
            }
        }
    }
    if (strip == 0) strip_nyquist<N, C, NB, HT>(h0, sc, t_rev, lds);
    __syncthreads();
    strip_transform<N, C, Sync, NB>(tw, lds);
    __syncthreads();
    strip_store<N, C, NB>(z1, z2, strip, lds);
}

// cols_lds_strip for two frames (times ta, tb) of the same C columns, both
// in LDS at once as 2C buffers per field (frame a's columns, then frame
// b's): each texel pair's h0 values are loaded once for both frames, and the
// dispersion is evaluated once for the two row pairs (ty, N - ty) and both
// frames.  Block = 6*C*T threads (NB), the 6C transforms one
// per T-thread group; lds holds 6*C*strip_len(N).
template <int N, int C, class Sync, int NB = 6 * C * (N / 16)>
__device__ __forceinline__ void cols_lds_strip_f2(const float4* __restrict__ h0, float2* __restrict__ z1a,
                                                  float2* __restrict__ z2a, float2* __restrict__ z1b,
                                                  float2* __restrict__ z2b, const float2* __restrict__ tw,
                                                  const SpecConst& sc, double ta, double tb, int strip, float2* lds)
{
    constexpr int E = 16, T = N / E, NT = 6 * C * T, LEN = strip_len(N), CL = 2 * C;
    constexpr int ITERS = (C * (N / 2 + 1) + NB - 1) / NB;  // row-pair tasks per thread
    static_assert(NB >= NT, "block too small for the strip");
    static_assert(NB == NT || !std::is_same<Sync, BlockSync>::value, "block barriers need every thread active");
    static_assert((N / 2) % C == 0, "strips of C columns tile the half plane");
    // frame f's folds of the texel pair at row y, all sharing the sin/cos of
    // that frame: frame a to buffers 0..C-1 of each field, frame b to C..2C-1
    auto fold2 = [&](float4 p, float4 q, int nx, int y, int qx, int qy, float kinv, float sa, float ca, float sb,
                     float cb, int c) {
        float2 ay, ax, az;
        pair_fold_sc(sc, p, q, nx, y - N / 2, qx, qy, kinv, sa, ca, ay, ax, az);
        lds[(0 * CL + c) * LEN + lds_pad(y)] = ay;
        lds[(1 * CL + c) * LEN + lds_pad(y)] = ax;
        lds[(2 * CL + c) * LEN + lds_pad(y)] = az;
        pair_fold_sc(sc, p, q, nx, y - N / 2, qx, qy, kinv, sb, cb, ay, ax, az);
        lds[(0 * CL + C + c) * LEN + lds_pad(y)] = ay;
        lds[(1 * CL + C + c) * LEN + lds_pad(y)] = ax;
        lds[(2 * CL + C + c) * LEN + lds_pad(y)] = az;
    };
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int k = threadIdx.x + it * NB;
        if (k < C * (N / 2 + 1)) {
            // rows ra and rb = -ra mod N share |k|: one dispersion evaluation
            // and one sin/cos per frame for the four texels
            const int c   = k % C, ra = k / C;
            const int rb  = (N - ra) & (N - 1);
            const int tx  = strip * C + c;  // < N/2
            const int mxc = (N - tx) & (N - 1);
            const int nx = tx - N / 2, qx = mxc - N / 2;
            float        kinv, sa, ca, sb, cb;
            const double w = omega_fp64(sc, nx * nx + (ra - N / 2) * (ra - N / 2), kinv);
            sincos_turns(turns(w, ta), sa, ca);
            sincos_turns(turns(w, tb), sb, cb);
            fold2(h0[(size_t)ra * N + tx], h0[(size_t)rb * N + mxc], nx, ra, qx, rb - N / 2, kinv, sa, ca, sb, cb, c);
            if (rb != ra)
                fold2(h0[(size_t)rb * N + tx], h0[(size_t)ra * N + mxc], nx, rb, qx, ra - N / 2, kinv, sa, ca, sb, cb,
                      c);
        }
    }
    if (strip == 0) {
        strip_nyquist<N, C, NB, false, CL>(h0, sc, ta, lds);
        strip_nyquist<N, C, NB, false, CL>(h0, sc, tb, lds + C * LEN);
    }
    __syncthreads();
    strip_transform<N, 2 * C, Sync, NB>(tw, lds);
    __syncthreads();
    strip_store<N, C, NB, CL>(z1a, z2a, strip, lds);
    strip_store<N, C, NB, CL>(z1b, z2b, strip, lds + C * LEN);
}

// The same strip in two passes through half the LDS: block = 2*C*T threads,
// lds holds 2*C*lds_len(N).  Phase A folds all three spectra; height and Dx go
// to LDS, each thread keeps its Dz values in registers.  The first pass
// transforms the 2C height/Dx columns and stores Z1; the second moves Dz into
// LDS, transforms the C Dz columns (half the block) and stores Z2.  With half the LDS of
// cols_lds_strip, two blocks fit a CU and one block's transforms overlap the
// other's loads and stores.
template <int N, int C>
__device__ __forceinline__ void cols_lds2_strip(const float4* __restrict__ h0, float2* __restrict__ z1,
                                                float2* __restrict__ z2, const float2* __restrict__ tw,
                                                const SpecConst& sc, double t_rev, int strip, float2* lds)
{
    constexpr int E = 16, T = N / E, NT = 2 * C * T, LEN = lds_len(N);
    constexpr int ITERS = C * N / NT;  // = N / (2T) = 8
    static_assert(T <= 64 && 64 % T == 0 && (C * N) % NT == 0, "each transform group inside one wave");
    float2 dzr[ITERS];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int k   = threadIdx.x + it * NT;
        const int c   = k % C, ty = k / C;
        const int txc = min(strip * C + c, N / 2);
        const int mxc = (N - txc) & (N - 1);
        const int my  = (N - ty) & (N - 1);
        float2 ay, ax;
        pair_fold(sc, h0[(size_t)ty * N + txc], h0[(size_t)my * N + mxc], txc - N / 2, ty - N / 2, mxc - N / 2,
                  my - N / 2, t_rev, ay, ax, dzr[it]);
        lds[c * LEN + lds_pad(ty)]       = ay;
        lds[(C + c) * LEN + lds_pad(ty)] = ax;
    }
    auto transform = [&](int groups) {
        if ((int)threadIdx.x < groups * T) {  // whole waves
            const int g = threadIdx.x / T, j = threadIdx.x % T;
            float2*   buf = lds + g * LEN;
            float2    v[E];
#pragma unroll
            for (int r = 0; r < E; ++r) v[r] = buf[lds_pad(j + r * T)];
            fft_inverse<N, E>(v, j, buf, tw, WaveSync{});
            WaveSync{}();
#pragma unroll
            for (int r = 0; r < E; ++r) buf[lds_pad(j + r * T)] = v[r];
        }
    };
    __syncthreads();
    transform(2 * C);
    __syncthreads();
    // Z1 tile m = k / C (rows 2m, 2m+1) of column c = k % C
#pragma unroll
    for (int it = 0; it < ITERS / 2; ++it) {
        const int k = threadIdx.x + it * NT;
        const int c = k % C, m = k / C;
        const int tx = strip * C + c;
        if (tx <= N / 2) {
            const int    mxc = (N - tx) & (N - 1);
            const float2 y0  = lds[c * LEN + lds_pad(2 * m)];
            const float2 x0  = lds[(C + c) * LEN + lds_pad(2 * m)];
            const float2 y1  = lds[c * LEN + lds_pad(2 * m + 1)];
            const float2 x1  = lds[(C + c) * LEN + lds_pad(2 * m + 1)];
            float4*      t1  = ztile<N>(z1, m);
            t1[tx] = make_float4(y0.x - x0.y, y0.y + x0.x, y1.x - x1.y, y1.y + x1.x);
            if (mxc != tx) t1[zpos<N>(mxc)] = make_float4(y0.x + x0.y, x0.x - y0.y, y1.x + x1.y, x1.x - y1.y);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int k = threadIdx.x + it * NT;
        lds[(k % C) * LEN + lds_pad(k / C)] = dzr[it];
    }
    __syncthreads();
    transform(C);
    __syncthreads();
    // Z2 tile m = k / C, m < N/4: Z2 rows 2m, 2m+1, Z2 row q pairing G_dz rows 2q and 2q+1
#pragma unroll
    for (int it = 0; it < ITERS / 4; ++it) {
        const int k = threadIdx.x + it * NT;
        const int c = k % C, m = k / C;
        const int tx = strip * C + c;
        if (tx <= N / 2) {
            const int    mxc = (N - tx) & (N - 1);
            const float2 a   = lds[c * LEN + lds_pad(4 * m)];
            const float2 b   = lds[c * LEN + lds_pad(4 * m + 1)];
            const float2 d   = lds[c * LEN + lds_pad(4 * m + 2)];
            const float2 e   = lds[c * LEN + lds_pad(4 * m + 3)];
            float4*      t2  = ztile<N>(z2, m);
            t2[tx] = make_float4(a.x - b.y, a.y + b.x, d.x - e.y, d.y + e.x);
            if (mxc != tx) t2[zpos<N>(mxc)] = make_float4(a.x + b.y, b.x - a.y, d.x + e.y, e.x - d.y);
        }
    }
}

// cols_lds2_strip for N >= 2048 (transform groups of several waves, block
// barriers): block = 2*C*T threads, lds holds 2*C*lds_len(N), h0t is h0
// transposed (the strip's columns contiguous).  Phase A folds all three
// spectra; height and Dx go to LDS and each thread keeps its Dz values in
// registers.  The first pass transforms the 2C height/Dx columns at radix 16
// and stores Z1; the second moves Dz into LDS and transforms its C columns at
// radix 8 (N/8 threads each), so every thread of the block takes part in both
// passes and meets every barrier.  With half the LDS per column of
// cols_lds_strip, a strip holds twice the columns: at N = 4096 two columns, so
// each two-row tile row gets 32 contiguous bytes from one block instead of 16.
template <int N, int C>
__device__ __forceinline__ void cols_lds2b_strip(const float4* __restrict__ h0t, float2* __restrict__ z1,
                                                 float2* __restrict__ z2, const float2* __restrict__ tw16,
                                                 const float2* __restrict__ tw8, const SpecConst& sc, double t_rev,
                                                 int strip, float2* lds)
{
    constexpr int T = N / 16, NT = 2 * C * T, LEN = lds_len(N);
    constexpr int ITERS = C * N / NT;  // = 8
    static_assert(T % 64 == 0 && NT <= 1024 && C * (N / 8) == NT, "whole waves; both passes use every thread");
    float2 dzr[ITERS];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int k   = threadIdx.x + it * NT;
        const int c   = k % C, ty = k / C;
        const int txc = min(strip * C + c, N / 2);
        const int mxc = (N - txc) & (N - 1);
        const int my  = (N - ty) & (N - 1);
        float2 ay, ax;
        pair_fold(sc, h0t[(size_t)txc * N + ty], h0t[(size_t)mxc * N + my], txc - N / 2, ty - N / 2, mxc - N / 2,
                  my - N / 2, t_rev, ay, ax, dzr[it]);
        lds[c * LEN + lds_pad(ty)]       = ay;
        lds[(C + c) * LEN + lds_pad(ty)] = ax;
    }
    __syncthreads();
    {
        const int g = threadIdx.x / T, j = threadIdx.x % T;
        float2*   buf = lds + g * LEN;
        float2    v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = buf[lds_pad(j + r * T)];
        fft_inverse<N, 16>(v, j, buf, tw16, BlockSync{});
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; ++r) buf[lds_pad(j + r * T)] = v[r];
    }
    __syncthreads();
    // Z1 tile m = k / C (rows 2m, 2m+1) of column c = k % C
#pragma unroll
    for (int it = 0; it < ITERS / 2; ++it) {
        const int k = threadIdx.x + it * NT;
        const int c = k % C, m = k / C;
        const int tx = strip * C + c;
        if (tx <= N / 2) {
            const int    mxc = (N - tx) & (N - 1);
            const float2 y0  = lds[c * LEN + lds_pad(2 * m)];
            const float2 x0  = lds[(C + c) * LEN + lds_pad(2 * m)];
            const float2 y1  = lds[c * LEN + lds_pad(2 * m + 1)];
            const float2 x1  = lds[(C + c) * LEN + lds_pad(2 * m + 1)];
            float4*      t1  = ztile<N>(z1, m);
            t1[tx] = make_float4(y0.x - x0.y, y0.y + x0.x, y1.x - x1.y, y1.y + x1.x);
            if (mxc != tx) t1[zpos<N>(mxc)] = make_float4(y0.x + x0.y, x0.x - y0.y, y1.x + x1.y, x1.x - y1.y);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int k = threadIdx.x + it * NT;
        lds[(k % C) * LEN + lds_pad(k / C)] = dzr[it];
    }
    __syncthreads();
    {
        constexpr int T8 = N / 8;
        const int     g = threadIdx.x / T8, j = threadIdx.x % T8;
        float2*       buf = lds + g * LEN;
        float2        v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = buf[lds_pad(j + r * T8)];
        fft_inverse<N, 8>(v, j, buf, tw8, BlockSync{});
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) buf[lds_pad(j + r * T8)] = v[r];
    }
    __syncthreads();
    // Z2 tile m = k / C, m < N/4: Z2 rows 2m, 2m+1, Z2 row q pairing G_dz rows 2q and 2q+1
#pragma unroll
    for (int it = 0; it < ITERS / 4; ++it) {
        const int k = threadIdx.x + it * NT;
        const int c = k % C, m = k / C;
        const int tx = strip * C + c;
        if (tx <= N / 2) {
            const int    mxc = (N - tx) & (N - 1);
            const float2 a   = lds[c * LEN + lds_pad(4 * m)];
            const float2 b   = lds[c * LEN + lds_pad(4 * m + 1)];
            const float2 d   = lds[c * LEN + lds_pad(4 * m + 2)];
            const float2 e   = lds[c * LEN + lds_pad(4 * m + 3)];
            float4*      t2  = ztile<N>(z2, m);
            t2[tx] = make_float4(a.x - b.y, a.y + b.x, d.x - e.y, d.y + e.x);
            if (mxc != tx) t2[zpos<N>(mxc)] = make_float4(a.x + b.y, b.x - a.y, d.x + e.y, e.x - d.y);
        }
    }
}

// ------------------------------------------------------------------ row pass + normals + Jacobian: one band

// Row transforms fused with the normal map (nearest taps) and the Jacobian
// for one band of R rows [y0, y0 + R), y0 = band*R even (R even, dividing N;
// the caller checks).  A group of T = N/E
// threads (thread j holds x = j + r*T) walks the band with a three-row window
// in registers: for each new row it transforms Z1 (height + i Dx) and, on
// even rows, the Z2 row holding Dz of that row and the next, then applies
// (-1)^(x+y) / N^2 (inversion_cs.glsl:35-41).  The x-neighbours of the stencil
// pass through the group's LDS buffer buf (lds_len(N) entries); halo rows
// y0-1 and y0+R are transformed too (periodic).  Sobel taps as
// normal_map_cs.glsl:37-53; J = (1 - l dDx/dx)(1 - l dDz/dz) - l^2 dDx/dz dDz/dx
// for the render displacement (-l Dx, -l Dz) (grid_tes.glsl:60-62), central
// differences with spacing L/N.  All pointers are at the frame's data;
// normals or jac may be null.  RENDER adds the render export (disp, rnorm,
// either may be null): (-l Dx, s h, -l Dz, J) and the unit normal of the
// displaced surface, dP/dv x dP/du.
template <int N, int E, class Sync, bool RENDER = false, bool PAIR2 = false, bool ALT = false>
__device__ __forceinline__ void rows_nrm_band(const float2* __restrict__ z1, const float2* __restrict__ z2,
                                              float* __restrict__ height, float* __restrict__ dx,
                                              float* __restrict__ dz, float4* __restrict__ normals,
                                              float* __restrict__ jac, const float2* __restrict__ tw, float2* buf,
                                              int j, int band, int R, float lambda, float inv2d,
                                              float4* __restrict__ disp = nullptr, float4* __restrict__ rnorm = nullptr,
                                              float dscale = 0.0f)
{
    constexpr int T = N / E;
    constexpr float scale = 1.0f / ((float)N * (float)N);
    // With ALT, odd bands walk their rows upwards, so that the halo tiles two
    // neighbouring bands share are read by both near the same time (both
    // walks end at their common boundary and start at the band boundaries
    // they share with the other neighbours) and come from one L2 fetch.
    // The walk runs over walk rows w = y0 .. y0+R-1 in its own order: walk row
    // w is grid row arow(w) = w, or N-1-w on an upward walk; both take the
    // unwrapped -1 .. N.  The flip maps a two-row tile onto a two-row tile
    // with its halves exchanged (N-1 is odd), and a Z2 row (rows 2u, 2u+1 in
    // its real and imaginary parts) onto a Z2 row with the parts exchanged, so
    // the walk below reads tiles and Z2 parts in walk order.  Signs, output
    // rows and stencils use grid rows, with the neighbours named by grid
    // position, so that every row's arithmetic is that of a downward walk
    // (frames are bit-identical either way).  N=1024 / 2048 frames -0.9%; below
    // that the selects cost more than the halo reads (N=512 +2%).
    const bool flip = ALT && (band & 1) != 0;
    const int  fm   = flip ? -1 : 0;
    const int  y0   = flip ? N - (band + 1) * R : band * R;
    auto arow  = [&](int w) { return (w ^ fm) & (N - 1); };
    auto arow2 = [&](int u) { return (u ^ fm) & (N / 2 - 1); };  // Z2 rows

    // Rows are read from the two-row tiles of Z1 / Z2: load_row takes one grid
    // row (8 bytes of every 16), load_pair the tile holding walk rows w, w+1
    // (w even; the caller passes the tile index arow(w) >> 1) -- walk row w
    // into v, w+1 into w -- with one 16-byte load per position.  Thread j
    // holds x = j + r*T (natural order after the transform).
    auto load_row = [&](const float2* z, int y, float2 (&v)[E]) {
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = z[zidx<N>(y, zpos_elem<N, E>(j, r))];
    };
    auto load_pair = [&](const float2* z, int tile, float2 (&v)[E], float2 (&w)[E]) {
        const float4* t = ztile<N>(z, tile);
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const float4 q  = t[zpos_elem<N, E>(j, r)];
            const float2 lo = make_float2(q.x, q.y), hi = make_float2(q.z, q.w);
            v[r] = flip ? hi : lo;
            w[r] = flip ? lo : hi;
        }
    };
    auto fft = [&](float2 (&v)[E]) { fft_inverse<N, E>(v, j, buf, tw, Sync{}); };
    // a transformed Z2 row in walk order: even walk row in .x, odd in .y
    auto fft2 = [&](float2 (&v)[E]) {
        fft(v);
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = flip ? make_float2(v[r].y, v[r].x) : v[r];
    };
    // (-1)^(x+y) / N^2 of grid row y; x = j + r*T and T is even, so the parity is that of j + y
    auto sgn = [&](int y) { return ((j + y) & 1) ? -scale : scale; };
    // d/dz by central differences: walk neighbours (next - previous) times this
    const float vinv = flip ? -inv2d : inv2d;

    // The band reads walk rows y0-1 .. y0+R of Z1 and Z2 rows (y0-1)/2 .. (y0+R)/2
    // (row N wraps to 0), counted unwrapped: an even Z1 row y >= y0 other
    // than the last is loaded together with row y+1, which waits in p1 for
    // the next step.  With PAIR2 the Z2 rows b2 .. e2 are paired the same way
    // (pending row in p2); without it they are read one at a time, which saves
    // the registers of p2.
    const int b2 = y0 >> 1, e2 = b2 + R / 2;
    float hA[E], xA[E], zA[E], hB[E], xB[E], zB[E], zN[E];
    float2 v[E], p1[E], p2[E];
    {
        const int yA = (y0 - 1) & (N - 1);  // odd: Dz from the odd part of Z2 row yA >> 1
        load_row(z2, arow2(yA >> 1), v);
        fft2(v);
        const float sa = sgn(arow(yA));
#pragma unroll
        for (int r = 0; r < E; ++r) zA[r] = sa * v[r].y;
        load_row(z1, arow(yA), v);
        fft(v);
#pragma unroll
        for (int r = 0; r < E; ++r) {
            hA[r] = sa * v[r].x;
            xA[r] = sa * v[r].y;
        }
        const float sb = sgn(arow(y0));
        if (PAIR2 && (b2 & 1) == 0) load_pair(z2, arow2(b2) >> 1, v, p2);
        else load_row(z2, arow2(b2), v);
        fft2(v);
#pragma unroll
        for (int r = 0; r < E; ++r) {
            zB[r] = sb * v[r].x;
            zN[r] = -sb * v[r].y;  // Dz of walk row y0 + 1 (opposite sign)
        }
        load_pair(z1, arow(y0) >> 1, v, p1);  // R >= 2: walk row y0 + 1 is in the band
        fft(v);
#pragma unroll
        for (int r = 0; r < E; ++r) {
            hB[r] = sb * v[r].x;
            xB[r] = sb * v[r].y;
        }
    }

#pragma unroll 1
    for (int k = 0; k < R; ++k) {
        const int   y  = y0 + k;                 // walk row
        const int   yu = y + 1;                  // unwrapped, <= N
        const int   yC = yu & (N - 1);
        const float sc = sgn(arow(yu));
        float zC[E];
        if ((yC & 1) == 0) {
            const int u = yu >> 1;  // Z2 walk row, unwrapped
            if (PAIR2 && (u & 1)) {
#pragma unroll
                for (int r = 0; r < E; ++r) v[r] = p2[r];
            } else if (PAIR2 && u < e2) {
                load_pair(z2, arow2(u) >> 1, v, p2);
            } else {
                load_row(z2, arow2(yC >> 1), v);
            }
            fft2(v);
#pragma unroll
            for (int r = 0; r < E; ++r) {
                zC[r] = sc * v[r].x;
                zN[r] = -sc * v[r].y;
            }
        } else {
#pragma unroll
            for (int r = 0; r < E; ++r) zC[r] = zN[r];
        }
        if (yu & 1) {
#pragma unroll
            for (int r = 0; r < E; ++r) v[r] = p1[r];
        } else if (yu < y0 + R) {
            load_pair(z1, arow(yC) >> 1, v, p1);
        } else {
            load_row(z1, arow(yC), v);
        }
        fft(v);  // v = height + i Dx of walk row yC (unsigned)

        const size_t row = (size_t)arow(y) * N;
        if (normals) {
            // v_s = h(y-1) + 2 h(y) + h(y+1), d_s = h(y-1) - h(y+1); nx = v_s(x-1) - v_s(x+1),
            // nz = d_s(x-1) + 2 d_s(x) + d_s(x+1)  (normal_map_cs.glsl:49-51); up / dn are the
            // grid rows y-1 / y+1 (walk rows y+1 / y-1 on an upward walk)
            Sync{}();
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const float hC = sc * v[r].x;
                const float up = flip ? hC : hA[r], dn = flip ? hA[r] : hC;
                buf[lds_pad(j + r * T)] = make_float2(up + 2.0f * hB[r] + dn, up - dn);
            }
            Sync{}();
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const int    x  = j + r * T;
                const float2 lf = buf[lds_pad((x - 1) & (N - 1))];
                const float2 rt = buf[lds_pad((x + 1) & (N - 1))];
                const float2 md = buf[lds_pad(x)];
                const float  nx = lf.x - rt.x;
                const float  nz = lf.y + 2.0f * md.y + rt.y;
                const float  inv = rsqrtf(fmaf(nx, nx, fmaf(nz, nz, 1.0f)));
                st_out(&normals[row + x], make_float4(nx * inv, inv, nz * inv, 1.0f));
            }
        }
        if (jac || (RENDER && (disp || rnorm))) {
            float dhdx[E];
            if (RENDER && rnorm) {  // x-derivative of the height for the displaced-surface normal
                Sync{}();
#pragma unroll
                for (int r = 0; r < E; ++r) buf[lds_pad(j + r * T)] = make_float2(hB[r], 0.0f);
                Sync{}();
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    const int x = j + r * T;
                    dhdx[r] = (buf[lds_pad((x + 1) & (N - 1))].x - buf[lds_pad((x - 1) & (N - 1))].x) * inv2d;
                }
            }
            Sync{}();
#pragma unroll
            for (int r = 0; r < E; ++r) buf[lds_pad(j + r * T)] = make_float2(xB[r], zB[r]);
            Sync{}();
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const int    x    = j + r * T;
                const float2 lf   = buf[lds_pad((x - 1) & (N - 1))];
                const float2 rt   = buf[lds_pad((x + 1) & (N - 1))];
                const float  dxdx = (rt.x - lf.x) * inv2d;
                const float  dzdx = (rt.y - lf.y) * inv2d;
                const float  dxdz = (sc * v[r].y - xA[r]) * vinv;
                const float  dzdz = (zC[r] - zA[r]) * vinv;
                const float  J = (1.0f - lambda * dxdx) * (1.0f - lambda * dzdz) - lambda * lambda * dxdz * dzdx;
                if (jac) st_out(&jac[row + x], J);
                if (!RENDER) continue;
                if (disp) st_out(&disp[row + x], make_float4(-lambda * xB[r], dscale * hB[r], -lambda * zB[r], J));
                if (rnorm) {
                    // n = dP/dv x dP/du with dP/du = (1 - l dDx/dx, s dh/dx, -l dDz/dx),
                    // dP/dv = (-l dDx/dz, s dh/dz, 1 - l dDz/dz); its y component is J
                    const float dhdz = (sc * v[r].x - hA[r]) * vinv;
                    const float nx = -dscale * (lambda * dhdz * dzdx + (1.0f - lambda * dzdz) * dhdx[r]);
                    const float nz = -dscale * (lambda * dxdz * dhdx[r] + dhdz * (1.0f - lambda * dxdx));
                    const float inv = rsqrtf(fmaf(nx, nx, fmaf(J, J, nz * nz)));
                    st_out(&rnorm[row + x], make_float4(nx * inv, J * inv, nz * inv, 1.0f));
                }
            }
        }
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const int x = j + r * T;
            st_out(&height[row + x], hB[r]);
            st_out(&dx[row + x], xB[r]);
            st_out(&dz[row + x], zB[r]);
            hA[r] = hB[r];
            xA[r] = xB[r];
            zA[r] = zB[r];
            hB[r] = sc * v[r].x;
            xB[r] = sc * v[r].y;
            zB[r] = zC[r];
        }
    }
}

}  // namespace ocn
