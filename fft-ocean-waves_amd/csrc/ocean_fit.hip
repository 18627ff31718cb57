// ocean_fit.hip -- differentiable fit of the ocean height: loss over K frames
// against target heights, its adjoint down to dL/dh0 and dL/d(wind speed,
// wind angle, amplitude), and an Adam step, for one or many oceans per
// launch.  No reference counterpart: the reference only renders the forward
// path (tilde_h0_k_cs.glsl, tilde_h0_t_cs.glsl, butterfly_cs.glsl,
// inversion_cs.glsl), whose maths the forward part here repeats.
//
// Loss L = 1/(K N^2) sum_k sum_x (h_k(x) - T_k(x))^2.  h_k is real: with the
// inversion step's (-1)^(x+y) it is the inverse DFT, over centred wave numbers
// n, of the Hermitian part A_k = (H_k(n) + conj H_k(-n))/2 of its spectrum,
// divided by N^2.  By Parseval the loss is then 1/(K N^4) sum_k sum_n
// |A_k(n) - T^_k(n)|^2 with T^_k the DFT of the targets over the same n, and
// its gradient with respect to each texel's spectrum is pointwise.  So a fit
// step evaluates the spectrum of every texel pair and frame, as the forward
// path does, and needs no transform: T^ is made once per change of the
// targets.  Nothing per frame is stored between steps: h0 is recomputed from
// the unit noise xi and the current variables.
//
//   k_fit_xi, k_fit_pq, k_fit_phase  at creation: the unit noise, the fold's
//                   noise coefficients per texel pair, the phase table
//   k_fit_targets_rows/cols  per member and target frame, when the targets
//                   changed: the 2-D DFT T^ at the half-plane texels
//   k_fit_spectral  item (member, strip of C columns or a row part of it):
//                   h0 = s(V, theta, A) xi,
//                   per frame the folded spectrum, its residual against T^,
//                   the loss partial; in gradient mode the pull-back to dL/dh0
//                   and dL/ds per texel and the three parameter partials
//   k_fit_reduce    one block per member: loss, gradient, optional Adam or
//                   quasi-Newton step
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "ocean_device.hpp"
#include "ocean_fit.h"

namespace ocn {

namespace {

// Spectrum constants of the member at its current variables.
__device__ __forceinline__ SpecConst fit_spec(const FitMember& m, const FitVars& v)
{
    SpecConst sc{};
    sc.N = m.N;
    sc.kscale = m.kscale;
    sc.Lp2 = v.Lp2;
    sc.l2 = m.l2;
    sc.A = v.Af;
    sc.wx = v.wx;
    sc.wy = v.wy;
    sc.g = m.g;
    sc.conj_mode = m.conj_mode;
    sc.k2_d = m.k2_d;
    sc.g2_d = m.g2_d;
    return sc;
}

// e^{i phi} = (cos, sin) of the dispersion phase phi = w t of frame k.  With
// equally spaced times the kernels instead start from frame 0 and step by
// e^{i w dt} (one complex multiply per frame in place of the FP64 phase
// reduction and the sin/cos polynomials); up to 16 frames the recurrence adds
// at most ~2e-6 of relative error.
struct PhaseWalk {
    float2 cur, step;
    __device__ __forceinline__ void start(double w, double t0_rev, double dt_rev)
    {
        float s, c;
        sincos_turns(turns(w, t0_rev), s, c);
        cur = make_float2(c, s);
        sincos_turns(turns(w, dt_rev), s, c);
        step = make_float2(c, s);
    }
    // from a phase-table entry (cos, sin of w t_0; cos, sin of w dt)
    __device__ __forceinline__ void load(float4 q)
    {
        cur  = make_float2(q.x, q.y);
        step = make_float2(q.z, q.w);
    }
    // e^{i phi} of the current frame, then advance to the next
    __device__ __forceinline__ float2 next()
    {
        const float2 r = cur;
        cur = make_float2(fmaf(r.x, step.x, -r.y * step.y), fmaf(r.x, step.y, r.y * step.x));
        return r;
    }
};

__device__ __forceinline__ float2 phase_at(double w, double t_rev)
{
    float s, c;
    sincos_turns(turns(w, t_rev), s, c);
    return make_float2(c, s);
}

template <class T>
__device__ __forceinline__ T block_sum(T x, T* red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    const int w = threadIdx.x / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = x;
    __syncthreads();
    T t = 0;
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) t += red[i];
    return t;
}

// The fit's unit noise xi is stored in column strips: texel (x, y) at
// ((p / C) N + y) C + p % C with p = zpos(x) and C = fit_cols_per_block(N), so
// that a strip's C columns and, through zpos, its C mirror columns are
// contiguous runs down the rows.  The Philox counter stays the texel's
// row-major index, so the values are those of a row-major draw.
__host__ __device__ __forceinline__ size_t fit_xi_index(int N, int x, int y)
{
    const int C = fit_cols_per_block(N);
    const int p = x <= N / 2 ? x : 3 * N / 2 - x;
    return ((size_t)(p / C) * N + y) * C + p % C;
}

// ---------------------------------------------------------------- parameter gradient factors

// The parameter-gradient factors of one texel: dL/d(V, theta, A) gets
// dL/ds * ds/d(V, theta, A) with s = sqrt(P/2) unclamped, nothing where the
// clamp binds or s = 0.  s = c |d| with d = k_hat . w and c free of d, so
// ds/dtheta = c sgn(d) dd/dtheta needs no division by d (a d recomputed apart
// from s could round to 0 where s does not); at |d| < 1e-6 the angle term is
// dropped, as in the oracle.  ds/dV = s * dv_scale / |k|^2, ds/dA = s * da_scale.
// Under k -> -k (a texel and its mirror), |k|, c and s are unchanged and d,
// dd/dtheta change sign, so every factor here is shared by the pair.
struct ParamGeo {
    float s;      // the amplitude
    float rk2;    // 1 / |k|^2
    float dterm;  // c sgn(d) dd/dtheta, or 0
    bool  live;
};

// sqA = sqrt(A): c = sqrt(A exp(x) / |k|^4) / sqrt(2) is evaluated as
// sqA exp(x/2) / |k|^2 / sqrt(2), with 1/|k| from one reciprocal square root
// of the clamped |k|^2 (two transcendentals per texel instead of four).
__device__ __forceinline__ ParamGeo param_geo(const SpecConst& sc, float iLp2, float sqA, int N, int tx, int ty)
{
    const float kx = sc.kscale * (float)(tx - N / 2), ky = sc.kscale * (float)(ty - N / 2);
    const float k2   = fmaxf(kx * kx + ky * ky, 1e-10f);  // |k| clamped at 1e-5
    const float rk   = __builtin_amdgcn_rsqf(k2);
    const float rk2  = rk * rk;
    const float d    = (kx * sc.wx + ky * sc.wy) * rk;
    const float dp   = (ky * sc.wx - kx * sc.wy) * rk;
    const float c    = sqA * expf(-0.5f * (rk2 * iLp2 + k2 * sc.l2)) * rk2 * 0.70710678118654752f;
    ParamGeo    g;
    g.s     = c * fabsf(d);
    g.rk2   = rk2;
    g.dterm = fabsf(d) >= 1e-6f ? c * (d < 0.0f ? -dp : dp) : 0.0f;
    g.live  = g.s > 0.0f && g.s <= 4000.0f;
    return g;
}

// dL/ds of a texel: its dL/dh0 (ga for h0k, gb for h0mk) against the unit noise.
__device__ __forceinline__ float dl_ds(float2 ga, float2 gb, float4 xi)
{
    return ga.x * xi.x + ga.y * xi.y + gb.x * xi.z + gb.y * xi.w;
}

__device__ __forceinline__ void param_add(const ParamGeo& g, float dl, float dv_scale, float da_scale, float& dV,
                                          float& dT, float& dA)
{
    const float w = dl * g.s;
    dV = fmaf(w, dv_scale * g.rk2, dV);
    dT = fmaf(dl, g.dterm, dT);
    dA = fmaf(w, da_scale, dA);
}

// ---------------------------------------------------------------- the step in the spectral domain

// Loss and gradient of one column item, or of one of R equal parts of its
// texel pairs, without transforms.  h_k is real with (-1)^(x+y)/N^2 times the inverse DFT of its Hermitian-folded spectrum A_k
// over the texel grid, i.e. h_k = IDFT(A_k)/N^2 over centred wave numbers n.
// With T^_k the DFT of the targets over the same n (Parseval):
//   sum_x (h_k - T_k)^2 = sum_n |A_k(n) - T^_k(n)|^2 / N^2,
// so each texel pair (u, -u) adds 2 |D|^2 / N^2 with D = A_k(u) - T^_k(u)
// (|D(-u)| = |D(u)|: both A_k and T^_k are Hermitian), or |D|^2 / N^2 for the
// columns 0 and N/2, whose rows hold u and -u as texels of their own.  The
// gradient of L = sum / (K N^2) with respect to the spectrum H(u) of the
// texel is 2 D / (K N^4), and conj of it for -u: the pull-back takes it to
// dL/dh0, then to dL/ds and the parameter partials.  Loss partials hold
// sum_x (h_k - T_k)^2 (k_fit_reduce divides by K N^2).
template <int N, bool UNIFORM, int R>
__device__ __forceinline__ void fit_spectral(const FitMember& m, int item, int mode, float* red)
{
    constexpr int C = fit_cols_per_block(N), NT = kFitColsThreads, IT = C * N / NT / R;
    static_assert((C * N) % (NT * R) == 0, "whole texel pairs per thread");
    // block `item` covers part `item % R` of the strip's C * N texel pairs
    const int strip = item / R, it0 = (item % R) * IT * NT;
    const FitVars   vars = *m.vars;
    const SpecConst sc   = fit_spec(m, vars);
    const bool      grad = mode != kFitLoss;
    const bool      store_grad_h0 = mode == kFitGrad;
    const size_t    ne   = fit_phase_entries(N);
    const size_t    base = (size_t)strip * (N * C);
    const float     inv_nn = 1.0f / ((float)N * (float)N);
    const float     gsc  = 2.0f * m.inv_knn * inv_nn;
    const float     dv_scale = vars.dv_scale, da_scale = vars.da_scale;
    const float     iLp2 = vars.iLp2, sqA = vars.sqA;
    float lsum = 0.0f, dV = 0.0f, dT = 0.0f, dA = 0.0f;
    // A pair's loads (its noise coefficients, phase-table entry and the first
    // four frames' T^) are issued one pair ahead, during the previous pair's
    // arithmetic.  A pair off the edges takes its noise as the coefficients
    // (P~, Q~) of its entry in m.pq; a pair on row 0 or in the columns 0 and
    // N/2 takes both texels' unit noise from m.edge.  Every index stays inside
    // the arrays for the columns past the half plane too (the last strip,
    // tx < N, read from m.pq); those pairs are loaded and skipped.
    float4 na, nb{}, nph{};
    float2 nT[4];
    auto issue = [&](int it) {
        const int kk = it0 + threadIdx.x + it * NT;
        const int cc = kk % C, yy = kk / C, xx = strip * C + cc;
        if (xx <= N / 2 && (xx == 0 || xx == N / 2 || yy == 0)) {
            const int ei = fit_edge_index(N, xx, yy);
            na = m.edge[2 * ei];
            nb = m.edge[2 * ei + 1];
        } else {
            na = m.pq[base + kk];
        }
        if constexpr (UNIFORM) nph = m.phase[base + kk];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            nT[q] = q < m.K ? m.that[(size_t)q * ne + base + kk] : make_float2(0.0f, 0.0f);
    };
    issue(0);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const float4 a = na, b = nb, ph = nph;
        float2       Tq[4] = {nT[0], nT[1], nT[2], nT[3]};
        if (it + 1 < IT) issue(it + 1);
        const int k  = it0 + threadIdx.x + it * NT;
        const int c  = k % C, ty = k / C;
        const int tx = strip * C + c;
        if (tx > N / 2) continue;  // past the half plane (the last strip)
        const int  mx = (N - tx) & (N - 1), my = (N - ty) & (N - 1);
        const bool self = tx == 0 || tx == N / 2;
        // -u lies in the same column for tx = 0, N/2 (a u of its own); off the
        // Nyquist row its centred index is u's negated and shares u's factors
        const bool paired = !self && ty != 0;
        // The amplitude is even in k; the mirror's differs on the Nyquist row or
        // column only.  In gradient mode it is taken from the parameter-gradient
        // factors (the same amplitude to rounding), which the terms after the
        // frames reuse.
        ParamGeo gu{}, gm{};
        float    s, sm;
        if (grad) {
            gu = param_geo(sc, iLp2, sqA, N, tx, ty);
            s  = fminf(gu.s, 4000.0f);
            if (!self && !paired) {
                gm = param_geo(sc, iLp2, sqA, N, mx, my);
                sm = fminf(gm.s, 4000.0f);
            } else {
                sm = (tx != 0 && ty != 0) ? s : phillips_amp(sc, mx, my);
            }
        } else {
            s  = phillips_amp(sc, tx, ty);
            sm = (tx != 0 && ty != 0) ? s : phillips_amp(sc, mx, my);
        }
        PhaseWalk walk;
        double    w = 0.0;
        if constexpr (UNIFORM) {
            walk.load(ph);
        } else {
            const int nx = tx - N / 2, ny = ty - N / 2;
            float     kinv;
            w = omega_fp64(sc, nx * nx + ny * ny, kinv);
        }
        const float wt = self ? inv_nn : 2.0f * inv_nn;
        // With H = h0k e + C(h0mk) conj(e) at u and G the same at -u, the fold
        // (H + conj G)/2 is P e + Q conj(e) = U cos + V sin with P, Q formed once
        // per pair: 4 FMAs per frame.  Off the edges s = sm and P = s P~,
        // Q = s Q~ with the noise coefficients of the pair (k_fit_pq).
        const bool edge = !paired;
        float2     P, Q;
        if (edge) {
            const float bpi = sc.conj_mode ? -a.w : a.w, bqi = sc.conj_mode ? -b.w : b.w;
            P = make_float2(0.5f * (s * a.x + sm * b.z), 0.5f * (s * a.y - sm * bqi));
            Q = make_float2(0.5f * (s * a.z + sm * b.x), 0.5f * (s * bpi - sm * b.y));
        } else {
            P = make_float2(s * a.x, s * a.y);
            Q = make_float2(s * a.z, s * a.w);
        }
        const float2 U = make_float2(P.x + Q.x, P.y + Q.y), Vs = make_float2(Q.y - P.y, P.x - Q.x);
        // The pull-back of G = gsc D through H = h0k e + C(h0mk) conj(e): dL/dh0k
        // (ga) gathers conj(e) G; at -u, where the gradient is conj G, dL/dh0k
        // (gam) gathers conj(e) conj(G).  dL/dh0mk gathers e G at u and e conj(G)
        // at -u for C = identity, conj(e) conj(G) and conj(e) G for C = conj.
        // All four follow from S1 = sum conj(e) D and S2 = sum e D.
        float2 S1{}, S2{};
        // frames in chunks of four: a chunk's T^ loads are issued before its arithmetic
        for (int f0 = 0; f0 < m.K; f0 += 4) {
            if (f0 > 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    Tq[q] = f0 + q < m.K ? m.that[(size_t)(f0 + q) * ne + base + k] : make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (f0 + q >= m.K) break;
                const float2 e  = UNIFORM ? walk.next() : phase_at(w, m.t_rev[f0 + q]);
                const float2 D  = make_float2(fmaf(U.x, e.x, fmaf(Vs.x, e.y, -Tq[q].x)),
                                              fmaf(U.y, e.x, fmaf(Vs.y, e.y, -Tq[q].y)));
                lsum = fmaf(wt, fmaf(D.x, D.x, D.y * D.y), lsum);
                if (grad) {
                    S1.x = fmaf(e.x, D.x, fmaf(e.y, D.y, S1.x));
                    S1.y = fmaf(e.x, D.y, fmaf(-e.y, D.x, S1.y));
                    S2.x = fmaf(e.x, D.x, fmaf(-e.y, D.y, S2.x));
                    S2.y = fmaf(e.x, D.y, fmaf(e.y, D.x, S2.y));
                }
            }
        }
        if (!grad) continue;
        const float2 ga  = make_float2(gsc * S1.x, gsc * S1.y);
        const float2 gam = make_float2(gsc * S2.x, -gsc * S2.y);
        const float2 gb  = sc.conj_mode ? gam : make_float2(gsc * S2.x, gsc * S2.y);
        const float2 gbm = sc.conj_mode ? ga : make_float2(gsc * S1.x, -gsc * S1.y);
        if (store_grad_h0) {
            m.grad_h0[(size_t)ty * N + tx] = make_float4(ga.x, ga.y, gb.x, gb.y);
            if (!self) m.grad_h0[(size_t)my * N + mx] = make_float4(gam.x, gam.y, gbm.x, gbm.y);
        }
        // dL/ds: the texels' gradients against their unit noise; for a pair off
        // the edges, ga.a + gb.a' + gam.b + gbm.b' = 2 gsc (S1.P~ + S2.Q~)
        if (gu.live) {
            const float dl = edge ? dl_ds(ga, gb, a)
                                  : 2.0f * gsc * (S1.x * a.x + S1.y * a.y + S2.x * a.z + S2.y * a.w);
            param_add(gu, dl, dv_scale, da_scale, dV, dT, dA);
        }
        if (self || paired) continue;
        if (gm.live) param_add(gm, dl_ds(gam, gbm, b), dv_scale, da_scale, dV, dT, dA);
    }
    lsum = block_sum(lsum, red);
    if (threadIdx.x == 0) m.loss_part[item] = lsum;
    if (grad) {
        dV = block_sum(dV, red);
        dT = block_sum(dT, red);
        dA = block_sum(dA, red);
        if (threadIdx.x == 0) {
            m.grad_part[3 * item + 0] = dV;
            m.grad_part[3 * item + 1] = dT;
            m.grad_part[3 * item + 2] = dA;
        }
    }
}

// T^ of one target frame, rows: R(t, y) = sum_x (-1)^(x+y) T(x, y) e^{-2 pi i t x / N}
// (the conjugate of the inverse transform of the real input), one row per
// block of N/8 threads; R in the member's Z buffer.
template <int N>
__device__ __forceinline__ void fit_targets_row(const float* __restrict__ tgt, float2* __restrict__ R,
                                                const float2* __restrict__ tw8, float2* lds)
{
    constexpr int E = 8, T = N / E;
    using Sync = typename std::conditional<(T <= 64), WaveSync, BlockSync>::type;
    const int y = blockIdx.x, j = threadIdx.x;
    float2    v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int x = j + r * T;
        const float t = tgt[(size_t)y * N + x];
        v[r] = make_float2(((x + y) & 1) ? -t : t, 0.0f);
    }
    fft_inverse<N, E>(v, j, lds, tw8, Sync{});
#pragma unroll
    for (int r = 0; r < E; ++r) R[(size_t)y * N + j + r * T] = make_float2(v[r].x, -v[r].y);
}

// ... and columns t = 0..N/2 of R: T^(t, s) = sum_y R(t, y) e^{-2 pi i s y / N},
// stored at the half-plane entries (strip t / C, c = t % C, row s).
template <int N>
__device__ __forceinline__ void fit_targets_col(const float2* __restrict__ R, float2* __restrict__ that,
                                                const float2* __restrict__ tw8, float2* lds)
{
    constexpr int E = 8, T = N / E, C = fit_cols_per_block(N);
    using Sync = typename std::conditional<(T <= 64), WaveSync, BlockSync>::type;
    const int t = blockIdx.x, j = threadIdx.x;  // t <= N/2
    float2    v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const float2 q = R[(size_t)(j + r * T) * N + t];
        v[r] = make_float2(q.x, -q.y);
    }
    fft_inverse<N, E>(v, j, lds, tw8, Sync{});
    const size_t col = (size_t)(t / C) * (N * C) + t % C;
#pragma unroll
    for (int r = 0; r < E; ++r) that[col + (size_t)(j + r * T) * C] = make_float2(v[r].x, -v[r].y);
}

}  // namespace

template <int N, int R>
__device__ __forceinline__ void fit_spectral_parts(const FitMember& m, int item, int mode, float* red)
{
    if (m.uniform) fit_spectral<N, true, R>(m, item, mode, red);
    else fit_spectral<N, false, R>(m, item, mode, red);
}

#define OCN_FIT_SWITCH(N_, CALL)                 \
    switch (N_) {                                \
    case 64: CALL(64); break;                    \
    case 128: CALL(128); break;                  \
    case 256: CALL(256); break;                  \
    case 512: CALL(512); break;                  \
    case 1024: CALL(1024); break;                \
    default: break;                              \
    }

__global__ __launch_bounds__(kFitColsThreads) void k_fit_spectral(const FitMember* __restrict__ members,
                                                                  const int2* __restrict__ items, int mode)
{
    __shared__ float red[kFitColsThreads / 64];
    const int2       it = items[xcd_logical(blockIdx.x, gridDim.x)];
    const FitMember& m  = members[it.x];
#define OCN_CALL(n) \
    if (m.row_parts == 4) fit_spectral_parts<n, 4>(m, it.y, mode, red); \
    else if (m.row_parts == 2) fit_spectral_parts<n, 2>(m, it.y, mode, red); \
    else fit_spectral_parts<n, 1>(m, it.y, mode, red)
    OCN_FIT_SWITCH(m.N, OCN_CALL)
#undef OCN_CALL
}

template <int N>
__global__ __launch_bounds__(N / 8) void k_fit_targets_rows(const float* __restrict__ tgt, float2* __restrict__ R,
                                                            const float2* __restrict__ tw8)
{
    __shared__ float2 lds[lds_len(N)];
    fit_targets_row<N>(tgt, R, tw8, lds);
}

template <int N>
__global__ __launch_bounds__(N / 8) void k_fit_targets_cols(const float2* __restrict__ R, float2* __restrict__ that,
                                                            const float2* __restrict__ tw8)
{
    __shared__ float2 lds[lds_len(N)];
    fit_targets_col<N>(R, that, tw8, lds);
}

__global__ __launch_bounds__(256) void k_fit_xi(float4* __restrict__ xi, int N, uint32_t key0, uint32_t key1,
                                                uint32_t ocean_id, uint32_t cascade_id)
{
    const int tx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ty = blockIdx.y;
    if (tx >= N) return;
    const uint32_t idx = (uint32_t)ty * (uint32_t)N + (uint32_t)tx;
    const uint4    r   = philox4x32_10(make_uint4(idx, ocean_id, cascade_id, 0u), make_uint2(key0, key1));
    xi[fit_xi_index(N, tx, ty)] = box_muller(u01(r.x), u01(r.y), u01(r.z), u01(r.w));
}

// One quasi-Newton step of one member (a single thread, FP64), after the loss
// L and gradient g at the current variables x were reduced.  BFGS on the
// inverse Hessian with a trust region (Nocedal & Wright, Numerical
// Optimization, 2nd ed., eq. 6.17 and 6.20), in scaled coordinates
// z = x / scale:
//   - first evaluation: x becomes the anchor, H = I, radius = cfg.radius;
//   - L above the anchor's loss (or not finite): the trial is rejected and the
//     radius shrinks to a quarter of the failed step;
//   - otherwise the trial is accepted: BFGS update with s = z - z_anchor,
//     y = g - g_anchor when s.y > 0 (H scaled by s.y / y.y before the first
//     update), the radius doubles (up to cfg.radius_max) after a step that
//     reached it, and the trial becomes the anchor;
// then the next trial is z_anchor + p, p = -H g_anchor cut to the radius.
// Wind speed and amplitude stay positive (floors 1e-3 and 1e-6).
__device__ void qn_step(FitVars& v, double L, const double g[3], const QnCfg& cfg)
{
    double z[3], gz[3];
    z[0] = v.V / cfg.scale[0];
    z[1] = v.theta / cfg.scale[1];
    z[2] = v.A / cfg.scale[2];
    for (int i = 0; i < 3; ++i) gz[i] = g[i] * cfg.scale[i];
    auto H = [&](int i, int j) -> double& {
        const int a = i < j ? i : j, b = i < j ? j : i;
        return v.qH[a == 0 ? b : a == 1 ? 2 + b : 5];
    };
    bool anchor = false;
    if (v.qn == 0) {
        for (int i = 0; i < 6; ++i) v.qH[i] = (i == 0 || i == 3 || i == 5) ? 1.0 : 0.0;
        v.qr   = cfg.radius;
        v.qk   = 0;
        anchor = true;
    } else if (!(L <= v.qL)) {
        double d2 = 0.0;
        for (int i = 0; i < 3; ++i) d2 += (z[i] - v.qz[i]) * (z[i] - v.qz[i]);
        v.qr = 0.25 * sqrt(d2);
    } else {
        double s[3], y[3], sy = 0.0, ss = 0.0, yy = 0.0;
        for (int i = 0; i < 3; ++i) {
            s[i] = z[i] - v.qz[i];
            y[i] = gz[i] - v.qg[i];
            sy += s[i] * y[i];
            ss += s[i] * s[i];
            yy += y[i] * y[i];
        }
        if (sy > 1e-12 * sqrt(ss * yy)) {
            if (v.qk == 0)
                for (int i = 0; i < 6; ++i) v.qH[i] *= sy / yy;
            // H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T
            const double rho = 1.0 / sy;
            double       Hy[3], yHy = 0.0;
            for (int i = 0; i < 3; ++i) {
                Hy[i] = 0.0;
                for (int j = 0; j < 3; ++j) Hy[i] += H(i, j) * y[j];
                yHy += y[i] * Hy[i];
            }
            double nh[6];
            for (int i = 0, k = 0; i < 3; ++i)
                for (int j = i; j < 3; ++j, ++k)
                    nh[k] = H(i, j) - rho * (s[i] * Hy[j] + Hy[i] * s[j]) + (rho * rho * yHy + rho) * s[i] * s[j];
            for (int k = 0; k < 6; ++k) v.qH[k] = nh[k];
            v.qk += 1;
        }
        if (sqrt(ss) >= 0.99 * v.qr) v.qr = fmin(2.0 * v.qr, cfg.radius_max);
        anchor = true;
    }
    if (anchor) {
        for (int i = 0; i < 3; ++i) {
            v.qz[i] = z[i];
            v.qg[i] = gz[i];
        }
        v.qL = L;
    }
    double p[3], pp = 0.0;
    for (int i = 0; i < 3; ++i) {
        p[i] = 0.0;
        for (int j = 0; j < 3; ++j) p[i] -= H(i, j) * v.qg[j];
        pp += p[i] * p[i];
    }
    const double f = pp > v.qr * v.qr ? v.qr / sqrt(pp) : 1.0;
    v.V     = fmax((v.qz[0] + f * p[0]) * cfg.scale[0], 1e-3);
    v.theta = (v.qz[1] + f * p[1]) * cfg.scale[1];
    v.A     = fmax((v.qz[2] + f * p[2]) * cfg.scale[2], 1e-6);
    v.qn += 1;
}

// One block per member: loss and gradient from the partial sums (FP64), then
// optionally one Adam step on (V, theta, A) (Kingma & Ba 2015, bias-corrected).
// V and A stay positive (floors 1e-3 and 1e-6).
__global__ __launch_bounds__(kFitReduceThreads) void k_fit_reduce(const FitMember* __restrict__ members, int mode,
                                                                  AdamCfg cfg, QnCfg qcfg)
{
    constexpr int     NW = kFitReduceThreads / 64;
    __shared__ double red[4 * NW];
    const FitMember&  m = members[blockIdx.x];
    // thread 0 loads the member's variables before the partial sums, so the
    // two loads' latencies overlap
    FitVars v;
    if (threadIdx.x == 0) v = *m.vars;
    double t[4] = {0.0, 0.0, 0.0, 0.0};  // loss, dL/dV, dL/dtheta, dL/dA
    for (int i = threadIdx.x; i < m.n_loss_part; i += blockDim.x) t[0] += (double)m.loss_part[i];
    if (mode != kFitLoss)
        for (int i = threadIdx.x; i < m.n_grad_part; i += blockDim.x) {
            t[1] += (double)m.grad_part[3 * i + 0];
            t[2] += (double)m.grad_part[3 * i + 1];
            t[3] += (double)m.grad_part[3 * i + 2];
        }
    // the four sums in one pass: across each wave, then over the waves in order
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < 4; ++j) t[j] += __shfl_xor(t[j], o);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[4 * (threadIdx.x / 64) + j] = t[j];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int j = 0; j < 4; ++j) sum[j] += red[4 * w + j];
    const double L = sum[0] / ((double)m.K * (double)m.N * (double)m.N);
    if (mode == kFitLoss) {
        m.vars->loss = L;
        return;
    }
    const double g[3] = {sum[1], sum[2], sum[3]};
    v.loss = L;
    for (int i = 0; i < 3; ++i) v.grad[i] = g[i];
    if (mode == kFitQn) {
        qn_step(v, L, g, qcfg);
    } else if (mode == kFitAdam) {
        v.step += 1;
        const double b1t = 1.0 - pow(cfg.beta1, (double)v.step), b2t = 1.0 - pow(cfg.beta2, (double)v.step);
        double* x[3] = {&v.V, &v.theta, &v.A};
        for (int i = 0; i < 3; ++i) {
            v.m[i] = cfg.beta1 * v.m[i] + (1.0 - cfg.beta1) * g[i];
            v.v[i] = cfg.beta2 * v.v[i] + (1.0 - cfg.beta2) * g[i] * g[i];
            *x[i] -= cfg.lr[i] * (v.m[i] / b1t) / (sqrt(v.v[i] / b2t) + cfg.eps);
        }
        v.V = fmax(v.V, 1e-3);
        v.A = fmax(v.A, 1e-6);
    }
    fit_vars_derive(v, m.g_d);
    *m.vars = v;
}

// The phase table of a member with equally spaced times: per column item and
// texel pair the start and step of the recurrence the step kernel takes
// (the same omega_fp64 and PhaseWalk::start it would evaluate).
__global__ __launch_bounds__(256) void k_fit_phase(FitMember m)
{
    const int    C = fit_cols_per_block(m.N), N = m.N;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)fit_cols_items_per_member(N) * N * C) return;
    const int strip = (int)(e / ((size_t)N * C)), k = (int)(e % ((size_t)N * C));
    const int c = k % C, ty = k / C;
    const int tx = min(strip * C + c, N / 2);
    const int nx = tx - N / 2, ny = ty - N / 2;
    SpecConst sc{};
    sc.kscale = m.kscale;
    sc.g      = m.g;
    sc.k2_d   = m.k2_d;
    sc.g2_d   = m.g2_d;
    float     kinv;
    PhaseWalk walk;
    walk.start(omega_fp64(sc, nx * nx + ny * ny, kinv), m.t_rev[0], m.dt_rev);
    const_cast<float4*>(m.phase)[e] = make_float4(walk.cur.x, walk.cur.y, walk.step.x, walk.step.y);
}

hipError_t launch_fit_phase(const FitMember& m, hipStream_t s)
{
    if (!m.uniform || !m.phase) return hipErrorInvalidValue;
    const size_t n = fit_phase_entries(m.N);
    hipLaunchKernelGGL(k_fit_phase, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m);
    return hipGetLastError();
}

// The noise coefficients of each half-plane texel pair, in the phase table's
// order: P~ = (a + conj(C b'))/2 and Q~ = (C a' + conj b)/2 with a = (xi.x, xi.y),
// a' = (xi.z, xi.w) of the texel and b, b' of its mirror, C the member's
// conjugation of h0mk; and, for the pairs on row 0 and in the columns 0 and
// N/2, both texels' unit noise (fit_edge_index).  Made once at creation.
__global__ __launch_bounds__(256) void k_fit_pq(FitMember m)
{
    const int    C = fit_cols_per_block(m.N), N = m.N;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= fit_phase_entries(N)) return;
    const int strip = (int)(e / ((size_t)N * C)), k = (int)(e % ((size_t)N * C));
    const int c = k % C, ty = k / C, tx = strip * C + c;
    float4*   pq = const_cast<float4*>(m.pq);
    if (tx > N / 2) {
        pq[e] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const int    mx = (N - tx) & (N - 1), my = (N - ty) & (N - 1);
    const float4 a = m.xi[fit_xi_index(N, tx, ty)], b = m.xi[fit_xi_index(N, mx, my)];
    const float  aw = m.conj_mode ? -a.w : a.w, bw = m.conj_mode ? -b.w : b.w;
    pq[e] = make_float4(0.5f * (a.x + b.z), 0.5f * (a.y - bw), 0.5f * (a.z + b.x), 0.5f * (aw - b.y));
    if (tx == 0 || tx == N / 2 || ty == 0) {
        float4*   edge = const_cast<float4*>(m.edge);
        const int ei   = fit_edge_index(N, tx, ty);
        edge[2 * ei]     = a;
        edge[2 * ei + 1] = b;
    }
}

hipError_t launch_fit_pq(const FitMember& m, hipStream_t s)
{
    if (!m.pq || !m.edge) return hipErrorInvalidValue;
    const size_t n = fit_phase_entries(m.N);
    hipLaunchKernelGGL(k_fit_pq, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m);
    return hipGetLastError();
}

template <int N>
static hipError_t launch_fit_targets_n(const FitMember& m, hipStream_t s)
{
    const size_t nn = (size_t)N * N, ne = fit_phase_entries(N);
    for (int f = 0; f < m.K; ++f) {
        hipLaunchKernelGGL(k_fit_targets_rows<N>, dim3(N), dim3(N / 8), 0, s, m.targets + f * nn, m.scratch, m.tw8);
        hipLaunchKernelGGL(k_fit_targets_cols<N>, dim3(N / 2 + 1), dim3(N / 8), 0, s, m.scratch, m.that + f * ne, m.tw8);
    }
    return hipGetLastError();
}

hipError_t launch_fit_targets(const FitMember& m, hipStream_t s)
{
#define OCN_CALL(n) return launch_fit_targets_n<n>(m, s)
    OCN_FIT_SWITCH(m.N, OCN_CALL)
#undef OCN_CALL
    return hipErrorInvalidValue;
}

hipError_t launch_fit_xi(const FitMember& m, hipStream_t s)
{
    dim3 grid((m.N + 255) / 256, m.N);
    hipLaunchKernelGGL(k_fit_xi, grid, dim3(256), 0, s, const_cast<float4*>(m.xi), m.N, m.key0, m.key1, m.ocean_id,
                       m.cascade_id);
    return hipGetLastError();
}

hipError_t launch_fit(const FitMember* members, int n_members, const int2* cols_items, int n_cols, int mode,
                      const AdamCfg& adam, const QnCfg& qn, hipStream_t s)
{
    if (n_members < 1 || n_cols < 1) return hipErrorInvalidValue;
    // one launch over the column items: loss and gradient in the spectral domain
    hipLaunchKernelGGL(k_fit_spectral, dim3(n_cols), dim3(kFitColsThreads), 0, s, members, cols_items, mode);
    hipLaunchKernelGGL(k_fit_reduce, dim3(n_members), dim3(kFitReduceThreads), 0, s, members, mode, adam, qn);
    return hipGetLastError();
}

}  // namespace ocn
