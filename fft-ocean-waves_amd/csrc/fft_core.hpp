// fft_core.hpp -- register/LDS Stockham inverse FFT building blocks for gfx950.
//
// One length-N inverse DFT (sign +, unnormalised: X[m] = sum_n x[n] e^{+2 pi i n m / N})
// is computed by a group of T = N / E threads.  Thread j of the group holds E
// complex values in registers, element r being index j + r*T.  The transform
// runs as a sequence of Stockham stages (radix E, then one smaller radix R_last
// when N is not a power of E); between stages the group exchanges data through
// an LDS buffer of padded length lds_len(N).  After the last stage, register r
// again holds output index j + r*T (natural order), so loads and stores of the
// same row are coalesced across consecutive j.
//
// This replaces the reference's 2*log2(N) radix-2 butterfly dispatches per
// field (butterfly_cs.glsl:54-144, driven from main.cpp:612-664) and its
// bit-reversal / twiddle precompute (main.cpp:29-64, twiddle_factors_cs.glsl);
// the result is the same unnormalised inverse DFT.
#pragma once

#include <hip/hip_runtime.h>

namespace ocn {

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }
__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// Stage plan of an (N, E) transform: `full` stages of radix E, then one stage
// of radix `last` (1 = none).
template <int N, int E>
struct Plan {
    static_assert((N & (N - 1)) == 0 && (E & (E - 1)) == 0, "powers of two only");
    static_assert(N >= E, "N must be at least E");
    static constexpr int T     = N / E;
    static constexpr int full  = ilog2(N) / ilog2(E);
    static constexpr int last  = N / ipow(E, full);
    static constexpr int nstages = full + (last > 1 ? 1 : 0);
};

// Number of float2 twiddles the table of an (N, E) plan holds.
__host__ __device__ constexpr int twiddle_count_rec(int N, int E, int ns)
{
    // stage with stride ns and radix R = min(E, N/ns) contributes ns*(R-1) entries (ns > 1)
    return ns >= N ? 0
                   : ((ns > 1 ? ns * ((N / ns < E ? N / ns : E) - 1) : 0) +
                      twiddle_count_rec(N, E, ns * (N / ns < E ? N / ns : E)));
}
template <int N, int E>
__host__ __device__ constexpr int twiddle_count() { return twiddle_count_rec(N, E, 1); }

// Offset of the stage with stride `ns` in the twiddle table.
__host__ __device__ constexpr int twiddle_offset_rec(int N, int E, int ns_cur, int ns_target)
{
    return ns_cur >= ns_target ? 0
                               : ((ns_cur > 1 ? ns_cur * ((N / ns_cur < E ? N / ns_cur : E) - 1) : 0) +
                                  twiddle_offset_rec(N, E, ns_cur * (N / ns_cur < E ? N / ns_cur : E),
                                                     ns_target));
}

// LDS length (float2) of one transform's exchange buffer: one pad slot per 16
// entries spreads the strided Stockham writes over the LDS banks, and the odd
// length makes neighbouring buffers of a block start in different banks (with
// a length of a multiple of 32 float2, the same element of every buffer sits
// in one bank, and threads that walk buffers side by side conflict).
__host__ __device__ constexpr int lds_len(int N) { return N + N / 16 + 1; }
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }

// ------------------------------------------------------------------ complex helpers

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }     // * i
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }    // * -i

// a * e^{2 pi i K / 16}, K a compile-time constant; the multiples of pi/4 are
// done without general multiplies.
template <int K>
__device__ __forceinline__ float2 tw16(float2 a)
{
    constexpr float C1 = 0.92387953251128675613f;  // cos(pi/8)
    constexpr float S1 = 0.38268343236508977173f;  // sin(pi/8)
    constexpr float H  = 0.70710678118654752440f;  // sqrt(2)/2
    constexpr int k = K & 15;
    if constexpr (k == 0) return a;
    else if constexpr (k == 4) return mul_i(a);
    else if constexpr (k == 8) return make_float2(-a.x, -a.y);
    else if constexpr (k == 12) return mul_mi(a);
    else if constexpr (k == 2) return make_float2(H * (a.x - a.y), H * (a.x + a.y));
    else if constexpr (k == 6) return make_float2(-H * (a.x + a.y), H * (a.x - a.y));
    else if constexpr (k == 10) return make_float2(H * (a.y - a.x), -H * (a.x + a.y));
    else if constexpr (k == 14) return make_float2(H * (a.x + a.y), H * (a.y - a.x));
    else {
        constexpr float c = (k == 1 || k == 15) ? C1 : (k == 3 || k == 13) ? S1 : (k == 5 || k == 11) ? -S1 : -C1;
        constexpr float s = (k == 1 || k == 7) ? S1 : (k == 3 || k == 5) ? C1 : (k == 9 || k == 15) ? -S1 : -C1;
        return cmul(a, make_float2(c, s));
    }
}

// ------------------------------------------------------------------ in-register DFTs
// Inverse DFT of R values held in v[S*0], v[S*1], ..., v[S*(R-1)] (stride S),
// output in natural order in the same slots.  Radix-2 decimation in time,
// expanded at compile time.

template <int R, int S, int OFF, int E>
struct RegDft {
    __device__ __forceinline__ static void run(float2 (&v)[E])
    {
        // even part lives at OFF + 2*S*i, odd part at OFF + S + 2*S*i
        RegDft<R / 2, 2 * S, OFF, E>::run(v);
        RegDft<R / 2, 2 * S, OFF + S, E>::run(v);
        float2 t[R];
#pragma unroll
        for (int m = 0; m < R / 2; ++m) {
            t[m]         = v[OFF + 2 * S * m];
            t[m + R / 2] = v[OFF + S + 2 * S * m];
        }
        combine<0>(v, t);
    }
    template <int M>
    __device__ __forceinline__ static void combine(float2 (&v)[E], float2 (&t)[R])
    {
        if constexpr (M < R / 2) {
            float2 e = t[M];
            float2 o = tw16<M * (16 / R)>(t[M + R / 2]);
            v[OFF + S * M]           = cadd(e, o);
            v[OFF + S * (M + R / 2)] = csub(e, o);
            combine<M + 1>(v, t);
        }
    }
};
template <int S, int OFF, int E>
struct RegDft<1, S, OFF, E> {
    __device__ __forceinline__ static void run(float2 (&)[E]) {}
};

// Inverse DFT of a[0..R) in place.
template <int R>
__device__ __forceinline__ void dft_inv(float2 (&a)[R])
{
    static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 16, "radix 2..16");
    RegDft<R, 1, 0, R>::run(a);
}

// ------------------------------------------------------------------ Stockham transform

struct BlockSync {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
};

// For transforms whose thread group lies inside one wavefront (T <= 64 and
// groups aligned to the wave): a wave's LDS operations execute in program
// order, so ordering only has to hold in the compiler; the fences also make
// it wait for this wave's outstanding LDS accesses.
struct WaveSync {
    __device__ __forceinline__ void operator()() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// One Stockham stage of radix R with stride NS on the E registers of thread j.
// Virtual butterfly b (0 <= b < E/R) has index jj = j + b*T and owns registers
// b + r*(E/R).  tw points at this stage's table: tw[k*(R-1) + r-1] =
// e^{+2 pi i r k / (NS*R)}.
template <int N, int E, int R, int NS>
__device__ __forceinline__ void stage_compute(float2 (&v)[E], int j, const float2* __restrict__ tw)
{
    constexpr int T = N / E;
    constexpr int B = E / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int jj = j + b * T;
        float2 a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[b + r * B];
        if constexpr (NS > 1) {
            const int k = jj & (NS - 1);
            const float2* w = tw + k * (R - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w[r - 1]);
        }
        dft_inv<R>(a);
#pragma unroll
        for (int r = 0; r < R; ++r) v[b + r * B] = a[r];
    }
}

// Scatter the stage outputs to their Stockham positions, then gather the next
// stage's inputs (index j + r*T).
template <int N, int E, int R, int NS, class Sync>
__device__ __forceinline__ void stage_exchange(float2 (&v)[E], int j, float2* lds, Sync sync)
{
    constexpr int T = N / E;
    constexpr int B = E / R;
    sync();  // previous readers of this buffer are done
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int jj   = j + b * T;
        const int k    = jj & (NS - 1);
        const int base = (jj / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lds_pad(base + r * NS)] = v[b + r * B];
    }
    sync();
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = lds[lds_pad(j + r * T)];
}

template <int N, int E, int NS, class Sync>
__device__ __forceinline__ void fft_stages(float2 (&v)[E], int j, float2* lds, const float2* __restrict__ tw,
                                           Sync sync)
{
    if constexpr (NS < N) {
        constexpr int R = (N / NS < E) ? N / NS : E;
        constexpr int off = twiddle_offset_rec(N, E, 1, NS);
        stage_compute<N, E, R, NS>(v, j, tw + off);
        if constexpr (NS * R < N) {
            stage_exchange<N, E, R, NS>(v, j, lds, sync);
            fft_stages<N, E, NS * R>(v, j, lds, tw, sync);
        }
    }
}

// Full inverse transform of one length-N sequence held by a T-thread group.
// On entry and exit register r of thread j holds index j + r*T.  The last
// stage leaves outputs in natural order without a final exchange because its
// scatter pattern is exactly index j + r*T.
template <int N, int E, class Sync>
__device__ __forceinline__ void fft_inverse(float2 (&v)[E], int j, float2* lds, const float2* __restrict__ tw,
                                            Sync sync)
{
    fft_stages<N, E, 1>(v, j, lds, tw, sync);
}

}  // namespace ocn
