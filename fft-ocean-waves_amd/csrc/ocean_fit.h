// ocean_fit.h -- internal interface of the differentiable-fit kernels
// (ocean_fit.hip) for the C ABI (ocean_capi.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ocn {

constexpr int kFitMaxK       = 16;    // frames (times) per fit
constexpr int kFitMinN       = 64;
constexpr int kFitMaxN       = 1024;
constexpr int kFitColsWidth  = 1024;  // C * N: columns per column block times N
constexpr int kFitColsThreads = kFitColsWidth / 4;  // threads of a column item: 4 texel pairs each
constexpr int kFitReduceThreads = 256;

// Fitted variables of one member, in device memory: written by the host
// (set) and by the reduction kernel (loss, gradient, Adam state), read by the
// step kernel of the next fit step.
struct FitVars {
    double V, theta, A;    // wind speed, wind angle (direction (cos, sin)), Phillips amplitude
    double m[3], v[3];     // Adam moments of (V, theta, A)
    double loss;           // last loss
    double grad[3];        // last dL/d(V, theta, A)
    int    step;           // Adam steps taken
    int    qn;             // quasi-Newton evaluations taken
    // quasi-Newton state, in scaled coordinates z_i = x_i / scale_i: the anchor
    // (the best point accepted so far), its loss and gradient, the inverse
    // Hessian approximation (zz: xx, xy, xz, yy, yz, zz), the trust radius
    double qz[3], qg[3], qL, qH[6], qr;
    int    qk;             // BFGS updates taken
    int    pad_;
    // Single-precision constants of the step kernel derived from (V, theta, A)
    // by fit_vars_derive, which every writer of the variables calls:
    // Lp^2 = (V^2/g)^2 and its reciprocal, A, the wind direction, the factors
    // of dL/dV and dL/dA (2 g^2 / V^5, 1 / (2A)) and sqrt(A)
    float  Lp2, iLp2, Af, wx, wy, dv_scale, da_scale, sqA;
};

__host__ __device__ inline void fit_vars_derive(FitVars& v, double g)
{
    const double Lp = v.V * v.V / g;
    v.Lp2      = (float)(Lp * Lp);
    v.iLp2     = 1.0f / v.Lp2;
    v.Af       = (float)v.A;
    v.wx       = (float)cos(v.theta);
    v.wy       = (float)sin(v.theta);
    v.dv_scale = (float)(2.0 * g * g / (v.V * v.V * v.V * v.V * v.V));
    v.da_scale = (float)(0.5 / v.A);
    v.sqA      = (float)sqrt(v.A);
}

struct AdamCfg {
    double lr[3];
    double beta1, beta2, eps;
};

// Quasi-Newton (BFGS with a trust region) settings: the scales of the
// variables and the initial and largest trust radius in scaled units.
struct QnCfg {
    double scale[3];
    double radius, radius_max;
};

// Device-side descriptor of one fit member (one ocean, K frames).
struct FitMember {
    int     N, K;
    int     conj_mode;
    float   kscale;         // 2 pi / L
    float   l2;             // suppression^2
    float   g;
    float   inv_knn;        // 1 / (K N^2)
    double  k2_d, g2_d;     // (2 pi / L)^2, g^2
    double  g_d;            // g
    double  t_rev[kFitMaxK];  // t_k / (2 pi)
    double  dt_rev;         // (t_{k+1} - t_k) / (2 pi) when `uniform`
    int     uniform;        // 1: the times are equally spaced (phases by recurrence)
    int     row_parts;      // blocks per column strip (1, 2 or 4), each over a contiguous part of its texel pairs
    const float4* xi;       // unit-amplitude noise, N*N
    // Per half-plane texel pair, in the phase table's order: the noise
    // coefficients (P~, Q~) of the fold (launch_fit_pq); and, for the pairs on
    // row 0 and in the columns 0 and N/2, both texels' unit noise, two float4
    // per pair at fit_edge_index.
    const float4* pq;
    const float4* edge;
    // Equally spaced times: per column item and texel pair of its strip, the
    // dispersion phase e^{i w t_0} and its step e^{i w dt} as (cos, sin,
    // cos, sin), written once at creation (launch_fit_phase); entry
    // (strip * N + ty) * C + c.  Null for other time sets.
    const float4* phase;
    const float2* tw8;      // Stockham twiddles (N, 8) and (N, 16)
    const float2* tw16;
    float2* scratch;        // N*N: the row transforms of a target frame (launch_fit_targets)
    const float* targets;   // K * N*N
    // Per target frame k, the sign-shifted 2-D DFT of the targets,
    // T^_k(n) = sum_x T_k(x) e^{-2 pi i n.x / N} over centred wave numbers n,
    // at the half-plane texels of the column items: entry
    // k * fit_phase_entries(N) + (strip * N + ty) * C + c (the phase table's
    // order).  Written by launch_fit_targets whenever the targets change.
    float2* that;
    float4* grad_h0;        // N*N: dL/dh0 as (dRe h0k, dIm h0k, dRe h0mk, dIm h0mk)
    float*  loss_part;      // one partial sum per column item and row part
    float*  grad_part;      // three partial sums per column item and row part
    int     n_loss_part, n_grad_part;
    FitVars* vars;
    uint32_t key0, key1, ocean_id, cascade_id;  // Philox noise of xi
};

__host__ __device__ constexpr int fit_cols_per_block(int N) { return kFitColsWidth / N; }
__host__ __device__ inline int fit_cols_items_per_member(int N) { return (N / 2) / fit_cols_per_block(N) + 1; }
constexpr int kFitMaxRowParts = 4;  // a step block covers at least one texel pair per thread
// half-plane texel pairs on row 0 and in the columns 0 and N/2, and the index
// of one (tx <= N/2; ty == 0 or tx == 0 or tx == N/2)
__host__ __device__ constexpr int fit_edge_pairs(int N) { return N / 2 + 2 * N - 1; }
__host__ __device__ inline int fit_edge_index(int N, int tx, int ty)
{
    return ty == 0 ? tx : tx == 0 ? N / 2 + ty : N / 2 + N - 1 + ty;
}
// float4 entries of a member's phase table
__host__ __device__ inline size_t fit_phase_entries(int N) { return (size_t)fit_cols_items_per_member(N) * N * fit_cols_per_block(N); }

// Modes of the fit launch sequence.
enum FitMode { kFitLoss = 0, kFitGrad = 1, kFitAdam = 2, kFitQn = 3 };

hipError_t launch_fit_xi(const FitMember& m, hipStream_t s);
// T^ of every target frame of member m (m.scratch holds the row transforms).
hipError_t launch_fit_targets(const FitMember& m, hipStream_t s);
// Fills m.pq and m.edge from m.xi.
hipError_t launch_fit_pq(const FitMember& m, hipStream_t s);
// Fills m.phase (a member with equally spaced times) from its constants.
hipError_t launch_fit_phase(const FitMember& m, hipStream_t s);
// cols_items: int2 (member, strip * row_parts + part): one block each for the step kernel
hipError_t launch_fit(const FitMember* members, int n_members, const int2* cols_items, int n_cols, int mode,
                      const AdamCfg& adam, const QnCfg& qn, hipStream_t s);

}  // namespace ocn
