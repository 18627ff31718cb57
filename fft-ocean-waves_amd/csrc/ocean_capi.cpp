// ocean_capi.cpp -- the extern "C" boundary (include/ocean_c.h) over the
// gfx950 kernels.  Host-side validation, the handle's device buffers, FFT
// plans, schedules (with the autotuner) and launch sequencing.  Citations
// name the reference code each entry point replaces (diharaw/fft-ocean-waves,
// src/main.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ocean_c.h"
#include "ocean_batch.h"
#include "ocean_fit.h"
#include <rocprofiler-sdk-roctx/roctx.h>

#include "ocean_kernels.h"

namespace {

// Host-side trace ranges (roctx) around the ABI's setup, frame, batch, fit and
// autotune calls, shown by rocprofv3 --marker-trace on the host timeline; a
// range costs a few tens of nanoseconds when no tool is attached.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&)            = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

thread_local std::string g_err;

int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int status, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return status;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(OCN_EHIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

int validate(const ocn_params* p)
{
    if (!p) return fail(OCN_EINVAL, "params is NULL");
    if (!is_pow2(p->N) || p->N < 64 || p->N > 4096)
        return fail(OCN_EINVAL, "N=%d must be a power of two in [64, 4096]", p->N);
    if (!(std::isfinite(p->L) && p->L > 0.0)) return fail(OCN_EINVAL, "L must be finite and > 0");
    if (!(std::isfinite(p->wind_speed) && p->wind_speed >= 0.0))
        return fail(OCN_EINVAL, "wind_speed must be finite and >= 0");
    if (!(std::isfinite(p->wind_dir_x) && std::isfinite(p->wind_dir_y)) ||
        p->wind_dir_x * p->wind_dir_x + p->wind_dir_y * p->wind_dir_y <= 0.0)
        return fail(OCN_EINVAL, "wind direction must be finite and non-zero");
    if (!(std::isfinite(p->amplitude) && p->amplitude >= 0.0))
        return fail(OCN_EINVAL, "amplitude must be finite and >= 0");
    if (!(std::isfinite(p->suppression) && p->suppression >= 0.0))
        return fail(OCN_EINVAL, "suppression must be finite and >= 0");
    if (!std::isfinite(p->choppiness)) return fail(OCN_EINVAL, "choppiness must be finite");
    if (!(std::isfinite(p->gravity) && p->gravity > 0.0)) return fail(OCN_EINVAL, "gravity must be finite and > 0");
    if (!std::isfinite(p->displacement_scale)) return fail(OCN_EINVAL, "displacement_scale must be finite");
    if (p->conj_mode != OCN_CONJ_REFERENCE && p->conj_mode != OCN_CONJ_TEXTBOOK)
        return fail(OCN_EINVAL, "conj_mode must be OCN_CONJ_REFERENCE or OCN_CONJ_TEXTBOOK");
    if (p->normal_mode != OCN_NORMALS_NEAREST && p->normal_mode != OCN_NORMALS_BILINEAR)
        return fail(OCN_EINVAL, "normal_mode must be OCN_NORMALS_NEAREST or OCN_NORMALS_BILINEAR");
    return OCN_OK;
}

constexpr double kPi = 3.14159265358979323846264338327950288;

// Twiddle table of the (N, E) Stockham plan used by fft_core.hpp: for each
// stage with stride ns > 1 and radix R = min(E, N/ns), entries
// k*(R-1) + (r-1) = e^{+2 pi i r k / (ns R)}, evaluated in FP64, rounded once.
std::vector<float> stockham_twiddles(int N, int E)
{
    std::vector<float> t;
    for (int ns = 1; ns < N;) {
        const int R = (N / ns < E) ? N / ns : E;
        if (ns > 1)
            for (int k = 0; k < ns; ++k)
                for (int r = 1; r < R; ++r) {
                    const double a = 2.0 * kPi * (double)r * (double)k / ((double)ns * R);
                    t.push_back((float)std::cos(a));
                    t.push_back((float)std::sin(a));
                }
        ns *= R;
    }
    return t;
}

ocn::SpecConst spec_const(const ocn_params& p)
{
    ocn::SpecConst s{};
    const double n  = std::sqrt(p.wind_dir_x * p.wind_dir_x + p.wind_dir_y * p.wind_dir_y);
    const double Lp = p.wind_speed * p.wind_speed / p.gravity;
    const double ks = 2.0 * kPi / p.L;
    s.N          = p.N;
    s.kscale     = (float)ks;
    s.Lp2        = (float)(Lp * Lp);
    s.l2         = (float)(p.suppression * p.suppression);
    s.A          = (float)p.amplitude;
    s.wx         = (float)(p.wind_dir_x / n);
    s.wy         = (float)(p.wind_dir_y / n);
    s.g          = (float)p.gravity;
    s.key0       = (uint32_t)p.seed;
    s.key1       = (uint32_t)(p.seed >> 32);
    s.ocean_id   = p.ocean_id;
    s.cascade_id = p.cascade_id;
    s.conj_mode  = p.conj_mode;
    s.k2_d       = ks * ks;
    s.g2_d       = p.gravity * p.gravity;
    return s;
}

// Schedules the autotuner found, per (device, N, frames-per-launch class,
// normal mode): the fused row variants only run with nearest-tap normals.
std::mutex                                               g_sched_mu;
std::map<std::tuple<int, int, int, int>, std::pair<int, int>> g_sched;

int frames_class(int nframes)
{
    return nframes <= 1 ? 1 : nframes <= 4 ? 4 : nframes <= 8 ? 8 : nframes <= 16 ? 16 : ocn::kMaxFrames;
}

// Built-in schedules: the fastest (column variant, row variant) measured on
// MI355X per N for one frame per launch and for batches of frames; the
// autotuner may replace them.  Column variants: 0 (16,1) 1 (16,2) 2 (16,4)
// 3 (16,8) 4 (8,2) 5 (8,4) 6 (8,8) as (radix, columns per block) of k_cols,
// 7..9 k_cols_lds with 1, 2, 4 columns per block, 10..12 k_cols_lds2 with 2, 4,
// 8 columns per block, 13..14 k_cols_lds with 1, 2 columns and twice the
// threads, 15..16 k_cols_ldsb (N >= 2048) with 1, 2 columns.  Row variants:
// 0 radix 16,
// 1 radix 8 (each followed by the normal-map kernel), 2..5 the fused
// row + normals walk with bands of 4, 8, 16, 32 rows, 6..7 the fused
// parallel bands of 4 and 2 rows, 8 the parallel bands of 2 rows with twice
// the threads.  Single frames at N = 256..1024 take the parallel bands after
// the column strips with twice the threads (graph-timed sweep of every pair
// in one box: N=1024 31.7 -> 27.4 us, N=512 19.9 -> 12.7 us, N=256 11.9 ->
// 8.8 us against the earlier single-frame defaults).
std::pair<int, int> default_schedule(int N, int nframes)
{
    const bool batch = nframes > 1;
    // Multi-frame launches: the row pass's band height follows the frames per
    // launch (sweep of every pair at 2..32 frames per launch over a 600-frame
    // sequence).  The walks of one launch fill the GPU best at about four per
    // CU: a few frames need short (parallel) bands, 16-32 frames the long
    // walks.  Launches whose intermediates fit the 256 MB Infinity Cache
    // beside h0 (up to 20 N=1024 frames, 4 N=2048 frames) are the fastest per
    // frame, since the row pass then reads them from there.
    // Column strips over pairs of frames (variant 19, k_cols_lds_f2 with 4
    // columns: one h0 load and dispersion per texel pair for two frames) at
    // N <= 512 where they measured faster per frame over the sequence: N=512
    // from 8 frames per launch (8: 4.59 -> 4.41 us, 32: 3.31-3.46 -> 3.25),
    // N=256 from 16 (32: 1.09 -> 1.07), N=128 at 32 (0.73 -> 0.71), N=64 from 8.
    int cv, rv;
    switch (N) {
    case 64: cv = !batch ? 14 : nframes >= 8 ? 19 : 9; rv = batch ? 1 : 8; break;
    case 128: cv = !batch ? 14 : nframes >= 32 ? 19 : 9; rv = batch ? 2 : 8; break;
    case 256: cv = !batch ? 13 : nframes >= 16 ? 19 : 9; rv = 8; break;
    case 512: cv = !batch ? 14 : nframes <= 4 ? 9 : 19; rv = batch ? 7 : 8; break;
    case 1024:
        cv = batch ? 9 : 14;
        rv = !batch ? 7 : nframes <= 6 ? 7 : nframes <= 9 ? 3 : nframes <= 16 ? 4 : 5;
        break;
    // N=2048: LDS-staged block-synchronised columns for single frames and
    // launches of a few frames (single 96.6 -> 79.4 us, 2 frames 90.4 -> 83.7
    // us per frame), k_cols radix 16 with 4 columns for 32
    case 2048:
        cv = nframes < 32 ? 16 : 2;
        rv = !batch ? 2 : nframes <= 2 ? 3 : nframes <= 4 ? 4 : 5;
        break;
    // 4096: k_cols_ldsb (one 4096-point transform over four waves, block
    // barriers at the stage exchanges): 8-frame launches 394 -> 353 us per
    // frame, single frames 448 -> 406 us, against k_cols radix 16; from 4
    // frames per launch its two-pass form with two columns per strip
    // (k_cols_lds2b): 8-frame launches 356 -> 330 us per frame
    default: cv = nframes >= 4 ? 17 : 15; rv = batch ? 5 : 4; break;
    }
    if (!ocn::schedule_valid(N, cv, rv)) {
        cv = 0;
        rv = 0;
        while (cv < ocn::kNumColsVariants && !ocn::schedule_valid(N, cv, rv)) ++cv;
    }
    return {cv, rv};
}

}  // namespace

struct ocn_handle {
    ocn_params p{};
    int        device   = 0;
    bool       h0_ready = false;
    int        zcap     = 0;        // frames the intermediate buffers hold
    int        cols_variant = 0;
    int        rows_variant = 0;
    bool       tuned = false;       // schedule set by ocn_autotune / ocn_set_schedule
    float4*    h0 = nullptr;        // N*N
    float4*    h0t = nullptr;       // N*N, h0 transposed (N >= 2048: read by the column strips)
    float2*    z1 = nullptr;        // zcap*N*N: intermediate; frame 0 doubles as butterfly scratch
    float2*    z2 = nullptr;        // zcap*N*N/2
    float2*    tw[3] = {nullptr, nullptr, nullptr};  // Stockham twiddles, radix 4 / 8 / 16
    float2*    spec[3] = {nullptr, nullptr, nullptr};  // tilde_h0_t textures (lazy)
    uint8_t*   noise = nullptr;     // 4*256*256 (lazy)
    float*     height = nullptr;
    float*     dx = nullptr;
    float*     dz = nullptr;
    float*     jac = nullptr;
    float4*    normals = nullptr;
};

namespace {

void free_all(ocn_handle* h)
{
    void* ptrs[] = {h->h0, h->h0t, h->z1, h->z2, h->tw[0], h->tw[1], h->tw[2], h->spec[0], h->spec[1], h->spec[2],
                    h->noise, h->height, h->dx, h->dz, h->jac, h->normals};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
}

int check_handle(const ocn_handle* h)
{
    if (!h) return fail(OCN_EINVAL, "handle is NULL");
    return OCN_OK;
}

int set_device(const ocn_handle* h)
{
    hipError_t e = hipSetDevice(h->device);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipSetDevice");
}

template <class T>
int dalloc(T** p, size_t count, const char* what)
{
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(OCN_ENOMEM, "hipMalloc(%s, %zu bytes): %s", what, sizeof(T) * count, hipGetErrorString(e));
    }
    return OCN_OK;
}

int upload_twiddles(ocn_handle* h)
{
    const int radices[3] = {4, 8, 16};
    for (int i = 0; i < 3; ++i) {
        std::vector<float> t = stockham_twiddles(h->p.N, radices[i]);
        if (h->tw[i]) (void)hipFree(h->tw[i]);
        h->tw[i] = nullptr;
        int rc = dalloc(&h->tw[i], t.size() / 2 + 1, "twiddles");
        if (rc) return rc;
        if (!t.empty()) {
            hipError_t e = hipMemcpy(h->tw[i], t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpy(twiddles)");
        }
    }
    return OCN_OK;
}

// Sizes the intermediate buffers for launches of up to `frames` frames.  Only
// setup calls reach this (ocn_create, ocn_reserve_frames, ocn_autotune): the
// per-frame entry points never allocate or free, they split their frames into
// launches of at most zcap frames.  Growing waits for the device (the old
// buffers may still be in use by queued work) before freeing them.
int reserve_z(ocn_handle* h, int frames)
{
    if (frames <= h->zcap) return OCN_OK;
    const size_t nn = (size_t)h->p.N * h->p.N;
    float2 *z1 = nullptr, *z2 = nullptr;
    int rc;
    if ((rc = dalloc(&z1, nn * frames, "z1"))) return rc;
    if ((rc = dalloc(&z2, nn / 2 * frames, "z2"))) {
        (void)hipFree(z1);
        return rc;
    }
    if (h->z1 || h->z2) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            (void)hipFree(z1);
            (void)hipFree(z2);
            return hip_fail(e, "hipDeviceSynchronize (reserve)");
        }
        if (h->z1 && (e = hipFree(h->z1)) != hipSuccess) return hip_fail(e, "hipFree(z1)");
        if (h->z2 && (e = hipFree(h->z2)) != hipSuccess) return hip_fail(e, "hipFree(z2)");
    }
    h->z1   = z1;
    h->z2   = z2;
    h->zcap = frames;
    return OCN_OK;
}

void pick_schedule(ocn_handle* h, int nframes, int* cv, int* rv)
{
    if (h->tuned) {
        *cv = h->cols_variant;
        *rv = h->rows_variant;
        return;
    }
    {
        std::lock_guard<std::mutex> lk(g_sched_mu);
        auto it = g_sched.find({h->device, h->p.N, frames_class(nframes), h->p.normal_mode});
        if (it != g_sched.end()) {
            *cv = it->second.first;
            *rv = it->second.second;
            return;
        }
    }
    const auto d = default_schedule(h->p.N, nframes);
    *cv = d.first;
    *rv = d.second;
}

void frame_args(const ocn_handle* h, float2* z1, float2* z2, const double* t, int nframes, const ocn_fields* out,
                size_t out_stride, ocn::FrameArgs* a)
{
    const size_t nn = (size_t)h->p.N * h->p.N;
    a->h0 = h->h0;
    a->h0t = h->h0t;
    a->z1 = z1;
    a->z2 = z2;
    for (int i = 0; i < 3; ++i) a->tw[i] = h->tw[i];
    a->sc = spec_const(h->p);
    for (int f = 0; f < nframes; ++f) a->fb.t_rev[f] = t[f] / (2.0 * kPi);
    a->fb.z1_stride  = nn;
    a->fb.z2_stride  = nn / 2;
    a->fb.out_stride = out_stride;
    a->fb.nframes    = nframes;
    a->nframes       = nframes;
    a->height        = out->height;
    a->dx            = out->dx;
    a->dz            = out->dz;
    a->normals       = reinterpret_cast<float4*>(out->normals);
    a->jacobian      = out->jacobian;
    a->disp          = reinterpret_cast<float4*>(out->displacement);
    a->rnorm         = reinterpret_cast<float4*>(out->render_normals);
    a->normal_mode   = h->p.normal_mode;
    a->lambda        = (float)h->p.choppiness;
    a->dscale        = (float)h->p.displacement_scale;
    a->inv2d         = (float)((double)h->p.N / (2.0 * h->p.L));
}

int ready_for_frame(const ocn_handle* h, double t)
{
    if (!std::isfinite(t)) return fail(OCN_EINVAL, "t must be finite");
    if (!h->h0_ready) return fail(OCN_ESTATE, "the spectrum does not exist yet: call ocn_tilde_h0_k first");
    return OCN_OK;
}

int check_fields(const ocn_fields* out)
{
    if (!out) return fail(OCN_EINVAL, "fields is NULL");
    if (!out->height || !out->dx || !out->dz) return fail(OCN_EINVAL, "height, dx and dz are required");
    return OCN_OK;
}

// Launches `nframes` (<= kMaxFrames) frames with the given stages, with the
// intermediates at z1 / z2 (room for nframes frames).
int run_frames_z(ocn_handle* h, float2* z1, float2* z2, const double* t, int nframes, const ocn_fields* out,
                 size_t out_stride, int stages, int cv, int rv, hipStream_t s)
{
    ocn::FrameArgs a{};
    frame_args(h, z1, z2, t, nframes, out, out_stride, &a);
    a.cols_variant = cv;
    a.rows_variant = rv;
    hipError_t e = ocn::launch_frame(h->p.N, a, stages, s);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "update launch");
}

// The same on the handle's own intermediates.
int run_frames(ocn_handle* h, const double* t, int nframes, const ocn_fields* out, size_t out_stride, int stages,
               int cv, int rv, hipStream_t s)
{
    if (nframes > h->zcap) return fail(OCN_ESTATE, "internal: %d frames exceed the reserved %d", nframes, h->zcap);
    return run_frames_z(h, h->z1, h->z2, t, nframes, out, out_stride, stages, cv, rv, s);
}

// Frames t[0..nframes) into consecutive output slices in launches of at
// most `chunk` frames, intermediates at z1 / z2 (room for `chunk` frames).
// Each launch takes the schedule for its own frame count, so a short last
// launch gets the short bands that fill the GPU.
int frames_in_chunks(ocn_handle* h, float2* z1, float2* z2, const double* t_host, int nframes,
                     const ocn_fields* out, int stage, int chunk, hipStream_t s)
{
    const size_t nn = (size_t)h->p.N * h->p.N;
    int cv, rv, rc;
    for (int f0 = 0; f0 < nframes; f0 += chunk) {
        const int nf = std::min(chunk, nframes - f0);
        pick_schedule(h, nf, &cv, &rv);
        ocn_fields sl{out->height + f0 * nn,
                      out->dx + f0 * nn,
                      out->dz + f0 * nn,
                      out->normals ? out->normals + f0 * nn * 4 : nullptr,
                      out->jacobian ? out->jacobian + f0 * nn : nullptr,
                      out->displacement ? out->displacement + f0 * nn * 4 : nullptr,
                      out->render_normals ? out->render_normals + f0 * nn * 4 : nullptr};
        if ((rc = run_frames_z(h, z1, z2, t_host + f0, nf, &sl, nn, stage, cv, rv, s))) return rc;
    }
    return OCN_OK;
}

// Checks shared by the frame entry points (before any HIP call).
int check_frames(const ocn_handle* h, const double* t_host, int nframes, const ocn_fields* out)
{
    int rc = check_handle(h);
    if (rc || (rc = check_fields(out))) return rc;
    if (!t_host || nframes < 0) return fail(OCN_EINVAL, "t_host must be non-NULL and nframes >= 0");
    for (int f = 0; f < nframes; ++f)
        if ((rc = ready_for_frame(h, t_host[f]))) return rc;
    return OCN_OK;
}

}  // namespace

// ====================================================================== C ABI

// ====================================================================== batches

struct ocn_batch {
    int                                    device = 0;
    std::vector<ocn_params>                params;
    std::vector<ocn::BatchMember>          members;  // host copy of the device descriptors
    ocn::BatchMember*                      d_members = nullptr;
    int2*                                  d_cols    = nullptr;
    int2*                                  d_rows    = nullptr;
    int2*                                  d_wide    = nullptr;  // row items of the N = 4096 members
    int                                    n_cols    = 0;
    int                                    n_rows    = 0;
    int                                    n_wide    = 0;
    char*                                  pool      = nullptr;  // every member's buffers
    char*                                  render_pool = nullptr;    // render exports (on request)
    bool                                   render      = false;
    std::map<int, std::array<float2*, 2>>  tw;  // per N: radix-8 and radix-16 Stockham twiddles
    bool                                   h0_ready  = false;
    hipGraph_t                             graph     = nullptr;
    hipGraphExec_t                         exec      = nullptr;
    hipGraphNode_t                         node_cols = nullptr;
    hipGraphNode_t                         node_rows = nullptr;
    hipGraphNode_t                         node_wide = nullptr;
};

namespace {

void batch_free(ocn_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void* ptrs[] = {b->d_members, b->d_cols, b->d_rows, b->d_wide, b->pool, b->render_pool};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    for (auto& kv : b->tw)
        for (float2* q : kv.second)
            if (q) (void)hipFree(q);
}

int upload_table(float2** dst, const std::vector<float>& t)
{
    int rc = dalloc(dst, t.size() / 2 + 1, "twiddles");
    if (rc || t.empty()) return rc;
    hipError_t e = hipMemcpy(*dst, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipMemcpy(twiddles)");
}

template <class T>
int upload_vector(T** dst, const std::vector<T>& v, const char* what)
{
    int rc = dalloc(dst, v.size(), what);
    if (rc) return rc;
    hipError_t e = hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return e == hipSuccess ? OCN_OK : hip_fail(e, what);
}

// Order of work items for a launch whose blocks are spread over the 8 XCDs in
// contiguous ranges (xcd_logical): items sorted by decreasing cost and dealt
// round-robin, so every XCD's range gets a like share of the work and runs its
// largest items first; the last blocks of the launch are then the smallest.
template <class Cost>
std::vector<int2> balance_items(const std::vector<int2>& items, Cost cost)
{
    constexpr int kXcds = 8;
    std::vector<int2> sorted(items);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int2 a, int2 b) { return cost(a) > cost(b); });
    std::vector<std::vector<int2>> bins(kXcds);
    for (size_t i = 0; i < sorted.size(); ++i) bins[i % kXcds].push_back(sorted[i]);
    std::vector<int2> out;
    out.reserve(items.size());
    for (const auto& bin : bins) out.insert(out.end(), bin.begin(), bin.end());
    return out;
}

// Row work items from every member's band height: (member, first band) of
// the shared row launch, balanced over the XCDs, and (member, band) of the
// wide launch.  Replaces the device tables and drops a built graph (its row
// nodes change shape); no launch may be reading the old tables.
int batch_build_rows(ocn_batch* b)
{
    std::vector<int2> rows, wide;
    for (size_t i = 0; i < b->members.size(); ++i) {
        const int N = b->params[i].N, R = b->members[i].band, gpb = ocn::batch_rows_groups_per_block(N);
        for (int k = 0; k < ocn::batch_rows_items_per_member(N, R); ++k)
            (N == ocn::kBatchWideN ? wide : rows).push_back(make_int2((int)i, k * gpb));
    }
    rows = balance_items(rows, [&](int2 it) {
        const int N = b->params[it.x].N, R = b->members[it.x].band;
        const int bands = std::min(ocn::batch_rows_groups_per_block(N), N / R - it.y);
        return (double)bands * (R + 2) * N;  // rows transformed by the item
    });
    for (int2** d : {&b->d_rows, &b->d_wide})
        if (*d) {
            (void)hipFree(*d);
            *d = nullptr;
        }
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec  = nullptr;
        b->graph = nullptr;
    }
    b->n_rows = (int)rows.size();
    b->n_wide = (int)wide.size();
    int rc    = OCN_OK;
    if (!rows.empty()) rc = upload_vector(&b->d_rows, rows, "batch row items");
    if (!rc && !wide.empty()) rc = upload_vector(&b->d_wide, wide, "batch wide row items");
    return rc;
}

// Column work items from every member's strip width: (member, strip),
// member-major (a member's strips share Z lines).  Replaces the device table
// and drops a built graph; no launch may be reading the old table.
int batch_build_cols(ocn_batch* b)
{
    std::vector<int2> cols;
    for (size_t i = 0; i < b->members.size(); ++i)
        for (int k = 0; k < ocn::batch_cols_items_per_member(b->params[i].N, b->members[i].cols); ++k)
            cols.push_back(make_int2((int)i, k));
    if (b->d_cols) {
        (void)hipFree(b->d_cols);
        b->d_cols = nullptr;
    }
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec  = nullptr;
        b->graph = nullptr;
    }
    b->n_cols = (int)cols.size();
    return upload_vector(&b->d_cols, cols, "batch column items");
}

int check_batch(const ocn_batch* b)
{
    if (!b) return fail(OCN_EINVAL, "batch is NULL");
    return OCN_OK;
}

int batch_ready(const ocn_batch* b, double t)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!std::isfinite(t)) return fail(OCN_EINVAL, "t must be finite");
    if (!b->h0_ready) return fail(OCN_ESTATE, "the spectra do not exist yet: call ocn_batch_tilde_h0_k first");
    hipError_t e = hipSetDevice(b->device);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipSetDevice");
}

}  // namespace

// ====================================================================== fits

struct ocn_fit {
    int                                   device = 0;
    int                                   K      = 0;
    std::vector<ocn_params>               params;
    std::vector<ocn::FitMember>           members;  // host copy of the device descriptors
    ocn::FitMember*                       d_members = nullptr;
    ocn::FitVars*                         d_vars    = nullptr;
    int2*                                 d_cols    = nullptr;
    int                                   n_cols    = 0;
    char*                                 pool      = nullptr;
    std::map<int, std::array<float2*, 2>> tw;  // per N: radix-8 and radix-16 Stockham twiddles
    // Members whose targets may have changed since their spectra T^ were
    // made: set by ocn_fit_targets (which hands out the writable buffer) and
    // ocn_fit_upload_targets, cleared when the next fit call remakes T^.
    mutable std::vector<char>             targets_dirty;
};

namespace {

void fit_free(ocn_fit* f)
{
    void* ptrs[] = {f->d_members, f->d_vars, f->d_cols, f->pool};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    for (auto& kv : f->tw)
        for (float2* q : kv.second)
            if (q) (void)hipFree(q);
}

int check_fit(const ocn_fit* f, int member)
{
    if (!f) return fail(OCN_EINVAL, "fit is NULL");
    if (member < 0 || member >= (int)f->members.size())
        return fail(OCN_EINVAL, "member %d out of range [0, %zu)", member, f->members.size());
    return OCN_OK;
}

ocn::FitVars initial_vars(const ocn_params& p)
{
    ocn::FitVars v{};
    v.V     = p.wind_speed;
    v.theta = std::atan2(p.wind_dir_y, p.wind_dir_x);
    v.A     = p.amplitude;
    ocn::fit_vars_derive(v, p.gravity);
    return v;
}

int fit_run(ocn_fit* f, int mode, const ocn::AdamCfg& adam, const ocn::QnCfg& qn, void* stream)
{
    if (!f) return fail(OCN_EINVAL, "fit is NULL");
    hipError_t e = hipSetDevice(f->device);
    for (size_t i = 0; e == hipSuccess && i < f->members.size(); ++i)
        if (f->targets_dirty[i]) {
            e = ocn::launch_fit_targets(f->members[i], (hipStream_t)stream);
            if (e == hipSuccess) f->targets_dirty[i] = 0;
        }
    if (e == hipSuccess)
        e = ocn::launch_fit(f->d_members, (int)f->members.size(), f->d_cols, f->n_cols, mode, adam, qn,
                            (hipStream_t)stream);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "fit launch");
}

}  // namespace

// One imported memory object of a graphics API (ocn_external_import).
struct ocn_external {
    int                 device = 0;
    hipExternalMemory_t mem    = nullptr;
};

extern "C" {

int ocn_abi_version(void) { return OCN_ABI_VERSION; }

const char* ocn_last_error(void) { return g_err.c_str(); }

const char* ocn_status_string(int status)
{
    switch (status) {
    case OCN_OK: return "OCN_OK";
    case OCN_EINVAL: return "OCN_EINVAL";
    case OCN_ENOMEM: return "OCN_ENOMEM";
    case OCN_EHIP: return "OCN_EHIP";
    case OCN_ESTATE: return "OCN_ESTATE";
    case OCN_ENOTSUP: return "OCN_ENOTSUP";
    default: return "OCN_E(unknown)";
    }
}

int ocn_params_default(ocn_params* p)
{
    if (!p) return fail(OCN_EINVAL, "params is NULL");
    std::memset(p, 0, sizeof *p);
    p->N           = 256;    // DISPLACEMENT_MAP_SIZE (main.cpp:17, 1645)
    p->L           = 1000.0; // m_L (main.cpp:1646)
    p->wind_speed  = 80.0;   // m_wind_speed (main.cpp:1641)
    p->wind_dir_x  = 1.0;    // m_wind_direction (main.cpp:1644)
    p->wind_dir_y  = 1.0;
    p->amplitude   = 2.0;    // m_amplitude (main.cpp:1642)
    p->suppression = 0.1;    // m_suppression_factor (main.cpp:1643)
    p->choppiness  = 0.75;   // m_choppiness (main.cpp:1633)
    p->gravity     = 9.81;   // g (tilde_h0_k_cs.glsl:26)
    p->displacement_scale = 0.5;  // m_displacement_scale (main.cpp:1632)
    p->seed        = 42;
    p->conj_mode   = OCN_CONJ_REFERENCE;
    p->normal_mode = OCN_NORMALS_NEAREST;
    return OCN_OK;
}

int ocn_create(const ocn_params* p, ocn_handle** out)
{
    if (!out) return fail(OCN_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = validate(p);
    if (rc) return rc;
    ocn_handle* h = new (std::nothrow) ocn_handle();
    if (!h) return fail(OCN_ENOMEM, "host allocation of the handle failed");
    h->p = *p;
    hipError_t e = hipGetDevice(&h->device);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "hipGetDevice");
    }
    const auto sch = default_schedule(p->N, 1);
    h->cols_variant = sch.first;
    h->rows_variant = sch.second;
    const size_t nn = (size_t)p->N * p->N;
    if ((rc = dalloc(&h->h0, nn, "h0")) || (p->N >= 2048 && (rc = dalloc(&h->h0t, nn, "h0 transposed"))) ||
        (rc = reserve_z(h, 1)) || (rc = dalloc(&h->height, nn, "height")) ||
        (rc = dalloc(&h->dx, nn, "dx")) || (rc = dalloc(&h->dz, nn, "dz")) ||
        (rc = dalloc(&h->jac, nn, "jacobian")) || (rc = dalloc(&h->normals, nn, "normals")) ||
        (rc = upload_twiddles(h))) {
        free_all(h);
        delete h;
        return rc;
    }
    // set-up of the static scratch of ocn_update_frames_noalloc on this device
    // (its address is looked up once per device, never in the frame call)
    if ((e = ocn::static_scratch_resolve(h->device)) != hipSuccess) {
        rc = hip_fail(e, "hipGetSymbolAddress(static scratch)");
        free_all(h);
        delete h;
        return rc;
    }
    *out = h;
    return OCN_OK;
}

int ocn_destroy(ocn_handle* h)
{
    if (!h) return OCN_OK;
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    free_all(h);
    delete h;
    return OCN_OK;
}

int ocn_set_params(ocn_handle* h, const ocn_params* p)
{
    int rc = check_handle(h);
    if (rc || (rc = validate(p))) return rc;
    if (p->N != h->p.N) return fail(OCN_EINVAL, "N cannot change (handle N=%d, new N=%d)", h->p.N, p->N);
    h->p        = *p;
    h->h0_ready = false;
    return OCN_OK;
}

int ocn_get_params(const ocn_handle* h, ocn_params* p)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!p) return fail(OCN_EINVAL, "params is NULL");
    *p = h->p;
    return OCN_OK;
}

int ocn_tilde_h0_k(ocn_handle* h, void* stream)
{
    TraceRange trace("ocn_tilde_h0_k");
    int rc = check_handle(h);
    if (rc || (rc = set_device(h))) return rc;
    hipError_t e = ocn::launch_h0(h->p.N, h->h0, h->h0t, nullptr, spec_const(h->p), (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "tilde_h0_k launch");
    h->h0_ready = true;
    return OCN_OK;
}

int ocn_tilde_h0_k_from_noise(ocn_handle* h, const uint8_t* noise_host, void* stream)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!noise_host) return fail(OCN_EINVAL, "noise is NULL");
    if ((rc = set_device(h))) return rc;
    if (!h->noise && (rc = dalloc(&h->noise, 4 * 65536, "noise"))) return rc;
    hipError_t e = hipMemcpyAsync(h->noise, noise_host, 4 * 65536, hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(noise)");
    e = hipStreamSynchronize((hipStream_t)stream);  // noise_host may be freed by the caller on return
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    e = ocn::launch_h0(h->p.N, h->h0, h->h0t, h->noise, spec_const(h->p), (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "tilde_h0_k launch");
    h->h0_ready = true;
    return OCN_OK;
}

int ocn_generate_twiddle_factors(ocn_handle* h)
{
    int rc = check_handle(h);
    if (rc || (rc = set_device(h))) return rc;
    return upload_twiddles(h);
}

int ocn_tilde_h0_t(ocn_handle* h, double t, void* stream)
{
    int rc = check_handle(h);
    if (rc || (rc = ready_for_frame(h, t)) || (rc = set_device(h))) return rc;
    const size_t nn = (size_t)h->p.N * h->p.N;
    for (int f = 0; f < 3; ++f)
        if (!h->spec[f] && (rc = dalloc(&h->spec[f], nn, "tilde_h0_t"))) return rc;
    hipError_t e = ocn::launch_spectrum_t(h->p.N, h->h0, h->spec[0], h->spec[1], h->spec[2], spec_const(h->p),
                                          t / (2.0 * kPi), (hipStream_t)stream);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "tilde_h0_t launch");
}

int ocn_butterfly_fft_buffers(ocn_handle* h, const void* src, float* dst, void* stream)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!src || !dst) return fail(OCN_EINVAL, "src and dst are required");
    const size_t zbytes = sizeof(float2) * (size_t)h->p.N * h->p.N * h->zcap;
    auto in_z1 = [&](const void* q) {
        return (const char*)q >= (const char*)h->z1 && (const char*)q < (const char*)h->z1 + zbytes;
    };
    if (in_z1(src) || in_z1(dst)) return fail(OCN_EINVAL, "src/dst alias the handle's scratch");
    if ((rc = set_device(h))) return rc;
    hipError_t e = ocn::launch_ifft2(h->p.N, (const float2*)src, h->z1, dst, h->tw, (hipStream_t)stream);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "butterfly_fft launch");
}

int ocn_butterfly_fft(ocn_handle* h, int field, void* stream)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (field < OCN_FIELD_DY || field > OCN_FIELD_DZ) return fail(OCN_EINVAL, "field %d out of range", field);
    if (!h->spec[field]) return fail(OCN_ESTATE, "no spectrum texture yet: call ocn_tilde_h0_t first");
    float* dst = field == OCN_FIELD_DY ? h->height : field == OCN_FIELD_DX ? h->dx : h->dz;
    return ocn_butterfly_fft_buffers(h, h->spec[field], dst, stream);
}

int ocn_generate_normal_map(ocn_handle* h, void* stream)
{
    int rc = check_handle(h);
    if (rc || (rc = set_device(h))) return rc;
    hipError_t e = ocn::launch_normals(h->p.N, h->height, h->dx, h->dz, h->normals, h->jac, h->p.normal_mode,
                                       (float)h->p.choppiness, (float)((double)h->p.N / (2.0 * h->p.L)),
                                       (hipStream_t)stream);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "generate_normal_map launch");
}

int ocn_update_stage(ocn_handle* h, const double* t_host, int nframes, const ocn_fields* out, int stage,
                     void* stream)
{
    TraceRange trace("ocn_update_stage");
    int rc = check_frames(h, t_host, nframes, out);
    if (rc) return rc;
    if (stage < OCN_STAGE_COLUMNS || stage > OCN_STAGE_ALL)
        return fail(OCN_EINVAL, "stage %d is not a combination of OCN_STAGE_* bits", stage);
    if (nframes == 0) return OCN_OK;
    if ((rc = set_device(h))) return rc;
    // launches of at most the reserved frame count: nothing is allocated here
    const int chunk = std::min(ocn::kMaxFrames, h->zcap);
    return frames_in_chunks(h, h->z1, h->z2, t_host, nframes, out, stage, chunk, (hipStream_t)stream);
}

int ocn_update_frames_noalloc(ocn_handle* h, const double* t_host, int nframes, const ocn_fields* out, void* stream)
{
    TraceRange trace("ocn_update_frames_noalloc");
    int rc = check_frames(h, t_host, nframes, out);
    if (rc) return rc;
    const size_t nn        = (size_t)h->p.N * h->p.N;
    const size_t per_frame = nn * (sizeof(float2) + sizeof(float2) / 2);  // Z1 + Z2: 12 bytes per texel
    if (per_frame > ocn::kStaticScratchBytes)
        return fail(OCN_EINVAL,
                    "ocn_update_frames_noalloc: one N=%d frame needs %zu bytes of intermediates, more than the "
                    "%zu-byte static scratch (N <= 2048 fit)",
                    h->p.N, per_frame, ocn::kStaticScratchBytes);
    if (nframes == 0) return OCN_OK;
    if ((rc = set_device(h))) return rc;
    char* base = static_cast<char*>(ocn::static_scratch_ptr(h->device));
    if (!base) return fail(OCN_ESTATE, "the static scratch of device %d was not resolved at ocn_create", h->device);
    const int chunk = (int)std::min<size_t>(ocn::kMaxFrames, ocn::kStaticScratchBytes / per_frame);
    float2*   z1    = reinterpret_cast<float2*>(base);
    float2*   z2    = reinterpret_cast<float2*>(base + (size_t)chunk * nn * sizeof(float2));
    return frames_in_chunks(h, z1, z2, t_host, nframes, out, OCN_STAGE_ALL, chunk, (hipStream_t)stream);
}

int ocn_update_into(ocn_handle* h, double t, const ocn_fields* out, void* stream)
{
    TraceRange trace("ocn_update_into");
    return ocn_update_stage(h, &t, 1, out, OCN_STAGE_ALL, stream);
}

int ocn_update(ocn_handle* h, double t, void* stream)
{
    int rc = check_handle(h);
    if (rc) return rc;
    ocn_fields f{h->height, h->dx, h->dz, reinterpret_cast<float*>(h->normals), h->jac};
    return ocn_update_into(h, t, &f, stream);
}

int ocn_update_frames(ocn_handle* h, const double* t_host, int nframes, const ocn_fields* out, void* stream)
{
    return ocn_update_stage(h, t_host, nframes, out, OCN_STAGE_ALL, stream);
}

int ocn_reserve_frames(ocn_handle* h, int nframes)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (nframes < 1 || nframes > ocn::kMaxFrames)
        return fail(OCN_EINVAL, "nframes=%d must be in [1, %d]", nframes, ocn::kMaxFrames);
    if ((rc = set_device(h))) return rc;
    return reserve_z(h, nframes);
}

int ocn_frames_reserved(const ocn_handle* h, int* nframes)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!nframes) return fail(OCN_EINVAL, "nframes is NULL");
    *nframes = std::min(ocn::kMaxFrames, h->zcap);
    return OCN_OK;
}

int ocn_set_schedule(ocn_handle* h, int cols_variant, int rows_variant)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!ocn::schedule_valid(h->p.N, cols_variant, rows_variant))
        return fail(OCN_EINVAL, "schedule (%d, %d) is not available for N=%d", cols_variant, rows_variant, h->p.N);
    h->cols_variant = cols_variant;
    h->rows_variant = rows_variant;
    h->tuned        = true;
    return OCN_OK;
}

int ocn_get_schedule(const ocn_handle* h, int nframes, int* cols_variant, int* rows_variant, int* cols_radix,
                     int* cols_per_block, int* rows_radix, int* rows_band)
{
    int rc = check_handle(h);
    if (rc) return rc;
    int cv, rv;
    pick_schedule(const_cast<ocn_handle*>(h), nframes, &cv, &rv);
    int E, C;
    ocn::cols_variant_shape(cv, &E, &C);
    if (cols_variant) *cols_variant = cv;
    if (rows_variant) *rows_variant = rv;
    if (cols_radix) *cols_radix = E;
    if (cols_per_block) *cols_per_block = C;
    if (rows_radix) *rows_radix = ocn::rows_variant_radix(rv);
    if (rows_band) *rows_band = h->p.normal_mode == OCN_NORMALS_NEAREST ? ocn::rows_variant_band(rv) : 0;
    return OCN_OK;
}

int ocn_autotune(ocn_handle* h, int nframes, int reps, void* stream)
{
    TraceRange trace("ocn_autotune");
    int rc = check_handle(h);
    if (rc) return rc;
    if (nframes < 1 || nframes > ocn::kMaxFrames || reps < 1 || reps > 10000)
        return fail(OCN_EINVAL, "nframes must be in [1, %d] and reps in [1, 10000]", ocn::kMaxFrames);
    if ((rc = set_device(h)) || (rc = reserve_z(h, nframes))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t nn = (size_t)h->p.N * h->p.N;
    float* outb = nullptr;
    if ((rc = dalloc(&outb, nn * 8 * nframes, "autotune outputs"))) return rc;
    ocn_fields out{outb, outb + nn * nframes, outb + 2 * nn * nframes, outb + 3 * nn * nframes,
                   outb + 7 * nn * nframes};
    std::vector<double> t(nframes);
    for (int f = 0; f < nframes; ++f) t[f] = 0.5 + f / 60.0;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    int   bc = -1, br = -1;
    for (int cv = 0; cv < ocn::kNumColsVariants && rc == OCN_OK; ++cv)
        for (int rv = 0; rv < ocn::kNumRowsVariants && rc == OCN_OK; ++rv) {
            if (!ocn::schedule_valid(h->p.N, cv, rv)) continue;
            // bilinear normals always run the separate normal-map kernel, so a
            // fused row variant would time the same kernels as variant 0
            if (h->p.normal_mode != OCN_NORMALS_NEAREST && ocn::rows_variant_band(rv) > 0) continue;
            const int stages = OCN_STAGE_ALL;
            if ((rc = run_frames(h, t.data(), nframes, &out, nn, stages, cv, rv, s))) break;  // warm-up
            (void)hipEventRecord(e0, s);
            for (int i = 0; i < reps && rc == OCN_OK; ++i)
                rc = run_frames(h, t.data(), nframes, &out, nn, stages, cv, rv, s);
            (void)hipEventRecord(e1, s);
            hipError_t e = hipEventSynchronize(e1);
            if (e != hipSuccess) {
                rc = hip_fail(e, "autotune");
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) {
                best = ms;
                bc   = cv;
                br   = rv;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(outb);
    if (rc) return rc;
    if (bc < 0) return fail(OCN_ENOTSUP, "no schedule available for N=%d", h->p.N);
    {
        std::lock_guard<std::mutex> lk(g_sched_mu);
        g_sched[{h->device, h->p.N, frames_class(nframes), h->p.normal_mode}] = {bc, br};
    }
    h->tuned = false;  // use the tuned table entry for this frame count class
    return OCN_OK;
}

int ocn_autotune_save(const char* path)
{
    if (!path) return fail(OCN_EINVAL, "path is NULL");
    std::vector<std::pair<std::tuple<int, int, int, int>, std::pair<int, int>>> rows;
    {
        std::lock_guard<std::mutex> lk(g_sched_mu);
        rows.assign(g_sched.begin(), g_sched.end());
    }
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(OCN_EINVAL, "cannot open %s for writing", path);
    std::fprintf(fp, "ocn-autotune %d\n# device N frames_class normal_mode cols_variant rows_variant\n",
                 OCN_ABI_VERSION);
    for (const auto& r : rows)
        std::fprintf(fp, "%d %d %d %d %d %d\n", std::get<0>(r.first), std::get<1>(r.first), std::get<2>(r.first),
                     std::get<3>(r.first), r.second.first, r.second.second);
    const bool ok = std::fclose(fp) == 0;
    return ok ? OCN_OK : fail(OCN_EINVAL, "writing %s failed", path);
}

int ocn_autotune_load(const char* path)
{
    if (!path) return fail(OCN_EINVAL, "path is NULL");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(OCN_EINVAL, "cannot open %s", path);
    std::map<std::tuple<int, int, int, int>, std::pair<int, int>> got;
    char line[256];
    int  version = -1, lineno = 0, rc = OCN_OK;
    while (rc == OCN_OK && std::fgets(line, sizeof line, fp)) {
        ++lineno;
        if (lineno == 1) {
            if (std::sscanf(line, "ocn-autotune %d", &version) != 1 || version != OCN_ABI_VERSION)
                rc = fail(OCN_EINVAL, "%s: not an autotune file of ABI %d", path, OCN_ABI_VERSION);
            continue;
        }
        if (line[0] == '#' || line[0] == '\n') continue;
        int  dev, N, cls, mode, cv, rv;
        char extra;
        if (std::sscanf(line, "%d %d %d %d %d %d %c", &dev, &N, &cls, &mode, &cv, &rv, &extra) != 6) {
            rc = fail(OCN_EINVAL, "%s:%d: expected six integers", path, lineno);
            break;
        }
        const bool ok = dev >= 0 && N >= 64 && N <= 4096 && (N & (N - 1)) == 0 && frames_class(cls) == cls &&
                        (mode == OCN_NORMALS_NEAREST || mode == OCN_NORMALS_BILINEAR) &&
                        ocn::schedule_valid(N, cv, rv) &&
                        (mode == OCN_NORMALS_NEAREST || ocn::rows_variant_band(rv) == 0);
        if (!ok) {
            rc = fail(OCN_EINVAL, "%s:%d: not a valid (device, N, class, normal mode, schedule)", path, lineno);
            break;
        }
        got[{dev, N, cls, mode}] = {cv, rv};
    }
    std::fclose(fp);
    if (rc == OCN_OK && version < 0) rc = fail(OCN_EINVAL, "%s is empty", path);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_sched_mu);
    for (const auto& kv : got) g_sched[kv.first] = kv.second;
    return OCN_OK;
}

int ocn_get_fields(const ocn_handle* h, ocn_fields* out)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!out) return fail(OCN_EINVAL, "out is NULL");
    *out = ocn_fields{h->height, h->dx, h->dz, reinterpret_cast<float*>(h->normals), h->jac};
    return OCN_OK;
}

int ocn_get_spectrum(const ocn_handle* h, void** h0, void** hy, void** hx, void** hz)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (h0) *h0 = h->h0;
    if (hy) *hy = h->spec[0];
    if (hx) *hx = h->spec[1];
    if (hz) *hz = h->spec[2];
    return OCN_OK;
}

int ocn_copy_fields_to_host(ocn_handle* h, float* height, float* dx, float* dz, float* normals, float* jacobian,
                            void* stream)
{
    int rc = check_handle(h);
    if (rc || (rc = set_device(h))) return rc;
    const size_t nn = (size_t)h->p.N * h->p.N;
    hipStream_t  s  = (hipStream_t)stream;
    const struct {
        float*      dst;
        const void* src;
        size_t      bytes;
    } copies[] = {{height, h->height, nn * 4}, {dx, h->dx, nn * 4}, {dz, h->dz, nn * 4},
                  {normals, h->normals, nn * 16}, {jacobian, h->jac, nn * 4}};
    for (const auto& c : copies) {
        if (!c.dst) continue;
        hipError_t e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(fields)");
    }
    hipError_t e = hipStreamSynchronize(s);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipStreamSynchronize");
}

int ocn_bit_reversed_indices(int N, int32_t* out)
{
    if (!is_pow2(N) || N < 2 || N > 65536) return fail(OCN_EINVAL, "N=%d must be a power of two in [2, 65536]", N);
    if (!out) return fail(OCN_EINVAL, "out is NULL");
    int nb = 0;
    while ((1 << nb) < N) ++nb;
    for (int i = 0; i < N; ++i) {
        uint32_t v = (uint32_t)i, r = 0;
        for (int b = 0; b < nb; ++b, v >>= 1) r = (r << 1) | (v & 1u);
        out[i] = (int32_t)r;
    }
    return OCN_OK;
}

int ocn_twiddle_factors(int N, float* out)
{
    if (!is_pow2(N) || N < 2 || N > 65536) return fail(OCN_EINVAL, "N=%d must be a power of two in [2, 65536]", N);
    if (!out) return fail(OCN_EINVAL, "out is NULL");
    int nb = 0;
    while ((1 << nb) < N) ++nb;
    std::vector<int32_t> br(N);
    ocn_bit_reversed_indices(N, br.data());
    for (int s = 0; s < nb; ++s) {
        const int span = 1 << s, span2 = 2 << s;
        for (int j = 0; j < N; ++j) {
            // twiddle index k = (j * N / 2^(s+1)) mod N, w = e^{+2 pi i k / N}
            const long long k = ((long long)j * (N / span2)) % N;
            const double a = 2.0 * kPi * (double)k / (double)N;
            const bool top = (j % span2) < span;
            int ia, ib;
            if (s == 0) {
                ia = top ? br[j] : br[j - 1];
                ib = top ? br[j + 1] : br[j];
            } else {
                ia = top ? j : j - span;
                ib = top ? j + span : j;
            }
            float* o = out + ((size_t)s * N + j) * 4;
            o[0] = (float)std::cos(a);
            o[1] = (float)std::sin(a);
            o[2] = (float)ia;
            o[3] = (float)ib;
        }
    }
    return OCN_OK;
}

// ---------------------------------------------------------------- batches

int ocn_batch_create(const ocn_params* params, int count, ocn_batch** out)
{
    if (!out) return fail(OCN_EINVAL, "out is NULL");
    *out = nullptr;
    if (!params || count < 1 || count > OCN_BATCH_MAX_MEMBERS)
        return fail(OCN_EINVAL, "params must be non-NULL and count in [1, %d]", OCN_BATCH_MAX_MEMBERS);
    for (int i = 0; i < count; ++i) {
        if (validate(&params[i])) return fail(OCN_EINVAL, "member %d: %s", i, std::string(g_err).c_str());
        if (params[i].N < ocn::kBatchMinN || params[i].N > ocn::kBatchMaxN)
            return fail(OCN_EINVAL, "member %d: batched N must be in [%d, %d]", i, ocn::kBatchMinN, ocn::kBatchMaxN);
        if (params[i].normal_mode != OCN_NORMALS_NEAREST)
            return fail(OCN_EINVAL, "member %d: the batched frame fuses the nearest-tap normal map only", i);
    }
    ocn_batch* b = new (std::nothrow) ocn_batch();
    if (!b) return fail(OCN_ENOMEM, "host allocation of the batch failed");
    b->params.assign(params, params + count);
    int        rc = OCN_OK;
    hipError_t e  = hipGetDevice(&b->device);
    if (e != hipSuccess) rc = hip_fail(e, "hipGetDevice");
    // Per member, 256-byte aligned: h0 16, z1 8, z2 4, height/dx/dz/jacobian 4 each, normals 16 bytes per texel.
    auto   al    = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t total = 0;
    for (const ocn_params& p : b->params) {
        const size_t nn = (size_t)p.N * p.N;
        total += al(nn * 16) + al(nn * 8) + al(nn * 4) + 4 * al(nn * 4) + al(nn * 16);
    }
    if (!rc) rc = dalloc(&b->pool, total, "batch pool");
    for (const ocn_params& p : b->params) {
        if (rc || b->tw.count(p.N)) continue;
        std::array<float2*, 2>& t = b->tw[p.N];
        t = {nullptr, nullptr};
        if (!(rc = upload_table(&t[0], stockham_twiddles(p.N, 8)))) rc = upload_table(&t[1], stockham_twiddles(p.N, 16));
    }
    if (!rc) {
        char*             q = b->pool;
        std::vector<int2> cols;
        for (int i = 0; i < count; ++i) {
            const ocn_params& p  = b->params[i];
            const size_t      nn = (size_t)p.N * p.N;
            ocn::BatchMember  m{};
            m.h0      = reinterpret_cast<const float4*>(q), q += al(nn * 16);
            m.z1      = reinterpret_cast<float2*>(q), q += al(nn * 8);
            m.z2      = reinterpret_cast<float2*>(q), q += al(nn * 4);
            m.height  = reinterpret_cast<float*>(q), q += al(nn * 4);
            m.dx      = reinterpret_cast<float*>(q), q += al(nn * 4);
            m.dz      = reinterpret_cast<float*>(q), q += al(nn * 4);
            m.jac     = reinterpret_cast<float*>(q), q += al(nn * 4);
            m.normals = reinterpret_cast<float4*>(q), q += al(nn * 16);
            m.tw8     = b->tw[p.N][0];
            m.tw16    = b->tw[p.N][1];
            m.sc      = spec_const(p);
            m.lambda  = (float)p.choppiness;
            m.inv2d   = (float)((double)p.N / (2.0 * p.L));
            m.dscale  = (float)p.displacement_scale;
            m.band    = ocn::batch_rows_band(p.N);
            m.cols    = ocn::batch_cols_per_block(p.N);
            b->members.push_back(m);
        }
        if (!(rc = upload_vector(&b->d_members, b->members, "batch members")) && !(rc = batch_build_cols(b)))
            rc = batch_build_rows(b);
    }
    if (rc) {
        batch_free(b);
        delete b;
        return rc;
    }
    *out = b;
    return OCN_OK;
}

int ocn_batch_destroy(ocn_batch* b)
{
    if (!b) return OCN_OK;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    batch_free(b);
    delete b;
    return OCN_OK;
}

int ocn_batch_size(const ocn_batch* b, int* count)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!count) return fail(OCN_EINVAL, "count is NULL");
    *count = (int)b->params.size();
    return OCN_OK;
}

int ocn_batch_get_params(const ocn_batch* b, int member, ocn_params* p)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!p) return fail(OCN_EINVAL, "p is NULL");
    if (member < 0 || member >= (int)b->params.size())
        return fail(OCN_EINVAL, "member %d out of range [0, %zu)", member, b->params.size());
    *p = b->params[member];
    return OCN_OK;
}

int ocn_batch_tilde_h0_k(ocn_batch* b, void* stream)
{
    int rc = check_batch(b);
    if (rc) return rc;
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    for (size_t i = 0; i < b->members.size(); ++i) {
        e = ocn::launch_h0(b->params[i].N, const_cast<float4*>(b->members[i].h0), nullptr, nullptr, b->members[i].sc,
                           (hipStream_t)stream);
        if (e != hipSuccess) return hip_fail(e, "batch tilde_h0_k launch");
    }
    b->h0_ready = true;
    return OCN_OK;
}

int ocn_batch_update(ocn_batch* b, double t, void* stream)
{
    TraceRange trace("ocn_batch_update");
    int rc = batch_ready(b, t);
    if (rc) return rc;
    hipError_t e = ocn::launch_batch_frame(b->d_members, b->d_cols, b->n_cols, b->d_rows, b->n_rows, b->d_wide,
                                           b->n_wide, t / (2.0 * kPi), b->render, (hipStream_t)stream);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "batch update launch");
}

int ocn_batch_update_graph(ocn_batch* b, double t, void* stream)
{
    TraceRange trace("ocn_batch_update_graph");
    int rc = batch_ready(b, t);
    if (rc) return rc;
    // The graph holds kernel nodes only: columns, then the row pass (and the
    // row pass of the N = 4096 members, beside it); t reaches the column node
    // through its parameters on every replay.
    double              t_rev       = t / (2.0 * kPi);
    void*               cols_args[] = {&b->d_members, &b->d_cols, &t_rev};
    void*               rows_args[] = {&b->d_members, &b->d_rows};
    void*               wide_args[] = {&b->d_members, &b->d_wide};
    hipKernelNodeParams kc{};
    kc.func           = const_cast<void*>(ocn::batch_cols_kernel());
    kc.gridDim        = dim3(b->n_cols);
    kc.blockDim       = dim3(ocn::kBatchColsThreads);
    kc.sharedMemBytes = 0;
    kc.kernelParams   = cols_args;
    kc.extra          = nullptr;
    hipError_t e;
    if (!b->exec) {
        hipKernelNodeParams kr = kc, kw = kc;
        kr.func                = const_cast<void*>(ocn::batch_rows_kernel(b->render));
        kr.gridDim             = dim3(b->n_rows);
        kr.blockDim            = dim3(ocn::kBatchRowsThreads);
        kr.kernelParams        = rows_args;
        kw.func                = const_cast<void*>(ocn::batch_wide_rows_kernel(b->render));
        kw.gridDim             = dim3(b->n_wide);
        kw.blockDim            = dim3(ocn::kBatchWideRowsThreads);
        kw.kernelParams        = wide_args;
        if ((e = hipGraphCreate(&b->graph, 0)) != hipSuccess) {
            b->graph = nullptr;
            return hip_fail(e, "hipGraphCreate");
        }
        e = hipGraphAddKernelNode(&b->node_cols, b->graph, nullptr, 0, &kc);
        if (e == hipSuccess && b->n_rows > 0) e = hipGraphAddKernelNode(&b->node_rows, b->graph, &b->node_cols, 1, &kr);
        if (e == hipSuccess && b->n_wide > 0) e = hipGraphAddKernelNode(&b->node_wide, b->graph, &b->node_cols, 1, &kw);
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
            b->exec  = nullptr;
            return hip_fail(e, "batch graph construction");
        }
    } else if ((e = hipGraphExecKernelNodeSetParams(b->exec, b->node_cols, &kc)) != hipSuccess) {
        return hip_fail(e, "hipGraphExecKernelNodeSetParams");
    }
    e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipGraphLaunch");
}

int ocn_batch_get_fields(const ocn_batch* b, int member, ocn_fields* out)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!out) return fail(OCN_EINVAL, "out is NULL");
    if (member < 0 || member >= (int)b->members.size())
        return fail(OCN_EINVAL, "member %d out of range [0, %zu)", member, b->members.size());
    const ocn::BatchMember& m = b->members[member];
    *out = ocn_fields{m.height, m.dx, m.dz, reinterpret_cast<float*>(m.normals), m.jac,
                      reinterpret_cast<float*>(m.disp), reinterpret_cast<float*>(m.rnorm)};
    return OCN_OK;
}

// Caller-provided outputs of one member (the batch counterpart of
// ocn_update_into): the row pass writes that member's textures at the given
// device addresses from the next frame on.
int ocn_batch_set_member_fields(ocn_batch* b, int member, const ocn_fields* f)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!f) return fail(OCN_EINVAL, "fields is NULL");
    if (member < 0 || member >= (int)b->members.size())
        return fail(OCN_EINVAL, "member %d out of range [0, %zu)", member, b->members.size());
    if (!f->height || !f->dx || !f->dz) return fail(OCN_EINVAL, "height, dx and dz must be non-NULL");
    if (((uintptr_t)f->normals & 15) || ((uintptr_t)f->displacement & 15) || ((uintptr_t)f->render_normals & 15))
        return fail(OCN_EINVAL, "float4 fields must be 16-byte aligned");
    if ((f->displacement || f->render_normals) && !b->render)
        return fail(OCN_ESTATE, "the render export is off (ocn_batch_set_render_export)");
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();  // no launch may read the member table while it changes
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
    ocn::BatchMember& m = b->members[member];
    m.height = f->height;
    m.dx     = f->dx;
    m.dz     = f->dz;
    m.normals = reinterpret_cast<float4*>(f->normals);
    m.jac     = f->jacobian;
    if (b->render) {
        if (f->displacement) m.disp = reinterpret_cast<float4*>(f->displacement);
        if (f->render_normals) m.rnorm = reinterpret_cast<float4*>(f->render_normals);
    }
    e = hipMemcpy(b->d_members + member, &m, sizeof(m), hipMemcpyHostToDevice);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipMemcpy(batch member)");
}

// Band height of the row pass for every member of size N (the batch's
// schedule at that N; the outputs do not depend on it).
int ocn_batch_set_schedule(ocn_batch* b, int N, int rows_band)
{
    int rc = check_batch(b);
    if (rc) return rc;
    bool present = false;
    for (const ocn_params& p : b->params) present |= p.N == N;
    if (!present) return fail(OCN_EINVAL, "no member has N=%d", N);
    if (!ocn::batch_band_valid(N, rows_band))
        return fail(OCN_EINVAL, "rows_band %d is not a valid band height at N=%d", rows_band, N);
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();  // no launch may read the tables while they change
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
    for (size_t i = 0; i < b->members.size(); ++i)
        if (b->params[i].N == N) b->members[i].band = rows_band;
    e = hipMemcpy(b->d_members, b->members.data(), b->members.size() * sizeof(ocn::BatchMember),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(batch members)");
    return batch_build_rows(b);
}

// Columns per strip of the column pass for every member of size N.
int ocn_batch_set_column_strips(ocn_batch* b, int N, int cols_per_block)
{
    int rc = check_batch(b);
    if (rc) return rc;
    bool present = false;
    for (const ocn_params& p : b->params) present |= p.N == N;
    if (!present) return fail(OCN_EINVAL, "no member has N=%d", N);
    if (!ocn::batch_cols_valid(N, cols_per_block))
        return fail(OCN_EINVAL, "cols_per_block %d is not a valid strip width at N=%d", cols_per_block, N);
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();  // no launch may read the tables while they change
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
    for (size_t i = 0; i < b->members.size(); ++i)
        if (b->params[i].N == N) b->members[i].cols = cols_per_block;
    e = hipMemcpy(b->d_members, b->members.data(), b->members.size() * sizeof(ocn::BatchMember),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(batch members)");
    return batch_build_cols(b);
}

int ocn_batch_get_column_strips(const ocn_batch* b, int member, int* cols_per_block)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!cols_per_block) return fail(OCN_EINVAL, "cols_per_block is NULL");
    if (member < 0 || member >= (int)b->members.size())
        return fail(OCN_EINVAL, "member %d out of range [0, %zu)", member, b->members.size());
    *cols_per_block = b->members[member].cols;
    return OCN_OK;
}

int ocn_batch_get_schedule(const ocn_batch* b, int member, int* rows_band)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!rows_band) return fail(OCN_EINVAL, "rows_band is NULL");
    if (member < 0 || member >= (int)b->members.size())
        return fail(OCN_EINVAL, "member %d out of range [0, %zu)", member, b->members.size());
    *rows_band = b->members[member].band;
    return OCN_OK;
}

// Per N present, every valid (column strip width, band height) pair timed
// over `reps` whole frames of the batch (HIP events on `stream`, one warm-up
// frame first); the fastest pair is kept, N by N.
int ocn_batch_autotune(ocn_batch* b, int reps, void* stream)
{
    TraceRange trace("ocn_batch_autotune");
    int rc = batch_ready(b, 0.0);
    if (rc) return rc;
    if (reps < 1 || reps > 1000) return fail(OCN_EINVAL, "reps must be in [1, 1000]");
    std::vector<int> sizes;
    for (const ocn_params& p : b->params)
        if (std::find(sizes.begin(), sizes.end(), p.N) == sizes.end()) sizes.push_back(p.N);
    hipStream_t s  = (hipStream_t)stream;
    hipEvent_t  e0 = nullptr, e1 = nullptr;
    hipError_t  e  = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int N : sizes) {
        int   best_band = 0, best_cols = 0;
        float best_ms   = 0.0f;
        const int C0 = ocn::batch_cols_per_block(N);
        for (int C : {C0, C0 / 2})
        for (int R = 2; R <= 32 && e == hipSuccess && !rc; R *= 2) {
            if (!ocn::batch_band_valid(N, R) || !ocn::batch_cols_valid(N, C)) continue;
            if ((rc = ocn_batch_set_schedule(b, N, R)) || (rc = ocn_batch_set_column_strips(b, N, C))) break;
            auto frame = [&](int i) {
                return ocn::launch_batch_frame(b->d_members, b->d_cols, b->n_cols, b->d_rows, b->n_rows, b->d_wide,
                                               b->n_wide, 0.25 * i, b->render, s);
            };
            e = frame(0);
            if (e == hipSuccess) e = hipEventRecord(e0, s);
            for (int i = 1; i <= reps && e == hipSuccess; ++i) e = frame(i);
            if (e == hipSuccess) e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess && (best_band == 0 || ms < best_ms)) {
                best_band = R;
                best_cols = C;
                best_ms   = ms;
            }
        }
        if (e != hipSuccess || rc) break;
        if ((rc = ocn_batch_set_schedule(b, N, best_band)) || (rc = ocn_batch_set_column_strips(b, N, best_cols)))
            break;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    return e == hipSuccess ? OCN_OK : hip_fail(e, "batch autotune");
}

// grid_tes.glsl:58-62 / main.cpp:1632-1633 for every member: the render export
// of the single-ocean fields (ocn_fields.displacement / render_normals).
int ocn_batch_set_render_export(ocn_batch* b, int enable)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (enable != 0 && enable != 1) return fail(OCN_EINVAL, "enable must be 0 or 1");
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();  // no launch may read the member table while it changes
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    if (enable && !b->render_pool) {
        size_t total = 0;
        for (const ocn_params& p : b->params) total += 2 * al((size_t)p.N * p.N * 16);
        if ((rc = dalloc(&b->render_pool, total, "batch render pool"))) return rc;
    }
    char* q = b->render_pool;
    for (size_t i = 0; i < b->members.size(); ++i) {
        const size_t bytes = al((size_t)b->params[i].N * b->params[i].N * 16);
        b->members[i].disp  = enable ? reinterpret_cast<float4*>(q) : nullptr;
        b->members[i].rnorm = enable ? reinterpret_cast<float4*>(q + bytes) : nullptr;
        if (q) q += 2 * bytes;
    }
    e = hipMemcpy(b->d_members, b->members.data(), b->members.size() * sizeof(ocn::BatchMember),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(batch members)");
    if (b->exec && b->render != (enable != 0)) {  // the graph's row node changes kernel: rebuilt on next use
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec  = nullptr;
        b->graph = nullptr;
    }
    b->render = enable != 0;
    return OCN_OK;
}

int ocn_batch_get_h0(const ocn_batch* b, int member, void** h0)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!h0) return fail(OCN_EINVAL, "h0 is NULL");
    if (member < 0 || member >= (int)b->members.size())
        return fail(OCN_EINVAL, "member %d out of range [0, %zu)", member, b->members.size());
    *h0 = const_cast<float4*>(b->members[member].h0);
    return OCN_OK;
}

int ocn_batch_copy_fields_to_host(ocn_batch* b, int member, float* height, float* dx, float* dz, float* normals,
                                  float* jacobian, void* stream)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (member < 0 || member >= (int)b->members.size())
        return fail(OCN_EINVAL, "member %d out of range [0, %zu)", member, b->members.size());
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const ocn::BatchMember& m  = b->members[member];
    const size_t            nn = (size_t)b->params[member].N * b->params[member].N;
    hipStream_t             s  = (hipStream_t)stream;
    const struct {
        float*      dst;
        const void* src;
        size_t      bytes;
    } copies[] = {{height, m.height, nn * 4}, {dx, m.dx, nn * 4}, {dz, m.dz, nn * 4}, {normals, m.normals, nn * 16},
                  {jacobian, m.jac, nn * 4}};
    for (const auto& c : copies) {
        if (!c.dst) continue;
        if ((e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s)) != hipSuccess)
            return hip_fail(e, "hipMemcpyAsync(batch fields)");
    }
    e = hipStreamSynchronize(s);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipStreamSynchronize");
}

// ---------------------------------------------------------------- fits

int ocn_fit_create(const ocn_params* params, int count, int K, const double* times, ocn_fit** out)
{
    if (!out) return fail(OCN_EINVAL, "out is NULL");
    *out = nullptr;
    if (!params || count < 1 || count > OCN_BATCH_MAX_MEMBERS)
        return fail(OCN_EINVAL, "params must be non-NULL and count in [1, %d]", OCN_BATCH_MAX_MEMBERS);
    if (!times || K < 1 || K > OCN_FIT_MAX_FRAMES)
        return fail(OCN_EINVAL, "times must be non-NULL and K in [1, %d]", OCN_FIT_MAX_FRAMES);
    for (int k = 0; k < K; ++k)
        if (!std::isfinite(times[k])) return fail(OCN_EINVAL, "times[%d] is not finite", k);
    for (int i = 0; i < count; ++i) {
        if (validate(&params[i])) return fail(OCN_EINVAL, "member %d: %s", i, std::string(g_err).c_str());
        if (params[i].N < ocn::kFitMinN || params[i].N > ocn::kFitMaxN)
            return fail(OCN_EINVAL, "member %d: fitted N must be in [%d, %d]", i, ocn::kFitMinN, ocn::kFitMaxN);
    }
    // Equally spaced times (to 1e-12 of their scale) let the kernels step the
    // dispersion phase from frame to frame instead of evaluating it per frame.
    bool   uniform_times = K >= 2;
    double uniform_dt    = K >= 2 ? (times[K - 1] - times[0]) / (K - 1) : 0.0;
    {
        double scale = 1.0;
        for (int k = 0; k < K; ++k) scale = std::max(scale, std::fabs(times[k]));
        for (int k = 0; k < K && uniform_times; ++k)
            uniform_times = std::fabs(times[k] - (times[0] + k * uniform_dt)) <= 1e-12 * scale;
    }
    ocn_fit* f = new (std::nothrow) ocn_fit();
    if (!f) return fail(OCN_ENOMEM, "host allocation of the fit failed");
    f->K = K;
    f->params.assign(params, params + count);
    int        rc = OCN_OK;
    hipError_t e  = hipGetDevice(&f->device);
    if (e != hipSuccess) rc = hip_fail(e, "hipGetDevice");
    // Per member, 256-byte aligned: xi 16, scratch 8, targets 4 K, dL/dh0 16 bytes per texel, partials,
    // the targets' spectra T^ (8 bytes per frame and texel pair of the half plane), with equally spaced
    // times the phase table (16 bytes per texel pair of the half plane), the noise coefficients (16 bytes
    // per texel pair of the half plane) and the unit noise of the edge pairs (32 bytes each).
    // Small fits split each column strip's rows over up to four blocks so the
    // step launch has at least two blocks per CU of the 256.
    int row_parts = 1;
    {
        long items = 0;
        for (const ocn_params& p : f->params) items += ocn::fit_cols_items_per_member(p.N);
        while (row_parts < ocn::kFitMaxRowParts && items * row_parts < 512) row_parts *= 2;
    }
    auto   al    = [](size_t x) { return (x + 255) & ~size_t(255); };
    auto   phase_bytes = [&](int N) { return uniform_times ? al(sizeof(float4) * ocn::fit_phase_entries(N)) : 0; };
    auto   bytes = [&](int N) {
        const size_t nn = (size_t)N * N;
        return al(nn * 16) + al(nn * 8) + al(nn * 4 * K) + al(nn * 16) +
               al(sizeof(float) * ocn::fit_cols_items_per_member(N) * row_parts) +
               al(sizeof(float) * 3 * ocn::fit_cols_items_per_member(N) * row_parts) +
               al(sizeof(float2) * K * ocn::fit_phase_entries(N)) + phase_bytes(N) +
               al(sizeof(float4) * ocn::fit_phase_entries(N)) + al(2 * sizeof(float4) * ocn::fit_edge_pairs(N));
    };
    size_t total = 0;
    for (const ocn_params& p : f->params) total += bytes(p.N);
    if (!rc) rc = dalloc(&f->pool, total, "fit pool");
    if (!rc && (e = hipMemset(f->pool, 0, total)) != hipSuccess) rc = hip_fail(e, "hipMemset(fit pool)");
    for (const ocn_params& p : f->params) {
        if (rc || f->tw.count(p.N)) continue;
        std::array<float2*, 2>& t = f->tw[p.N];
        t = {nullptr, nullptr};
        if (!(rc = upload_table(&t[0], stockham_twiddles(p.N, 8)))) rc = upload_table(&t[1], stockham_twiddles(p.N, 16));
    }
    if (!rc) rc = dalloc(&f->d_vars, (size_t)count, "fit variables");
    if (!rc) {
        std::vector<ocn::FitVars> vars;
        char*                     q = f->pool;
        std::vector<int2>         cols;
        for (int i = 0; i < count; ++i) {
            const ocn_params& p  = f->params[i];
            const size_t      nn = (size_t)p.N * p.N;
            ocn::FitMember    m{};
            m.N         = p.N;
            m.K         = K;
            m.conj_mode = p.conj_mode;
            m.kscale    = (float)(2.0 * kPi / p.L);
            m.l2        = (float)(p.suppression * p.suppression);
            m.g         = (float)p.gravity;
            m.inv_knn   = (float)(1.0 / ((double)K * (double)nn));
            m.k2_d      = (2.0 * kPi / p.L) * (2.0 * kPi / p.L);
            m.g2_d      = p.gravity * p.gravity;
            m.g_d       = p.gravity;
            for (int k = 0; k < K; ++k) m.t_rev[k] = times[k] / (2.0 * kPi);
            m.uniform = uniform_times ? 1 : 0;
            m.dt_rev  = uniform_dt / (2.0 * kPi);
            m.xi        = reinterpret_cast<const float4*>(q), q += al(nn * 16);
            m.scratch   = reinterpret_cast<float2*>(q), q += al(nn * 8);
            m.targets   = reinterpret_cast<const float*>(q), q += al(nn * 4 * K);
            m.grad_h0   = reinterpret_cast<float4*>(q), q += al(nn * 16);
            m.row_parts   = row_parts;
            m.n_loss_part = ocn::fit_cols_items_per_member(p.N) * row_parts;  // one partial per step block
            m.n_grad_part = m.n_loss_part;
            m.loss_part = reinterpret_cast<float*>(q), q += al(sizeof(float) * m.n_loss_part);
            m.grad_part = reinterpret_cast<float*>(q), q += al(sizeof(float) * 3 * m.n_grad_part);
            m.that      = reinterpret_cast<float2*>(q), q += al(sizeof(float2) * K * ocn::fit_phase_entries(p.N));
            m.phase     = uniform_times ? reinterpret_cast<const float4*>(q) : nullptr, q += phase_bytes(p.N);
            m.pq        = reinterpret_cast<const float4*>(q), q += al(sizeof(float4) * ocn::fit_phase_entries(p.N));
            m.edge      = reinterpret_cast<const float4*>(q), q += al(2 * sizeof(float4) * ocn::fit_edge_pairs(p.N));
            m.tw8       = f->tw[p.N][0];
            m.tw16      = f->tw[p.N][1];
            m.vars      = f->d_vars + i;
            m.key0      = (uint32_t)p.seed;
            m.key1      = (uint32_t)(p.seed >> 32);
            m.ocean_id  = p.ocean_id;
            m.cascade_id = p.cascade_id;
            f->members.push_back(m);
            vars.push_back(initial_vars(p));
            for (int k = 0; k < m.n_grad_part; ++k) cols.push_back(make_int2(i, k));
        }
        f->n_cols = (int)cols.size();
        f->targets_dirty.assign(count, 1);
        if (!(rc = upload_vector(&f->d_members, f->members, "fit members")) &&
            !(rc = upload_vector(&f->d_cols, cols, "fit column items"))) {
            e = hipMemcpy(f->d_vars, vars.data(), sizeof(ocn::FitVars) * vars.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy(fit variables)");
        }
        for (int i = 0; !rc && i < count; ++i) {
            if ((e = ocn::launch_fit_xi(f->members[i], nullptr)) != hipSuccess) rc = hip_fail(e, "fit noise launch");
            else if ((e = ocn::launch_fit_pq(f->members[i], nullptr)) != hipSuccess)
                rc = hip_fail(e, "fit noise coefficients launch");
            else if (uniform_times && (e = ocn::launch_fit_phase(f->members[i], nullptr)) != hipSuccess)
                rc = hip_fail(e, "fit phase launch");
        }
        if (!rc && (e = hipDeviceSynchronize()) != hipSuccess) rc = hip_fail(e, "hipDeviceSynchronize");
    }
    if (rc) {
        fit_free(f);
        delete f;
        return rc;
    }
    *out = f;
    return OCN_OK;
}

int ocn_fit_destroy(ocn_fit* f)
{
    if (!f) return OCN_OK;
    (void)hipSetDevice(f->device);
    (void)hipDeviceSynchronize();
    fit_free(f);
    delete f;
    return OCN_OK;
}

int ocn_fit_targets(const ocn_fit* f, int member, float** targets)
{
    int rc = check_fit(f, member);
    if (rc) return rc;
    if (!targets) return fail(OCN_EINVAL, "targets is NULL");
    *targets = const_cast<float*>(f->members[member].targets);
    f->targets_dirty[member] = 1;  // the caller may write them through this pointer
    return OCN_OK;
}

int ocn_fit_upload_targets(ocn_fit* f, int member, const float* targets, void* stream)
{
    int rc = check_fit(f, member);
    if (rc) return rc;
    if (!targets) return fail(OCN_EINVAL, "targets is NULL");
    const ocn::FitMember& m = f->members[member];
    const size_t          bytes = sizeof(float) * (size_t)m.K * m.N * m.N;
    hipError_t            e = hipSetDevice(f->device);
    if (e == hipSuccess)
        e = hipMemcpyAsync(const_cast<float*>(m.targets), targets, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e == hipSuccess) f->targets_dirty[member] = 1;
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipMemcpyAsync(fit targets)");
}

int ocn_fit_grad_h0(const ocn_fit* f, int member, void** grad_h0)
{
    int rc = check_fit(f, member);
    if (rc) return rc;
    if (!grad_h0) return fail(OCN_EINVAL, "grad_h0 is NULL");
    *grad_h0 = f->members[member].grad_h0;
    return OCN_OK;
}

int ocn_fit_set_vars(ocn_fit* f, int member, const double vars[3])
{
    int rc = check_fit(f, member);
    if (rc) return rc;
    if (!vars) return fail(OCN_EINVAL, "vars is NULL");
    if (!(std::isfinite(vars[0]) && vars[0] > 0.0 && std::isfinite(vars[1]) && std::isfinite(vars[2]) &&
          vars[2] > 0.0))
        return fail(OCN_EINVAL, "wind speed and amplitude must be finite and > 0, the angle finite");
    ocn::FitVars v{};
    v.V     = vars[0];
    v.theta = vars[1];
    v.A     = vars[2];
    ocn::fit_vars_derive(v, f->params[member].gravity);
    hipError_t e = hipSetDevice(f->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(f->members[member].vars, &v, sizeof(v), hipMemcpyHostToDevice);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipMemcpy(fit variables)");
}

int ocn_fit_forward(ocn_fit* f, void* stream)
{
    TraceRange trace("ocn_fit_forward");
    return fit_run(f, ocn::kFitLoss, ocn::AdamCfg{}, ocn::QnCfg{}, stream);
}

int ocn_fit_backward(ocn_fit* f, void* stream)
{
    TraceRange trace("ocn_fit_backward");
    return fit_run(f, ocn::kFitGrad, ocn::AdamCfg{}, ocn::QnCfg{}, stream);
}

int ocn_fit_step(ocn_fit* f, const ocn_adam* adam, void* stream)
{
    TraceRange trace("ocn_fit_step");
    if (!adam) return fail(OCN_EINVAL, "adam is NULL");
    for (int i = 0; i < 3; ++i)
        if (!(std::isfinite(adam->lr[i]) && adam->lr[i] >= 0.0)) return fail(OCN_EINVAL, "lr[%d] must be >= 0", i);
    if (!(adam->beta1 >= 0.0 && adam->beta1 < 1.0 && adam->beta2 >= 0.0 && adam->beta2 < 1.0 && adam->eps > 0.0))
        return fail(OCN_EINVAL, "Adam needs beta1, beta2 in [0, 1) and eps > 0");
    ocn::AdamCfg c{};
    for (int i = 0; i < 3; ++i) c.lr[i] = adam->lr[i];
    c.beta1 = adam->beta1;
    c.beta2 = adam->beta2;
    c.eps   = adam->eps;
    return fit_run(f, ocn::kFitAdam, c, ocn::QnCfg{}, stream);
}

int ocn_fit_step_qn(ocn_fit* f, const ocn_qn* qn, void* stream)
{
    TraceRange trace("ocn_fit_step_qn");
    if (!qn) return fail(OCN_EINVAL, "qn is NULL");
    for (int i = 0; i < 3; ++i)
        if (!(std::isfinite(qn->scale[i]) && qn->scale[i] > 0.0)) return fail(OCN_EINVAL, "scale[%d] must be > 0", i);
    if (!(std::isfinite(qn->radius) && qn->radius > 0.0 && std::isfinite(qn->radius_max) &&
          qn->radius_max >= qn->radius))
        return fail(OCN_EINVAL, "the quasi-Newton step needs 0 < radius <= radius_max, both finite");
    ocn::QnCfg c{};
    for (int i = 0; i < 3; ++i) c.scale[i] = qn->scale[i];
    c.radius     = qn->radius;
    c.radius_max = qn->radius_max;
    return fit_run(f, ocn::kFitQn, ocn::AdamCfg{}, c, stream);
}

int ocn_fit_get(ocn_fit* f, int member, double vars[3], double grad[3], double* loss, void* stream)
{
    int rc = check_fit(f, member);
    if (rc) return rc;
    ocn::FitVars v{};
    hipError_t   e = hipSetDevice(f->device);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e == hipSuccess) e = hipMemcpy(&v, f->members[member].vars, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(fit variables)");
    if (vars) {
        vars[0] = v.V;
        vars[1] = v.theta;
        vars[2] = v.A;
    }
    if (grad)
        for (int i = 0; i < 3; ++i) grad[i] = v.grad[i];
    if (loss) *loss = v.loss;
    return OCN_OK;
}

// ---------------------------------------------------------------- graphics interop

// Stands in for the GL texture storage of main.cpp:1096-1099: the memory a
// renderer exports is mapped, and the frames are written into it.
int ocn_external_import(int device, int fd, uint64_t size, uint64_t offset, uint64_t bytes, ocn_external** out,
                        void** dptr)
{
    if (!out || !dptr) return fail(OCN_EINVAL, "out/dptr is NULL");
    *out  = nullptr;
    *dptr = nullptr;
    if (fd < 0) return fail(OCN_EINVAL, "fd %d is not a file descriptor", fd);
    if (size == 0 || bytes == 0) return fail(OCN_EINVAL, "size and bytes must be > 0");
    if (offset > size || bytes > size - offset)
        return fail(OCN_EINVAL, "[%llu, %llu + %llu) exceeds the %llu-byte object", (unsigned long long)offset,
                    (unsigned long long)offset, (unsigned long long)bytes, (unsigned long long)size);
    if (offset % 256) return fail(OCN_EINVAL, "offset %llu is not a multiple of 256", (unsigned long long)offset);
    int        count = 0;
    hipError_t e     = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= count) return fail(OCN_EINVAL, "device %d not in [0, %d)", device, count);
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipExternalMemoryHandleDesc hd{};
    hd.type      = hipExternalMemoryHandleTypeOpaqueFd;
    hd.handle.fd = fd;
    hd.size      = size;
    hipExternalMemory_t mem = nullptr;
    if ((e = hipImportExternalMemory(&mem, &hd)) != hipSuccess) return hip_fail(e, "hipImportExternalMemory");
    hipExternalMemoryBufferDesc bd{};
    bd.offset = offset;
    bd.size   = bytes;
    void* p   = nullptr;
    if ((e = hipExternalMemoryGetMappedBuffer(&p, mem, &bd)) != hipSuccess || !p) {
        (void)hipDestroyExternalMemory(mem);
        return e != hipSuccess ? hip_fail(e, "hipExternalMemoryGetMappedBuffer")
                               : fail(OCN_EHIP, "hipExternalMemoryGetMappedBuffer returned NULL");
    }
    auto* x = new (std::nothrow) ocn_external;
    if (!x) {
        (void)hipDestroyExternalMemory(mem);
        return fail(OCN_ENOMEM, "host allocation of the import record");
    }
    x->device = device;
    x->mem    = mem;
    *out      = x;
    *dptr     = p;
    return OCN_OK;
}

// The mapped address lies inside the imported object, whose release
// unmaps it; it is not freed on its own.
int ocn_external_release(ocn_external* x)
{
    if (!x) return OCN_OK;
    hipError_t e = hipSetDevice(x->device);
    if (e == hipSuccess) e = hipDestroyExternalMemory(x->mem);
    delete x;
    return e == hipSuccess ? OCN_OK : hip_fail(e, "hipDestroyExternalMemory");
}

}  // extern "C"
