// ocean_capi.cpp -- the extern "C" boundary (include/ocean_c.h) over the
// gfx950 kernels.  Host-side validation, the handle's device buffers, FFT
// plans and launch sequencing.  Citations name the reference code each entry
// point replaces (diharaw/fft-ocean-waves, src/main.cpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/ocean_c.h"
#include "ocean_kernels.h"

namespace {

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
    s.N          = p.N;
    s.kscale     = (float)(2.0 * kPi / p.L);
    s.Lp2        = (float)(Lp * Lp);
    s.l2         = (float)(p.suppression * p.suppression);
    s.A          = (float)p.amplitude;
    s.wx         = (float)(p.wind_dir_x / n);
    s.wy         = (float)(p.wind_dir_y / n);
    s.key0       = (uint32_t)p.seed;
    s.key1       = (uint32_t)(p.seed >> 32);
    s.ocean_id   = p.ocean_id;
    s.cascade_id = p.cascade_id;
    s.conj_mode  = p.conj_mode;
    s.kscale_d   = 2.0 * kPi / p.L;
    s.g_d        = p.gravity;
    return s;
}

}  // namespace

struct ocn_handle {
    ocn_params p{};
    int        device = 0;
    bool       h0_ready = false;
    float4*    h0 = nullptr;      // N*N
    float2*    z1 = nullptr;      // N*N: fused-pass intermediate / butterfly scratch
    float2*    z2 = nullptr;      // N*N/2
    float2*    tw = nullptr;      // Stockham twiddles
    float2*    spec[3] = {nullptr, nullptr, nullptr};  // tilde_h0_t textures (lazy)
    uint8_t*   noise = nullptr;   // 4*256*256 (lazy)
    float*     height = nullptr;
    float*     dx = nullptr;
    float*     dz = nullptr;
    float*     jac = nullptr;
    float4*    normals = nullptr;
};

namespace {

void free_all(ocn_handle* h)
{
    void* ptrs[] = {h->h0, h->z1, h->z2, h->tw, h->spec[0], h->spec[1], h->spec[2], h->noise,
                    h->height, h->dx, h->dz, h->jac, h->normals};
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
    std::vector<float> t = stockham_twiddles(h->p.N, ocn::fft_radix());
    if (h->tw) (void)hipFree(h->tw);
    h->tw = nullptr;
    int rc = dalloc(&h->tw, t.size() / 2 + 1, "twiddles");
    if (rc) return rc;
    if (!t.empty()) {
        hipError_t e = hipMemcpy(h->tw, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy(twiddles)");
    }
    return OCN_OK;
}

int frame_args(const ocn_handle* h, double t, const ocn_fields* out, ocn::FrameArgs* a)
{
    a->h0          = h->h0;
    a->z1          = h->z1;
    a->z2          = h->z2;
    a->tw          = h->tw;
    a->sc          = spec_const(h->p);
    a->t_rev       = t / (2.0 * kPi);
    a->height      = out->height;
    a->dx          = out->dx;
    a->dz          = out->dz;
    a->normals     = reinterpret_cast<float4*>(out->normals);
    a->jacobian    = out->jacobian;
    a->normal_mode = h->p.normal_mode;
    a->lambda      = (float)h->p.choppiness;
    a->inv2d       = (float)((double)h->p.N / (2.0 * h->p.L));
    return OCN_OK;
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

}  // namespace

// ====================================================================== C ABI

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
    const size_t nn = (size_t)p->N * p->N;
    if ((rc = dalloc(&h->h0, nn, "h0")) || (rc = dalloc(&h->z1, nn, "z1")) || (rc = dalloc(&h->z2, nn / 2, "z2")) ||
        (rc = dalloc(&h->height, nn, "height")) || (rc = dalloc(&h->dx, nn, "dx")) ||
        (rc = dalloc(&h->dz, nn, "dz")) || (rc = dalloc(&h->jac, nn, "jacobian")) ||
        (rc = dalloc(&h->normals, nn, "normals")) || (rc = upload_twiddles(h))) {
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
    int rc = check_handle(h);
    if (rc || (rc = set_device(h))) return rc;
    hipError_t e = ocn::launch_h0(h->p.N, h->h0, nullptr, spec_const(h->p), (hipStream_t)stream);
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
    e = ocn::launch_h0(h->p.N, h->h0, h->noise, spec_const(h->p), (hipStream_t)stream);
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
    if (src == (const void*)h->z1 || (const void*)dst == (const void*)h->z1)
        return fail(OCN_EINVAL, "src/dst alias the handle's scratch");
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

int ocn_update_stage(ocn_handle* h, double t, const ocn_fields* out, int stage, void* stream)
{
    int rc = check_handle(h);
    if (rc || (rc = check_fields(out)) || (rc = ready_for_frame(h, t))) return rc;
    if (stage < OCN_STAGE_COLUMNS || stage > OCN_STAGE_ALL)
        return fail(OCN_EINVAL, "stage %d is not a combination of OCN_STAGE_* bits", stage);
    if ((rc = set_device(h))) return rc;
    ocn::FrameArgs a{};
    frame_args(h, t, out, &a);
    hipError_t e = ocn::launch_frame(h->p.N, a, stage, (hipStream_t)stream);
    return e == hipSuccess ? OCN_OK : hip_fail(e, "update launch");
}

int ocn_update_into(ocn_handle* h, double t, const ocn_fields* out, void* stream)
{
    return ocn_update_stage(h, t, out, OCN_STAGE_ALL, stream);
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
    int rc = check_handle(h);
    if (rc || (rc = check_fields(out))) return rc;
    if (!t_host || nframes < 0) return fail(OCN_EINVAL, "t_host must be non-NULL and nframes >= 0");
    for (int f = 0; f < nframes; ++f)
        if ((rc = ready_for_frame(h, t_host[f]))) return rc;
    const size_t nn = (size_t)h->p.N * h->p.N;
    for (int f = 0; f < nframes; ++f) {
        ocn_fields s{out->height + f * nn, out->dx + f * nn, out->dz + f * nn,
                     out->normals ? out->normals + f * nn * 4 : nullptr,
                     out->jacobian ? out->jacobian + f * nn : nullptr};
        if ((rc = ocn_update_into(h, t_host[f], &s, stream))) return rc;
    }
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

}  // extern "C"
