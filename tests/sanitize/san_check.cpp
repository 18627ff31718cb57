// san_check.cpp -- host AddressSanitizer + UndefinedBehaviorSanitizer run over
// the FP64 oracle (oracle/ocean_oracle.c) and the host side of the C ABI
// (fft-ocean-waves_amd/csrc/ocean_capi.cpp), both compiled here with
// -fsanitize=address,undefined.  No GPU is touched: the oracle is CPU code, and
// every ABI call below is rejected by the library's host-side validation
// before any HIP call (or is a host utility).  Prints one line per check and
// "san_check OK" at the end; any sanitizer report aborts the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/ocean_c.h"
#include "../../oracle/ocean_oracle.h"

static int g_fail = 0;

static void expect(bool ok, const char* what)
{
    std::printf("%s %s\n", ok ? "ok" : "FAIL", what);
    if (!ok) ++g_fail;
}

static void expect_status(int got, int want, const char* what)
{
    const bool ok = got == want;
    std::printf("%s %s: %s%s%s\n", ok ? "ok" : "FAIL", what, ocn_status_string(got), got ? " -- " : "",
                got ? ocn_last_error() : "");
    if (!ok) ++g_fail;
}

static orc_params oracle_params(int N)
{
    orc_params p{};
    p.N = N;
    p.L = 3.0 * N;
    p.wind_speed = 21.0;
    p.wind_dir_x = 1.0;
    p.wind_dir_y = 0.35;
    p.amplitude = 3.0;
    p.suppression = p.L / 2000.0;
    p.choppiness = 1.2;
    p.gravity = 9.81;
    p.seed = 42;
    return p;
}

static void oracle_checks()
{
    const int N = 64;
    const size_t nn = (size_t)N * N;
    orc_params p = oracle_params(N);
    std::vector<int32_t> br(N);
    expect(orc_bit_reversed_indices(N, br.data()) == 0 && br[1] == N / 2, "oracle bit reversal");
    std::vector<double> tw(4 * (size_t)orc_log2n(N) * N);
    expect(orc_twiddle_factors(N, tw.data()) == 0, "oracle twiddles");
    std::vector<double> h0(4 * nn), hy(2 * nn), hx(2 * nn), hz(2 * nn), scratch(2 * nn), out(nn);
    expect(orc_h0(&p, h0.data()) == 0, "oracle h0 (Philox)");
    std::vector<uint8_t> noise(4 * 256 * 256);
    for (size_t i = 0; i < noise.size(); ++i) noise[i] = (uint8_t)((i * 2654435761u) >> 24);
    std::vector<double> h0n(4 * nn);
    expect(orc_h0_from_noise(&p, noise.data(), h0n.data()) == 0, "oracle h0 (noise images)");
    expect(orc_spectrum_t(&p, h0.data(), 1.5, hy.data(), hx.data(), hz.data()) == 0, "oracle spectrum");
    for (int nt : {1, 3}) {
        std::vector<double> d(hy);
        expect(orc_butterfly_fft2d(N, tw.data(), d.data(), scratch.data(), nt) == 0, "oracle butterflies");
        orc_inversion(N, d.data(), out.data());
    }
    std::vector<double> nrm(4 * nn), jac(nn), disp(4 * nn), rn(4 * nn);
    for (int mode : {0, 1}) orc_normal_map(N, out.data(), mode, nrm.data());
    orc_jacobian(N, p.L, p.choppiness, out.data(), out.data(), jac.data());
    orc_render(N, p.L, p.choppiness, 0.5, out.data(), out.data(), out.data(), disp.data(), rn.data());
    std::vector<double> fh(nn), fx(nn), fz(nn), fn(4 * nn), fj(nn);
    for (int nt : {1, 2, 5}) {
        p.normal_mode = nt == 2;
        p.conj_mode = nt == 5;
        expect(orc_frame(&p, h0.data(), 2.25, fh.data(), fx.data(), fz.data(), fn.data(), fj.data(), nt) == 0,
               "oracle frame (threads)");
    }
    expect(orc_frame(&p, h0.data(), 2.25, fh.data(), nullptr, nullptr, nullptr, nullptr, 2) == 0,
           "oracle frame (height only)");
    const double times[3] = {0.5, 1.0, 1.5};
    std::vector<double> tg(3 * nn, 0.1), gh(4 * nn);
    double loss = 0.0, grad[3];
    expect(orc_fit_eval(&p, h0.data(), 3, times, tg.data(), &loss, gh.data(), 2) == 0 && std::isfinite(loss),
           "oracle fit loss + adjoint");
    expect(orc_fit_param_grad(&p, gh.data(), grad) == 0 && std::isfinite(grad[0]), "oracle parameter gradient");
    orc_params bad = p;
    bad.N = 100;
    expect(orc_h0(&bad, h0.data()) != 0, "oracle rejects N=100");
    expect(orc_frame(&p, nullptr, 0.0, fh.data(), nullptr, nullptr, nullptr, nullptr, 1) != 0, "oracle NULL h0");
}

static void abi_checks()
{
    ocn_params p;
    expect_status(ocn_params_default(nullptr), OCN_EINVAL, "params_default(NULL)");
    expect_status(ocn_params_default(&p), OCN_OK, "params_default");
    ocn_handle* h = nullptr;
    expect_status(ocn_create(nullptr, &h), OCN_EINVAL, "create(NULL params)");
    expect_status(ocn_create(&p, nullptr), OCN_EINVAL, "create(NULL out)");
    const double nan = std::numeric_limits<double>::quiet_NaN();
    struct Bad {
        const char* what;
        void (*set)(ocn_params&);
    } bads[] = {
        {"N=100", [](ocn_params& q) { q.N = 100; }},
        {"N=8192", [](ocn_params& q) { q.N = 8192; }},
        {"N=32", [](ocn_params& q) { q.N = 32; }},
        {"L=NaN", [](ocn_params& q) { q.L = std::numeric_limits<double>::quiet_NaN(); }},
        {"L<0", [](ocn_params& q) { q.L = -5.0; }},
        {"wind speed inf", [](ocn_params& q) { q.wind_speed = std::numeric_limits<double>::infinity(); }},
        {"zero wind direction", [](ocn_params& q) { q.wind_dir_x = q.wind_dir_y = 0.0; }},
        {"amplitude<0", [](ocn_params& q) { q.amplitude = -1.0; }},
        {"gravity 0", [](ocn_params& q) { q.gravity = 0.0; }},
        {"conj_mode 7", [](ocn_params& q) { q.conj_mode = 7; }},
        {"normal_mode -1", [](ocn_params& q) { q.normal_mode = -1; }},
    };
    (void)nan;
    for (const Bad& b : bads) {
        ocn_params q = p;
        b.set(q);
        expect_status(ocn_create(&q, &h), OCN_EINVAL, b.what);
        expect(h == nullptr, "no handle on failure");
    }
    expect_status(ocn_destroy(nullptr), OCN_OK, "destroy(NULL)");
    double t[2] = {0.0, 1.0};
    ocn_fields f{};
    int iv = 0;
    void* vp = nullptr;
    float fb[4];
    expect_status(ocn_set_params(nullptr, &p), OCN_EINVAL, "set_params(NULL)");
    expect_status(ocn_get_params(nullptr, &p), OCN_EINVAL, "get_params(NULL)");
    expect_status(ocn_tilde_h0_k(nullptr, nullptr), OCN_EINVAL, "tilde_h0_k(NULL)");
    expect_status(ocn_tilde_h0_k_from_noise(nullptr, nullptr, nullptr), OCN_EINVAL, "tilde_h0_k_from_noise(NULL)");
    expect_status(ocn_generate_twiddle_factors(nullptr), OCN_EINVAL, "generate_twiddle_factors(NULL)");
    expect_status(ocn_tilde_h0_t(nullptr, 1.0, nullptr), OCN_EINVAL, "tilde_h0_t(NULL)");
    expect_status(ocn_butterfly_fft(nullptr, 0, nullptr), OCN_EINVAL, "butterfly_fft(NULL)");
    expect_status(ocn_butterfly_fft_buffers(nullptr, fb, fb, nullptr), OCN_EINVAL, "butterfly_fft_buffers(NULL)");
    expect_status(ocn_generate_normal_map(nullptr, nullptr), OCN_EINVAL, "generate_normal_map(NULL)");
    expect_status(ocn_update(nullptr, 0.0, nullptr), OCN_EINVAL, "update(NULL)");
    expect_status(ocn_update_into(nullptr, 0.0, &f, nullptr), OCN_EINVAL, "update_into(NULL)");
    expect_status(ocn_update_stage(nullptr, t, 2, &f, OCN_STAGE_ALL, nullptr), OCN_EINVAL, "update_stage(NULL)");
    expect_status(ocn_update_frames(nullptr, t, 2, &f, nullptr), OCN_EINVAL, "update_frames(NULL)");
    expect_status(ocn_reserve_frames(nullptr, 32), OCN_EINVAL, "reserve_frames(NULL)");
    expect_status(ocn_frames_reserved(nullptr, &iv), OCN_EINVAL, "frames_reserved(NULL)");
    expect_status(ocn_autotune(nullptr, 1, 1, nullptr), OCN_EINVAL, "autotune(NULL)");
    expect_status(ocn_set_schedule(nullptr, 0, 0), OCN_EINVAL, "set_schedule(NULL)");
    expect_status(ocn_get_schedule(nullptr, 1, &iv, &iv, &iv, &iv, &iv, &iv), OCN_EINVAL, "get_schedule(NULL)");
    expect_status(ocn_get_fields(nullptr, &f), OCN_EINVAL, "get_fields(NULL)");
    expect_status(ocn_get_spectrum(nullptr, &vp, &vp, &vp, &vp), OCN_EINVAL, "get_spectrum(NULL)");
    expect_status(ocn_copy_fields_to_host(nullptr, fb, fb, fb, fb, fb, nullptr), OCN_EINVAL, "copy_fields(NULL)");
    for (int s : {OCN_OK, OCN_EINVAL, OCN_ENOMEM, OCN_EHIP, OCN_ESTATE, OCN_ENOTSUP, -99})
        expect(std::strlen(ocn_status_string(s)) > 0, "status string");

    // host utilities at the largest sizes, into buffers of exactly the documented size
    for (int N : {2, 64, 4096, 65536}) {
        std::vector<int32_t> br(N);
        expect_status(ocn_bit_reversed_indices(N, br.data()), OCN_OK, "bit_reversed_indices");
    }
    for (int N : {2, 256, 4096}) {
        int nb = 0;
        while ((1 << nb) < N) ++nb;
        std::vector<float> tw(4 * (size_t)nb * N);
        expect_status(ocn_twiddle_factors(N, tw.data()), OCN_OK, "twiddle_factors");
    }
    expect_status(ocn_bit_reversed_indices(3, nullptr), OCN_EINVAL, "bit_reversed_indices(3)");
    expect_status(ocn_twiddle_factors(256, nullptr), OCN_EINVAL, "twiddle_factors(NULL)");

    // batches
    ocn_batch* b = nullptr;
    std::vector<ocn_params> mem(3, p);
    expect_status(ocn_batch_create(nullptr, 1, &b), OCN_EINVAL, "batch_create(NULL)");
    expect_status(ocn_batch_create(mem.data(), 0, &b), OCN_EINVAL, "batch_create(0 members)");
    expect_status(ocn_batch_create(mem.data(), OCN_BATCH_MAX_MEMBERS + 1, &b), OCN_EINVAL, "batch_create(too many)");
    mem[2].N = 8192;
    expect_status(ocn_batch_create(mem.data(), 3, &b), OCN_EINVAL, "batch_create(N=8192 member)");
    mem[2] = p;
    mem[1].normal_mode = OCN_NORMALS_BILINEAR;
    expect_status(ocn_batch_create(mem.data(), 3, &b), OCN_EINVAL, "batch_create(bilinear member)");
    expect_status(ocn_batch_destroy(nullptr), OCN_OK, "batch_destroy(NULL)");
    expect_status(ocn_batch_size(nullptr, &iv), OCN_EINVAL, "batch_size(NULL)");
    expect_status(ocn_batch_tilde_h0_k(nullptr, nullptr), OCN_EINVAL, "batch_tilde_h0_k(NULL)");
    expect_status(ocn_batch_update(nullptr, 0.0, nullptr), OCN_EINVAL, "batch_update(NULL)");
    expect_status(ocn_batch_update_graph(nullptr, 0.0, nullptr), OCN_EINVAL, "batch_update_graph(NULL)");
    expect_status(ocn_batch_get_fields(nullptr, 0, &f), OCN_EINVAL, "batch_get_fields(NULL)");
    expect_status(ocn_batch_set_member_fields(nullptr, 0, &f), OCN_EINVAL, "batch_set_member_fields(NULL)");
    expect_status(ocn_batch_set_schedule(nullptr, 256, 8), OCN_EINVAL, "batch_set_schedule(NULL)");
    expect_status(ocn_batch_autotune(nullptr, 3, nullptr), OCN_EINVAL, "batch_autotune(NULL)");

    // fits
    ocn_fit* fit = nullptr;
    std::vector<ocn_params> fm(2, p);
    const double times[3] = {0.5, 1.0, std::numeric_limits<double>::quiet_NaN()};
    expect_status(ocn_fit_create(fm.data(), 2, 0, times, &fit), OCN_EINVAL, "fit_create(K=0)");
    expect_status(ocn_fit_create(fm.data(), 2, OCN_FIT_MAX_FRAMES + 1, times, &fit), OCN_EINVAL, "fit_create(K=17)");
    expect_status(ocn_fit_create(fm.data(), 2, 3, times, &fit), OCN_EINVAL, "fit_create(NaN time)");
    expect_status(ocn_fit_create(fm.data(), 2, 2, nullptr, &fit), OCN_EINVAL, "fit_create(NULL times)");
    fm[1].N = 2048;
    expect_status(ocn_fit_create(fm.data(), 2, 2, times, &fit), OCN_EINVAL, "fit_create(N=2048 member)");
    expect_status(ocn_fit_destroy(nullptr), OCN_OK, "fit_destroy(NULL)");
    expect_status(ocn_fit_forward(nullptr, nullptr), OCN_EINVAL, "fit_forward(NULL)");
    expect_status(ocn_fit_backward(nullptr, nullptr), OCN_EINVAL, "fit_backward(NULL)");
    expect_status(ocn_fit_step(nullptr, nullptr, nullptr), OCN_EINVAL, "fit_step(NULL)");
    expect_status(ocn_fit_step_qn(nullptr, nullptr, nullptr), OCN_EINVAL, "fit_step_qn(NULL)");
    ocn_qn qn{{1.0, 0.05, 0.1}, 2.0, 1.0};
    expect_status(ocn_fit_step_qn(nullptr, &qn, nullptr), OCN_EINVAL, "fit_step_qn(radius > radius_max)");
    qn.radius = 1.0;
    qn.scale[1] = std::numeric_limits<double>::quiet_NaN();
    expect_status(ocn_fit_step_qn(nullptr, &qn, nullptr), OCN_EINVAL, "fit_step_qn(NaN scale)");
    double vars[3], g3[3], l1;
    expect_status(ocn_fit_get(nullptr, 0, vars, g3, &l1, nullptr), OCN_EINVAL, "fit_get(NULL)");

    // graphics interop
    ocn_external* x = nullptr;
    expect_status(ocn_external_import(0, -1, 1 << 20, 0, 1 << 20, &x, &vp), OCN_EINVAL, "external_import(fd -1)");
    expect_status(ocn_external_import(0, 3, 1 << 20, 100, 1000, &x, &vp), OCN_EINVAL, "external_import(misaligned)");
    expect_status(ocn_external_import(0, 3, 4096, 4096, 1 << 20, &x, &vp), OCN_EINVAL, "external_import(past end)");
    expect_status(ocn_external_release(nullptr), OCN_OK, "external_release(NULL)");
}

int main()
{
    oracle_checks();
    abi_checks();
    if (g_fail) {
        std::printf("san_check: %d checks failed\n", g_fail);
        return 1;
    }
    std::printf("san_check OK\n");
    return 0;
}
