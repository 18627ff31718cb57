// ocean_api_example.cpp -- drives the ocean library through the C++ host API
// (include/ocean.hpp), the way the reference's init()/update() drive its
// passes (diharaw/fft-ocean-waves src/main.cpp:199-255).
//
//   ocean_api_example --validate        argument checks only (no GPU needed)
//   ocean_api_example --batch-fit       a two-member batch against single
//       oceans, then a noiseless fit of (wind speed, angle, amplitude) to
//       heights of known variables; prints both results, exit 0 on success.
//   ocean_api_example OUT.bin N L V wx wy A l lambda seed t
//       one frame, fused and pass by pass; writes the fused frame's height,
//       dx, dz, normals, jacobian (float32, in that order) to OUT.bin and
//       prints the largest difference between the two paths.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ocean.hpp"

static int validate_only()
{
    int failures = 0;
    auto expect_einval = [&](const char* what, ocn_params p) {
        try {
            ocean::Ocean o(p);
            std::printf("FAIL %s: accepted\n", what);
            ++failures;
        } catch (const ocean::Error& e) {
            const bool ok = e.status() == OCN_EINVAL;
            std::printf("%s %s: %s\n", ok ? "ok" : "FAIL", what, e.what());
            failures += ok ? 0 : 1;
        }
    };
    ocn_params p = ocean::default_params();
    p.N = 100;
    expect_einval("N=100", p);
    p = ocean::default_params();
    p.L = -1.0;
    expect_einval("L<0", p);
    p = ocean::default_params();
    p.wind_dir_x = p.wind_dir_y = 0.0;
    expect_einval("zero wind", p);
    p = ocean::default_params();
    p.conj_mode = 9;
    expect_einval("conj_mode", p);
    return failures ? 1 : 0;
}

static int batch_and_fit()
{
    ocn_params a = ocean::default_params();
    a.N = 64;
    a.L = 100.0;
    a.wind_speed = 12.0;
    ocn_params b = a;
    b.N = 128;
    b.L = 400.0;
    b.wind_speed = 20.0;
    b.cascade_id = 1;
    // batch of two cascades against the single-ocean path
    ocean::Batch batch({a, b});
    batch.tilde_h0_k();
    batch.update(1.25);
    double worst = 0.0;
    for (int i = 0; i < batch.size(); ++i) {
        ocean::Ocean one(i == 0 ? a : b);
        one.init();
        one.update(1.25);
        const ocean::HostFrame x = one.download(), y = batch.download(i);
        double d = 0.0, hmax = 0.0;
        for (size_t k = 0; k < x.height.size(); ++k) {
            d    = std::fmax(d, std::fabs(x.height[k] - y.height[k]));
            hmax = std::fmax(hmax, std::fabs(x.height[k]));
        }
        worst = std::fmax(worst, d / hmax);
        std::printf("batch member %d N=%d: max|height| %.4e, max diff vs single %.3e\n", i, i == 0 ? a.N : b.N, hmax,
                    d);
    }
    // the render export, a band height and the autotuner leave the frame unchanged
    const ocean::HostFrame before = batch.download(1);
    batch.set_render_export(true);
    batch.set_schedule(128, 4);
    batch.update(1.25);
    const bool export_on = batch.fields(0).displacement != nullptr && batch.schedule(1) == 4;
    const bool same_band = batch.download(1).height == before.height;
    batch.autotune(2);
    batch.update_graph(1.25);
    const bool same_tuned = batch.download(1).height == before.height;
    std::printf("batch: render export %s, band 4 frame %s, autotuned band %d frame %s\n", export_on ? "on" : "OFF",
                same_band ? "identical" : "DIFFERS", batch.schedule(1), same_tuned ? "identical" : "DIFFERS");
    if (!(export_on && same_band && same_tuned)) return 1;
    // noiseless fit: targets are heights of member a at two times
    const std::vector<double> times = {0.5, 1.5};
    std::vector<float>        targets;
    {
        ocean::Ocean hidden(a);
        hidden.init();
        for (double t : times) {
            hidden.update(t);
            const ocean::HostFrame f = hidden.download();
            targets.insert(targets.end(), f.height.begin(), f.height.end());
        }
    }
    ocn_params start = a;
    start.wind_speed = 10.0;
    start.amplitude  = 0.8 * a.amplitude;
    ocean::Fit fit({start}, times);
    fit.upload_targets(0, targets);
    fit.forward();
    const double loss0 = fit.result(0).loss;
    ocn_adam adam{{0.1, 0.01, 0.05 * a.amplitude}, 0.9, 0.999, 1e-8};
    for (int s = 0; s < 300; ++s) fit.step(adam);
    const ocean::FitResult r = fit.result(0);
    std::printf("fit: loss %.3e -> %.3e, V %.4f (true %.4f), theta %.5f (true %.5f), A %.5g (true %.5g)\n", loss0,
                r.loss, r.vars[0], a.wind_speed, r.vars[1], std::atan2(a.wind_dir_y, a.wind_dir_x), r.vars[2],
                a.amplitude);
    const bool ok = worst <= 1e-5 && r.loss < 1e-3 * loss0 && std::fabs(r.vars[0] - a.wind_speed) < 0.2 &&
                    std::fabs(r.vars[2] / a.amplitude - 1.0) < 0.02;
    return ok ? 0 : 1;
}

int main(int argc, char** argv)
{
    if (argc == 2 && std::strcmp(argv[1], "--validate") == 0) return validate_only();
    if (argc == 2 && std::strcmp(argv[1], "--batch-fit") == 0) {
        try {
            return batch_and_fit();
        } catch (const ocean::Error& e) {
            std::fprintf(stderr, "ocean error: %s\n", e.what());
            return 1;
        }
    }
    if (argc != 12) {
        std::fprintf(stderr, "usage: %s OUT.bin N L V wx wy A l lambda seed t | --validate\n", argv[0]);
        return 2;
    }
    ocn_params p   = ocean::default_params();
    p.N            = std::atoi(argv[2]);
    p.L            = std::atof(argv[3]);
    p.wind_speed   = std::atof(argv[4]);
    p.wind_dir_x   = std::atof(argv[5]);
    p.wind_dir_y   = std::atof(argv[6]);
    p.amplitude    = std::atof(argv[7]);
    p.suppression  = std::atof(argv[8]);
    p.choppiness   = std::atof(argv[9]);
    p.seed         = std::strtoull(argv[10], nullptr, 10);
    const double t = std::atof(argv[11]);
    try {
        ocean::Ocean ocean(p);
        ocean.init();  // tilde_h0_k + twiddles, as init() does
        const std::vector<int32_t> br = ocean.generate_bit_reversed_indices();
        if (br.size() != (size_t)p.N || br[1] != p.N / 2) {
            std::fprintf(stderr, "bit reversal table wrong\n");
            return 1;
        }
        ocean.update(t);  // fused frame
        const ocean::HostFrame fused = ocean.download();
        ocean.tilde_h0_t(t);  // the same frame pass by pass
        ocean.butterfly_fft(ocean::Field::Dy);
        ocean.butterfly_fft(ocean::Field::Dx);
        ocean.butterfly_fft(ocean::Field::Dz);
        ocean.generate_normal_map();
        const ocean::HostFrame passes = ocean.download();
        double dmax = 0.0, hmax = 0.0;
        for (size_t i = 0; i < fused.height.size(); ++i) {
            dmax = std::fmax(dmax, std::fabs(fused.height[i] - passes.height[i]));
            hmax = std::fmax(hmax, std::fabs(fused.height[i]));
        }
        std::printf("N=%d max|height|=%.6e fused-vs-passes max diff=%.3e\n", p.N, hmax, dmax);
        FILE* f = std::fopen(argv[1], "wb");
        if (!f) return 1;
        for (const std::vector<float>* v : {&fused.height, &fused.dx, &fused.dz, &fused.normals, &fused.jacobian})
            std::fwrite(v->data(), sizeof(float), v->size(), f);
        std::fclose(f);
        return dmax <= 1e-4 * hmax ? 0 : 1;
    } catch (const ocean::Error& e) {
        std::fprintf(stderr, "ocean error: %s\n", e.what());
        return 1;
    }
}
