// ocean_api_example.cpp -- drives the ocean library through the C++ host API
// (include/ocean.hpp), the way the reference's init()/update() drive its
// passes (diharaw/fft-ocean-waves src/main.cpp:199-255).
//
//   ocean_api_example --validate        argument checks only (no GPU needed)
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

int main(int argc, char** argv)
{
    if (argc == 2 && std::strcmp(argv[1], "--validate") == 0) return validate_only();
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
