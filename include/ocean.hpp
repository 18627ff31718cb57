// ocean.hpp -- C++ host API of the MI355X ocean library, mirroring the
// reference's pass structure (diharaw/fft-ocean-waves src/main.cpp):
//
//   FFTOceanWaves::init()    main.cpp:199-225  -> ocean::Ocean(params) + init()
//     tilde_h0_k()           main.cpp:553-583  -> tilde_h0_k()
//     generate_bit_reversed_indices()  :733-744 -> generate_bit_reversed_indices()
//     generate_twiddle_factors()       :711-729 -> generate_twiddle_factors()
//   FFTOceanWaves::update()  main.cpp:229-255  -> update(t)
//     tilde_h0_t()           main.cpp:587-608  -> tilde_h0_t(t)
//     butterfly_fft(src,dst) main.cpp:612-683  -> butterfly_fft(field)
//     generate_normal_map()  main.cpp:687-707  -> generate_normal_map()
//   textures m_dy/m_dx/m_dz/m_normal_map (main.cpp:1096-1099) -> fields()
//
// Header-only; every call goes through the C ABI in ocean_c.h (libocean.so),
// which never throws.  This layer turns non-OK statuses into ocean::Error.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ocean_c.h"

namespace ocean {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& msg)
        : std::runtime_error(std::string(ocn_status_string(status)) + ": " + msg), status_(status)
    {
    }
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status)
{
    if (status != OCN_OK) throw Error(status, ocn_last_error());
}

// Parameter block with the reference's defaults (main.cpp:1632-1646).
inline ocn_params default_params()
{
    ocn_params p;
    check(ocn_params_default(&p));
    return p;
}

enum class Field : int { Dy = OCN_FIELD_DY, Dx = OCN_FIELD_DX, Dz = OCN_FIELD_DZ };

// Host copies of one frame's outputs (row-major [ty][tx]).
struct HostFrame {
    std::vector<float> height, dx, dz, normals, jacobian;
};

class Ocean {
public:
    // create_textures (main.cpp:1083-1099): all device resources of one ocean
    // on the current HIP device.  `stream` is a hipStream_t (nullptr = null stream).
    explicit Ocean(const ocn_params& p, void* stream = nullptr) : stream_(stream)
    {
        check(ocn_create(&p, &h_));
    }
    ~Ocean() { ocn_destroy(h_); }
    Ocean(const Ocean&)            = delete;
    Ocean& operator=(const Ocean&) = delete;
    Ocean(Ocean&& o) noexcept : h_(o.h_), stream_(o.stream_) { o.h_ = nullptr; }

    // The FFT precomputations of init() (main.cpp:217-220).
    void init()
    {
        tilde_h0_k();
        generate_twiddle_factors();
    }

    void tilde_h0_k() { check(ocn_tilde_h0_k(h_, stream_)); }
    void tilde_h0_k(const std::vector<uint8_t>& noise4x256x256)
    {
        if (noise4x256x256.size() != 4u * 256u * 256u) throw Error(OCN_EINVAL, "noise must be 4 x 256 x 256 bytes");
        check(ocn_tilde_h0_k_from_noise(h_, noise4x256x256.data(), stream_));
    }
    std::vector<int32_t> generate_bit_reversed_indices() const
    {
        std::vector<int32_t> idx(params().N);
        check(ocn_bit_reversed_indices(params().N, idx.data()));
        return idx;
    }
    void generate_twiddle_factors() { check(ocn_generate_twiddle_factors(h_)); }

    // update() (main.cpp:240-244), unfused pass by pass ...
    void tilde_h0_t(double t) { check(ocn_tilde_h0_t(h_, t, stream_)); }
    void butterfly_fft(Field f) { check(ocn_butterfly_fft(h_, static_cast<int>(f), stream_)); }
    void generate_normal_map() { check(ocn_generate_normal_map(h_, stream_)); }
    // ... or fused (the same outputs in three kernels).
    void update(double t) { check(ocn_update(h_, t, stream_)); }
    void update_into(double t, const ocn_fields& out) { check(ocn_update_into(h_, t, &out, stream_)); }
    void update_frames(const std::vector<double>& t, const ocn_fields& out)
    {
        check(ocn_update_frames(h_, t.data(), static_cast<int>(t.size()), &out, stream_));
    }

    void set_params(const ocn_params& p) { check(ocn_set_params(h_, &p)); }
    ocn_params params() const
    {
        ocn_params p;
        check(ocn_get_params(h_, &p));
        return p;
    }
    void autotune(int nframes = 1, int reps = 20) { check(ocn_autotune(h_, nframes, reps, stream_)); }

    // Device pointers of the output textures.
    ocn_fields fields() const
    {
        ocn_fields f;
        check(ocn_get_fields(h_, &f));
        return f;
    }
    // Synchronises the stream and copies the outputs to the host.
    HostFrame download() const
    {
        const size_t nn = static_cast<size_t>(params().N) * params().N;
        HostFrame f{std::vector<float>(nn), std::vector<float>(nn), std::vector<float>(nn),
                    std::vector<float>(nn * 4), std::vector<float>(nn)};
        check(ocn_copy_fields_to_host(h_, f.height.data(), f.dx.data(), f.dz.data(), f.normals.data(),
                                      f.jacobian.data(), stream_));
        return f;
    }

    ocn_handle* handle() const { return h_; }

private:
    ocn_handle* h_ = nullptr;
    void*       stream_;
};

}  // namespace ocean
