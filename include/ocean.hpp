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
// The autotuner's table as a file (ocn_autotune_save / ocn_autotune_load).
inline void autotune_save(const std::string& path) { check(ocn_autotune_save(path.c_str())); }
inline void autotune_load(const std::string& path) { check(ocn_autotune_load(path.c_str())); }

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
    // Setup: intermediates for launches of up to nframes frames (update_frames
    // never allocates; it launches at most this many frames at a time).
    void reserve_frames(int nframes) { check(ocn_reserve_frames(h_, nframes)); }
    int frames_reserved() const
    {
        int n = 0;
        check(ocn_frames_reserved(h_, &n));
        return n;
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

// Many oceans / cascades (uneven N) advanced together: one column launch and
// one row launch per frame (or a replayed two-node HIP graph).  No reference
// counterpart; each member's outputs equal what Ocean::update_into writes.
class Batch {
public:
    Batch(const std::vector<ocn_params>& members, void* stream = nullptr) : stream_(stream)
    {
        check(ocn_batch_create(members.data(), static_cast<int>(members.size()), &b_));
    }
    ~Batch() { ocn_batch_destroy(b_); }
    Batch(const Batch&)            = delete;
    Batch& operator=(const Batch&) = delete;

    void       tilde_h0_k() { check(ocn_batch_tilde_h0_k(b_, stream_)); }
    void       update(double t) { check(ocn_batch_update(b_, t, stream_)); }
    void       update_graph(double t) { check(ocn_batch_update_graph(b_, t, stream_)); }
    // Every member's render export (fields().displacement / render_normals) on or off.
    void       set_render_export(bool enable) { check(ocn_batch_set_render_export(b_, enable ? 1 : 0)); }
    // Rows per band of the row pass for the members of size N, and the autotuner over it.
    void set_schedule(int N, int rows_band) { check(ocn_batch_set_schedule(b_, N, rows_band)); }
    int  schedule(int member) const
    {
        int r = 0;
        check(ocn_batch_get_schedule(b_, member, &r));
        return r;
    }
    void set_column_strips(int N, int cols_per_block) { check(ocn_batch_set_column_strips(b_, N, cols_per_block)); }
    int column_strips(int member) const
    {
        int c = 0;
        check(ocn_batch_get_column_strips(b_, member, &c));
        return c;
    }
    void autotune(int reps = 10) { check(ocn_batch_autotune(b_, reps, stream_)); }
    // One member's outputs at caller-provided device addresses (e.g. ExternalMemory).
    void set_member_fields(int member, const ocn_fields& f) { check(ocn_batch_set_member_fields(b_, member, &f)); }
    ocn_fields fields(int member) const
    {
        ocn_fields f;
        check(ocn_batch_get_fields(b_, member, &f));
        return f;
    }
    int size() const
    {
        int n = 0;
        check(ocn_batch_size(b_, &n));
        return n;
    }
    // Synchronises the stream and copies one member's outputs to the host.
    HostFrame download(int member) const
    {
        ocn_params p;
        check(ocn_batch_get_params(b_, member, &p));
        const size_t nn = static_cast<size_t>(p.N) * p.N;
        HostFrame f{std::vector<float>(nn), std::vector<float>(nn), std::vector<float>(nn),
                    std::vector<float>(nn * 4), std::vector<float>(nn)};
        check(ocn_batch_copy_fields_to_host(b_, member, f.height.data(), f.dx.data(), f.dz.data(), f.normals.data(),
                                            f.jacobian.data(), stream_));
        return f;
    }

private:
    ocn_batch* b_ = nullptr;
    void*      stream_;
};

// The differentiable fit of (wind speed, wind angle, amplitude) to target
// heights: a forward operator (loss) and a backward operator (loss and
// gradients, the spectrum recomputed), plus device-side Adam and
// quasi-Newton steps.  No reference counterpart.
struct FitResult {
    double vars[3];  // wind speed, wind angle (rad), amplitude
    double grad[3];  // dL/d(vars)
    double loss;
};

class Fit {
public:
    Fit(const std::vector<ocn_params>& members, const std::vector<double>& times, void* stream = nullptr)
        : stream_(stream)
    {
        check(ocn_fit_create(members.data(), static_cast<int>(members.size()), static_cast<int>(times.size()),
                             times.data(), &f_));
        for (const ocn_params& p : members) target_sizes_.push_back(times.size() * p.N * p.N);
    }
    ~Fit() { ocn_fit_destroy(f_); }
    Fit(const Fit&)            = delete;
    Fit& operator=(const Fit&) = delete;

    float* targets(int member) const
    {
        float* t = nullptr;
        check(ocn_fit_targets(f_, member, &t));
        return t;
    }
    // K frames of N x N target heights, frame-major.
    void upload_targets(int member, const std::vector<float>& host)
    {
        if (member < 0 || member >= static_cast<int>(target_sizes_.size()) || host.size() != target_sizes_[member])
            throw Error(OCN_EINVAL, "targets must hold K * N * N floats of an existing member");
        check(ocn_fit_upload_targets(f_, member, host.data(), stream_));
    }
    void forward() { check(ocn_fit_forward(f_, stream_)); }
    void backward() { check(ocn_fit_backward(f_, stream_)); }
    void step(const ocn_adam& a) { check(ocn_fit_step(f_, &a, stream_)); }
    void step_qn(const ocn_qn& q) { check(ocn_fit_step_qn(f_, &q, stream_)); }
    void set_vars(int member, double V, double theta, double A)
    {
        const double v[3] = {V, theta, A};
        check(ocn_fit_set_vars(f_, member, v));
    }
    FitResult result(int member) const
    {
        FitResult r{};
        check(ocn_fit_get(f_, member, r.vars, r.grad, &r.loss, stream_));
        return r;
    }
    void* grad_h0(int member) const
    {
        void* g = nullptr;
        check(ocn_fit_grad_h0(f_, member, &g));
        return g;
    }

private:
    ocn_fit*            f_ = nullptr;
    void*               stream_;
    std::vector<size_t> target_sizes_;
};

// Memory a renderer exported as a POSIX file descriptor (a Vulkan
// VkDeviceMemory from vkGetMemoryFdKHR, or an OpenGL memory object), mapped on
// the GPU for the lifetime of this object; point ocn_fields into data() and
// the frames land in the renderer's textures (main.cpp:1096-1099 in the
// reference).  The descriptor is not closed.
class ExternalMemory {
public:
    ExternalMemory(int fd, uint64_t size, uint64_t offset = 0, uint64_t bytes = 0, int device = 0)
    {
        check(ocn_external_import(device, fd, size, offset, bytes ? bytes : size - offset, &e_, &ptr_));
    }
    ~ExternalMemory() { ocn_external_release(e_); }
    ExternalMemory(const ExternalMemory&)            = delete;
    ExternalMemory& operator=(const ExternalMemory&) = delete;

    void* data() const { return ptr_; }

private:
    ocn_external* e_   = nullptr;
    void*         ptr_ = nullptr;
};

}  // namespace ocean
