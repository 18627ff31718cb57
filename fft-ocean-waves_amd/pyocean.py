"""ctypes binding of libocean.so (include/ocean_c.h) plus a minimal HIP runtime
wrapper for device buffers, streams and events.

This is the Python mirror of the reference's ocean passes
(diharaw/fft-ocean-waves src/main.cpp: tilde_h0_k :553, tilde_h0_t :587,
butterfly_fft :612, generate_normal_map :687, generate_twiddle_factors :711,
generate_bit_reversed_indices :733, called from init() :199 / update() :229).
Every compute call goes through the C ABI to the gfx950 kernels; there is no
CPU fallback: loading fails loudly when libocean.so is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libocean.so")

OCN_OK, OCN_EINVAL, OCN_ENOMEM, OCN_EHIP, OCN_ESTATE, OCN_ENOTSUP = 0, -1, -2, -3, -4, -5
OCN_CONJ_REFERENCE, OCN_CONJ_TEXTBOOK = 0, 1
OCN_NORMALS_NEAREST, OCN_NORMALS_BILINEAR = 0, 1
OCN_FIELD_DY, OCN_FIELD_DX, OCN_FIELD_DZ = 0, 1, 2
OCN_STAGE_COLUMNS, OCN_STAGE_ROWS, OCN_STAGE_NORMALS, OCN_STAGE_ALL = 1, 2, 4, 7

# Every function include/ocean_c.h declares (checked by tests/test_capi.py).
EXPORTS = [
    "ocn_params_default", "ocn_create", "ocn_destroy", "ocn_set_params", "ocn_get_params",
    "ocn_last_error", "ocn_status_string", "ocn_abi_version",
    "ocn_tilde_h0_k", "ocn_tilde_h0_k_from_noise", "ocn_generate_twiddle_factors",
    "ocn_tilde_h0_t", "ocn_butterfly_fft", "ocn_butterfly_fft_buffers", "ocn_generate_normal_map",
    "ocn_update", "ocn_update_into", "ocn_update_stage", "ocn_update_frames",
    "ocn_reserve_frames", "ocn_frames_reserved", "ocn_update_frames_noalloc",
    "ocn_autotune", "ocn_autotune_save", "ocn_autotune_load", "ocn_set_schedule", "ocn_get_schedule",
    "ocn_get_fields", "ocn_get_spectrum", "ocn_copy_fields_to_host",
    "ocn_batch_create", "ocn_batch_destroy", "ocn_batch_size", "ocn_batch_get_params",
    "ocn_batch_tilde_h0_k", "ocn_batch_update", "ocn_batch_update_graph",
    "ocn_batch_get_fields", "ocn_batch_get_h0", "ocn_batch_copy_fields_to_host", "ocn_batch_set_render_export",
    "ocn_batch_set_member_fields", "ocn_batch_set_schedule", "ocn_batch_get_schedule", "ocn_batch_autotune",
    "ocn_batch_set_column_strips", "ocn_batch_get_column_strips",
    "ocn_fit_upload_targets",
    "ocn_fit_create", "ocn_fit_destroy", "ocn_fit_targets", "ocn_fit_set_vars",
    "ocn_fit_forward", "ocn_fit_backward", "ocn_fit_step", "ocn_fit_step_qn", "ocn_fit_get", "ocn_fit_grad_h0",
    "ocn_external_import", "ocn_external_release",
    "ocn_bit_reversed_indices", "ocn_twiddle_factors",
]


class OcnParams(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32),
        ("conj_mode", ctypes.c_int32),
        ("normal_mode", ctypes.c_int32),
        ("ocean_id", ctypes.c_uint32),
        ("cascade_id", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("L", ctypes.c_double),
        ("wind_speed", ctypes.c_double),
        ("wind_dir_x", ctypes.c_double),
        ("wind_dir_y", ctypes.c_double),
        ("amplitude", ctypes.c_double),
        ("suppression", ctypes.c_double),
        ("choppiness", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("displacement_scale", ctypes.c_double),
    ]


class OcnFields(ctypes.Structure):
    _fields_ = [("height", ctypes.c_void_p), ("dx", ctypes.c_void_p), ("dz", ctypes.c_void_p),
                ("normals", ctypes.c_void_p), ("jacobian", ctypes.c_void_p),
                ("displacement", ctypes.c_void_p), ("render_normals", ctypes.c_void_p)]


class OcnAdam(ctypes.Structure):
    _fields_ = [("lr", ctypes.c_double * 3), ("beta1", ctypes.c_double), ("beta2", ctypes.c_double),
                ("eps", ctypes.c_double)]


class OcnQn(ctypes.Structure):
    _fields_ = [("scale", ctypes.c_double * 3), ("radius", ctypes.c_double), ("radius_max", ctypes.c_double)]


class OceanError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (status, msg))
        self.status = status


_lib = None


def lib():
    """Loads libocean.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libocean.so is missing: build it with `make -C fft-ocean-waves_amd` "
                               "or __graft_entry__.build(); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        P = ctypes.POINTER(OcnParams)
        F = ctypes.POINTER(OcnFields)
        sig = {
            "ocn_params_default": [P],
            "ocn_create": [P, ctypes.POINTER(vp)],
            "ocn_destroy": [vp],
            "ocn_set_params": [vp, P],
            "ocn_get_params": [vp, P],
            "ocn_abi_version": [],
            "ocn_tilde_h0_k": [vp, vp],
            "ocn_tilde_h0_k_from_noise": [vp, ctypes.POINTER(ctypes.c_uint8), vp],
            "ocn_generate_twiddle_factors": [vp],
            "ocn_tilde_h0_t": [vp, dbl, vp],
            "ocn_butterfly_fft": [vp, i32, vp],
            "ocn_butterfly_fft_buffers": [vp, vp, vp, vp],
            "ocn_generate_normal_map": [vp, vp],
            "ocn_update": [vp, dbl, vp],
            "ocn_update_into": [vp, dbl, F, vp],
            "ocn_update_stage": [vp, ctypes.POINTER(ctypes.c_double), i32, F, i32, vp],
            "ocn_update_frames": [vp, ctypes.POINTER(ctypes.c_double), i32, F, vp],
            "ocn_reserve_frames": [vp, i32],
            "ocn_update_frames_noalloc": [vp, ctypes.POINTER(ctypes.c_double), i32, F, vp],
            "ocn_frames_reserved": [vp, ctypes.POINTER(i32)],
            "ocn_autotune": [vp, i32, i32, vp],
            "ocn_autotune_save": [ctypes.c_char_p],
            "ocn_autotune_load": [ctypes.c_char_p],
            "ocn_set_schedule": [vp, i32, i32],
            "ocn_get_schedule": [vp, i32] + [ctypes.POINTER(i32)] * 6,
            "ocn_get_fields": [vp, F],
            "ocn_get_spectrum": [vp] + [ctypes.POINTER(vp)] * 4,
            "ocn_copy_fields_to_host": [vp] * 7,
            "ocn_bit_reversed_indices": [i32, ctypes.POINTER(ctypes.c_int32)],
            "ocn_twiddle_factors": [i32, ctypes.POINTER(ctypes.c_float)],
            "ocn_batch_create": [P, i32, ctypes.POINTER(vp)],
            "ocn_batch_destroy": [vp],
            "ocn_batch_size": [vp, ctypes.POINTER(i32)],
            "ocn_batch_get_params": [vp, i32, P],
            "ocn_batch_tilde_h0_k": [vp, vp],
            "ocn_batch_update": [vp, dbl, vp],
            "ocn_batch_update_graph": [vp, dbl, vp],
            "ocn_batch_get_fields": [vp, i32, F],
            "ocn_batch_get_h0": [vp, i32, ctypes.POINTER(vp)],
            "ocn_batch_copy_fields_to_host": [vp, i32] + [vp] * 6,
            "ocn_batch_set_render_export": [vp, i32],
            "ocn_batch_set_member_fields": [vp, i32, F],
            "ocn_batch_set_schedule": [vp, i32, i32],
            "ocn_batch_get_schedule": [vp, i32, ctypes.POINTER(i32)],
            "ocn_batch_autotune": [vp, i32, vp],
            "ocn_batch_set_column_strips": [vp, i32, i32],
            "ocn_batch_get_column_strips": [vp, i32, ctypes.POINTER(i32)],
            "ocn_fit_upload_targets": [vp, i32, vp, vp],
            "ocn_fit_create": [P, i32, i32, ctypes.POINTER(dbl), ctypes.POINTER(vp)],
            "ocn_fit_destroy": [vp],
            "ocn_fit_targets": [vp, i32, ctypes.POINTER(vp)],
            "ocn_fit_set_vars": [vp, i32, ctypes.POINTER(dbl)],
            "ocn_fit_forward": [vp, vp],
            "ocn_fit_backward": [vp, vp],
            "ocn_fit_step": [vp, ctypes.POINTER(OcnAdam), vp],
            "ocn_fit_step_qn": [vp, ctypes.POINTER(OcnQn), vp],
            "ocn_fit_get": [vp, i32, ctypes.POINTER(dbl), ctypes.POINTER(dbl), ctypes.POINTER(dbl), vp],
            "ocn_fit_grad_h0": [vp, i32, ctypes.POINTER(vp)],
            "ocn_external_import": [i32, i32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(vp),
                                    ctypes.POINTER(vp)],
            "ocn_external_release": [vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.ocn_last_error.argtypes = []
        L.ocn_last_error.restype = ctypes.c_char_p
        L.ocn_status_string.argtypes = [ctypes.c_int]
        L.ocn_status_string.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != OCN_OK:
        L = lib()
        raise OceanError(L.ocn_status_string(rc).decode(), L.ocn_last_error().decode())
    return rc


def default_params():
    p = OcnParams()
    check(lib().ocn_params_default(ctypes.byref(p)))
    return p


def make_params(N=256, L=1000.0, wind_speed=40.0, wind_dir=(1.0, 1.0), amplitude=4.0, suppression=None,
                choppiness=1.0, gravity=9.81, seed=42, ocean_id=0, cascade_id=0,
                conj_mode=OCN_CONJ_REFERENCE, normal_mode=OCN_NORMALS_NEAREST, displacement_scale=0.5):
    """Parameter block; suppression defaults to L/2000 (the small-wave cutoff rule)."""
    p = default_params()
    p.N, p.L, p.wind_speed = int(N), float(L), float(wind_speed)
    p.wind_dir_x, p.wind_dir_y = float(wind_dir[0]), float(wind_dir[1])
    p.amplitude = float(amplitude)
    p.suppression = float(L / 2000.0 if suppression is None else suppression)
    p.choppiness, p.gravity = float(choppiness), float(gravity)
    p.seed, p.ocean_id, p.cascade_id = int(seed), int(ocean_id), int(cascade_id)
    p.conj_mode, p.normal_mode = int(conj_mode), int(normal_mode)
    p.displacement_scale = float(displacement_scale)
    return p


def bit_reversed_indices(N):
    out = np.zeros(N, dtype=np.int32)
    check(lib().ocn_bit_reversed_indices(N, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return out


def twiddle_factors(N):
    nb = int(round(np.log2(N)))
    out = np.zeros((nb, N, 4), dtype=np.float32)
    check(lib().ocn_twiddle_factors(N, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


# ---------------------------------------------------------------- HIP runtime (ctypes)

class Hip:
    """The few HIP runtime calls the tests and the bench need."""
    H2D, D2H, D2D = 1, 2, 3

    def __init__(self):
        lib()  # libocean.so pulls in the runtime
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        self.rt = ctypes.CDLL(path)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        r = self.rt
        for name, args in {
            "hipMalloc": [ctypes.POINTER(vp), sz], "hipFree": [vp],
            "hipMemcpy": [vp, vp, sz, ctypes.c_int], "hipMemset": [vp, ctypes.c_int, sz],
            "hipDeviceSynchronize": [], "hipGetDeviceCount": [ctypes.POINTER(ctypes.c_int)],
            "hipStreamCreate": [ctypes.POINTER(vp)], "hipStreamDestroy": [vp], "hipStreamSynchronize": [vp],
            "hipEventCreate": [ctypes.POINTER(vp)], "hipEventDestroy": [vp], "hipEventRecord": [vp, vp],
            "hipEventSynchronize": [vp],
            "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), vp, vp],
            "hipGetLastError": [],
            "hipStreamBeginCapture": [vp, ctypes.c_int], "hipStreamEndCapture": [vp, ctypes.POINTER(vp)],
            "hipGraphInstantiate": [ctypes.POINTER(vp), vp, vp, vp, sz], "hipGraphLaunch": [vp, vp],
            "hipGraphExecDestroy": [vp], "hipGraphDestroy": [vp],
        }.items():
            fn = getattr(r, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        r.hipGetErrorString.argtypes = [ctypes.c_int]
        r.hipGetErrorString.restype = ctypes.c_char_p

    def ok(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed: %s (%d)" % (what, self.rt.hipGetErrorString(rc).decode(), rc))

    def device_count(self):
        n = ctypes.c_int(0)
        rc = self.rt.hipGetDeviceCount(ctypes.byref(n))
        return n.value if rc == 0 else 0

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.rt.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc(%d)" % nbytes)
        return p.value

    def free(self, p):
        self.ok(self.rt.hipFree(p), "hipFree")

    def upload(self, dptr, arr):
        a = np.ascontiguousarray(arr)
        self.ok(self.rt.hipMemcpy(dptr, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy H2D")

    def download(self, dptr, shape, dtype=np.float32):
        out = np.empty(shape, dtype=dtype)
        self.ok(self.rt.hipMemcpy(out.ctypes.data, dptr, out.nbytes, self.D2H), "hipMemcpy D2H")
        return out

    def memset(self, dptr, val, nbytes):
        self.ok(self.rt.hipMemset(dptr, val, nbytes), "hipMemset")

    def sync(self):
        self.ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def stream(self):
        s = ctypes.c_void_p()
        self.ok(self.rt.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def stream_sync(self, s):
        self.ok(self.rt.hipStreamSynchronize(s), "hipStreamSynchronize")

    def stream_destroy(self, s):
        self.ok(self.rt.hipStreamDestroy(s), "hipStreamDestroy")

    def capture(self, stream, fn):
        """Records what fn() launches on `stream` into a graph (thread-local
        capture mode); returns (graph, exec)."""
        self.ok(self.rt.hipStreamBeginCapture(stream, 1), "hipStreamBeginCapture")
        g = ctypes.c_void_p()
        try:
            fn()
        finally:
            self.ok(self.rt.hipStreamEndCapture(stream, ctypes.byref(g)), "hipStreamEndCapture")
        ex = ctypes.c_void_p()
        self.ok(self.rt.hipGraphInstantiate(ctypes.byref(ex), g, None, None, 0), "hipGraphInstantiate")
        return g.value, ex.value

    def graph_launch(self, ex, stream):
        self.ok(self.rt.hipGraphLaunch(ex, stream), "hipGraphLaunch")

    def graph_destroy(self, g, ex):
        self.ok(self.rt.hipGraphExecDestroy(ex), "hipGraphExecDestroy")
        self.ok(self.rt.hipGraphDestroy(g), "hipGraphDestroy")

    def event(self):
        e = ctypes.c_void_p()
        self.ok(self.rt.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e.value

    def record(self, ev, stream=None):
        self.ok(self.rt.hipEventRecord(ev, stream), "hipEventRecord")

    def elapsed_ms(self, e0, e1):
        self.ok(self.rt.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = ctypes.c_float()
        self.ok(self.rt.hipEventElapsedTime(ctypes.byref(ms), e0, e1), "hipEventElapsedTime")
        return ms.value

    def event_destroy(self, ev):
        self.ok(self.rt.hipEventDestroy(ev), "hipEventDestroy")


# ---------------------------------------------------------------- Ocean (reference pass structure)

class Ocean:
    """One ocean on the GPU, with the reference's passes as methods."""

    def __init__(self, params, stream=None):
        self.p = params
        self.N = params.N
        self.stream = stream
        h = ctypes.c_void_p()
        check(lib().ocn_create(ctypes.byref(params), ctypes.byref(h)))
        self.h = h.value
        self.hip = Hip()

    def close(self):
        if self.h:
            check(lib().ocn_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # init() passes (main.cpp:217-220)
    def tilde_h0_k(self):
        check(lib().ocn_tilde_h0_k(self.h, self.stream))

    def tilde_h0_k_from_noise(self, noise4):
        n = np.ascontiguousarray(noise4, dtype=np.uint8)
        assert n.shape == (4, 256, 256)
        check(lib().ocn_tilde_h0_k_from_noise(self.h, n.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                              self.stream))

    def generate_twiddle_factors(self):
        check(lib().ocn_generate_twiddle_factors(self.h))

    def set_params(self, params):
        check(lib().ocn_set_params(self.h, ctypes.byref(params)))
        self.p = params

    # update() passes (main.cpp:240-244)
    def tilde_h0_t(self, t):
        check(lib().ocn_tilde_h0_t(self.h, float(t), self.stream))

    def butterfly_fft(self, field):
        check(lib().ocn_butterfly_fft(self.h, int(field), self.stream))

    def butterfly_fft_buffers(self, src_dptr, dst_dptr):
        check(lib().ocn_butterfly_fft_buffers(self.h, src_dptr, dst_dptr, self.stream))

    def generate_normal_map(self):
        check(lib().ocn_generate_normal_map(self.h, self.stream))

    def update(self, t):
        check(lib().ocn_update(self.h, float(t), self.stream))

    def update_into(self, t, fields):
        check(lib().ocn_update_into(self.h, float(t), ctypes.byref(fields), self.stream))

    def update_stage(self, times, fields, stage):
        ts = np.ascontiguousarray(np.atleast_1d(times), dtype=np.float64)
        check(lib().ocn_update_stage(self.h, ts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(ts),
                                     ctypes.byref(fields), int(stage), self.stream))

    def update_frames(self, times, fields):
        ts = np.ascontiguousarray(times, dtype=np.float64)
        check(lib().ocn_update_frames(self.h, ts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(ts),
                                      ctypes.byref(fields), self.stream))

    def reserve_frames(self, nframes):
        """Sizes the intermediates for launches of up to nframes frames (a
        setup call: allocates, and waits for the device when it grows)."""
        check(lib().ocn_reserve_frames(self.h, int(nframes)))

    def frames_reserved(self):
        v = ctypes.c_int()
        check(lib().ocn_frames_reserved(self.h, ctypes.byref(v)))
        return v.value

    # schedules
    def autotune(self, nframes=1, reps=20):
        check(lib().ocn_autotune(self.h, int(nframes), int(reps), self.stream))

    def set_schedule(self, cols_variant, rows_variant):
        check(lib().ocn_set_schedule(self.h, int(cols_variant), int(rows_variant)))

    def schedule(self, nframes=1):
        v = [ctypes.c_int() for _ in range(6)]
        check(lib().ocn_get_schedule(self.h, int(nframes), *[ctypes.byref(x) for x in v]))
        keys = ("cols_variant", "rows_variant", "cols_radix", "cols_per_block", "rows_radix", "rows_band")
        return dict(zip(keys, [x.value for x in v]))

    # buffers
    def fields(self):
        f = OcnFields()
        check(lib().ocn_get_fields(self.h, ctypes.byref(f)))
        return f

    def spectrum_ptrs(self):
        ptrs = [ctypes.c_void_p() for _ in range(4)]
        check(lib().ocn_get_spectrum(self.h, *[ctypes.byref(q) for q in ptrs]))
        return [q.value for q in ptrs]

    def download_fields(self):
        """Synchronises and copies the handle's outputs to host numpy arrays."""
        self.hip.sync()
        f, N = self.fields(), self.N
        return {
            "height": self.hip.download(f.height, (N, N)),
            "dx": self.hip.download(f.dx, (N, N)),
            "dz": self.hip.download(f.dz, (N, N)),
            "normals": self.hip.download(f.normals, (N, N, 4)),
            "jacobian": self.hip.download(f.jacobian, (N, N)),
        }

    def download_h0(self):
        self.hip.sync()
        h0 = self.spectrum_ptrs()[0]
        return self.hip.download(h0, (self.N, self.N, 4))

    def download_spectrum_t(self):
        self.hip.sync()
        _, hy, hx, hz = self.spectrum_ptrs()
        N = self.N
        return [self.hip.download(q, (N, N), np.complex64) for q in (hy, hx, hz)]


# ---------------------------------------------------------------- batches of oceans / cascades

class External:
    """Memory a graphics API exported as a POSIX file descriptor, mapped on the
    GPU (ocn_external_import); `ptr` is its device address.  The frames are
    written there by pointing OcnFields into it.  fd is not closed."""

    def __init__(self, fd, size, offset=0, nbytes=None, device=0):
        self.h = ctypes.c_void_p()
        p = ctypes.c_void_p()
        nbytes = size - offset if nbytes is None else nbytes
        check(lib().ocn_external_import(int(device), int(fd), int(size), int(offset), int(nbytes),
                                        ctypes.byref(self.h), ctypes.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    def close(self):
        if self.h:
            check(lib().ocn_external_release(self.h))
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def texture_fields(base, N, frames=1, align=256):
    """OcnFields for the reference's five output textures laid out one after
    another from device address `base` (height, dx, dz as R32F, normals as
    RGBA32F, Jacobian R32F, each `frames` slices deep as ocn_update_frames
    writes them; each start rounded up to `align` bytes), and the bytes the
    block spans."""
    sizes = (4, 4, 4, 16, 4)
    ptrs, off = [], 0
    for b in sizes:
        off = (off + align - 1) // align * align
        ptrs.append(base + off)
        off += frames * N * N * b
    return OcnFields(*ptrs), off


def cascade_params(n_oceans=64, cascades=((2000.0, 256), (250.0, 512), (30.0, 1024)), seed=42,
                   wind_speed=(5.0, 30.0), choppiness=(0.8, 1.5), amplitude=4.0):
    """Members of a batch of independent oceans, each with several cascades.

    Ocean o draws its wind speed ~ U(wind_speed), direction angle ~ U(0, 2 pi)
    and choppiness ~ U(choppiness) from numpy's Philox generator keyed by
    seed + o; all its cascades share them.  Cascade c of ocean o has the
    given (L, N), suppression L/2000, and noise seed `seed` with
    ocean_id = o, cascade_id = c in the Philox counter, so every member's
    noise is independent.  Member order: ocean-major."""
    out = []
    for o in range(n_oceans):
        rng = np.random.Generator(np.random.Philox(key=seed + o))
        v = float(rng.uniform(*wind_speed))
        th = float(rng.uniform(0.0, 2.0 * np.pi))
        lam = float(rng.uniform(*choppiness))
        for c, (L, N) in enumerate(cascades):
            out.append(make_params(N=N, L=L, wind_speed=v, wind_dir=(np.cos(th), np.sin(th)), amplitude=amplitude,
                                   choppiness=lam, seed=seed, ocean_id=o, cascade_id=c))
    return out


class Batch:
    """Many oceans / cascades (uneven N) computed by one pair of launches per frame."""

    def __init__(self, members, stream=None):
        arr = (OcnParams * len(members))(*members)
        h = ctypes.c_void_p()
        check(lib().ocn_batch_create(arr, len(members), ctypes.byref(h)))
        self.h = h.value
        self.members = list(members)
        self.stream = stream
        self.hip = Hip()

    def close(self):
        if self.h:
            check(lib().ocn_batch_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        n = ctypes.c_int()
        check(lib().ocn_batch_size(self.h, ctypes.byref(n)))
        return n.value

    def params(self, member):
        p = OcnParams()
        check(lib().ocn_batch_get_params(self.h, int(member), ctypes.byref(p)))
        return p

    def tilde_h0_k(self):
        check(lib().ocn_batch_tilde_h0_k(self.h, self.stream))

    def update(self, t):
        check(lib().ocn_batch_update(self.h, float(t), self.stream))

    def update_graph(self, t):
        check(lib().ocn_batch_update_graph(self.h, float(t), self.stream))

    def fields(self, member):
        f = OcnFields()
        check(lib().ocn_batch_get_fields(self.h, int(member), ctypes.byref(f)))
        return f

    def h0_ptr(self, member):
        q = ctypes.c_void_p()
        check(lib().ocn_batch_get_h0(self.h, int(member), ctypes.byref(q)))
        return q.value

    def download_fields(self, member):
        self.hip.sync()
        f, N = self.fields(member), self.members[member].N
        return {
            "height": self.hip.download(f.height, (N, N)),
            "dx": self.hip.download(f.dx, (N, N)),
            "dz": self.hip.download(f.dz, (N, N)),
            "normals": self.hip.download(f.normals, (N, N, 4)),
            "jacobian": self.hip.download(f.jacobian, (N, N)),
        }

    def download_h0(self, member):
        self.hip.sync()
        N = self.members[member].N
        return self.hip.download(self.h0_ptr(member), (N, N, 4))

    def set_render_export(self, enable=True):
        check(lib().ocn_batch_set_render_export(self.h, 1 if enable else 0))

    def set_schedule(self, N, rows_band):
        check(lib().ocn_batch_set_schedule(self.h, int(N), int(rows_band)))

    def schedule(self, member):
        r = ctypes.c_int()
        check(lib().ocn_batch_get_schedule(self.h, int(member), ctypes.byref(r)))
        return r.value

    def autotune(self, reps=10):
        check(lib().ocn_batch_autotune(self.h, int(reps), self.stream))

    def set_column_strips(self, N, cols_per_block):
        check(lib().ocn_batch_set_column_strips(self.h, int(N), int(cols_per_block)))

    def column_strips(self, member):
        v = ctypes.c_int()
        check(lib().ocn_batch_get_column_strips(self.h, int(member), ctypes.byref(v)))
        return v.value

    def set_member_fields(self, member, fields):
        """Outputs of one member written at caller-provided device addresses."""
        check(lib().ocn_batch_set_member_fields(self.h, int(member), ctypes.byref(fields)))

    def download_render(self, member):
        """The member's render export (displacement, render_normals), N x N x 4 each."""
        self.hip.sync()
        f, N = self.fields(member), self.members[member].N
        if not f.displacement:
            raise OceanError("OCN_ESTATE", "the render export is off")
        return {"displacement": self.hip.download(f.displacement, (N, N, 4)),
                "render_normals": self.hip.download(f.render_normals, (N, N, 4))}


def autotune_save(path):
    """Writes the autotuner's table (every ocn_autotune result of this process)."""
    check(lib().ocn_autotune_save(os.fsencode(path)))


def autotune_load(path):
    """Records the entries of a file autotune_save wrote."""
    check(lib().ocn_autotune_load(os.fsencode(path)))


# ---------------------------------------------------------------- differentiable fits

def adam(lr=(0.2, 0.02, 0.05), beta1=0.9, beta2=0.999, eps=1e-8):
    """Adam settings; the default rates (m/s, rad, amplitude per step) fit the
    N=512 workload in 200 steps."""
    a = OcnAdam()
    for i in range(3):
        a.lr[i] = float(lr[i])
    a.beta1, a.beta2, a.eps = float(beta1), float(beta2), float(eps)
    return a


def qn(scale=(1.0, 0.05, 0.1), radius=1.0, radius_max=20.0):
    """Quasi-Newton settings: scales of (wind speed m/s, angle rad, amplitude),
    the units of the trust region, and its initial and largest radius."""
    q = OcnQn()
    for i in range(3):
        q.scale[i] = float(scale[i])
    q.radius, q.radius_max = float(radius), float(radius_max)
    return q


def fit_batch_params(n_oceans=256, N=256, L=500.0, seed=42, start=(17.5, 0.0, 3.5)):
    """Hidden and starting members of a batched fit of independent oceans.
    Ocean o draws its hidden wind speed ~ U(5, 30) m/s, angle ~ U(0, 2 pi) and
    amplitude ~ U(2, 5) from numpy's Philox generator keyed by seed + o; its
    noise seed is seed + o with ocean_id o; suppression L/2000.  Every ocean
    starts from `start` (speed, angle, amplitude).  Returns (hidden, starts)."""
    hidden, starts = [], []
    for o in range(n_oceans):
        rng = np.random.Generator(np.random.Philox(key=seed + o))
        v, th, a = float(rng.uniform(5.0, 30.0)), float(rng.uniform(0.0, 2.0 * np.pi)), float(rng.uniform(2.0, 5.0))
        common = dict(N=N, L=L, seed=seed + o, ocean_id=o)
        hidden.append(make_params(wind_speed=v, wind_dir=(np.cos(th), np.sin(th)), amplitude=a, **common))
        starts.append(make_params(wind_speed=start[0], wind_dir=(np.cos(start[1]), np.sin(start[1])),
                                  amplitude=start[2], **common))
    return hidden, starts


class Fit:
    """Fits (wind speed, wind angle, amplitude) of one or many oceans to target
    heights at K shared times: forward (loss) and backward (loss + gradients)
    operators and a device-side Adam step."""

    def __init__(self, members, times, stream=None):
        arr = (OcnParams * len(members))(*members)
        ts = np.ascontiguousarray(times, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().ocn_fit_create(arr, len(members), len(ts), ts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   ctypes.byref(h)))
        self.h = h.value
        self.members = list(members)
        self.times = ts
        self.stream = stream
        self.hip = Hip()

    def close(self):
        if self.h:
            check(lib().ocn_fit_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def targets_ptr(self, member):
        q = ctypes.c_void_p()
        check(lib().ocn_fit_targets(self.h, int(member), ctypes.byref(q)))
        return q.value

    def upload_targets(self, member, targets):
        N = self.members[member].N
        t = np.ascontiguousarray(targets, dtype=np.float32)
        assert t.shape == (len(self.times), N, N)
        check(lib().ocn_fit_upload_targets(self.h, int(member), t.ctypes.data, self.stream))

    def set_vars(self, member, V, theta, A):
        v = (ctypes.c_double * 3)(V, theta, A)
        check(lib().ocn_fit_set_vars(self.h, int(member), v))

    def forward(self):
        check(lib().ocn_fit_forward(self.h, self.stream))

    def backward(self):
        check(lib().ocn_fit_backward(self.h, self.stream))

    def step(self, adam_cfg):
        check(lib().ocn_fit_step(self.h, ctypes.byref(adam_cfg), self.stream))

    def step_qn(self, qn_cfg):
        check(lib().ocn_fit_step_qn(self.h, ctypes.byref(qn_cfg), self.stream))

    def get(self, member):
        """(vars (V, theta, A), grad, loss) of one member; synchronises."""
        v = (ctypes.c_double * 3)()
        g = (ctypes.c_double * 3)()
        l = ctypes.c_double()
        check(lib().ocn_fit_get(self.h, int(member), v, g, ctypes.byref(l), self.stream))
        return np.array(v[:]), np.array(g[:]), l.value

    def grad_h0(self, member):
        q = ctypes.c_void_p()
        check(lib().ocn_fit_grad_h0(self.h, int(member), ctypes.byref(q)))
        self.hip.sync()
        N = self.members[member].N
        return self.hip.download(q.value, (N, N, 4))
