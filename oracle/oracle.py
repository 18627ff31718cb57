"""ctypes wrapper around the FP64 CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check its frame, and by bench.py's cpu_baseline leg.  The product library never
imports it.  Each wrapper names the reference shader/host code it restates
(see ocean_oracle.c for the file:line citations and for what pins each one).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int32),
        ("L", ctypes.c_double),
        ("wind_speed", ctypes.c_double),
        ("wind_dir_x", ctypes.c_double),
        ("wind_dir_y", ctypes.c_double),
        ("amplitude", ctypes.c_double),
        ("suppression", ctypes.c_double),
        ("choppiness", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("ocean_id", ctypes.c_uint32),
        ("cascade_id", ctypes.c_uint32),
        ("conj_mode", ctypes.c_int32),
        ("normal_mode", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_reverse_bits.argtypes = [ctypes.c_uint32, ctypes.c_int]
        L.orc_reverse_bits.restype = ctypes.c_uint32
        L.orc_log2n.argtypes = [ctypes.c_int]
        L.orc_bit_reversed_indices.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
        L.orc_twiddle_factors.argtypes = [ctypes.c_int, dp]
        L.orc_butterfly_fft2d.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int]
        L.orc_inversion.argtypes = [ctypes.c_int, dp, dp]
        L.orc_inversion.restype = None
        L.orc_normal_map.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp]
        L.orc_normal_map.restype = None
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.orc_philox4x32_10.restype = None
        L.orc_u01.argtypes = [ctypes.c_uint32]
        L.orc_u01.restype = ctypes.c_double
        L.orc_h0.argtypes = [ctypes.POINTER(OrcParams), dp]
        L.orc_h0_from_noise.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(ctypes.c_uint8), dp]
        L.orc_phillips_amp.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int, ctypes.c_int]
        L.orc_phillips_amp.restype = ctypes.c_double
        L.orc_spectrum_t.argtypes = [ctypes.POINTER(OrcParams), dp, ctypes.c_double, dp, dp, dp]
        L.orc_jacobian.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.orc_jacobian.restype = None
        L.orc_frame.argtypes = [ctypes.POINTER(OrcParams), dp, ctypes.c_double, dp, dp, dp, dp, dp,
                                ctypes.c_int]
        L.orc_render.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp, dp, dp, dp]
        L.orc_render.restype = None
        L.orc_fit_eval.argtypes = [ctypes.POINTER(OrcParams), dp, ctypes.c_int, dp, dp, dp, dp, ctypes.c_int]
        L.orc_fit_param_grad.argtypes = [ctypes.POINTER(OrcParams), dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(N=256, L=1000.0, wind_speed=40.0, wind_dir=(1.0, 1.0), amplitude=4.0,
                suppression=None, choppiness=1.0, gravity=9.81, seed=42, ocean_id=0,
                cascade_id=0, conj_mode=0, normal_mode=0):
    """Parameter block; suppression defaults to L/2000 (the small-wave cutoff rule)."""
    if suppression is None:
        suppression = L / 2000.0
    return OrcParams(int(N), float(L), float(wind_speed), float(wind_dir[0]), float(wind_dir[1]),
                     float(amplitude), float(suppression), float(choppiness), float(gravity),
                     int(seed), int(ocean_id), int(cascade_id), int(conj_mode), int(normal_mode))


def reverse_bits(v, nbits):
    return int(lib().orc_reverse_bits(v, nbits))


def bit_reversed_indices(N):
    out = np.zeros(N, dtype=np.int32)
    rc = lib().orc_bit_reversed_indices(N, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert rc == 0
    return out


def twiddle_factors(N):
    nb = int(lib().orc_log2n(N))
    out = np.zeros((nb, N, 4), dtype=np.float64)
    assert lib().orc_twiddle_factors(N, _dp(out)) == 0
    return out


def butterfly_fft2d(data, nthreads=1):
    """Reference radix-2 ping-pong IFFT (unnormalised, e^{+}) of a complex NxN array."""
    N = data.shape[0]
    d = np.ascontiguousarray(data, dtype=np.complex128).copy()
    scratch = np.zeros_like(d)
    tw = twiddle_factors(N)
    rc = lib().orc_butterfly_fft2d(N, _dp(tw), d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   scratch.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nthreads)
    assert rc == 0
    return d


def inversion(data):
    N = data.shape[0]
    d = np.ascontiguousarray(data, dtype=np.complex128)
    out = np.zeros((N, N), dtype=np.float64)
    lib().orc_inversion(N, d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _dp(out))
    return out


def normal_map(h, mode=0):
    N = h.shape[0]
    hh = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros((N, N, 4), dtype=np.float64)
    lib().orc_normal_map(N, _dp(hh), mode, _dp(out))
    return out


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def u01(x):
    return float(lib().orc_u01(x))


def h0(params):
    N = params.N
    out = np.zeros((N, N, 4), dtype=np.float64)
    assert lib().orc_h0(ctypes.byref(params), _dp(out)) == 0
    return out


def h0_from_noise(params, noise4):
    N = params.N
    n = np.ascontiguousarray(noise4, dtype=np.uint8)
    assert n.shape == (4, 256, 256)
    out = np.zeros((N, N, 4), dtype=np.float64)
    assert lib().orc_h0_from_noise(ctypes.byref(params), n.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                   _dp(out)) == 0
    return out


def phillips_amp(params, tx, ty):
    return float(lib().orc_phillips_amp(ctypes.byref(params), tx, ty))


def spectrum_t(params, h0arr, t):
    N = params.N
    a = np.ascontiguousarray(h0arr, dtype=np.float64)
    hy = np.zeros((N, N), dtype=np.complex128)
    hx = np.zeros_like(hy)
    hz = np.zeros_like(hy)
    cp = lambda z: z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib().orc_spectrum_t(ctypes.byref(params), _dp(a), float(t), cp(hy), cp(hx), cp(hz)) == 0
    return hy, hx, hz


def jacobian(N, L, lam, dx, dz):
    out = np.zeros((N, N), dtype=np.float64)
    a = np.ascontiguousarray(dx, dtype=np.float64)
    b = np.ascontiguousarray(dz, dtype=np.float64)
    lib().orc_jacobian(N, float(L), float(lam), _dp(a), _dp(b), _dp(out))
    return out


def frame(params, h0arr, t, nthreads=1, want=("height", "dx", "dz", "normals", "jacobian")):
    """Whole reference frame; returns a dict of float64 arrays."""
    N = params.N
    a = np.ascontiguousarray(h0arr, dtype=np.float64)
    res = {}
    ptrs = []
    for name, shape in (("height", (N, N)), ("dx", (N, N)), ("dz", (N, N)),
                        ("normals", (N, N, 4)), ("jacobian", (N, N))):
        if name in want:
            res[name] = np.zeros(shape, dtype=np.float64)
            ptrs.append(_dp(res[name]))
        else:
            ptrs.append(None)
    rc = lib().orc_frame(ctypes.byref(params), _dp(a), float(t), *ptrs, int(nthreads))
    assert rc == 0
    return res


def fit_eval(params, h0arr, times, targets, want_grad=True, nthreads=1):
    """Loss of K height frames against targets (K, N, N) and, if asked,
    dL/dh0 (N, N, 4) by the FP64 adjoint."""
    N = params.N
    a = np.ascontiguousarray(h0arr, dtype=np.float64)
    ts = np.ascontiguousarray(times, dtype=np.float64)
    tg = np.ascontiguousarray(targets, dtype=np.float64)
    assert tg.shape == (len(ts), N, N)
    loss = ctypes.c_double()
    g = np.zeros((N, N, 4), dtype=np.float64) if want_grad else None
    rc = lib().orc_fit_eval(ctypes.byref(params), _dp(a), len(ts), _dp(ts), _dp(tg), ctypes.byref(loss),
                            _dp(g) if want_grad else None, int(nthreads))
    assert rc == 0
    return (loss.value, g) if want_grad else loss.value


def fit_param_grad(params, grad_h0):
    g = np.ascontiguousarray(grad_h0, dtype=np.float64)
    out = np.zeros(3, dtype=np.float64)
    assert lib().orc_fit_param_grad(ctypes.byref(params), _dp(g), _dp(out)) == 0
    return out


def render(N, L, lam, scale, h, dx, dz):
    """Render export: (displacement (N, N, 4), normals of the displaced surface (N, N, 4))."""
    a, b, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (h, dx, dz))
    disp = np.zeros((N, N, 4), dtype=np.float64)
    rn = np.zeros((N, N, 4), dtype=np.float64)
    lib().orc_render(N, float(L), float(lam), float(scale), _dp(a), _dp(b), _dp(c), _dp(disp), _dp(rn))
    return disp, rn
