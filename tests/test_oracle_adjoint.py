"""Pins of the FP64 oracle's differentiable fit (oracle/ocean_oracle.c:
orc_fit_eval, orc_fit_param_grad).  No reference counterpart exists: the loss
is pinned by an independent numpy inverse DFT, the adjoint by FP64 central
differences and the directional-derivative identity <grad L, d> = dL(h0 + e d)/de
(which holds over every texel at once), in both conj modes."""
import math

import numpy as np
import pytest

import oracle


def _params(N=32, conj=0, V=22.0, theta=math.radians(30.0), A=3.5, L=200.0, **kw):
    return oracle.make_params(N=N, L=L, wind_speed=V, wind_dir=(math.cos(theta), math.sin(theta)), amplitude=A,
                              conj_mode=conj, **kw)


def _heights(p, h0, times):
    return np.stack([oracle.frame(p, h0, t, want=("height",))["height"] for t in times])


def _targets(p_hidden, times, sigma=0.05, seed=7):
    rng = np.random.Generator(np.random.Philox(key=seed))
    h = _heights(p_hidden, oracle.h0(p_hidden), times)
    return h + sigma * rng.standard_normal(h.shape)


def test_loss_matches_numpy_inverse_dft():
    p = _params(N=32)
    times = [0.5, 1.0, 2.5]
    h0 = oracle.h0(p)
    tg = _targets(_params(N=32, V=18.0, A=3.0), times)
    N = p.N
    perm = np.where((np.add.outer(np.arange(N), np.arange(N)) & 1) == 1, -1.0, 1.0)
    ref = 0.0
    for k, t in enumerate(times):
        hy = oracle.spectrum_t(p, h0, t)[0]
        h = perm * np.fft.ifft2(hy).real          # (-1)^(x+y) Re(sum H e^{+}) / N^2
        ref += ((h - tg[k]) ** 2).sum()
    ref /= len(times) * N * N
    loss = oracle.fit_eval(p, h0, times, tg, want_grad=False)
    assert abs(loss - ref) <= 1e-12 * ref


def test_loss_and_gradient_vanish_on_exact_targets():
    p = _params(N=16)
    times = [0.3, 0.9]
    h0 = oracle.h0(p)
    loss, g = oracle.fit_eval(p, h0, times, _heights(p, h0, times))
    assert loss < 1e-28 and np.abs(g).max() < 1e-14


@pytest.mark.parametrize("conj", [0, 1])
def test_h0_gradient_directional_identity(conj):
    """<dL/dh0, d> equals the exact directional derivative 2/(K N^2) sum (h - T) F(d),
    F(d) the heights of the direction d used as a spectrum: every texel and
    component of the adjoint at once."""
    p = _params(N=32, conj=conj)
    times = [0.5, 1.5, 4.0]
    h0 = oracle.h0(p)
    tg = _targets(_params(N=32, conj=conj, V=25.0), times)
    loss, g = oracle.fit_eval(p, h0, times, tg)
    rng = np.random.Generator(np.random.Philox(key=11))
    d = rng.standard_normal(h0.shape)
    res = _heights(p, h0, times) - tg
    Fd = _heights(p, d, times)
    K, N = len(times), p.N
    exact = 2.0 / (K * N * N) * (res * Fd).sum()
    got = (g * d).sum()
    assert abs(got - exact) <= 1e-10 * abs(exact)


@pytest.mark.parametrize("conj", [0, 1])
def test_h0_gradient_matches_central_differences(conj):
    p = _params(N=16, conj=conj)
    times = [0.5, 2.0]
    h0 = oracle.h0(p)
    tg = _targets(_params(N=16, conj=conj, V=15.0, A=4.0), times)
    _, g = oracle.fit_eval(p, h0, times, tg)
    # the loss is quadratic in h0, so central differences are exact up to
    # rounding for any step; a step of 1e-3 keeps the loss rounding (about
    # 1e-15 L / step) far below the tolerance
    eps = 1e-3
    for (ty, tx) in [(8, 8), (8, 9), (3, 12), (0, 5), (12, 0), (15, 15)]:
        for c in range(4):
            hp, hm = h0.copy(), h0.copy()
            hp[ty, tx, c] += eps
            hm[ty, tx, c] -= eps
            fd = (oracle.fit_eval(p, hp, times, tg, want_grad=False) -
                  oracle.fit_eval(p, hm, times, tg, want_grad=False)) / (2 * eps)
            assert abs(g[ty, tx, c] - fd) <= 1e-6 * np.abs(g).max(), (ty, tx, c, g[ty, tx, c], fd)


@pytest.mark.parametrize("conj", [0, 1])
def test_parameter_gradient_matches_central_differences(conj):
    """dL/d(V, theta, A) through the Phillips amplitude against FP64 central
    differences of the loss with h = 1e-6 (h0 regenerated from Philox)."""
    V, th, A = 20.0, math.radians(40.0), 3.0
    times = [0.5, 1.0, 1.5, 2.0]
    tg = _targets(_params(N=64, conj=conj, V=22.0, theta=math.radians(30.0), A=3.5), times)

    def loss(V_, th_, A_):
        q = _params(N=64, conj=conj, V=V_, theta=th_, A=A_)
        return oracle.fit_eval(q, oracle.h0(q), times, tg, want_grad=False)

    p = _params(N=64, conj=conj, V=V, theta=th, A=A)
    _, gh = oracle.fit_eval(p, oracle.h0(p), times, tg)
    grad = oracle.fit_param_grad(p, gh)
    h = 1e-6
    fd = np.array([(loss(V + h, th, A) - loss(V - h, th, A)) / (2 * h),
                   (loss(V, th + h, A) - loss(V, th - h, A)) / (2 * h),
                   (loss(V, th, A + h) - loss(V, th, A - h)) / (2 * h)])
    assert np.all(np.abs(fd) > 0)
    np.testing.assert_allclose(grad, fd, rtol=1e-5)


def test_amplitude_gradient_closed_form():
    """h0 is proportional to sqrt(A), so the heights scale as sqrt(A) and
    A dL/dA = 1/(K N^2) sum (h - T) h exactly."""
    p = _params(N=32, A=2.0)
    times = [0.7, 1.4]
    h0 = oracle.h0(p)
    tg = _targets(_params(N=32), times)
    _, g = oracle.fit_eval(p, h0, times, tg)
    grad = oracle.fit_param_grad(p, g)
    h = _heights(p, h0, times)
    K, N = len(times), p.N
    ref = ((h - tg) * h).sum() / (K * N * N) / p.amplitude
    assert abs(grad[2] - ref) <= 1e-10 * abs(ref)


@pytest.mark.parametrize("theta_deg", [45.0, 90.0, 135.0])
def test_angle_gradient_at_lattice_kinks_matches_central_differences(theta_deg):
    """At 45/90/135 degrees whole lattice lines are perpendicular to the wind,
    where s = c |k_hat . w| has a kink in theta.  Central differences measure
    the symmetric derivative there, which the adjoint must reproduce (those
    texels contribute no angle term)."""
    th = math.radians(theta_deg)
    times = [0.5, 1.0, 1.5]
    tg = _targets(_params(N=32, V=14.0, theta=math.radians(20.0), A=3.0), times)

    def loss(t_):
        q = _params(N=32, V=12.0, theta=t_, A=2.5)
        return oracle.fit_eval(q, oracle.h0(q), times, tg, want_grad=False)

    p = _params(N=32, V=12.0, theta=th, A=2.5)
    _, gh = oracle.fit_eval(p, oracle.h0(p), times, tg)
    g = oracle.fit_param_grad(p, gh)[1]
    h = 1e-6
    fd = (loss(th + h) - loss(th - h)) / (2 * h)
    assert abs(g - fd) <= 1e-5 * abs(fd), (g, fd)
