"""GPU parity of the differentiable fit (ocn_fit_*): loss, dL/d(V, theta, A)
and dL/dh0 from the FP32 adjoint kernels against the FP64 oracle adjoint
(itself pinned by FP64 central differences in test_oracle_adjoint.py), with
relative error <= 1e-3 on the gradients; batched members against single fits;
the 200-step Adam fit of the N=512 workload; the quasi-Newton step against a
host restatement of its update rule."""
import math
import os
import sys

import numpy as np
import pytest

import oracle
import pyocean

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
GRAD_RTOL = 1e-3
HIDDEN = (22.0, math.radians(30.0), 3.5)
TIMES8 = [0.5 * k for k in range(1, 9)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    h = pyocean.Hip()
    assert h.device_count() >= 1, "no HIP device visible"
    return h


def _kw(N, L, V, th, A, conj=0, **extra):
    return dict(N=N, L=L, wind_speed=V, wind_dir=(math.cos(th), math.sin(th)), amplitude=A, conj_mode=conj,
                **extra)


def _targets(N, L, times, conj=0, sigma=0.05, key=1234, **extra):
    """Heights at the hidden variables (FP64 oracle) plus N(0, sigma^2) noise."""
    op = oracle.make_params(**_kw(N, L, *HIDDEN, conj=conj, **extra))
    h0 = oracle.h0(op)
    h = np.stack([oracle.frame(op, h0, t, nthreads=NTHREADS, want=("height",))["height"] for t in times])
    rng = np.random.Generator(np.random.Philox(key=key))
    return (h + sigma * rng.standard_normal(h.shape)).astype(np.float32)


def _oracle_grads(kw, times, targets):
    op = oracle.make_params(**kw)
    loss, gh = oracle.fit_eval(op, oracle.h0(op), times, targets.astype(np.float64), nthreads=NTHREADS)
    return loss, oracle.fit_param_grad(op, gh), gh


START = (18.0, math.radians(15.0), 2.5)
# The 45-degree starts put the wind exactly perpendicular to the lattice
# diagonals, where k_hat . w is zero or a rounding residue: the angle gradient
# must stay finite and exact there.
CASES = [
    ("c4_n512_k8", 512, 1000.0, TIMES8, 0, START),
    ("n512_k8_textbook", 512, 1000.0, TIMES8, 1, START),
    ("n64_k3", 64, 100.0, [0.25, 1.0, 3.0], 0, START),
    ("n1024_k2", 1024, 2000.0, [1.0, 2.0], 0, START),
    ("n128_k1_textbook", 128, 250.0, [1.5], 1, START),
    ("n256_k4_wind45", 256, 500.0, [0.5, 1.0, 1.5, 2.0], 0, (12.0, math.pi / 4, 3.0)),
    ("n512_k8_wind135_textbook", 512, 1000.0, TIMES8, 1, (25.0, 3 * math.pi / 4, 4.0)),
    ("n64_k16_max_frames", 64, 100.0, [0.25 * k for k in range(1, 17)], 0, START),
    ("n128_k15_odd", 128, 250.0, [0.3 * k for k in range(1, 16)], 1, START),
]


@pytest.mark.parametrize("name,N,L,times,conj,start", CASES, ids=[c[0] for c in CASES])
def test_fit_gradients_match_oracle_adjoint(hip, name, N, L, times, conj, start):
    kw = _kw(N, L, *start, conj=conj)
    tg = _targets(N, L, times, conj=conj)
    _say("fit %s: N=%d K=%d" % (name, N, len(times)))
    with pyocean.Fit([pyocean.make_params(**kw)], times) as f:
        f.upload_targets(0, tg)
        f.forward()
        _, _, loss_fwd = f.get(0)
        f.backward()
        x, g, loss = f.get(0)
        gh = f.grad_h0(0)
    np.testing.assert_allclose(x, start, rtol=0, atol=0)   # backward does not move the variables
    loss_ref, g_ref, gh_ref = _oracle_grads(kw, times, tg)
    _say("  loss %.6e ref %.6e; grad %s ref %s" % (loss, loss_ref, g, g_ref))
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    assert abs(loss_fwd - loss) <= 1e-6 * loss
    np.testing.assert_allclose(g, g_ref, rtol=GRAD_RTOL, atol=1e-6 * np.abs(g_ref).max())
    # dL/dh0 over every texel: relative L2 and max-abs against the largest entry
    err = gh.astype(np.float64) - gh_ref
    assert np.sqrt((err ** 2).sum() / (gh_ref ** 2).sum()) <= GRAD_RTOL
    assert np.abs(err).max() <= GRAD_RTOL * np.abs(gh_ref).max()


# A fit's step splits each column strip over 1, 2 or 4 blocks, chosen by the
# fit's total size, so a member's loss and parameter gradients (sums over its
# texels) are summed in another order in a batch than alone, and the kernel
# instances of the three splits round dL/dh0 differently in the last bits:
# a member agrees with the same member fitted alone to float rounding.
SUM_RTOL = 1e-6


def _assert_same_member(together, g, loss, gh, msg):
    assert abs(together[2] - loss) <= SUM_RTOL * abs(loss), msg
    np.testing.assert_allclose(together[1], g, rtol=10 * SUM_RTOL, atol=SUM_RTOL * np.abs(g).max(), err_msg=msg)
    np.testing.assert_allclose(together[3], gh, rtol=10 * SUM_RTOL, atol=SUM_RTOL * np.abs(gh).max(), err_msg=msg)


MEMBER_SETS = {
    # 45 column strips: four blocks per strip in the fit and alone
    "mixed_n": [(64, 100.0, (12.0, 0.3, 2.0), 0), (256, 500.0, (25.0, -1.0, 4.0), 1), (128, 250.0, (8.0, 2.5, 3.0), 0)],
    # 258 strips: two blocks per strip in the fit, four alone
    "two_n512": [(512, 1000.0, (18.0, 0.4, 2.5), 0), (512, 1000.0, (24.0, 2.0, 3.5), 1)],
}


@pytest.mark.parametrize("set_name", sorted(MEMBER_SETS))
def test_fit_members_equal_single_fits(hip, set_name):
    """Members in one fit give the single-member results to float rounding."""
    times = [0.5, 1.0, 1.5]
    specs = MEMBER_SETS[set_name]
    _say("fit members %s" % set_name)
    members, tgs = [], []
    for i, (N, L, v, conj) in enumerate(specs):
        members.append(pyocean.make_params(**_kw(N, L, *v, conj=conj, ocean_id=i, seed=42 + i)))
        tgs.append(_targets(N, L, times, conj=conj, key=100 + i, ocean_id=i, seed=42 + i))
    with pyocean.Fit(members, times) as f:
        for i in range(len(members)):
            f.upload_targets(i, tgs[i])
        f.backward()
        together = [f.get(i) + (f.grad_h0(i),) for i in range(len(members))]
    for i, m in enumerate(members):
        with pyocean.Fit([m], times) as f1:
            f1.upload_targets(0, tgs[i])
            f1.backward()
            x, g, loss = f1.get(0)
            gh = f1.grad_h0(0)
        _assert_same_member(together[i], g, loss, gh, "member %d" % i)


def test_fit_sees_changed_targets(hip):
    """The loss follows the targets through every way of changing them: a
    second upload_targets, and a write through targets_ptr after asking for
    the pointer again (the targets' 2-D DFT is remade on the next fit call)."""
    N, L, times = 128, 250.0, [0.5, 1.0, 1.5]
    kw = _kw(N, L, *START)
    tg1 = _targets(N, L, times, key=11)
    tg2 = _targets(N, L, times, key=22, sigma=0.2)
    ref1 = _oracle_grads(kw, times, tg1)[0]
    ref2 = _oracle_grads(kw, times, tg2)[0]
    losses = []
    with pyocean.Fit([pyocean.make_params(**kw)], times) as f:
        for step in ("upload 1", "upload 2", "pointer 1", "pointer 2"):
            _say("fit targets: %s" % step)
            tg = tg1 if step.endswith("1") else tg2
            if step.startswith("upload"):
                f.upload_targets(0, tg)
            else:
                hip.sync()
                hip.upload(f.targets_ptr(0), tg)
            f.forward()
            losses.append(f.get(0)[2])
    _say("  losses %s, oracle %.6e %.6e" % (losses, ref1, ref2))
    for loss, ref in zip(losses, (ref1, ref2, ref1, ref2)):
        assert abs(loss - ref) <= 1e-4 * ref


def _orc_from_lib(p):
    return oracle.make_params(N=p.N, L=p.L, wind_speed=p.wind_speed, wind_dir=(p.wind_dir_x, p.wind_dir_y),
                              amplitude=p.amplitude, suppression=p.suppression, choppiness=p.choppiness,
                              gravity=p.gravity, seed=p.seed, ocean_id=p.ocean_id, cascade_id=p.cascade_id,
                              conj_mode=p.conj_mode, normal_mode=p.normal_mode)


def test_c5_batched_fit_at_size(hip):
    """The batched-fit workload at its stated size: 256 oceans at N=256,
    L=500 m, K=4 (t = 0.5..2.0 s), every ocean with its own hidden (V, theta,
    A) and noise seed, all at the common start (17.5 m/s, 0 rad, 3.5).
    Targets are each ocean's heights at its hidden variables plus
    N(0, 0.05^2) noise (made on the GPU by a batch of the hidden oceans; both
    sides of every comparison get the same targets).  One forward+backward of
    all 256 members in one set of launches: members 0, 128 and 255 against
    the FP64 oracle adjoint (loss, dL/d(V, theta, A), dL/dh0 at rel <= 1e-3),
    and every member equal to a fit of that member alone to float
    rounding."""
    times = [0.5, 1.0, 1.5, 2.0]
    M = 256
    hidden, starts = pyocean.fit_batch_params(M)
    N = starts[0].N
    assert (N, starts[0].L, len(times)) == (256, 500.0, 4)
    rng = np.random.Generator(np.random.Philox(key=777))
    tgs = np.empty((M, len(times), N, N), dtype=np.float32)
    _say("c5: targets of %d oceans" % M)
    with pyocean.Batch(hidden) as b:
        b.tilde_h0_k()
        for k, t in enumerate(times):
            b.update(t)
            hip.sync()
            for o in range(M):
                tgs[o, k] = hip.download(b.fields(o).height, (N, N))
    tgs += (0.05 * rng.standard_normal(tgs.shape)).astype(np.float32)
    _say("c5: forward+backward of %d members" % M)
    with pyocean.Fit(starts, times) as f:
        for o in range(M):
            f.upload_targets(o, tgs[o])
        f.forward()
        loss_fwd = [f.get(o)[2] for o in range(M)]
        f.backward()
        together = [f.get(o) + (f.grad_h0(o),) for o in range(M)]
    for o in range(M):
        x, g, loss, gh = together[o]
        assert np.all(np.isfinite(g)) and np.isfinite(loss), o
        assert abs(loss_fwd[o] - loss) <= 1e-6 * loss, o
    for o in (0, 128, 255):
        op = _orc_from_lib(starts[o])
        loss_ref, gh_ref = oracle.fit_eval(op, oracle.h0(op), times, tgs[o].astype(np.float64), nthreads=NTHREADS)
        g_ref = oracle.fit_param_grad(op, gh_ref)
        _, g, loss, gh = together[o]
        _say("  member %d: loss %.6e ref %.6e; grad %s ref %s" % (o, loss, loss_ref, g, g_ref))
        assert abs(loss - loss_ref) <= 1e-4 * loss_ref
        np.testing.assert_allclose(g, g_ref, rtol=GRAD_RTOL, atol=1e-6 * np.abs(g_ref).max())
        err = gh.astype(np.float64) - gh_ref
        assert np.sqrt((err ** 2).sum() / (gh_ref ** 2).sum()) <= GRAD_RTOL
        assert np.abs(err).max() <= GRAD_RTOL * np.abs(gh_ref).max()
    _say("c5: %d single-member fits" % M)
    for o in range(M):
        with pyocean.Fit([starts[o]], times) as f1:
            f1.upload_targets(0, tgs[o])
            f1.backward()
            x, g, loss = f1.get(0)
            gh = f1.grad_h0(0)
        _assert_same_member(together[o], g, loss, gh, "member %d" % o)


def test_fit_adam_recovers_hidden_variables(hip):
    """The N=512 workload: targets from (22 m/s, 30 deg, 3.5) plus N(0, 0.05^2)
    noise at t = 0.5..4.0 s; 200 Adam steps from (18 m/s, 15 deg, 2.5)."""
    N, L = 512, 1000.0
    tg = _targets(N, L, TIMES8)
    start = (18.0, math.radians(15.0), 2.5)
    with pyocean.Fit([pyocean.make_params(**_kw(N, L, *start))], TIMES8) as f:
        f.upload_targets(0, tg)
        cfg = pyocean.adam()
        for _ in range(200):
            f.step(cfg)
        x, g, loss = f.get(0)
    _say("adam: V %.4f theta %.4f deg A %.4f loss %.6e" % (x[0], math.degrees(x[1]), x[2], loss))
    # the loss floor is the target noise variance 0.0025
    assert loss < 0.0026
    assert abs(x[0] - HIDDEN[0]) < 0.3
    assert abs(math.degrees(x[1] - HIDDEN[1])) < 0.3
    assert abs(x[2] - HIDDEN[2]) < 0.05


def test_batched_fit_stays_finite_and_descends(hip):
    """16 oceans at N=64 with hidden variables drawn as in the batched workload
    (the angle anywhere in [0, 2 pi), so winds perpendicular to lattice vectors
    occur), 150 Adam steps from one common start: every variable and gradient
    stays finite and every loss falls."""
    times = [0.5, 1.0, 1.5, 2.0]
    hidden, starts = pyocean.fit_batch_params(16, N=64, L=125.0)
    with pyocean.Fit(starts, times) as f:
        for o, hp in enumerate(hidden):
            op = oracle.make_params(N=64, L=125.0, wind_speed=hp.wind_speed, wind_dir=(hp.wind_dir_x, hp.wind_dir_y),
                                    amplitude=hp.amplitude, seed=hp.seed, ocean_id=hp.ocean_id)
            h0 = oracle.h0(op)
            tg = np.stack([oracle.frame(op, h0, t, want=("height",))["height"] for t in times]).astype(np.float32)
            f.upload_targets(o, tg)
        f.forward()
        loss0 = [f.get(o)[2] for o in range(16)]
        cfg = pyocean.adam()
        for _ in range(150):
            f.step(cfg)
        res = [f.get(o) for o in range(16)]
    for o, (x, g, loss) in enumerate(res):
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(g)) and np.isfinite(loss), (o, x, g, loss)
        assert loss < loss0[o], (o, loss, loss0[o])


def test_fit_rejects_bad_arguments(hip):
    p = pyocean.make_params(N=2048)
    with pytest.raises(pyocean.OceanError):
        pyocean.Fit([p], [1.0])
    with pytest.raises(pyocean.OceanError):
        pyocean.Fit([pyocean.make_params(N=64)], [float(t) for t in range(17)])
    with pyocean.Fit([pyocean.make_params(N=64)], [1.0]) as f:
        with pytest.raises(pyocean.OceanError):
            f.set_vars(0, -1.0, 0.0, 1.0)
        with pytest.raises(pyocean.OceanError):
            f.step(pyocean.adam(beta1=1.0))
        with pytest.raises(pyocean.OceanError):
            f.step_qn(pyocean.qn(radius=0.0))


def _qn_replay(x_eval, grads, losses, cfg):
    """Host restatement of the quasi-Newton step (ocn_fit_step_qn): BFGS on the
    inverse Hessian with a trust region (Nocedal & Wright, Numerical
    Optimization, eq. 6.17 and 6.20) in scaled coordinates x / scale, fed with
    the losses and gradients the device evaluated at each trial point x_eval[k];
    returns the next trial point after every evaluation."""
    D = np.array(cfg.scale[:], dtype=np.float64)
    r0, rmax = cfg.radius, cfg.radius_max
    out, state = [], None
    for x, g, L in zip(x_eval, grads, losses):
        z, gz = np.asarray(x) / D, np.asarray(g) * D
        if state is None:
            state = dict(z=z, g=gz, L=L, H=np.eye(3), r=r0, k=0)
        elif not L <= state["L"]:
            state["r"] = 0.25 * np.linalg.norm(z - state["z"])
        else:
            s, y = z - state["z"], gz - state["g"]
            sy = float(s @ y)
            if sy > 1e-12 * np.linalg.norm(s) * np.linalg.norm(y):
                H = state["H"] * (sy / float(y @ y)) if state["k"] == 0 else state["H"]
                rho = 1.0 / sy
                I = np.eye(3)
                state["H"] = (I - rho * np.outer(s, y)) @ H @ (I - rho * np.outer(y, s)) + rho * np.outer(s, s)
                state["k"] += 1
            if np.linalg.norm(s) >= 0.99 * state["r"]:
                state["r"] = min(2.0 * state["r"], rmax)
            state.update(z=z, g=gz, L=L)
        p = -state["H"] @ state["g"]
        n = np.linalg.norm(p)
        if n > state["r"]:
            p *= state["r"] / n
        nxt = (state["z"] + p) * D
        nxt[0], nxt[2] = max(nxt[0], 1e-3), max(nxt[2], 1e-6)
        out.append(nxt)
    return out


def test_fit_quasi_newton_step_follows_its_update_rule(hip):
    """Two N=128 oceans, 40 quasi-Newton steps from offset starts: after every
    step the device's next trial point equals the host restatement of the
    BFGS / trust-region rule fed with the losses and gradients the device
    reported (relative 1e-9), the run takes accepted and rejected trials and
    BFGS updates, and the loss ends below its start."""
    N, L = 128, 250.0
    times = [0.5, 1.0, 1.5, 2.0]
    tg = _targets(N, L, times)
    starts = [(18.0, math.radians(24.0), 2.8), (25.0, math.radians(40.0), 4.5)]
    cfg = pyocean.qn()
    with pyocean.Fit([pyocean.make_params(**_kw(N, L, *s)) for s in starts], times) as f:
        for o in range(2):
            f.upload_targets(o, tg)
        trials = [[np.array(s, dtype=np.float64)] for s in starts]
        grads, losses = [[], []], [[], []]
        for _ in range(40):
            f.step_qn(cfg)
            for o in range(2):
                x, g, loss = f.get(o)
                trials[o].append(x)
                grads[o].append(g)
                losses[o].append(loss)
    for o in range(2):
        want = _qn_replay(trials[o][:-1], grads[o], losses[o], cfg)
        for k, (w, got) in enumerate(zip(want, trials[o][1:])):
            np.testing.assert_allclose(got, w, rtol=1e-9, atol=1e-12, err_msg="member %d step %d" % (o, k))
        L = np.array(losses[o])
        best = np.minimum.accumulate(L)
        rejected = int(np.sum(L[1:] > best[:-1]))
        _say("qn member %d: loss %.6e -> best %.6e, %d rejected trials" % (o, L[0], best[-1], rejected))
        assert best[-1] < 0.5 * L[0]
        assert rejected >= 1 and rejected < 30


def test_fit_adam_then_quasi_newton_reaches_the_noise_floor(hip):
    """The N=512 workload of test_fit_adam_recovers_hidden_variables: 60 Adam
    steps, then 140 quasi-Newton steps, recover the hidden variables at least
    as closely as 200 Adam steps (wind speed within 0.1 m/s, angle within
    0.1 deg, amplitude within 1%), at the noise floor."""
    N, L = 512, 1000.0
    tg = _targets(N, L, TIMES8)
    start = (18.0, math.radians(15.0), 2.5)
    with pyocean.Fit([pyocean.make_params(**_kw(N, L, *start))], TIMES8) as f:
        f.upload_targets(0, tg)
        a, q = pyocean.adam(), pyocean.qn()
        for _ in range(60):
            f.step(a)
        for _ in range(140):
            f.step_qn(q)
        f.forward()
        x, g, loss = f.get(0)
    _say("adam + qn: V %.4f theta %.4f deg A %.4f loss %.6e" % (x[0], math.degrees(x[1]), x[2], loss))
    assert loss < 0.0026
    assert abs(x[0] - HIDDEN[0]) < 0.1
    assert abs(math.degrees(x[1] - HIDDEN[1])) < 0.1
    assert abs(x[2] / HIDDEN[2] - 1.0) < 0.01
