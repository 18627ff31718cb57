"""Benchmark of the Tessendorf ocean path on one MI355X.

Workload: the plain N=1024 ocean (L=2000 m, wind 25 m/s at 30 deg, A=4,
lambda=1.3, l=L/2000, Philox seed 42), frames at t_f = f/60 s, each frame's
height, Dx, Dz, normals and Jacobian written into its own output slice (600
slices, 19.3 GiB, reused cyclically for more than 600 steps).  A step is one
frame: spectrum evolution, three inverse 2-D FFTs with the inversion step,
normals and Jacobian, all on the GPU with inputs resident in HBM.

Prints one JSON line (metric, value in frames/s, roofline of the dominant
kernel measured with HIP events, CPU baseline of the FP64 oracle).

Beside it, the line carries the batched-cascade comparison (64 oceans x 3
cascades in two launches per frame against per-ocean launches) and the
differentiable fit (forward and forward+backward cost, a 200-step Adam fit).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S] [--cascade-frames F]
  python bench.py --mode copy      (device-to-device copy used to calibrate
                                    rocprofv3 FETCH_SIZE / WRITE_SIZE)
  python bench.py --mode cascades  (the batched-cascade comparison alone)
  python bench.py --mode fit       (the differentiable fit alone)
  python bench.py --mode batched-fit (256 oceans fitted together)
  python bench.py --mode single    (one ocn_update per frame, eager and graph-replayed)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fft-ocean-waves_amd"))

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
FRAMES = 600
N = 1024

# Canonical (algorithmic) bytes per texel.  Whole frame: h0 pair in (16) +
# one intermediate round trip of three complex FP32 fields (24 + 24) + height,
# Dx, Dz (12) + normals (16) + Jacobian (4) = 96.  Per kernel of the fused
# frame: columns = 16 + 24; rows = 24 + 12 with a separate normal-map kernel
# (16 + 4, plus re-reading height/Dx/Dz, which the canonical count does not
# credit), or 24 + 12 + 16 + 4 = 56 when the row kernel also writes the
# normals and the Jacobian.
CANON_FRAME = 96
CANON_STAGE = {"k_cols": 40, "k_rows": 36, "k_normals": 20, "k_rows_nrm": 56}
# Frames per kernel launch of the sequence (ocn_reserve_frames), by N.  The
# row pass writes its outputs with non-temporal stores, so the intermediates a
# launch's column pass wrote stay in the 256 MB Infinity Cache for its row
# pass as long as they fit there beside h0: 16 N=1024 frames (192 MB) ran at
# 12.4 us per frame over the 600-frame sequence, against 13.8 in 8-frame and
# 15.0 in 32-frame launches (384 MB, read back from HBM).  N=2048 runs
# fastest in 4-frame launches, N=4096 in 8-frame ones, smaller N in 32-frame
# ones.
def launch_frames_for(n):
    return 16 if n == 1024 else 4 if n == 2048 else 8 if n == 4096 else 32


LAUNCH_FRAMES = launch_frames_for(1024)

# HBM bytes per texel-frame measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (each scaled by what the counter reads for a device-to-device copy of known
# size: fetch x 2.0, write x 1.0) for this workload, keyed by kernel and the
# (column, row) schedule it ran under; see
# profiles/k_rows_nrm_k_cols_lds_n1024_batched_pmc_traffic.md.  Reported only
# when the schedule in use matches, otherwise null.
MEASURED_TRAFFIC = {
    ("k_rows_nrm", 9, 4, 16): 13.5 + 32.1,
    ("k_cols_lds", 9, 4, 16): 1.9 + 12.6,
    ("k_rows_nrm", 9, 3, 8): 15.8 + 32.0,
    ("k_cols_lds", 9, 3, 8): 2.9 + 12.9,
    ("k_rows_nrm", 9, 5, 32): 12.9 + 32.0,
    ("k_cols_lds", 9, 5, 32): 1.4 + 12.6,
}


def kernel_names(sched):
    """Kernel names of a (column, row) schedule as ocn_get_schedule reports it."""
    cv, rv = sched["cols_variant"], sched["rows_variant"]
    cols = ("k_cols" if cv < 7 else "k_cols_lds2" if 10 <= cv < 13 else "k_cols_lds_f2" if cv >= 19
            else "k_cols_lds2b" if cv >= 17 else "k_cols_ldsb" if cv >= 15 else "k_cols_lds")
    rows = "k_rows" if sched["rows_band"] == 0 else "k_rows_nrm_par" if rv >= 6 else "k_rows_nrm"
    return cols, rows


def workload_params():
    import pyocean
    th = math.radians(30.0)
    return pyocean.make_params(N=N, L=2000.0, wind_speed=25.0, wind_dir=(math.cos(th), math.sin(th)),
                               amplitude=4.0, suppression=1.0, choppiness=1.3, seed=42)


def host_cpu():
    """CPU model (first 'model name' of /proc/cpuinfo) and logical CPUs."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_share():
    """Threads for the CPU baseline: the CPU share this process may use.  On a
    shared GPU host that is what OMP_NUM_THREADS states (the host's share for
    one GPU), else the CPUs of this process's affinity mask."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        omp = 0
    return max(1, min(aff, omp) if omp > 0 else aff)


def cpu_baseline(seconds):
    """The FP64 oracle (a scalar C port of the reference shaders) timed on
    this host: whole N=1024 frames (spectrum, three radix-2 butterfly IFFTs
    with inversion, normals, Jacobian), once on one thread and once with every
    pass split by rows over the host's CPU share, each until its share of
    `seconds` has passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    th = math.radians(30.0)
    op = oracle.make_params(N=N, L=2000.0, wind_speed=25.0, wind_dir=(math.cos(th), math.sin(th)),
                            amplitude=4.0, suppression=1.0, choppiness=1.3, seed=42)
    h0 = oracle.h0(op)
    model, logical = host_cpu()
    cores = cpu_share()

    def run(nthreads, budget):
        oracle.frame(op, h0, 0.0, nthreads=nthreads)  # warm-up
        n, t0 = 0, time.perf_counter()
        while True:
            oracle.frame(op, h0, n / 60.0, nthreads=nthreads)
            n += 1
            el = time.perf_counter() - t0
            if el >= budget:
                return n, el

    n1, el1 = run(1, seconds / 3.0)
    nt, elt = run(cores, seconds * 2.0 / 3.0)
    desc = ("whole N=1024 frames (spectrum, 3 radix-2 butterfly IFFTs + inversion, normals, Jacobian) of the "
            "FP64 oracle")
    return {"value": round(nt / elt, 4), "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": "%d %s in %.1f s, every pass split by rows over %d threads" % (nt, desc, elt, cores),
            "cpu_model": model, "logical_cpus": logical,
            "single_thread": {"value": round(n1 / el1, 4), "unit": "frames/s", "cores": 1,
                              "sample": "%d frames in %.1f s, 1 thread" % (n1, el1)}}


def cascades_bench(hip, frames, warmup):
    """The batched-cascade workload (64 oceans x 3 cascades, L/N = 2000/256,
    250/512, 30/1024, wind U(5,30) m/s, direction U(0,2pi), lambda U(0.8,1.5))
    computed three ways on one stream, each timed with HIP events: the batch
    (two launches per frame for all 192 cascades) replayed as a HIP graph and
    launched eagerly, with the built-in and then the autotuned schedule, and
    per-ocean launches (one fused single-ocean update per cascade, the
    schedule it picks alone) from a host loop and replayed from a captured
    graph of all 192 updates.  The speedups use the fastest batch figure."""
    import pyocean
    members = pyocean.cascade_params()
    texels = sum(m.N * m.N for m in members)
    stream = hip.stream()
    e0, e1 = hip.event(), hip.event()

    def timed(fn, n):
        hip.record(e0, stream)
        for i in range(n):
            fn(i)
        hip.record(e1, stream)
        return hip.elapsed_ms(e0, e1) / n

    out = {"members": len(members), "texels_per_frame": texels}
    with pyocean.Batch(members, stream=stream) as b:
        b.tilde_h0_k()
        for i in range(warmup):
            b.update_graph(i / 60.0)
        out["batch_graph_ms"] = timed(lambda i: b.update_graph(i / 60.0), frames)
        for i in range(warmup):
            b.update(i / 60.0)
        out["batch_eager_ms"] = timed(lambda i: b.update(i / 60.0), frames)
        # the batch's own autotuner over the column strip width and the row
        # pass's band height per N
        b.autotune(reps=5)
        firsts = {}
        for i, m in enumerate(members):
            firsts.setdefault(m.N, i)
        out["autotuned"] = {str(N): {"cols_per_block": b.column_strips(i), "rows_band": b.schedule(i)}
                            for N, i in sorted(firsts.items())}
        for i in range(warmup):
            b.update_graph(i / 60.0)
        out["batch_graph_autotuned_ms"] = timed(lambda i: b.update_graph(i / 60.0), frames)
    hip.stream_sync(stream)
    ocs = [pyocean.Ocean(m, stream=stream) for m in members]
    try:
        for oc in ocs:
            oc.tilde_h0_k()
            oc.update(0.0)
        hip.stream_sync(stream)
        po = max(3, frames // 10)

        def per_ocean(i):
            for oc in ocs:
                oc.update(i / 60.0)

        per_ocean(0)
        out["per_ocean_loop_ms"] = timed(per_ocean, po)
        g, ex = hip.capture(stream, lambda: per_ocean(1))
        try:
            for _ in range(3):
                hip.graph_launch(ex, stream)
            out["per_ocean_graph_ms"] = timed(lambda i: hip.graph_launch(ex, stream), po)
        finally:
            hip.stream_sync(stream)
            hip.graph_destroy(g, ex)
    finally:
        for oc in ocs:
            oc.close()
    hip.stream_sync(stream)
    hip.stream_destroy(stream)
    for k in ("batch_graph_ms", "batch_eager_ms", "batch_graph_autotuned_ms", "per_ocean_loop_ms",
              "per_ocean_graph_ms"):
        out[k] = round(out[k], 4)
    best = min(out["batch_graph_ms"], out["batch_eager_ms"], out["batch_graph_autotuned_ms"])
    out["batch_frames_per_s"] = round(1e3 / best, 2)
    out["batch_canonical_GBps"] = round(CANON_FRAME * texels / (best * 1e-3) / 1e9, 1)
    out["speedup_vs_per_ocean_loop"] = round(out["per_ocean_loop_ms"] / best, 3)
    out["speedup_vs_per_ocean_graph"] = round(out["per_ocean_graph_ms"] / best, 3)
    return out


def fit_bench(hip, steps=200):
    """The differentiable fit.  N=512, L=1000 m, l=0.5: targets are the heights
    at (22 m/s, 30 deg, A=3.5) at t = 0.5..4.0 s (K=8) plus N(0, 0.05^2) noise;
    the forward (loss) and forward+backward (loss, gradients, spectrum
    recomputed) operators are timed with HIP events over graph replays of 20
    calls each (host launch cost excluded), then `steps` Adam steps run from
    (18 m/s, 15 deg, 2.5)."""
    import pyocean
    N, L = 512, 1000.0
    times = [0.5 * k for k in range(1, 9)]
    hidden = (22.0, math.radians(30.0), 3.5)
    start = (18.0, math.radians(15.0), 2.5)

    def params(v, **kw):
        return pyocean.make_params(N=N, L=L, wind_speed=v[0], wind_dir=(math.cos(v[1]), math.sin(v[1])),
                                   amplitude=v[2], **kw)

    stream = hip.stream()
    tg = np.empty((len(times), N, N), dtype=np.float32)
    with pyocean.Ocean(params(hidden), stream=stream) as oc:
        oc.tilde_h0_k()
        for k, t in enumerate(times):
            oc.update(t)
            tg[k] = oc.download_fields()["height"]
    rng = np.random.Generator(np.random.Philox(key=1234))
    tg += (0.05 * rng.standard_normal(tg.shape)).astype(np.float32)
    e0, e1 = hip.event(), hip.event()
    out = {"workload": "N=512, L=1000 m, K=8 frames t=0.5..4.0 s, hidden (22 m/s, 30 deg, A=3.5), "
                       "target noise N(0, 0.05^2), start (18 m/s, 15 deg, 2.5)"}
    with pyocean.Fit([params(start)], times, stream=stream) as f:
        f.upload_targets(0, tg)
        for name, op in (("forward", f.forward), ("forward_backward", f.backward)):
            g, ex = hip.capture(stream, lambda: [op() for _ in range(20)])
            try:
                hip.graph_launch(ex, stream)
                hip.record(e0, stream)
                for _ in range(5):
                    hip.graph_launch(ex, stream)
                hip.record(e1, stream)
                out[name + "_us"] = round(hip.elapsed_ms(e0, e1) * 1e3 / 100, 2)
            finally:
                hip.stream_sync(stream)
                hip.graph_destroy(g, ex)
        out["ratio_fwd_bwd_to_fwd"] = round(out["forward_backward_us"] / out["forward_us"], 3)
        cfg = pyocean.adam()
        hip.record(e0, stream)
        for _ in range(steps):
            f.step(cfg)
        hip.record(e1, stream)
        out["adam_steps"] = steps
        out["adam_steps_per_s"] = round(steps / (hip.elapsed_ms(e0, e1) * 1e-3), 1)
        x, _, loss = f.get(0)
    out["final"] = {"wind_speed": round(x[0], 4), "wind_angle_deg": round(math.degrees(x[1]), 4),
                    "amplitude": round(x[2], 4), "loss": loss}
    out["abs_error"] = {"wind_speed": round(abs(x[0] - hidden[0]), 4),
                        "wind_angle_deg": round(abs(math.degrees(x[1] - hidden[1])), 4),
                        "amplitude": round(abs(x[2] - hidden[2]), 4)}
    hip.stream_destroy(stream)
    return out


def batched_fit_bench(hip, steps=200, n_oceans=256, adam_steps=60):
    """The batched fit: 256 independent oceans at N=256, L=500 m, each with
    hidden (V ~ U(5,30) m/s, angle ~ U(0,2pi), A ~ U(2,5)), targets = its heights
    at t = 0.5, 1.0, 1.5, 2.0 s plus N(0, 0.05^2) noise, all starting from
    (17.5 m/s, 0 rad, 3.5); `steps` optimizer steps of all oceans, each step one
    column, one row, one adjoint-column and one reduction launch for all
    members: `adam_steps` Adam steps, then quasi-Newton (BFGS, trust region)
    steps.  Reports steps/s and the final parameter errors (the angle modulo
    pi: the Phillips spectrum depends on (k.w)^2, so w and -w are the same
    ocean), and the same for `steps` Adam steps alone.

    Convergence is measured against each ocean's noise floor, the loss of its
    noisy targets at the hidden parameters: a fit whose loss reaches it
    explains the data as well as the truth does, whatever its parameters.
    Adam alone crosses the angle landscape well but creeps along the valley
    where wind speed and amplitude trade off (strong winds), which BFGS
    follows in a few steps once it has the curvature; BFGS alone stops in the
    nearest angular minimum.  Low-wind oceans have heights of a few cm under
    5 cm of noise, where (V, A) hardly change the loss and their errors say
    little; errors are also reported over the oceans whose clean heights have
    an RMS of at least 2 sigma.  Remedy for the oceans caught in a local
    minimum: members whose loss exceeds 1.2 x the batch median (all oceans
    share the noise level) are fitted again, with the same steps, from the
    start with their angle rotated by 90 degrees, and the lower loss is kept."""
    import pyocean
    times = [0.5, 1.0, 1.5, 2.0]
    hidden, starts = pyocean.fit_batch_params(n_oceans)
    N = starts[0].N
    stream = hip.stream()
    rng = np.random.Generator(np.random.Philox(key=777))
    tgs = np.empty((n_oceans, len(times), N, N), dtype=np.float32)
    with pyocean.Batch(hidden, stream=stream) as b:
        b.tilde_h0_k()
        for k, t in enumerate(times):
            b.update(t)
            hip.stream_sync(stream)
            for o in range(n_oceans):
                tgs[o, k] = hip.download(b.fields(o).height, (N, N))
    signal_rms = np.sqrt((tgs.astype(np.float64) ** 2).mean(axis=(1, 2, 3)))
    tgs += (0.05 * rng.standard_normal(tgs.shape)).astype(np.float32)
    adam, qn = pyocean.adam(), pyocean.qn()
    e0, e1 = hip.event(), hip.event()

    def run(members, owners, n_adam):
        """`n_adam` Adam steps, then quasi-Newton steps up to `steps`, of a fit
        of `members` against their owners' targets; a last forward pass gives
        the loss at the final variables."""
        with pyocean.Fit(members, times, stream=stream) as f:
            for i, o in enumerate(owners):
                f.upload_targets(i, tgs[o])
            hip.record(e0, stream)
            for k in range(steps):
                if k < n_adam:
                    f.step(adam)
                else:
                    f.step_qn(qn)
            hip.record(e1, stream)
            f.forward()
            ms = hip.elapsed_ms(e0, e1)
            return [f.get(i) for i in range(len(members))], ms

    def with_angle(p, theta):
        q = pyocean.make_params(N=p.N, L=p.L, wind_speed=p.wind_speed, wind_dir=(math.cos(theta), math.sin(theta)),
                                amplitude=p.amplitude, seed=p.seed, ocean_id=p.ocean_id)
        return q

    # the noise floor: the loss of each ocean's targets at its hidden parameters
    with pyocean.Fit(hidden, times, stream=stream) as f:
        for o in range(n_oceans):
            f.upload_targets(o, tgs[o])
        f.forward()
        floor = np.array([f.get(o)[2] for o in range(n_oceans)])

    def errors(rr, idx):
        dv = np.array([abs(rr[o][0][0] - hidden[o].wind_speed) for o in idx])
        dth = np.array([abs((rr[o][0][1] - math.atan2(hidden[o].wind_dir_y, hidden[o].wind_dir_x) + math.pi / 2)
                            % math.pi - math.pi / 2) for o in idx])
        da = np.array([abs(rr[o][0][2] / hidden[o].amplitude - 1.0) for o in idx])
        st = lambda a, f=1.0: {"median": round(f * float(np.median(a)), 4), "p90": round(f * float(np.percentile(a, 90)), 4),
                               "max": round(f * float(a.max()), 4)}
        return {"wind_speed_abs": st(dv), "angle_deg_mod_180": st(dth, 180.0 / math.pi), "amplitude_rel": st(da)}

    allo = list(range(n_oceans))
    ident = [o for o in allo if signal_rms[o] >= 0.1]
    res_adam, _ = run(starts, allo, steps)
    res, ms = run(starts, allo, adam_steps)
    loss = np.array([r[2] for r in res])
    stuck = [o for o in allo if loss[o] > 1.2 * np.median(loss)]
    final = list(res)
    ms_restart = 0.0
    if stuck:
        rs, ms_restart = run([with_angle(starts[o], res[o][0][1] + math.pi / 2) for o in stuck], stuck, adam_steps)
        for o, r in zip(stuck, rs):
            if r[2] < final[o][2]:
                final[o] = r
    hip.stream_destroy(stream)
    loss_f = np.array([r[2] for r in final])
    return {
        "workload": "%d oceans, N=%d, L=500 m, K=4 (t=0.5..2.0 s), hidden V~U(5,30), angle~U(0,2pi), A~U(2,5), "
                    "target noise N(0, 0.05^2), start (17.5 m/s, 0 rad, 3.5)" % (n_oceans, N),
        "steps": steps,
        "optimizer": "%d Adam steps (lr %s), then %d quasi-Newton steps (scales %s, radius %g..%g)" % (
            adam_steps, list(adam.lr), steps - adam_steps, list(qn.scale), qn.radius, qn.radius_max),
        "ms_per_step": round(ms / steps, 4),
        "fit_steps_per_s": round(steps / (ms * 1e-3), 1),
        "ocean_steps_per_s": round(n_oceans * steps / (ms * 1e-3), 1),
        "at_noise_floor_after_steps": int((loss <= 1.01 * floor).sum()),
        "errors_after_steps": errors(res, allo),
        "restart": {"rule": "loss > 1.2 x batch median: the same %d steps from the angle + 90 deg" % steps,
                    "members": len(stuck), "ms_per_step": round(ms_restart / steps, 4) if stuck else 0.0},
        "at_noise_floor": int((loss_f <= 1.01 * floor).sum()),
        "final_errors": errors(final, allo),
        "identifiable_oceans": len(ident),
        "final_errors_identifiable": errors(final, ident),
        "final_loss_over_floor": {"median": round(float(np.median(loss_f / floor)), 4),
                                  "max": round(float((loss_f / floor).max()), 4)},
        "adam_only": {"at_noise_floor": int((np.array([r[2] for r in res_adam]) <= 1.01 * floor).sum()),
                      "errors_identifiable": errors(res_adam, ident)},
    }


def single_frame_bench(hip, frames=600, warmup=60, graph_frames=60):
    """The drop-in use: the reference runs exactly one frame per update()
    (tilde_h0_t, three butterfly_fft, generate_normal_map; main.cpp:240-244),
    always into the same textures.  Here one ocn_update per frame into the
    handle's own fields at t = f/60 on the N=1024 workload, timed with HIP
    events on the launch stream: eagerly (host launch cost included) and as
    replays of a graph that holds `graph_frames` consecutive single-frame
    updates (launch cost amortised)."""
    import pyocean
    nn = N * N
    stream = hip.stream()
    e0, e1 = hip.event(), hip.event()
    out = {"workload": "plain single ocean N=1024 as in the default mode, one ocn_update per frame into the "
                       "handle's own textures (no frame batching), t=f/60 s"}
    with pyocean.Ocean(workload_params(), stream=stream) as oc:
        oc.tilde_h0_k()
        for f in range(warmup):
            oc.update(f / 60.0)
        hip.stream_sync(stream)
        hip.record(e0, stream)
        for f in range(frames):
            oc.update(f / 60.0)
        hip.record(e1, stream)
        eager_ms = hip.elapsed_ms(e0, e1) / frames
        g, ex = hip.capture(stream, lambda: [oc.update(f / 60.0) for f in range(graph_frames)])
        try:
            for _ in range(3):
                hip.graph_launch(ex, stream)
            reps = max(1, frames // graph_frames)
            hip.record(e0, stream)
            for _ in range(reps):
                hip.graph_launch(ex, stream)
            hip.record(e1, stream)
            graph_ms = hip.elapsed_ms(e0, e1) / (reps * graph_frames)
        finally:
            hip.stream_sync(stream)
            hip.graph_destroy(g, ex)
        sched = oc.schedule(1)
    hip.stream_destroy(stream)
    best = min(eager_ms, graph_ms)
    gbps = CANON_FRAME * nn / (best * 1e-3) / 1e9
    out.update({"frames": frames, "eager_us_per_frame": round(eager_ms * 1e3, 3),
                "graph_us_per_frame": round(graph_ms * 1e3, 3), "graph_frames_per_replay": graph_frames,
                "frames_per_s": round(1e3 / best, 1), "canonical_GBps": round(gbps, 1),
                "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4), "schedule": sched})
    return out


def render_bench(hip, steps, warmup):
    """The N=1024 frame sequence with the render export on: besides the five
    fields, every frame writes the displacement (-l Dx, s h, -l Dz, J) and the
    displaced-surface normal (float4 each), what the reference's renderer reads
    from its displacement and normal textures (grid_tes.glsl:60-64).  Frames
    go to min(steps, 600) output slices as in the default mode."""
    import pyocean
    nn = N * N
    slots = min(FRAMES, max(1, steps))
    names = ("height", "dx", "dz", "normals", "jacobian", "displacement", "render_normals")
    sizes = dict(zip(names, (4, 4, 4, 16, 4, 16, 16)))
    bufs = {k: hip.malloc(slots * nn * sizes[k]) for k in names}
    stream = hip.stream()
    oc = pyocean.Ocean(workload_params(), stream=stream)
    oc.reserve_frames(min(LAUNCH_FRAMES, slots))  # setup: no allocation once frames run
    oc.tilde_h0_k()
    slot_fields = [pyocean.OcnFields(*[bufs[k] + s * nn * sizes[k] for k in names]) for s in range(slots)]

    def frames(f0, count):
        f = f0
        while f < f0 + count:
            s = f % slots
            n = min(slots - s, f0 + count - f)
            oc.update_frames(np.arange(f, f + n, dtype=np.float64) / 60.0, slot_fields[s])
            f += n

    frames(0, warmup)
    hip.stream_sync(stream)
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    frames(0, steps)
    hip.record(e1, stream)
    hip.stream_sync(stream)
    ms = hip.elapsed_ms(e0, e1)
    oc.close()
    for v in bufs.values():
        hip.free(v)
    hip.stream_destroy(stream)
    canon = CANON_FRAME + 32  # + displacement and displaced-surface normal, float4 each
    gbps = canon * nn / (ms * 1e-3 / steps) / 1e9
    return {"workload": "plain single ocean N=1024 as in the default mode, render export on, %d slices" % slots,
            "steps": steps, "frames_per_s": round(steps / (ms / 1e3), 2), "us_per_frame": round(ms * 1e3 / steps, 3),
            "canonical_bytes_per_texel": canon, "canonical_GBps": round(gbps, 1),
            "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4)}


def run_copy(hip):
    """Device-to-device copies of 256 MiB (rocprofv3 counter calibration)."""
    nbytes = 256 << 20
    a, b = hip.malloc(nbytes), hip.malloc(nbytes)
    hip.memset(a, 1, nbytes)
    for _ in range(5):
        hip.ok(hip.rt.hipMemcpy(b, a, nbytes, hip.D2D), "hipMemcpy D2D")
    hip.sync()
    hip.free(a)
    hip.free(b)
    print(json.dumps({"mode": "copy", "bytes_per_copy": nbytes, "copies": 5}), flush=True)


def main():
    global N, LAUNCH_FRAMES
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--mode", choices=["frames", "copy", "cascades", "fit", "batched-fit", "render", "single"],
                    default="frames")
    ap.add_argument("--cascade-frames", type=int, default=60,
                    help="frames of the batched-cascade comparison (0 skips it)")
    ap.add_argument("--fit", type=int, default=1, help="1: time the differentiable fit too (0 skips it)")
    ap.add_argument("--single", type=int, default=1, help="1: time single-frame updates too (0 skips them)")
    ap.add_argument("--launch-frames", type=int, default=0,
                    help="frames per kernel launch of the sequence (ocn_reserve_frames), 1..32; "
                         "0: by N (16 at N=1024)")
    ap.add_argument("--size", type=int, default=N,
                    help="grid size N of the frame sequence (profiling other sizes; the headline is N=1024)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the ocean path runs on a single MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.size not in (64, 128, 256, 512, 1024, 2048, 4096):
        print("bench.py: --size must be a power of two in [64, 4096]", file=sys.stderr)
        sys.exit(2)
    N = args.size
    if not 0 <= args.launch_frames <= 32:
        print("bench.py: --launch-frames must be in [1, 32] (0: by N)", file=sys.stderr)
        sys.exit(2)
    LAUNCH_FRAMES = args.launch_frames or launch_frames_for(N)
    import pyocean
    hip = pyocean.Hip()
    if hip.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.mode == "copy":
        run_copy(hip)
        return
    if args.mode == "fit":
        print(json.dumps({"mode": "fit", **fit_bench(hip)}), flush=True)
        return
    if args.mode == "batched-fit":
        print(json.dumps({"mode": "batched-fit", **batched_fit_bench(hip)}), flush=True)
        return
    if args.mode == "render":
        print(json.dumps({"mode": "render", **render_bench(hip, args.steps, args.warmup)}), flush=True)
        return
    if args.mode == "single":
        print(json.dumps({"mode": "single", **single_frame_bench(hip)}), flush=True)
        return
    if args.mode == "cascades":
        print(json.dumps({"mode": "cascades", **cascades_bench(hip, max(1, args.cascade_frames), 10)}), flush=True)
        return

    nn = N * N
    # output slices: one per frame up to the 600 of the sequence, within 32 GiB
    slots = min(FRAMES, max(1, args.steps), max(LAUNCH_FRAMES, (32 << 30) // (28 * nn)))
    sizes = {"height": 4, "dx": 4, "dz": 4, "normals": 16, "jacobian": 4}
    bufs = {k: hip.malloc(slots * nn * b) for k, b in sizes.items()}
    stream = hip.stream()
    oc = pyocean.Ocean(workload_params(), stream=stream)
    # Setup, before any timing: the intermediates for the launch size.  The
    # frame calls themselves never allocate, so warm-up and timed calls of any
    # length run the same kernels on the same buffers.
    oc.reserve_frames(min(LAUNCH_FRAMES, slots))
    oc.tilde_h0_k()
    hip.stream_sync(stream)

    def fields(slot):
        return pyocean.OcnFields(*[bufs[k] + slot * nn * sizes[k] for k in ("height", "dx", "dz", "normals",
                                                                            "jacobian")])

    slot_fields = [fields(s) for s in range(slots)]

    def frames(f0, count):
        # consecutive frames go to consecutive slots; one ABI call per run of slots
        f = f0
        while f < f0 + count:
            s = f % slots
            n = min(slots - s, f0 + count - f)
            oc.update_frames(np.arange(f, f + n, dtype=np.float64) / 60.0, slot_fields[s])
            f += n

    frames(0, args.warmup)
    hip.stream_sync(stream)
    hip.sync()

    e0, e1 = hip.event(), hip.event()
    t_wall = time.perf_counter()
    hip.record(e0, stream)
    frames(0, args.steps)
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1)
    hip.sync()
    t_wall = time.perf_counter() - t_wall
    fps = args.steps / (ms / 1e3)

    # Per-kernel durations: HIP events on the launch stream around `reps`
    # launches of each kernel alone, each launch covering LAUNCH_FRAMES frames
    # exactly as the timed region launches them.
    stage_ms = {}
    lf = min(LAUNCH_FRAMES, slots)
    tl = np.arange(lf, dtype=np.float64) / 60.0
    reps = 40
    sched = oc.schedule(lf)
    fused = sched["rows_band"] > 0  # normal map + Jacobian inside the row kernel
    cols_name, rows_name = kernel_names(sched)
    stages = [(cols_name, pyocean.OCN_STAGE_COLUMNS), (rows_name, pyocean.OCN_STAGE_ROWS)]
    if not fused:
        stages.append(("k_normals", pyocean.OCN_STAGE_NORMALS))
    for name, stage in stages:
        for _ in range(3):
            oc.update_stage(tl, slot_fields[0], stage)
        hip.record(e0, stream)
        for i in range(reps):
            oc.update_stage(tl + i, slot_fields[0], stage)
        hip.record(e1, stream)
        stage_ms[name] = hip.elapsed_ms(e0, e1) / reps
    dominant = max(stage_ms, key=stage_ms.get)
    dom_bytes = CANON_STAGE["k_cols" if dominant.startswith("k_cols") else
                            "k_rows_nrm" if dominant.startswith("k_rows_nrm") else dominant] * nn * lf
    achieved = dom_bytes / (stage_ms[dominant] * 1e-3) / 1e9
    frame_gbps = CANON_FRAME * nn / (ms * 1e-3 / args.steps) / 1e9
    tb = MEASURED_TRAFFIC.get((dominant, sched["cols_variant"], sched["rows_variant"], lf))
    traffic = int(round(tb * nn * lf)) if tb is not None else None

    oc.close()
    for v in bufs.values():
        hip.free(v)
    hip.stream_destroy(stream)

    single = single_frame_bench(hip) if args.single else None
    casc = cascades_bench(hip, args.cascade_frames, 10) if args.cascade_frames > 0 else None
    fit = fit_bench(hip) if args.fit else None
    bfit = batched_fit_bench(hip) if args.fit else None
    cpu = cpu_baseline(args.cpu_seconds) if args.cpu_seconds > 0 else None
    out = {
        "metric": "ocean-frames/s + HBM GB/s (% peak) at N=1024; batched & fwd+bwd vs plain on GPU",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox4x32-10 seed 42 + Box-Muller h0, generated on the GPU)",
        "config": {
            "workload": "plain single ocean N=%dx%d, L=2000 m, wind 25 m/s at 30 deg, A=4, lambda=1.3, "
                        "l=1 m, frames t=f/60 s; height/Dx/Dz/normals/Jacobian per frame into %d slices"
                        % (N, N, slots),
            "N": N, "frames_in_sequence": FRAMES, "output_slices": slots,
        },
        "frame": {
            "canonical_bytes_per_frame": CANON_FRAME * nn,
            "canonical_GBps": round(frame_gbps, 1),
            "hbm_frac": round(frame_gbps / HBM_PEAK_GBPS, 4),
            "wall_s_timed_region": round(t_wall, 3),
            "kernel_ms_per_launch": {k: round(v, 5) for k, v in stage_ms.items()},
            "frames_per_launch": lf,
            "schedule": sched,
        },
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dom_bytes,
        },
        "cpu_baseline": cpu,
        "single_frame": single,
        "batched_cascades": casc,
        "fit": fit,
        "batched_fit": bfit,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
