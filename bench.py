"""Benchmark: RK4 steps/s of kielflow's pseudo-spectral step on one MI355X.

Workload: the 256^3 equidistant box (L = 16 per side) with M = 1.5,
Q = -0.3, 1e-3 seeded density noise, Boltzmann electrons solved to a
relative residual of 1e-10 every stage, 2/3 dealiasing, penalization on,
dt = 1e-3.  `value` is the mixed-precision rate (FFT, RHS and Poisson sweeps
in FP32; RK4 state, accumulation, Poisson residual and Newton correction in
FP64); the FP64 rate of the same workload, their ratio and the per-field
relative L2 distance between the two runs are reported alongside.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

L_BOX = 16.0
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    p.add_argument("--fp64-only", action="store_true",
                   help="run only the FP64 leg of the workload and report its rate as value (for profiling "
                        "the FP64 kernels; implies --no-cpu --no-adjoint --no-batch)")
    p.add_argument("--cpu-n", type=int, default=64, help="box size of the CPU sample")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost leg")
    p.add_argument("--adjoint-steps", type=int, default=200)
    p.add_argument("--no-batch", action="store_true", help="skip the batched-boxes leg")
    p.add_argument("--no-config1", action="store_true", help="skip the 128^3 SiSt FP64 leg")
    p.add_argument("--config1-steps", type=int, default=20)
    p.add_argument("--only", choices=("adjoint", "config1", "batch"), default=None,
                   help="run only this leg and print its JSON (for profiling)")
    p.add_argument("--summarize-trace", nargs=3, metavar=("DIR", "TITLE", "COMMAND"),
                   help="print the text summary of a rocprofv3 --kernel-trace --stats csv run in DIR and exit")
    p.add_argument("--step-marker", default=None, help="kernel name fragment launched --per times per step")
    p.add_argument("--per", type=int, default=1)
    p.add_argument("--last", type=int, default=None, help="steps at the end of the trace to average over")
    p.add_argument("--first", type=int, default=None, help="average over the --last steps after this marked step")
    p.add_argument("--summarize-pmc", nargs=6, metavar=("FETCH", "WRITE", "CF", "CW", "TITLE", "COMMANDS"),
                   help="print the HBM traffic summary of four --pmc runs (COMMANDS separated by ';') and exit")
    p.add_argument("--smoothing", type=str, default="",
                   help="Poisson smoothing sweeps 'pre,post,coarse' (library default when empty)")
    p.add_argument("--long", type=int, default=0,
                   help="only run the mixed and FP64 steppers side by side for this many steps, printing the "
                        "per-field rel-L2 every 100 steps, and exit")
    p.add_argument("--batch-steps", type=int, default=100)
    p.add_argument("--batch-eager", action="store_true",
                   help="batch leg without graphs (host-launched steps), for kernel traces")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run the plain device copy (1 GiB, 5 times) that the HBM byte counters are "
                        "calibrated against, and exit")
    return p.parse_args()


def copy_calibration():
    """Five launches of the library's streaming copy kernel over 1 GiB: each
    reads and writes exactly 2^30 bytes, so FETCH_SIZE / WRITE_SIZE per
    dispatch of this kernel give the counters' scale."""
    import kielflow_amd as kf
    ctx = kf.Context(0)
    nbytes = 1 << 30
    a = ctx.zeros((nbytes // 8,), np.float64)
    b = ctx.zeros((nbytes // 8,), np.float64)
    for _ in range(5):
        b.copy_from(a)
    ctx.sync()
    print(json.dumps({"copy_calibration_bytes_per_launch": nbytes, "launches": 5}))


SMOOTHING = None


def run_precision(kf, ctx, g, axes, prec, nd_d, mask_d, Fn0, steps, warmup, prof_family=None):
    P = kf.KF_MIXED if prec == "mixed" else kf.KF_FP64
    ps = kf.Poisson(g, P, tol=1e-10, max_it=60)
    if SMOOTHING:
        ps.set_smoothing(*SMOOTHING)
    rhs = kf.Rhs(g, ps, M=1.5, tau=0.1, theta=100.0, sv_on=False, pen_on=True, precision=P)
    phi = g.rfield()
    rhs.set_fields(phi, nd_d, mask_d)
    F = [g.cfield(), g.cfield(), g.cfield(), g.cfield()]
    F[3].copy_from(Fn0)
    rk = kf.Rk4(rhs, 1e-3)
    rk.step(F, warmup)
    ctx.sync()
    n0, c0 = ps.newton_counts()
    t0 = time.perf_counter()
    rk.step(F, steps)
    ctx.sync()
    t1 = time.perf_counter()
    n1, c1 = ps.newton_counts()
    out = {"steps_per_s": steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps,
           "poisson_failures": rhs.poisson_failures(),
           "newton_corrections_per_solve": round((c1 - c0) / max(n1 - n0, 1), 3)}
    # state after warmup + timed steps (compared across precisions)
    out["state"] = [f.get() for f in F]
    out["phi"] = phi.get()
    for fam in (prof_family or ()):
        # the same steps again with one kernel family bracketed by HIP events
        # on the library's stream
        kf.api.profile_enable(fam)
        rk.step(F, steps)
        ctx.sync()
        n, ms = kf.api.profile_read()
        kf.api.profile_enable(-1)
        out.setdefault("prof", {})[fam] = (n, ms)
    for o in (rk, rhs, ps):
        o.close()
    return out


def cpu_baseline(n_sample, n_target):
    """The FP64 oracle (a C restatement of the reference path, no FFTW) timed
    on the host: one RHS evaluation of the same configuration (Poisson to a
    relative residual of 1e-10) on an n_sample^3 box, after one warm-up
    evaluation, median of 3.  Once on one thread (the reference is
    single-threaded: no OpenMP regions, no FFTW threads) and once on all host
    threads.  A RK4 step is 4 RHS evaluations plus negligible AXPYs; the
    rate is scaled to n_target^3 by the n^3 log2 n work ratio."""
    import kielflow_amd as kf
    from oracle import oracle as O
    O.build()
    og = O.Grid.cube(n_sample, L_BOX)
    axes = kf.cube(n_sample, L_BOX)
    ln_n, nd, mask = kf.setup_host(axes, seed=0, amp=1e-3, Q=-0.3, a=2.0 * L_BOX / n_sample)
    st0 = [np.zeros(og.cshape, complex) for _ in range(3)] + [og.r2c(ln_n)]
    scale = (n_target / n_sample) ** 3 * (np.log2(n_target) / np.log2(n_sample))

    def timed(threads):
        O.set_threads(threads)
        used = O.get_threads()
        ts = []
        for rep in range(4):
            r = O.Rhs(og, M=1.5, tau=0.1, theta=100.0, sv_on=False, pen_on=True,
                      poisson_mode=O.POISSON_FAS, poisson_tol=1e-10, poisson_max=200)
            r.set_fields(np.zeros(og.rshape), nd, mask)
            t0 = time.perf_counter()
            r.eval(0.0, st0)
            if rep:
                ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), used

    t1, c1 = timed(1)
    all_threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    tn, cn = timed(all_threads)
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    step1 = 4 * t1 * scale
    stepn = 4 * tn * scale
    return {"value": round(1.0 / step1, 6), "unit": "steps/s", "cores": c1, "kind": "port",
            "sample": (f"one RHS evaluation (Poisson to 1e-10) of the same configuration at {n_sample}^3, "
                       f"median of 3 after a warm-up: {t1:.3f} s on {c1} thread of {cpu}; a {n_target}^3 "
                       f"RK4 step estimated as 4 RHS x {scale:.1f} (n^3 log2 n ratio)"),
            "extrapolation_factor": round(4 * scale, 2),
            "all_cores": {"value": round(1.0 / stepn, 6), "unit": "steps/s", "cores": cn,
                          "rhs_seconds": round(tn, 3)},
            "host_threads_available": os.cpu_count()}


def adjoint_cost(kf, ctx, n=64, steps=200, k=20, m=3.0):
    """Forward run vs gradient run on the 64^3 hyperbolic-sinh stretched box
    (m = 3): 200 FP64 RK4 steps, dt = 1e-3, M = 1, Q = -0.1, J = grain x-drag
    at the final step.  The headline gradient run checkpoints every k = 20
    steps and recomputes each segment's stage intermediates once (the capture
    budget is set to one segment's); a second run keeps the whole trajectory's
    intermediates in HBM (about 35 GB here), the fastest mode when it fits.
    The timed runs include their allocations."""
    axes = [kf.axis(kf.HYST, n, -0.5 * L_BOX, L_BOX, m) for _ in range(3)]
    x0 = kf.api.axis_values(axes[0], n // 2)[0]
    x1 = kf.api.axis_values(axes[0], n // 2 + 1)[0]
    Q = -0.1
    ln_n, nd, mask = kf.setup_host(axes, seed=0, amp=1e-3, Q=Q, a=2.0 * (x1 - x0))
    g = kf.Grid(ctx, axes)
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-10, max_it=60)
    rhs = kf.Rhs(g, ps, M=1.0, tau=0.1, theta=100.0, sv_on=False, pen_on=True)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    fft = kf.FFT(g, kf.KF_FP64)
    Fn0 = g.cfield()
    fft.r2c(ctx.to_device(ln_n), Fn0)
    dV = ctx.to_device(kf.api.cell_volumes(axes))
    dG = ctx.to_device(nd / Q)
    # transposed Poisson solves to the forward solves' tolerance
    adj = kf.Adjoint(rhs, 1e-3, tol=1e-10)

    def fresh():
        phi.upload(np.zeros(g.rshape))
        y = [g.cfield() for _ in range(4)]
        y[3].copy_from(Fn0)
        return y

    rk = kf.Rk4(rhs, 1e-3)
    rk.step(fresh(), 2)  # warm-up
    y = fresh()
    ctx.sync()
    t0 = time.perf_counter()
    rk.step(y, steps)
    ctx.sync()
    t_fwd = time.perf_counter() - t0
    J_fwd = adj.objective(y, dV)
    names = {kf.api.KF_ADJ_RECOMPUTE: "recomputed per step", kf.api.KF_ADJ_SEGMENTS: "segments recomputed once",
             kf.api.KF_ADJ_TRAJECTORY: "whole trajectory kept by the forward sweep"}
    per_step = adj.step_capture_bytes()

    def gradient(budget):
        adj.set_capture_budget(budget)
        y = fresh()
        dF = g.cfield()
        ctx.sync()
        t0 = time.perf_counter()
        J, dM, dQ = adj.run(y, steps, k, dV, dG, dF)
        ctx.sync()
        t = time.perf_counter() - t0
        mode, lin_it = adj.run_info()
        dF.free()
        return t, J, dM, dQ, names[mode], lin_it

    # checkpoints every k steps, each segment recomputed once (as stated for
    # this workload), then the whole trajectory kept
    t_seg, J, dM, dQ, kept, lin_it = gradient(k * per_step)
    t_traj, J2, dM2, dQ2, kept2, _ = gradient(64 << 30)
    for o in (adj, rk, rhs, ps, fft):
        o.close()
    return {"workload": f"{n}^3 HySt (m={m}) FP64, {steps} RK4 steps, checkpoints every {k}",
            "forward_s": round(t_fwd, 3), "gradient_run_s": round(t_seg, 3),
            "gradient_over_forward": round(t_seg / t_fwd, 3),
            "reverse_sweep_over_forward": round((t_seg - t_fwd) / t_fwd, 3),
            "J": J, "J_forward_only": J_fwd, "dJ_dM": dM, "dJ_dQ": dQ,
            "intermediates": kept, "transposed_poisson_iterations": lin_it,
            "stage_bytes_per_step": per_step,
            "trajectory_mode": {"gradient_run_s": round(t_traj, 3),
                                "gradient_over_forward": round(t_traj / t_fwd, 3), "intermediates": kept2,
                                "dJ_dM_rel_diff": abs(dM2 - dM) / max(abs(dM), 1e-300),
                                "dJ_dQ_rel_diff": abs(dQ2 - dQ) / max(abs(dQ), 1e-300)}}


def config1_rate(kf, ctx, n=128, steps=20, warm=2):
    """FP64 RK4 steps/s on the 128^3 sine-stretched box (S'(0) = 1/3),
    M = 1.2, Q = -0.2, 1e-3 density noise, spectral viscosity and
    penalization on, Poisson to 1e-10, dt from cfl_all before every step."""
    m = L_BOX * (1 - 1 / 3) / (2 * np.pi)
    axes = [kf.axis(kf.SIST, n, -0.5 * L_BOX, L_BOX, m) for _ in range(3)]
    x0 = kf.api.axis_values(axes[0], n // 2)[0]
    x1 = kf.api.axis_values(axes[0], n // 2 + 1)[0]
    ln_n, nd, mask = kf.setup_host(axes, seed=0, amp=1e-3, Q=-0.2, a=2.0 * (x1 - x0))
    g = kf.Grid(ctx, axes)
    ps = kf.Poisson(g, kf.KF_FP64, tol=1e-10, max_it=80)
    rhs = kf.Rhs(g, ps, M=1.2, tau=0.1, theta=100.0, sv_on=True, pen_on=True)
    phi = g.rfield()
    rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
    fft = kf.FFT(g, kf.KF_FP64)
    F = [g.cfield() for _ in range(4)]
    fft.r2c(ctx.to_device(ln_n), F[3])
    rk = kf.Rk4(rhs, 1e-3)
    dt = 1e-3

    def step():
        nonlocal dt
        dt = g.cfl(dt, F[0], F[1], F[2])
        rk.set_dt(dt)
        rk.step(F, 1)

    for _ in range(warm):
        step()
    ctx.sync()
    n0, c0 = ps.newton_counts()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    t = time.perf_counter() - t0
    n1, c1 = ps.newton_counts()
    fails = rhs.poisson_failures()
    for o in (rk, fft, rhs, ps):
        o.close()
    g.close()
    return {"workload": f"{n}^3 SiSt (S'(0) = 1/3) FP64 RK4, spectral viscosity and penalization on, "
                        "dt from cfl_all every step", "steps_per_s": round(steps / t, 3),
            "ms_per_step": round(1e3 * t / steps, 3), "steps": steps, "last_dt": dt, "poisson_failures": fails,
            "newton_corrections_per_solve": round((c1 - c0) / max(1, n1 - n0), 3)}


def batch_sweep(kf, ctx, nbox=64, steps=100, warm=5, seed=0, nstreams=4, poisson_iterations=3, check_boxes=4,
                eager=False):
    """Parameter sweep of independent boxes: sizes uniform in {32, 48, 64, 96}^3,
    M uniform in [0.5, 2], Q uniform in [-0.5, -0.05], 1e-3 noise, mixed
    precision, one captured graph per box and step over `nstreams` streams
    (eager: no graphs, every step launched from the host, for kernel traces).
    Error: the first box of each size against an FP64 solo run."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([32, 48, 64, 96], nbox)
    Ms = rng.uniform(0.5, 2.0, nbox)
    Qs = rng.uniform(-0.5, -0.05, nbox)
    boxes, fields = [], []
    for b in range(nbox):
        n = int(sizes[b])
        axes = kf.cube(n, L_BOX)
        x0 = kf.api.axis_values(axes[0], n // 2)[0]
        x1 = kf.api.axis_values(axes[0], n // 2 + 1)[0]
        fields.append((axes, kf.setup_host(axes, seed=1000 + b, amp=1e-3, Q=Qs[b], a=2.0 * (x1 - x0))))
        boxes.append((axes, kf.api.Parameters(Ms[b], 0.1, 100.0, 0.0, 0.0)))
    B = kf.Batch(ctx, boxes, precision=kf.KF_MIXED, nstreams=nstreams)
    for b, (axes, (ln_n, nd, mask)) in enumerate(fields):
        B.set_fields(b, ctx.to_device(nd), ctx.to_device(mask))
        F, _ = B.box(b)
        g = kf.Grid(ctx, axes)
        fft = kf.FFT(g, kf.KF_FP64)
        fft.r2c(ctx.to_device(ln_n), F[3])
        fft.close()
        g.close()
    B.steps(warm)
    B.sync()
    if not eager:
        B.capture(poisson_iterations)
    ctx.sync()
    t0 = time.perf_counter()
    B.steps(steps)
    B.sync()
    t = time.perf_counter() - t0
    fails = B.poisson_failures()
    errs = {}
    for n in (32, 48, 64, 96):
        idx = [b for b in range(nbox) if sizes[b] == n]
        if not idx or len(errs) >= check_boxes:
            continue
        b = idx[0]
        axes, (ln_n, nd, mask) = fields[b]
        g = kf.Grid(ctx, axes)
        ps = kf.Poisson(g, kf.KF_FP64, tol=1e-10, max_it=100)
        rhs = kf.Rhs(g, ps, M=Ms[b], tau=0.1, theta=100.0, sv_on=False, pen_on=True)
        phi = g.rfield()
        rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
        F = [g.cfield() for _ in range(4)]
        fft = kf.FFT(g, kf.KF_FP64)
        fft.r2c(ctx.to_device(ln_n), F[3])
        rk = kf.Rk4(rhs, 1e-3)
        rk.step(F, warm + steps)
        Fb, _ = B.box(b)
        errs[f"box{b}_n{n}"] = {f: rel_l2(Fb[i].get(), F[i].get()) for i, f in
                                 enumerate(("FUx", "FUy", "FUz", "ln_n"))}
        for o in (rk, fft, rhs, ps, g):
            o.close()
    B.close()
    pts = float(sum(int(n) ** 3 for n in sizes))
    return {"workload": f"{nbox} boxes, sizes uniform in {{32,48,64,96}}^3 (seed {seed}), M in [0.5,2], "
                        f"Q in [-0.5,-0.05], mixed precision, {steps} "
                        f"{'eager' if eager else 'graph-replayed'} RK4 steps",
            "box_steps_per_s": round(nbox * steps / t, 2), "batch_steps_per_s": round(steps / t, 3),
            "grid_point_steps_per_s": pts * steps / t, "sizes": {str(n): int((sizes == n).sum()) for n in
                                                                (32, 48, 64, 96)},
            "streams": nstreams, "poisson_newton_iterations_in_graph": poisson_iterations,
            "poisson_failures": fails, "rel_l2_vs_fp64_solo": errs}


def long_run(kf, n, steps, every=100):
    """Mixed and FP64 runs of the bench workload side by side; per-field
    rel-L2 of mixed against FP64 every `every` steps."""
    ctx = kf.Context(0)
    axes = kf.cube(n, L_BOX)
    ln_n, nd, mask = kf.setup_host(axes, seed=0, amp=1e-3, Q=-0.3, a=2.0 * L_BOX / n)
    g = kf.Grid(ctx, axes)
    fft = kf.FFT(g, kf.KF_FP64)
    runs = {}
    for name, P in (("mixed", kf.KF_MIXED), ("fp64", kf.KF_FP64)):
        ps = kf.Poisson(g, P, tol=1e-10, max_it=60)
        rhs = kf.Rhs(g, ps, M=1.5, tau=0.1, theta=100.0, sv_on=False, pen_on=True, precision=P)
        phi = g.rfield()
        rhs.set_fields(phi, ctx.to_device(nd), ctx.to_device(mask))
        F = [g.cfield() for _ in range(4)]
        fft.r2c(ctx.to_device(ln_n), F[3])
        runs[name] = dict(ps=ps, rhs=rhs, phi=phi, F=F, rk=kf.Rk4(rhs, 1e-3), t=0.0)
    rows = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        ms = {}
        for name, r in runs.items():
            ctx.sync()
            t0 = time.perf_counter()
            r["rk"].step(r["F"], k)
            ctx.sync()
            dt_ = time.perf_counter() - t0
            r["t"] += dt_
            ms[name] = 1e3 * dt_ / k
        done += k
        a, b = runs["mixed"], runs["fp64"]
        row = {"step": done, "mixed_ms_per_step": round(ms["mixed"], 3), "fp64_ms_per_step": round(ms["fp64"], 3)}
        for i, f in enumerate(("FUx", "FUy", "FUz", "ln_n")):
            row[f] = rel_l2(a["F"][i].get(), b["F"][i].get())
        row["phi"] = rel_l2(a["phi"].get(), b["phi"].get())
        rows.append(row)
        print(json.dumps(row), flush=True)
    out = {"workload": f"{n}^3 equidistant box, M=1.5, Q=-0.3, dt=1e-3, {steps} RK4 steps",
           "mixed_steps_per_s": steps / runs["mixed"]["t"], "fp64_steps_per_s": steps / runs["fp64"]["t"],
           "poisson_failures": {k: r["rhs"].poisson_failures() for k, r in runs.items()},
           "max_rel_l2": max(max(v for kk, v in r.items() if kk in ("FUx", "FUy", "FUz", "ln_n", "phi"))
                             for r in rows)}
    print(json.dumps(out), flush=True)


def hbm_traffic(kernel, functor, grid=None):
    """HBM bytes per dispatch of a kernel from the committed counter summary
    (profiles/hbm_traffic_256_mixed.txt: FETCH_SIZE + WRITE_SIZE per
    dispatch, calibrated on the library's copy kernel), launch-weighted over
    the matching lines; (None, None) when absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_256_mixed.txt")
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None, None
    tot, cnt = 0.0, 0
    kernels = (kernel,) if isinstance(kernel, str) else tuple(kernel)
    for ln in lines:
        if not any(k in ln for k in kernels) or (functor and functor not in ln):
            continue
        parts = ln.split()
        try:
            g, nlaunch, total = parts[-5], int(parts[-4]), float(parts[-1])
        except (ValueError, IndexError):
            continue
        if grid is not None and g != grid:
            continue
        tot += total * nlaunch
        cnt += nlaunch
    if not cnt:
        return None, None
    return round(tot / cnt * 1e6), "profiles/hbm_traffic_256_mixed.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE)"


def _csv_rows(d, suffix):
    import csv
    import glob
    paths = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not paths:
        raise SystemExit(f"no *{suffix} under {d}")
    rows = []
    for pth in paths:
        with open(pth) as f:
            rows.extend(csv.DictReader(f))
    return rows


def summarize_trace(d, title, command, marker=None, per=1, last=None, top=30, first=None):
    """Text summary of a `rocprofv3 --kernel-trace --stats --output-format
    csv -d D` run: the tool's per-kernel statistics (top entries), and, with
    a step marker (a kernel name fragment launched `per` times per step), the
    per-step wall time, kernel time, launches and idle time over `last`
    marked steps: the last ones of the trace, or those after step `first`."""
    stats = _csv_rows(d, "kernel_stats.csv")
    tot_ns = sum(float(r["TotalDurationNs"]) for r in stats)
    ndisp = sum(int(r["Calls"]) for r in stats)
    out = [title, "", "Command (from the repo root):", "  " + command, "",
           f"total kernel time {tot_ns / 1e6:.2f} ms over {ndisp} dispatches", ""]
    k = 0
    if marker:
        tr = sorted(_csv_rows(d, "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
        ends = [i for i, r in enumerate(tr) if marker in r["Kernel_Name"]][per - 1::per]
        nst = len(ends)
        k = min(last or nst - 1, nst - 1)
        lo = nst - 1 - k if first is None else min(first, nst - 1 - k)
        if k >= 1:
            sel = tr[ends[lo] + 1:ends[lo + k] + 1]
            t0, t1 = int(sel[0]["Start_Timestamp"]), int(sel[-1]["End_Timestamp"])
            busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in sel)
            which = f"steps {lo + 2}..{lo + k + 1}" if first is not None else f"the last {k}"
            out += [f"steps in the trace: {nst} (marker '{marker}', {per} per step); {which}:",
                    f"  wall {(t1 - t0) / 1e6 / k:.3f} ms/step, kernels {busy / 1e6 / k:.3f} ms/step, "
                    f"idle {(t1 - t0 - busy) / 1e6 / k:.3f} ms/step, {len(sel) / k:.1f} launches/step", ""]
    def short(n):
        return n.replace("void ", "").replace("kf::fft::", "").replace("kf::", "")[:110]

    if marker and k >= 1:
        import collections
        agg = collections.defaultdict(lambda: [0, 0])
        for r in sel:
            a = agg[r["Kernel_Name"]]
            a[0] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            a[1] += 1
        tot = sum(v[0] for v in agg.values())
        out.append(f"per step over those {k} steps:")
        out.append(f"{'%':>6} {'ms/step':>9} {'calls/step':>10} {'avg us':>9}  kernel")
        for name, (ns, c) in sorted(agg.items(), key=lambda kv: -kv[1][0])[:top]:
            out.append(f"{100 * ns / tot:6.2f} {ns / 1e6 / k:9.3f} {c / k:10.1f} {ns / c / 1e3:9.2f}  {short(name)}")
        return "\n".join(out) + "\n"
    out.append(f"{'%':>6} {'total ms':>9} {'calls':>6} {'avg us':>9}  kernel")
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        out.append(f"{100 * float(r['TotalDurationNs']) / tot_ns:6.2f} {float(r['TotalDurationNs']) / 1e6:9.3f} "
                   f"{int(r['Calls']):6d} {float(r['AverageNs']) / 1e3:9.2f}  {short(r['Name'])}")
    return "\n".join(out) + "\n"


def summarize_pmc(fetch_dir, write_dir, cf_dir, cw_dir, title, commands, min_mb=20.0):
    """HBM bytes per dispatch from two counter runs (FETCH_SIZE, WRITE_SIZE,
    KiB per dispatch), each scaled by 2^30 bytes over the same counter on the
    library's 1 GiB copy kernel (k_copy16); one line per kernel and grid
    size, in the format hbm_traffic() reads."""
    import collections

    def per_kernel(d, counter):
        acc = collections.defaultdict(lambda: [0.0, 0])
        for r in _csv_rows(d, "counter_collection.csv"):
            if r.get("Counter_Name") != counter:
                continue
            grid = r.get("Grid_Size") or r.get("Grid_Size_X")
            a = acc[(r["Kernel_Name"], grid)]
            a[0] += float(r["Counter_Value"])
            a[1] += 1
        return acc

    def calib(d, counter):
        vals = [v[0] / v[1] for (k, _), v in per_kernel(d, counter).items() if "k_copy16" in k]
        if not vals:
            raise SystemExit(f"no k_copy16 dispatch in {d}")
        return vals[0]

    cf, cw = calib(cf_dir, "FETCH_SIZE"), calib(cw_dir, "WRITE_SIZE")
    ff, fw = (1 << 30) / (cf * 1024), (1 << 30) / (cw * 1024)
    fe, wr = per_kernel(fetch_dir, "FETCH_SIZE"), per_kernel(write_dir, "WRITE_SIZE")
    out = [title, "", "Commands (from the repo root; each counter pass is its own run):"]
    out += ["  " + c for c in commands]
    out += ["", f"Calibration: the library's copy kernel k_copy16 over 1 GiB reads FETCH_SIZE = {cf:.0f} KiB and",
            f"WRITE_SIZE = {cw:.0f} KiB per dispatch; factors applied below: fetch x {ff:.4f}, write x {fw:.4f}.", "",
            f"{'kernel':<82}{'grid':>9} {'n':>4} {'fetch MB':>9} {'write MB':>9} {'total MB':>9}"]
    lines = []
    for key, (fsum, fn) in fe.items():
        wsum, wn = wr.get(key, [0.0, 0])
        fmb = fsum / fn * 1024 * ff / 1e6
        wmb = (wsum / wn * 1024 * fw / 1e6) if wn else 0.0
        if fmb + wmb < min_mb:
            continue
        name = key[0].replace("void ", "").replace("kf::fft::", "").replace("kf::", "")
        name = name if len(name) <= 79 else name[:76] + "..."
        lines.append((fmb + wmb, f"{name:<82}{key[1]:>9} {fn:>4} {fmb:9.1f} {wmb:9.1f} {fmb + wmb:9.1f}"))
    out += [ln for _, ln in sorted(lines, key=lambda t: -t[0])]
    return "\n".join(out) + "\n"


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (independent boxes "
              "would shard one per GPU with no exchange; a single box is not split across GPUs).",
              file=sys.stderr)
        sys.exit(2)
    global SMOOTHING
    if args.smoothing:
        SMOOTHING = tuple(int(v) for v in args.smoothing.split(","))
    if args.copy_calibration:
        copy_calibration()
        return
    if args.summarize_trace:
        d, title, cmd = args.summarize_trace
        sys.stdout.write(summarize_trace(d, title, cmd, args.step_marker, args.per, args.last, first=args.first))
        return
    if args.summarize_pmc:
        f, w, cf, cw, title, cmds = args.summarize_pmc
        sys.stdout.write(summarize_pmc(f, w, cf, cw, title, [c.strip() for c in cmds.split(";")]))
        return
    if args.long:
        import kielflow_amd as kf
        long_run(kf, args.n, args.long)
        return
    import kielflow_amd as kf
    if args.only:
        ctx = kf.Context(0)
        if args.only == "adjoint":
            out = adjoint_cost(kf, ctx, steps=args.adjoint_steps)
        elif args.only == "config1":
            out = config1_rate(kf, ctx, steps=args.config1_steps)
        else:
            out = batch_sweep(kf, ctx, steps=args.batch_steps, eager=args.batch_eager)
        print(json.dumps(out))
        return
    n = args.n
    axes = kf.cube(n, L_BOX)
    a = 2.0 * L_BOX / n
    ln_n, nd, mask = kf.setup_host(axes, seed=0, amp=1e-3, Q=-0.3, a=a)
    ctx = kf.Context(0)
    g = kf.Grid(ctx, axes)
    fft = kf.FFT(g, kf.KF_FP64)
    d_ln = ctx.to_device(ln_n)
    Fn0 = g.cfield()
    fft.r2c(d_ln, Fn0)
    nd_d, mask_d = ctx.to_device(nd), ctx.to_device(mask)
    del d_ln

    if args.fp64_only:
        r = run_precision(kf, ctx, g, axes, "fp64", nd_d, mask_d, Fn0, args.steps, args.warmup)
        print(json.dumps({"metric": "RK4 steps/s at 256^3, FP64 leg only", "value": round(r["steps_per_s"], 3),
                          "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": round(r["ms_per_step"], 3), "dtype": "f64",
                          "poisson_failures": r["poisson_failures"]}))
        return
    fams = (kf.api.PROF_FFT_ZC2R, kf.api.PROF_FFT_Y, kf.api.PROF_RB0)
    mixed = run_precision(kf, ctx, g, axes, "mixed", nd_d, mask_d, Fn0, args.steps, args.warmup, fams)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_precision(kf, ctx, g, axes, "fp64", nd_d, mask_d, Fn0, args.steps, args.warmup)

    def roofline(fam, alg_bytes, kernel, traffic_key):
        launches, ms = mixed["prof"][fam]
        avg_ms = ms / max(launches, 1)
        achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
        traffic, tsrc = hbm_traffic(*traffic_key)
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
                "avg_launch_us": round(avg_ms * 1e3, 2), "launches": launches,
                "alg_bytes_per_launch": int(alg_bytes), "traffic_source": tsrc}

    nc = n * n * (n // 2 + 1)
    C, Rb = 8 * nc, 4 * n ** 3  # bytes of one FP32 complex / real field
    # The dominant kernel family: the z passes of the inverse transforms
    # (half spectrum -> real rows), timed together.  One RHS evaluation runs
    # 8 of them; with their fused loads and stores they move 25 C + 28 R
    # bytes (C, R: one FP32 complex / real field; an FP64 field is twice):
    #   rho (FP64 transform of the FP64 stage input): 2C + n_d (2R) in,
    #   ln n (R) and rho (2R) out                                  = 2C + 5R
    #   velocities: C in, U_b out, 3 times                         = 3 (C + R)
    #   per field (k_zpen): the three staged derivative rows and the
    #   linear terms' rows in (4C), U_0..2, f (U_a or ln n) and the
    #   mask in (5R), the blend's forward z transform out (C), 4 times
    #                                                              = 4 (5C + 5R)
    roof = roofline(kf.api.PROF_FFT_ZC2R, (25 * C + 28 * Rb) / 8,
                    "FFT z passes of the inverse transforms (k_zc2r, k_zpen: products, penalization blend and "
                    "forward z transform fused; FP32, and the FP64 one of rho; Nz = 256)",
                    (("k_zc2r<float, 128", "k_zpen<float, 128",
                      "k_zc2r<double, 128, 16, RowLoadC<double>, ZStoreK<double, 6>"), None))
    # the FFT y passes (strided, length 256).  Per RHS evaluation: 4 FP32
    # ones that read a staged field once and write it and its i k_y multiple
    # (3 C), 13 FP32 ones that read and write one complex field (2 C), and
    # the FP64 inverse pass of rho's transform (4 C): 42 C over 18
    roof_y = roofline(kf.api.PROF_FFT_Y, 42 * C / 18, "FFT y passes (k_strided, k_strided_ik; FP32, and the FP64 "
                      "one of rho; length 256)",
                      (("k_strided<float, 256", "k_strided_ik<float, 256",
                        "k_strided<double, 256, 8, 1, LoadMasked<double>"), "LoadMasked"))
    # the finest-level red-black half sweep of the Poisson multigrid: the
    # other colour's x, this colour's b and e in, this colour's x out,
    # 4 * (N/2) * 4 bytes (FP32)
    roof_rb = roofline(kf.api.PROF_RB0, 8 * n ** 3, "k_rb_split (red-black half sweep, finest level, FP32)",
                       ("k_rb_split<float>", None, str(n ** 3 // 2)))

    extra = {"poisson_failures_mixed": mixed["poisson_failures"]}
    if fp64 is not None:
        errs = {f: rel_l2(mixed["state"][i], fp64["state"][i]) for i, f in
                enumerate(("FUx", "FUy", "FUz", "ln_n"))}
        errs["phi"] = rel_l2(mixed["phi"], fp64["phi"])
        extra.update({"fp64_steps_per_s": round(fp64["steps_per_s"], 3),
                      "fp64_ms_per_step": round(fp64["ms_per_step"], 3),
                      "mixed_over_fp64": round(mixed["steps_per_s"] / fp64["steps_per_s"], 3),
                      "newton_corrections_per_solve": {"mixed": mixed["newton_corrections_per_solve"],
                                                       "fp64": fp64["newton_corrections_per_solve"]},
                      "rel_l2_mixed_vs_fp64_after_steps": errs,
                      "poisson_failures_fp64": fp64["poisson_failures"]})

    if not args.no_adjoint:
        extra["adjoint"] = adjoint_cost(kf, ctx, steps=args.adjoint_steps)
    if not args.no_config1:
        extra["config1"] = config1_rate(kf, ctx, steps=args.config1_steps)
    if not args.no_batch:
        extra["batch"] = batch_sweep(kf, ctx, steps=args.batch_steps)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n, n)
    line = {
        "metric": "RK4 steps/s at 256^3 (FP64 vs mixed, rel-L2<=1e-5); adjoint/forward cost",
        "value": round(mixed["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 1e-3 density noise, Philox4x32-10)",
        "config": {"workload": f"{n}^3 equidistant box RK4 step, mixed precision (FP32 FFT/RHS/sweeps, "
                               "FP64 state, accumulation, Poisson residual)",
                   "n": n, "L": L_BOX, "dt": 1e-3, "M": 1.5, "Q": -0.3, "tau": 0.1, "theta": 100.0,
                   "penalization": True, "spectral_viscosity": False, "poisson_tol": 1e-10},
        "roofline": roof,
        "roofline_fft_y": roof_y, "roofline_poisson_sweep": roof_rb,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
