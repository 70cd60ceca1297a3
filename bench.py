#!/usr/bin/env python3
"""Benchmark of the CT-operator hot path on one MI355X.

Workload (the 256^3 cone-beam configuration): 256^3 volume on [-1,1]^3 with
20 seeded random ellipsoids, circular orbit, 360 views over 2*pi, 512x384
flat panel, source-to-centre 4, source-to-detector 8.  One step is one
forward projection A (Siddon) plus one backprojection A^T (exact
voxel-driven adjoint) of the whole volume, inputs resident in HBM.

Prints ONE JSON line.  value = proj+backproj GUPS = 2*N_vox*n_views /
(t_A + t_At) / 1e9 over the timed steps (FP32 plain path).  Beside it:
  cgls      -- 50 CGLS iterations of the FP32 plain path (Siddon + exact
               adjoint, FP32 storage), of the precision change alone (FP16 /
               BF16 storage), of the operator change alone (matched adjoint)
               and of the mixed path (both), device-resident with FP64
               reductions and graph-replayed iterations: iterations/s and the
               relative L2 difference to the FP64 GPU solve;
  roofline  -- the dominant kernel's algorithmic bytes per launch over its
               measured duration, against 8 TB/s;
  framelet  -- 27-band 3-D framelet analysis and synthesis of the volume,
               bandwidth against their algorithmic bytes;
  batch     -- 64 uneven 2-D fan problems in persistent batched launches;
  c4        -- the 512^3, 720-view, 1024x768 configuration in BF16 storage:
               forward, exact and matched adjoint;
  cpu_baseline -- the oracle's restatement of the reference's CPU path on
               a bounded sample (one view at full size, median of 5 views),
               on 1 core and on every core, with the 1-core figure rescaled
               to the reference's own kernels by a ratio measured in the
               build container;
  c1        -- the 2-D fan configuration (256^2, 360 x 512) timed in full on
               the CPU restatement (A, A^T, 20 CGLS) and on the GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "proj+backproj GUPS & CGLS iters/s at 256³: mixed vs FP32 plain @ rel err≤1e-3"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_info():
    model, cores = platform.processor() or "unknown", os.cpu_count()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return model, cores


def _median_s(fn, runs=5):
    """Median wall time of `runs` calls after one untimed warm-up call."""
    import statistics
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def _calibration():
    """Reference / restatement speed ratios on one core, measured in the build
    container (tests/golden/cpu_calibration.json), or None."""
    try:
        with open(os.path.join(ROOT, "tests", "golden", "cpu_calibration.json")) as fh:
            return json.load(fh)
    except OSError:
        return None


def cpu_baseline(p, X, views=5, adj_views=5):
    """The oracle's C restatement of the reference's CPU path (the reference
    code does not travel to this host) on a bounded sample of the workload:
    one Siddon forward (median over `views` distinct views) plus one exact
    adjoint (median over the first `adj_views` of them) of the full 256^3
    volume for one view, after one untimed warm-up of each, on 1 core (the
    reference is single-threaded) and on every core of this process (OpenMP
    over rays / voxels): about 35 s of single-core work.  The restatement follows the reference's float
    operations, and the reference / restatement ratio measured on one core
    in the build container rescales the 1-core figure to the reference's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from framelab_amd import geometry as G
    nvox = p.nx * p.ny * p.nz
    picks = [int(k * p.nv / views) + 7 for k in range(views)]

    def one_view(v):
        return G.cone_para(p.na, p.nb, 1, N=(p.nx, p.ny, p.nz), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                           Lz=G.full_cone_height(), sd_phi=np.array([p.sd_phi[v]], np.float32))

    subs = [one_view(v) for v in picks]
    model, cores = cpu_info()
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    out, times = {}, {}
    for threads in (1, ncpu):
        O.set_threads(threads)
        O.atx_cone_f32(O.siddon_cone_f32(X, subs[0]), subs[0])   # warm-up
        tA, tAt = [], []
        for k, sub in enumerate(subs):
            t0 = time.perf_counter()
            y = O.siddon_cone_f32(X, sub)
            t1 = time.perf_counter()
            tA.append(t1 - t0)
            if k < adj_views:
                O.atx_cone_f32(y, sub)
                tAt.append(time.perf_counter() - t1)
        a, at = float(np.median(tA)), float(np.median(tAt))
        times[threads] = (a, at)
        out[threads] = {"value": 2.0 * nvox / (a + at) / 1e9, "unit": "GUPS", "cores": threads, "kind": "port",
                        "sample": f"one forward + one exact adjoint of the full 256^3 volume for one view of 512x384 "
                                  f"cells after one untimed warm-up (forward: median over {views} views, t_A={a:.2f}s; adjoint: median over "
                                  f"{min(views, adj_views)}, t_At={at:.2f}s); host: {model}, {cores} logical CPUs"}
    O.set_threads(1)
    res = dict(out[1])
    res["all_cores"] = out[ncpu]
    cal = _calibration()
    if cal:
        k = cal["kernels"]
        rf = k["cone_siddon_forward_256cube_1view"]["reference_over_restatement"]
        ra = k["cone_exact_adjoint_256cube_1view"]["reference_over_restatement"]
        tA1, tAt1 = times[1]
        res["reference_equivalent_1core"] = {
            "value": 2.0 * nvox / (tA1 * rf + tAt1 * ra) / 1e9, "unit": "GUPS",
            "ratios": {"forward": rf, "exact_adjoint": ra}, "calibrated_on": cal["host"]}
    return res


def cpu_c1(dev):
    """The 2-D fan configuration timed in full on the CPU restatement: 256^2
    image of 10 seeded ellipses, 360 views x 512 bins, SO 4, SD 8; Siddon
    forward A, exact pixel-driven adjoint A^T (medians of 5 after a warm-up),
    and 20 CGLS iterations on the noiseless sinogram, on 1 core and on every
    core; the same workload on the GPU beside it (A and A^T medians of 20
    after a warm-up)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import torch
    import oracle as O
    from framelab_amd import geometry as G
    from framelab_amd.phantoms import ellipses_2d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls
    p = G.fan_para(512, 360, N=(256, 256), SO=4.0, OD=4.0, Ly=G.full_fan_width())
    X = ellipses_2d(256, 10, seed=1).astype(np.float32)
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    out = {"workload": "fan 256^2, 360 views x 512 bins, SO=4 SD=8: A, A^T (CPU medians of 5, GPU medians of 20, each after a warm-up) and 20 CGLS iterations"}
    y = O.siddon_fan_f32(X, p)
    for threads in (1, ncpu):
        O.set_threads(threads)
        tA = _median_s(lambda: O.siddon_fan_f32(X, p))
        tAt = _median_s(lambda: O.atx_fan_f32(y, p))
        A = lambda v: O.siddon_fan_f32(v.astype(np.float32), p).astype(np.float64)
        At = lambda w: O.atx_fan_f32(w.astype(np.float32), p).astype(np.float64)
        b = At(y)
        t0 = time.perf_counter()
        O.cgls(A, At, b, np.zeros_like(b), 0.0, 20)
        tc = time.perf_counter() - t0
        out[f"cpu_{threads}_cores"] = {"t_A_s": tA, "t_At_s": tAt, "cgls_20_s": tc, "cgls_iters_per_s": 20 / tc}
    O.set_threads(1)
    P = Projector(p)
    x = torch.from_numpy(X).to(dev)
    yg = P.forward(x)
    P.adjoint(yg)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    tg = []
    for _ in range(20):
        evs[0].record()
        P.forward(x, out=yg)
        evs[1].record()
        xb = P.adjoint(yg)
        evs[2].record()
        torch.cuda.synchronize()
        tg.append((evs[0].elapsed_time(evs[1]), evs[1].elapsed_time(evs[2])))
    b = P.adjoint(yg)
    cgls(P, b, torch.zeros_like(b), 0.0, 20, graph=True, log=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cgls(P, b, torch.zeros_like(b), 0.0, 20, graph=True, log=False)
    torch.cuda.synchronize()
    tc = time.perf_counter() - t0
    out["gpu"] = {"t_A_ms": float(np.median([t[0] for t in tg])), "t_At_ms": float(np.median([t[1] for t in tg])),
                  "cgls_20_s": tc, "cgls_iters_per_s": 20 / tc}
    del xb
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated against a 1 GiB copy),
    or None when the kernel has no entry."""
    path = os.path.join(ROOT, "profiles", "cone256_hbm_traffic.tsv")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("#"):
                    continue
                name, rd, wr = line.rstrip("\n").split("\t")
                if name == kernel:
                    return float(rd) + float(wr)
    except OSError:
        pass
    return None


VALU_PEAK_FP32, VALU_PEAK_INT = 0.239, 0.469   # wave-level VALU instructions per SIMD per cycle (measured)
SIMDS, CLOCK_HZ = 1024, 2.4e9


def valu_issue(kernel, ms):
    """VALU issue rate of `kernel` over a live-timed launch of `ms`
    milliseconds: its wave-level VALU instruction count per launch from the
    committed counter summary (SQ_INSTS_VALU), divided by SIMDs x cycles.
    No instruction mix issues faster than the integer peak, so the fraction
    of that peak is a lower bound on the share of VALU issue capacity used.
    None when the kernel has no entry."""
    path = os.path.join(ROOT, "profiles", "cone256_valu.tsv")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("#"):
                    continue
                name, insts, lanes = line.rstrip("\n").split("\t")
                if name == kernel:
                    rate = float(insts) / (SIMDS * CLOCK_HZ * ms * 1e-3)
                    return {"bound": "valu-issue", "kernel": kernel, "instructions_per_launch": float(insts),
                            "achieved": rate, "unit": "VALU instructions per SIMD per cycle",
                            "peak_int": VALU_PEAK_INT, "peak_fp32": VALU_PEAK_FP32,
                            "frac": rate / VALU_PEAK_INT, "lane_efficiency": float(lanes),
                            "source": "profiles/cone256_valu_pmc.md (rocprofv3 --pmc SQ_INSTS_VALU ...)"}
    except OSError:
        pass
    return None


def time_cgls(P, b, u0, iters, **kw):
    """Seconds per CGLS iteration (the init's normal-operator pass excluded)."""
    import torch
    from framelab_amd.solvers import cgls
    cgls(P, b, u0, 0.0, 1, log=False, **kw)                  # warm: workspace + graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cgls(P, b, u0, 0.0, 0, log=False, **kw)
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    u, _, _ = cgls(P, b, u0, 0.0, iters, log=False, **kw)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    return (t - t_init) / iters, u


def cgls_leg(P, x, iters):
    """CGLS iterations/s at 256^3 (graph-replayed iterations, FP64 dots), each
    path against the FP64 GPU solve (FP64 vectors over the FP64 projector
    pair) after the same number of iterations on the same noiseless data:
      fp32_plain    -- FP32 storage, Siddon + exact adjoint (the reference's pair);
      fp16_exact / bf16_exact -- the same operators with FP16 / BF16 storage of
                       p and A p (the precision change alone);
      fp32_matched  -- FP32 storage with the matched adjoint (the operator
                       change alone);
      mixed         -- FP16 storage and the matched adjoint.
    precision_only_speedup and operator_swap_speedup separate the two
    sources of mixed_vs_fp32_speedup.  The mixed iteration's kernel budget
    is the plain iteration's time over 3 (the 3x target)."""
    import torch
    from framelab_amd.projector import ADJ_MATCHED
    from framelab_amd.solvers import cgls_f64
    x64 = x.double()
    b64 = P.adjoint_f64(P.forward_f64(x64))
    cgls_f64(P, b64, torch.zeros_like(x64), iters=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u64, g64, _ = cgls_f64(P, b64, torch.zeros_like(x64), iters=iters)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    ref = u64.float()
    del b64, x64
    b_exact = P.adjoint(P.forward(x))
    b_matched = P.adjoint(P.forward(x), kind=ADJ_MATCHED)
    u0 = torch.zeros_like(x)
    legs = {
        "fp32_plain": (b_exact, {}),
        "fp16_exact": (b_exact, {"storage": torch.float16}),
        "bf16_exact": (b_exact, {"storage": torch.bfloat16}),
        "fp32_matched": (b_matched, {"matched": True}),
        "mixed": (b_matched, {"matched": True, "storage": torch.float16}),
    }
    out = {"iters": iters, "fp64_iters_per_s": iters / t64,
           "fp64_vs_phantom_rel_l2": float((ref - x).norm() / x.norm())}
    for name, (b, kw) in legs.items():
        s, u = time_cgls(P, b, u0, iters, graph=True, **kw)
        out[name] = {"iters_per_s": 1.0 / s, "rel_l2_vs_fp64": float((u - ref).norm() / ref.norm())}
        log(f"bench: cgls {name}: {1.0 / s:.2f} it/s, rel L2 vs FP64 {out[name]['rel_l2_vs_fp64']:.2e}")
    plain = out["fp32_plain"]["iters_per_s"]
    out["precision_only_speedup"] = out["fp16_exact"]["iters_per_s"] / plain
    out["operator_swap_speedup"] = out["fp32_matched"]["iters_per_s"] / plain
    out["mixed_vs_fp32_speedup"] = out["mixed"]["iters_per_s"] / plain
    out["mixed_rel_l2_vs_fp64"] = out["mixed"]["rel_l2_vs_fp64"]
    # one mixed iteration: its forward and adjoint kernels timed alone, the
    # rest (vector updates, reductions) by difference
    xl = x.to(torch.float16)
    yl = torch.empty(P.sino_shape, dtype=torch.float16, device=x.device)
    xo = torch.empty_like(x)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    P.forward(xl, out=yl)
    P.adjoint(yl, out=xo, kind=ADJ_MATCHED)
    ev[0].record()
    P.forward(xl, out=yl)
    ev[1].record()
    P.adjoint(yl, out=xo, kind=ADJ_MATCHED)
    ev[2].record()
    torch.cuda.synchronize()
    t_it = 1000.0 / out["mixed"]["iters_per_s"]
    tf, ta = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    out["mixed_iteration_ms"] = {"total": t_it, "forward_fp16": tf, "matched_adjoint_fp16": ta,
                                 "vector_updates_and_reductions": t_it - tf - ta,
                                 "budget_for_3x": 1000.0 / (3.0 * plain)}
    out["paths"] = {"fp32_plain": "FP32 storage, Siddon + exact adjoint", "mixed": "FP16 storage of p and A p, FP32 "
                    "accumulation, Siddon + matched adjoint", "reductions": "FP64 on the GPU, graph-replayed iterations"}
    return out


def precision_leg(P, x, outer=30, inner=3, budget_s=90.0):
    """Split Bregman with the 27-band 3-D linear frame (mu 1, lam 1e-3,
    `outer` x `inner`, graph-replayed CGLS with the matched adjoint, FP64
    reductions) with the projector input and sinogram stored in BF16, in
    FP16 and in FP32, each iterate against the FP64 solve's (FP64 pair and
    FP64 framelets) after every outer step, on the 256^3 / 360-view
    geometry (an FP64 solve at 512^3 / 720 views takes minutes).  The FP64
    solve runs as one call; `outer` is cut so that its estimate from one
    timed FP64 projector pair stays within `budget_s`."""
    import torch
    from framelab_amd.solvers import split_bregman, split_bregman_f64
    t0 = time.perf_counter()
    f64 = P.forward_f64(x.double())
    P.adjoint_f64(f64)
    torch.cuda.synchronize()
    pair = time.perf_counter() - t0
    per_outer = (inner + 1) * pair
    fit = max(1, min(outer, int(budget_s / per_outer)))
    log(f"bench: FP64 projector pair {pair * 1e3:.0f} ms; FP64 split Bregman {fit} outer x {inner} inner "
        f"(about {fit * per_outer:.0f} s)")
    outer = fit
    f32 = f64.float()
    u0 = torch.zeros_like(x)
    out = {"workload": "cone 256^3, 360 views, 512x384; split Bregman, 27-band linear frame, mu=1, lam=1e-3, "
                       f"{outer} outer x {inner} inner, matched adjoint; rel L2 per outer iterate vs FP64"}
    t0 = time.perf_counter()
    _, _, ref = split_bregman_f64(P, "linear", f64, u0.double(), 1e-3, 1.0, outer, inner, history=True)
    torch.cuda.synchronize()
    out["fp64_seconds"] = time.perf_counter() - t0
    out["fp64_pair_seconds"] = pair
    log(f"bench: FP64 split Bregman {out['fp64_seconds']:.1f} s")
    rnorm = ref.flatten(1).norm(dim=1)
    for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16), ("fp16", torch.float16)):
        t0 = time.perf_counter()
        _, _, h = split_bregman(P, "linear", f32, u0, 1e-3, 1.0, outer, inner, graph=True, matched=True, storage=dt,
                                history=True)
        torch.cuda.synchronize()
        out[f"{name}_seconds"] = time.perf_counter() - t0
        errs = ((h.double() - ref).flatten(1).norm(dim=1) / rnorm).tolist()
        out[f"{name}_rel_l2_vs_fp64"] = errs
        log(f"bench: split Bregman {name} storage vs FP64: outer 1 {errs[0]:.2e}, outer {outer} {errs[-1]:.2e}, "
            f"max {max(errs):.2e}")
        del h
    return out


def batch_problems(count=64, seed=5):
    """64 independent fan problems of seeded uneven sizes: image side from
    {128,192,256,384,512}, views U[90,720], bins U[256,1024], 6-15 ellipses,
    source-to-centre 4, source-to-detector 8, full fan."""
    import numpy as np
    from framelab_amd import geometry as G
    from framelab_amd.phantoms import ellipses_2d
    rng = np.random.default_rng(seed)
    paras, xs = [], []
    for i in range(count):
        N = int(rng.choice([128, 192, 256, 384, 512]))
        nv = int(rng.integers(90, 721))
        nd = int(rng.integers(256, 1025))
        paras.append(G.fan_para(nd, nv, N=(N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width()))
        xs.append(ellipses_2d(N, int(rng.integers(6, 16)), seed=seed * 1000 + i).astype(np.float32))
    return paras, xs


def bench_batch(dev, cgls_iters=30, outer=5, inner=3):
    """Batched leg: aggregate proj+backproj GUPS of one persistent launch per
    operator against a loop of single-problem calls; batched CGLS (graph
    replay, FP64 per-problem reductions) iterations/s against the loop, with
    each problem's relative L2 difference to its single-plan reconstruction
    and to the phantom; split Bregman with the 9-band linear framelet."""
    import numpy as np
    import torch
    from framelab_amd.batch import ProjectorBatch
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls

    paras, xs = batch_problems()
    B = ProjectorBatch(paras)
    Ps = [Projector(p) for p in paras]
    x = B.pack([torch.from_numpy(v) for v in xs]).to(dev)
    y = torch.empty(B.n_sino, device=dev)
    xb = torch.empty(B.n_vol, device=dev)
    upd = float(sum(p.nv * p.nx * p.ny for p in paras))

    def t_batched(reps=3):
        B.forward(x, out=y)
        B.adjoint(y, out=xb)
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e[0].record()
        for _ in range(reps):
            B.forward(x, out=y)
            B.adjoint(y, out=xb)
        e[1].record()
        torch.cuda.synchronize()
        return e[0].elapsed_time(e[1]) / reps

    t_default = t_batched()
    tuned = B.autotune()
    t_tuned = t_batched()
    xi = [x[B.vol_offsets[i]:B.vol_offsets[i + 1]].view(B.vol_shapes[i]) for i in range(len(paras))]
    yi = [torch.empty(B.sino_shapes[i], device=dev) for i in range(len(paras))]
    for i, P in enumerate(Ps):
        P.adjoint(P.forward(xi[i], out=yi[i]))
    e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    e[0].record()
    for i, P in enumerate(Ps):
        P.adjoint(P.forward(xi[i], out=yi[i]))
    e[1].record()
    torch.cuda.synchronize()
    t_loop = e[0].elapsed_time(e[1])

    # CGLS on the noiseless normal equations, all problems in lock step
    f = B.forward(x)
    b = B.adjoint(f)
    u0 = torch.zeros_like(b)
    B.cgls(b, u0, iters=1, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.cgls(b, u0, iters=0, graph=True)
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    u, glog, done = B.cgls(b, u0, iters=cgls_iters, graph=True)
    torch.cuda.synchronize()
    s_batch = (time.perf_counter() - t0 - t_init) / cgls_iters
    us = B.unpack(u)
    bs = B.unpack(b)
    rel_single, rel_phantom = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    singles = [cgls(P, bs[i], torch.zeros_like(bs[i]), 0.0, cgls_iters, log=False, graph=True)[0]
               for i, P in enumerate(Ps)]
    torch.cuda.synchronize()
    s_loop = (time.perf_counter() - t0) / cgls_iters
    for i in range(len(paras)):
        rel_single.append(float((us[i] - singles[i]).norm() / singles[i].norm()))
        rel_phantom.append(float((us[i] - xi[i]).norm() / xi[i].norm()))

    # split Bregman, 9-band linear framelet
    B.split_bregman(f, u0, frame="linear", lam=1e-3, mu=1.0, outer=1, inner=inner, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ub, res = B.split_bregman(f, u0, frame="linear", lam=1e-3, mu=1.0, outer=outer, inner=inner, graph=True)
    torch.cuda.synchronize()
    t_sb = time.perf_counter() - t0
    ubs = B.unpack(ub)
    rel_sb = [float((ubs[i] - xi[i]).norm() / xi[i].norm()) for i in range(len(paras))]
    sides = [p.nx for p in paras]
    return {
        "workload": "64 fan problems, side in {128..512}, views U[90,720], bins U[256,1024], SO=4 SD=8",
        "problems": len(paras), "sides_hist": {str(k): sides.count(k) for k in sorted(set(sides))},
        "updates_per_operator": upd,
        "proj_backproj_gups_batched": 2 * upd / (t_tuned * 1e-3) / 1e9,
        "proj_backproj_gups_batched_untuned": 2 * upd / (t_default * 1e-3) / 1e9,
        "proj_backproj_gups_single_loop": 2 * upd / (t_loop * 1e-3) / 1e9,
        "autotuned": tuned,
        "cgls_iters": cgls_iters,
        "cgls_iters_per_s_batched": 1.0 / s_batch,
        "cgls_iters_per_s_single_loop": 1.0 / s_loop,
        "cgls_rel_l2_vs_single_max": max(rel_single),
        "cgls_rel_l2_vs_phantom_median": float(np.median(rel_phantom)),
        "cgls_rel_l2_vs_phantom_max": max(rel_phantom),
        "split_bregman": {"outer": outer, "inner": inner, "frame": "linear (9 bands)", "lam": 1e-3, "mu": 1.0,
                          "seconds": t_sb, "rel_l2_vs_phantom_median": float(np.median(rel_sb)),
                          "final_resid_max": float(res[:, -1].max())},
    }


def framelet_leg(x, nvox, reps=3):
    """Event-timed 3-D framelet analysis (W u) and synthesis (W^T alpha) of
    the volume x, with achieved bandwidth against the algorithmic bytes."""
    import torch
    from framelab_amd.framelets import FrameletSystem
    W = FrameletSystem(dim=3, type="linear", level=1)
    out = {"frame": "linear B-spline, 3-D, 27 bands, 1 level", "volume": list(x.shape)}
    for cdt, name, sb in ((torch.float32, "f32", 4), (torch.bfloat16, "bf16", 2)):
        a = W.forwardTransform(x, coef_dtype=cdt)
        u = W.adjointTransform(a)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record()
        for _ in range(reps):
            W.forwardTransform(x, coef_dtype=cdt)
        ev[1].record()
        for _ in range(reps):
            u = W.adjointTransform(a)
        ev[2].record()
        torch.cuda.synchronize()
        rel = float((u - x).norm() / x.norm())   # W^T W = I
        nbytes = nvox * (4 + W.nbands * sb)
        for leg, t in (("analysis", ev[0].elapsed_time(ev[1]) / reps),
                       ("synthesis", ev[1].elapsed_time(ev[2]) / reps)):
            gbs = nbytes / (t * 1e-3) / 1e9
            out[f"{name}_{leg}_ms"] = t
            out[f"{name}_{leg}_gbs"] = gbs
            out[f"{name}_{leg}_frac_hbm"] = gbs / HBM_PEAK_GBS
        out[f"{name}_perfect_reconstruction_rel"] = rel
        del a
    log(f"bench: framelet analysis {out['f32_analysis_gbs']:.0f} GB/s, synthesis "
        f"{out['f32_synthesis_gbs']:.0f} GB/s (FP32 coefficients)")
    return out


def c4_leg():
    """The largest configuration, 512^3 with 720 views and a 1024x768 panel,
    in BF16 storage (FP32 arithmetic): one Siddon forward, one exact and one
    matched adjoint, event-timed after a warm-up of each."""
    import torch
    from framelab_amd import geometry as G
    from framelab_amd.projector import ADJ_EXACT, ADJ_MATCHED, Projector
    from framelab_amd.phantoms import ellipsoids_3d_device
    dev = torch.device("cuda:0")
    N, nv, na, nb = 512, 720, 1024, 768
    p = G.cone_para(na, nb, nv, N=(N, N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width(), Lz=G.full_cone_height())
    P = Projector(p)
    x = ellipsoids_3d_device(N, count=20, seed=2, device=dev).to(torch.bfloat16)
    y = torch.empty(P.sino_shape, dtype=torch.bfloat16, device=dev)
    xo = torch.empty(P.vol_shape, dtype=torch.bfloat16, device=dev)
    P.forward(x, out=y)
    P.adjoint(y, out=xo, kind=ADJ_EXACT)
    P.adjoint(y, out=xo, kind=ADJ_MATCHED)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    P.forward(x, out=y)
    ev[1].record()
    P.adjoint(y, out=xo, kind=ADJ_EXACT)
    ev[2].record()
    P.adjoint(y, out=xo, kind=ADJ_MATCHED)
    ev[3].record()
    torch.cuda.synchronize()
    tA, tE, tM = (ev[k].elapsed_time(ev[k + 1]) for k in range(3))
    upd = float(N) ** 3 * nv
    out = {"workload": "cone-beam 512^3, 720 views, 1024x768 panel, BF16 storage, FP32 arithmetic",
           "t_A_ms": tA, "t_At_exact_ms": tE, "t_At_matched_ms": tM,
           "proj_backproj_gups_exact": 2 * upd / ((tA + tE) * 1e-3) / 1e9,
           "proj_backproj_gups_matched": 2 * upd / ((tA + tM) * 1e-3) / 1e9}
    log(f"bench: 512^3 BF16: A {tA:.1f} ms, At exact {tE:.1f} ms, matched {tM:.1f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cgls-iters", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-cgls", action="store_true", help="skip the CGLS leg")
    ap.add_argument("--cpu-views", type=int, default=5)
    ap.add_argument("--no-c1", action="store_true", help="skip the 2-D fan CPU/GPU leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 2D leg")
    ap.add_argument("--no-c4", action="store_true", help="skip the 512^3 BF16 leg")
    ap.add_argument("--no-precision", action="store_true", help="skip the split-Bregman precision leg")
    ap.add_argument("--profile", action="store_true",
                    help="counter run: one forward, one exact and one matched adjoint at 256^3, one "
                         "framelet analysis and synthesis (27 bands, FP32), then a "
                         "plain 1 GiB device copy (the calibration kernel for HBM byte counters) and the "
                         "FMA-chain kernel (the calibration for VALU instruction counters); no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU sharding is "
              "described, not run, here", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from framelab_amd import geometry as G
    from framelab_amd.phantoms import ellipsoids_3d
    from framelab_amd.projector import ADJ_EXACT, ADJ_MATCHED, Projector

    dev = torch.device("cuda:0")
    N, nv, na, nb = 256, 360, 512, 384
    p = G.cone_para(na, nb, nv, N=(N, N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                    Lz=G.full_cone_height())
    log("bench: building phantom")
    Xh = ellipsoids_3d(N, count=20, seed=2)
    P = Projector(p)
    x = torch.from_numpy(Xh).to(dev)
    y = torch.empty(P.sino_shape, dtype=torch.float32, device=dev)
    xb = torch.empty(P.vol_shape, dtype=torch.float32, device=dev)
    nvox = N ** 3
    ndet = na * nb

    def step(kind=ADJ_EXACT, xin=x, yb=y, xo=xb, evs=None):
        if evs is not None:
            evs[0].record()
        P.forward(xin, out=yb)
        if evs is not None:
            evs[1].record()
        P.adjoint(yb, out=xo, kind=kind)
        if evs is not None:
            evs[2].record()

    if args.profile:
        P.forward(x, out=y)
        P.adjoint(y, out=xb, kind=ADJ_EXACT)
        P.adjoint(y, out=xb, kind=ADJ_MATCHED)
        from framelab_amd.framelets import FrameletSystem
        W3 = FrameletSystem(dim=3, type="linear", level=1)
        W3.adjointTransform(W3.forwardTransform(x))
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        # VALU calibration: 8192 workgroups of FP32 FMA chains, then of integer chains
        import ctypes as C
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from diag import diag
        out = torch.empty(8192 * 256, dtype=torch.float32, device=dev)
        for kind in (0, 1):
            diag.check(diag.lib().fl_calibrate_valu(C.c_void_p(out.data_ptr()), out.numel(), 5000, kind,
                                                    C.c_void_p(torch.cuda.current_stream().cuda_stream)),
                       "fl_calibrate_valu")
        torch.cuda.synchronize()
        log("bench: profile run done (forward, exact adjoint, matched adjoint, framelet analysis and "
            "synthesis, 1 GiB copy, VALU calibration)")
        return

    log("bench: warmup")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs=evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tA = [e[0].elapsed_time(e[1]) for e in evs]
    tAt = [e[1].elapsed_time(e[2]) for e in evs]
    mA, mAt = float(np.median(tA)), float(np.median(tAt))
    gups = 2.0 * nvox * nv / ((mA + mAt) * 1e-3) / 1e9
    log(f"bench: A {mA:.2f} ms, At(exact) {mAt:.2f} ms, wall/step {wall / args.steps * 1e3:.2f} ms")

    # per-operator algorithmic bytes (per-view streaming model): nv*(Nvox*sx + Ndet*sy)
    B_op = nv * (nvox * 4 + ndet * 4)
    dom_ms, dom_name = ((mAt, "k_adj_cone_zc<float,float,4>") if mAt >= mA
                        else (mA, "k_fwd_cone<float,float,16,16,8,4,4,4>"))
    achieved = B_op / (dom_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom_name), "kernel": dom_name,
            "algorithmic_bytes_per_launch": B_op, "kernel_ms": dom_ms,
            "traffic_source": "profiles/cone256_hbm_traffic_pmc.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"}

    valu = valu_issue(dom_name, dom_ms)

    # matched adjoint and low-precision storage variants (reported beside)
    extra = {"t_A_ms": mA, "t_At_exact_ms": mAt}
    evm = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    P.adjoint(y, out=xb, kind=ADJ_MATCHED)
    evm[0].record()
    P.adjoint(y, out=xb, kind=ADJ_MATCHED)
    evm[1].record()
    torch.cuda.synchronize()
    extra["t_At_matched_ms"] = evm[0].elapsed_time(evm[1])
    for dt, name in ((torch.bfloat16, "bf16"), (torch.float16, "fp16")):
        xl = x.to(dt)
        yl = torch.empty(P.sino_shape, dtype=dt, device=dev)
        xo = torch.empty(P.vol_shape, dtype=dt, device=dev)
        step(xin=xl, yb=yl, xo=xo)
        e3 = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        step(xin=xl, yb=yl, xo=xo, evs=e3, kind=ADJ_MATCHED)
        torch.cuda.synchronize()
        a, b = e3[0].elapsed_time(e3[1]), e3[1].elapsed_time(e3[2])
        extra[f"{name}_t_A_ms"] = a
        extra[f"{name}_t_At_matched_ms"] = b
        extra[f"{name}_matched_gups"] = 2.0 * nvox * nv / ((a + b) * 1e-3) / 1e9

    # framelet analysis / synthesis on the same volume (3-D linear B-spline
    # frame: 27 bands, one level), FP32 and BF16 coefficient storage.
    # Algorithmic bytes per voxel: s_x + 27 s_b (read the volume, write every
    # band; synthesis the reverse).
    framelet = framelet_leg(x, nvox)

    result = {
        "metric": METRIC, "value": gups, "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": mA + mAt, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cone-beam 256^3, 360 views over 2pi, 512x384 panel, SO=4 SD=8; "
                               "one Siddon forward + one exact voxel-driven adjoint (FP32 plain)",
                   "volume": [N, N, N], "views": nv, "detector": [na, nb],
                   "phantom": "20 seeded ellipsoids (seed 2)"},
        "roofline": roof,
        "valu": valu,
        "extra": extra,
        "framelet": framelet,
    }

    if not args.no_cgls:
        log("bench: cgls")
        result["cgls"] = cgls_leg(P, x, args.cgls_iters)

    if not args.no_precision:
        log("bench: precision (split Bregman vs FP64)")
        result["precision"] = precision_leg(P, x)

    if not args.no_batch:
        log("bench: batch")
        result["batch"] = bench_batch(dev)
        log(f"bench: batch {result['batch']['proj_backproj_gups_batched']:.1f} GUPS "
            f"(loop {result['batch']['proj_backproj_gups_single_loop']:.1f}), cgls "
            f"{result['batch']['cgls_iters_per_s_batched']:.2f} it/s (loop "
            f"{result['batch']['cgls_iters_per_s_single_loop']:.2f})")

    if not args.no_c4:
        log("bench: 512^3 BF16")
        result["c4"] = c4_leg()

    if not args.no_cpu:
        log("bench: cpu baseline")
        result["cpu_baseline"] = cpu_baseline(p, Xh, views=args.cpu_views)
        if not args.no_c1:
            log("bench: 2-D fan on CPU and GPU")
            result["c1"] = cpu_c1(dev)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
