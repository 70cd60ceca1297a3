#!/usr/bin/env python3
"""Benchmark of the CT-operator hot path on one MI355X.

Workload (the 256^3 cone-beam configuration): 256^3 volume on [-1,1]^3 with
20 seeded random ellipsoids, circular orbit, 360 views over 2*pi, 512x384
flat panel, source-to-centre 4, source-to-detector 8.  One step is one
forward projection A (Siddon) plus one backprojection A^T (exact
voxel-driven adjoint) of the whole volume, inputs resident in HBM.

Prints ONE JSON line.  value = proj+backproj GUPS = 2*N_vox*n_views /
(t_A + t_At) / 1e9 over the timed steps (FP32 plain path).  Beside it:
  cgls      -- CGLS iterations/s of the FP32 plain path (Siddon + exact
               adjoint, FP32 storage) and of the mixed path (FP16 storage of
               the projector input and sinogram, FP32 accumulation, matched
               adjoint), both device-resident with FP64 reductions and
               graph-replayed iterations, and the mixed reconstruction's
               relative L2 difference to the FP32 one;
  roofline  -- the dominant kernel's algorithmic bytes per launch over its
               measured duration, against 8 TB/s;
  cpu_baseline -- the reference's own single-threaded CPU kernels on a
               bounded sample of views.
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


def cpu_baseline(p, X, views=3):
    """The reference's own single-threaded CPU kernels (oracle/_ref), or the
    oracle's restatement when _ref is absent, on `views` views of the same
    volume: Siddon forward + exact adjoint, reported as proj+backproj GUPS."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from framelab_amd import geometry as G
    sub = G.cone_para(p.na, p.nb, views, N=(p.nx, p.ny, p.nz), SO=4.0, OD=4.0,
                      Ly=G.full_fan_width(), Lz=G.full_cone_height(),
                      sd_phi=p.sd_phi[:: max(1, p.nv // views)][:views])
    kind = "reference" if O.has_ref() else "port"
    fwd = O.ref_siddon_cone if kind == "reference" else O.siddon_cone_f32
    adj = O.ref_atx_cone_cpu if kind == "reference" else O.atx_cone_f32
    if kind == "port":
        os.environ["OMP_NUM_THREADS"] = "1"
    t0 = time.perf_counter()
    y = fwd(X, sub)
    t1 = time.perf_counter()
    adj(y, sub)
    t2 = time.perf_counter()
    nvox = p.nx * p.ny * p.nz
    gups = 2.0 * nvox * views / ((t1 - t0) + (t2 - t1)) / 1e9
    model, cores = cpu_info()
    return {"value": gups, "unit": "GUPS", "cores": 1, "kind": kind,
            "sample": f"{views} of {p.nv} views at full 256^3 / 512x384 size, one forward + one adjoint "
                      f"(t_A={t1 - t0:.2f}s, t_At={t2 - t1:.2f}s); host: {model}, {cores} logical CPUs"}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cgls-iters", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-cgls", action="store_true", help="skip the CGLS leg")
    ap.add_argument("--cpu-views", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU sharding is "
              "described, not run, in this build", file=sys.stderr)
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
    mA, mAt = float(np.mean(tA)), float(np.mean(tAt))
    gups = 2.0 * nvox * nv / ((mA + mAt) * 1e-3) / 1e9
    log(f"bench: A {mA:.2f} ms, At(exact) {mAt:.2f} ms, wall/step {wall / args.steps * 1e3:.2f} ms")

    # per-operator algorithmic bytes (per-view streaming model): nv*(Nvox*sx + Ndet*sy)
    B_op = nv * (nvox * 4 + ndet * 4)
    dom_ms, dom_name = (mAt, "k_adj_cone_fp<float,float>") if mAt >= mA else (mA, "k_fwd_cone<float,float,64,4>")
    achieved = B_op / (dom_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom_name,
            "algorithmic_bytes_per_launch": B_op, "kernel_ms": dom_ms}

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

    result = {
        "metric": METRIC, "value": gups, "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": mA + mAt, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cone-beam 256^3, 360 views over 2pi, 512x384 panel, SO=4 SD=8; "
                               "one Siddon forward + one exact voxel-driven adjoint (FP32 plain)",
                   "volume": [N, N, N], "views": nv, "detector": [na, nb],
                   "phantom": "20 seeded ellipsoids (seed 2)"},
        "roofline": roof,
        "extra": extra,
    }

    if not args.no_cgls:
        log("bench: cgls")
        b_rhs = P.adjoint(P.forward(x))
        u0 = torch.zeros_like(x)
        it = args.cgls_iters
        s32, u32 = time_cgls(P, b_rhs, u0, it, graph=True)
        bm = P.adjoint(P.forward(x), kind=ADJ_MATCHED)
        smx, umx = time_cgls(P, bm, u0, it, graph=True, matched=True, storage=torch.float16)
        s16, u16 = time_cgls(P, b_rhs, u0, it, graph=True, storage=torch.float16)
        rel = float((umx - u32).norm() / u32.norm())
        rel16 = float((u16 - u32).norm() / u32.norm())
        result["cgls"] = {
            "iters": it,
            "fp32_plain_iters_per_s": 1.0 / s32,
            "mixed_iters_per_s": 1.0 / smx,
            "mixed_vs_fp32_speedup": s32 / smx,
            "mixed_rel_l2_vs_fp32": rel,
            "fp16_exact_adjoint_iters_per_s": 1.0 / s16,
            "fp16_exact_adjoint_rel_l2_vs_fp32": rel16,
            "fp32_plain": "FP32 storage, Siddon + exact adjoint, FP64 dots, graph-replayed",
            "mixed": "FP16 storage of p and A p, FP32 accumulation, Siddon + matched adjoint, FP64 dots, "
                     "graph-replayed",
        }
        log(f"bench: cgls fp32 {1 / s32:.2f} it/s, mixed {1 / smx:.2f} it/s, rel {rel:.2e}")

    if not args.no_cpu:
        log("bench: cpu baseline")
        result["cpu_baseline"] = cpu_baseline(p, Xh, views=args.cpu_views)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
