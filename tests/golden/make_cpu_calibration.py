"""Writes tests/golden/cpu_calibration.json: the speed of the reference's own
CPU kernels (oracle/_ref, compiled from the reference sources; this container
only) against the oracle's C restatement of the same algorithms, both on one
core of this host.  The GPU box times only the restatement (the reference
code does not travel there); this ratio converts its figure to the
reference's.  Median of 5 runs each:
  * fan Siddon forward, 256^2, 360 views x 512 bins (Ax_fan_mf_cpu_siddon);
  * cone Siddon forward, 256^3, one view of 512 x 384 (Ax_cone_mf_cpu_siddon);
  * cone exact adjoint, 256^3, one view (Atx_cone_mf_cpu_new).
The fan exact adjoint's reference source includes MATLAB's mex.h and is not
built here, so it has no ratio.

Run:  python tests/golden/make_cpu_calibration.py
"""
import json
import os
import platform
import statistics
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]


def median_time(fn, *args, runs=5):
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn(*args)
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def main():
    import numpy as np
    import oracle as O
    from framelab_amd import geometry as G
    from framelab_amd.phantoms import ellipses_2d, ellipsoids_3d
    if not O.has_ref():
        sys.exit("oracle/_ref is not built: run __graft_entry__.build() where the reference checkout exists")
    O.set_threads(1)
    fan = G.fan_para(512, 360, N=(256, 256), SO=4.0, OD=4.0, Ly=G.full_fan_width())
    Xf = ellipses_2d(256, 10, seed=1).astype(np.float32)
    cone = G.cone_para(512, 384, 1, N=(256, 256, 256), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                       Lz=G.full_cone_height(), sd_phi=np.array([0.3], np.float32))
    Xc = ellipsoids_3d(256, count=20, seed=2)
    Yc = O.siddon_cone_f32(Xc, cone)
    rows = {}
    for name, ref, port, args in (
            ("fan_siddon_forward_256sq_360x512", O.ref_siddon_fan, O.siddon_fan_f32, (Xf, fan)),
            ("cone_siddon_forward_256cube_1view", O.ref_siddon_cone, O.siddon_cone_f32, (Xc, cone)),
            ("cone_exact_adjoint_256cube_1view", O.ref_atx_cone_cpu, O.atx_cone_f32, (Yc, cone))):
        tr, tp = median_time(ref, *args), median_time(port, *args)
        rows[name] = {"reference_s": tr, "restatement_s": tp, "reference_over_restatement": tr / tp}
        print(f"{name}: reference {tr:.3f} s, restatement {tp:.3f} s, ratio {tr / tp:.3f}", flush=True)
    model = platform.processor()
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    out = {"host": model, "cores_used": 1, "compiler": "gcc -O2 -ffp-contract=off (both)", "runs": 5,
           "statistic": "median", "kernels": rows}
    with open(os.path.join(HERE, "cpu_calibration.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
