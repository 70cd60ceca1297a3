#!/usr/bin/env python3
"""Write tests/golden/golden_projectors.npz from the reference's own CPU kernels.

Runs in the build container only (it needs oracle/_ref, compiled by
oracle/Makefile from the reference checkout).  Inputs are regenerated from the
stored seeds and geometry parameters; only inputs' seeds, geometry and the
reference outputs are stored (data, not code).

Cases:
  sort_alpha   merges of ascending lists with engineered cross-list ties
  fan_siddon   Ax_fan_mf_cpu_siddon on 3 small fan geometries
  fan_gao      Ax_fan_mf_cpu_new (Gao column walk)
  cone_siddon  Ax_cone_mf_cpu_siddon (circle, helix, non-unit dz)
  cone_adj     Atx_cone_mf_cpu_new (dz=1, circular)
  find_l_3d    find_l_3d on random and special arguments, incl. the KATs of
               matlab/mex/src2/test_find_l_3d.cpp
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from cases import cone_cases, fan_cases, find_l_3d_args, merge_cases  # noqa: E402


def main():
    if not O.has_ref():
        raise SystemExit("oracle/_ref not built: run `make -C oracle` where the reference checkout exists")
    out = {}
    for k, (a, b) in enumerate(merge_cases()):
        out[f"merge_{k}"] = O.ref_sort_alpha(a, b)
    for k, (p, X) in enumerate(fan_cases()):
        out[f"fan_siddon_{k}"] = O.ref_siddon_fan(X, p)
        out[f"fan_gao_{k}"] = O.ref_gao_fan(X, p)
    for k, (p, X, Y, adj_ok) in enumerate(cone_cases()):
        out[f"cone_siddon_{k}"] = O.ref_siddon_cone(X, p)
        if adj_ok:
            out[f"cone_adj_{k}"] = O.ref_atx_cone_cpu(Y, p)
    args = find_l_3d_args()
    out["find_l_3d"] = np.array([O.ref().ref_find_l_3d(*map(float, a)) for a in args], np.float32)
    np.savez_compressed(os.path.join(HERE, "golden_projectors.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
