"""Writes the committed golden fixtures of tests/golden/ from the CPU oracle.

The reference (Fortran, MPI) cannot be built in this image and its own tests
hold no numeric vectors for this path, so these fixtures are the oracle's
results on seeded synthetic inputs, frozen once the oracle had passed its
closed-form and invariant checks (tests/test_oracle.py).  They guard the
oracle against regressions and let the GPU tests check long integrations
without re-running the oracle.

    python tests/golden/make_golden.py

warm_bubble_40x40x80.npz -- the 40x40x80 dry warm bubble (DX=DY=500 m,
  DZ=250 m, bubble +3 K r=2 km at z=2 km, u=10 m/s, theta noise 0.01 K,
  UD3, dt=1 s, one tracer QV=0), after 1, 10 and 60 large steps: per field
  the interior sum, L2 norm and max |x|, and after 60 steps the x-z slice
  through the domain centre.
stage_16x16x24.npz -- one HEVI stage (UD3 and CD4) on a 16x16x24 case:
  per output the sum, L2 norm and max |x| over the elements the stage
  defines.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

FIELDS = ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")
STAGE_OUTS = ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK", "mflx_hi", "tflx_hi")


def stats(a):
    a = a[~np.isnan(a)]
    return np.array([a.sum(), np.sqrt((a * a).sum()), np.abs(a).max()])


def interior(g, a):
    return a[g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]


def warm_bubble_fixture():
    import pyoracle as po
    from scale_amd import inputs
    case = inputs.warm_bubble_40x40x80()
    g = case.grid
    m = po.Model(case)
    out = {}
    done = 0
    for n in (1, 10, 60):
        m.step(n - done)
        done = n
        for f in FIELDS:
            out[f"{f}_{n}"] = stats(interior(g, m.state[f]))
        out[f"QTRC_{n}"] = stats(interior(g, m.qtrc[..., 0]))
    jc = (g.JS + g.JE) // 2
    for f in FIELDS:
        out[f"{f}_slice_60"] = np.ascontiguousarray(m.state[f][g.KS:g.KE + 1, g.IS:g.IE + 1, jc])
    return out


def stage_fixture():
    import helpers as H
    import pyoracle as po
    out = {}
    for scheme in (3, 4):
        case = H.small_case(kmax=24, imax=16, jmax=16, seed=2400 + scheme)
        inp = H.stage_inputs(case, seed=17)
        ref = po.tstep_short_hevi(case, inp, 0.5, scheme=scheme)
        for k in STAGE_OUTS:
            out[f"{k}_{scheme}"] = stats(ref[k])
    return out


def main():
    import pyoracle as po
    po.set_threads(min(os.cpu_count() or 1, 8))
    np.savez(os.path.join(HERE, "warm_bubble_40x40x80.npz"), **warm_bubble_fixture())
    np.savez(os.path.join(HERE, "stage_16x16x24.npz"), **stage_fixture())
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
