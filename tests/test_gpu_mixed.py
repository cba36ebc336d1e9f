"""GPU tests of the mixed-precision dynamics (FP32 prognostics, fluxes and
tendencies; FP64 reference state, pressure terms, vertical solve and
budgets), against the FP64 GPU path, which itself matches the oracle.

Tolerances: potential temperature within 2e-3 K of the FP64 run after 60
steps of the 40x40x80 warm bubble (bench.py --theta-check measures the
512x512x128 shear workload); every prognostic within 1e-3 of its FP64
counterpart relative to the field's range, with the error of each printed;
total mass conserved to 1e-6 relative.  The FP32 storage of RHOT bounds the
accuracy: its rounding (about 2e-5 at RHOT ~ 360) enters the pressure
perturbation RT2P * (RHOT - REF_rhot), which is evaluated in FP64 but from
the FP32 state, as in the reference's SINGLE build; over 60 s that gives
MOMZ errors of O(1e-4) of its range."""
import numpy as np
import pytest

from scale_amd import hip, inputs

pytestmark = pytest.mark.gpu

FIELDS = ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")


def interior(g, a):
    return a[g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]


def run(case, steps, mixed):
    ctx = hip.model_from_case(case, mixed_precision=mixed)
    try:
        m0 = ctx.model_mass("DENS")
        ctx.model_large_step(steps)
        st = {f: ctx.model_get(f) for f in FIELDS}
        m1 = ctx.model_mass("DENS")
    finally:
        ctx.close()
    return st, m0, m1


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_mixed_precision_warm_bubble(scheme):
    print(f"[mixed] warm bubble 40x40x80, 60 steps, scheme {scheme}", flush=True)
    case = inputs.warm_bubble_40x40x80()
    case = inputs.Case(**{**case.__dict__, "scheme": scheme, "qtrc": case.qtrc[..., :0]})
    g = case.grid
    ref, _, _ = run(case, 60, False)
    got, m0, m1 = run(case, 60, True)
    th_ref = interior(g, ref["RHOT"]) / interior(g, ref["DENS"])
    th_got = interior(g, got["RHOT"]) / interior(g, got["DENS"])
    dth = np.abs(th_got - th_ref).max()
    print(f"  max |theta_mixed - theta_fp64| = {dth:.3e} K", flush=True)
    assert dth <= 2e-3, dth
    for f in FIELDS:
        a, b = interior(g, ref[f]), interior(g, got[f])
        rng = max(a.max() - a.min(), np.abs(a).max() * 1e-3)
        err = np.abs(a - b).max() / rng
        print(f"  {f}: max |mixed - fp64| / range = {err:.3e}", flush=True)
        assert err <= 1e-3, f
    assert abs(m1 - m0) <= 1e-6 * abs(m0), (m0, m1)


def test_mixed_precision_rejects_terrain_and_tracers():
    case = inputs.warm_bubble_40x40x80()
    ctx = hip.Context(case.grid, case.scheme, case.scheme_tracer)
    try:
        ctx.set_metrics(case.metric)
        with pytest.raises(hip.ScaleHipError):
            ctx.model_create(case.dt, case.dt, 1, mixed_precision=True)
    finally:
        ctx.close()
