"""GPU tests of the mixed-precision dynamics (FP32 prognostics, fluxes and
tendencies; FP64 reference state, pressure terms, vertical solve and
budgets), against the FP64 GPU path, which itself matches the oracle.

DENS and RHOT are stored in FP32 as deviations from per-level profiles, so
their rounding scales with the deviation rather than with RHOT ~ 360.
Tolerances: potential temperature within 2e-4 K of the FP64 run after 60
steps of the 40x40x80 warm bubble (measured 2.5e-5 K with UD3, 5.0e-5 K
with CD4; bench.py --theta-check measures the 512x512x128 shear workload
over 600 steps); every prognostic within 5e-5 of its FP64 counterpart
relative to the field's range (measured at most 6e-6), with the error of
each printed; total mass conserved to 1e-6 relative."""
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
    assert dth <= 2e-4, dth
    for f in FIELDS:
        a, b = interior(g, ref[f]), interior(g, got[f])
        rng = max(a.max() - a.min(), np.abs(a).max() * 1e-3)
        err = np.abs(a - b).max() / rng
        print(f"  {f}: max |mixed - fp64| / range = {err:.3e}", flush=True)
        assert err <= 5e-5, f
    assert abs(m1 - m0) <= 1e-6 * abs(m0), (m0, m1)


def test_mixed_precision_with_fct_tracers():
    """The stretched shear workload at 48x40x32 with 3 FCT tracers, 40 large
    steps in mixed precision against FP64: the tracers advance in FP64 on the
    FP64 copies of the FP32 dynamics, so each tracer's mass is conserved to
    1e-13 relative, and each tracer stays within 2e-4 of its range of the FP64
    run (the difference comes from the FP32 mass fluxes and density)."""
    print("[mixed] shear 48x40x32, 3 FCT tracers, 40 steps", flush=True)
    case = inputs.shear_512x512x128(kmax=32, imax=48, jmax=40)
    g = case.grid
    out = {}
    for mixed in (False, True):
        ctx = hip.model_from_case(case, mixed_precision=mixed)
        try:
            q0 = [ctx.model_mass(hip.QTRC_BASE + iq, weight_dens=True) for iq in range(3)]
            ctx.model_large_step(40)
            q1 = [ctx.model_mass(hip.QTRC_BASE + iq, weight_dens=True) for iq in range(3)]
            for a, b in zip(q0, q1):
                assert abs(b - a) <= 1e-13 * abs(a), (mixed, a, b)
            out[mixed] = [ctx.model_get(hip.QTRC_BASE + iq) for iq in range(3)]
        finally:
            ctx.close()
    for iq in range(3):
        a, b = interior(g, out[False][iq]), interior(g, out[True][iq])
        rng = max(a.max() - a.min(), np.abs(a).max() * 1e-3)
        err = np.abs(a - b).max() / rng
        print(f"  QTRC{iq}: max |mixed - fp64| / range = {err:.3e}", flush=True)
        assert err <= 2e-4, iq


def test_mixed_precision_rejects_terrain():
    import helpers as H
    case = inputs.warm_bubble_40x40x80()
    case = inputs.Case(**{**case.__dict__, "metric": H.terrain_metrics(case, amp=0.02)})
    ctx = hip.Context(case.grid, case.scheme, case.scheme_tracer)
    try:
        ctx.set_metrics(case.metric)
        with pytest.raises(hip.ScaleHipError):
            ctx.model_create(case.dt, case.dt, 1, mixed_precision=True)
    finally:
        ctx.close()
