"""Ensembles (scale_hip_ensemble_*): members of different grid sizes stepped
together, the filter and the four HEVI stages of every member in one launch
each, optionally replayed from a captured hipGraph, with autotuned launch
shapes.  Every member must come out bit-identical to the same model stepped
alone (scale_hip_model_large_step), which itself matches the oracle; one
member is also checked against the oracle directly."""
import numpy as np
import pytest

import pyoracle as po
from scale_amd import hip, inputs

pytestmark = pytest.mark.gpu

FIELDS = ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")


def member_cases(nstep_dt=None):
    """Three members: 24x20x16, 40x36x20 and 70x20x24 (two i-tiles, the
    second partial), seeded theta noise and winds of their own, two FCT
    tracers each."""
    out = []
    for m, (im, jm, km) in enumerate([(24, 20, 16), (40, 36, 20), (70, 20, 24)]):
        out.append(inputs.make_case(kmax=km, imax=im, jmax=jm, dx=500.0, dy=500.0, dz=250.0, seed=7100 + m,
                                    noise=0.5, u0=5.0 + 4.0 * m, v0=-1.5 + m, ntracer=2,
                                    tracer_kind=["blob", "tophat"], fct=True, bubble_r=2000.0, bubble_zc=2000.0,
                                    momz_noise=0.05, dt=1.0))
    return out


def state_of(ctx, nq):
    st = {k: ctx.model_get(k) for k in FIELDS}
    for iq in range(nq):
        st[f"Q{iq}"] = ctx.model_get(hip.QTRC_BASE + iq)
    return st


def single_runs(cases, steps, **kw):
    res = []
    for case in cases:
        ctx = hip.model_from_case(case, **kw)
        try:
            ctx.model_large_step(steps)
            ctx.synchronize()
            res.append(state_of(ctx, case.qtrc.shape[-1]))
        finally:
            ctx.close()
    return res


def assert_identical(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), f"{what}: {k} differs (max |diff| {np.nanmax(np.abs(a[k] - b[k])):.3e})"


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_ensemble_members_match_single_runs(mixed):
    """3 direct ensemble steps, then 3 replays of the captured graph: each
    member bit-identical to its own model after 3 and after 6 large steps."""
    print(f"[ensemble] 3 members, mixed={mixed}: direct then graph", flush=True)
    cases = member_cases()
    ref3 = single_runs(cases, 3, mixed_precision=mixed)
    ref6 = single_runs(cases, 6, mixed_precision=mixed)
    ctxs = [hip.model_from_case(c, mixed_precision=mixed) for c in cases]
    try:
        with hip.Ensemble(ctxs) as ens:
            ens.large_step(3)
            ens.synchronize()
            for m, (c, ctx) in enumerate(zip(cases, ctxs)):
                assert_identical(state_of(ctx, c.qtrc.shape[-1]), ref3[m], f"member {m} after 3 steps")
            ens.set_graph(True)
            ens.large_step(3)
            ens.synchronize()
            for m, (c, ctx) in enumerate(zip(cases, ctxs)):
                assert_identical(state_of(ctx, c.qtrc.shape[-1]), ref6[m], f"member {m} after 3 + 3 graph steps")
    finally:
        for ctx in ctxs:
            ctx.close()


def test_ensemble_shapes_and_tuner():
    """Every launch shape gives the same bits; the tuner picks one of them,
    leaves the members' states as it found them, and the steps after it
    still match the single runs."""
    print("[ensemble] shapes and tuner", flush=True)
    cases = member_cases()
    ref2 = single_runs(cases, 2)
    ctxs = [hip.model_from_case(c) for c in cases]
    try:
        with hip.Ensemble(ctxs) as ens:
            before = [state_of(ctx, c.qtrc.shape[-1]) for c, ctx in zip(cases, ctxs)]
            best, report = ens.tune(steps=1)
            print(report, flush=True)
            assert best > 0.0 and "chosen" in report
            for m, (c, ctx) in enumerate(zip(cases, ctxs)):
                assert_identical(state_of(ctx, c.qtrc.shape[-1]), before[m], f"member {m} after tuning")
            sb, nb, tj = ens.get_shape()
            assert sb in (0, 1, 2) and nb in (0, 2) and tj in (4, 8)
            ens.set_shape(1, 2, 4)
            ens.large_step(1)
            ens.set_shape(2, 0, 8)
            ens.large_step(1)
            ens.synchronize()
            for m, (c, ctx) in enumerate(zip(cases, ctxs)):
                assert_identical(state_of(ctx, c.qtrc.shape[-1]), ref2[m], f"member {m} with changed shapes")
    finally:
        for ctx in ctxs:
            ctx.close()


def test_ensemble_member_matches_oracle_with_substeps():
    """dtl = 2 dts (two short steps per large step): the smallest member of
    an ensemble against the oracle after 3 large steps, 1e-12 relative."""
    print("[ensemble] nstep = 2 against the oracle", flush=True)
    cases = member_cases()
    ctxs = [hip.model_from_case(c, dtl=2.0) for c in cases]
    try:
        with hip.Ensemble(ctxs) as ens:
            ens.large_step(3)
            ens.synchronize()
        case, ctx = cases[0], ctxs[0]
        m = po.Model(case, dtl=2.0)
        m.step(3)
        g = case.grid
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        for k in FIELDS:
            got, ref = ctx.model_get(k), m.state[k]
            assert np.abs(got[sl] - ref[sl]).max() <= 1e-12 * np.abs(ref[sl]).max(), k
        for iq in range(2):
            got, ref = ctx.model_get(hip.QTRC_BASE + iq), m.qtrc[..., iq]
            assert np.abs(got[sl] - ref[sl]).max() <= 1e-12 * max(np.abs(ref[sl]).max(), 1e-30), iq
    finally:
        for ctx in ctxs:
            ctx.close()


def test_ensemble_rejects_mixed_configurations():
    cases = member_cases()[:2]
    a = hip.model_from_case(cases[0])
    b = hip.model_from_case(cases[1], mixed_precision=True)
    try:
        with pytest.raises(hip.ScaleHipError):
            hip.Ensemble([a, b])
    finally:
        a.close()
        b.close()
