"""Terrain-following metrics generated on the device (scale_hip_set_terrain)
against the oracle's restatement of the reference's metric code
(orc_terrain_metrics, pinned in test_oracle.py), bit for bit, and a model
on a bell mountain whose metrics the device computed, against the oracle
model on the oracle's metrics."""
import numpy as np
import pytest

import pyoracle as po
from scale_amd import hip, inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("stretched", [None, (60.0, 400.0)], ids=["uniform", "stretched"])
def test_terrain_metrics_bitwise(stretched):
    """A bell mountain h = 800 m, a = 2 km (mod_mktopo.F90:238-246) on a
    40x36x20 grid: GSQRT, J13G, J23G (7 components) and PHI."""
    print(f"[terrain] metrics, stretched={stretched}", flush=True)
    case = inputs.make_case(kmax=20, imax=40, jmax=36, dx=500.0, dy=500.0, dz=250.0, seed=5, ntracer=0,
                            stretched=stretched)
    g = case.grid
    zs = inputs.bell_topography(g, height=800.0, half_width=2000.0)
    ref = po.terrain_metrics(g, zs, case.fz, case.metric)
    with hip.Context(g) as ctx:
        ctx.set_terrain(case.metric, case.fz, zs)
        got = ctx.get_metrics()
    for k in ("GSQRT", "J13G", "J23G", "PHI"):
        mask = ~np.isnan(ref[k])
        assert mask.sum() > 0
        assert np.array_equal(ref[k][mask], got[k][mask]), f"{k}: max |diff| {np.abs(ref[k][mask] - got[k][mask]).max():.3e}"


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_model_on_device_terrain_matches_oracle(scheme):
    """5 large steps over a bell mountain (h = 500 m, a = 2.5 km) with two FCT
    tracers: the model on device-generated metrics against the oracle model
    on the oracle's metrics, within 1e-12 of each field's scale."""
    print(f"[terrain] model over a bell mountain, scheme {scheme}", flush=True)
    case = inputs.make_case(kmax=20, imax=24, jmax=20, dx=500.0, dy=500.0, dz=250.0, seed=6, ntracer=2,
                            tracer_kind=["blob", "tophat"], fct=True, u0=10.0, noise=0.05, scheme=scheme,
                            bubble_r=2000.0, bubble_zc=2000.0)
    g = case.grid
    zs = inputs.bell_topography(g, height=500.0, half_width=2500.0)
    om = po.terrain_metrics(g, zs, case.fz, case.metric)
    ocase = inputs.Case(**{**case.__dict__, "metric": {**case.metric, **om}})
    m = po.Model(ocase)
    m.step(5)
    ctx = hip.model_from_case(case, zsfc=zs)
    try:
        ctx.model_large_step(5)
        ctx.synchronize()
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        for k in hip.STATE:
            got, ref = ctx.model_get(k), m.state[k]
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  {k}: max |diff| {d:.3e}", flush=True)
            assert d <= 1e-12 * np.abs(ref[sl]).max(), k
        for iq in range(2):
            got, ref = ctx.model_get(hip.QTRC_BASE + iq), m.qtrc[..., iq]
            assert np.abs(got[sl] - ref[sl]).max() <= 1e-12 * max(np.abs(ref[sl]).max(), 1e-30), iq
    finally:
        ctx.close()
