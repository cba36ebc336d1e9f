"""GPU tests of the adjoint over terrain-following grids: the transpose of
the stage and of whole large steps, checked by the dot-product test
<L dx, y> = <dx, L^T y> against the oracle's tangent-linear model (the
oracle on dual numbers, whose tangent test_tlm_oracle.py pins by central
differences on the same terrain stage).

Tolerances: the GPU adjoint sums with FP64 atomics (reordered sums), and the
GPU forward over terrain differs from the oracle's by the last bits of pow();
the two sides of the dot product must agree within 1e-10 relative.
"""
import numpy as np
import pytest

from scale_amd import hip, inputs

pytestmark = pytest.mark.gpu

O = pytest.importorskip("pyoracle")
T = pytest.importorskip("pytlm")
STATE = hip.STATE

from test_tlm_oracle import stage_directions, terrain_stage_case  # noqa: E402


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_terrain_stage_adjoint_dot_product(scheme):
    """One HEVI stage on random smooth GSQRT, J13G, J23G and MAPF with
    Coriolis (every metric term of the stage, the map-factor part of the
    Coriolis/metric term included): the GPU stage adjoint through the C ABI
    against the dual-number stage, on 3 seeded directions and weights."""
    print(f"[terrain adjoint] stage 8x8x10 on random metrics, scheme {scheme}", flush=True)
    case, inp = terrain_stage_case(scheme)
    g = case.grid
    m = case.metric
    dt = 0.5
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    rng = np.random.default_rng(40 + scheme)
    with hip.Context(g, scheme, scheme) as ctx:
        for trial in range(3):
            t = stage_directions(case, rng)
            y = {}
            for k in T.STAGE_OUT:
                a = np.zeros(g.shape, order="F")
                a[sl] = rng.standard_normal(a[sl].shape)
                y[k] = a
            _, tan = T.stage_tangent(case, inp, dt, tangents=t)
            lhs = sum(float(np.sum(tan[k][sl] * y[k][sl])) for k in T.STAGE_OUT)
            got = ctx.ATMOS_DYN_Tstep_short_fvm_hevi_adj(
                y, *[inp[k + "0"] for k in STATE], *[inp[k] for k in STATE], *[inp[k + "_t"] for k in STATE],
                inp["DPRES0"], inp["RT2P"], inp["CORIOLI"], inp["num_diff"], case.wdamp_coef,
                m["CDZ"], m["FDZ"], m["FDX"], m["FDY"], m["RCDZ"], m["RCDX"], m["RCDY"], m["RFDZ"], m["RFDX"], m["RFDY"],
                m["PHI"], m["GSQRT"], m["J13G"], m["J23G"], m["J33G"], m["MAPF"], inp["REF_dens"], inp["REF_rhot"], dt)
            rhs = sum(float(np.sum(got[key] * t[key])) for key in t)
            print(f"  trial {trial}: <L dx, y> = {lhs:.15e}, <dx, L^T y> (GPU) = {rhs:.15e}", flush=True)
            assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)


def test_terrain_stage_adjoint_odd_width():
    """The same on 65 x 4 x 10 (UD3): x rows of 66 faces across the 64-lane
    waves of the shuffle-summed scatters, y rows and columns shorter than, or
    a partial multiple of, the 8-point line windows of the J13/J23, mass-flux
    and momentum scatters."""
    print("[terrain adjoint] stage 65x4x10 on random metrics, UD3", flush=True)
    case, inp = terrain_stage_case(3, kmax=10, imax=65, jmax=4)
    g = case.grid
    m = case.metric
    dt = 0.5
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    rng = np.random.default_rng(49)
    with hip.Context(g, 3, 3) as ctx:
        t = stage_directions(case, rng)
        y = {}
        for k in T.STAGE_OUT:
            a = np.zeros(g.shape, order="F")
            a[sl] = rng.standard_normal(a[sl].shape)
            y[k] = a
        _, tan = T.stage_tangent(case, inp, dt, tangents=t)
        lhs = sum(float(np.sum(tan[k][sl] * y[k][sl])) for k in T.STAGE_OUT)
        got = ctx.ATMOS_DYN_Tstep_short_fvm_hevi_adj(
            y, *[inp[k + "0"] for k in STATE], *[inp[k] for k in STATE], *[inp[k + "_t"] for k in STATE],
            inp["DPRES0"], inp["RT2P"], inp["CORIOLI"], inp["num_diff"], case.wdamp_coef,
            m["CDZ"], m["FDZ"], m["FDX"], m["FDY"], m["RCDZ"], m["RCDX"], m["RCDY"], m["RFDZ"], m["RFDX"], m["RFDY"],
            m["PHI"], m["GSQRT"], m["J13G"], m["J23G"], m["J33G"], m["MAPF"], inp["REF_dens"], inp["REF_rhot"], dt)
        rhs = sum(float(np.sum(got[key] * t[key])) for key in t)
    print(f"  <L dx, y> = {lhs:.15e}, <dx, L^T y> (GPU) = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_terrain_model_adjoint_dot_product(scheme):
    """6 large steps over a bell mountain (h = 500 m, a = 2.5 km; metrics
    generated on the device) with 2 FCT tracers, checkpoints every 4: <L dx, y>
    from the oracle's tangent-linear model on the oracle's metrics against
    <dx, L^T y> from the GPU's checkpointed forward and reverse sweep."""
    print(f"[terrain adjoint] model 24x20x16 over a bell mountain, scheme {scheme}", flush=True)
    case = inputs.make_case(kmax=16, imax=24, jmax=20, v0=2.0, momz_noise=0.05, noise=0.1, ntracer=2,
                            tracer_kind=["blob", "tophat"], fct=True, scheme=scheme, u0=8.0)
    g = case.grid
    zs = inputs.bell_topography(g, height=500.0, half_width=2500.0)
    om = O.terrain_metrics(g, zs, case.fz, case.metric)
    ocase = inputs.Case(**{**case.__dict__, "metric": {**case.metric, **om}})
    N = 6
    rng = np.random.default_rng(51 + scheme)
    ctx = hip.model_from_case(case, zsfc=zs)
    try:
        ctx.model_forward_checkpointed(N, 4)
        for trial in range(3):
            d = {}
            for k in STATE:
                a = np.zeros(g.shape, order="F")
                ke = g.KE - 1 if k == "MOMZ" else g.KE
                s = (slice(g.KS, ke + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
                a[s] = rng.standard_normal(a[s].shape) * np.abs(case.state[k]).max()
                d[k] = inputs.fill_ghost_and_halo(g, a, k)
            y = {k: np.asfortranarray(rng.standard_normal(g.shape)) for k in STATE}
            t = T.TangentModel(ocase, ocase.state, d)
            t.step(N)
            tan = t.tangents()
            lhs = sum(float(np.sum(tan[k] * y[k])) for k in STATE)
            xb0 = ctx.model_adjoint(y)
            grad = inputs.interior_gradient(g, xb0)
            rhs = sum(float(np.sum(grad[k] * d[k])) for k in STATE)
            print(f"  trial {trial}: <L dx, y> = {lhs:.15e}, <dx, L^T y> (GPU) = {rhs:.15e}", flush=True)
            assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)
    finally:
        ctx.close()


def test_batch_adjoint_equals_member_runs():
    """Three members of uneven sizes (two over bell mountains of different
    heights, one flat) with 2 FCT tracers, 4 large steps with checkpoints
    every 2: the batched forward leaves every member's state bit for bit as
    its own checkpointed forward does, and the batched reverse sweep returns
    each member's own model_adjoint within 1e-10 of each field's largest
    adjoint (the atomic sums reorder)."""
    print("[terrain adjoint] batch of 3 uneven members against their own runs", flush=True)
    specs = [dict(kmax=16, imax=24, jmax=20, h=500.0), dict(kmax=12, imax=16, jmax=16, h=300.0),
             dict(kmax=16, imax=20, jmax=24, h=0.0)]
    members = []
    for q, sp in enumerate(specs):
        case = inputs.make_case(kmax=sp["kmax"], imax=sp["imax"], jmax=sp["jmax"], seed=60 + q, bv=0.012, u0=10.0,
                                v0=1.0, bubble=False, noise=0.1, momz_noise=0.05, ntracer=2,
                                tracer_kind=["blob", "tophat"], fct=True)
        zs = inputs.bell_topography(case.grid, height=sp["h"], half_width=2500.0) if sp["h"] > 0 else None
        members.append((case, zs))
    N, every = 4, 2
    rng = np.random.default_rng(70)
    ys = [{k: np.asfortranarray(rng.standard_normal(c.grid.shape)) for k in STATE} for c, _ in members]
    ref_state, ref_adj = [], []
    for (case, zs), y in zip(members, ys):
        ctx = hip.model_from_case(case, zsfc=zs)
        try:
            ctx.model_forward_checkpointed(N, every)
            ref_state.append({k: ctx.model_get(k) for k in STATE})
            ref_adj.append(ctx.model_adjoint(y))
        finally:
            ctx.close()
    ctxs = [hip.model_from_case(case, zsfc=zs) for case, zs in members]
    try:
        hip.batch_forward_checkpointed(ctxs, N, every)
        for m, ctx in enumerate(ctxs):
            for k in STATE:
                assert np.array_equal(ctx.model_get(k), ref_state[m][k]), (m, k)
        got = hip.batch_adjoint(ctxs, ys)
    finally:
        for ctx in ctxs:
            ctx.close()
    for m in range(len(members)):
        for k in STATE:
            ref = ref_adj[m][k]
            scale = max(float(np.abs(ref).max()), 1e-300)
            err = float(np.abs(got[m][k] - ref).max())
            assert err <= 1e-10 * scale, (m, k, err / scale)
