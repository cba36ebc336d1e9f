"""GPU parity of the code paths the 512x512x128 shear workload runs (DZ
stretched 50..500 m, linear shear, 3 FCT tracers), plus the integrator and
sub-stepping options of the large step.

* One HEVI stage on a stretched grid, where the face weights F2H
  (HEVI.F90:62) differ from 1/2: bit-identical to the oracle.
* The numerical filter with each tile height (4, 8 and 16 rows) forced
  through the context's knob: bit-identical.
* Large steps of a reduced shear workload (same DZ range, shear, noise and
  tracers), which run the fused RK4 combination of the stage kernel and the
  FCT tracer step: within 1e-12 of each field's scale.
* nstep = 2 (dtl = 2 dts: the sub-stepped short loop and the mass-flux
  average of LARGE.F90:960-1031), the RK3 short integrators and the Euler
  tracer integrator: within 1e-12 of the oracle.
* The full 512x512x128 workload: 3 large steps conserve mass and tracer
  mass to 1e-13, produce no NaN, and two runs are bit-identical.

Multi-step runs pass through pow() in the large-step preparation, whose GPU
and glibc results may differ in the last bit, hence 1e-12 rather than
bitwise (the single calls are bitwise).
"""
import numpy as np
import pytest

import helpers as H
import pyoracle as po
from scale_amd import inputs
from test_gpu_parity import OUTS, oracle_stage, run_stage_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from scale_amd import hip as shp
    shp.load_library()
    return shp


def interior_slice(g):
    return (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))


def check_model(ctx, m, case, nq, tol=1.0e-12):
    g = case.grid
    sl = interior_slice(g)
    for k in H.VARS:
        got = ctx.model_get(k)
        ref = m.state[k]
        scale = np.abs(ref[sl]).max()
        d = np.abs(got[sl] - ref[sl]).max()
        print(f"  {k}: max|diff| {d:.3e} (scale {scale:.3e})", flush=True)
        assert d <= tol * scale, k
    for iq in range(nq):
        got = ctx.model_get(ctx_qtrc(iq))
        ref = m.qtrc[..., iq]
        d = np.abs(got[sl] - ref[sl]).max()
        print(f"  QTRC{iq}: max|diff| {d:.3e}", flush=True)
        assert d <= tol * max(np.abs(ref[sl]).max(), 1e-30), iq


def ctx_qtrc(iq):
    from scale_amd import hip as shp
    return shp.QTRC_BASE + iq


def stretched_f2h_is_nontrivial(case):
    """F2H(k,1) = CDZ(k+1)/(CDZ(k)+CDZ(k+1)) differs from 1/2 on the grid."""
    cdz = case.metric["CDZ"]
    g = case.grid
    f = cdz[g.KS + 1:g.KE + 1] / (cdz[g.KS:g.KE] + cdz[g.KS + 1:g.KE + 1])
    return np.abs(f - 0.5).max() > 1e-3


@pytest.mark.parametrize("divdmp", [False, True], ids=["plain", "divdmp"])
@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_hevi_stage_bitwise_stretched(hip, scheme, divdmp):
    """hevi_tile (flat terrain) on DZ stretched 50..500 m: bit-identical."""
    print(f"[hevi stage] stretched dz, scheme={scheme} divdmp={divdmp}", flush=True)
    case = H.small_case(seed=3000 + scheme, stretched=(50.0, 500.0), kmax=24)
    assert stretched_f2h_is_nontrivial(case)
    inp = H.stage_inputs(case, seed=31)
    if divdmp:
        inp["DDIV"] = np.asfortranarray(np.random.Generator(np.random.PCG64(5)).standard_normal(case.grid.shape) * 1e-3)
        inp["divdmp_coef"] = 0.1
    ref = oracle_stage(case, inp, 1.0, None, scheme)
    got = run_stage_gpu(hip, case, inp, 1.0, None, scheme)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


@pytest.mark.parametrize("tj", [4, 8, 16])
def test_numfilter_bitwise_tile_heights(hip, tj):
    """numfilter_tile<TJ> for every tile height on a stretched grid with two
    i-tiles (one partial) and partial j-tiles: bit-identical."""
    print(f"[numfilter] TJ={tj}", flush=True)
    case = H.small_case(seed=97, stretched=(50.0, 500.0), kmax=20, imax=70, jmax=20)
    g = case.grid
    st = H.perturbed(case, 6)
    ref = po.numfilter_coef(case, st, 2.0)
    got = np.full(ref.shape, np.nan, order="F")
    m = case.metric
    with hip.Context(g) as ctx:
        ctx.tune_set(hip.TUNE_NF_TJ, tj)
        assert ctx.tune_get(hip.TUNE_NF_TJ) == tj
        ctx.ATMOS_DYN_numfilter_coef(got, st["DENS"], st["MOMZ"], st["MOMX"], st["MOMY"], st["RHOT"],
                                     m["CDZ"], m["CDX"], m["CDY"], m["FDZ"], m["FDX"], m["FDY"], 2.0,
                                     case.ref["REF_dens"], case.ref["REF_pott"], 1.0e-4, 1, 1.0, True)
    H.compare(ref, got, "num_diff")


def test_tune_knob_rejects_bad_values(hip):
    case = H.small_case()
    with hip.Context(case.grid) as ctx:
        with pytest.raises(hip.ScaleHipError):
            ctx.tune_set(hip.TUNE_NF_TJ, 5)
        with pytest.raises(hip.ScaleHipError):
            ctx.tune_set(999, 0)


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_model_parity_stretched_shear_fct(hip, scheme):
    """The shear workload at 24x20x32 (DZ 50..500 m, shear 0..20 m/s over
    10 km, theta noise 0.1 K, dt 2 s, 3 FCT tracers): 5 large steps within
    1e-12 of the oracle."""
    print(f"[model] stretched shear 24x20x32, 3 FCT tracers, scheme {scheme}, 5 steps", flush=True)
    case = inputs.shear_512x512x128(kmax=32, imax=24, jmax=20, scheme=scheme)
    assert stretched_f2h_is_nontrivial(case)
    m = po.Model(case)
    m.step(5)
    ctx = hip.model_from_case(case)
    try:
        ctx.model_large_step(5)
        ctx.synchronize()
        check_model(ctx, m, case, 3)
    finally:
        ctx.close()


@pytest.mark.parametrize("short", ["hevi", "heve"])
def test_model_parity_nstep2(hip, short):
    """dtl = 2 dts: two short steps per large step, the tracers advanced with
    the averaged mass flux (LARGE.F90:960-1031): 4 large steps within 1e-12."""
    print(f"[model] nstep = 2, {short}", flush=True)
    dts = 0.5 if short == "hevi" else 0.1
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.01, ntracer=2, tracer_kind=["blob", "tophat"], fct=True,
                            bubble_r=2000.0, bubble_zc=2000.0, momz_noise=0.05, v0=-2.0, dt=dts,
                            stretched=(100.0, 400.0))
    m = po.Model(case, dtl=2 * dts, short=short)
    m.step(4)
    ctx = hip.model_from_case(case, dtl=2 * dts, short=short)
    try:
        ctx.model_large_step(4)
        ctx.synchronize()
        check_model(ctx, m, case, 2)
    finally:
        ctx.close()


@pytest.mark.parametrize("tinteg", ["RK3", "RK3WS2002"])
@pytest.mark.parametrize("nstep", [1, 3])
def test_model_parity_rk3(hip, tinteg, nstep):
    """ATMOS_DYN_TINTEG_SHORT_TYPE = RK3 (Heun) and RK3WS2002
    (tinteg_short_rk3.F90:188-576) with FCT tracers: 4 large steps within 1e-12."""
    print(f"[model] {tinteg}, nstep {nstep}", flush=True)
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.01, ntracer=1, tracer_kind=["tophat"], fct=True,
                            bubble_r=2000.0, bubble_zc=2000.0, momz_noise=0.05, v0=-2.0, dt=0.5,
                            stretched=(100.0, 400.0))
    m = po.Model(case, dtl=nstep * 0.5, tinteg_short=tinteg)
    m.step(4)
    ctx = hip.model_from_case(case, dtl=nstep * 0.5, tinteg_short=tinteg)
    try:
        ctx.model_large_step(4)
        ctx.synchronize()
        check_model(ctx, m, case, 1)
    finally:
        ctx.close()


@pytest.mark.parametrize("fct", [False, True], ids=["nofct", "fct"])
def test_model_parity_tracer_euler(hip, fct):
    """ATMOS_DYN_TINTEG_TRACER_TYPE = EULER (tinteg_tracer_euler.F90:80-133): 5 large steps within 1e-12."""
    print(f"[model] tracer Euler, fct={fct}", flush=True)
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.01, ntracer=2, tracer_kind=["blob", "tophat"], fct=fct,
                            bubble_r=2000.0, bubble_zc=2000.0, momz_noise=0.05, v0=-2.0, u0=12.0)
    m = po.Model(case, tinteg_tracer="EULER")
    m.step(5)
    ctx = hip.model_from_case(case, tinteg_tracer="EULER")
    try:
        ctx.model_large_step(5)
        ctx.synchronize()
        check_model(ctx, m, case, 2)
    finally:
        ctx.close()


@pytest.mark.parametrize("variant", ["terrain", "heve", "heve_terrain"])
def test_model_parity_halo3_terrain_and_heve(hip, variant):
    """A lateral halo of 3 (the UD5 tracer scheme's) with terrain metrics
    (hevi_march, tracer kernels with metrics) and with the HEVE short step:
    5 large steps within 1e-12."""
    print(f"[model] halo 3, UD5 tracers, {variant}", flush=True)
    short = "heve" if variant.startswith("heve") else "hevi"
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.01, ntracer=2, tracer_kind=["blob", "tophat"], fct=True,
                            bubble_r=2000.0, bubble_zc=2000.0, momz_noise=0.05, v0=-2.0, scheme_tracer=5, halo=3,
                            dt=0.25 if short == "heve" else 1.0)
    if variant.endswith("terrain"):
        case = inputs.Case(**{**case.__dict__, "metric": H.terrain_metrics(case, amp=0.02)})
    m = po.Model(case, short=short)
    m.step(5)
    ctx = hip.model_from_case(case, short=short)
    try:
        ctx.model_large_step(5)
        ctx.synchronize()
        check_model(ctx, m, case, 2)
    finally:
        ctx.close()


def test_full_size_shear_properties(hip):
    """The 512x512x128 shear workload itself (3 FCT tracers): 3 large steps
    conserve total mass and each tracer's mass to 1e-13 relative, leave no
    NaN in the interior, and a second run from the same state is
    bit-identical."""
    print("[full size] 512x512x128 shear, 3 FCT tracers, 3 steps, twice", flush=True)
    case = inputs.shear_512x512x128()
    g = case.grid
    sl = interior_slice(g)
    runs = []
    for rep in range(2):
        ctx = hip.model_from_case(case)
        try:
            m0 = ctx.model_mass("DENS")
            q0 = [ctx.model_mass(hip.QTRC_BASE + iq, weight_dens=True) for iq in range(3)]
            ctx.model_large_step(3)
            m1 = ctx.model_mass("DENS")
            q1 = [ctx.model_mass(hip.QTRC_BASE + iq, weight_dens=True) for iq in range(3)]
            print(f"  run {rep}: mass rel. change {abs(m1 - m0) / m0:.2e}, tracers "
                  + ", ".join(f"{abs(b - a) / a:.2e}" for a, b in zip(q0, q1)), flush=True)
            assert abs(m1 - m0) <= 1e-13 * abs(m0)
            for a, b in zip(q0, q1):
                assert abs(b - a) <= 1e-13 * abs(a)
            out = {}
            for k in H.VARS:
                a = ctx.model_get(k)
                assert not np.isnan(a[sl]).any(), k
                out[k] = a[sl].copy()
            for iq in range(3):
                a = ctx.model_get(hip.QTRC_BASE + iq)
                assert not np.isnan(a[sl]).any(), iq
                out[f"Q{iq}"] = a[sl].copy()
            runs.append(out)
        finally:
            ctx.close()
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k
