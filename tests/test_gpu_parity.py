"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.

Single calls (one HEVI stage, numfilter, one tracer stage) perform the same
IEEE-754 operations in the same order on both sides, so they are required to
be bit-identical on every element the reference defines.  Multi-step
integrations additionally go through pow() in the large-step preparation
(RT2P, DPRES0), whose GPU and glibc implementations may differ in the last
bit; they are held to the FP64 tolerance of the workload specification
(1e-12 relative after N steps).
"""
import numpy as np
import pytest

import helpers as H
import pyoracle as po
from scale_amd import inputs

pytestmark = pytest.mark.gpu

OUTS = ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK", "mflx_hi", "tflx_hi")


@pytest.fixture(scope="module")
def hip():
    from scale_amd import hip as shp
    shp.load_library()
    return shp


def run_stage_gpu(hip, case, inp, dtrk, metric=None, scheme=None, heve=False, fct=(False, False), along=True):
    g = case.grid
    m = metric or case.metric
    out = {k: po.nan_like_outputs(g) for k in OUTS[:5]}
    out["mflx_hi"] = po.nan_like_outputs(g, 3)
    out["tflx_hi"] = po.nan_like_outputs(g, 3)
    with hip.Context(g, scheme or case.scheme, case.scheme_tracer) as ctx:
        stage = ctx.ATMOS_DYN_Tstep_short_fvm_heve if heve else ctx.ATMOS_DYN_Tstep_short_fvm_hevi
        stage(
            out["DENS_RK"], out["MOMZ_RK"], out["MOMX_RK"], out["MOMY_RK"], out["RHOT_RK"], None, out["mflx_hi"], out["tflx_hi"],
            inp["DENS0"], inp["MOMZ0"], inp["MOMX0"], inp["MOMY0"], inp["RHOT0"],
            inp["DENS"], inp["MOMZ"], inp["MOMX"], inp["MOMY"], inp["RHOT"],
            inp["DENS_t"], inp["MOMZ_t"], inp["MOMX_t"], inp["MOMY_t"], inp["RHOT_t"], None, None,
            inp["DPRES0"], inp["RT2P"], inp.get("CORIOLI", case.corioli), inp["num_diff"], inp.get("wdamp_coef", case.wdamp_coef),
            inp.get("divdmp_coef", 0.0), inp["DDIV"], fct[0], fct[1], along,
            m["CDZ"], m["FDZ"], m["FDX"], m["FDY"], m["RCDZ"], m["RCDX"], m["RCDY"], m["RFDZ"], m["RFDX"], m["RFDY"],
            m["PHI"], m["GSQRT"], m["J13G"], m["J23G"], m["J33G"], m["MAPF"], inp["REF_dens"], inp["REF_rhot"],
            False, False, False, False, dtrk, False)
    return out


def oracle_stage(case, inp, dtrk, metric=None, scheme=None, heve=False, **kw):
    if metric is not None:
        case = inputs.Case(**{**case.__dict__, "metric": metric})
    if heve:
        return po.tstep_short_heve(case, inp, dtrk, scheme=scheme, **kw)
    return po.tstep_short_hevi(case, inp, dtrk, scheme=scheme)


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_hevi_stage_bitwise(hip, scheme, terrain):
    print(f"[hevi stage] scheme={scheme} terrain={terrain}", flush=True)
    case = H.small_case(seed=1000 + scheme)
    inp = H.stage_inputs(case, seed=21, coriolis=terrain)
    metric = H.terrain_metrics(case) if terrain else None
    if terrain:
        inp["wdamp_coef"] = np.linspace(0.0, 0.01, case.grid.KA)
    ref = oracle_stage(case, inp, 0.5, metric, scheme)
    got = run_stage_gpu(hip, case, inp, 0.5, metric, scheme)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


@pytest.mark.parametrize("seg,parts", [(0, ""), (1, ""), (3, ""), (8, ""), (64, ""), (8, "15"), (3, "3,12")],
                         ids=["march", "seg1", "seg3", "seg8", "seg64", "seg8-one_launch", "seg3-two_launches"])
def test_hevi_stage_terrain_segments(hip, seg, parts, monkeypatch):
    """The terrain stage split in launches (hevi_expl over column segments of
    seg levels, each starting from the z fluxes of the face below it, one
    launch per section of the point by default, then hevi_solve) and in one
    (seg = 0: hevi_march): bit for bit the oracle's, with segment lengths
    that do and do not divide the 20 levels and the sections in one, two or
    four launches."""
    monkeypatch.setenv("SCALE_HIP_TERRAIN_SEG", str(seg))
    if parts:
        monkeypatch.setenv("SCALE_HIP_TERRAIN_PARTS", parts)
    print(f"[hevi stage] terrain, segments of {seg} levels, sections {parts or 'default'}", flush=True)
    case = H.small_case(seed=1003)
    inp = H.stage_inputs(case, seed=23, coriolis=True)
    metric = H.terrain_metrics(case)
    inp["wdamp_coef"] = np.linspace(0.0, 0.01, case.grid.KA)
    ref = oracle_stage(case, inp, 0.5, metric, 3)
    got = run_stage_gpu(hip, case, inp, 0.5, metric, 3)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
@pytest.mark.parametrize("divdmp", [False, True], ids=["plain", "divdmp"])
@pytest.mark.parametrize("seg", [3, 8], ids=["seg3", "seg8"])
def test_hevi_stage_flat_split(hip, scheme, divdmp, seg, monkeypatch):
    """The flat stage as split launches (the path of grids too small for the
    tile kernel to fill the device): bit for bit the oracle's."""
    monkeypatch.setenv("SCALE_HIP_FLAT_SPLIT", "1")
    monkeypatch.setenv("SCALE_HIP_TERRAIN_SEG", str(seg))
    print(f"[hevi stage] flat, split launches, scheme={scheme} divdmp={divdmp} seg={seg}", flush=True)
    case = H.small_case(seed=1100 + scheme, stretched=(50.0, 500.0), kmax=20)
    inp = H.stage_inputs(case, seed=25, coriolis=True)
    if divdmp:
        inp["divdmp_coef"] = 0.1
        inp["DDIV"] = np.asfortranarray(np.random.Generator(np.random.PCG64(4)).standard_normal(case.grid.shape) * 1e-3)
    ref = oracle_stage(case, inp, 0.5, None, scheme)
    got = run_stage_gpu(hip, case, inp, 0.5, None, scheme)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_hevi_stage_divdmp(hip, terrain):
    print(f"[hevi stage] divergence damping terrain={terrain}", flush=True)
    case = H.small_case(seed=77)
    inp = H.stage_inputs(case, seed=5)
    g = case.grid
    inp["DDIV"] = np.asfortranarray(np.random.Generator(np.random.PCG64(3)).standard_normal(g.shape) * 1e-3)
    inp["divdmp_coef"] = 0.1
    metric = H.terrain_metrics(case) if terrain else None
    ref = oracle_stage(case, inp, 1.0, metric)
    got = run_stage_gpu(hip, case, inp, 1.0, metric)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_heve_stage_bitwise(hip, scheme, terrain):
    """One explicit HEVE stage (HEVE.F90:109-1516) through the drop-in entry
    point, bit-identical to the oracle."""
    print(f"[heve stage] scheme={scheme} terrain={terrain}", flush=True)
    case = H.small_case(seed=2000 + scheme)
    inp = H.stage_inputs(case, seed=22, coriolis=terrain)
    metric = H.terrain_metrics(case) if terrain else None
    if terrain:
        inp["wdamp_coef"] = np.linspace(0.0, 0.01, case.grid.KA)
    ref = oracle_stage(case, inp, 0.5, metric, scheme, heve=True)
    got = run_stage_gpu(hip, case, inp, 0.5, metric, scheme, heve=True)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_heve_stage_divdmp(hip, terrain):
    print(f"[heve stage] divergence damping terrain={terrain}", flush=True)
    case = H.small_case(seed=78)
    inp = H.stage_inputs(case, seed=6)
    g = case.grid
    inp["DDIV"] = np.asfortranarray(np.random.Generator(np.random.PCG64(4)).standard_normal(g.shape) * 1e-3)
    inp["divdmp_coef"] = 0.1
    metric = H.terrain_metrics(case) if terrain else None
    ref = oracle_stage(case, inp, 1.0, metric, heve=True)
    got = run_stage_gpu(hip, case, inp, 1.0, metric, heve=True)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


@pytest.mark.parametrize("along", [True, False], ids=["along_stream", "7point"])
@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_heve_stage_fct_theta_bitwise(hip, terrain, along):
    """HEVE with FLAG_FCT_T (HEVE.F90:1388-1475) on a 10 % theta top-hat:
    the limited RHOT_RK and the halo-exchanged DENS_RK and mflx_hi are
    bit-identical to the oracle."""
    print(f"[heve stage] FCT theta terrain={terrain} along={along}", flush=True)
    case = H.small_case(seed=41, u0=15.0, v0=-8.0)
    g = case.grid
    inp = H.stage_inputs(case, seed=3, coriolis=terrain)
    for key in ("RHOT0", "RHOT"):
        a = np.array(inp[key], order="F", copy=True)
        a[g.KS + 3:g.KS + 9, g.IS + 2:g.IS + 7, g.JS + 2:g.JS + 6] *= 1.1
        inputs.halo_fill(g, a)
        inp[key] = a
    metric = H.terrain_metrics(case, amp=0.02) if terrain else None
    ref = oracle_stage(case, inp, 2.0, metric, 4, heve=True, fct_t=True, along_stream=along)
    got = run_stage_gpu(hip, case, inp, 2.0, metric, 4, heve=True, fct=(False, True), along=along)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


def _momentum_tophat_inputs(case, seed, coriolis):
    """Stage inputs with boxes of stronger wind in MOMX, MOMY and MOMZ (state
    and state0), so that the momentum limiter is active."""
    g = case.grid
    inp = H.stage_inputs(case, seed=seed, coriolis=coriolis)
    boxes = {"MOMX": (slice(g.KS + 3, g.KS + 9), slice(g.IS + 2, g.IS + 6), slice(g.JS + 2, g.JS + 6), 1.5),
             "MOMY": (slice(g.KS + 4, g.KS + 10), slice(g.IS + 5, g.IS + 9), slice(g.JS + 3, g.JS + 7), 1.4),
             "MOMZ": (slice(g.KS + 5, g.KS + 11), slice(g.IS + 3, g.IS + 8), slice(g.JS + 1, g.JS + 5), None)}
    for v, (sk, si, sj, f) in boxes.items():
        for key in (v + "0", v):
            a = np.array(inp[key], order="F", copy=True)
            if f is None:
                a[sk, si, sj] += 2.0 * inp["DENS0"][sk, si, sj]
            else:
                a[sk, si, sj] *= f
            inputs.halo_fill(g, a)
            inp[key] = a
    return inp


@pytest.mark.parametrize("along", [True, False], ids=["along_stream", "7point"])
@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_heve_stage_fct_momentum_bitwise(hip, scheme, terrain, along):
    """HEVE with FLAG_FCT_MOMENTUM (HEVE.F90:386-450, 716-810, 958-1057,
    1207-1309) on boxes of stronger wind: the limited MOMZ_RK, MOMX_RK and
    MOMY_RK and every other output are bit-identical to the oracle."""
    print(f"[heve stage] FCT momentum scheme={scheme} terrain={terrain} along={along}", flush=True)
    case = H.small_case(seed=43, u0=15.0, v0=-8.0)
    inp = _momentum_tophat_inputs(case, 4, terrain)
    metric = H.terrain_metrics(case, amp=0.02) if terrain else None
    ref = oracle_stage(case, inp, 2.0, metric, scheme, heve=True, fct_momentum=True, along_stream=along)
    plain = oracle_stage(case, inp, 2.0, metric, scheme, heve=True)
    g = case.grid
    sl = (slice(g.KS, g.KE), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    for v in ("MOMZ_RK", "MOMX_RK", "MOMY_RK"):  # the limiter acts
        assert np.abs(ref[v][sl] - plain[v][sl]).max() > 1e-6 * np.abs(plain[v][sl]).max(), v
    got = run_stage_gpu(hip, case, inp, 2.0, metric, scheme, heve=True, fct=(True, False), along=along)
    for k in OUTS:
        H.compare(ref[k], got[k], k)


@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_heve_stage_fct_momentum_and_theta_bitwise(hip, terrain):
    """Both limiters of the HEVE stage in one call, bit-identical to the oracle."""
    print(f"[heve stage] FCT momentum + theta terrain={terrain}", flush=True)
    case = H.small_case(seed=44, u0=15.0, v0=-8.0)
    g = case.grid
    inp = _momentum_tophat_inputs(case, 5, terrain)
    for key in ("RHOT0", "RHOT"):
        a = np.array(inp[key], order="F", copy=True)
        a[g.KS + 3:g.KS + 9, g.IS + 2:g.IS + 7, g.JS + 2:g.JS + 6] *= 1.1
        inputs.halo_fill(g, a)
        inp[key] = a
    metric = H.terrain_metrics(case, amp=0.02) if terrain else None
    ref = oracle_stage(case, inp, 2.0, metric, 3, heve=True, fct_momentum=True, fct_t=True)
    got = run_stage_gpu(hip, case, inp, 2.0, metric, 3, heve=True, fct=(True, True))
    for k in OUTS:
        H.compare(ref[k], got[k], k)


def test_numfilter_bitwise(hip):
    print("[numfilter]", flush=True)
    case = H.small_case(seed=99, stretched=(50.0, 400.0))
    g = case.grid
    st = H.perturbed(case, 4)
    ref = po.numfilter_coef(case, st, 1.5)
    got = np.full(ref.shape, np.nan, order="F")
    m = case.metric
    with hip.Context(g) as ctx:
        ctx.ATMOS_DYN_numfilter_coef(got, st["DENS"], st["MOMZ"], st["MOMX"], st["MOMY"], st["RHOT"],
                                     m["CDZ"], m["CDX"], m["CDY"], m["FDZ"], m["FDX"], m["FDY"], 1.5,
                                     case.ref["REF_dens"], case.ref["REF_pott"], 1.0e-4, 1, 1.0, True)
    H.compare(ref, got, "num_diff")


@pytest.mark.parametrize("is_qv", [False, True], ids=["tracer", "qv"])
def test_numfilter_q_bitwise(hip, is_qv):
    """ATMOS_DYN_numfilter_coef_q (COMMON.F90:989-1200) through the drop-in
    entry point: bit-identical to the oracle."""
    print(f"[numfilter_q] is_qv={is_qv}", flush=True)
    case = H.small_case(seed=98, stretched=(50.0, 400.0))
    g = case.grid
    st = H.perturbed(case, 5)
    rng = np.random.default_rng(3)
    q = np.asfortranarray(1e-3 * (1.0 + 0.2 * rng.standard_normal(g.shape)))
    ref_qv = np.asfortranarray(1e-3 * (1.0 + 0.05 * rng.standard_normal(g.shape))) if is_qv else None
    ref = po.numfilter_coef_q(case, st["DENS"], q, 3.0, nd_coef=2.0e-3, nd_sfc_fact=0.7, is_qv=is_qv, REF_qv=ref_qv)
    got = np.full(ref.shape, np.nan, order="F")
    m = case.metric
    with hip.Context(g) as ctx:
        nd = np.zeros((g.KA, g.IA, g.JA, 5, 3), order="F")  # builds the filter tables from the grid
        ctx.ATMOS_DYN_numfilter_coef(nd, st["DENS"], st["MOMZ"], st["MOMX"], st["MOMY"], st["RHOT"],
                                     m["CDZ"], m["CDX"], m["CDY"], m["FDZ"], m["FDX"], m["FDY"], 1.5,
                                     case.ref["REF_dens"], case.ref["REF_pott"], 1.0e-4, 1, 1.0, True)
        ctx.ATMOS_DYN_numfilter_coef_q(got, st["DENS"], q, is_qv, m["CDZ"], m["CDX"], m["CDY"], 3.0,
                                       ref_qv if is_qv else g.zeros(), 1, 2.0e-3, 1, 0.7, True)
    H.compare(ref, got, "num_diff_q")


def test_model_parity_tracer_filter(hip):
    """Large steps with the tracer filter on (ATMOS_DYN_NUMERICAL_DIFF_COEF_TRACER
    > 0, LARGE.F90:1049-1061) against the oracle: 5 steps, 1e-12 relative."""
    print("[model] tracer filter 1e-2, 5 steps", flush=True)
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.01, ntracer=2, tracer_kind=["blob", "tophat"], fct=True,
                            bubble_r=2000.0, bubble_zc=2000.0, momz_noise=0.05, v0=-2.0)
    g = case.grid
    m = po.Model(case, nd_coef_q=1.0e-2)
    m.step(5)
    ctx = hip.model_from_case(case, nd_coef_q=1.0e-2)
    try:
        ctx.model_large_step(5)
        ctx.synchronize()
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        for iq in range(2):
            got = ctx.model_get(hip.QTRC_BASE + iq)
            ref = m.qtrc[..., iq]
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  QTRC{iq}: max|diff| {d:.3e}", flush=True)
            assert d <= 1.0e-12 * max(np.abs(ref[sl]).max(), 1e-30)
    finally:
        ctx.close()


@pytest.mark.parametrize("fct,along", [(False, True), (True, True), (True, False)], ids=["nofct", "fct_along", "fct_7pt"])
@pytest.mark.parametrize("scheme,halo", [(3, 2), (4, 2), (31, 2), (1, 2), (2, 2), (5, 3), (6, 3), (3, 3)],
                         ids=["UD3", "CD4", "UD3KOREN1993", "UD1", "CD2", "UD5_halo3", "CD6_halo3", "UD3_halo3"])
def test_tracer_bitwise(hip, fct, along, scheme, halo):
    print(f"[tracer] fct={fct} along={along} scheme={scheme} halo={halo}", flush=True)
    case = H.small_case(seed=55, halo=halo)
    g = case.grid
    prep = inputs.large_step_prep(case)
    del prep
    st = H.perturbed(case, 8)
    rng = np.random.Generator(np.random.PCG64(9))
    mflx = np.asfortranarray(rng.standard_normal(g.shape + (3,)) * np.array([0.3, 8.0, 5.0]))
    for n in range(3):
        inputs.halo_fill(g, mflx[..., n])
    q = np.asfortranarray(case.qtrc[..., 1])
    q0 = np.asfortranarray(case.qtrc[..., 1] * (1.0 + 0.01 * rng.standard_normal(g.shape)))
    inputs.halo_fill(g, q0)
    nd = np.asfortranarray(1e-5 * rng.standard_normal(g.shape + (3,)))  # num_diff_q (unused by UD1)
    rq = np.zeros(g.shape, order="F")
    mapf_xy = np.asfortranarray(case.metric["MAPF"][:, :, :, 0])
    ref = np.full(g.shape, np.nan, order="F")
    po.tstep_tracer(case, ref, q, q0, rq, case.state["DENS"], st["DENS"], mflx, nd, 2.0, fct, along, scheme=scheme)
    got = np.full(g.shape, np.nan, order="F")
    with hip.Context(g, case.scheme, scheme) as ctx:
        m = case.metric
        ctx.ATMOS_DYN_Tstep_tracer_fvm_heve(got, q, q0, rq, case.state["DENS"], st["DENS"], mflx, nd, m["GSQRT"], mapf_xy,
                                            m["CDZ"], m["RCDZ"], m["RCDX"], m["RCDY"], 2.0, fct, along)
    H.compare(ref, got, "QTRCo")


@pytest.mark.parametrize("scheme_tracer", [5, 6], ids=["UD5", "CD6"])
def test_model_parity_halo3(hip, scheme_tracer):
    """Large steps on a grid with a lateral halo of 3 (IHALO = JHALO = 3, as
    the UD5/CD6 tracer schemes need) with FCT tracers and the tracer filter,
    against the oracle: 5 steps, 1e-12 relative for the dynamics and tracers."""
    print(f"[model] halo 3, tracer scheme {scheme_tracer}, 5 steps", flush=True)
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.01, ntracer=2, tracer_kind=["blob", "tophat"], fct=True,
                            bubble_r=2000.0, bubble_zc=2000.0, momz_noise=0.05, v0=-2.0, scheme_tracer=scheme_tracer,
                            halo=3)
    g = case.grid
    assert g.IS == 3
    m = po.Model(case, nd_coef_q=1.0e-2)
    m.step(5)
    ctx = hip.model_from_case(case, nd_coef_q=1.0e-2)
    try:
        ctx.model_large_step(5)
        ctx.synchronize()
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        fields = [(k, ctx.model_get(k), m.state[k]) for k in H.VARS]
        fields += [(f"QTRC{iq}", ctx.model_get(hip.QTRC_BASE + iq), m.qtrc[..., iq]) for iq in range(2)]
        for name, got, ref in fields:
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  {name}: max|diff| {d:.3e}", flush=True)
            assert d <= 1.0e-12 * max(np.abs(ref[sl]).max(), 1e-30), name
    finally:
        ctx.close()


@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_model_parity_divergence_damping(hip, terrain):
    """Large steps with divergence damping (ATMOS_DYN_divergence each short
    step, LARGE.F90:860-872) and Rayleigh damping near the top, against the
    oracle: 5 steps, 1e-12 relative to each field's scale."""
    print(f"[model] divdmp 0.2 + wdamp, terrain={terrain}, 5 steps", flush=True)
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.01, ntracer=0, bubble_r=2000.0, bubble_zc=2000.0,
                            momz_noise=0.05, u0=6.0, v0=-2.0)
    g = case.grid
    wd = hip.ATMOS_DYN_wdamp_setup(g, 100.0, 3000.0, case.fz)
    assert wd.max() > 0.0
    metric = H.terrain_metrics(case, amp=0.02) if terrain else case.metric
    case = inputs.Case(**{**case.__dict__, "wdamp_coef": np.asfortranarray(wd), "metric": metric})
    m = po.Model(case, divdmp_coef=0.2)
    m.step(5)
    ctx = hip.model_from_case(case, divdmp_coef=0.2)
    try:
        ctx.model_large_step(5)
        ctx.synchronize()
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        for k in H.VARS:
            got = ctx.model_get(k)
            ref = m.state[k]
            scale = np.abs(ref[sl]).max()
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  {k}: max|diff| {d:.3e} (scale {scale:.3e})", flush=True)
            assert d <= 1.0e-12 * scale, k
    finally:
        ctx.close()


@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_model_parity_heve(hip, terrain):
    """Large steps with the HEVE short step (ATMOS_DYN_TYPE = FVM-HEVE: four
    explicit stages and the RK4 combination) and FCT tracers against the
    oracle: 5 steps at dt = 0.25 s (inside the explicit acoustic limit of
    dz = 250 m), 1e-12 relative to each field's scale."""
    print(f"[model] HEVE, terrain={terrain}, 5 steps", flush=True)
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.01, ntracer=1, tracer_kind=["blob"], fct=True,
                            bubble_r=2000.0, bubble_zc=2000.0, momz_noise=0.05, v0=-2.0, dt=0.25)
    g = case.grid
    if terrain:
        case = inputs.Case(**{**case.__dict__, "metric": H.terrain_metrics(case, amp=0.02)})
    m = po.Model(case, short="heve")
    m.step(5)
    ctx = hip.model_from_case(case, short="heve")
    try:
        ctx.model_large_step(5)
        ctx.synchronize()
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        for k in H.VARS:
            got = ctx.model_get(k)
            ref = m.state[k]
            scale = np.abs(ref[sl]).max()
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  {k}: max|diff| {d:.3e} (scale {scale:.3e})", flush=True)
            assert d <= 1.0e-12 * scale, k
        got = ctx.model_get(hip.QTRC_BASE)
        ref = m.qtrc[..., 0]
        assert np.abs(got[sl] - ref[sl]).max() <= 1.0e-12 * np.abs(ref[sl]).max()
    finally:
        ctx.close()


@pytest.mark.parametrize("split", ["0", "1"], ids=["tile", "split_launches"])
def test_model_parity_warm_bubble(hip, split, monkeypatch):
    """Device-resident large steps (dynamics RK4 + RK3 tracers with FCT) against
    the oracle integrator, 10 steps of a reduced warm bubble: with the flat
    stage on the tile kernel (the fourth stage fused with the RK4
    combination) and as split launches (the combination its own kernel)."""
    monkeypatch.setenv("SCALE_HIP_FLAT_SPLIT", split)
    print(f"[model] warm bubble 24x20x16, 10 steps, split={split}", flush=True)
    case = inputs.make_case(kmax=24, imax=20, jmax=16, noise=0.01, ntracer=2, tracer_kind=["blob", "tophat"], fct=True,
                            bubble_r=2000.0, bubble_zc=2000.0)
    m = po.Model(case)
    m.step(10)
    ctx = hip.model_from_case(case)
    try:
        ctx.model_large_step(10)
        ctx.synchronize()
        g = case.grid
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        for k in H.VARS:
            got = ctx.model_get(k)
            ref = m.state[k]
            scale = np.abs(ref[sl]).max()
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  {k}: max|diff| {d:.3e} (scale {scale:.3e})", flush=True)
            assert d <= 1.0e-12 * scale, k
        for iq in range(2):
            got = ctx.model_get(hip.QTRC_BASE + iq)
            ref = m.qtrc[..., iq]
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  QTRC{iq}: max|diff| {d:.3e}", flush=True)
            assert d <= 1.0e-12 * max(np.abs(ref[sl]).max(), 1e-30)
    finally:
        ctx.close()


def test_model_mass_conservation(hip):
    print("[model] mass conservation", flush=True)
    case = inputs.make_case(kmax=32, imax=24, jmax=24, noise=0.1, ntracer=2, tracer_kind=["blob", "tophat"], fct=True)
    ctx = hip.model_from_case(case)
    try:
        m0 = ctx.model_mass("DENS")
        q0 = [ctx.model_mass(hip.QTRC_BASE + iq, weight_dens=True) for iq in range(2)]
        ctx.model_large_step(20)
        m1 = ctx.model_mass("DENS")
        q1 = [ctx.model_mass(hip.QTRC_BASE + iq, weight_dens=True) for iq in range(2)]
        assert abs(m1 - m0) <= 1.0e-13 * abs(m0)
        for a, b in zip(q0, q1):
            assert abs(b - a) <= 1.0e-13 * abs(a)
    finally:
        ctx.close()


def test_model_deterministic(hip):
    print("[model] determinism", flush=True)
    case = H.small_case(seed=5)
    outs = []
    for _ in range(2):
        ctx = hip.model_from_case(case)
        ctx.model_large_step(3)
        outs.append({k: ctx.model_get(k) for k in H.VARS})
        ctx.close()
    for k in H.VARS:
        assert np.array_equal(outs[0][k], outs[1][k]), k
