"""The tiled tracer kernels (LDS tiles of 64 x 8 columns marching k: the FCT
ratio pass with phi_lo on a one-cell ring, the FCT update, and the stage
without FCT) against the oracle, bit
for bit, on grids with several tiles in i and j, partial tiles at the east
and north edges, and a stretched vertical grid; and against the per-point
kernels (tuning knob SCALE_HIP_TUNE_TRACER_TILE = 0), bit for bit.  The small
test grids force the tiles with the knob's value 2."""
import numpy as np
import pytest

import helpers as H
import pyoracle as po
from scale_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from scale_amd import hip as shp
    shp.load_library()
    return shp


def tracer_case(scheme, imax, jmax, kmax, seed):
    case = H.small_case(kmax=kmax, imax=imax, jmax=jmax, seed=seed, stretched=(60.0, 400.0), scheme_tracer=scheme)
    g = case.grid
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    st = H.perturbed(case, seed + 2)
    mflx = np.asfortranarray(rng.standard_normal(g.shape + (3,)) * np.array([0.3, 8.0, 5.0]))
    for n in range(3):
        inputs.halo_fill(g, mflx[..., n])
    q = np.asfortranarray(case.qtrc[..., 1] * (1.0 + 0.02 * rng.standard_normal(g.shape)))
    inputs.halo_fill(g, q)
    q0 = np.asfortranarray(case.qtrc[..., 1] * (1.0 + 0.01 * rng.standard_normal(g.shape)))
    inputs.halo_fill(g, q0)
    nd = np.asfortranarray(1e-5 * rng.standard_normal(g.shape + (3,)))
    rq = np.asfortranarray(1e-7 * rng.standard_normal(g.shape))
    return case, dict(QTRC=q, QTRC0=q0, RHOQ_t=rq, DENS0=case.state["DENS"], DENS=st["DENS"], mflx=mflx, nd=nd)


def run(hip, case, inp, fct, along, tile):
    g = case.grid
    m = case.metric
    out = np.full(g.shape, np.nan, order="F")
    mapf_xy = np.asfortranarray(m["MAPF"][:, :, :, 0])
    with hip.Context(g, case.scheme, case.scheme_tracer) as ctx:
        ctx.tune_set(hip.TUNE_TRACER_TILE, 2 if tile else 0)
        ctx.ATMOS_DYN_Tstep_tracer_fvm_heve(out, inp["QTRC"], inp["QTRC0"], inp["RHOQ_t"], inp["DENS0"], inp["DENS"],
                                            inp["mflx"], inp["nd"], m["GSQRT"], mapf_xy, m["CDZ"], m["RCDZ"], m["RCDX"],
                                            m["RCDY"], 1.5, fct, along)
    return out


@pytest.mark.parametrize("fct,along", [(True, True), (True, False), (False, False)], ids=["fct_along", "fct_7pt", "nofct"])
@pytest.mark.parametrize("scheme", [3, 4, 31, 1, 2], ids=["UD3", "CD4", "UD3KOREN1993", "UD1", "CD2"])
def test_tracer_tile_bitwise(hip, scheme, fct, along):
    """70 x 20 x 24 interior: two i-tiles (the second 6 columns wide) and
    three j-tiles (the last 4 rows high)."""
    print(f"[tracer tile] scheme={scheme} fct={fct} along={along}", flush=True)
    case, inp = tracer_case(scheme, 70, 20, 24, 4100 + scheme)
    ref = np.full(case.grid.shape, np.nan, order="F")
    po.tstep_tracer(case, ref, inp["QTRC"], inp["QTRC0"], inp["RHOQ_t"], inp["DENS0"], inp["DENS"], inp["mflx"],
                    inp["nd"], 1.5, fct, along, scheme=scheme)
    got = run(hip, case, inp, fct, along, tile=True)
    H.compare(ref, got, "QTRCo (tiles)")
    old = run(hip, case, inp, fct, along, tile=False)
    assert np.array_equal(got, old, equal_nan=True)


def test_tracer_tile_wide_grid(hip):
    """A grid with full 64-column tiles only (128 x 16) and an FCT top-hat:
    bit-identical to the per-point kernels."""
    print("[tracer tile] 128 x 16 x 20, UD3, along-stream FCT", flush=True)
    case, inp = tracer_case(3, 128, 16, 20, 4200)
    got = run(hip, case, inp, True, True, tile=True)
    old = run(hip, case, inp, True, True, tile=False)
    g = case.grid
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    assert not np.isnan(got[sl]).any()
    assert np.array_equal(got, old, equal_nan=True)
