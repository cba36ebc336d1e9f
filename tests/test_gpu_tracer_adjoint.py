"""GPU adjoint of the tracer stage and of the three-stage tracer integration
with the FCT limiter frozen, through the C ABI, against the oracle's
tangent-linear model (scale_oracle.c on dual numbers with R+/R- held at their
forward values): <L dx, y> = <dx, L^T y>.

dx perturbs QTRC, QTRC0, RHOQ_t and num_diff_q on the whole array (the stage
reads QTRC and QTRC0 in the halo); y weights the interior, the points the
stage writes.  The adjoint sums in another order than the tangent, so the
two sides agree to 1e-11 relative rather than bit for bit.
"""
import numpy as np
import pytest

import helpers as H
from scale_amd import inputs

pytestmark = pytest.mark.gpu

T = pytest.importorskip("pytlm")


@pytest.fixture(scope="module")
def hip():
    from scale_amd import hip as shp
    shp.load_library()
    return shp


def stage_case(scheme, halo, terrain, seed=55):
    case = H.small_case(kmax=12, imax=10, jmax=8, seed=seed, halo=halo, scheme_tracer=scheme)
    g = case.grid
    rng = np.random.default_rng(9)
    mflx = np.asfortranarray(rng.standard_normal(g.shape + (3,)) * np.array([0.3, 8.0, 5.0]))
    for n in range(3):
        inputs.halo_fill(g, mflx[..., n])
    q = np.asfortranarray(case.qtrc[..., 1])
    q0 = np.asfortranarray(q * (1.0 + 0.01 * rng.standard_normal(g.shape)))
    inputs.halo_fill(g, q0)
    nd = np.asfortranarray(1e-5 * rng.standard_normal(g.shape + (3,)))
    rq = np.asfortranarray(1e-6 * rng.standard_normal(g.shape))
    st = H.perturbed(case, 8)
    metric = H.terrain_metrics(case) if terrain else case.metric
    x = dict(QTRC=q, QTRC0=q0, RHOQ_t=rq, DENS0=case.state["DENS"], DENS=st["DENS"], mflx_hi=mflx, num_diff_q=nd)
    return case, metric, x


def directions(case, rng, x):
    g = case.grid
    return {"QTRC": rng.standard_normal(g.shape) * np.abs(x["QTRC"]).max(),
            "QTRC0": rng.standard_normal(g.shape) * np.abs(x["QTRC0"]).max(),
            "RHOQ_t": rng.standard_normal(g.shape) * 1e-6,
            "num_diff": rng.standard_normal(g.shape + (3,)) * 1e-5}


def weights(case, rng):
    g = case.grid
    y = np.zeros(g.shape, order="F")
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    y[sl] = rng.standard_normal(y[sl].shape)
    return y, sl


def dot_check(case, metric, x, ctx, rk3, fct, along, seed):
    g = case.grid
    rng = np.random.default_rng(seed)
    mapf_xy = np.asfortranarray(metric["MAPF"][:, :, :, 0])
    met = (metric["GSQRT"], mapf_xy, metric["CDZ"], metric["RCDZ"], metric["RCDX"], metric["RCDY"])
    dt = 2.0
    for trial in range(2):
        d = directions(case, rng, x)
        y, sl = weights(case, rng)
        tl = T.tracer_rk3_tangent if rk3 else T.tracer_stage_tangent
        _, tan = tl(case, **x, dtl=dt, fct=fct, along_stream=along, metric=metric, tangents=d)
        lhs = float(np.sum(tan[sl] * y[sl]))
        if rk3:
            b = ctx.ATMOS_DYN_Tinteg_tracer_rk3_adj(y, x["QTRC"], x["QTRC0"], x["RHOQ_t"], x["DENS0"], x["DENS"],
                                                    x["mflx_hi"], x["num_diff_q"], *met, dt, fct, along)
        else:
            b = ctx.ATMOS_DYN_Tstep_tracer_fvm_heve_adj(y, x["QTRC"], x["QTRC0"], x["DENS0"], x["DENS"], x["mflx_hi"],
                                                       x["num_diff_q"], *met, dt, fct, along)
        terms = {k: float(np.sum(b[k] * d[k])) for k in ("QTRC", "QTRC0", "RHOQ_t", "num_diff")}
        rhs = sum(terms.values())
        print(f"  trial {trial}: <L dx, y> = {lhs:.15e}, <dx, L^T y> = {rhs:.15e} "
              f"({', '.join(f'{k} {v:.3e}' for k, v in terms.items())})", flush=True)
        assert all(np.isfinite(b[k]).all() for k in b)
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("fct,along", [(True, True), (True, False), (False, True)], ids=["fct_along", "fct_7pt", "nofct"])
@pytest.mark.parametrize("scheme,halo", [(3, 2), (4, 2), (1, 2), (2, 2), (5, 3), (6, 3)],
                         ids=["UD3", "CD4", "UD1", "CD2", "UD5_halo3", "CD6_halo3"])
def test_tracer_stage_adjoint_dot_product(hip, scheme, halo, fct, along):
    print(f"[tracer adjoint] stage scheme={scheme} halo={halo} fct={fct} along={along}", flush=True)
    case, metric, x = stage_case(scheme, halo, terrain=False)
    with hip.Context(case.grid, case.scheme, scheme) as ctx:
        dot_check(case, metric, x, ctx, False, fct, along, seed=scheme)


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_tracer_stage_adjoint_terrain(hip, scheme):
    print(f"[tracer adjoint] stage on general metrics, scheme={scheme}", flush=True)
    case, metric, x = stage_case(scheme, 2, terrain=True)
    with hip.Context(case.grid, case.scheme, scheme) as ctx:
        dot_check(case, metric, x, ctx, False, True, True, seed=20 + scheme)


@pytest.mark.parametrize("scheme,halo", [(3, 2), (4, 2), (5, 3)], ids=["UD3", "CD4", "UD5_halo3"])
def test_tracer_rk3_adjoint_dot_product(hip, scheme, halo):
    print(f"[tracer adjoint] RK3 integration scheme={scheme} halo={halo}", flush=True)
    case, metric, x = stage_case(scheme, halo, terrain=False)
    with hip.Context(case.grid, case.scheme, scheme) as ctx:
        dot_check(case, metric, x, ctx, True, True, True, seed=30 + scheme)


def test_tracer_adjoint_rejects_koren(hip):
    print("[tracer adjoint] UD3KOREN1993 rejected", flush=True)
    case, metric, x = stage_case(31, 2, terrain=False)
    mapf_xy = np.asfortranarray(metric["MAPF"][:, :, :, 0])
    with hip.Context(case.grid, case.scheme, 31) as ctx:
        with pytest.raises(RuntimeError, match="UD3KOREN1993"):
            ctx.ATMOS_DYN_Tstep_tracer_fvm_heve_adj(x["QTRC"], x["QTRC"], x["QTRC0"], x["DENS0"], x["DENS"],
                                                    x["mflx_hi"], x["num_diff_q"], metric["GSQRT"], mapf_xy,
                                                    metric["CDZ"], metric["RCDZ"], metric["RCDX"], metric["RCDY"],
                                                    2.0, True, True)
