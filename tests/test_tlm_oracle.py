"""The oracle's tangent-linear model (scale_oracle.c on dual numbers, oracle/
dual.h) and the dot-product test of the oracle's adjoint against it.

* The dual build computes the oracle's own values, bit for bit (3 large
  steps with FCT tracers, UD3 and CD4).
* Its tangent is the derivative: central differences of the plain oracle
  agree with it to 1e-7 at eps = 1e-7 (larger steps cross the upwind kinks).
* <L dx, y> = <dx, L^T y> for 3 large steps, L from the TLM and L^T from the
  oracle adjoint (scale_oracle_adj.c), to 1e-12 relative, on several seeded
  directions and weights.
* The tracer stage on dual numbers (R+/R- of the FCT limiter frozen) computes
  the oracle's values bit for bit, and its tangent is the derivative of the
  stage with R+/R- fixed at the oracle's own ratios (central differences,
  1e-7), not the derivative of the free limiter.
"""
import ctypes as C

import numpy as np
import pytest

from scale_amd import inputs

O = pytest.importorskip("pyoracle")
T = pytest.importorskip("pytlm")

STATE = O.STATE


def case_of(scheme, seed=3):
    case = inputs.make_case(kmax=10, imax=8, jmax=8, v0=2.0, momz_noise=0.05, noise=0.1, ntracer=1,
                            tracer_kind=["tophat"], fct=True, scheme=scheme)
    g = case.grid
    rng = np.random.default_rng(seed)
    for k, a in case.state.items():
        a[g.KS:g.KE + 1] *= 1.0 + 1e-3 * rng.standard_normal(a[g.KS:g.KE + 1].shape)
        inputs.fill_ghost_and_halo(g, a, k)
    return case


def direction(case, rng):
    """A perturbation of the interior prognostics (MOMZ below KE), filled into
    the ghost levels and halo as the model fills the state."""
    g = case.grid
    d = {}
    for k in STATE:
        a = np.zeros(g.shape, order="F")
        ke = g.KE - 1 if k == "MOMZ" else g.KE
        sl = (slice(g.KS, ke + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        a[sl] = rng.standard_normal(a[sl].shape) * np.abs(case.state[k]).max()
        d[k] = inputs.fill_ghost_and_halo(g, a, k)
    return d


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_dual_oracle_values_are_the_oracle(scheme):
    case = case_of(scheme)
    m = O.Model(case)
    m.step(3)
    t = T.TangentModel(case, case.state, direction(case, np.random.default_rng(1)))
    t.step(3)
    vals = t.values()
    for k in STATE:
        assert np.array_equal(vals[k], m.state[k], equal_nan=True), k
    q, _ = t.qtrc()
    assert np.array_equal(q, m.qtrc, equal_nan=True)


def test_tangent_is_the_derivative():
    """Central differences of the plain oracle along dx agree with L dx to
    1e-7 (max-norm relative, per field) at eps = 1e-7 over 1 and 2 large
    steps.  Larger steps eps*dx move some face velocity across zero, where
    the upwind switch sign(vel) makes the step non-differentiable: from
    eps = 1e-6 up the difference quotient then misses L dx by ~1e-3 at the
    points next to that face, whatever eps (no convergence), which is the
    kink, not the tangent."""
    case = case_of(3)
    d = direction(case, np.random.default_rng(2))
    for nst in (1, 2):
        t = T.TangentModel(case, case.state, d)
        t.step(nst)
        tan = t.tangents()

        def run(eps):
            m = O.Model(case)
            m.state = {k: np.asfortranarray(case.state[k] + eps * d[k]) for k in STATE}
            m.step(nst)
            return m.state

        eps = 1e-7
        p, mm = run(eps), run(-eps)
        for k in STATE:
            err = np.abs((p[k] - mm[k]) / (2 * eps) - tan[k]).max() / np.abs(tan[k]).max()
            print(f"{nst} step(s), {k}: max |FD - L dx| / max |L dx| = {err:.2e}")
            assert err < 1e-7, (nst, k, err)


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_dot_product_oracle_adjoint(scheme):
    """<L dx, y> against <dx, L^T y> over 3 large steps (FCT tracer on)."""
    case = case_of(scheme)
    g = case.grid
    N = 3
    traj = [{k: np.array(v, order="F", copy=True) for k, v in case.state.items()}]
    for _ in range(N):
        m = O.Model(case)
        m.state = {k: np.array(v, order="F", copy=True) for k, v in traj[-1].items()}
        m.step(1)
        traj.append(m.state)
    rng = np.random.default_rng(11 + scheme)
    for trial in range(3):
        d = direction(case, rng)
        y = {k: np.asfortranarray(rng.standard_normal(g.shape)) for k in STATE}
        t = T.TangentModel(case, case.state, d)
        t.step(N)
        tan = t.tangents()
        lhs = sum(float(np.sum(tan[k] * y[k])) for k in STATE)
        xb = y
        for n in reversed(range(N)):
            m = O.Model(case)
            m.state = traj[n]
            xb = O.adj_large_step(m, xb)
        grad = inputs.interior_gradient(g, xb)
        rhs = sum(float(np.sum(grad[k] * d[k])) for k in STATE)
        print(f"scheme {scheme} trial {trial}: <L dx, y> = {lhs:.15e}, <dx, L^T y> = {rhs:.15e}")
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


# ---- tracer stage with the FCT limiter frozen ----------------------------------

def tracer_inputs(seed=9, scheme=3):
    """One tracer stage's inputs on the tophat tracer (the limiter active):
    random face mass fluxes, a perturbed QTRC0, num_diff_q and RHOQ_t."""
    case = inputs.make_case(kmax=10, imax=8, jmax=8, ntracer=2, tracer_kind=["blob", "tophat"], fct=True,
                            scheme_tracer=scheme)
    g = case.grid
    rng = np.random.default_rng(seed)
    mflx = np.asfortranarray(rng.standard_normal(g.shape + (3,)) * np.array([0.3, 8.0, 5.0]))
    for n in range(3):
        inputs.halo_fill(g, mflx[..., n])
    q = np.asfortranarray(case.qtrc[..., 1])
    q0 = np.asfortranarray(q * (1.0 + 0.01 * rng.standard_normal(g.shape)))
    inputs.halo_fill(g, q0)
    nd = np.asfortranarray(1e-5 * rng.standard_normal(g.shape + (3,)))
    rq = np.asfortranarray(1e-6 * rng.standard_normal(g.shape))
    dens = np.asfortranarray(case.state["DENS"] * (1.0 + 1e-3 * rng.standard_normal(g.shape)))
    inputs.halo_fill(g, dens)
    return case, dict(QTRC=q, QTRC0=q0, RHOQ_t=rq, DENS0=case.state["DENS"], DENS=dens, mflx_hi=mflx, num_diff_q=nd)


def tracer_directions(case, rng, x):
    g = case.grid
    return {"QTRC": rng.standard_normal(g.shape) * np.abs(x["QTRC"]).max(),
            "QTRC0": rng.standard_normal(g.shape) * np.abs(x["QTRC0"]).max(),
            "RHOQ_t": rng.standard_normal(g.shape) * 1e-6,
            "num_diff": rng.standard_normal(g.shape + (3,)) * 1e-5}


@pytest.mark.parametrize("fct,along", [(True, True), (True, False), (False, True)], ids=["fct_along", "fct_7pt", "nofct"])
def test_dual_tracer_stage_values_are_the_oracle(fct, along):
    case, x = tracer_inputs()
    ref = np.full(case.grid.shape, np.nan, order="F")
    O.tstep_tracer(case, ref, x["QTRC"], x["QTRC0"], x["RHOQ_t"], x["DENS0"], x["DENS"], x["mflx_hi"],
                   x["num_diff_q"], 2.0, fct, along)
    d = tracer_directions(case, np.random.default_rng(4), x)
    val, _ = T.tracer_stage_tangent(case, **x, dtl=2.0, fct=fct, along_stream=along, tangents=d)
    assert np.array_equal(val, ref, equal_nan=True)


@pytest.mark.parametrize("along", [True, False], ids=["along", "7pt"])
def test_frozen_limiter_tangent(along):
    """The tracer stage's tangent with R+/R- frozen is the derivative of the
    stage with R+/R- fixed: the oracle's own ratios at the base point are
    captured (orc_fct_save_r) and imposed (orc_fct_fixed_r) on central
    differences of the plain oracle, which then agree with the tangent to
    1e-7 at eps = 1e-7 (the sign switch of each anti-diffusive flux is the
    only kink left).  Without the imposed ratios the differences miss the
    frozen tangent by far more where the limiter is active, so the check
    tells a frozen limiter from a differentiated one."""
    case, x = tracer_inputs()
    g = case.grid
    lib = O.lib()
    save = np.zeros(2 * int(np.prod(g.shape)))
    sp = C.c_void_p.in_dll(lib, "orc_fct_save_r")
    fp = C.c_void_p.in_dll(lib, "orc_fct_fixed_r")
    dt = 2.0

    def run(eps, dd):
        out = np.full(g.shape, np.nan, order="F")
        xs = dict(x)
        xs["QTRC"] = np.asfortranarray(x["QTRC"] + eps * dd["QTRC"])
        xs["QTRC0"] = np.asfortranarray(x["QTRC0"] + eps * dd["QTRC0"])
        xs["RHOQ_t"] = np.asfortranarray(x["RHOQ_t"] + eps * dd["RHOQ_t"])
        xs["num_diff_q"] = np.asfortranarray(x["num_diff_q"] + eps * dd["num_diff"])
        O.tstep_tracer(case, out, xs["QTRC"], xs["QTRC0"], xs["RHOQ_t"], x["DENS0"], x["DENS"], x["mflx_hi"],
                       xs["num_diff_q"], dt, True, along)
        return out

    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    dd = tracer_directions(case, np.random.default_rng(5), x)
    _, tan = T.tracer_stage_tangent(case, **x, dtl=dt, fct=True, along_stream=along, tangents=dd)
    eps = 1e-7
    sp.value = save.ctypes.data
    try:
        run(0.0, dd)
    finally:
        sp.value = None
    limited = np.nanmin(save) < 1.0
    assert limited, "the limiter should be active on the tophat"
    fp.value = save.ctypes.data
    try:
        fd = (run(eps, dd) - run(-eps, dd)) / (2 * eps)
    finally:
        fp.value = None
    err = np.abs(fd[sl] - tan[sl]).max() / np.abs(tan[sl]).max()
    free = (run(eps, dd) - run(-eps, dd)) / (2 * eps)
    err_free = np.abs(free[sl] - tan[sl]).max() / np.abs(tan[sl]).max()
    print(f"along={along}: frozen FD vs tangent {err:.2e}; unfrozen FD vs tangent {err_free:.2e}")
    assert err < 1e-7, err
    assert err_free > 1e-4, err_free


def terrain_stage_case(scheme, kmax=10, imax=8, jmax=8):
    """A HEVI stage on random smooth GSQRT, J13G, J23G and MAPF (every metric
    term of the stage non-trivial), with Coriolis."""
    import dataclasses

    import helpers as H
    case = H.small_case(kmax=kmax, imax=imax, jmax=jmax, seed=77, scheme=scheme, ntracer=0)
    case = dataclasses.replace(case, metric=H.terrain_metrics(case, seed=8))
    inp = H.stage_inputs(case, seed=9, coriolis=True)
    return case, inp


def stage_directions(case, rng):
    """Directions of every input the stage adjoint differentiates, over the
    whole arrays (halos and ghost levels included: the stage reads them)."""
    g = case.grid
    t = {}
    for k in STATE:
        for key in (k + "0", k):
            t[key] = np.asfortranarray(rng.standard_normal(g.shape) * (np.abs(case.state[k]).max() + 1.0))
    t["num_diff"] = np.asfortranarray(rng.standard_normal(g.shape + (5, 3)))
    for key in ("DPRES0", "RT2P", "REF_rhot"):
        t[key] = np.asfortranarray(rng.standard_normal(g.shape))
    return t


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_dual_stage_on_terrain(scheme):
    """The HEVI stage on dual numbers over terrain metrics: its values are the
    oracle's bit for bit, and its tangent matches central differences of the
    plain oracle stage (1e-7 relative at eps = 1e-7 of each input's scale)."""
    case, inp = terrain_stage_case(scheme)
    g = case.grid
    dt = 0.5
    ref = O.tstep_short_hevi(case, inp, dt)
    rng = np.random.default_rng(12)
    t = stage_directions(case, rng)
    val, tan = T.stage_tangent(case, inp, dt, tangents=t)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    for k in T.STAGE_OUT:
        assert np.array_equal(val[k][sl], ref[k][sl]), k
    eps = 1e-7

    def shifted(sign):
        return {key: (np.asfortranarray(v + sign * eps * t[key]) if key in t else v) for key, v in inp.items()}

    fp = O.tstep_short_hevi(case, shifted(+1), dt)
    fm = O.tstep_short_hevi(case, shifted(-1), dt)
    for k in T.STAGE_OUT:
        fd = (fp[k][sl] - fm[k][sl]) / (2 * eps)
        scale = np.abs(tan[k][sl]).max()
        assert scale > 0, k
        err = np.abs(fd - tan[k][sl]).max() / scale
        assert err < 1e-7, (k, err)
