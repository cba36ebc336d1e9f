"""Pins of the CPU oracle (oracle/liboracle.so) against facts that do not come
from re-typing its own formulas: textbook closed forms of the stencils, an
independent linear solver, the discrete implicit equations of the HEVI column
(as the reference's own check_equation states them, HEVI.F90:1131-1271),
exact conservation, symmetry of the discretisation under x<->y, translation
invariance, and the reference unit tests' invariants
(scalelib/test/unit/test_atmos_dyn.f90: test_const :382-454, test_conserve
:456-592, test_cwc :594-670, test_fctminmax :672-763, test_undef :307-380) and
its halo known-answer test (scalelib/test/unit/test_comm.f90:254-330).
"""
import numpy as np
import pytest

import helpers as H
import pyoracle as po
from scale_amd import inputs

UD1, UD3, CD4, UD3K = 1, 3, 4, 31


# ---------------------------------------------------------------------------
# face-value stencils: finite-volume reconstructions from cell averages
def cell_averages(coef, xs, h):
    """Exact cell averages of the polynomial sum c_n x^n over [x-h/2, x+h/2]."""
    out = []
    for x in xs:
        a, b = x - h / 2, x + h / 2
        out.append(sum(c * (b ** (n + 1) - a ** (n + 1)) / (n + 1) for n, c in enumerate(coef)) / h)
    return out


def poly(coef, x):
    return sum(c * x ** n for n, c in enumerate(coef))


# UD3KOREN1993 (ud3Koren1993.F90): Koren's (1993) limited third-order upwind
# reconstruction.  Pins: constancy, exactness for linear data (limiter value 1
# at the smoothness ratio r = 1), the TVD bound (face value between the two
# cells it separates), first order at extrema, and the mirror symmetry
# between the two wind directions (which fixes the limiter's argument order).
@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("vel", [3.0, -3.0])
def test_koren_constant_field(kind, vel):
    assert po.face_value(UD3K, kind, 1.25, 1.25, 1.25, 1.25, vel) == 1.25


@pytest.mark.parametrize("seed", range(4))
def test_koren_exact_for_linear_data(seed):
    rng = np.random.default_rng(seed)
    a, d = rng.normal(), rng.uniform(0.5, 2.0) * rng.choice([-1, 1])
    qm1, q0, q1, q2 = (a + d * n for n in range(4))
    # upwind from the left: q0 + d/2, upwind from the right: q1 - d/2; both the face midpoint
    assert po.face_value(UD3K, 0, qm1, q0, q1, q2, 1.0) == pytest.approx(a + 1.5 * d, rel=1e-14, abs=1e-14)
    assert po.face_value(UD3K, 0, qm1, q0, q1, q2, -1.0) == pytest.approx(a + 1.5 * d, rel=1e-14, abs=1e-14)


def test_koren_tvd_bound_and_extrema():
    rng = np.random.default_rng(11)
    for _ in range(2000):
        qm1, q0, q1, q2 = rng.normal(size=4)
        for vel in (1.0, -1.0):
            v = po.face_value(UD3K, 0, qm1, q0, q1, q2, vel)
            lo, hi = min(q0, q1), max(q0, q1)
            assert lo - 1e-14 <= v <= hi + 1e-14
        # a local maximum at the upwind cell: first order (the upwind value)
        if q0 > max(qm1, q1):
            assert po.face_value(UD3K, 0, qm1, q0, q1, q2, 1.0) == q0


def test_koren_wind_mirror_symmetry():
    rng = np.random.default_rng(12)
    for _ in range(500):
        qm1, q0, q1, q2 = rng.normal(size=4)
        assert po.face_value(UD3K, 0, qm1, q0, q1, q2, -2.0) == po.face_value(UD3K, 0, q2, q1, q0, qm1, 2.0)


def test_koren_boundary_faces():
    """Lowest face: upwind from below is first order (the KS value); highest
    face: upwind from above is first order (the KE value).  The limited side
    of the highest face differences toward the boundary cell as
    ud3Koren1993.F90:242-249 is written, so on linear data it returns the
    value half a cell below q0 rather than the face midpoint."""
    rng = np.random.default_rng(13)
    for _ in range(200):
        q0, q1, q2 = rng.normal(size=3)
        assert po.face_value(UD3K, 1, 0.0, q0, q1, q2, 1.0) == q0
        assert po.face_value(UD3K, 2, q0, q1, q2, 0.0, -1.0) == q2
    a, d = 0.3, 0.5
    assert po.face_value(UD3K, 2, a, a + d, a + 2 * d, 0.0, 1.0) == pytest.approx(a + 0.5 * d, rel=1e-14)


@pytest.mark.parametrize("seed", range(5))
def test_cd4_exact_for_cubics(seed):
    rng = np.random.default_rng(seed)
    coef = rng.normal(size=4)
    h = 0.7
    qm1, q0, q1, q2 = cell_averages(coef, [-h, 0.0, h, 2 * h], h)
    face = poly(coef, h / 2)
    for vel in (1.0, -1.0):
        v = po.face_value(CD4, 0, qm1, q0, q1, q2, vel)
        assert v == pytest.approx(face, rel=1e-13, abs=1e-13)


@pytest.mark.parametrize("seed", range(5))
def test_ud3_exact_for_quadratics_and_upwind_biased(seed):
    rng = np.random.default_rng(seed)
    coef = rng.normal(size=3)
    h = 0.5
    qm1, q0, q1, q2 = cell_averages(coef, [-h, 0.0, h, 2 * h], h)
    face = poly(coef, h / 2)
    assert po.face_value(UD3, 0, qm1, q0, q1, q2, 1.0) == pytest.approx(face, rel=1e-13, abs=1e-13)
    assert po.face_value(UD3, 0, qm1, q0, q1, q2, -1.0) == pytest.approx(face, rel=1e-13, abs=1e-13)
    # for a cubic the two upwind choices differ and each uses its upwind stencil only
    coef4 = rng.normal(size=4)
    qm1, q0, q1, q2 = cell_averages(coef4, [-h, 0.0, h, 2 * h], h)
    up = po.face_value(UD3, 0, qm1, q0, q1, q2, 2.5)
    assert up == pytest.approx((-qm1 + 5 * q0 + 2 * q1) / 6, rel=1e-14)
    dn = po.face_value(UD3, 0, qm1, q0, q1, q2, -2.5)
    assert dn == pytest.approx((2 * q0 + 5 * q1 - q2) / 6, rel=1e-14)
    # the upwind value does not depend on the far downwind cell and vice versa
    # (mathematically; the reference's grouping cancels it to round-off)
    assert po.face_value(UD3, 0, qm1, q0, q1, q2 + 10.0, 1.0) == pytest.approx(up, rel=1e-13)
    assert po.face_value(UD3, 0, qm1 + 10.0, q0, q1, q2, -1.0) == pytest.approx(dn, rel=1e-13)


def test_ud1_is_donor_cell():
    assert po.face_value(UD1, 0, 9.0, 1.25, 3.5, 9.0, 2.0) == 1.25
    assert po.face_value(UD1, 0, 9.0, 1.25, 3.5, 9.0, -2.0) == 3.5
    # sign(1,-0.0) = -1: a negative zero velocity picks the downwind cell as the reference does
    assert po.face_value(UD1, 0, 9.0, 1.25, 3.5, 9.0, -0.0) == 3.5


def test_boundary_faces():
    rng = np.random.default_rng(3)
    coef = rng.normal(size=3)
    h = 1.0
    q0, q1, q2 = cell_averages(coef, [0.0, h, 2 * h], h)
    face = poly(coef, h / 2)
    # UD3 near the bottom: second order from above (vel > 0), third order from below (vel < 0)
    assert po.face_value(UD3, 1, 0, q0, q1, q2, 1.0) == pytest.approx(0.5 * (q0 + q1), rel=1e-15)
    assert po.face_value(UD3, 1, 0, q0, q1, q2, -1.0) == pytest.approx(face, rel=1e-13)
    assert po.face_value(CD4, 1, 0, q0, q1, q2, -1.0) == pytest.approx(0.5 * (q0 + q1), rel=1e-15)
    qm1, q0, q1 = cell_averages(coef, [-h, 0.0, h], h)
    assert po.face_value(UD3, 2, qm1, q0, q1, 0, 1.0) == pytest.approx(face, rel=1e-13)
    assert po.face_value(UD3, 2, qm1, q0, q1, 0, -1.0) == pytest.approx(0.5 * (q0 + q1), rel=1e-15)


# ---------------------------------------------------------------------------
def test_thomas_matches_numpy_solve():
    case = H.small_case(kmax=30, imax=4, jmax=4)
    g = case.grid
    rng = np.random.default_rng(0)
    n = g.KE - g.KS  # KMAX-1 unknowns
    F1 = np.zeros(g.KA); F2 = np.zeros(g.KA); F3 = np.zeros(g.KA)
    rows = range(g.KS, g.KE)
    for k in rows:
        F1[k] = rng.uniform(-1, 1); F3[k] = rng.uniform(-1, 1); F2[k] = 3.0 + rng.uniform(0, 1)
    A = np.zeros((n, n))
    for r, k in enumerate(rows):
        A[r, r] = F2[k]
        if r + 1 < n:
            A[r, r + 1] = F1[k]
        if r > 0:
            A[r, r - 1] = F3[k]
    rhs = rng.normal(size=n)
    x = po.solve_direct(g, rhs, F1, F2, F3)
    assert np.allclose(x, np.linalg.solve(A, rhs), rtol=1e-13, atol=1e-14)


# ---------------------------------------------------------------------------
def _uniform_column_case(stretched=None):
    """Horizontally uniform state with a theta perturbation profile and w = 0;
    num_diff = 0 so the explicit tendencies vanish and the stage reduces to
    the vertical implicit solve."""
    case = inputs.make_case(kmax=24, imax=6, jmax=6, noise=0.0, bubble=False, u0=0.0, v0=0.0, ntracer=0,
                            stretched=stretched)
    g = case.grid
    z = np.linspace(0, 1, g.KA)
    rhot = case.state["RHOT"] * (1.0 + 0.002 * np.sin(7 * z))[:, None, None]
    case.state["RHOT"] = np.asfortranarray(rhot)
    return case


@pytest.mark.parametrize("stretched", [None, (60.0, 300.0)], ids=["uniform_dz", "stretched_dz"])
def test_hevi_column_satisfies_implicit_equations(stretched):
    """The solved column satisfies the discrete implicit equations for rho,
    rho*theta and rho*w (the system check_equation verifies, HEVI.F90:1131-1271),
    written here from the equations, not from the solver's bands."""
    case = _uniform_column_case(stretched)
    g = case.grid
    prep = inputs.large_step_prep(case)
    # RT2P/DPRES0 from a different state so the pressure perturbation is nonzero
    rhot_ref = np.asfortranarray(case.state["RHOT"] * 0.999)
    prep["REF_rhot"] = rhot_ref
    inp = {}
    for v in H.VARS:
        inp[v] = case.state[v]; inp[v + "0"] = case.state[v]; inp[v + "_t"] = g.zeros()
    inp.update(prep)
    inp["num_diff"] = np.zeros(g.shape + (5, 3), order="F")
    inp["DDIV"] = g.zeros()
    inp["REF_dens"] = np.asfortranarray(case.ref["REF_dens"] * 0.998)
    dt = 3.0
    out = po.tstep_short_hevi(case, inp, dt)
    m = case.metric
    KS, KE = g.KS, g.KE
    i, j = g.IS + 1, g.JS + 2
    C = out["MOMZ_RK"][:, i, j]  # MOMZ0 = MOMZ = 0 so MOMZ_RK = C
    Dn = out["DENS_RK"][:, i, j]
    Rn = out["RHOT_RK"][:, i, j]
    D = case.state["DENS"][:, i, j]
    R = case.state["RHOT"][:, i, j]
    P = R / D
    cdz, rcdz, rfdz = m["CDZ"], m["RCDZ"], m["RFDZ"]
    # theta at w-levels: UD3 with sign(1, w=0) = +1 (UD3.F90:98-153)
    PT = np.zeros(g.KA)
    for k in range(KS + 1, KE - 1):
        PT[k] = (-P[k - 1] + 5 * P[k] + 2 * P[k + 1]) / 6
    PT[KS] = 0.5 * (P[KS] + P[KS + 1])
    PT[KE - 1] = (2 * P[KE] + 5 * P[KE - 1] - P[KE - 2]) / 6
    Cf = np.zeros(g.KA)
    Cf[KS:KE] = C[KS:KE]
    for k in range(KS, KE + 1):
        lhs = (Dn[k] - D[k]) / dt
        rhs = -(Cf[k] - Cf[k - 1]) * rcdz[k]
        assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-15)
        lhs = (Rn[k] - R[k]) / dt
        rhs = -(Cf[k] * PT[k] - Cf[k - 1] * PT[k - 1]) * rcdz[k]
        assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-13)
    rt2p, dp0, rr = prep["RT2P"][:, i, j], prep["DPRES0"][:, i, j], rhot_ref[:, i, j]
    refd = inp["REF_dens"][:, i, j]
    dpres_n = dp0 + rt2p * (Rn - rr)
    for k in range(KS, KE):
        # explicit density anomaly at the w-level with the F2H weights (HEVI.F90:62,716-717) ...
        f1 = cdz[k] / (cdz[k] + cdz[k + 1])
        f2 = cdz[k + 1] / (cdz[k] + cdz[k + 1])
        # ... and the implicit density increment as the mass-weighted mean of the
        # two layers, which is what B = g dt^2 / (CDZ(k+1)+CDZ(k)) encodes
        # (HEVI.F90:698-709, comment at :711); equal to F2H on a uniform grid
        w1 = cdz[k + 1] / (cdz[k] + cdz[k + 1])
        w2 = cdz[k] / (cdz[k] + cdz[k + 1])
        lhs = Cf[k] / dt
        rhs = (-(dpres_n[k + 1] - dpres_n[k]) * rfdz[k]
               - 9.80665 * (f1 * (D[k + 1] - refd[k + 1]) + f2 * (D[k] - refd[k]))
               - 9.80665 * (w1 * (Dn[k + 1] - D[k + 1]) + w2 * (Dn[k] - D[k])))
        assert lhs == pytest.approx(rhs, rel=1e-8, abs=1e-10 * max(1.0, abs(rhs)))


# ---------------------------------------------------------------------------
def _stage(case, inp, dt=0.7, scheme=None):
    return po.tstep_short_hevi(case, inp, dt, scheme=scheme)


@pytest.mark.parametrize("scheme", [UD3, CD4])
def test_stage_conserves_mass_and_rhot(scheme):
    case = H.small_case(seed=31)
    inp = H.stage_inputs(case, seed=8, tend=False)
    out = _stage(case, inp, scheme=scheme)
    g = case.grid
    m = po.Model(case)
    for v in ("DENS", "RHOT"):
        a0 = m.mass(np.asfortranarray(inp[v + "0"]))
        a1 = m.mass(np.asfortranarray(np.where(np.isnan(out[v + "_RK"]), 0.0, out[v + "_RK"])))
        assert abs(a1 - a0) <= 2e-15 * abs(a0) * g.KA, v


def _transpose_case_inputs(case, inp):
    """x<->y mirror: swap i and j, MOMX and MOMY, and the x/y components."""
    T = lambda a: np.asfortranarray(np.swapaxes(a, 1, 2))
    ti = {}
    for k, v in inp.items():
        if not isinstance(v, np.ndarray) or v.ndim < 3:
            ti[k] = v
            continue
        ti[k] = T(v)
    for s in ("", "0", "_t"):
        ti["MOMX" + s], ti["MOMY" + s] = T(inp["MOMY" + s]), T(inp["MOMX" + s])
    nd = T(inp["num_diff"])
    nd2 = nd.copy(order="F")
    # swap the XDIR/YDIR components and the MOMX/MOMY variables (index 2 <-> 3)
    for v in range(5):
        vv = {2: 3, 3: 2}.get(v, v)
        nd2[..., vv, 0] = nd[..., v, 0]
        nd2[..., vv, 1] = nd[..., v, 2]
        nd2[..., vv, 2] = nd[..., v, 1]
    ti["num_diff"] = np.asfortranarray(nd2)
    return ti


def test_stage_xy_mirror_symmetry():
    case = H.small_case(kmax=16, imax=10, jmax=10, seed=17)
    inp = H.stage_inputs(case, seed=4)
    out = _stage(case, inp)
    tcase = inputs.Case(**{**case.__dict__})
    tinp = _transpose_case_inputs(case, inp)
    tout = _stage(tcase, tinp)
    T = lambda a: np.swapaxes(a, 1, 2)
    g = case.grid
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    pairs = [("DENS_RK", "DENS_RK"), ("RHOT_RK", "RHOT_RK"), ("MOMZ_RK", "MOMZ_RK"), ("MOMX_RK", "MOMY_RK"), ("MOMY_RK", "MOMX_RK")]
    for a, b in pairs:
        x, y = out[a][sl], T(tout[b])[sl]
        assert np.allclose(x, y, rtol=1e-13, atol=1e-13 * np.abs(x).max()), (a, b, np.abs(x - y).max())
    mx, my = out["mflx_hi"], T(tout["mflx_hi"])
    assert np.allclose(mx[..., 0][sl], my[..., 0][sl], rtol=1e-13, atol=1e-16)
    assert np.allclose(mx[..., 1][sl], my[..., 2][sl], rtol=1e-13, atol=1e-16)
    assert np.allclose(mx[..., 2][sl], my[..., 1][sl], rtol=1e-13, atol=1e-16)


def test_stage_translation_invariance():
    case = H.small_case(kmax=12, imax=10, jmax=8, seed=2)
    inp = H.stage_inputs(case, seed=6)
    out = _stage(case, inp)
    g = case.grid
    si, sj = 3, 5

    def shift(a):
        a = np.array(a, order="F", copy=True)
        core = a[:, g.IS:g.IE + 1, g.JS:g.JE + 1]
        a[:, g.IS:g.IE + 1, g.JS:g.JE + 1] = np.roll(np.roll(core, si, axis=1), sj, axis=2)
        inputs.halo_fill(g, a) if a.ndim == 3 else None
        return a

    sinp = {}
    for k, v in inp.items():
        if isinstance(v, np.ndarray) and v.ndim == 3 and v.shape == g.shape:
            sinp[k] = shift(v)
        elif k == "num_diff":
            nd = np.array(v, order="F", copy=True)
            for a_ in range(5):
                for d_ in range(3):
                    nd[..., a_, d_] = shift(v[..., a_, d_])
            sinp[k] = nd
        else:
            sinp[k] = v
    sout = _stage(case, sinp)
    for k in ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK"):
        a = out[k][g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]
        b = sout[k][g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]
        assert np.array_equal(np.roll(np.roll(a, si, axis=1), sj, axis=2), b), k


def test_stage_rest_state_stays_at_rest():
    """Reference state, no motion, no perturbation: every tendency vanishes."""
    case = inputs.make_case(kmax=20, imax=8, jmax=8, noise=0.0, bubble=False, u0=0.0, ntracer=0)
    g = case.grid
    inp = {}
    for v in H.VARS:
        inp[v] = case.state[v]; inp[v + "0"] = case.state[v]; inp[v + "_t"] = g.zeros()
    prep = inputs.large_step_prep(case)
    inp.update(prep)
    inp["num_diff"] = po.numfilter_coef(case, case.state, 1.0)
    inp["DDIV"] = g.zeros()
    inp["REF_dens"] = case.ref["REF_dens"]
    # reference pressure consistent with RHOT so DPRES0 = 0
    inp["DPRES0"] = g.zeros()
    out = _stage(case, inp, dt=2.0)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    assert np.abs(out["MOMZ_RK"][sl]).max() < 1e-10
    assert np.abs(out["MOMX_RK"][sl]).max() < 1e-10
    assert np.abs(out["MOMY_RK"][sl]).max() < 1e-10
    assert np.allclose(out["DENS_RK"][sl], case.state["DENS"][sl], rtol=1e-14, atol=0)


# ---------------------------------------------------------------------------
# HEVE (scale_atmos_dyn_tstep_short_fvm_heve.F90): the fully explicit short
# step.  Its tendency is the dt -> 0 limit of the HEVI stage started from
# state = state0 (the implicit corrections are O(dt)), so HEVI -- pinned above
# against its own implicit equations -- pins every HEVE term, vertical pressure
# gradient, buoyancy and the MOMZ part of the vertical mass flux included.
def _heve_inputs(case, seed=8):
    inp = H.stage_inputs(case, seed=seed, tend=True, coriolis=True)
    for v in H.VARS:
        inp[v] = inp[v + "0"]
    return inp


@pytest.mark.parametrize("scheme", [UD3, CD4])
@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_heve_is_the_explicit_limit_of_hevi(scheme, terrain):
    case = H.small_case(seed=31)
    if terrain:
        case = inputs.Case(**{**case.__dict__, "metric": H.terrain_metrics(case)})
    g = case.grid
    inp = _heve_inputs(case)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    he = po.tstep_short_heve(case, inp, 1.0, scheme=scheme)
    dts = (4.0e-3, 2.0e-3, 1.0e-3)
    hi = [po.tstep_short_hevi(case, inp, dt, scheme=scheme) for dt in dts]
    for v in H.VARS:
        fe = he[v + "_RK"][sl] - inp[v + "0"][sl]
        s = np.abs(fe).max()
        f = [(o[v + "_RK"][sl] - inp[v + "0"][sl]) / dt for o, dt in zip(hi, dts)]
        e = [np.abs(x - fe).max() / s for x in f]
        if v in ("MOMX", "MOMY"):  # explicit in both schemes: equal up to rounding
            assert max(e) < 1e-11, (v, e)
            continue
        # first order in dt ...
        assert 1.8 < e[0] / e[1] < 2.2 and 1.8 < e[1] / e[2] < 2.2, (v, e)
        r = [np.abs(2 * f[n + 1] - f[n] - fe).max() / s for n in range(2)]
        assert r[1] < 0.1 * e[2], (v, r, e)
        if scheme == CD4:
            # ... and the Richardson extrapolates converge on HEVE at second
            # order, so the limit is HEVE's tendency itself.  Not so for UD3:
            # HEVI upwinds the explicit vertical theta flux by the sign of its
            # explicit mass flux (J13/J23/filter part, HEVI.F90:604-608) and the
            # implicit part by MOMZ (:691), HEVE by the sign of their sum, so
            # the UD3 limits differ by that switching term.
            assert 3.5 < r[0] / r[1] < 4.5, (v, r, e)


@pytest.mark.parametrize("scheme", [UD3, CD4])
def test_heve_conserves_mass_and_rhot(scheme):
    case = H.small_case(seed=31)
    inp = H.stage_inputs(case, seed=8, tend=False)
    out = po.tstep_short_heve(case, inp, 0.7, scheme=scheme)
    g = case.grid
    m = po.Model(case)
    for v in ("DENS", "RHOT"):
        a0 = m.mass(np.asfortranarray(inp[v + "0"]))
        a1 = m.mass(np.asfortranarray(np.where(np.isnan(out[v + "_RK"]), 0.0, out[v + "_RK"])))
        assert abs(a1 - a0) <= 2e-15 * abs(a0) * g.KA, v


@pytest.mark.parametrize("fct", [False, True], ids=["plain", "fct_momentum_t"])
def test_heve_xy_mirror_symmetry(fct):
    case = H.small_case(kmax=16, imax=10, jmax=10, seed=17)
    inp = H.stage_inputs(case, seed=4)
    kw = dict(fct_momentum=fct, fct_t=fct, along_stream=False)
    out = po.tstep_short_heve(case, inp, 0.7, **kw)
    tout = po.tstep_short_heve(inputs.Case(**{**case.__dict__}), _transpose_case_inputs(case, inp), 0.7, **kw)
    T = lambda a: np.swapaxes(a, 1, 2)
    g = case.grid
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    for a, b in [("DENS_RK", "DENS_RK"), ("RHOT_RK", "RHOT_RK"), ("MOMZ_RK", "MOMZ_RK"),
                 ("MOMX_RK", "MOMY_RK"), ("MOMY_RK", "MOMX_RK")]:
        x, y = out[a][sl], T(tout[b])[sl]
        assert np.allclose(x, y, rtol=1e-13, atol=1e-13 * np.abs(x).max()), (a, b, np.abs(x - y).max())


def test_heve_divergence_damping_matches_hevi():
    """With unit map factors HEVI's horizontal damping rows (HEVI.F90:912,
    :1032) reduce to HEVE's (HEVE.F90:933, :1182) and are explicit in both, so
    the change that divdmp_coef makes to MOMX/MOMY is the same; the vertical
    row enters HEVI's implicit solve and agrees in the dt -> 0 limit."""
    case = H.small_case(seed=5)
    g = case.grid
    inp = _heve_inputs(case, seed=3)
    rng = np.random.Generator(np.random.PCG64(9))
    ddiv = g.zeros()
    ddiv[:, g.IS:g.IE + 1, g.JS:g.JE + 1] = 1e-3 * rng.standard_normal((g.KA, g.IE - g.IS + 1, g.JE - g.JS + 1))
    inputs.halo_fill(g, ddiv)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    slz = (slice(g.KS, g.KE), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    dt = 1.0e-3
    delta = {}
    for name, fn in (("heve", po.tstep_short_heve), ("hevi", po.tstep_short_hevi)):
        o = [fn(case, dict(inp, DDIV=ddiv, divdmp_coef=c), dt, scheme=CD4) for c in (0.0, 0.1)]
        delta[name] = {v: o[1][v + "_RK"] - o[0][v + "_RK"] for v in ("MOMZ", "MOMX", "MOMY")}
    for v in ("MOMX", "MOMY"):
        x, y = delta["heve"][v][sl], delta["hevi"][v][sl]
        assert np.abs(x).max() > 1e-8, v
        assert np.allclose(x, y, rtol=1e-9, atol=1e-9 * np.abs(x).max()), v
    x, y = delta["heve"]["MOMZ"][slz], delta["hevi"]["MOMZ"][slz]
    assert np.abs(x).max() > 1e-8
    assert np.abs(x - y).max() < 1e-2 * np.abs(x).max()


def _momentum_shear_inputs(axis):
    """A horizontally uniform atmosphere at hydrostatic rest with the
    non-divergent flow u = 12, v = V(x) (axis 0) or v = 12, u = U(y) (axis 1),
    V and U top-hats between -4 and 6 m/s.  The density stays unchanged, the
    pressure gradient and Coriolis terms vanish, and the momentum equation of
    the top-hat component is linear advection by the uniform component."""
    case = inputs.make_case(kmax=20, imax=12, jmax=10, noise=0.0, bubble=False, u0=0.0, ntracer=0)
    g = case.grid
    inp = {}
    for v in H.VARS:
        inp[v] = np.array(case.state[v], order="F", copy=True)
        inp[v + "_t"] = g.zeros()
    inp.update(inputs.large_step_prep(case))
    inp["num_diff"] = np.asfortranarray(np.zeros((g.KA, g.IA, g.JA, 5, 3)))
    inp["DDIV"] = g.zeros()
    inp["REF_dens"] = case.ref["REF_dens"]
    inp["DPRES0"] = g.zeros()
    D = inp["DENS"]
    uni, hat = g.zeros(), g.zeros()
    uni[:] = 12.0 * D
    hat[:] = -4.0 * D
    b = (slice(None), slice(g.IS + 3, g.IS + 7), slice(None)) if axis == 0 else \
        (slice(None), slice(None), slice(g.JS + 3, g.JS + 7))
    hat[b] = 6.0 * D[b]
    for a in (uni, hat):
        inputs.halo_fill(g, a)
    inp["MOMX"], inp["MOMY"] = (uni, hat) if axis == 0 else (hat, uni)
    inp["MOMZ"] = g.zeros()
    for v in H.VARS:
        inp[v + "0"] = inp[v]
    return case, inp


@pytest.mark.parametrize("along", [True, False], ids=["along_stream", "7point"])
@pytest.mark.parametrize("scheme", [UD3, CD4])
@pytest.mark.parametrize("axis", [0, 1], ids=["v_of_x", "u_of_y"])
def test_heve_fct_momentum_bounds_horizontal_velocity(axis, scheme, along):
    """FLAG_FCT_MOMENTUM (HEVE.F90:386-450, 716-810, 958-1057, 1207-1309):
    for the advected top-hat the density is steady and the flow
    non-divergent, so the UD1 solution is monotone and the limited velocity
    stays inside the initial range to round-off; without FCT the scheme
    over- and undershoots by more than half a m/s."""
    case, inp = _momentum_shear_inputs(axis)
    g = case.grid
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    name = "MOMY" if axis == 0 else "MOMX"
    D = inp["DENS0"][sl]
    q0 = inp[name + "0"][sl] / D
    plain = po.tstep_short_heve(case, inp, 8.0, scheme=scheme)
    fct = po.tstep_short_heve(case, inp, 8.0, scheme=scheme, fct_momentum=True, along_stream=along)
    assert np.array_equal(plain["DENS_RK"][sl], inp["DENS0"][sl])
    qp = plain[name + "_RK"][sl] / D
    qf = fct[name + "_RK"][sl] / D
    assert qp.max() > q0.max() + 0.5 and qp.min() < q0.min() - 0.5  # the control
    assert qf.max() <= q0.max() + 1e-12 and qf.min() >= q0.min() - 1e-12
    assert np.abs(qf - q0).max() > 1.0  # the top-hat moved
    other = "MOMX" if axis == 0 else "MOMY"
    assert np.allclose(fct[other + "_RK"][sl] / D, 12.0, rtol=1e-13, atol=0)


@pytest.mark.parametrize("scheme", [UD3, CD4])
def test_heve_fct_momentum_limits_vertical_velocity(scheme):
    """A 3 m/s box of w in the uniform flow u = 12: without FCT w over- and
    undershoots its range by 0.2 m/s or more; with it the excursion is below
    1 % of that (the residual comes from the box's own vertical mass
    divergence, which changes DENS_RK under the box's edges)."""
    case, inp = _momentum_shear_inputs(0)
    g = case.grid
    D = inp["DENS0"]
    Dz = 0.5 * (D[:-1] + D[1:])
    mz = g.zeros()
    b = (slice(g.KS + 5, g.KS + 11), slice(g.IS + 3, g.IS + 8), slice(g.JS + 2, g.JS + 6))
    mz[b] = 3.0 * Dz[b]
    inputs.halo_fill(g, mz)
    inp["MOMZ"] = inp["MOMZ0"] = mz
    inp["MOMY"] = inp["MOMY0"] = g.zeros()
    inp["wdamp_coef"] = np.zeros(g.KA)
    slz = (slice(g.KS, g.KE), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    q0 = (mz[:-1] / Dz)[slz]
    ex = []
    for fct in (False, True):
        o = po.tstep_short_heve(case, inp, 8.0, scheme=scheme, fct_momentum=fct)
        Dn = np.array(o["DENS_RK"], order="F", copy=True)
        inputs.halo_fill(g, Dn)
        q = (o["MOMZ_RK"][:-1] / (0.5 * (Dn[:-1] + Dn[1:])))[slz]
        ex.append((q.max() - q0.max(), q0.min() - q.min()))
    assert min(ex[0]) > 0.15, ex
    assert max(ex[1]) < 1e-2 * min(ex[0]), ex


@pytest.mark.parametrize("terrain", [False, True], ids=["flat", "terrain"])
def test_heve_fct_momentum_is_conservative(terrain):
    """The FCT correction of each momentum component is a flux divergence:
    weighted by each point's volume over its map factors it sums to zero
    over the periodic domain (the limited anti-diffusive fluxes vanish at
    the bottom and top), it is not zero, and DENS_RK, RHOT_RK, mflx_hi and
    tflx_hi are unchanged."""
    case = H.small_case(seed=31)
    if terrain:
        case = inputs.Case(**{**case.__dict__, "metric": H.terrain_metrics(case)})
    g, m = case.grid, case.metric
    inp = _heve_inputs(case)
    plain = po.tstep_short_heve(case, inp, 2.0, scheme=UD3)
    fct = po.tstep_short_heve(case, inp, 2.0, scheme=UD3, fct_momentum=True)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    for k in ("DENS_RK", "RHOT_RK", "tflx_hi"):
        assert np.array_equal(plain[k][sl], fct[k][sl]), k
    assert np.array_equal(plain["mflx_hi"][sl], fct["mflx_hi"][sl])
    GS, MP = m["GSQRT"], m["MAPF"]
    ii, jj = slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1)
    for v, kk, c, mc, rz, rx, ry in (("MOMZ", slice(g.KS, g.KE), 1, 0, "RFDZ", "RCDX", "RCDY"),
                                     ("MOMX", slice(g.KS, g.KE + 1), 4, 1, "RCDZ", "RFDX", "RCDY"),
                                     ("MOMY", slice(g.KS, g.KE + 1), 5, 2, "RCDZ", "RCDX", "RFDY")):
        d = fct[v + "_RK"][kk, ii, jj] - plain[v + "_RK"][kk, ii, jj]
        w = GS[kk, ii, jj, c] / (MP[ii, jj, 0, mc] * MP[ii, jj, 1, mc])[None] \
            / (m[rz][kk][:, None, None] * m[rx][ii][None, :, None] * m[ry][jj][None, None, :])
        scale = np.abs(d * w).sum()
        assert np.abs(d).max() > 1e-6 * np.abs(plain[v + "_RK"][kk, ii, jj]).max(), v
        assert abs((d * w).sum()) < 1e-12 * scale, (v, (d * w).sum(), scale)


def _theta_tophat_inputs():
    case = H.small_case(seed=41, u0=15.0, v0=-8.0)
    g = case.grid
    inp = H.stage_inputs(case, seed=3, tend=False)
    for key in ("RHOT0", "RHOT"):
        a = np.array(inp[key], order="F", copy=True)
        a[g.KS + 3:g.KS + 9, g.IS + 2:g.IS + 7, g.JS + 2:g.JS + 6] *= 1.1
        inputs.halo_fill(g, a)
        inp[key] = a
    return case, inp


@pytest.mark.parametrize("along", [True, False], ids=["along_stream", "7point"])
def test_heve_fct_t_bounds_theta_and_conserves(along):
    """FLAG_FCT_T (HEVE.F90:1388-1475): with a 10 % theta top-hat the HEVE
    stage without FCT over- and undershoots the initial theta range; with
    it, theta = RHOT_RK/DENS_RK stays inside that range (the reference's
    test_fctminmax property for theta), RHOT stays conserved, and every
    other output is unchanged."""
    case, inp = _theta_tophat_inputs()
    g = case.grid
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    th0 = (inp["RHOT0"] / inp["DENS0"])[sl]
    m = po.Model(case)
    plain = po.tstep_short_heve(case, inp, 2.0, scheme=CD4)
    fct = po.tstep_short_heve(case, inp, 2.0, scheme=CD4, fct_t=True, along_stream=along)
    th_p = plain["RHOT_RK"][sl] / plain["DENS_RK"][sl]
    th_f = fct["RHOT_RK"][sl] / fct["DENS_RK"][sl]
    assert th_p.max() > th0.max() + 0.1 and th_p.min() < th0.min() - 0.01  # the control
    assert th_f.max() <= th0.max() + 1e-12 and th_f.min() >= th0.min() - 1e-12
    r0 = m.mass(np.asfortranarray(inp["RHOT0"]))
    r1 = m.mass(np.asfortranarray(np.where(np.isnan(fct["RHOT_RK"]), 0.0, fct["RHOT_RK"])))
    assert abs(r1 - r0) <= 2e-15 * r0 * g.KA
    for k in ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "tflx_hi"):
        assert np.array_equal(plain[k][sl], fct[k][sl]), k


def test_heve_rest_state_stays_at_rest():
    case = inputs.make_case(kmax=20, imax=8, jmax=8, noise=0.0, bubble=False, u0=0.0, ntracer=0)
    g = case.grid
    inp = {}
    for v in H.VARS:
        inp[v] = case.state[v]; inp[v + "0"] = case.state[v]; inp[v + "_t"] = g.zeros()
    inp.update(inputs.large_step_prep(case))
    inp["num_diff"] = po.numfilter_coef(case, case.state, 1.0)
    inp["DDIV"] = g.zeros()
    inp["REF_dens"] = case.ref["REF_dens"]
    inp["DPRES0"] = g.zeros()
    out = po.tstep_short_heve(case, inp, 2.0)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    for v in ("MOMZ", "MOMX", "MOMY"):
        assert np.abs(out[v + "_RK"][sl]).max() < 1e-10, v
    assert np.allclose(out["DENS_RK"][sl], case.state["DENS"][sl], rtol=1e-14, atol=0)


def test_heve_model_conserves_mass_and_converges_to_hevi():
    """Large steps with the HEVE short step conserve mass to round-off, and
    over the same 1 s of model time the HEVE and HEVI integrations differ
    by O(dt): halving the step halves the difference (they differ only in
    the implicit treatment of the vertical acoustic terms)."""
    diffs = []
    for dt, n in ((0.1, 10), (0.05, 20)):
        case = inputs.make_case(kmax=20, imax=12, jmax=10, noise=0.01, ntracer=1, tracer_kind=["blob"], fct=True,
                                bubble_r=1500.0, bubble_zc=1500.0, momz_noise=0.05, dt=dt)
        g = case.grid
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        me, mi = po.Model(case, short="heve"), po.Model(case, short="hevi")
        m0 = me.mass(me.state["DENS"])
        me.step(n)
        mi.step(n)
        assert abs(me.mass(me.state["DENS"]) - m0) <= 1e-13 * m0
        diffs.append({k: np.abs(me.state[k][sl] - mi.state[k][sl]).max() / np.abs(mi.state[k][sl]).max()
                      for k in H.VARS if k != "MOMY"})
    for k, d in diffs[0].items():
        assert 1.6 < d / diffs[1][k] < 2.4, (k, d, diffs[1][k])


# ---------------------------------------------------------------------------
def test_numfilter_two_grid_wave_closed_form():
    """For q = A(-1)^i in x the 3rd-difference flux is 8A(-1)^i/dx^3 and
    num_diff = ND_COEF*dx/(2 DT) * A(-1)^i (density, x direction)."""
    case = inputs.make_case(kmax=10, imax=8, jmax=6, noise=0.0, bubble=False, u0=0.0, ntracer=0)
    g = case.grid
    A = 1.0e-3
    st = {k: np.array(v, order="F", copy=True) for k, v in case.state.items()}
    sign = (-1.0) ** np.arange(g.IA)
    st["DENS"] = np.asfortranarray(case.ref["REF_dens"] + A * sign[None, :, None])
    dt, coef = 2.0, 1.0e-4
    nd = po.numfilter_coef(case, st, dt, nd_coef=coef)
    want = coef * g.dx / (2 * dt) * A * sign
    for k in range(g.KS + 2, g.KE + 1):
        assert np.allclose(nd[k, g.IS:g.IE + 1, g.JS:g.JE + 1, 0, 1], want[g.IS:g.IE + 1, None], rtol=1e-12)
    # a field that varies only in x has no y or z flux
    assert np.abs(nd[g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1, 0, 2]).max() < 1e-18


def test_numfilter_annihilates_quadratics():
    case = inputs.make_case(kmax=10, imax=12, jmax=6, noise=0.0, bubble=False, u0=0.0, ntracer=0)
    g = case.grid
    st = {k: np.array(v, order="F", copy=True) for k, v in case.state.items()}
    i = np.arange(g.IA, dtype=float)
    st["DENS"] = np.asfortranarray(case.ref["REF_dens"] + (1e-3 + 2e-4 * i - 3e-5 * i * i)[None, :, None])
    nd = po.numfilter_coef(case, st, 1.0)
    inner = nd[g.KS:g.KE + 1, g.IS + 1:g.IE - 1, g.JS:g.JE + 1, 0, 1]
    assert np.abs(inner).max() < 1e-15


def test_numfilter_q_is_the_theta_filter_on_the_tracer():
    """num_diff_q of a tracer equals the RHOT rows of num_diff when the tracer
    is theta and REF_pott = 0: the same third differences, density weights and
    surface rows (COMMON.F90:880-962 and :1108-1184), bit for bit."""
    case = inputs.make_case(kmax=12, imax=10, jmax=10, v0=2.0, momz_noise=0.05, noise=0.1, ntracer=0)
    g = case.grid
    st = case.state
    case0 = inputs.Case(**{**case.__dict__, "ref": {**case.ref, "REF_pott": np.zeros(g.shape, order="F")}})
    nd = po.numfilter_coef(case0, st, 1.5, nd_coef=3e-4, nd_sfc_fact=0.4)
    theta = np.asfortranarray(st["RHOT"] / st["DENS"] - 0.0)
    ndq = po.numfilter_coef_q(case, st["DENS"], theta, 1.5, nd_coef=3e-4, nd_sfc_fact=0.4)
    for d in range(3):
        assert np.array_equal(ndq[..., d], nd[..., 4, d]), d
    assert np.abs(ndq).max() > 0.0


def test_numfilter_q_qv_filters_the_anomaly():
    """QV with ND_USE_RS differences QTRC - REF_qv (COMMON.F90:1085-1098)."""
    case = inputs.make_case(kmax=12, imax=10, jmax=8, noise=0.1, ntracer=0)
    g = case.grid
    rng = np.random.default_rng(4)
    q = np.asfortranarray(1e-3 * (1.0 + 0.1 * rng.standard_normal(g.shape)))
    ref = np.asfortranarray(1e-3 * (1.0 + 0.1 * rng.standard_normal(g.shape)))
    dens = case.state["DENS"]
    a = po.numfilter_coef_q(case, dens, q, 1.0, is_qv=True, REF_qv=ref)
    b = po.numfilter_coef_q(case, dens, np.asfortranarray(q - ref), 1.0)
    assert np.array_equal(a, b)
    assert np.abs(po.numfilter_coef_q(case, dens, q, 1.0, is_qv=True, REF_qv=q)).max() == 0.0


def test_model_tracer_filter_conserves_tracer_mass():
    """With the tracer filter on, tracer mass stays conserved (flux form)."""
    case = inputs.make_case(kmax=12, imax=10, jmax=8, noise=0.05, ntracer=1, tracer_kind="blob", fct=True)
    m = po.Model(case, nd_coef_q=1e-2)
    q0 = m.mass(np.asfortranarray(m.qtrc[..., 0] * m.state["DENS"]))
    m.step(4)
    q1 = m.mass(np.asfortranarray(m.qtrc[..., 0] * m.state["DENS"]))
    assert abs(q1 - q0) <= 1e-13 * abs(q0)
    ref = po.Model(case)
    ref.step(4)
    assert np.abs(ref.qtrc - m.qtrc).max() > 0.0  # the filter acted


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_fct_weights_partition_unity(seed):
    rng = np.random.default_rng(seed)
    rw, ru, rv = rng.normal(size=3) * rng.choice([1e-3, 1.0, 1e3], size=3)
    f = po.get_fact_fct(rw, ru, rv)
    assert f.sum() == pytest.approx(1.0, abs=1e-14)
    assert (f >= -1e-15).all()


def test_fct_weights_axis_aligned_and_zero_flow():
    f = po.get_fact_fct(0.0, 2.0, 0.0)  # pure +x flow: all weight on fact(0,1,0)
    assert f[0, 2, 1] == 1.0 and np.count_nonzero(f) == 1
    f = po.get_fact_fct(0.0, 0.0, 0.0)  # no flow: fact(0,0,0) = 1
    assert f[0, 1, 1] == 1.0 and np.count_nonzero(f) == 1
    f = po.get_fact_fct(-3.0, 0.0, 0.0)  # vertical flow: fact(1,0,0)
    assert f[1, 1, 1] == 1.0 and np.count_nonzero(f) == 1


def _consistent_tracer_setup(seed=0, dt=4.0):
    """DENS = DENS0 - dt*div(mflx) so the continuity equation holds discretely."""
    case = inputs.make_case(kmax=16, imax=12, jmax=10, noise=0.0, bubble=False, ntracer=0)
    g = case.grid
    rng = np.random.default_rng(seed)
    mflx = np.zeros(g.shape + (3,), order="F")
    mflx[g.KS:g.KE, :, :, 0] = 0.5 * rng.normal(size=(g.KE - g.KS, g.IA, g.JA))
    mflx[g.KS:g.KE + 1, :, :, 1] = 10 * rng.normal(size=(g.KE - g.KS + 1, g.IA, g.JA))
    mflx[g.KS:g.KE + 1, :, :, 2] = 10 * rng.normal(size=(g.KE - g.KS + 1, g.IA, g.JA))
    for n in range(3):
        inputs.halo_fill(g, mflx[..., n])
    m = case.metric
    D0 = case.state["DENS"]
    D = D0.copy(order="F")
    KS, KE, IS, IE, JS, JE = g.KS, g.KE, g.IS, g.IE, g.JS, g.JE
    for k in range(KS, KE + 1):
        div = ((mflx[k, IS:IE + 1, JS:JE + 1, 0] - mflx[k - 1, IS:IE + 1, JS:JE + 1, 0]) * m["RCDZ"][k]
               + (mflx[k, IS:IE + 1, JS:JE + 1, 1] - mflx[k, IS - 1:IE, JS:JE + 1, 1]) * m["RCDX"][IS:IE + 1, None]
               + (mflx[k, IS:IE + 1, JS:JE + 1, 2] - mflx[k, IS:IE + 1, JS - 1:JE, 2]) * m["RCDY"][None, JS:JE + 1])
        D[k, IS:IE + 1, JS:JE + 1] = D0[k, IS:IE + 1, JS:JE + 1] - dt * div
    inputs.halo_fill(g, D)
    return case, mflx, D0, D, dt


@pytest.mark.parametrize("fct", [False, True])
def test_tracer_constant_stays_constant(fct):
    """test_cwc: a uniform tracer is preserved when the mass fluxes satisfy continuity."""
    case, mflx, D0, D, dt = _consistent_tracer_setup(1)
    g = case.grid
    q = np.full(g.shape, 0.01, order="F")
    out = np.full(g.shape, np.nan, order="F")
    po.tstep_tracer(case, out, q, q, g.zeros(), D0, D, mflx, np.zeros(g.shape + (3,), order="F"), dt, fct)
    inner = out[g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]
    assert np.allclose(inner, 0.01, rtol=1e-13, atol=0)


@pytest.mark.parametrize("along", [True, False])
def test_fct_keeps_two_level_field_in_bounds(along):
    """test_fctminmax: q in {0, 0.1} stays within [0, 0.1] up to 1e-14 relative."""
    case, mflx, D0, D, dt = _consistent_tracer_setup(2)
    g = case.grid
    rng = np.random.default_rng(5)
    q = np.asfortranarray(np.where(rng.uniform(size=g.shape) < 0.5, 0.0, 0.1))
    inputs.halo_fill(g, q)
    out = np.full(g.shape, np.nan, order="F")
    nd = np.zeros(g.shape + (3,), order="F")
    po.tstep_tracer(case, out, q, q, g.zeros(), D0, D, mflx, nd, dt, True, along)
    inner = out[g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]
    eps = 1.0e-14
    assert inner.min() >= 0.0 - eps and inner.max() <= 0.1 * (1 + eps)
    # without the limiter the high-order scheme overshoots (the test is not vacuous)
    out2 = np.full(g.shape, np.nan, order="F")
    po.tstep_tracer(case, out2, q, q, g.zeros(), D0, D, mflx, nd, dt, False)
    inner2 = out2[g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]
    assert inner2.min() < -1e-6 or inner2.max() > 0.1 + 1e-6


@pytest.mark.parametrize("fct", [False, True])
def test_tracer_mass_conserved(fct):
    case, mflx, D0, D, dt = _consistent_tracer_setup(3)
    g = case.grid
    rng = np.random.default_rng(9)
    q = np.asfortranarray(np.abs(rng.normal(size=g.shape)) * 1e-3)
    inputs.halo_fill(g, q)
    out = np.full(g.shape, np.nan, order="F")
    po.tstep_tracer(case, out, q, q, g.zeros(), D0, D, mflx, np.zeros(g.shape + (3,), order="F"), dt, fct)
    m = po.Model(case)
    before = m.mass(np.asfortranarray(q * D0))
    after = m.mass(np.asfortranarray(np.nan_to_num(out) * D))
    assert abs(after - before) <= 1e-14 * abs(before)


@pytest.mark.parametrize("scheme", [po.CD2, po.UD1], ids=["CD2", "UD1"])
@pytest.mark.parametrize("axis", [0, 1, 2], ids=["x", "y", "z"])
def test_tracer_cd2_ud1_fourier_closed_form(scheme, axis):
    """The CD2 and UD1 tracer schemes (cd2.F90, ud1.F90) advecting a sine wave
    by a uniform mass flux rho*u over a uniform density: one step is
    q - dt u/dx * (q(n+1) - q(n-1))/2 for CD2 and q - dt u/dx * (q(n) - q(n-1))
    for UD1 (u > 0), i.e. for q = sin(2 pi n / N) the closed forms
    sin(2 pi/N) cos(2 pi n/N) and 2 sin(pi/N) cos(2 pi (n - 1/2)/N).  In z the
    levels two away from the top and bottom follow the same form."""
    case = inputs.make_case(kmax=16, imax=12, jmax=10, noise=0.0, bubble=False, u0=0.0, ntracer=0)
    g, m = case.grid, case.metric
    D = np.full(g.shape, 1.2, order="F")
    kk, ii, jj = np.meshgrid(np.arange(g.KA), np.arange(g.IA), np.arange(g.JA), indexing="ij")
    n = (ii - g.IS, jj - g.JS, kk - g.KS)[axis]
    N = (g.IE - g.IS + 1, g.JE - g.JS + 1, g.KE - g.KS + 1)[axis]
    q = np.asfortranarray(1.0 + 1e-3 * np.sin(2 * np.pi * n / N))
    u, dt = 3.0, 0.5
    mflx = np.zeros(g.shape + (3,), order="F")
    comp = (1, 2, 0)[axis]
    mflx[..., comp] = 1.2 * u
    if axis == 2:  # no flux through the bottom and the top
        mflx[:g.KS, :, :, 0] = 0.0
        mflx[g.KE:, :, :, 0] = 0.0
    nd = np.zeros(g.shape + (3,), order="F")
    out = np.full(g.shape, np.nan, order="F")
    po.tstep_tracer(case, out, q, q, g.zeros(), D, D, mflx, nd, dt, False, scheme=scheme)
    dx = (1.0 / m["RCDX"][g.IS], 1.0 / m["RCDY"][g.JS], 1.0 / m["RCDZ"][g.KS])[axis]
    if scheme == po.CD2:
        dq = np.sin(2 * np.pi / N) * np.cos(2 * np.pi * n / N)
    else:
        dq = 2 * np.sin(np.pi / N) * np.cos(2 * np.pi * (n - 0.5) / N)
    expect = q - dt * u / dx * 1e-3 * dq
    sl = [slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1)]
    if axis == 2:
        sl[0] = slice(g.KS + 2, g.KE - 1)
    sl = tuple(sl)
    assert np.abs(out[sl] - expect[sl]).max() < 1e-14, np.abs(out[sl] - expect[sl]).max()
    assert np.abs(out[sl] - q[sl]).max() > 1e-6  # the wave moved


# (scheme, kind, vel sign) -> highest polynomial degree reproduced exactly from
# cell averages: the stencil's order of accuracy minus one
_FACE6_DEGREE = {
    (5, 0, 1): 4, (5, 0, -1): 4, (6, 0, 1): 5, (6, 0, -1): 5,      # 5-point upwind / 6-point central
    (5, 1, 1): 1, (5, 1, -1): 2, (6, 1, 1): 1, (6, 1, -1): 1,      # KS: F2 / 3-point upwind
    (5, 3, 1): 2, (5, 3, -1): 4, (6, 3, 1): 3, (6, 3, -1): 3,      # KS+1: 3- / 5-point upwind, CD4
    (5, 2, 1): 2, (5, 2, -1): 1, (6, 2, 1): 1, (6, 2, -1): 1,      # KE-1
    (5, 4, 1): 4, (5, 4, -1): 2, (6, 4, 1): 3, (6, 4, -1): 3,      # KE-2
}


@pytest.mark.parametrize("key", sorted(_FACE6_DEGREE), ids=lambda k: f"s{k[0]}_kind{k[1]}_v{k[2]:+d}")
def test_face6_reproduces_polynomials_from_cell_averages(key):
    """The UD5 and CD6 face values (ud5.F90:236-302, cd6.F90:195-234) are
    finite-volume reconstructions: fed the cell averages of a polynomial on a
    unit grid they return its value at the face exactly up to the stencil's
    degree, and miss it one degree higher.  A wrong coefficient, a reversed
    stencil or an upwind side taken from the wrong sign of the velocity breaks
    one of these."""
    s, kind, sgn = key
    deg = _FACE6_DEGREE[key]
    x = np.arange(6, dtype=float)                       # cell centres of q[0..5]
    face = {0: 2.5, 1: 0.5, 3: 1.5, 2: 4.5, 4: 3.5}[kind]
    x0 = 2.5 if kind == 0 else 0.0                      # centre the monomials on the face region

    def cell_avg(p):
        return (((x + 0.5 - x0) ** (p + 1)) - ((x - 0.5 - x0) ** (p + 1))) / (p + 1)

    for p in range(deg + 2):
        got = po.face_value6(s, kind, cell_avg(p), 1.7 * sgn)
        exact = (face - x0) ** p
        if p <= deg:
            assert abs(got - exact) < 1e-12, (p, got, exact)
        else:
            assert abs(got - exact) > 1e-6, (p, got, exact)


@pytest.mark.parametrize("scheme", [po.UD5, po.CD6], ids=["UD5", "CD6"])
@pytest.mark.parametrize("fct", [False, True], ids=["nofct", "fct"])
def test_tracer_ud5_cd6_conserve_mass_and_constants(scheme, fct):
    """UD5/CD6 tracer steps on a grid with a lateral halo of 3: the flux form
    conserves the tracer mass sum(DENS q) to round-off over a periodic domain
    with closed top and bottom, and a constant field with a density that
    follows the mass-flux divergence stays constant."""
    case = H.small_case(kmax=12, imax=10, jmax=8, seed=7, halo=3)
    g = case.grid
    assert g.IS == 3 and g.IA == g.imax + 6
    rng = np.random.Generator(np.random.PCG64(17))
    mflx = np.asfortranarray(rng.standard_normal(g.shape + (3,)) * np.array([0.3, 8.0, 5.0]))
    mflx[:g.KS, :, :, 0] = 0.0
    mflx[g.KE:, :, :, 0] = 0.0
    for n in range(3):
        inputs.halo_fill(g, mflx[..., n])
    q = np.asfortranarray(1.0 + 0.1 * rng.standard_normal(g.shape))
    inputs.halo_fill(g, q)
    nd = np.asfortranarray(1e-5 * rng.standard_normal(g.shape + (3,)))
    for n in range(3):
        inputs.halo_fill(g, nd[..., n])
    D0 = case.state["DENS"]
    dt = 2.0
    # the density consistent with the fluxes: D = D0 - dt * div(mflx)
    m = case.metric
    D = D0.copy(order="F")
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    kz, ix, jy = (np.arange(g.KS, g.KE + 1)[:, None, None], np.arange(g.IS, g.IE + 1)[None, :, None],
                  np.arange(g.JS, g.JE + 1)[None, None, :])
    div = ((mflx[kz, ix, jy, 0] - mflx[kz - 1, ix, jy, 0]) * m["RCDZ"][kz]
           + (mflx[kz, ix, jy, 1] - mflx[kz, ix - 1, jy, 1]) * m["RCDX"][ix]
           + (mflx[kz, ix, jy, 2] - mflx[kz, ix, jy - 1, 2]) * m["RCDY"][jy])
    D[sl] = D0[sl] - dt * div
    inputs.halo_fill(g, D)
    out = np.full(g.shape, np.nan, order="F")
    po.tstep_tracer(case, out, q, q, g.zeros(), D0, D, mflx, nd, dt, fct, scheme=scheme)
    vol = 1.0 / (m["RCDZ"][kz] * m["RCDX"][ix] * m["RCDY"][jy])
    mass0 = (D0[sl] * q[sl] * vol).sum()
    mass1 = (D[sl] * out[sl] * vol).sum()
    assert abs(mass1 - mass0) < 1e-12 * abs(mass0)
    one = np.asfortranarray(np.ones(g.shape))
    po.tstep_tracer(case, out, one, one, g.zeros(), D0, D, mflx, np.zeros_like(nd), dt, fct, scheme=scheme)
    assert np.abs(out[sl] - 1.0).max() < 1e-12


def test_tracer_ud1_monotone_and_without_num_diff():
    """UD1 as the tracer scheme: with continuity-consistent fluxes a two-level
    field stays within its bounds without any limiter (the monotone scheme),
    and the num_diff argument does not enter (the UD1 routines take none,
    UD1.F90:135-249), while CD2 does use it."""
    case, mflx, D0, D, dt = _consistent_tracer_setup(2)
    g = case.grid
    rng = np.random.default_rng(6)
    q = np.asfortranarray(np.where(rng.uniform(size=g.shape) < 0.5, 0.0, 0.1))
    inputs.halo_fill(g, q)
    nd0 = np.zeros(g.shape + (3,), order="F")
    nd1 = np.asfortranarray(1e-4 * rng.standard_normal(g.shape + (3,)))
    inner = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    outs = {}
    for scheme in (po.UD1, po.CD2):
        for tag, nd in (("0", nd0), ("1", nd1)):
            o = np.full(g.shape, np.nan, order="F")
            po.tstep_tracer(case, o, q, q, g.zeros(), D0, D, mflx, nd, dt, False, scheme=scheme)
            outs[scheme, tag] = o[inner]
    u = outs[po.UD1, "0"]
    assert u.min() >= -1e-15 and u.max() <= 0.1 * (1 + 1e-14)
    assert np.array_equal(outs[po.UD1, "0"], outs[po.UD1, "1"])
    assert not np.array_equal(outs[po.CD2, "0"], outs[po.CD2, "1"])
    c = outs[po.CD2, "0"]
    assert c.min() < -1e-6 or c.max() > 0.1 + 1e-6  # the central scheme is not monotone


# ---------------------------------------------------------------------------
def test_halo_known_answer():
    """test_comm.f90:254-330 re-expressed for one periodic domain."""
    case = inputs.make_case(kmax=6, imax=7, jmax=5, ntracer=0)
    g = case.grid
    a = np.full(g.shape, np.nan, order="F")
    kk, ii, jj = np.meshgrid(np.arange(g.KA), np.arange(g.IA), np.arange(g.JA), indexing="ij")
    val = kk * 1e6 + ii * 1e3 + jj
    core = (slice(None), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    a[core] = val[core]
    po.halo_periodic(g, a)
    imax, jmax = g.IE - g.IS + 1, g.JE - g.JS + 1
    wrap_i = lambda i: i + imax if i < g.IS else (i - imax if i > g.IE else i)
    wrap_j = lambda j: j + jmax if j < g.JS else (j - jmax if j > g.JE else j)
    for i in range(g.IA):
        for j in range(g.JA):
            assert np.array_equal(a[:, i, j], val[:, wrap_i(i), wrap_j(j)])


# ---------------------------------------------------------------------------
def test_model_invariants():
    """test_conserve / test_undef over whole large steps (dynamics + FCT tracers)."""
    case = inputs.make_case(kmax=20, imax=12, jmax=12, noise=0.05, ntracer=2, tracer_kind=["blob", "tophat"], fct=True,
                            bubble_r=1500.0, bubble_zc=1500.0)
    m = po.Model(case)
    m0 = m.mass(m.state["DENS"])
    r0 = m.mass(m.state["RHOT"])
    q0 = [m.mass(np.asfortranarray(m.qtrc[..., iq] * m.state["DENS"])) for iq in range(2)]
    m.step(6)
    g = case.grid
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    for v in H.VARS:
        assert not np.isnan(m.state[v][sl]).any()
    assert abs(m.mass(m.state["DENS"]) - m0) <= 1e-13 * m0
    assert abs(m.mass(m.state["RHOT"]) - r0) <= 1e-13 * r0
    for iq in range(2):
        q1 = m.mass(np.asfortranarray(m.qtrc[..., iq] * m.state["DENS"]))
        assert abs(q1 - q0[iq]) <= 1e-13 * q0[iq]
    # FCT keeps the top-hat within its initial range
    qt = m.qtrc[..., 1][sl]
    assert qt.min() >= -1e-18 and qt.max() <= 1e-3 * (1 + 1e-13)


def test_model_horizontally_uniform_stays_uniform():
    """test_const: a horizontally uniform state stays horizontally uniform."""
    case = inputs.make_case(kmax=16, imax=8, jmax=8, noise=0.0, bubble=False, u0=7.0, v0=-2.0, ntracer=1,
                            tracer_kind="uniform", fct=True)
    m = po.Model(case)
    m.step(3)
    g = case.grid
    for v in H.VARS:
        a = m.state[v][g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]
        assert np.array_equal(a, np.broadcast_to(a[:, :1, :1], a.shape)), v


def test_model_constant_theta_preserved():
    """test_cwc: theta = const and q = const survive random momentum."""
    case = inputs.make_case(kmax=16, imax=10, jmax=8, noise=0.0, bubble=False, bv=0.0, u0=5.0, ntracer=1,
                            tracer_kind="uniform", fct=True, momz_noise=0.3)
    g = case.grid
    rng = np.random.default_rng(1)
    for v in ("MOMX", "MOMY"):
        a = case.state[v]
        a[:, g.IS:g.IE + 1, g.JS:g.JE + 1] += rng.normal(size=a[:, g.IS:g.IE + 1, g.JS:g.JE + 1].shape)
        inputs.halo_fill(g, a)
    m = po.Model(case)
    m.step(3)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    pott = m.state["RHOT"][sl] / m.state["DENS"][sl]
    assert np.allclose(pott, 300.0, rtol=1e-13, atol=0)
    assert np.allclose(m.qtrc[..., 0][sl], 1e-3, rtol=1e-13, atol=0)


# ---------------------------------------------------------------------------
# time integrators (scale_atmos_dyn_tinteg_short_rk3.F90:188-576,
# scale_atmos_dyn_tinteg_tracer_euler.F90:80-133)
def _explicit_large_step(tinteg_short, nstep, T=1.0):
    """One large step of T seconds in nstep short steps of the explicit HEVE
    stage (X_RK = X0 + dtrk f(X)) with the filter off: within a large step
    RT2P/DPRES0/REF_rhot are frozen, so the short steps integrate the ODE
    dX/dt = f(X) and the textbook order of the RK method shows."""
    case = inputs.make_case(kmax=16, imax=10, jmax=8, noise=0.01, ntracer=0, bubble_r=1500.0, bubble_zc=1500.0,
                            momz_noise=0.05, dt=T / nstep, scheme=CD4, v0=-2.0)
    m = po.Model(case, nd_coef=0.0, short="heve", dtl=T, tinteg_short=tinteg_short)
    m.step(1)
    return case, m.state


@pytest.mark.parametrize("tinteg,order", [("RK3", 3), ("RK3WS2002", 3), ("RK4", 4)])
def test_short_integrators_converge_at_their_order(tinteg, order):
    """Heun's RK3 (stages dt/3, 2dt/3, dt and (3 r1 + 3 r3 - 2 phi)/4),
    Wicher-Skamarock RK3 (dt/3, dt/2, dt) and RK4 are third, third and fourth
    order: against an RK4 run with 128 short steps, halving the step divides
    the error of every prognostic by about 2^order.  A wrong stage factor or
    combination weight drops the order to 1 or 2 (ratio 2 or 4)."""
    case, ref = _explicit_large_step("RK4", 128)
    g = case.grid
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    errs = []
    for n in (4, 8, 16):
        _, st = _explicit_large_step(tinteg, n)
        errs.append({k: np.abs(st[k][sl] - ref[k][sl]).max() / np.abs(ref[k][sl]).max() for k in H.VARS})
    lo = {3: 6.0, 4: 11.0}[order]
    hi = {3: 11.0, 4: 30.0}[order]
    for k in H.VARS:
        r = errs[1][k] / errs[2][k]
        assert lo < r < hi, (tinteg, k, [e[k] for e in errs])


def test_heun_and_ws2002_differ():
    """The two RK3 variants are different methods (same order, different
    error constants), so the integrator switch takes effect."""
    _, a = _explicit_large_step("RK3", 4)
    _, b = _explicit_large_step("RK3WS2002", 4)
    assert not np.array_equal(a["MOMX"], b["MOMX"])


@pytest.mark.parametrize("tinteg", ["RK3", "RK3WS2002"])
def test_rk3_model_conserves_mass(tinteg):
    """test_conserve over large steps with the RK3 short integrators (HEVI, FCT tracers)."""
    case = inputs.make_case(kmax=20, imax=12, jmax=12, noise=0.05, ntracer=1, tracer_kind=["tophat"], fct=True,
                            bubble_r=1500.0, bubble_zc=1500.0)
    m = po.Model(case, tinteg_short=tinteg)
    m0 = m.mass(m.state["DENS"])
    q0 = m.mass(np.asfortranarray(m.qtrc[..., 0] * m.state["DENS"]))
    m.step(4)
    assert abs(m.mass(m.state["DENS"]) - m0) <= 1e-13 * m0
    assert abs(m.mass(np.asfortranarray(m.qtrc[..., 0] * m.state["DENS"])) - q0) <= 1e-13 * q0


def _tracer_run(tinteg_tracer, n, T=8.0, fct=False):
    case = inputs.make_case(kmax=16, imax=12, jmax=10, noise=0.01, ntracer=1, tracer_kind=["blob"], bubble_r=1500.0,
                            bubble_zc=1500.0, momz_noise=0.05, dt=T / n, scheme=CD4, scheme_tracer=CD4, u0=12.0,
                            v0=-5.0, fct=fct)
    m = po.Model(case, tinteg_tracer=tinteg_tracer)
    q0 = m.mass(np.asfortranarray(m.qtrc[..., 0] * m.state["DENS"]))
    m.step(n)
    return case, m, q0


def test_tracer_euler_is_first_order():
    """Over fixed model time the forward-Euler tracer integrator differs from
    the third-order RK3WS2002 one by O(dt): halving the step halves the
    difference, while tracer mass is conserved to round-off."""
    diffs = []
    for n in (4, 8, 16):
        case, me, q0 = _tracer_run("EULER", n)
        _, mr, _ = _tracer_run("RK3WS2002", n)
        g = case.grid
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        q1 = me.mass(np.asfortranarray(me.qtrc[..., 0] * me.state["DENS"]))
        assert abs(q1 - q0) <= 1e-13 * q0
        diffs.append(np.abs(me.qtrc[..., 0][sl] - mr.qtrc[..., 0][sl]).max())
    assert 1.8 < diffs[0] / diffs[1] < 2.2 and 1.8 < diffs[1] / diffs[2] < 2.2, diffs


@pytest.mark.parametrize("fct", [False, True], ids=["nofct", "fct"])
def test_tracer_euler_fct_keeps_bounds(fct):
    """test_fctminmax for the Euler integrator: with FCT the top-hat tracer
    stays within its initial range; without it the same run leaves it."""
    case = inputs.make_case(kmax=16, imax=12, jmax=10, noise=0.01, ntracer=1, tracer_kind=["tophat"], bubble_r=1500.0,
                            bubble_zc=1500.0, momz_noise=0.05, dt=2.0, u0=12.0, v0=-5.0, fct=fct)
    m = po.Model(case, tinteg_tracer="EULER")
    m.step(5)
    g = case.grid
    qt = m.qtrc[..., 0][g.KS:g.KE + 1, g.IS:g.IE + 1, g.JS:g.JE + 1]
    inside = qt.min() >= -1e-18 and qt.max() <= 1e-3 * (1 + 1e-13)
    assert inside == fct, (qt.min(), qt.max())


# ---------------------------------------------------------------------------
# numerical filter: the closed-form response to 2-grid waves, every row
_ND_VARS = ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")
# levels where num_diff(:,:,:,v,d) is defined (COMMON.F90:541-962); others are 0
_ND_ROWS = {("DENS", 0): (-1, 0), ("MOMZ", 0): (1, -1), ("MOMX", 0): (0, -1), ("MOMY", 0): (0, -1),
            ("RHOT", 0): (-1, 0), ("MOMZ", 1): (0, -1), ("MOMZ", 2): (0, -1)}


@pytest.mark.parametrize("sfc", [1.0, 0.3])
@pytest.mark.parametrize("d", [0, 1, 2], ids=["z", "x", "y"])
@pytest.mark.parametrize("var", _ND_VARS)
def test_numfilter_two_grid_wave_every_row(var, d, sfc):
    """On a uniform grid a 2-grid wave A(-1)^m along direction d is an
    eigenfunction of the filter's third difference, the ghost-level mirror
    rows included (COMMON.F90:1600-1700): the difference is +-8A(-1)^m/D^3
    (+ for a field centred along d, whose difference sits on the faces; - for
    the field staggered along d), so num_diff = ND_COEF/(2^4 DT) D^4 times
    that, times the density at the flux point for the velocities and theta
    (uniform density rho0 here), and in x and y the surface rows KS and KS+1
    carry ND_SFC_FACT and (1 + ND_SFC_FACT)/2 (COMMON.F90:593-594).  Every
    other direction of the same variable is exactly 0.  A flipped sign, a
    shifted index, a wrong density weight, mirror row or surface factor
    changes some row by a factor other than 1."""
    case = inputs.make_case(kmax=12, imax=8, jmax=6, noise=0.0, bubble=False, u0=0.0, ntracer=0, dz=250.0)
    g = case.grid
    KS, KE = g.KS, g.KE
    rho0, th0, A = 1.1, 300.0, 1.0e-3
    st = {k: np.zeros(g.shape, order="F") for k in _ND_VARS}
    st["DENS"][:] = rho0
    st["RHOT"][:] = rho0 * th0
    ref = {"REF_dens": np.full(g.shape, rho0, order="F"), "REF_pott": np.full(g.shape, th0, order="F"),
           "REF_pres": case.ref["REF_pres"]}
    idx = np.indices(g.shape)[d]
    wave = (-1.0) ** idx
    if var == "DENS":
        st["DENS"] = np.asfortranarray(rho0 + A * wave)
    elif var == "RHOT":
        st["RHOT"] = np.asfortranarray(rho0 * (th0 + A * wave))
    else:
        st[var] = np.asfortranarray(rho0 * A * wave)
        if var == "MOMZ":
            st["MOMZ"][:KS] = 0.0
            st["MOMZ"][KE:] = 0.0
    case = inputs.Case(**{**case.__dict__, "ref": ref})
    dt, coef = 2.0, 1.0e-4
    nd = po.numfilter_coef(case, st, dt, nd_coef=coef, nd_sfc_fact=sfc)
    spacing = (250.0, g.dx, g.dy)[d]
    v = _ND_VARS.index(var)
    staggered = (var, d) in (("MOMZ", 0), ("MOMX", 1), ("MOMY", 2))
    weight = 1.0 if var == "DENS" else rho0
    want = (-1.0 if staggered else 1.0) * coef / (16.0 * dt) * spacing ** 4 * 8.0 * A / spacing ** 3 * weight * wave
    lo, hi = _ND_ROWS.get((var, d), (0, 0))
    ks, ke = KS + lo, KE + hi
    if d > 0:
        want = want.copy()
        want[KS] *= sfc
        want[KS + 1] *= (1.0 + sfc) * 0.5
    if (var, d) == ("MOMZ", 0):
        # the top row takes the wall value MOMZ(KE) = 0 in place of the wave's
        # (COMMON.F90:1696-1698): -3A - 3A - A instead of -8A
        want = want.copy()
        want[KE - 1] *= 7.0 / 8.0
    isl, jsl = slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1)
    got = nd[:, isl, jsl, v, d]
    assert np.allclose(got[ks:ke + 1], want[ks:ke + 1, isl, jsl], rtol=1e-10, atol=0), (var, d)
    assert not got[:ks].any() and not got[ke + 1:].any(), (var, d)
    for dd in range(3):
        if dd == d:
            continue
        other = nd[:, isl, jsl, v, dd]
        if var == "MOMZ" and dd == 0:
            # a horizontal wave in w is constant in z except for the wall value
            # MOMZ(KE) = 0, which the top row sees: -(w(KE-1))/dz^3 (COMMON.F90:1696-1698)
            top = -coef / (16.0 * dt) * 250.0 * A * rho0 * wave[KE - 1, isl, jsl]
            assert np.allclose(other[KE - 1], top, rtol=1e-10, atol=0), (var, d)
            other = np.delete(other, KE - 1, axis=0)
        assert not other.any(), (var, d, dd)


# ---------------------------------------------------------------------------
# ATMOS_DYN_divergence (COMMON.F90:1437-1515) and ATMOS_DYN_wdamp_setup (:323-381)
@pytest.mark.parametrize("stretched", [None, (100.0, 400.0)], ids=["uniform_dz", "stretched_dz"])
def test_divergence_of_a_polynomial_momentum_field(stretched):
    """MOMX = a x, MOMY = b y and MOMZ = c z (H - z) (zero at the surface and
    the top, as the walls require) have the divergence a + b + c (H - 2 z):
    the flux differences over a cell are exact for these polynomials, so every
    level KS..KE, the one-sided wall rows included, gives the closed form to
    round-off.  A sign or index slip in any term, or a wrong wall row, is O(1)."""
    case = inputs.make_case(kmax=12, imax=8, jmax=6, noise=0.0, bubble=False, u0=0.0, ntracer=0, dz=250.0,
                            stretched=stretched)
    g = case.grid
    KS, KE = g.KS, g.KE
    fz = case.fz  # FZ(0:KA): fz[k + 1] is the top face of level k
    zf = fz[1:]   # top face of each level (the w point)
    zc = 0.5 * (fz[1:] + fz[:-1])
    H = zf[KE]
    x_u = (np.arange(g.IA) - g.IS + 1.0) * g.dx   # u points: east face of cell i
    y_v = (np.arange(g.JA) - g.JS + 1.0) * g.dy
    a, b, c = 2.0e-3, -1.5e-3, 4.0e-7
    MOMX = np.asfortranarray(np.broadcast_to(a * x_u[None, :, None], g.shape).copy())
    MOMY = np.asfortranarray(np.broadcast_to(b * y_v[None, None, :], g.shape).copy())
    w = c * zf * (H - zf)
    w[:KS] = 0.0
    w[KE:] = 0.0
    MOMZ = np.asfortranarray(np.broadcast_to(w[:, None, None], g.shape).copy())
    dd = po.divergence(case, MOMZ, MOMX, MOMY)
    want = a + b + c * (H - 2.0 * zc)
    isl, jsl = slice(g.IS, g.IE + 2), slice(g.JS, g.JE + 2)
    for k in range(KS, KE + 1):
        assert np.allclose(dd[k, isl, jsl], want[k], rtol=1e-10, atol=1e-15), k


@pytest.mark.parametrize("height", [4000.0, 10000.0])
def test_wdamp_setup_properties(height):
    """The Rayleigh coefficient alpha/2 (1 - cos(pi (z - h)/(z_top - h)))
    at the w points: 0 at and below h, alpha = 1/tau exactly at the top face,
    alpha/2 where z is halfway, strictly increasing in between, ghost levels
    copied from KS / KE; h at or above the top switches it off."""
    case = inputs.make_case(kmax=40, imax=4, jmax=4, noise=0.0, bubble=False, ntracer=0, dz=500.0)
    g = case.grid
    KS, KE = g.KS, g.KE
    tau = 250.0
    wd = po.wdamp_setup(g, tau, height, case.fz)
    zf = case.fz[1:]
    above = zf[KS:KE + 1] > height
    assert not wd[KS:KE + 1][~above].any()
    assert wd[KE] == pytest.approx(1.0 / tau, rel=1e-15)
    ka = np.flatnonzero(above) + KS
    assert np.all(np.diff(wd[ka]) > 0.0)
    mid = 0.5 * (height + zf[KE])
    kmid = np.flatnonzero(np.isclose(zf, mid))
    assert kmid.size == 1 and wd[kmid[0]] == pytest.approx(0.5 / tau, rel=1e-14)
    assert np.all(wd[:KS] == wd[KS]) and np.all(wd[KE + 1:] == wd[KE])
    assert not po.wdamp_setup(g, tau, zf[KE], case.fz).any()


# ---------------------------------------------------------------------------
# HEVI horizontal momentum update (HEVI.F90:816-927, 936-1047): pressure
# gradient, Coriolis, map-factor metric term and momentum advection
def _quiet_inputs(case):
    """Stage inputs of a fluid at rest in the reference state: DENS = REF_dens,
    RHOT = REF_rhot (so DPRES = DPRES0 exactly), no momentum, no tendencies,
    no filter; the tests add one ingredient at a time."""
    g = case.grid
    prep = inputs.large_step_prep(case)
    inp = {}
    for v in H.VARS:
        base = case.state["DENS"] if v == "DENS" else (prep["REF_rhot"] if v == "RHOT" else g.zeros())
        inp[v + "0"] = np.array(base, order="F", copy=True)
        inp[v] = np.array(base, order="F", copy=True)
        inp[v + "_t"] = g.zeros()
    inp["RHOT0"][:] = case.state["RHOT"]
    inp["RHOT"][:] = case.state["RHOT"]
    inp["RT2P"] = prep["RT2P"]
    inp["REF_rhot"] = np.array(case.state["RHOT"], order="F", copy=True)
    inp["DPRES0"] = np.array(prep["DPRES0"], order="F", copy=True)
    inp["REF_dens"] = case.ref["REF_dens"]
    inp["num_diff"] = g.zeros(5, 3)
    inp["DDIV"] = g.zeros()
    return inp


def _calm_case(**kw):
    args = dict(kmax=10, imax=8, jmax=6, noise=0.0, bubble=False, u0=0.0, v0=0.0, ntracer=0, dz=300.0)
    args.update(kw)
    return inputs.make_case(**args)


def _coords(g):
    xc = (np.arange(g.IA) - g.IS + 0.5) * g.dx
    yc = (np.arange(g.JA) - g.JS + 0.5) * g.dy
    return xc, yc, xc + 0.5 * g.dx, yc + 0.5 * g.dy


def _tendency(out, inp, v, dtrk):
    return (out[v + "_RK"] - inp[v + "0"]) / dtrk


@pytest.mark.parametrize("scheme", [UD3, CD4])
def test_hevi_horizontal_pressure_gradient_closed_form(scheme):
    """DPRES = P(z) + q1 x + q2 x^2 + r1 y + r2 y^2 at rest: the MOMX and MOMY
    tendencies are -(q1 + 2 q2 x_u) and -(r1 + 2 r2 y_v), the exact gradient
    at the staggered points.  A sign error flips them; an index shift moves
    them by 2 q2 dx."""
    case = _calm_case()
    g = case.grid
    inp = _quiet_inputs(case)
    xc, yc, xu, yv = _coords(g)
    q1, q2, r1, r2 = 0.02, 3.0e-6, -0.015, -2.0e-6
    add = q1 * xc[:, None] + q2 * xc[:, None] ** 2 + r1 * yc[None, :] + r2 * yc[None, :] ** 2
    inp["DPRES0"] = np.asfortranarray(inp["DPRES0"] + add[None, :, :])
    dt = 0.6
    out = po.tstep_short_hevi(case, inp, dt, scheme=scheme)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    tx = _tendency(out, inp, "MOMX", dt)[sl]
    ty = _tendency(out, inp, "MOMY", dt)[sl]
    wx = -(q1 + 2.0 * q2 * xu[g.IS:g.IE + 1])[None, :, None]
    wy = -(r1 + 2.0 * r2 * yv[g.JS:g.JE + 1])[None, None, :]
    assert np.allclose(tx, np.broadcast_to(wx, tx.shape), rtol=1e-7, atol=0)
    assert np.allclose(ty, np.broadcast_to(wy, ty.shape), rtol=1e-7, atol=0)


def test_hevi_coriolis_closed_form():
    """Uniform v with f = f0 + beta x + gamma y: the MOMX tendency is
    f(x_u, y) rho v; uniform u: the MOMY tendency is -f(x, y_v) rho u
    (HEVI.F90:900-905, 1020-1025).  A sign error or an average over the
    wrong neighbours (j+1 in place of i+1) fails."""
    case = _calm_case()
    g = case.grid
    xc, yc, xu, yv = _coords(g)
    f0, beta, gamma = 1.0e-4, 2.0e-9, -3.0e-9
    cor = np.asfortranarray((f0 + beta * xc[:, None] + gamma * yc[None, :])[None])
    rho = case.state["DENS"]
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    dt = 0.5
    for comp, vel in (("MOMY", 7.0), ("MOMX", -5.0)):
        inp = _quiet_inputs(case)
        inp["CORIOLI"] = cor
        for key in (comp, comp + "0"):
            inp[key] = np.asfortranarray(rho * vel)
        out = po.tstep_short_hevi(case, inp, dt)
        if comp == "MOMY":  # MOMX tendency at (x_u, y)
            got = _tendency(out, inp, "MOMX", dt)[sl]
            want = (f0 + beta * xu[:, None] + gamma * yc[None, :])[None] * rho * vel
        else:               # MOMY tendency at (x, y_v)
            got = _tendency(out, inp, "MOMY", dt)[sl]
            want = -(f0 + beta * xc[:, None] + gamma * yv[None, :])[None] * rho * vel
        assert np.allclose(got, want[sl], rtol=1e-7, atol=0), comp


def test_hevi_map_factor_metric_term_closed_form():
    """A map factor linear in its reciprocal: m_y = 1/(1 + a x) (m_x = 1) with
    uniform v gives the MOMX tendency m_y(x_u) rho v^2 a; m_x = 1/(1 + b y)
    (m_y = 1) with uniform u gives the MOMY tendency m_x(y_v) rho u^2 b: the
    metric term cf of HEVI.F90:900-906, 1020-1026 (f = 0).  Swapping the
    factors, their indices or the sign fails."""
    case = _calm_case()
    g = case.grid
    xc, yc, xu, yv = _coords(g)
    rho = case.state["DENS"]
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    dt = 0.5
    a, b = 2.0e-6, -3.0e-6
    # MAPF(i,j,n,c): n = 0 x factor, 1 y factor; c = XY, UY, XV, UV
    xs = {0: xc, 1: xu, 2: xc, 3: xu}
    ys = {0: yc, 1: yc, 2: yv, 3: yv}
    for comp, vel in (("MOMY", 6.0), ("MOMX", -4.0)):
        mapf = np.ones((g.IA, g.JA, 2, 4), order="F")
        for c in range(4):
            if comp == "MOMY":
                mapf[:, :, 1, c] = (1.0 / (1.0 + a * xs[c]))[:, None]
            else:
                mapf[:, :, 0, c] = (1.0 / (1.0 + b * ys[c]))[None, :]
        cm = inputs.Case(**{**case.__dict__, "metric": {**case.metric, "MAPF": mapf}})
        inp = _quiet_inputs(cm)
        for key in (comp, comp + "0"):
            inp[key] = np.asfortranarray(rho * vel)
        out = po.tstep_short_hevi(cm, inp, dt)
        if comp == "MOMY":
            got = _tendency(out, inp, "MOMX", dt)[sl]
            want = (1.0 / (1.0 + a * xu))[None, :, None] * rho * vel * vel * a
        else:
            got = _tendency(out, inp, "MOMY", dt)[sl]
            want = (1.0 / (1.0 + b * yv))[None, None, :] * rho * vel * vel * b
        assert np.allclose(got, np.broadcast_to(want, rho.shape)[sl], rtol=1e-9, atol=0), comp


def _advection_error(scheme, n_per_wave, comp):
    """Max error of the MOMX (or MOMY) tendency of u = u0 + eps sin(k x)
    (or v in y) at constant density against the exact -d(rho u^2)/dx."""
    nx = n_per_wave
    L = 16000.0
    dx = L / nx
    if comp == "MOMX":
        case = _calm_case(imax=nx, jmax=4, dx=dx)
    else:
        case = _calm_case(imax=4, jmax=nx, dy=dx)
    g = case.grid
    inp = _quiet_inputs(case)
    rho = 1.15
    for v in ("DENS", "DENS0"):
        inp[v] = np.full(g.shape, rho, order="F")
    xc, yc, xu, yv = _coords(g)
    kw = 2 * np.pi / L
    u0, eps = 8.0, 2.0
    s = xu if comp == "MOMX" else yv
    vel = u0 + eps * np.sin(kw * s)
    shape = (1, -1, 1) if comp == "MOMX" else (1, 1, -1)
    for key in (comp, comp + "0"):
        inp[key] = np.asfortranarray(np.broadcast_to(rho * vel.reshape(shape), g.shape).copy())
    dt = 0.5
    out = po.tstep_short_hevi(case, inp, dt, scheme=scheme)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    got = _tendency(out, inp, comp, dt)[sl]
    exact = -rho * 2.0 * vel * eps * kw * np.cos(kw * s)
    return np.abs(got - np.broadcast_to(exact.reshape(shape), g.shape)[sl]).max() / np.abs(exact).max()


@pytest.mark.parametrize("comp", ["MOMX", "MOMY"])
@pytest.mark.parametrize("scheme", [UD3, CD4])
def test_hevi_momentum_advection_is_second_order(scheme, comp):
    """u = u0 + eps sin(k x) advecting itself at constant density: the
    flux-form tendency converges to -d(rho u^2)/dx at second order (the face
    velocity is a two-point average; the face value is 3rd/4th order), so
    doubling the resolution divides the error by about 4.  A shifted index or
    a wrong upwind sign leaves a first-order or O(1) error."""
    e = [_advection_error(scheme, n, comp) for n in (16, 32, 64)]
    assert e[0] < 0.05, e
    for a, b in zip(e, e[1:]):
        assert 3.5 < a / b < 4.5, e


# ---------------------------------------------------------------------------
# terrain-following metrics (scale_atmos_grid_cartesC_real.F90:391-545,
# scale_atmos_grid_cartesC_metric.F90:200-345) and the terrain pressure gradient
def _metric_case(**kw):
    args = dict(kmax=12, imax=12, jmax=10, noise=0.0, bubble=False, u0=0.0, ntracer=0, dz=250.0)
    args.update(kw)
    return inputs.make_case(**args)


def test_terrain_metrics_flat_and_uniform_lift():
    """Zsfc = 0 gives the flat metrics (G^1/2 = 1 to round-off, J13G = J23G = 0
    exactly); a uniform surface height h0 compresses every column by
    (H - h0)/H, with J13G = J23G = 0."""
    case = _metric_case()
    g = case.grid
    H_ = case.fz[g.KE + 1] - case.fz[g.KS]
    for h0 in (0.0, 600.0):
        m = po.terrain_metrics(g, np.full((g.IA, g.JA), h0, order="F"), case.fz, case.metric)
        assert np.allclose(m["GSQRT"], (H_ - h0) / H_, rtol=1e-13, atol=0)
        assert not m["J13G"].any() and not m["J23G"].any()
        cz = 0.5 * (case.fz[1:] + case.fz[:-1])
        assert np.allclose(m["PHI"], inputs.GRAV * ((H_ - h0) / H_ * cz + h0)[:, None, None], rtol=1e-14, atol=0)


@pytest.mark.parametrize("axis", [0, 1], ids=["x", "y"])
def test_terrain_metrics_linear_slope_closed_form(axis):
    """Zsfc = s x (or s y): z(x, zeta) = (H - Zs)/H zeta + Zs, so
    G^1/2 = (H - Zs)/H at every staggered point and J13G (J23G) = -s (1 - zeta/H)
    at each point's own zeta (cell centre or face), exactly, away from the
    periodic seam.  A sign, a wrong staggered position or a swapped
    centre/face height fails."""
    case = _metric_case()
    g = case.grid
    H_ = case.fz[g.KE + 1] - case.fz[g.KS]
    s = 0.03
    xc = (np.arange(g.IA) - g.IS + 0.5) * g.dx
    yc = (np.arange(g.JA) - g.JS + 0.5) * g.dy
    zs = np.asfortranarray(s * (xc[:, None] + 0 * yc[None, :]) if axis == 0 else s * (0 * xc[:, None] + yc[None, :]))
    m = po.terrain_metrics(g, zs, case.fz, case.metric)
    fz = case.fz
    cz = 0.5 * (fz[1:] + fz[:-1])
    zeta = {"c": cz, "f": fz[1:]}  # cell centre / top face of each level
    # staggered x / y offsets (in cells) and vertical position of each component
    comps = {0: (0.0, 0.0, "c"), 1: (0.0, 0.0, "f"), 2: (0.5, 0.0, "f"), 3: (0.0, 0.5, "f"),
             4: (0.5, 0.0, "c"), 5: (0.0, 0.5, "c"), 6: (0.5, 0.5, "c")}
    if axis == 0:
        isl, jsl = slice(g.IS + 1, g.IE - 1), slice(g.JS, g.JE + 1)
    else:
        isl, jsl = slice(g.IS, g.IE + 1), slice(g.JS + 1, g.JE - 1)
    kk = slice(g.KS, g.KE + 1)
    J = m["J13G"] if axis == 0 else m["J23G"]
    Jo = m["J23G"] if axis == 0 else m["J13G"]
    for c, (ox, oy, pos) in comps.items():
        z_ = zeta[pos][kk][:, None, None]
        pos_xy = (xc[isl] + ox * g.dx)[:, None] if axis == 0 else (yc[jsl] + oy * g.dy)[None, :]
        zs_c = s * pos_xy
        gs_want = (H_ - zs_c)[None] / H_ + 0 * z_
        assert np.allclose(m["GSQRT"][kk, isl, jsl, c] if c in (0, 4, 5, 6) else m["GSQRT"][kk, isl, jsl, c],
                           np.broadcast_to(gs_want, m["GSQRT"][kk, isl, jsl, c].shape), rtol=1e-12, atol=0), c
        j_want = -s * (1.0 - z_ / H_)
        assert np.allclose(J[kk, isl, jsl, c], np.broadcast_to(j_want, J[kk, isl, jsl, c].shape), rtol=1e-12,
                           atol=1e-16), c
        assert np.abs(Jo[kk, isl, jsl, c]).max() < 1e-16, c


def _terrain_pg_error(refine):
    """Max |MOMX tendency| of a fluid at rest whose pressure perturbation
    depends on the physical height only, over a bell mountain, relative to
    the terrain-following part G^1/2 dp/dx|zeta it must cancel."""
    n = refine
    case = _metric_case(kmax=12 * n, imax=24 * n, jmax=4, dx=1000.0 / n, dz=500.0 / n)
    g = case.grid
    zs = inputs.bell_topography(g, height=800.0, half_width=4000.0, cy=0.0, ry=1e12)  # a ridge along y
    m = po.terrain_metrics(g, zs, case.fz, case.metric)
    cm = inputs.Case(**{**case.__dict__, "metric": {**case.metric, **m}})
    inp = _quiet_inputs(cm)
    zreal = m["PHI"] / inputs.GRAV
    # hydrostatic with the lowest level's density, so that the ghost level
    # extrapolation DPRES(KS-1) = DPRES0(KS-1) - DENS(KS) (PHI(KS-1) - PHI(KS+1))
    # (HEVI.F90:357) continues it when DPRES0(KS-1) = DPRES0(KS+1), as the
    # large step leaves it for a hydrostatic reference state; a curvature term
    # e^{-z/Hs} is added on the interior levels and on their ghost copies
    rho_s = case.state["DENS"][g.KS, g.IS, g.JS]
    curv = 300.0 * np.exp(-zreal / 3000.0)
    p = 500.0 - rho_s * inputs.GRAV * zreal + curv
    p[g.KS - 1] = p[g.KS + 1] + curv[g.KS - 1] - curv[g.KS + 1]
    p[g.KE + 1] = p[g.KE - 1] + curv[g.KE + 1] - curv[g.KE - 1]
    inp["DPRES0"] = np.asfortranarray(p)
    dt = 0.5
    out = po.tstep_short_hevi(cm, inp, dt)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    tx = _tendency(out, inp, "MOMX", dt)[sl]
    # the terrain-following gradient alone (what a dropped J13G term would leave)
    tf = np.abs((p[:, g.IS + 1:g.IE + 2] - p[:, g.IS:g.IE + 1]) / g.dx)[g.KS:g.KE + 1]
    return np.abs(tx).max() / tf.max()


def test_hevi_terrain_pressure_gradient_is_second_order():
    """Over a bell mountain (the metrics of orc_terrain_metrics) a pressure
    that depends on physical height only has no horizontal gradient: the
    HEVI MOMX tendency (G^1/2 dp/dx|zeta plus the J13G term, HEVI.F90:880-895)
    is a small fraction of either part and falls by about 4 per halving of
    the grid spacing.  A sign or staggering error in the J13G term leaves an
    O(1) fraction."""
    e = [_terrain_pg_error(n) for n in (1, 2, 4)]
    assert e[0] < 0.05, e
    for a, b in zip(e, e[1:]):
        assert 3.0 < a / b < 5.0, e
