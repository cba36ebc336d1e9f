"""GPU tests of the adjoint (scale_hip_tstep_short_hevi_adj and the
checkpointed model forward/adjoint pair) against the oracle's adjoint, which
tests/test_adjoint_oracle.py pins by finite differences, and against finite
differences of the GPU forward itself.

Tolerances: the GPU adjoint accumulates with FP64 atomics, so its sums are
reordered; it must agree with the oracle's adjoint within 1e-10 of each
field's largest adjoint.  The gradient of J = sum of (theta - theta_ref)^2
over a central box must agree with central finite differences of the GPU
forward (eps = 1e-7 of each field's largest magnitude per unit of a
standard-normal direction) within 1e-6 relative.
"""
import numpy as np
import pytest

from scale_amd import hip, inputs

pytestmark = pytest.mark.gpu

O = pytest.importorskip("pyoracle")
STATE = hip.STATE

from test_adjoint_oracle import fill, interior_gradient, perturbed_case, stage_inputs, theta_box_objective  # noqa: E402


def close(gpu, ref, tol=1e-10):
    scale = max(float(np.abs(ref).max()), 1e-300)
    return float(np.abs(gpu - ref).max()) <= tol * scale


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_stage_adjoint_matches_oracle(scheme):
    print(f"[adjoint] stage 8x8x10 scheme {scheme}", flush=True)
    case = perturbed_case(scheme)
    g = case.grid
    inp = stage_inputs(case)
    rng = np.random.default_rng(17)
    out_b = {k + "_RK": np.asfortranarray(rng.standard_normal(g.shape)) for k in STATE}
    dt = 0.5
    ref = O.adj_tstep_short_hevi(case, inp, dt, out_b)
    m = case.metric
    with hip.Context(g, scheme, scheme) as ctx:
        got = ctx.ATMOS_DYN_Tstep_short_fvm_hevi_adj(
            out_b, *[inp[k + "0"] for k in STATE], *[inp[k] for k in STATE], *[inp[k + "_t"] for k in STATE],
            inp["DPRES0"], inp["RT2P"], inp["CORIOLI"], inp["num_diff"], inp["wdamp_coef"],
            m["CDZ"], m["FDZ"], m["FDX"], m["FDY"], m["RCDZ"], m["RCDX"], m["RCDY"], m["RFDZ"], m["RFDX"], m["RFDY"],
            m["PHI"], m["GSQRT"], m["J13G"], m["J23G"], m["J33G"], m["MAPF"], inp["REF_dens"], inp["REF_rhot"], dt)
    for key in ref:
        assert close(got[key], ref[key]), key


@pytest.mark.parametrize("imax,jmax,kmax", [(65, 3, 10), (130, 2, 9)], ids=["65x3x10", "130x2x9"])
def test_stage_adjoint_odd_widths(imax, jmax, kmax):
    """x rows of 66 and 131 faces, which cross the 64-lane waves of the
    shuffle-summed x scatters at varying offsets; y rows and columns shorter
    than, or a partial multiple of, the 8-point line windows."""
    print(f"[adjoint] stage {imax}x{jmax}x{kmax} UD3", flush=True)
    case = inputs.make_case(kmax=kmax, imax=imax, jmax=jmax, v0=2.0, momz_noise=0.05, noise=0.1, ntracer=0, scheme=3)
    g = case.grid
    rng = np.random.default_rng(23)
    for k, a in case.state.items():
        a[g.KS:g.KE + 1] *= 1.0 + 1e-3 * rng.standard_normal(a[g.KS:g.KE + 1].shape)
        fill(g, a, k)
    inp = stage_inputs(case)
    out_b = {k + "_RK": np.asfortranarray(rng.standard_normal(g.shape)) for k in STATE}
    dt = 0.5
    ref = O.adj_tstep_short_hevi(case, inp, dt, out_b)
    m = case.metric
    with hip.Context(g, 3, 3) as ctx:
        got = ctx.ATMOS_DYN_Tstep_short_fvm_hevi_adj(
            out_b, *[inp[k + "0"] for k in STATE], *[inp[k] for k in STATE], *[inp[k + "_t"] for k in STATE],
            inp["DPRES0"], inp["RT2P"], inp["CORIOLI"], inp["num_diff"], inp["wdamp_coef"],
            m["CDZ"], m["FDZ"], m["FDX"], m["FDY"], m["RCDZ"], m["RCDX"], m["RCDY"], m["RFDZ"], m["RFDX"], m["RFDY"],
            m["PHI"], m["GSQRT"], m["J13G"], m["J23G"], m["J33G"], m["MAPF"], inp["REF_dens"], inp["REF_rhot"], dt)
    for key in ref:
        assert close(got[key], ref[key]), key


@pytest.mark.parametrize("records,tape,split", [("1", 0, "0"), ("1", 1, "0"), ("0", 0, "0"), ("1", 1, "1")],
                         ids=["all_segments_taped", "last_segment_taped", "per_step_recompute",
                              "last_segment_taped-split_launches"])
def test_model_adjoint_matches_oracle(records, tape, split, monkeypatch):
    """The three reverse sweeps: every segment's tape kept by the forward (the
    default where memory holds them), only the last one's (the others recorded
    again from their checkpoints) and, with SCALE_HIP_ADJ_RECORDS=0, the
    per-step recomputation used when no records fit in device memory; and
    the forward's flat stage as split launches (the path of small grids)."""
    monkeypatch.setenv("SCALE_HIP_ADJ_RECORDS", records)
    monkeypatch.setenv("SCALE_HIP_FLAT_SPLIT", split)
    print(f"[adjoint] model 8x8x10, 3 steps, checkpoint every 2, records={records}, taped segments={tape}", flush=True)
    case = perturbed_case()
    N = 3
    traj = [{k: np.array(v, order="F", copy=True) for k, v in case.state.items()}]
    for _ in range(N):
        m = O.Model(case)
        m.state = {k: np.array(v, order="F", copy=True) for k, v in traj[-1].items()}
        m.step(1)
        traj.append(m.state)
    _, xbN = theta_box_objective(case, traj[N])
    xb = {k: v.copy(order="F") for k, v in xbN.items()}
    for n in reversed(range(N)):
        m = O.Model(case)
        m.state = traj[n]
        xb = O.adj_large_step(m, xb)
    ctx = hip.model_from_case(case)
    try:
        ctx.tune_set(hip.TUNE_ADJ_TAPE, tape)
        ctx.model_forward_checkpointed(N, 2)
        for k in STATE:  # the forward itself matches the oracle
            assert close(ctx.model_get(k), traj[N][k], 1e-12), k
        got = ctx.model_adjoint(xbN)
        # a second reverse sweep of the same run (the first one re-recorded
        # into the tape slots of the segments it did not have taped)
        again = ctx.model_adjoint(xbN)
    finally:
        ctx.close()
    for k in STATE:
        assert close(got[k], xb[k]), k
        assert close(again[k], xb[k]), k


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_gpu_adjoint_dot_product_against_tangent_linear(scheme):
    """The dot-product test of the GPU adjoint: <L dx, y> from the oracle's
    tangent-linear model (the oracle on dual numbers) against <dx, L^T y>
    from the GPU's checkpointed forward and reverse sweep, over 6 large steps
    (checkpoints every 4, 2 FCT tracers on) of a 24x20x16 state, on 3
    seeded directions and weights: within 1e-10 relative (the GPU adjoint
    sums with atomics, and its forward differs from the oracle's by the last
    bit of pow())."""
    T = pytest.importorskip("pytlm")
    print(f"[adjoint] dot product against the tangent-linear model, scheme {scheme}", flush=True)
    case = inputs.make_case(kmax=16, imax=24, jmax=20, v0=2.0, momz_noise=0.05, noise=0.1, ntracer=2,
                            tracer_kind=["blob", "tophat"], fct=True, scheme=scheme, u0=8.0)
    g = case.grid
    N = 6
    rng = np.random.default_rng(31 + scheme)
    for trial in range(3):
        d = {}
        for k in STATE:
            a = np.zeros(g.shape, order="F")
            ke = g.KE - 1 if k == "MOMZ" else g.KE
            sl = (slice(g.KS, ke + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
            a[sl] = rng.standard_normal(a[sl].shape) * np.abs(case.state[k]).max()
            d[k] = inputs.fill_ghost_and_halo(g, a, k)
        y = {k: np.asfortranarray(rng.standard_normal(g.shape)) for k in STATE}
        t = T.TangentModel(case, case.state, d)
        t.step(N)
        tan = t.tangents()
        lhs = sum(float(np.sum(tan[k] * y[k])) for k in STATE)
        ctx = hip.model_from_case(case)
        try:
            ctx.model_forward_checkpointed(N, 4)
            xb0 = ctx.model_adjoint(y)
        finally:
            ctx.close()
        grad = inputs.interior_gradient(g, xb0)
        rhs = sum(float(np.sum(grad[k] * d[k])) for k in STATE)
        print(f"  trial {trial}: <L dx, y> = {lhs:.15e}, <dx, L^T y> (GPU) = {rhs:.15e}", flush=True)
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)


def test_model_gradient_matches_finite_differences():
    print("[adjoint] warm bubble 24x24x32, 2 spin-up steps, 8 steps, checkpoint every 3", flush=True)
    case = inputs.make_case(kmax=32, imax=24, jmax=24, dz=250.0, ntracer=0, noise=0.01, u0=10.0)
    g = case.grid
    # spin up off the exactly-zero velocities of the initial state
    ctx = hip.model_from_case(case)
    ctx.model_large_step(2)
    x0 = {k: ctx.model_get(k) for k in STATE}
    ctx.close()
    N = 8

    def forward(state):
        c = hip.model_from_case(case)
        try:
            for k in STATE:
                c.model_set(k, state[k])
            c.model_large_step(N)
            return {k: c.model_get(k) for k in STATE}
        finally:
            c.close()

    ctx = hip.model_from_case(case)
    try:
        for k in STATE:
            ctx.model_set(k, x0[k])
        ctx.model_forward_checkpointed(N, 3)
        xN = {k: ctx.model_get(k) for k in STATE}
        J0, xbN = theta_box_objective(case, xN)
        xb0 = ctx.model_adjoint(xbN)
    finally:
        ctx.close()
    grad = interior_gradient(g, xb0)
    rng = np.random.default_rng(21)
    for trial in range(3):
        d = {}
        for k in STATE:
            a = np.zeros(g.shape, order="F")
            ke = g.KE - 1 if k == "MOMZ" else g.KE
            sl = (slice(g.KS, ke + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
            a[sl] = rng.standard_normal(a[sl].shape) * np.abs(x0[k]).max()
            d[k] = fill(g, a, k)
        ad = sum(float(np.sum(grad[k] * d[k])) for k in STATE)
        eps = 1e-7
        Jp = theta_box_objective(case, forward({k: np.asfortranarray(x0[k] + eps * d[k]) for k in STATE}))[0]
        Jm = theta_box_objective(case, forward({k: np.asfortranarray(x0[k] - eps * d[k]) for k in STATE}))[0]
        fd = (Jp - Jm) / (2 * eps)
        print(f"  direction {trial}: fd {fd:.12e} adjoint {ad:.12e} rel {abs(fd - ad) / abs(ad):.2e}", flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)


def test_adjoint_rejects_unsupported_configurations():
    case = inputs.warm_bubble_40x40x80()
    ctx = hip.model_from_case(case, mixed_precision=True)
    try:
        with pytest.raises(hip.ScaleHipError):
            ctx.model_forward_checkpointed(2, 1)
        with pytest.raises(hip.ScaleHipError):
            ctx.model_adjoint({k: np.zeros(case.grid.shape, order="F") for k in STATE})
    finally:
        ctx.close()
