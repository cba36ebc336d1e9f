"""Host coupling of the device-resident model: physics tendencies and restart.

* Physics tendencies.  The host hands the large step its physics tendencies
  DENS_tp .. RHOT_tp and RHOQ_tp (mod_atmos_dyn_driver.F90:327-388 passes
  them to ATMOS_DYN; the large step folds them into the short-step
  tendencies, LARGE.F90:443-839).  Large steps with smooth random tendencies
  on every prognostic and tracer match the oracle's to 1e-12 relative.
* Restart.  The reference checks its restart by running a case straight
  (restart/500m runA: HEVI, RK4, CD4 dynamics, UD3KOREN1993 tracers,
  TIME_DT = 2 s, TIME_DT_ATMOS_DYN = 1 s) and split in two at a restart file
  (runB1 + runB2), every restart variable compared for exact equality.  Here
  the same: N large steps straight against N/2, a restart file, a fresh
  context that reads it and N/2 more, bit for bit.
"""
import numpy as np
import pytest

import helpers as H
import pyoracle as po
from scale_amd import inputs

pytestmark = pytest.mark.gpu

STATE = ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")


@pytest.fixture(scope="module")
def hip():
    from scale_amd import hip as shp
    shp.load_library()
    return shp


def smooth_tendencies(case, seed):
    """Smooth random physics tendencies (periodic halos filled): a few
    low-wavenumber modes per field, scaled to each field's magnitude over 1e4 s."""
    g = case.grid
    rng = np.random.default_rng(seed)
    k = np.arange(g.KA)[:, None, None] / g.KA
    i = np.arange(g.IA)[None, :, None] / g.imax
    j = np.arange(g.JA)[None, None, :] / g.jmax

    def field(scale):
        a = np.zeros(g.shape)
        for _ in range(3):
            a += rng.standard_normal() * np.sin(2 * np.pi * (rng.integers(1, 3) * i + rng.uniform())) \
                 * np.cos(2 * np.pi * (rng.integers(1, 3) * j + rng.uniform())) * np.sin(np.pi * k * rng.integers(1, 3))
        a = np.asfortranarray(a * scale)
        inputs.halo_fill(g, a)
        return a

    tend = {f"{v}_tp": field(np.abs(case.state[v]).max() * 1e-4 + 1e-6) for v in STATE}
    qa = case.qtrc.shape[-1]
    rhoq = np.asfortranarray(np.stack([field(1e-7) for _ in range(qa)], axis=-1))
    return tend, rhoq


def test_model_physics_tendencies_match_oracle(hip):
    print("[coupling] physics tendencies on every prognostic and tracer, 5 large steps", flush=True)
    case = H.small_case(kmax=20, imax=16, jmax=12, seed=77, fct=True)
    tend, rhoq = smooth_tendencies(case, 5)
    ref_case = inputs.Case(**{**case.__dict__, "zero_tend": {**tend, "RHOQ_tp": rhoq}})
    m = po.Model(ref_case)
    m.rhoq_tp = np.array(rhoq, order="F", copy=True)
    m.step(5)
    ctx = hip.model_from_case(case)
    try:
        for k, v in tend.items():
            ctx.model_set(k, v)
        for iq in range(rhoq.shape[-1]):
            ctx.model_set(hip.RHOQ_TP_BASE + iq, np.asfortranarray(rhoq[..., iq]))
        ctx.model_large_step(5)
        g = case.grid
        sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
        for k in STATE:
            got, ref = ctx.model_get(k), m.state[k]
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  {k}: max|diff| {d:.3e}", flush=True)
            assert d <= 1e-12 * np.abs(ref[sl]).max(), k
        for iq in range(rhoq.shape[-1]):
            got, ref = ctx.model_get(hip.QTRC_BASE + iq), m.qtrc[..., iq]
            d = np.abs(got[sl] - ref[sl]).max()
            print(f"  QTRC{iq}: max|diff| {d:.3e}", flush=True)
            assert d <= 1e-12 * max(np.abs(ref[sl]).max(), 1e-30), iq
    finally:
        ctx.close()
    # and they act: the state differs from a run without them
    m0 = po.Model(case)
    m0.step(5)
    assert np.abs(m0.state["RHOT"][sl] - m.state["RHOT"][sl]).max() > 1e-9 * np.abs(m.state["RHOT"][sl]).max()


def restart_case():
    # restart/500m runA: CD4 dynamics, UD3KOREN1993 tracers, dt_dyn = 1 s, dt = 2 s (nstep = 2)
    return H.small_case(kmax=20, imax=16, jmax=12, seed=91, scheme=4, scheme_tracer=31, fct=True, dt=1.0)


@pytest.mark.parametrize("with_tendencies", [False, True], ids=["no_physics", "physics_tendencies"])
def test_restart_is_bitwise(hip, tmp_path, with_tendencies):
    print(f"[restart] 6 large steps straight vs 3 + restart file + 3, physics={with_tendencies}", flush=True)
    case = restart_case()
    tend, rhoq = smooth_tendencies(case, 8)
    qa = case.qtrc.shape[-1]

    def model():
        c = hip.model_from_case(case, dtl=2.0)
        if with_tendencies:  # the host sets its tendencies again after a restart
            for k, v in tend.items():
                c.model_set(k, v)
            for iq in range(qa):
                c.model_set(hip.RHOQ_TP_BASE + iq, np.asfortranarray(rhoq[..., iq]))
        return c

    def everything(c):
        return {**{k: c.model_get(k) for k in STATE}, **{f"QTRC{iq}": c.model_get(hip.QTRC_BASE + iq) for iq in range(qa)}}

    a = model()
    try:
        a.model_large_step(6)
        run_a = everything(a)
    finally:
        a.close()
    path = str(tmp_path / "restart.bin")
    b1 = model()
    try:
        b1.model_large_step(3)
        b1.model_restart_write(path)
        mid = everything(b1)
    finally:
        b1.close()
    b2 = model()
    try:
        for k in STATE:  # the fresh context starts elsewhere: the file must replace its state
            b2.model_set(k, case.state[k] * 1.01)
        b2.model_restart_read(path)
        back = everything(b2)
        for k in back:
            assert np.array_equal(back[k], mid[k]), k
        b2.model_large_step(3)
        run_b = everything(b2)
    finally:
        b2.close()
    for k in run_a:
        assert np.array_equal(run_a[k], run_b[k]), k
    assert not np.array_equal(run_a["RHOT"], mid["RHOT"])


def test_restart_rejects_other_grids_and_mixed_precision(hip, tmp_path):
    print("[restart] mismatched grid, bad file, mixed precision", flush=True)
    case = restart_case()
    path = str(tmp_path / "r.bin")
    c = hip.model_from_case(case, dtl=2.0)
    try:
        c.model_restart_write(path)
    finally:
        c.close()
    other = H.small_case(kmax=20, imax=16, jmax=8, seed=91, scheme=4, scheme_tracer=31, fct=True)
    c = hip.model_from_case(other)
    try:
        with pytest.raises(hip.ScaleHipError, match="does not match"):
            c.model_restart_read(path)
        bad = tmp_path / "bad.bin"
        bad.write_bytes(b"not a restart file")
        with pytest.raises(hip.ScaleHipError, match="not a restart file"):
            c.model_restart_read(str(bad))
    finally:
        c.close()
    c = hip.model_from_case(case, mixed_precision=True)
    try:
        with pytest.raises(hip.ScaleHipError, match="mixed precision"):
            c.model_restart_write(path)
    finally:
        c.close()
