"""GPU test of the Fortran drop-in (scale_amd/fortran/scale_atmos_dyn_hip.F90):
a Fortran program binds the module's procedures exactly as SCALE's
ATMOS_DYN_tinteg_short / _tinteg_tracer bind their procedure pointers, runs
one HEVI (or HEVE) stage and one FCT tracer stage on the GPU through the C ABI, and its
outputs must equal the CPU oracle's bit for bit.  This pins the Fortran side
of the boundary: argument order, by-reference scalars, LOGICAL conversion and
the (KA,IA,JA) layout."""
import os
import subprocess

import numpy as np
import pytest

import helpers as H
import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "dropin_driver")


def f(a):
    return np.asarray(a, dtype=np.float64).ravel(order="F").tobytes()


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
@pytest.mark.parametrize("short", ["hevi", "heve"])
def test_fortran_dropin_bitwise(tmp_path, scheme, short):
    print(f"[fortran drop-in] scheme={scheme} short={short}", flush=True)
    if not os.path.exists(DRIVER):
        pytest.fail("tests/fortran/dropin_driver is not built (run __graft_entry__.build())")
    case = H.small_case(seed=500 + scheme, scheme=scheme, scheme_tracer=scheme)
    g = case.grid
    m = case.metric
    inp = H.stage_inputs(case, seed=31, coriolis=True)
    inp["wdamp_coef"] = np.linspace(0.0, 0.01, g.KA)
    dtrk, dtl = 0.5, 2.0
    rng = np.random.Generator(np.random.PCG64(99))
    qtrc = np.asfortranarray(case.qtrc[..., 0])
    qtrco = np.asfortranarray(np.zeros(g.shape))
    rhoq_t = np.asfortranarray(1.0e-6 * rng.standard_normal(g.shape))
    nd_q = np.asfortranarray(np.zeros(g.shape + (3,)))

    hdr = np.array([g.KA, g.IA, g.JA, g.KS + 1, g.KE + 1, g.IS + 1, g.IE + 1, g.JS + 1, g.JE + 1,
                    scheme, scheme, 1, int(short == "heve")], dtype=np.int32)
    blob = [hdr.tobytes()]
    for v in H.VARS:
        blob.append(f(inp[v + "0"]))
    for v in H.VARS:
        blob.append(f(inp[v]))
    for v in H.VARS:
        blob.append(f(inp[v + "_t"]))
    blob += [f(inp["DPRES0"]), f(inp["RT2P"]), f(inp["CORIOLI"]), f(inp["num_diff"]), f(inp["wdamp_coef"]),
             f(0.0), f(inp["DDIV"])]
    blob += [f(m[k]) for k in ("CDZ", "FDZ", "FDX", "FDY", "RCDZ", "RCDX", "RCDY", "RFDZ", "RFDX", "RFDY")]
    blob += [f(m["PHI"]), f(m["GSQRT"]), f(m["J13G"]), f(m["J23G"]), f(m["J33G"]), f(m["MAPF"]),
             f(inp["REF_dens"]), f(inp["REF_rhot"]), f(dtrk)]
    blob += [f(qtrco), f(qtrc), f(rhoq_t), f(nd_q), f(dtl)]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(b"".join(blob))

    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.frombuffer(fout.read_bytes(), dtype=np.float64)
    n3 = g.KA * g.IA * g.JA
    got, off = {}, 0
    for k, n in (("DENS_RK", 1), ("MOMZ_RK", 1), ("MOMX_RK", 1), ("MOMY_RK", 1), ("RHOT_RK", 1),
                 ("mflx_hi", 3), ("tflx_hi", 3), ("QTRCo", 1)):
        shape = g.shape if n == 1 else g.shape + (n,)
        got[k] = raw[off:off + n * n3].reshape(shape, order="F")
        off += n * n3
    assert off == raw.size

    # oracle: same stage; mflx_hi starts from zeros as in the driver
    out = {k: po.nan_like_outputs(g) for k in ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK")}
    out["mflx_hi"] = np.asfortranarray(np.zeros(g.shape + (3,)))
    out["tflx_hi"] = po.nan_like_outputs(g, 3)
    stage = po.tstep_short_heve if short == "heve" else po.tstep_short_hevi
    ref = stage(case, inp, dtrk, scheme=scheme, out=out)
    for k in ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK", "tflx_hi"):
        H.compare(ref[k], got[k], k)

    # the driver writes mflx_hi after its periodic halo exchange
    mflx = np.array(ref["mflx_hi"], order="F", copy=True)
    for c in range(3):
        comp = np.asfortranarray(mflx[..., c])
        po.halo_periodic(g, comp)
        mflx[..., c] = comp
    assert np.array_equal(mflx, got["mflx_hi"])
    q_ref = np.array(qtrco, order="F", copy=True)
    po.tstep_tracer(case, q_ref, qtrc, qtrc, rhoq_t, inp["DENS0"], inp["DENS"], mflx, nd_q, dtl,
                    fct=True, along_stream=True, scheme=scheme)
    assert np.array_equal(q_ref, got["QTRCo"])


LARGE_DRIVER = os.path.join(ROOT, "tests", "fortran", "large_driver")


@pytest.mark.parametrize("scheme", [3, 4], ids=["UD3", "CD4"])
def test_fortran_large_step_dropin(tmp_path, scheme):
    """The large-step drop-in ATMOS_DYN_Tstep_large_fvm_heve_hip, called from
    Fortran with the dummies of the interface `large` for 3 large steps of a
    case with two FCT tracers: the state and the tracers within 1e-12 of the
    oracle model (pow() in the large-step preparation may round its last bit
    differently on the GPU)."""
    print(f"[fortran large-step drop-in] scheme={scheme}", flush=True)
    if not os.path.exists(LARGE_DRIVER):
        pytest.fail("tests/fortran/large_driver is not built (run __graft_entry__.build())")
    from scale_amd import inputs
    case = inputs.make_case(kmax=20, imax=16, jmax=12, noise=0.05, ntracer=2, tracer_kind=["blob", "tophat"], fct=True,
                            bubble_r=2000.0, bubble_zc=2000.0, momz_noise=0.05, v0=-2.0, scheme=scheme,
                            scheme_tracer=scheme)
    g = case.grid
    m = case.metric
    nsteps = 3
    hdr = np.array([g.KA, g.IA, g.JA, g.KS + 1, g.KE + 1, g.IS + 1, g.IE + 1, g.JS + 1, g.JE + 1,
                    scheme, scheme, 2, nsteps, 1], dtype=np.int32)
    blob = [hdr.tobytes()]
    blob += [f(case.state[k]) for k in H.VARS] + [f(case.qtrc)]
    blob += [f(case.ref["REF_dens"]), f(case.ref["REF_pott"]), f(case.ref["REF_pres"]), f(case.corioli),
             f(case.wdamp_coef)]
    blob += [f(m[k]) for k in ("CDZ", "CDX", "CDY", "FDZ", "FDX", "FDY", "RCDZ", "RCDX", "RCDY", "RFDZ", "RFDX", "RFDY")]
    blob += [f(m["PHI"]), f(m["GSQRT"]), f(m["J13G"]), f(m["J23G"]), f(m["J33G"]), f(m["MAPF"])]
    blob += [f(case.dt), f(case.dt), f(1.0e-4)]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(b"".join(blob))
    r = subprocess.run([LARGE_DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.frombuffer(fout.read_bytes(), dtype=np.float64)
    n3 = g.KA * g.IA * g.JA
    assert raw.size == 7 * n3
    ref = po.Model(case)
    ref.step(nsteps)
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    for q, k in enumerate(H.VARS):
        got = raw[q * n3:(q + 1) * n3].reshape(g.shape, order="F")
        d = np.abs(got[sl] - ref.state[k][sl]).max()
        print(f"  {k}: max |diff| {d:.3e}", flush=True)
        assert d <= 1e-12 * np.abs(ref.state[k][sl]).max(), k
    for iq in range(2):
        got = raw[(5 + iq) * n3:(6 + iq) * n3].reshape(g.shape, order="F")
        d = np.abs(got[sl] - ref.qtrc[..., iq][sl]).max()
        assert d <= 1e-12 * max(np.abs(ref.qtrc[..., iq][sl]).max(), 1e-30), iq
