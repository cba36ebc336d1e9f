"""Tangent-linear model of the CPU oracle (TEST INFRASTRUCTURE ONLY).

liboracle_tl.so is scale_oracle.c compiled with dual numbers (dual.h): every
real carries its derivative along one direction, so running the oracle's
large step on (state, tangent) pairs returns the state after the step and
L dx, the Jacobian of the step applied to dx, exactly.  The values it
computes are the oracle's own, bit for bit.  Used by the tests to check the
adjoints (oracle and GPU) by <L dx, y> = <dx, L^T y>.

A dual array is a (2, ...) float64 array in Fortran order: element [0, ...]
the value, [1, ...] the tangent, adjacent in memory as the C++ pair is.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_TL = os.path.join(HERE, "liboracle_tl.so")

_dp = C.POINTER(C.c_double)


class DualC(C.Structure):
    _fields_ = [("v", C.c_double), ("t", C.c_double)]


def _dual_struct(base):
    """The oracle struct with each double field a (value, tangent) pair."""
    fields = [(n, DualC if t is C.c_double else t) for n, t in base._fields_]
    return type(base.__name__ + "Dual", (C.Structure,), {"_fields_": fields})


OrcMetricD = _dual_struct(po.OrcMetric)
OrcConfigD = _dual_struct(po.OrcConfig)
OrcStateD = _dual_struct(po.OrcState)  # pointers and ints only: the same layout

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_TL):
            subprocess.run(["make", "-C", HERE, "-s"], check=True)
        _lib = C.CDLL(LIB_TL)
    return _lib


def dual(v, t=None):
    v = np.asarray(v, dtype=np.float64)
    a = np.zeros((2,) + v.shape, order="F")
    a[0] = v
    if t is not None:
        a[1] = t
    return a


def dptr(a):
    assert a.dtype == np.float64 and a.flags.f_contiguous and a.shape[0] == 2
    return a.ctypes.data_as(_dp)


def _metric_dual(m, keep):
    """The oracle metric struct on dual numbers (zero tangents)."""
    ms = OrcMetricD()
    for n in ("CDZ", "FDZ", "RCDZ", "RFDZ", "CDX", "FDX", "RCDX", "RFDX", "CDY", "FDY", "RCDY", "RFDY",
              "PHI", "GSQRT", "J13G", "J23G", "MAPF"):
        a = dual(m[n])
        keep.append(a)
        setattr(ms, n, dptr(a))
    ms.J33G = DualC(float(m["J33G"]), 0.0)
    return ms


STAGE_OUT = ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK")


def stage_tangent(case, inp: dict, dtrk: float, tangents=None, scheme=None):
    """One HEVI stage (orc_tstep_short_hevi) on dual numbers.  inp as for
    pyoracle.tstep_short_hevi; tangents maps input keys (DENS0..RHOT0,
    DENS..RHOT, DPRES0, RT2P, num_diff, REF_rhot, ...) to directions (absent
    keys: zero).  The metrics are case.metric, held fixed.  Returns
    ({DENS_RK..RHOT_RK: value}, {DENS_RK..RHOT_RK: tangent})."""
    g = case.grid
    t = tangents or {}
    keep = []

    def d(key, v=None):
        a = dual(inp[key] if v is None else v, t.get(key))
        keep.append(a)
        return dptr(a)

    ms = _metric_dual(case.metric, keep)
    outs = {k: dual(np.full(g.shape, np.nan)) for k in STAGE_OUT}
    mf, tf = dual(np.full(g.shape + (3,), np.nan)), dual(np.full(g.shape + (3,), np.nan))
    ins = ("DENS0", "MOMZ0", "MOMX0", "MOMY0", "RHOT0", "DENS", "MOMZ", "MOMX", "MOMY", "RHOT",
           "DENS_t", "MOMZ_t", "MOMX_t", "MOMY_t", "RHOT_t", "DPRES0", "RT2P")
    rc = lib().orc_tstep_short_hevi(
        C.byref(po.grid_struct(g)), C.c_int(scheme if scheme is not None else case.scheme),
        *[dptr(outs[k]) for k in STAGE_OUT], dptr(mf), dptr(tf), *[d(k) for k in ins],
        d("CORIOLI", inp.get("CORIOLI", case.corioli)), d("num_diff"),
        d("wdamp_coef", inp.get("wdamp_coef", case.wdamp_coef)), DualC(float(inp.get("divdmp_coef", 0.0)), 0.0),
        d("DDIV"), C.byref(ms), d("REF_dens"), d("REF_rhot"), DualC(float(dtrk), 0.0))
    if rc != 0:
        raise RuntimeError("orc_tstep_short_hevi (dual) failed")
    return ({k: np.asfortranarray(outs[k][0]) for k in STAGE_OUT},
            {k: np.asfortranarray(outs[k][1]) for k in STAGE_OUT})


class TangentModel:
    """The oracle's Model (po.Model) on dual numbers: step() advances the
    state and its tangent together."""

    def __init__(self, case, state: dict, dstate: dict, dqtrc=None, **model_kw):
        g = case.grid
        self.case = case
        self._keep = []
        pm = po.Model(case, **model_kw)
        self.cfg = OrcConfigD()
        for n, t in po.OrcConfig._fields_:
            val = getattr(pm.cfg, n)
            setattr(self.cfg, n, DualC(val, 0.0) if t is C.c_double else val)
        self.x = {k: dual(state[k], dstate.get(k)) for k in po.STATE}
        qa = case.qtrc.shape[-1]
        self.qa = qa
        q0 = case.qtrc if qa > 0 else np.zeros(g.shape + (1,))
        self.q = dual(q0, dqtrc)
        self.ms = _metric_dual(case.metric, self._keep)
        st = OrcStateD()
        for k in po.STATE:
            setattr(st, k, dptr(self.x[k]))
        st.QTRC = dptr(self.q)
        st.QA = qa
        for k in ("REF_dens", "REF_pott", "REF_pres"):
            a = dual(case.ref[k])
            self._keep.append(a)
            setattr(st, k, dptr(a))
        for k in ("DENS_tp", "MOMZ_tp", "MOMX_tp", "MOMY_tp", "RHOT_tp"):
            a = dual(case.zero_tend[k])
            self._keep.append(a)
            setattr(st, k, dptr(a))
        for k, v in (("RHOQ_tp", np.zeros(g.shape + (max(qa, 1),))), ("CORIOLI", case.corioli),
                     ("wdamp_coef", case.wdamp_coef), ("mflx_hi", np.full(g.shape + (3,), np.nan)),
                     ("tflx_hi", np.full(g.shape + (3,), np.nan))):
            a = dual(v)
            self._keep.append(a)
            setattr(st, k, dptr(a))
        self.st = st
        self.gs = po.grid_struct(g)

    def step(self, n=1):
        for _ in range(n):
            rc = lib().orc_large_step(C.byref(self.gs), C.byref(self.ms), C.byref(self.cfg), C.byref(self.st))
            if rc != 0:
                raise RuntimeError("orc_large_step (dual) failed")

    def values(self):
        return {k: np.asfortranarray(self.x[k][0]) for k in po.STATE}

    def tangents(self):
        return {k: np.asfortranarray(self.x[k][1]) for k in po.STATE}

    def qtrc(self):
        return np.asfortranarray(self.q[0]), np.asfortranarray(self.q[1])


def _tracer_call(fn, case, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff_q, dtl, fct, along_stream,
                 scheme, metric, tangents):
    g = case.grid
    m = metric if metric is not None else case.metric
    t = tangents or {}
    keep = []

    def d(v, key=None):
        a = dual(v, t.get(key))
        keep.append(a)
        return dptr(a)

    mapf_xy = np.asfortranarray(m["MAPF"][:, :, :, 0])
    q = dual(QTRC, t.get("QTRC"))
    args = [d(QTRC0, "QTRC0"), d(RHOQ_t, "RHOQ_t"), d(DENS0), d(DENS), d(mflx_hi), d(num_diff_q, "num_diff"),
            d(m["GSQRT"]), d(mapf_xy), d(m["CDZ"]), d(m["RCDZ"]), d(m["RCDX"]), d(m["RCDY"]), DualC(float(dtl), 0.0),
            C.c_int(int(fct)), C.c_int(int(along_stream))]
    sch = C.c_int(scheme if scheme is not None else case.scheme_tracer)
    if fn == "stage":
        out = dual(np.full(g.shape, np.nan))
        rc = lib().orc_tstep_tracer_heve(C.byref(po.grid_struct(g)), sch, dptr(out), dptr(q), *args)
    else:
        out = q
        rc = lib().orc_tinteg_tracer_rk3(C.byref(po.grid_struct(g)), sch, dptr(out), *args)
    if rc != 0:
        raise RuntimeError(f"tracer {fn} (dual) failed")
    return np.asfortranarray(out[0]), np.asfortranarray(out[1])


def tracer_stage_tangent(case, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff_q, dtl, fct, along_stream=True,
                         scheme=None, metric=None, tangents=None):
    """One tracer stage (orc_tstep_tracer_heve) on dual numbers: returns
    (QTRCo, its tangent) along the directions tangents["QTRC" | "QTRC0" |
    "RHOQ_t" | "num_diff"] (absent keys: zero), the mass fluxes, densities and
    metrics held fixed and the FCT ratios R+/R- frozen at their forward
    values."""
    return _tracer_call("stage", case, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff_q, dtl, fct, along_stream,
                        scheme, metric, tangents)


def tracer_rk3_tangent(case, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff_q, dtl, fct, along_stream=True,
                       scheme=None, metric=None, tangents=None):
    """The three-stage tracer integration (orc_tinteg_tracer_rk3) on dual
    numbers, as tracer_stage_tangent: returns (QTRC after, its tangent)."""
    return _tracer_call("rk3", case, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff_q, dtl, fct, along_stream,
                        scheme, metric, tangents)
