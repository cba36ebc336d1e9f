"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the scale_amd package.

Arrays are NumPy float64 in Fortran order with the reference's shapes
((KA,IA,JA), (KA,IA,JA,3), (KA,IA,JA,5,3), (IA,JA,2,4) ...).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

UD1, CD2, UD3, CD4, UD5, CD6 = 1, 2, 3, 4, 5, 6

_dp = C.POINTER(C.c_double)


class OrcGrid(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("KA", "IA", "JA", "KS", "KE", "IS", "IE", "JS", "JE")]


class OrcMetric(C.Structure):
    _fields_ = [(n, _dp) for n in ("CDZ", "FDZ", "RCDZ", "RFDZ", "CDX", "FDX", "RCDX", "RFDX",
                                   "CDY", "FDY", "RCDY", "RFDY", "PHI", "GSQRT", "J13G", "J23G", "MAPF")] + [("J33G", C.c_double)]


class OrcState(C.Structure):
    _fields_ = [("DENS", _dp), ("MOMZ", _dp), ("MOMX", _dp), ("MOMY", _dp), ("RHOT", _dp),
                ("QTRC", _dp), ("QA", C.c_int),
                ("REF_dens", _dp), ("REF_pott", _dp), ("REF_pres", _dp),
                ("DENS_tp", _dp), ("MOMZ_tp", _dp), ("MOMX_tp", _dp), ("MOMY_tp", _dp), ("RHOT_tp", _dp), ("RHOQ_tp", _dp),
                ("CORIOLI", _dp), ("wdamp_coef", _dp), ("mflx_hi", _dp), ("tflx_hi", _dp)]


class OrcConfig(C.Structure):
    _fields_ = [("scheme", C.c_int), ("scheme_tracer", C.c_int), ("tinteg_short", C.c_int),
                ("dtl", C.c_double), ("dts", C.c_double), ("nd_coef", C.c_double), ("nd_sfc_fact", C.c_double),
                ("nd_coef_q", C.c_double), ("nd_order", C.c_int), ("nd_use_rs", C.c_int),
                ("divdmp_coef", C.c_double), ("flag_fct_tracer", C.c_int), ("flag_fct_along_stream", C.c_int),
                ("short_type", C.c_int), ("tinteg_tracer", C.c_int)]

# ATMOS_DYN_TINTEG_SHORT_TYPE and ATMOS_DYN_TINTEG_TRACER_TYPE ids of the oracle
TINTEG_SHORT = {"RK4": 4, "RK3": 3, "RK3WS2002": 32}
TINTEG_TRACER = {"RK3WS2002": 0, "EULER": 1}


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_mass_sum.restype = C.c_double
        _lib.orc_face_value.restype = C.c_double
        _lib.orc_face_value.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5
        _lib.orc_face_value6.restype = C.c_double
        _lib.orc_face_value6.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.c_double]
        _lib.orc_get_threads.restype = C.c_int
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64
    assert a.flags.f_contiguous or a.ndim == 1, "oracle arrays must be Fortran-contiguous"
    return a.ctypes.data_as(_dp)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def grid_struct(g) -> OrcGrid:
    return OrcGrid(g.KA, g.IA, g.JA, g.KS, g.KE, g.IS, g.IE, g.JS, g.JE)


def metric_struct(m: dict) -> OrcMetric:
    s = OrcMetric()
    for n in ("CDZ", "FDZ", "RCDZ", "RFDZ", "CDX", "FDX", "RCDX", "RFDX", "CDY", "FDY", "RCDY", "RFDY",
              "PHI", "GSQRT", "J13G", "J23G", "MAPF"):
        setattr(s, n, ptr(m[n]))
    s.J33G = float(m["J33G"])
    return s


def nan_like_outputs(g, n=None):
    shape = (g.KA, g.IA, g.JA) if n is None else (g.KA, g.IA, g.JA, n)
    return np.full(shape, np.nan, order="F")


def tstep_short_hevi(case, inp: dict, dtrk: float, scheme: int | None = None, out: dict | None = None) -> dict:
    """One HEVI stage.  inp holds DENS0..RHOT0, DENS..RHOT, DENS_t..RHOT_t, DPRES0, RT2P,
    num_diff (KA,IA,JA,5,3), DDIV, REF_dens, REF_rhot; optional CORIOLI, wdamp_coef, divdmp_coef."""
    g = case.grid
    L = lib()
    if out is None:
        out = {k: nan_like_outputs(g) for k in ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK")}
        out["mflx_hi"] = nan_like_outputs(g, 3)
        out["tflx_hi"] = nan_like_outputs(g, 3)
    gs, ms = grid_struct(g), metric_struct(case.metric)
    cor = inp.get("CORIOLI", case.corioli)
    wd = inp.get("wdamp_coef", case.wdamp_coef)
    rc = L.orc_tstep_short_hevi(
        C.byref(gs), C.c_int(scheme if scheme is not None else case.scheme),
        *[ptr(out[k]) for k in ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK", "mflx_hi", "tflx_hi")],
        *[ptr(inp[k]) for k in ("DENS0", "MOMZ0", "MOMX0", "MOMY0", "RHOT0", "DENS", "MOMZ", "MOMX", "MOMY", "RHOT",
                                "DENS_t", "MOMZ_t", "MOMX_t", "MOMY_t", "RHOT_t", "DPRES0", "RT2P")],
        ptr(cor), ptr(inp["num_diff"]), ptr(wd), C.c_double(inp.get("divdmp_coef", 0.0)), ptr(inp["DDIV"]),
        C.byref(ms), ptr(inp["REF_dens"]), ptr(inp["REF_rhot"]), C.c_double(dtrk))
    if rc != 0:
        raise RuntimeError(f"orc_tstep_short_hevi failed: {rc}")
    return out


def tstep_short_heve(case, inp: dict, dtrk: float, scheme: int | None = None, out: dict | None = None,
                     fct_momentum=False, fct_t=False, along_stream=True) -> dict:
    """One HEVE stage (scale_atmos_dyn_tstep_short_fvm_heve.F90), same inputs as tstep_short_hevi.
    With fct_t the outputs DENS_RK and mflx_hi also get their periodic halos (the
    exchanges of HEVE.F90:1390-1401)."""
    g = case.grid
    if out is None:
        out = {k: nan_like_outputs(g) for k in ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK")}
        out["mflx_hi"] = nan_like_outputs(g, 3)
        out["tflx_hi"] = nan_like_outputs(g, 3)
    gs, ms = grid_struct(g), metric_struct(case.metric)
    rc = lib().orc_tstep_short_heve(
        C.byref(gs), C.c_int(scheme if scheme is not None else case.scheme),
        *[ptr(out[k]) for k in ("DENS_RK", "MOMZ_RK", "MOMX_RK", "MOMY_RK", "RHOT_RK", "mflx_hi", "tflx_hi")],
        *[ptr(inp[k]) for k in ("DENS0", "MOMZ0", "MOMX0", "MOMY0", "RHOT0", "DENS", "MOMZ", "MOMX", "MOMY", "RHOT",
                                "DENS_t", "MOMZ_t", "MOMX_t", "MOMY_t", "RHOT_t", "DPRES0", "RT2P")],
        ptr(inp.get("CORIOLI", case.corioli)), ptr(inp["num_diff"]), ptr(inp.get("wdamp_coef", case.wdamp_coef)),
        C.c_double(inp.get("divdmp_coef", 0.0)), ptr(inp["DDIV"]), C.c_int(int(fct_momentum)), C.c_int(int(fct_t)),
        C.c_int(int(along_stream)),
        C.byref(ms), ptr(inp["REF_dens"]), ptr(inp["REF_rhot"]), C.c_double(dtrk))
    if rc != 0:
        raise RuntimeError(f"orc_tstep_short_heve failed: {rc}")
    return out


def numfilter_coef(case, state: dict, dt: float, nd_coef=1.0e-4, nd_sfc_fact=1.0) -> np.ndarray:
    g = case.grid
    nd = np.full((g.KA, g.IA, g.JA, 5, 3), np.nan, order="F")
    gs, ms = grid_struct(g), metric_struct(case.metric)
    rc = lib().orc_numfilter_coef(C.byref(gs), ptr(nd), *[ptr(state[k]) for k in ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")],
                                  C.byref(ms), C.c_double(dt), ptr(case.ref["REF_dens"]), ptr(case.ref["REF_pott"]),
                                  C.c_double(nd_coef), C.c_int(1), C.c_double(nd_sfc_fact), C.c_int(1))
    if rc != 0:
        raise RuntimeError("orc_numfilter_coef failed")
    return nd


def numfilter_coef_q(case, DENS, QTRC, dt: float, nd_coef=1.0e-4, nd_sfc_fact=1.0, is_qv=False, REF_qv=None) -> np.ndarray:
    g = case.grid
    nd = np.full((g.KA, g.IA, g.JA, 3), np.nan, order="F")
    gs, ms = grid_struct(g), metric_struct(case.metric)
    rc = lib().orc_numfilter_coef_q(C.byref(gs), ptr(nd), ptr(DENS), ptr(QTRC), C.c_int(int(is_qv)), C.byref(ms),
                                    C.c_double(dt), ptr(REF_qv), C.c_double(nd_coef), C.c_int(1),
                                    C.c_double(nd_sfc_fact), C.c_int(1))
    if rc != 0:
        raise RuntimeError("orc_numfilter_coef_q failed")
    return nd


def tstep_tracer(case, QTRCo, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff_q, dtl, fct: bool,
                 along_stream: bool = True, scheme: int | None = None):
    g = case.grid
    gs = grid_struct(g)
    m = case.metric
    mapf_xy = np.asfortranarray(m["MAPF"][:, :, :, 0])
    rc = lib().orc_tstep_tracer_heve(C.byref(gs), C.c_int(scheme if scheme is not None else case.scheme_tracer),
                                     ptr(QTRCo), ptr(QTRC), ptr(QTRC0), ptr(RHOQ_t), ptr(DENS0), ptr(DENS),
                                     ptr(mflx_hi), ptr(num_diff_q), ptr(m["GSQRT"]), ptr(mapf_xy), ptr(m["CDZ"]),
                                     ptr(m["RCDZ"]), ptr(m["RCDX"]), ptr(m["RCDY"]), C.c_double(dtl),
                                     C.c_int(int(fct)), C.c_int(int(along_stream)))
    if rc != 0:
        raise RuntimeError("orc_tstep_tracer_heve failed")
    return QTRCo


def face_value(scheme, kind, qm1, q0, q1, q2, vel) -> float:
    return lib().orc_face_value(scheme, kind, qm1, q0, q1, q2, vel)


def face_value6(scheme, kind, q, vel) -> float:
    """UD5/CD6 face values: kind 0 interior (q = val(k-2..k+3)), 1 face KS and
    3 face KS+1 (q = val(KS..KS+5)), 2 face KE-1 and 4 face KE-2 (q = val(KE-5..KE))."""
    qa = (C.c_double * 6)(*[float(x) for x in q])
    return lib().orc_face_value6(scheme, kind, qa, vel)


def get_fact_fct(rw, ru, rv) -> np.ndarray:
    f = np.zeros(18)
    lib().orc_get_fact_fct(ptr(f), C.c_double(rw), C.c_double(ru), C.c_double(rv))
    return f.reshape((2, 3, 3), order="F")  # fact[a, b+1, c+1]


def solve_direct(g, Cvec, F1, F2, F3):
    Cvec = np.ascontiguousarray(Cvec, dtype=np.float64).copy()
    gs = grid_struct(g)
    lib().orc_solve_direct(C.byref(gs), ptr(Cvec), ptr(np.ascontiguousarray(F1)), ptr(np.ascontiguousarray(F2)),
                           ptr(np.ascontiguousarray(F3)))
    return Cvec


def wdamp_setup(g, wdamp_tau, wdamp_height, FZ):
    """ATMOS_DYN_wdamp_setup; FZ holds FZ(0:KA)."""
    out = np.zeros(g.KA)
    lib().orc_wdamp_setup(C.byref(grid_struct(g)), ptr(out), C.c_double(wdamp_tau), C.c_double(wdamp_height),
                          ptr(np.ascontiguousarray(FZ, dtype=np.float64)))
    return out


def terrain_metrics(g, Zsfc, fz, metric: dict) -> dict:
    """orc_terrain_metrics: GSQRT, J13G, J23G (KA,IA,JA,7) and PHI from the
    surface height Zsfc(IA,JA) and the flat grid's FZ(0:KA); metric supplies
    RCDZ, RFDZ, RCDX, RFDX, RCDY, RFDY."""
    fz = np.ascontiguousarray(fz, dtype=np.float64)
    cz = np.ascontiguousarray(0.5 * (fz[1:] + fz[:-1]))
    out = dict(GSQRT=np.full((g.KA, g.IA, g.JA, 7), np.nan, order="F"),
               J13G=np.full((g.KA, g.IA, g.JA, 7), np.nan, order="F"),
               J23G=np.full((g.KA, g.IA, g.JA, 7), np.nan, order="F"),
               PHI=np.full((g.KA, g.IA, g.JA), np.nan, order="F"))
    lib().orc_terrain_metrics(C.byref(grid_struct(g)), ptr(np.asfortranarray(Zsfc, dtype=np.float64)), ptr(cz), ptr(fz),
                              *[ptr(np.ascontiguousarray(metric[k])) for k in ("RCDZ", "RFDZ", "RCDX", "RFDX", "RCDY", "RFDY")],
                              ptr(out["GSQRT"]), ptr(out["J13G"]), ptr(out["J23G"]), ptr(out["PHI"]))
    return out


def divergence(case, MOMZ, MOMX, MOMY, metric=None) -> np.ndarray:
    """ATMOS_DYN_divergence (i = IS..IE+1, j = JS..JE+1, k = KS-1..KE+1); NaN elsewhere."""
    g = case.grid
    out = np.full(g.shape, np.nan, order="F")
    lib().orc_divergence(C.byref(grid_struct(g)), ptr(out), ptr(MOMZ), ptr(MOMX), ptr(MOMY),
                         C.byref(metric_struct(metric or case.metric)))
    return out


def halo_periodic(g, a):
    gs = grid_struct(g)
    lib().orc_halo_periodic(C.byref(gs), ptr(a))
    return a


# ---- adjoint checker (scale_oracle_adj.c) -----------------------------------

STATE = ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")


def adj_halo_periodic(g, ab):
    lib().orc_adj_halo_periodic(C.byref(grid_struct(g)), ptr(ab))
    return ab


def adj_tstep_short_hevi(case, inp: dict, dtrk: float, out_b: dict, scheme: int | None = None) -> dict:
    """Adjoint of one flat-terrain HEVI stage.  out_b holds DENS_RK..RHOT_RK adjoints;
    returns the adjoints of DENS0..RHOT0, DENS..RHOT, num_diff, DPRES0, RT2P, REF_rhot."""
    g = case.grid
    z = lambda *extra: np.zeros((g.KA, g.IA, g.JA) + extra, order="F")
    res = {k + "0": z() for k in STATE}
    res.update({k: z() for k in STATE})
    res.update(num_diff=z(5, 3), DPRES0=z(), RT2P=z(), REF_rhot=z())
    gs, ms = grid_struct(g), metric_struct(case.metric)
    rc = lib().orc_adj_tstep_short_hevi(
        C.byref(gs), C.c_int(scheme if scheme is not None else case.scheme), C.byref(ms),
        *[ptr(inp[k]) for k in ("DENS0", "MOMZ0", "MOMX0", "MOMY0", "RHOT0", "DENS", "MOMZ", "MOMX", "MOMY", "RHOT",
                                "DENS_t", "MOMZ_t", "MOMX_t", "MOMY_t", "RHOT_t", "DPRES0", "RT2P")],
        ptr(inp.get("CORIOLI", case.corioli)), ptr(inp["num_diff"]), ptr(inp.get("wdamp_coef", case.wdamp_coef)),
        ptr(inp["REF_dens"]), ptr(inp["REF_rhot"]), C.c_double(dtrk),
        *[ptr(out_b[k + "_RK"]) for k in STATE],
        *[ptr(res[k + "0"]) for k in STATE], *[ptr(res[k]) for k in STATE],
        ptr(res["num_diff"]), ptr(res["DPRES0"]), ptr(res["RT2P"]), ptr(res["REF_rhot"]))
    if rc != 0:
        raise RuntimeError(f"orc_adj_tstep_short_hevi failed: {rc}")
    return res


def adj_numfilter_coef(case, state: dict, dt: float, nd_b: np.ndarray, nd_coef=1.0e-4, nd_sfc_fact=1.0) -> dict:
    g = case.grid
    nd_b = np.array(nd_b, order="F", copy=True)
    res = {k: np.zeros((g.KA, g.IA, g.JA), order="F") for k in STATE}
    gs, ms = grid_struct(g), metric_struct(case.metric)
    rc = lib().orc_adj_numfilter_coef(C.byref(gs), C.byref(ms), C.c_double(dt), *[ptr(state[k]) for k in STATE],
                                      ptr(case.ref["REF_dens"]), ptr(case.ref["REF_pott"]), C.c_double(nd_coef),
                                      C.c_double(nd_sfc_fact), ptr(nd_b), *[ptr(res[k]) for k in STATE])
    if rc != 0:
        raise RuntimeError("orc_adj_numfilter_coef failed")
    return res


def adj_large_step(model: "Model", xb: dict) -> dict:
    """Adjoint of one Model.step() from the model's current state: xb holds the
    adjoint of the state after the step (whole arrays); returns the adjoint of
    the state before it.  The model state is not changed."""
    g = model.case.grid
    out = {k: np.array(xb[k], order="F", copy=True) for k in STATE}
    arr = (_dp * 5)(*[ptr(out[k]) for k in STATE])
    gs, ms = grid_struct(g), metric_struct(model.case.metric)
    st = model._st()
    rc = lib().orc_adj_large_step(C.byref(gs), C.byref(ms), C.byref(model.cfg), C.byref(st), arr)
    if rc != 0:
        raise RuntimeError("orc_adj_large_step failed (flat terrain, one short step, divdmp 0 only)")
    return out


class Model:
    """Whole-model integrator (LARGE.F90 on-path subset) over a Case; short =
    "hevi" or "heve" selects the stage of the RK4 short step (ATMOS_DYN_TYPE)."""

    def __init__(self, case, nd_coef=1.0e-4, divdmp_coef=0.0, fct=None, along_stream=True, nd_coef_q=0.0,
                 short="hevi", dtl=None, tinteg_short="RK4", tinteg_tracer="RK3WS2002"):
        g = case.grid
        self.case = case
        self.state = {k: np.array(v, order="F", copy=True) for k, v in case.state.items()}
        qa = case.qtrc.shape[-1]
        self.qa = qa
        self.qtrc = np.array(case.qtrc if qa > 0 else np.zeros(g.shape + (1,)), order="F", copy=True)
        self.mflx_hi = np.full(g.shape + (3,), np.nan, order="F")
        self.tflx_hi = np.full(g.shape + (3,), np.nan, order="F")
        self.rhoq_tp = np.zeros(g.shape + (max(qa, 1),), order="F")
        # dtl: the large step (case.dt, the short step, by default: nstep = 1)
        self.cfg = OrcConfig(case.scheme, case.scheme_tracer, TINTEG_SHORT[tinteg_short],
                             case.dt if dtl is None else dtl, case.dt, nd_coef, 1.0, nd_coef_q, 1, 1,
                             divdmp_coef, int(case.fct if fct is None else fct), int(along_stream),
                             {"hevi": 0, "heve": 1}[short], TINTEG_TRACER[tinteg_tracer])
        self._keep = []

    def _st(self):
        c = self.case
        s = OrcState()
        for k in ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT"):
            setattr(s, k, ptr(self.state[k]))
        s.QTRC = ptr(self.qtrc)
        s.QA = self.qa
        s.REF_dens, s.REF_pott, s.REF_pres = (ptr(c.ref[k]) for k in ("REF_dens", "REF_pott", "REF_pres"))
        for k in ("DENS_tp", "MOMZ_tp", "MOMX_tp", "MOMY_tp", "RHOT_tp"):
            setattr(s, k, ptr(c.zero_tend[k]))
        s.RHOQ_tp = ptr(self.rhoq_tp)
        s.CORIOLI = ptr(c.corioli)
        s.wdamp_coef = ptr(c.wdamp_coef)
        s.mflx_hi = ptr(self.mflx_hi)
        s.tflx_hi = ptr(self.tflx_hi)
        return s

    def step(self, n=1):
        g = self.case.grid
        gs, ms = grid_struct(g), metric_struct(self.case.metric)
        st = self._st()
        for _ in range(n):
            rc = lib().orc_large_step(C.byref(gs), C.byref(ms), C.byref(self.cfg), C.byref(st))
            if rc != 0:
                raise RuntimeError("orc_large_step failed")

    def mass(self, a):
        g = self.case.grid
        return lib().orc_mass_sum(C.byref(grid_struct(g)), C.byref(metric_struct(self.case.metric)), ptr(a))
