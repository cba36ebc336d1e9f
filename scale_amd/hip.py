"""Python front end of libscalehip.so, mirroring the reference's plug-in
interfaces for the short-step dynamical core:

* ``ATMOS_DYN_Tstep_short_fvm_hevi`` -- abstract interface ``short`` of
  scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short.F90:38-128;
* ``ATMOS_DYN_Tstep_tracer_fvm_heve`` -- interface ``step`` of
  scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_tracer.F90:34-64;
* ``ATMOS_DYN_numfilter_coef`` --
  scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:385-985;
* ``Model`` -- the device-resident large step of
  scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_large_fvm_heve.F90:171-1130.

Arguments keep the reference's names, order and shapes (NumPy float64 arrays in
Fortran order).  Errors raise ``ScaleHipError`` with the library's message, the
counterpart of the reference's LOG_ERROR + PRC_abort.  There is no CPU
fallback: without a loadable libscalehip.so every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import inputs

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscalehip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

UD3, CD4 = 3, 4
UD3KOREN1993, UD1, CD2, UD5, CD6 = 31, 1, 2, 5, 6  # tracer schemes (UD5 and CD6 need a lateral halo of 3)
FIELD = dict(DENS=0, MOMZ=1, MOMX=2, MOMY=3, RHOT=4, REF_dens=10, REF_pott=11, REF_pres=12, CORIOLI=13,
             wdamp_coef=14, mflx_hi=20, tflx_hi=21, DENS_tp=30, MOMZ_tp=31, MOMX_tp=32, MOMY_tp=33, RHOT_tp=34)
QTRC_BASE, RHOQ_TP_BASE = 100, 200
STATE = ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")


class ScaleHipError(RuntimeError):
    pass


class HipGrid(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("KA", "IA", "JA", "KS", "KE", "IS", "IE", "JS", "JE",
                                          "KHALO", "IHALO", "JHALO", "flux_scheme", "flux_scheme_tracer", "device")]


class DynConfig(C.Structure):
    _fields_ = [("dtl", C.c_double), ("dts", C.c_double), ("nd_coef", C.c_double), ("nd_sfc_fact", C.c_double),
                ("divdmp_coef", C.c_double), ("nd_order", C.c_int32), ("fct_tracer", C.c_int32),
                ("fct_along_stream", C.c_int32), ("mixed_precision", C.c_int32), ("nd_coef_q", C.c_double),
                ("short_type", C.c_int32), ("tinteg_short", C.c_int32), ("tinteg_tracer", C.c_int32)]


# ATMOS_DYN_TINTEG_SHORT_TYPE / ATMOS_DYN_TINTEG_TRACER_TYPE ids of scale_hip_dyn_config
TINTEG_SHORT = {"RK4": 4, "RK3": 3, "RK3WS2002": 32}
TINTEG_TRACER = {"RK3WS2002": 0, "EULER": 1}
TUNE_NF_TJ = 1  # scale_hip_tune_set knob: rows of the numerical-filter tile (0 = from the grid)
TUNE_TRACER_TILE = 2  # knob: FCT tracer stage on LDS tiles: 1 (default) on large grids, 2 always, 0 never
TUNE_ADJ_TAPE = 3  # knob: trailing segments whose adjoint tapes stay in HBM (0: half the free memory's worth)


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libscalehip.so; raises if it is missing (no silent fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ScaleHipError(f"{path} not built; run __graft_entry__.build() or python scale_amd/build.py")
        _lib = C.CDLL(path)
        _lib.scale_hip_last_error.restype = C.c_char_p
        _lib.scale_hip_version.restype = C.c_char_p
    return _lib


def _check(rc: int):
    if rc != 0:
        msg = load_library().scale_hip_last_error().decode()
        raise ScaleHipError(f"libscalehip error {rc}: {msg}")


def _p(a):
    if a is None:
        return None
    if not isinstance(a, np.ndarray) or a.dtype != np.float64:
        raise TypeError("arrays must be float64 NumPy arrays")
    if not (a.flags.f_contiguous or a.ndim <= 1):
        raise ValueError("arrays must be Fortran-contiguous (the reference's layout)")
    return a.ctypes.data_as(_dp)


def _d(x):
    return C.byref(C.c_double(float(x)))


def _l(x):
    return C.byref(C.c_int32(1 if x else 0))


def ATMOS_DYN_wdamp_setup(grid, wdamp_tau, wdamp_height, FZ):
    """Rayleigh damping coefficient wdamp_coef(KA) from FZ(0:KA)
    (scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:323-381); host only."""
    g = HipGrid(grid.KA, grid.IA, grid.JA, grid.KS + 1, grid.KE + 1, grid.IS + 1, grid.IE + 1, grid.JS + 1,
                grid.JE + 1, grid.KS, grid.IS, grid.JS, UD3, UD3, 0)
    out = np.zeros(grid.KA)
    fz = np.ascontiguousarray(FZ, dtype=np.float64)
    _check(load_library().scale_hip_wdamp_setup(C.byref(g), _p(out), _d(wdamp_tau), _d(wdamp_height), _p(fz)))
    return out


class Context:
    """One device context (scale_hip_init / scale_hip_finalize)."""

    def __init__(self, grid, scheme=UD3, scheme_tracer=UD3, device=0):
        L = load_library()
        g = HipGrid(grid.KA, grid.IA, grid.JA, grid.KS + 1, grid.KE + 1, grid.IS + 1, grid.IE + 1, grid.JS + 1,
                    grid.JE + 1, grid.KS, grid.IS, grid.JS, scheme, scheme_tracer, device)
        h = C.c_void_p()
        _check(L.scale_hip_init(C.byref(g), C.byref(h)))
        self._h = h
        self.grid = grid
        self.has_model = False

    def close(self):
        if self._h:
            load_library().scale_hip_finalize(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- drop-in calls (host Fortran arrays) ----
    def ATMOS_DYN_Tstep_short_fvm_hevi(self, DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, PROG_RK, mflx_hi, tflx_hi,
                                       DENS0, MOMZ0, MOMX0, MOMY0, RHOT0, DENS, MOMZ, MOMX, MOMY, RHOT,
                                       DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, PROG0, PROG,
                                       DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV,
                                       FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM,
                                       CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,
                                       PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot,
                                       BND_W, BND_E, BND_S, BND_N, dtrk, last):
        L = load_library()
        _check(L.scale_hip_tstep_short_hevi(
            self._h, _p(DENS_RK), _p(MOMZ_RK), _p(MOMX_RK), _p(MOMY_RK), _p(RHOT_RK), _p(PROG_RK), _p(mflx_hi), _p(tflx_hi),
            _p(DENS0), _p(MOMZ0), _p(MOMX0), _p(MOMY0), _p(RHOT0), _p(DENS), _p(MOMZ), _p(MOMX), _p(MOMY), _p(RHOT),
            _p(DENS_t), _p(MOMZ_t), _p(MOMX_t), _p(MOMY_t), _p(RHOT_t), _p(PROG0), _p(PROG),
            _p(DPRES0), _p(RT2P), _p(CORIOLI), _p(num_diff), _p(wdamp_coef), _d(divdmp_coef), _p(DDIV),
            _l(FLAG_FCT_MOMENTUM), _l(FLAG_FCT_T), _l(FLAG_FCT_ALONG_STREAM),
            _p(CDZ), _p(FDZ), _p(FDX), _p(FDY), _p(RCDZ), _p(RCDX), _p(RCDY), _p(RFDZ), _p(RFDX), _p(RFDY),
            _p(PHI), _p(GSQRT), _p(J13G), _p(J23G), _d(J33G), _p(MAPF), _p(REF_dens), _p(REF_rhot),
            _l(BND_W), _l(BND_E), _l(BND_S), _l(BND_N), _d(dtrk), _l(last)))

    def ATMOS_DYN_Tstep_short_fvm_heve(self, DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, PROG_RK, mflx_hi, tflx_hi,
                                       DENS0, MOMZ0, MOMX0, MOMY0, RHOT0, DENS, MOMZ, MOMX, MOMY, RHOT,
                                       DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, PROG0, PROG,
                                       DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV,
                                       FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM,
                                       CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,
                                       PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot,
                                       BND_W, BND_E, BND_S, BND_N, dtrk, last):
        """One explicit HEVE stage (scale_atmos_dyn_tstep_short_fvm_heve.F90:109-1516);
        FLAG_FCT_MOMENTUM and FLAG_FCT_T are applied as in the reference."""
        L = load_library()
        _check(L.scale_hip_tstep_short_heve(
            self._h, _p(DENS_RK), _p(MOMZ_RK), _p(MOMX_RK), _p(MOMY_RK), _p(RHOT_RK), _p(PROG_RK), _p(mflx_hi), _p(tflx_hi),
            _p(DENS0), _p(MOMZ0), _p(MOMX0), _p(MOMY0), _p(RHOT0), _p(DENS), _p(MOMZ), _p(MOMX), _p(MOMY), _p(RHOT),
            _p(DENS_t), _p(MOMZ_t), _p(MOMX_t), _p(MOMY_t), _p(RHOT_t), _p(PROG0), _p(PROG),
            _p(DPRES0), _p(RT2P), _p(CORIOLI), _p(num_diff), _p(wdamp_coef), _d(divdmp_coef), _p(DDIV),
            _l(FLAG_FCT_MOMENTUM), _l(FLAG_FCT_T), _l(FLAG_FCT_ALONG_STREAM),
            _p(CDZ), _p(FDZ), _p(FDX), _p(FDY), _p(RCDZ), _p(RCDX), _p(RCDY), _p(RFDZ), _p(RFDX), _p(RFDY),
            _p(PHI), _p(GSQRT), _p(J13G), _p(J23G), _d(J33G), _p(MAPF), _p(REF_dens), _p(REF_rhot),
            _l(BND_W), _l(BND_E), _l(BND_S), _l(BND_N), _d(dtrk), _l(last)))

    def ATMOS_DYN_Tstep_tracer_fvm_heve(self, QTRCo, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff,
                                        GSQRT, MAPF, CDZ, RCDZ, RCDX, RCDY, dtl, FLAG_FCT_TRACER, FLAG_FCT_ALONG_STREAM):
        L = load_library()
        _check(L.scale_hip_tstep_tracer_heve(
            self._h, _p(QTRCo), _p(QTRC), _p(QTRC0), _p(RHOQ_t), _p(DENS0), _p(DENS), _p(mflx_hi), _p(num_diff),
            _p(GSQRT), _p(MAPF), _p(CDZ), _p(RCDZ), _p(RCDX), _p(RCDY), _d(dtl), _l(FLAG_FCT_TRACER),
            _l(FLAG_FCT_ALONG_STREAM)))

    def ATMOS_DYN_Tstep_tracer_fvm_heve_adj(self, QTRCo_b, QTRC, QTRC0, DENS0, DENS, mflx_hi, num_diff,
                                            GSQRT, MAPF, CDZ, RCDZ, RCDX, RCDY, dtl, FLAG_FCT_TRACER,
                                            FLAG_FCT_ALONG_STREAM) -> dict:
        """Adjoint of one tracer stage with the FCT limiter frozen at its
        forward R+/R- (scale_hip_tstep_tracer_heve_adj).  QTRCo_b is read at
        the interior; returns the adjoints QTRC, QTRC0, RHOQ_t and num_diff
        (KA,IA,JA,3) of the stage's inputs on the whole array."""
        g = self.grid
        z = lambda *e: np.zeros((g.KA, g.IA, g.JA) + e, order="F")
        res = dict(QTRC=z(), QTRC0=z(), RHOQ_t=z(), num_diff=z(3))
        _check(load_library().scale_hip_tstep_tracer_heve_adj(
            self._h, _p(np.asfortranarray(QTRCo_b)), _p(res["QTRC"]), _p(res["QTRC0"]), _p(res["RHOQ_t"]),
            _p(res["num_diff"]), _p(QTRC), _p(QTRC0), _p(DENS0), _p(DENS), _p(mflx_hi), _p(num_diff),
            _p(GSQRT), _p(MAPF), _p(CDZ), _p(RCDZ), _p(RCDX), _p(RCDY), _d(dtl), _l(FLAG_FCT_TRACER),
            _l(FLAG_FCT_ALONG_STREAM)))
        return res

    def ATMOS_DYN_Tinteg_tracer_rk3_adj(self, QTRC_b, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff,
                                        GSQRT, MAPF, CDZ, RCDZ, RCDX, RCDY, dtl, FLAG_FCT_TRACER,
                                        FLAG_FCT_ALONG_STREAM) -> dict:
        """Adjoint of the three-stage tracer integration
        ATMOS_DYN_tinteg_tracer_rk3 (scale_atmos_dyn_tinteg_tracer_rk3.F90:
        92-208) with the last stage's FCT limiter frozen: the stages of dtl/3
        and dtl/2 on DENS0 + (DENS - DENS0)/3 and /2 (no FCT), each followed
        by the periodic halo fill, and the stage of dtl on DENS.  The two
        intermediate fields are recomputed with the forward stage (the last
        stage's limiter reads the second); the three stage adjoints run in
        reverse with the halo fills transposed between them.  QTRC_b is read
        at the interior; returns the adjoints of QTRC, QTRC0, RHOQ_t and
        num_diff as ATMOS_DYN_Tstep_tracer_fvm_heve_adj does."""
        g = self.grid
        sl = (slice(g.KS, g.KE + 1), slice(g.IS - 1, g.IE + 2), slice(g.JS - 1, g.JE + 2))
        dens_rk = [np.array(DENS, order="F", copy=True), np.array(DENS, order="F", copy=True)]
        dens_rk[0][sl] = DENS0[sl] + (DENS[sl] - DENS0[sl]) / 3.0  # TRK3:143-147
        dens_rk[1][sl] = DENS0[sl] + (DENS[sl] - DENS0[sl]) * 0.5  # TRK3:171-175
        met = (GSQRT, MAPF, CDZ, RCDZ, RCDX, RCDY)
        fields = [np.asfortranarray(QTRC)]
        for n, dt in enumerate((dtl / 3.0, dtl / 2.0)):
            out = np.zeros((g.KA, g.IA, g.JA), order="F")
            self.ATMOS_DYN_Tstep_tracer_fvm_heve(out, fields[-1], QTRC0, RHOQ_t, DENS0, dens_rk[n], mflx_hi, num_diff,
                                                 *met, dt, False, FLAG_FCT_ALONG_STREAM)
            inputs.halo_fill(g, out)
            fields.append(out)
        stages = [(fields[2], DENS, dtl, FLAG_FCT_TRACER), (fields[1], dens_rk[1], dtl / 2.0, False),
                  (fields[0], dens_rk[0], dtl / 3.0, False)]
        res = None
        ob = np.asfortranarray(QTRC_b)
        for n, (q, d, dt, fct) in enumerate(stages):
            r = self.ATMOS_DYN_Tstep_tracer_fvm_heve_adj(ob, q, QTRC0, DENS0, d, mflx_hi, num_diff, *met, dt, fct,
                                                         FLAG_FCT_ALONG_STREAM)
            if res is None:
                res = r
            else:
                for k in ("QTRC0", "RHOQ_t", "num_diff"):
                    res[k] += r[k]
                res["QTRC"] = r["QTRC"]
            if n < 2:
                ob = inputs.halo_fill_adjoint(g, r["QTRC"].copy())
        return res

    def ATMOS_DYN_numfilter_coef(self, num_diff, DENS, MOMZ, MOMX, MOMY, RHOT, CDZ, CDX, CDY, FDZ, FDX, FDY, DT,
                                 REF_dens, REF_pott, ND_COEF, ND_ORDER, ND_SFC_FACT, ND_USE_RS):
        L = load_library()
        _check(L.scale_hip_numfilter_coef(
            self._h, _p(num_diff), _p(DENS), _p(MOMZ), _p(MOMX), _p(MOMY), _p(RHOT), _p(CDZ), _p(CDX), _p(CDY),
            _p(FDZ), _p(FDX), _p(FDY), _d(DT), _p(REF_dens), _p(REF_pott), _d(ND_COEF),
            C.byref(C.c_int32(int(ND_ORDER))), _d(ND_SFC_FACT), _l(ND_USE_RS)))

    def ATMOS_DYN_numfilter_coef_q(self, num_diff_q, DENS, QTRC, is_qv, CDZ, CDX, CDY, dt, REF_qv, iq,
                                   ND_COEF, ND_ORDER, ND_SFC_FACT, ND_USE_RS):
        _check(load_library().scale_hip_numfilter_coef_q(
            self._h, _p(num_diff_q), _p(DENS), _p(QTRC), _l(is_qv), _p(CDZ), _p(CDX), _p(CDY), _d(dt),
            _p(REF_qv), C.byref(C.c_int32(int(iq))), _d(ND_COEF), C.byref(C.c_int32(int(ND_ORDER))),
            _d(ND_SFC_FACT), _l(ND_USE_RS)))

    def ATMOS_DYN_Tstep_short_fvm_hevi_adj(self, out_b: dict, DENS0, MOMZ0, MOMX0, MOMY0, RHOT0,
                                           DENS, MOMZ, MOMX, MOMY, RHOT, DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t,
                                           DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef,
                                           CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,
                                           PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot, dtrk) -> dict:
        """Adjoint of one HEVI stage (flat or terrain-following metrics, FP64).  out_b holds the adjoints
        DENS_RK..RHOT_RK of the stage outputs; returns the adjoints of the inputs
        (DENS0..RHOT0, DENS..RHOT, num_diff, DPRES0, RT2P, REF_rhot)."""
        g = self.grid
        z = lambda *e: np.zeros((g.KA, g.IA, g.JA) + e, order="F")
        res = {k + "0": z() for k in STATE}
        res.update({k: z() for k in STATE})
        res.update(num_diff=z(5, 3), DPRES0=z(), RT2P=z(), REF_rhot=z())
        _check(load_library().scale_hip_tstep_short_hevi_adj(
            self._h, *[_p(np.asfortranarray(out_b[k + "_RK"])) for k in STATE],
            *[_p(res[k + "0"]) for k in STATE], *[_p(res[k]) for k in STATE],
            _p(res["num_diff"]), _p(res["DPRES0"]), _p(res["RT2P"]), _p(res["REF_rhot"]),
            _p(DENS0), _p(MOMZ0), _p(MOMX0), _p(MOMY0), _p(RHOT0), _p(DENS), _p(MOMZ), _p(MOMX), _p(MOMY), _p(RHOT),
            _p(DENS_t), _p(MOMZ_t), _p(MOMX_t), _p(MOMY_t), _p(RHOT_t), _p(DPRES0), _p(RT2P), _p(CORIOLI),
            _p(num_diff), _p(wdamp_coef),
            _p(CDZ), _p(FDZ), _p(FDX), _p(FDY), _p(RCDZ), _p(RCDX), _p(RCDY), _p(RFDZ), _p(RFDX), _p(RFDY),
            _p(PHI), _p(GSQRT), _p(J13G), _p(J23G), _d(J33G), _p(MAPF), _p(REF_dens), _p(REF_rhot), _d(dtrk)))
        return res

    # ---- device-resident model ----
    def set_metrics(self, m: dict):
        _check(load_library().scale_hip_set_metrics(
            self._h, *[_p(m[k]) for k in ("CDZ", "FDZ", "RCDZ", "RFDZ", "CDX", "FDX", "RCDX", "RFDX",
                                          "CDY", "FDY", "RCDY", "RFDY", "PHI", "GSQRT", "J13G", "J23G")],
            _d(m["J33G"]), _p(m["MAPF"])))

    def set_terrain(self, m: dict, fz, zsfc):
        """Terrain-following metrics computed on the device (scale_hip_set_terrain)
        from the surface height zsfc(IA,JA), the reference face heights fz(0:KA)
        (cell centres halfway) and the 1-D grid arrays and MAPF of m."""
        fz = np.ascontiguousarray(fz, dtype=np.float64)
        cz = np.ascontiguousarray(0.5 * (fz[1:] + fz[:-1]))
        _check(load_library().scale_hip_set_terrain(
            self._h, *[_p(np.ascontiguousarray(m[k])) for k in ("CDZ", "FDZ", "RCDZ", "RFDZ", "CDX", "FDX", "RCDX",
                                                                  "RFDX", "CDY", "FDY", "RCDY", "RFDY")],
            _p(cz), _p(fz), _p(np.asfortranarray(zsfc, dtype=np.float64)), _p(np.asfortranarray(m["MAPF"]))))

    def get_metrics(self):
        """GSQRT, J13G, J23G (KA,IA,JA,7) and PHI (KA,IA,JA) as the context holds them."""
        g = self.grid
        out = dict(GSQRT=np.zeros((g.KA, g.IA, g.JA, 7), order="F"), J13G=np.zeros((g.KA, g.IA, g.JA, 7), order="F"),
                   J23G=np.zeros((g.KA, g.IA, g.JA, 7), order="F"), PHI=np.zeros((g.KA, g.IA, g.JA), order="F"))
        _check(load_library().scale_hip_get_metrics(self._h, _p(out["GSQRT"]), _p(out["J13G"]), _p(out["J23G"]),
                                                    _p(out["PHI"])))
        return out

    def model_create(self, dtl, dts, QA, nd_coef=1.0e-4, nd_sfc_fact=1.0, fct_tracer=False, fct_along_stream=True,
                     mixed_precision=False, divdmp_coef=0.0, nd_coef_q=0.0, short="hevi", tinteg_short="RK4",
                     tinteg_tracer="RK3WS2002"):
        cfg = DynConfig(dtl, dts, nd_coef, nd_sfc_fact, divdmp_coef, 1, int(fct_tracer), int(fct_along_stream),
                        int(mixed_precision), nd_coef_q, {"hevi": 0, "heve": 1}[short], TINTEG_SHORT[tinteg_short],
                        TINTEG_TRACER[tinteg_tracer])
        _check(load_library().scale_hip_model_create(self._h, C.byref(cfg), C.c_int32(QA)))
        self.has_model = True
        self.QA = QA

    def model_set(self, name, arr):
        fid = name if isinstance(name, int) else FIELD[name]
        _check(load_library().scale_hip_model_set(self._h, C.c_int32(fid), _p(np.asfortranarray(arr, dtype=np.float64))))

    def model_get(self, name, out=None):
        g = self.grid
        fid = name if isinstance(name, int) else FIELD[name]
        ncomp = 3 if fid in (20, 21) else 1
        if out is None:
            out = np.zeros((g.KA, g.IA, g.JA) + ((3,) if ncomp == 3 else ()), order="F")
        _check(load_library().scale_hip_model_get(self._h, C.c_int32(fid), _p(out)))
        return out

    def model_restart_write(self, path):
        """The model's restart file (scale_hip_model_restart_write)."""
        _check(load_library().scale_hip_model_restart_write(self._h, os.fsencode(path)))

    def model_restart_read(self, path):
        """The model's state and tracers from a restart file (scale_hip_model_restart_read)."""
        _check(load_library().scale_hip_model_restart_read(self._h, os.fsencode(path)))

    def model_large_step(self, n=1):
        _check(load_library().scale_hip_model_large_step(self._h, C.c_int32(n)))

    def model_prepare(self):
        _check(load_library().scale_hip_model_prepare(self._h))

    def model_short_step(self, n=1):
        _check(load_library().scale_hip_model_short_step(self._h, C.c_int32(n)))

    def model_forward_checkpointed(self, n, every):
        """n large steps keeping the state every `every` steps for model_adjoint."""
        _check(load_library().scale_hip_model_forward_checkpointed(self._h, C.c_int32(n), C.c_int32(every)))

    def model_adjoint(self, xb: dict) -> dict:
        """Adjoint of the last model_forward_checkpointed run: xb holds the adjoints
        of DENS..RHOT after it (whole arrays); returns their adjoints at its start."""
        out = {k: np.array(xb[k], dtype=np.float64, order="F", copy=True) for k in STATE}
        _check(load_library().scale_hip_model_adjoint(self._h, *[_p(out[k]) for k in STATE]))
        return out

    def model_mass(self, name, weight_dens=False):
        fid = name if isinstance(name, int) else FIELD[name]
        out = C.c_double()
        _check(load_library().scale_hip_model_mass(self._h, C.c_int32(fid), C.c_int32(int(weight_dens)), C.byref(out)))
        return out.value

    def profile(self, on=True):
        """on: True/1 times every launch, 2 only the outermost of nested timers
        (whole calls such as model_large_step), False stops profiling."""
        _check(load_library().scale_hip_profile_enable(self._h, C.c_int32(int(on))))

    def profile_query(self, name):
        cnt = C.c_int64()
        ms = C.c_double()
        _check(load_library().scale_hip_profile_query(self._h, name.encode(), C.byref(cnt), C.byref(ms)))
        return cnt.value, ms.value

    def tune_set(self, knob, value):
        """Set a launch-shape knob (scale_hip_tune_set); results do not depend on it."""
        _check(load_library().scale_hip_tune_set(self._h, C.c_int32(int(knob)), C.c_int32(int(value))))

    def tune_get(self, knob):
        v = C.c_int32()
        _check(load_library().scale_hip_tune_get(self._h, C.c_int32(int(knob)), C.byref(v)))
        return v.value

    def synchronize(self):
        _check(load_library().scale_hip_synchronize(self._h))

    def hbm_copy_probe(self, n, reps=10):
        """Mean time [ms] of one streaming copy of n doubles (16*n bytes of traffic)."""
        ms = C.c_double()
        _check(load_library().scale_hip_hbm_copy_probe(self._h, C.c_int64(int(n)), C.c_int32(int(reps)), C.byref(ms)))
        return ms.value

    HBM_PROBES = ("copy", "copy_nontemporal", "read", "write", "hipMemcpyDtoD")

    def hbm_probe(self, kind, n, reps=10):
        """Mean time [ms] of one launch of bandwidth probe `kind` (an index or a
        name of HBM_PROBES) over n doubles."""
        k = self.HBM_PROBES.index(kind) if isinstance(kind, str) else int(kind)
        ms = C.c_double()
        _check(load_library().scale_hip_hbm_probe(self._h, C.c_int32(k), C.c_int64(int(n)), C.c_int32(int(reps)),
                                                  C.byref(ms)))
        return ms.value

    def stream_ptr(self) -> int:
        s = C.c_void_p()
        _check(load_library().scale_hip_stream(self._h, C.byref(s)))
        return s.value or 0


def model_from_case(case, ctx: Context | None = None, fct=None, nd_coef=1.0e-4, device=0, mixed_precision=False,
                    divdmp_coef=0.0, nd_coef_q=0.0, short="hevi", dtl=None, tinteg_short="RK4",
                    tinteg_tracer="RK3WS2002", zsfc=None) -> Context:
    """Create a context holding the device-resident model of an inputs.Case
    (mixed_precision: FP32 dynamics with FP64 solve and reference state, the
    tracers advanced in FP64; short = "hevi" or "heve", ATMOS_DYN_TYPE; dtl:
    the large step, case.dt (the short step) by default, so nstep =
    ceil(dtl/dt)); zsfc: a surface height (IA,JA) whose terrain-following
    metrics the device computes, instead of case.metric's)."""
    if ctx is None:
        ctx = Context(case.grid, case.scheme, case.scheme_tracer, device)
    if zsfc is None:
        ctx.set_metrics(case.metric)
    else:  # terrain-following metrics generated on the device from the surface height
        ctx.set_terrain(case.metric, case.fz, zsfc)
    qa = case.qtrc.shape[-1]
    ctx.model_create(case.dt if dtl is None else dtl, case.dt, qa, nd_coef=nd_coef,
                     fct_tracer=case.fct if fct is None else fct, mixed_precision=mixed_precision,
                     divdmp_coef=divdmp_coef, nd_coef_q=nd_coef_q, short=short, tinteg_short=tinteg_short,
                     tinteg_tracer=tinteg_tracer)
    for k, v in case.state.items():
        ctx.model_set(k, v)
    for k, v in case.ref.items():
        ctx.model_set(k, v)
    ctx.model_set("CORIOLI", case.corioli)
    ctx.model_set("wdamp_coef", case.wdamp_coef)
    for iq in range(qa):
        ctx.model_set(QTRC_BASE + iq, np.asfortranarray(case.qtrc[..., iq]))
    return ctx

def batch_forward_checkpointed(members, nsteps, every):
    """The checkpointed forward of several member models at once
    (scale_hip_batch_forward_checkpointed): each member on its own stream,
    advanced step by step across the members; the members split the device
    memory for tapes."""
    n = len(members)
    arr = (C.c_void_p * n)(*[m._h.value for m in members])
    _check(load_library().scale_hip_batch_forward_checkpointed(arr, C.c_int32(n), C.c_int32(int(nsteps)),
                                                               C.c_int32(int(every))))


def batch_adjoint(members, xbs):
    """Adjoints of the last batch_forward_checkpointed run
    (scale_hip_batch_adjoint): xbs[m] holds member m's adjoints of DENS..RHOT
    after the run (whole arrays); returns their adjoints at its start, one
    dict per member."""
    n = len(members)
    outs = [{k: np.array(xb[k], dtype=np.float64, order="F", copy=True) for k in STATE} for xb in xbs]
    arr = (C.c_void_p * n)(*[m._h.value for m in members])
    ptrs = [(C.POINTER(C.c_double) * n)(*[_p(o[k]) for o in outs]) for k in STATE]
    _check(load_library().scale_hip_batch_adjoint(arr, C.c_int32(n), *ptrs))
    return outs


class Ensemble:
    """Member contexts (each with its model) stepped together
    (scale_hip_ensemble_*): the filter and the HEVI stages of all members run
    as one launch each; optional hipGraph replay and autotuned launch shapes.
    The member contexts stay owned by the caller."""

    def __init__(self, members):
        L = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = C.c_void_p()
        _check(L.scale_hip_ensemble_create(arr, C.c_int32(len(self.members)), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            load_library().scale_hip_ensemble_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def large_step(self, n=1):
        _check(load_library().scale_hip_ensemble_large_step(self._h, C.c_int32(int(n))))

    def synchronize(self):
        _check(load_library().scale_hip_ensemble_synchronize(self._h))

    def set_graph(self, on=True):
        _check(load_library().scale_hip_ensemble_set_graph(self._h, C.c_int32(1 if on else 0)))

    def set_shape(self, stage_bpc, nf_bpc, nf_tj):
        _check(load_library().scale_hip_ensemble_set_shape(self._h, C.c_int32(stage_bpc), C.c_int32(nf_bpc),
                                                           C.c_int32(nf_tj)))

    def get_shape(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        _check(load_library().scale_hip_ensemble_get_shape(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def tune(self, steps=2):
        """Autotune the launch shapes; returns (best ms per large step, report)."""
        best = C.c_double()
        buf = C.create_string_buffer(4096)
        _check(load_library().scale_hip_ensemble_tune(self._h, C.c_int32(int(steps)), C.byref(best), buf, 4096))
        return best.value, buf.value.decode()
