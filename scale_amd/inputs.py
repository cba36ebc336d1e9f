"""Synthetic workloads for the HEVI short step and FCT tracer advection.

Every 3-D field is a NumPy float64 array of shape (KA, IA, JA) in Fortran
order, i.e. the memory layout of the reference's REAL(RP) :: X(KA,IA,JA)
with k fastest and a width-2 periodic halo in x and y plus 2 ghost levels
in z.  Indices used here are 0-based: KS = IS = JS = 2.

Formulas follow the reference's initial-state code (input generation only,
not the hot path):

* vertical grid: FZ/CZ/CDZ/FDZ as in
  scalelib/src/atmosphere/grid/scale_atmos_grid_cartesC.F90:700-816;
* hydrostatic column: the Newton iteration of
  scalelib/src/atmosphere/common/scale_atmos_hydrostatic.F90:584-694
  (criteria sqrt(eps), dry air);
* reference-state ghost levels: scale_atmos_refstate.F90:831-844;
* cos^2 bubble: scale-rm/src/preprocess/mod_mkinit.F90:599;
* ghost-level fill of the prognostics: LARGE.F90:908-927.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

GRAV = 9.80665
RDRY = 287.04
CPDRY = 1004.64
CVDRY = CPDRY - RDRY
P00 = 1.0e5
EPS = float(np.finfo(np.float64).eps)
HALO = 2


@dataclasses.dataclass
class Grid:
    kmax: int
    imax: int
    jmax: int
    dx: float
    dy: float
    dz: np.ndarray  # (kmax,) layer thicknesses, bottom to top
    hx: int = HALO  # lateral halo width (IHALO = JHALO): 2, or 3 for the UD5/CD6 tracer schemes

    @property
    def KA(self):
        return self.kmax + 2 * HALO

    @property
    def IA(self):
        return self.imax + 2 * self.hx

    @property
    def JA(self):
        return self.jmax + 2 * self.hx

    @property
    def KS(self):
        return HALO

    @property
    def KE(self):
        return HALO + self.kmax - 1

    @property
    def IS(self):
        return self.hx

    @property
    def IE(self):
        return self.hx + self.imax - 1

    @property
    def JS(self):
        return self.hx

    @property
    def JE(self):
        return self.hx + self.jmax - 1

    @property
    def shape(self):
        return (self.KA, self.IA, self.JA)

    def zeros(self, *extra):
        return np.zeros(self.shape + tuple(extra), dtype=np.float64, order="F")


def halo_fill(g: Grid, a: np.ndarray) -> np.ndarray:
    """Periodic lateral halo (one-rank COMM_vars8) in place; all k levels."""
    IS, IE, JS, JE, h = g.IS, g.IE, g.JS, g.JE, g.hx
    a[:, :IS, JS:JE + 1] = a[:, IE - h + 1:IE + 1, JS:JE + 1]
    a[:, IE + 1:, JS:JE + 1] = a[:, IS:IS + h, JS:JE + 1]
    a[:, :, :JS] = a[:, :, JE - h + 1:JE + 1]
    a[:, :, JE + 1:] = a[:, :, JS:JS + h]
    return a


def halo_fill_adjoint(g: Grid, ab: np.ndarray) -> np.ndarray:
    """Transpose of halo_fill, in place: each halo value is added to the
    point it was copied from, the y rows first (they were filled last and
    their sources include the x halo), then the x strips; the halo ends zero."""
    IS, IE, JS, JE, h = g.IS, g.IE, g.JS, g.JE, g.hx
    ab[:, :, JS:JS + h] += ab[:, :, JE + 1:]
    ab[:, :, JE + 1:] = 0.0
    ab[:, :, JE - h + 1:JE + 1] += ab[:, :, :JS]
    ab[:, :, :JS] = 0.0
    ab[:, IS:IS + h, JS:JE + 1] += ab[:, IE + 1:, JS:JE + 1]
    ab[:, IE + 1:, JS:JE + 1] = 0.0
    ab[:, IE - h + 1:IE + 1, JS:JE + 1] += ab[:, :IS, JS:JE + 1]
    ab[:, :IS, JS:JE + 1] = 0.0
    return ab


def fill_ghost_and_halo(g: Grid, a: np.ndarray, var: str) -> np.ndarray:
    """Ghost levels and periodic halo of a prognostic as a large step leaves
    them (LARGE.F90:908-927, then COMM_vars8), in place: MOMZ is zero below KS
    and from KE up, the other fields repeat KS below and KE above."""
    if var == "MOMZ":
        a[:g.KS] = 0.0
        a[g.KE:] = 0.0
    else:
        a[:g.KS] = a[g.KS]
        a[g.KE + 1:] = a[g.KE]
    return halo_fill(g, a)


def interior_gradient(g: Grid, xb: dict) -> dict:
    """Gradient with respect to the interior prognostics (MOMZ on KS..KE-1)
    from adjoints of whole fields: the transpose of fill_ghost_and_halo."""
    grad = {}
    for k in ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT"):
        b = halo_fill_adjoint(g, np.array(xb[k], order="F", copy=True))
        if k != "MOMZ":
            b[g.KS] += b[:g.KS].sum(axis=0)
            b[g.KE] += b[g.KE + 1:].sum(axis=0)
        b[:g.KS] = 0.0
        b[g.KE + 1:] = 0.0
        if k == "MOMZ":
            b[g.KE] = 0.0
        grad[k] = b
    return grad


def vertical_grid(g: Grid):
    """FZ(0:KA), CZ, CDZ, FDZ, RCDZ, RFDZ (0-based: FZ[kF] for kF=0..KA)."""
    KA, KS, KE = g.KA, g.KS, g.KE
    fz = np.zeros(KA + 1)
    # Fortran FZ(KS-1) = 0 at the surface; Fortran index kF = kC + 1
    fz[KS] = 0.0
    for kk in range(g.kmax):
        fz[KS + 1 + kk] = fz[KS + kk] + g.dz[kk]
    for kF in range(KS - 1, -1, -1):
        fz[kF] = fz[kF + 1] - g.dz[0]
    for kF in range(KE + 2, KA + 1):
        fz[kF] = fz[kF - 1] + g.dz[-1]
    cz = 0.5 * (fz[1:] + fz[:-1])          # CZ(k) for C index k
    cdz = fz[1:] - fz[:-1]
    fdz = cz[1:] - cz[:-1]                 # (KA-1)
    return fz, cz, cdz, fdz, 1.0 / cdz, 1.0 / fdz


def hydrostatic_column(g: Grid, cz: np.ndarray, fz: np.ndarray, pott: np.ndarray, pres_sfc=1.0e5, pott_sfc=300.0):
    """Dry hydrostatic density/pressure column (buildrho_1D with iteration)."""
    KS, KE, KA = g.KS, g.KE, g.KA
    cpovcv = CPDRY / CVDRY
    criteria = math.sqrt(EPS)
    dens = np.zeros(KA)
    pres = np.zeros(KA)
    dens_sfc = P00 / RDRY / pott_sfc * (pres_sfc / P00) ** (CVDRY / CPDRY)
    dz0 = cz[KS] - fz[KS]  # CZ(KS) - FZ(KS-1)
    # lowest cell from the surface by the same Newton step
    d = dens_sfc
    for _ in range(100):
        p = P00 * (d * RDRY * pott[KS] / P00) ** cpovcv
        dhyd = (pres_sfc - p) / dz0 - GRAV * 0.5 * (dens_sfc + d)
        dgrd = -cpovcv * p / d / dz0 - GRAV * 0.5
        dn = d - dhyd / dgrd
        if abs(dn - d) <= criteria:
            d = dn
            break
        d = dn
    dens[KS] = d
    pres[KS] = P00 * (d * RDRY * pott[KS] / P00) ** cpovcv
    for k in range(KS + 1, KE + 1):
        dz = cz[k] - cz[k - 1]
        dens_s = 0.0
        dk = dens[k - 1]
        for _ in range(100):
            if abs(dk - dens_s) <= criteria:
                break
            dens_s = dk
            pk = P00 * (dens_s * RDRY * pott[k] / P00) ** cpovcv
            dhyd = (pres[k - 1] - pk) / dz - GRAV * 0.5 * (dens[k - 1] + dens_s)
            dgrd = -cpovcv * pk / dens_s / dz - GRAV * 0.5
            dk = dens_s - dhyd / dgrd
        dens[k] = dk
        pres[k] = P00 * (dk * RDRY * pott[k] / P00) ** cpovcv
    dens[:KS] = dens[KS]
    dens[KE + 1:] = dens[KE]
    pres[:KS] = pres[KS]
    pres[KE + 1:] = pres[KE]
    return dens, pres


@dataclasses.dataclass
class Case:
    """Everything one workload needs: grid, metrics, reference state, prognostics."""
    grid: Grid
    metric: dict
    ref: dict
    state: dict
    qtrc: np.ndarray  # (KA, IA, JA, QA) Fortran order
    zero_tend: dict
    corioli: np.ndarray  # (1, IA, JA)
    wdamp_coef: np.ndarray  # (KA,)
    dt: float
    scheme: int = 3
    scheme_tracer: int = 3
    fct: bool = True
    fz: np.ndarray = None


def make_case(kmax=24, imax=16, jmax=16, dx=500.0, dy=500.0, dz=250.0, *, seed=20261202,
              bv=0.01, pott_sfc=300.0, u_profile="uniform", u0=10.0, v0=0.0, shear_depth=10000.0,
              bubble=True, bubble_dtheta=3.0, bubble_r=2000.0, bubble_zc=2000.0,
              noise=0.01, ntracer=1, tracer_kind="zero", dt=1.0, scheme=3, scheme_tracer=3,
              fct=True, stretched=None, momz_noise=0.0, halo=HALO) -> Case:
    """Build one synthetic workload.

    ``stretched=(dz_bottom, dz_top)`` makes layer thickness grow linearly
    bottom to top; otherwise ``dz`` is uniform.
    """
    if stretched is not None:
        dzb, dzt = stretched
        dzs = np.array([dzb + (dzt - dzb) * kk / max(kmax - 1, 1) for kk in range(kmax)])
    else:
        dzs = np.full(kmax, float(dz))
    g = Grid(kmax, imax, jmax, float(dx), float(dy), dzs, hx=halo)
    KA, IA, JA = g.shape
    KS, KE, IS, IE, JS, JE = g.KS, g.KE, g.IS, g.IE, g.JS, g.JE
    fz, cz, cdz, fdz, rcdz, rfdz = vertical_grid(g)
    cdx = np.full(IA, g.dx); fdx = np.full(IA - 1, g.dx)
    cdy = np.full(JA, g.dy); fdy = np.full(JA - 1, g.dy)
    metric = dict(
        CDZ=cdz, FDZ=fdz, RCDZ=rcdz, RFDZ=rfdz,
        CDX=cdx, FDX=fdx, RCDX=1.0 / cdx, RFDX=1.0 / fdx,
        CDY=cdy, FDY=fdy, RCDY=1.0 / cdy, RFDY=1.0 / fdy,
        PHI=np.asfortranarray(np.broadcast_to((GRAV * cz)[:, None, None], g.shape).copy()),
        GSQRT=g.zeros(7) + 1.0, J13G=g.zeros(7), J23G=g.zeros(7),
        MAPF=np.ones((IA, JA, 2, 4), order="F"), J33G=1.0,
    )
    # base state: theta = pott_sfc * exp(N^2/g z) (mod_mkinit.F90:1827-1846)
    pott_b = pott_sfc * np.exp(bv * bv / GRAV * cz)
    dens_b, pres_b = hydrostatic_column(g, cz, fz, pott_b, pott_sfc=pott_sfc)
    phi = GRAV * cz
    ref_pres = pres_b.copy()
    ref_pres[KS - 1] = pres_b[KS + 1] - dens_b[KS] * (phi[KS - 1] - phi[KS + 1])
    ref_pres[KE + 1] = pres_b[KE - 1] - dens_b[KE] * (phi[KE + 1] - phi[KE - 1])
    ref_pott = pott_b.copy()
    ref_pott[:KS] = pott_b[KS]
    ref_pott[KE + 1:] = pott_b[KE]
    bc = lambda col: np.asfortranarray(np.broadcast_to(col[:, None, None], g.shape).copy())
    ref = dict(REF_dens=bc(dens_b), REF_pott=bc(ref_pott), REF_pres=bc(ref_pres))

    rng = np.random.Generator(np.random.PCG64(seed))
    x = (np.arange(IA) - IS + 0.5) * g.dx
    y = (np.arange(JA) - JS + 0.5) * g.dy
    X, Y, Z = np.meshgrid(x, y, cz, indexing="ij")  # (IA, JA, KA)
    X = np.transpose(X, (2, 0, 1)); Y = np.transpose(Y, (2, 0, 1)); Z = np.transpose(Z, (2, 0, 1))
    Lx, Ly = imax * g.dx, jmax * g.dy
    theta = bc(pott_b)
    if bubble:
        r2 = ((X - 0.5 * Lx) / bubble_r) ** 2 + ((Y - 0.5 * Ly) / bubble_r) ** 2 + ((Z - bubble_zc) / bubble_r) ** 2
        theta = theta + bubble_dtheta * np.cos(0.5 * np.pi * np.sqrt(np.minimum(r2, 1.0))) ** 2
    if noise > 0.0:
        nz = g.zeros()
        nz[KS:KE + 1, IS:IE + 1, JS:JE + 1] = rng.normal(0.0, noise, size=(kmax, imax, jmax)).astype(np.float64)
        theta = theta + nz
    dens = bc(dens_b)
    if u_profile == "uniform":
        ucol = np.full(KA, u0)
    else:  # linear shear 0..u0 over shear_depth
        ucol = u0 * np.minimum(np.maximum(cz, 0.0), shear_depth) / shear_depth
    momx = np.asfortranarray(dens * ucol[:, None, None])
    momy = np.asfortranarray(dens * v0)
    momz = g.zeros()
    if momz_noise > 0.0:
        momz[KS:KE, IS:IE + 1, JS:JE + 1] = rng.normal(0.0, momz_noise, size=(kmax - 1, imax, jmax))
    rhot = np.asfortranarray(dens * theta)
    state = dict(DENS=dens, MOMZ=momz, MOMX=momx, MOMY=momy, RHOT=rhot)
    for v in state.values():
        halo_fill(g, v)
        v[:KS] = v[KS]
        v[KE + 1:] = v[KE]
    momz[:KS] = 0.0
    momz[KE:] = 0.0
    qa = max(ntracer, 0)
    qtrc = np.zeros(g.shape + (max(qa, 1),), order="F")
    kinds = tracer_kind if isinstance(tracer_kind, (list, tuple)) else [tracer_kind] * qa
    for iq in range(qa):
        kind = kinds[iq]
        if kind == "zero":
            q = np.zeros(g.shape)
        elif kind == "blob":
            r2 = ((X - 0.3 * Lx) / (0.15 * Lx)) ** 2 + ((Y - 0.5 * Ly) / (0.15 * Ly)) ** 2 + ((Z - 0.2 * cz[KE]) / (0.15 * cz[KE] + 1.0)) ** 2
            q = 1.0e-3 * np.cos(0.5 * np.pi * np.sqrt(np.minimum(r2, 1.0))) ** 2
        elif kind == "tophat":
            q = np.where((np.abs(X - 0.5 * Lx) < 0.2 * Lx) & (np.abs(Y - 0.5 * Ly) < 0.2 * Ly) & (Z < 0.3 * cz[KE]), 1.0e-3, 0.0)
        elif kind == "uniform":
            q = np.full(g.shape, 1.0e-3)
        elif kind == "noise":
            q = 1.0e-3 * (1.0 + 0.1 * rng.standard_normal(g.shape))
        else:
            raise ValueError(kind)
        qtrc[..., iq] = q
        halo_fill(g, qtrc[..., iq])
    zero_tend = {k: g.zeros() for k in ("DENS_tp", "MOMZ_tp", "MOMX_tp", "MOMY_tp", "RHOT_tp")}
    zero_tend["RHOQ_tp"] = g.zeros(max(qa, 1))
    return Case(grid=g, metric=metric, ref=ref, state=state, qtrc=qtrc[..., :qa] if qa > 0 else qtrc[..., :0],
                zero_tend=zero_tend, corioli=np.zeros((1, IA, JA), order="F"), wdamp_coef=np.zeros(KA),
                dt=float(dt), scheme=scheme, scheme_tracer=scheme_tracer, fct=fct, fz=fz)


def bell_topography(g: Grid, height=500.0, half_width=5000.0, cx=None, cy=None, ry=None) -> np.ndarray:
    """Surface height Zsfc(IA,JA) of a bell-shaped mountain
    height / (1 + ((x-cx)/a)^2 + ((y-cy)/b)^2) (scale-rm/src/preprocess/
    mod_mktopo.F90:238-246, BELL_eachnode = .false.), centred in the domain by
    default, evaluated on the interior and given a periodic halo (the
    topography's COMM_vars8 fill, scale_topography.F90:107-131)."""
    xc = (np.arange(g.IA) - g.IS + 0.5) * g.dx
    yc = (np.arange(g.JA) - g.JS + 0.5) * g.dy
    cx = 0.5 * g.imax * g.dx if cx is None else cx
    cy = 0.5 * g.jmax * g.dy if cy is None else cy
    ry = half_width if ry is None else ry
    dist = ((xc[:, None] - cx) / half_width) ** 2 + ((yc[None, :] - cy) / ry) ** 2
    z = np.asfortranarray(height / (1.0 + dist))
    halo_fill(g, z[None])
    return z


def mountain_members(n=32, imax=128, jmax=128, kmax=64, ntracer=5, height=500.0, half_width=5000.0, seed=20261202 + 4000):
    """Members of a batched adjoint over a bell mountain: n members of
    imax x jmax x kmax (DX = DY = 500 m, DZ = 250 m, dt = 1 s), stable
    stratification N = 0.012 1/s, uniform u = 10 m/s, no bubble, theta noise
    N(0, 0.1 K) from each member's own seed, ntracer FCT tracers; each with the
    surface height of a bell mountain (height, half_width) centred in its
    domain.  Returns [(case, zsfc)]."""
    kinds = ["blob", "tophat", "uniform", "blob", "tophat"]
    out = []
    for m in range(n):
        case = make_case(kmax=kmax, imax=imax, jmax=jmax, dx=500.0, dy=500.0, dz=250.0, seed=seed + m, bv=0.012,
                         u0=10.0, bubble=False, noise=0.1, ntracer=ntracer,
                         tracer_kind=[kinds[q % len(kinds)] for q in range(ntracer)], fct=True, dt=1.0)
        out.append((case, bell_topography(case.grid, height=height, half_width=half_width)))
    return out


def large_step_prep(case: Case):
    """RT2P, DPRES0, REF_rhot of LARGE.F90:443-471 (dry form) over the whole array."""
    g = case.grid
    KS, KE = g.KS, g.KE
    rhot = case.state["RHOT"]
    cpovcv = CPDRY / CVDRY
    rt2p = g.zeros(); dpres0 = g.zeros(); ref_rhot = g.zeros()
    pres = P00 * (RDRY * rhot[KS:KE + 1] / P00) ** cpovcv
    rt2p[KS:KE + 1] = cpovcv * pres / rhot[KS:KE + 1]
    dpres0[KS:KE + 1] = pres - case.ref["REF_pres"][KS:KE + 1]
    ref_rhot[KS:KE + 1] = rhot[KS:KE + 1]
    rp = case.ref["REF_pres"]
    dpres0[KS - 1] = dpres0[KS + 1] + (rp[KS + 1] - rp[KS - 1])
    dpres0[KE + 1] = dpres0[KE - 1] + (rp[KE - 1] - rp[KE + 1])
    return dict(RT2P=rt2p, DPRES0=dpres0, REF_rhot=ref_rhot)


# named workloads (sizes and settings of the benchmark configurations)
def warm_bubble_40x40x80(**kw) -> Case:
    """Dry warm bubble, 40x40x80, DX=DY=500 m, DZ=250 m, u=10 m/s, UD3, dt=1 s, 1 tracer QV=0."""
    args = dict(kmax=80, imax=40, jmax=40, dx=500.0, dy=500.0, dz=250.0, seed=20261202,
                bv=0.01, u0=10.0, bubble=True, bubble_dtheta=3.0, bubble_r=2000.0, bubble_zc=2000.0,
                noise=0.01, ntracer=1, tracer_kind="zero", dt=1.0, scheme=3, scheme_tracer=3, fct=False)
    args.update(kw)
    return make_case(**args)


def shear_512x512x128(**kw) -> Case:
    """512x512x128, DX=1 km, DZ 50..500 m, linear shear 0..20 m/s over 10 km,
    theta noise 0.1 K, 3 tracers (cos^2 blob, top-hat, uniform) with FCT, dt=2 s."""
    args = dict(kmax=128, imax=512, jmax=512, dx=1000.0, dy=1000.0, stretched=(50.0, 500.0),
                seed=20261202 + 1000, bv=0.01, u_profile="shear", u0=20.0, shear_depth=10000.0,
                bubble=False, noise=0.1, ntracer=3, tracer_kind=["blob", "tophat", "uniform"], dt=2.0,
                scheme=3, scheme_tracer=3, fct=True)
    args.update(kw)
    return make_case(**args)


def ensemble_members(scale=1):
    """The 64-member ensemble of uneven domains: 24 members at 96x96x48, 24 at
    128x128x64 and 16 at 192x192x64 (DX = DY = 500 m, DZ = 250 m, dt = 1 s),
    each with theta noise N(0, 0.5 K) from its own seed and a uniform wind u
    drawn from U(5, 15) m/s, one top-hat tracer with FCT.  scale > 1 divides
    the horizontal sizes (for tests)."""
    rng = np.random.default_rng(20261202 + 3000)
    winds = rng.uniform(5.0, 15.0, size=64)
    sizes = [(96, 48)] * 24 + [(128, 64)] * 24 + [(192, 64)] * 16
    out = []
    for m, (n, k) in enumerate(sizes):
        out.append(make_case(kmax=k, imax=n // scale, jmax=n // scale, dx=500.0, dy=500.0, dz=250.0,
                             seed=20261202 + 3000 + m, noise=0.5, u0=float(winds[m]), bubble=False, ntracer=1,
                             tracer_kind=["tophat"], fct=True, dt=1.0))
    return out
