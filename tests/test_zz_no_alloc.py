"""scale_hip_tstep_tracer_heve_noalloc (device arrays, no allocation) against
scale_hip_tstep_tracer_heve (host arrays) through the C ABI: bit-identical
on two calls in a row, the second on a smaller grid, plus the size limit of
its work arrays.  Device memory for the inputs and outputs is the test's own
(hipMalloc through the HIP runtime the library itself is linked against),
as a caller that cannot allocate while it steps would hold it."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
from scale_amd import inputs

pytestmark = pytest.mark.gpu

NOALLOC_MAX_N3 = 1 << 22


@pytest.fixture(scope="module")
def hip():
    from scale_amd import hip as shp
    shp.load_library()
    return shp


def tracer_inputs(case, seed, terrain):
    g = case.grid
    rng = np.random.Generator(np.random.PCG64(seed))
    st = H.perturbed(case, seed + 1)
    mflx = np.asfortranarray(rng.standard_normal(g.shape + (3,)) * np.array([0.3, 8.0, 5.0]))
    for n in range(3):
        inputs.halo_fill(g, mflx[..., n])
    q = np.asfortranarray(case.qtrc[..., 1])
    q0 = np.asfortranarray(case.qtrc[..., 1] * (1.0 + 0.01 * rng.standard_normal(g.shape)))
    inputs.halo_fill(g, q0)
    nd = np.asfortranarray(1e-5 * rng.standard_normal(g.shape + (3,)))
    rq = np.asfortranarray(1e-6 * rng.standard_normal(g.shape))
    m = H.terrain_metrics(case, seed=seed + 2) if terrain else case.metric
    return dict(QTRC=q, QTRC0=q0, RHOQ_t=rq, DENS0=case.state["DENS"], DENS=st["DENS"], mflx_hi=mflx, num_diff=nd,
                GSQRT=m["GSQRT"], MAPF=np.asfortranarray(m["MAPF"][:, :, :, 0]), CDZ=m["CDZ"], RCDZ=m["RCDZ"],
                RCDX=m["RCDX"], RCDY=m["RCDY"])


def hip_runtime():
    """The HIP runtime libscalehip.so is linked against (already loaded)."""
    from scale_amd import hip as shp
    shp.load_library()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    lib_dir = "/opt/rocm/lib"
    path = next((p for p in paths if p.startswith(lib_dir)), None) or next(iter(paths))
    rt = C.CDLL(path)
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipFree.argtypes = [C.c_void_p]
    return rt


class DevArray:
    """A device copy of a host array in the library's device layout:
    Fortran (KA,IA,JA[,n]) -> i fastest, then j, then k, components as
    consecutive planes; raw (the (IA,JA,n) map factor) and 1-D arrays as
    they are, since their Fortran order is already i fastest."""

    def __init__(self, rt, a, raw=False):
        self.rt = rt
        if raw or a.ndim < 3:
            h = np.ascontiguousarray(a.ravel(order="F"))
        elif a.ndim == 3:
            h = np.ascontiguousarray(a.transpose(0, 2, 1))
        elif a.ndim == 4:
            h = np.ascontiguousarray(a.transpose(3, 0, 2, 1))
        self.shape = h.shape
        self.nbytes = h.nbytes
        self.ptr = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        assert rt.hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, 1) == 0  # host to device

    def dp(self):
        return C.cast(self.ptr, C.POINTER(C.c_double))

    def get3(self):
        h = np.empty(self.shape)
        assert self.rt.hipMemcpy(h.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return np.asfortranarray(h.transpose(0, 2, 1))

    def free(self):
        if self.ptr:
            self.rt.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def run_pair(hip, case, inp, fct, along, dtl):
    """(host drop-in result, device no-allocation result), both Fortran layout."""
    g = case.grid
    poison = np.full(g.shape, np.nan, order="F")
    ref = poison.copy(order="F")
    with hip.Context(g, case.scheme, case.scheme_tracer) as ctx:
        ctx.ATMOS_DYN_Tstep_tracer_fvm_heve(ref, inp["QTRC"], inp["QTRC0"], inp["RHOQ_t"], inp["DENS0"], inp["DENS"],
                                            inp["mflx_hi"], inp["num_diff"], inp["GSQRT"], inp["MAPF"], inp["CDZ"],
                                            inp["RCDZ"], inp["RCDX"], inp["RCDY"], dtl, fct, along)
    rt = hip_runtime()
    with hip.Context(g, case.scheme, case.scheme_tracer) as ctx:
        dv = {k: DevArray(rt, v, raw=(k == "MAPF")) for k, v in inp.items()}
        out = DevArray(rt, poison)
        try:
            L = hip.load_library()
            rc = L.scale_hip_tstep_tracer_heve_noalloc(
                ctx._h, out.dp(), dv["QTRC"].dp(), dv["QTRC0"].dp(), dv["RHOQ_t"].dp(), dv["DENS0"].dp(),
                dv["DENS"].dp(), dv["mflx_hi"].dp(), dv["num_diff"].dp(), dv["GSQRT"].dp(), dv["MAPF"].dp(),
                dv["CDZ"].dp(), dv["RCDZ"].dp(), dv["RCDX"].dp(), dv["RCDY"].dp(),
                C.byref(C.c_double(dtl)), C.byref(C.c_int32(int(fct))), C.byref(C.c_int32(int(along))))
            hip._check(rc)
            ctx.synchronize()
            got = out.get3()
        finally:
            for b in list(dv.values()) + [out]:
                b.free()
    return ref, got


def test_noalloc_tracer_matches_dropin_twice_second_smaller(hip):
    """Call 1: 24x20x26 with terrain metrics, UD3, along-stream FCT.  Call 2
    (right after, same device work arrays): a smaller 12x10x20 flat grid,
    7-point FCT.  Both bit-identical to the host-array entry point."""
    big = H.small_case(kmax=26, imax=24, jmax=20, seed=801)
    small = H.small_case(kmax=20, imax=12, jmax=10, seed=802)
    assert small.grid.KA * small.grid.IA * small.grid.JA < big.grid.KA * big.grid.IA * big.grid.JA
    for n, (case, terrain, fct, along) in enumerate([(big, True, True, True), (small, False, True, False)]):
        print(f"[noalloc] call {n + 1}: grid {case.grid.shape}, terrain={terrain}, fct={fct}, along={along}", flush=True)
        inp = tracer_inputs(case, 900 + n, terrain)
        ref, got = run_pair(hip, case, inp, fct, along, 2.0)
        assert np.isnan(ref).sum() < ref.size
        assert np.array_equal(ref, got, equal_nan=True), f"call {n + 1}: max |diff| {np.nanmax(np.abs(ref - got)):.3e}"


def test_noalloc_tracer_without_fct_and_other_scheme(hip):
    """No FCT (no work arrays) and the CD4 tracer scheme: bit-identical."""
    case = H.small_case(kmax=20, imax=16, jmax=12, seed=803, scheme_tracer=4)
    print(f"[noalloc] no FCT, CD4, grid {case.grid.shape}", flush=True)
    inp = tracer_inputs(case, 950, False)
    ref, got = run_pair(hip, case, inp, False, True, 1.5)
    assert np.array_equal(ref, got, equal_nan=True)


def test_noalloc_tracer_rejects_oversized_grid_and_aliasing(hip):
    """FCT on a grid with KA*IA*JA above the work arrays' limit, and an
    output aliasing an input, are rejected before anything is launched."""
    g = inputs.Grid(kmax=252, imax=128, jmax=128, dx=500.0, dy=500.0, dz=np.full(252, 100.0))
    assert g.KA * g.IA * g.JA > NOALLOC_MAX_N3
    print(f"[noalloc] rejection checks, grid {g.shape}", flush=True)
    L = hip.load_library()
    rt = hip_runtime()
    t_out, t_in = DevArray(rt, np.zeros(16)), DevArray(rt, np.zeros(16))
    try:
        o, p = t_out.dp(), t_in.dp()
        with hip.Context(g) as ctx:
            rc = L.scale_hip_tstep_tracer_heve_noalloc(ctx._h, o, p, p, p, p, p, p, p, p, p, p, p, p, p,
                                                       C.byref(C.c_double(1.0)), C.byref(C.c_int32(1)),
                                                       C.byref(C.c_int32(1)))
            assert rc != 0 and b"exceeds the maximum" in L.scale_hip_last_error()
        small = H.small_case(kmax=20, imax=12, jmax=10, seed=804)
        with hip.Context(small.grid) as ctx:
            rc = L.scale_hip_tstep_tracer_heve_noalloc(ctx._h, o, p, o, p, p, p, p, p, p, p, p, p, p, p,
                                                       C.byref(C.c_double(1.0)), C.byref(C.c_int32(1)),
                                                       C.byref(C.c_int32(1)))
            assert rc != 0 and b"alias" in L.scale_hip_last_error()
    finally:
        t_out.free()
        t_in.free()
