// capi.hip -- extern "C" entry points of libscalehip.so (declared in
// include/scale_hip.h): context setup, the drop-in Tstep_short / Tstep_tracer /
// numfilter calls on host Fortran arrays, and the device-resident model.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/scale_hip.h"
#include "common.h"
#include "model.h"
#include "shp_internal.h"

struct scale_hip_ctx {
  shp::Ctx c;
  // drop-in device buffers (allocated on first use)
  std::vector<double*> dbuf;
  bool dropin_ready = false;
};

namespace {

thread_local std::string g_err;

int fail(const std::string& msg, int code = -1) {
  g_err = msg;
  return code;
}
int hipfail(hipError_t e, const char* where) {
  return fail(std::string(where) + ": " + hipGetErrorString(e), -2);
}
#define CK(x)                                      \
  do {                                             \
    hipError_t e_ = (x);                           \
    if (e_ != hipSuccess) return hipfail(e_, #x);  \
  } while (0)

using shp::Ctx;
using shp::Dims;

// Work arrays of scale_hip_tstep_tracer_heve_noalloc: phi_lo, R+ and R- of
// the FCT limiter, 3 fields of at most NOALLOC_MAX_N3 doubles each (96 MiB).
// Each device holds its own copy from the first kernel launch until the
// process ends.  Host code uses the per-device address g_noalloc_ptr[dev]
// taken once by scale_hip_init, never the symbol itself.
constexpr size_t NOALLOC_MAX_N3 = size_t(1) << 22;
constexpr size_t NOALLOC_FIELDS = 3;
static __device__ __align__(256) double scratch[NOALLOC_FIELDS * NOALLOC_MAX_N3];
constexpr int NOALLOC_MAX_DEV = 64;
double* g_noalloc_ptr[NOALLOC_MAX_DEV] = {};
std::mutex g_noalloc_mu;

// device address of this device's scratch (the device must be current)
hipError_t noalloc_setup(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[dev]) return hipSuccess;
  double* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_noalloc_ptr[dev] = ptr;
  return hipSuccess;
}

// device allocation of n doubles, zeroed on the context's stream, ahead of
// any later work of the context (elements no kernel writes, such as the
// ghost levels of work fields, then hold zeros, not earlier contents)
double* dalloc(Ctx& c, size_t n, std::vector<void*>* track) {
  void* p = nullptr;
  if (hipMalloc(&p, n * sizeof(double) + 256) != hipSuccess) return nullptr;
  if (hipMemsetAsync(p, 0, n * sizeof(double) + 256, c.stream) != hipSuccess) { (void)hipFree(p); return nullptr; }
  if (track) track->push_back(p);
  return (double*)p;
}

int ensure_stage(Ctx& c, size_t elems) {
  if (c.stage_elems >= elems) return 0;
  if (c.stage) hipFree(c.stage);
  c.stage = nullptr;
  c.stage_elems = 0;
  if (hipMalloc((void**)&c.stage, elems * sizeof(double)) != hipSuccess) return fail("hipMalloc of the staging buffer failed");
  c.stage_elems = elems;
  return 0;
}

// host Fortran array (ncomp x KA*IA*JA) -> device layout
int upload3(Ctx& c, double* dst, const double* host, int ncomp) {
  if (!host) return fail("null array argument");
  const size_t n = c.d.n3 * (size_t)ncomp;
  if (int r = ensure_stage(c, n)) return r;
  CK(hipMemcpyAsync(c.stage, host, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(shp::to_device_layout(c, dst, c.stage, ncomp, c.stream));
  return 0;
}
int download3(Ctx& c, double* host, const double* src, int ncomp) {
  const size_t n = c.d.n3 * (size_t)ncomp;
  if (int r = ensure_stage(c, n)) return r;
  CK(shp::to_fortran_layout(c, c.stage, src, ncomp, c.stream));
  CK(hipMemcpyAsync(host, c.stage, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CK(hipStreamSynchronize(c.stream));
  return 0;
}
int upload_raw(Ctx& c, double* dst, const double* host, size_t n) {
  if (!host) return fail("null array argument");
  CK(hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  return 0;
}

bool all_equal(const double* a, size_t n, double v) {
  for (size_t q = 0; q < n; ++q)
    if (a[q] != v) return false;
  return true;
}

// 1-D grid arrays: CDZ FDZ RCDZ RFDZ (KA) | CDX FDX RCDX RFDX (IA) | CDY FDY RCDY RFDY (JA)
int set_grid1d(Ctx& c, const double* CDZ, const double* FDZ, const double* RCDZ, const double* RFDZ,
               const double* CDX, const double* FDX, const double* RCDX, const double* RFDX,
               const double* CDY, const double* FDY, const double* RCDY, const double* RFDY) {
  const Dims& d = c.d;
  std::vector<double> h(4 * (size_t)(d.KA + d.IA + d.JA), 0.0);
  double* p = h.data();
  auto put = [&](const double* s, int n, int len) {
    if (s) memcpy(p, s, sizeof(double) * (size_t)len);
    p += n;
  };
  put(CDZ, d.KA, d.KA); put(FDZ, d.KA, d.KA - 1); put(RCDZ, d.KA, d.KA); put(RFDZ, d.KA, d.KA - 1);
  put(CDX, d.IA, d.IA); put(FDX, d.IA, d.IA - 1); put(RCDX, d.IA, d.IA); put(RFDX, d.IA, d.IA - 1);
  put(CDY, d.JA, d.JA); put(FDY, d.JA, d.JA - 1); put(RCDY, d.JA, d.JA); put(RFDY, d.JA, d.JA - 1);
  CK(hipMemcpyAsync(c.g1buf, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(hipStreamSynchronize(c.stream));
  if (CDZ) c.hCDZ.assign(CDZ, CDZ + d.KA);
  if (FDZ) c.hFDZ.assign(FDZ, FDZ + d.KA - 1);
  if (CDX) c.hCDX.assign(CDX, CDX + d.IA);
  if (FDX) c.hFDX.assign(FDX, FDX + d.IA - 1);
  if (CDY) c.hCDY.assign(CDY, CDY + d.JA);
  if (FDY) c.hFDY.assign(FDY, FDY + d.JA - 1);
  return 0;
}

// metric arrays; decides the flat (TERRAIN=false) specialisation
int set_metric3d(Ctx& c, const double* PHI, const double* GSQRT, const double* J13G, const double* J23G, double J33G, const double* MAPF) {
  const Dims& d = c.d;
  if (PHI) { if (int r = upload3(c, c.PHI, PHI, 1)) return r; }
  const bool flat = GSQRT && J13G && J23G && MAPF && all_equal(GSQRT, 7 * d.n3, 1.0) && all_equal(J13G, 7 * d.n3, 0.0) &&
                    all_equal(J23G, 7 * d.n3, 0.0) && all_equal(MAPF, 8 * d.pl, 1.0);
  if (!flat) {
    if (!GSQRT || !J13G || !J23G || !MAPF) return fail("metric arrays are required");
    if (!c.GSQRT) {
      c.GSQRT = dalloc(c, 7 * d.n3, nullptr);
      c.J13G = dalloc(c, 7 * d.n3, nullptr);
      c.J23G = dalloc(c, 7 * d.n3, nullptr);
      if (!c.GSQRT || !c.J13G || !c.J23G) return fail("hipMalloc of metric arrays failed");
    }
    if (int r = upload3(c, c.GSQRT, GSQRT, 7)) return r;
    if (int r = upload3(c, c.J13G, J13G, 7)) return r;
    if (int r = upload3(c, c.J23G, J23G, 7)) return r;
  }
  if (MAPF) { if (int r = upload_raw(c, c.MAPF, MAPF, 8 * d.pl)) return r; }
  c.terrain = !flat;
  c.J33G = J33G;
  c.metrics_set = true;
  CK(hipStreamSynchronize(c.stream));
  return 0;
}

// numerical-filter tables (ATMOS_DYN_filter_setup, COMMON.F90:92-319) and
// nd_coef_* (COMMON.F90:494-511) for a given DT and ND_COEF
int set_numfilter_tables(Ctx& c, double DT, double ND_COEF, double sfc) {
  if (c.nd_coef == ND_COEF && c.nd_dt == DT && c.nd_sfc_fact == sfc && c.CNZ3) return 0;
  const Dims& d = c.d;
  const int KA = d.KA, IA = d.IA, JA = d.JA, KS = d.KS, KE = d.KE, IS = d.IS, IE = d.IE, JS = d.JS, JE = d.JE;
  if ((int)c.hCDZ.size() != KA || (int)c.hCDX.size() != IA || (int)c.hCDY.size() != JA || (int)c.hFDZ.size() != KA - 1)
    return fail("numfilter: CDZ/CDX/CDY/FDZ/FDX/FDY must be set");
  const double* CDZ = c.hCDZ.data(); const double* FDZ = c.hFDZ.data();
  const double* CDX = c.hCDX.data(); const double* FDX = c.hFDX.data();
  const double* CDY = c.hCDY.data(); const double* FDY = c.hFDY.data();
  std::vector<double> z(3 * KA * 2, 0.0), x(3 * IA * 2, 0.0), y(3 * JA * 2, 0.0);
#define CNZ3(a, k, b) z[(a)-1 + 3 * ((size_t)(k) + (size_t)KA * ((b)-1))]
#define CNX3(a, i, b) x[(a)-1 + 3 * ((size_t)(i) + (size_t)IA * ((b)-1))]
#define CNY3(a, j, b) y[(a)-1 + 3 * ((size_t)(j) + (size_t)JA * ((b)-1))]
  for (int k = KS + 1; k <= KE - 1; ++k) {
    CNZ3(1, k, 1) = 1.0 / (FDZ[k] * CDZ[k] * FDZ[k - 1]);
    CNZ3(2, k, 1) = 1.0 / (FDZ[k] * CDZ[k] * FDZ[k - 1]) + 1.0 / (FDZ[k - 1] * CDZ[k] * FDZ[k - 1]) + 1.0 / (FDZ[k - 1] * CDZ[k - 1] * FDZ[k - 1]);
  }
  for (int k = KS + 2; k <= KE; ++k)
    CNZ3(3, k, 1) = 1.0 / (FDZ[k - 1] * CDZ[k] * FDZ[k - 1]) + 1.0 / (FDZ[k - 1] * CDZ[k - 1] * FDZ[k - 1]) + 1.0 / (FDZ[k - 1] * CDZ[k - 1] * FDZ[k - 2]);
  CNZ3(1, KS - 1, 1) = 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS]);
  CNZ3(1, KS, 1) = 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]);
  CNZ3(2, KS, 1) = 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS]);
  CNZ3(3, KS, 1) = 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS + 1]);
  CNZ3(3, KS + 1, 1) = 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]);
  CNZ3(1, KE, 1) = 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]);
  CNZ3(2, KE, 1) = 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE - 1] * FDZ[KE - 1]);
  CNZ3(1, KE + 1, 1) = 1.0 / (FDZ[KE - 2] * CDZ[KE - 1] * FDZ[KE - 1]);
  CNZ3(2, KE + 1, 1) = 1.0 / (FDZ[KE - 2] * CDZ[KE - 1] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE - 1] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]);
  CNZ3(3, KE + 1, 1) = 1.0 / (FDZ[KE - 1] * CDZ[KE + 1] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]);
  for (int k = KS + 1; k <= KE - 1; ++k) {
    CNZ3(1, k, 2) = 1.0 / (CDZ[k + 1] * FDZ[k] * CDZ[k]);
    CNZ3(2, k, 2) = 1.0 / (CDZ[k + 1] * FDZ[k] * CDZ[k]) + 1.0 / (CDZ[k] * FDZ[k] * CDZ[k]) + 1.0 / (CDZ[k] * FDZ[k - 1] * CDZ[k]);
    CNZ3(3, k, 2) = 1.0 / (CDZ[k] * FDZ[k] * CDZ[k]) + 1.0 / (CDZ[k] * FDZ[k - 1] * CDZ[k]) + 1.0 / (CDZ[k] * FDZ[k - 1] * CDZ[k - 1]);
  }
  CNZ3(1, KS - 1, 2) = 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]);
  CNZ3(1, KS, 2) = 1.0 / (CDZ[KS + 1] * FDZ[KS] * CDZ[KS]);
  CNZ3(2, KS, 2) = 1.0 / (CDZ[KS + 1] * FDZ[KS] * CDZ[KS]) + 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]) + 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]);
  CNZ3(3, KS, 2) = 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]) + 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]) + 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS + 1]);
  CNZ3(1, KE, 2) = 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE]);
  CNZ3(2, KE, 2) = 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE]) + 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE]) + 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE]);
  CNZ3(3, KE, 2) = 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE]) + 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE]) + 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE - 1]);
  CNZ3(1, KE + 1, 2) = 1.0 / (CDZ[KE - 2] * FDZ[KE - 2] * CDZ[KE - 1]);
  CNZ3(2, KE + 1, 2) = 1.0 / (CDZ[KE - 2] * FDZ[KE - 2] * CDZ[KE - 1]) + 1.0 / (CDZ[KE - 1] * FDZ[KE - 2] * CDZ[KE - 1]) + 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE - 1]);
  CNZ3(3, KE + 1, 2) = 1.0 / (CDZ[KE - 1] * FDZ[KE - 2] * CDZ[KE - 1]) + 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE - 1]) + 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE]);
  CNX3(1, IS - 1, 1) = 1.0 / (FDX[IS - 1] * CDX[IS - 1] * FDX[IS - 2]);
  for (int i = IS; i <= IE + 1; ++i) {
    CNX3(1, i, 1) = 1.0 / (FDX[i] * CDX[i] * FDX[i - 1]);
    CNX3(2, i, 1) = 1.0 / (FDX[i] * CDX[i] * FDX[i - 1]) + 1.0 / (FDX[i - 1] * CDX[i] * FDX[i - 1]) + 1.0 / (FDX[i - 1] * CDX[i - 1] * FDX[i - 1]);
    CNX3(3, i, 1) = 1.0 / (FDX[i - 1] * CDX[i] * FDX[i - 1]) + 1.0 / (FDX[i - 1] * CDX[i - 1] * FDX[i - 1]) + 1.0 / (FDX[i - 1] * CDX[i - 1] * FDX[i - 2]);
  }
  for (int i = IS - 1; i <= IE + 1; ++i) {
    CNX3(1, i, 2) = 1.0 / (CDX[i + 1] * FDX[i] * CDX[i]);
    CNX3(2, i, 2) = 1.0 / (CDX[i + 1] * FDX[i] * CDX[i]) + 1.0 / (CDX[i] * FDX[i] * CDX[i]) + 1.0 / (CDX[i] * FDX[i - 1] * CDX[i]);
    CNX3(3, i, 2) = 1.0 / (CDX[i] * FDX[i] * CDX[i]) + 1.0 / (CDX[i] * FDX[i - 1] * CDX[i]) + 1.0 / (CDX[i] * FDX[i - 1] * CDX[i - 1]);
  }
  CNY3(1, JS - 1, 1) = 1.0 / (FDY[JS - 1] * CDY[JS - 1] * FDY[JS - 2]);
  for (int j = JS; j <= JE + 1; ++j) {
    CNY3(1, j, 1) = 1.0 / (FDY[j] * CDY[j] * FDY[j - 1]);
    CNY3(2, j, 1) = 1.0 / (FDY[j] * CDY[j] * FDY[j - 1]) + 1.0 / (FDY[j - 1] * CDY[j] * FDY[j - 1]) + 1.0 / (FDY[j - 1] * CDY[j - 1] * FDY[j - 1]);
    CNY3(3, j, 1) = 1.0 / (FDY[j - 1] * CDY[j] * FDY[j - 1]) + 1.0 / (FDY[j - 1] * CDY[j - 1] * FDY[j - 1]) + 1.0 / (FDY[j - 1] * CDY[j - 1] * FDY[j - 2]);
  }
  for (int j = JS - 1; j <= JE + 1; ++j) {
    CNY3(1, j, 2) = 1.0 / (CDY[j + 1] * FDY[j] * CDY[j]);
    CNY3(2, j, 2) = 1.0 / (CDY[j + 1] * FDY[j] * CDY[j]) + 1.0 / (CDY[j] * FDY[j] * CDY[j]) + 1.0 / (CDY[j] * FDY[j - 1] * CDY[j]);
    CNY3(3, j, 2) = 1.0 / (CDY[j] * FDY[j] * CDY[j]) + 1.0 / (CDY[j] * FDY[j - 1] * CDY[j]) + 1.0 / (CDY[j] * FDY[j - 1] * CDY[j - 1]);
  }
#undef CNZ3
#undef CNX3
#undef CNY3
  // nd_coef_* with x**4 evaluated as (x*x)*(x*x), as a real**integer power is
  const double DIFF4 = ND_COEF / ((double)(1 << 4) * DT);
  auto p4 = [](double v) { double v2 = v * v; return v2 * v2; };
  std::vector<double> ndc(2 * (size_t)KA + 2 * (size_t)IA + 2 * (size_t)JA, 0.0);
  double* nd_cdz = ndc.data(); double* nd_fdz = nd_cdz + KA; double* nd_cdx = nd_fdz + KA;
  double* nd_fdx = nd_cdx + IA; double* nd_cdy = nd_fdx + IA; double* nd_fdy = nd_cdy + JA;
  for (int k = KS - 1; k <= KE; ++k) nd_cdz[k] = DIFF4 * p4(CDZ[k]);
  for (int k = KS + 1; k <= KE - 1; ++k) nd_fdz[k] = DIFF4 * p4(FDZ[k]);
  for (int i = IS; i <= IE; ++i) { nd_cdx[i] = DIFF4 * p4(CDX[i]); nd_fdx[i] = DIFF4 * p4(FDX[i]); }
  for (int j = JS; j <= JE; ++j) { nd_cdy[j] = DIFF4 * p4(CDY[j]); nd_fdy[j] = DIFF4 * p4(FDY[j]); }
  if (!c.CNZ3) {
    c.CNZ3 = dalloc(c, z.size(), nullptr); c.CNX3 = dalloc(c, x.size(), nullptr);
    c.CNY3 = dalloc(c, y.size(), nullptr); c.ndc = dalloc(c, ndc.size(), nullptr);
    if (!c.CNZ3 || !c.CNX3 || !c.CNY3 || !c.ndc) return fail("hipMalloc of filter tables failed");
  }
  CK(hipMemcpyAsync(c.CNZ3, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(hipMemcpyAsync(c.CNX3, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(hipMemcpyAsync(c.CNY3, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(hipMemcpyAsync(c.ndc, ndc.data(), ndc.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(hipStreamSynchronize(c.stream));
  c.nd_coef = ND_COEF; c.nd_dt = DT; c.nd_sfc_fact = sfc;
  return 0;
}

// nd_coef_cdz/cdx/cdy of the tracer filter (COMMON.F90:1032-1043) for ND_COEF_Q and dt
int set_ndq_table(Ctx& c, double dt, double ND_COEF, double sfc) {
  if (c.ndqc && c.ndq_coef == ND_COEF && c.ndq_dt == dt && c.ndq_sfc == sfc) return 0;
  const int KA = c.d.KA, IA = c.d.IA, JA = c.d.JA, KS = c.d.KS, KE = c.d.KE, IS = c.d.IS, IE = c.d.IE, JS = c.d.JS, JE = c.d.JE;
  if (c.hCDZ.empty() || c.hCDX.empty() || c.hCDY.empty()) return fail("the grid (CDZ, CDX, CDY) has not been given");
  const double DIFF4 = ND_COEF / ((double)(1 << 4) * dt);
  auto p4 = [](double v) { double v2 = v * v; return v2 * v2; };
  std::vector<double> t((size_t)KA + IA + JA, 0.0);
  for (int k = KS - 1; k <= KE; ++k) t[k] = DIFF4 * p4(c.hCDZ[k]);
  for (int i = IS; i <= IE; ++i) t[KA + i] = DIFF4 * p4(c.hCDX[i]);
  for (int j = JS; j <= JE; ++j) t[KA + IA + j] = DIFF4 * p4(c.hCDY[j]);
  if (!c.ndqc && !(c.ndqc = dalloc(c, t.size(), nullptr))) return fail("hipMalloc of the tracer filter table failed");
  CK(hipMemcpyAsync(c.ndqc, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(hipStreamSynchronize(c.stream));
  c.ndq_coef = ND_COEF; c.ndq_dt = dt; c.ndq_sfc = sfc;
  return 0;
}

// drop-in buffers: 0..4 s0, 5..9 s, 10..14 t, 15 DPRES0, 16 RT2P, 17 REF_dens, 18 REF_rhot, 19 DDIV,
// 20 num_diff(15), 21..25 outputs, 26 mflx(3), 27 tflx(3), 28 CORIOLI(IA*JA), 29 wdamp(KA)
enum { B_S0 = 0, B_S = 5, B_T = 10, B_DPRES0 = 15, B_RT2P, B_REFD, B_REFR, B_DDIV, B_ND, B_OUT, B_MF = 26, B_TF, B_COR, B_WD, B_N };

int ensure_dropin(scale_hip_ctx* h) {
  if (h->dropin_ready) return 0;
  Ctx& c = h->c;
  const size_t n3 = c.d.n3;
  h->dbuf.assign(B_N, nullptr);
  for (int q = 0; q < B_N; ++q) {
    size_t n = n3;
    if (q == B_ND) n = 15 * n3;
    else if (q == B_MF || q == B_TF) n = 3 * n3;
    else if (q == B_COR) n = c.d.pl;
    else if (q == B_WD) n = (size_t)c.d.KA;
    else if (q > B_OUT + 4 && q < B_MF) continue;
    h->dbuf[q] = dalloc(c, n, nullptr);
    if (!h->dbuf[q]) return fail("hipMalloc of drop-in buffers failed");
  }
  h->dropin_ready = true;
  return 0;
}

int check_ctx(scale_hip_ctx* h) {
  if (!h) return fail("null context");
  CK(hipSetDevice(h->c.device));
  return 0;
}

}  // namespace

extern "C" {

const char* scale_hip_last_error(void) { return g_err.c_str(); }
const char* scale_hip_version(void) { return "scale-hip 0.1 (gfx950)"; }

int scale_hip_init(const scale_hip_grid* g, scale_hip_ctx** out) {
  if (!g || !out) return fail("null argument");
  *out = nullptr;
  // KHALO = 2; IHALO = JHALO = 2, or 3 (needed by the UD5/CD6 tracer schemes)
  if (g->KHALO != 2 || g->IHALO != g->JHALO || (g->IHALO != 2 && g->IHALO != 3))
    return fail("KHALO must be 2 and IHALO = JHALO must be 2 or 3");
  const int H = g->IHALO;
  const int KS = g->KS - 1, KE = g->KE - 1, IS = g->IS - 1, IE = g->IE - 1, JS = g->JS - 1, JE = g->JE - 1;
  if (KS != 2 || IS != H || JS != H || KE + 3 != g->KA || IE + H + 1 != g->IA || JE + H + 1 != g->JA)
    return fail("grid bounds must be KS=3, IS=JS=IHALO+1 and KA=KE+2, IA=IE+IHALO, JA=JE+IHALO");
  if (KE - KS + 1 < 4) return fail("KMAX must be >= 4");
  if (IE - IS + 1 < H || JE - JS + 1 < H) return fail("IMAX and JMAX must be >= IHALO");
  if (g->flux_scheme != SCALE_HIP_UD3 && g->flux_scheme != SCALE_HIP_CD4)
    return fail("the dynamics flux scheme must be UD3 (3) or CD4 (4)");
  if (g->flux_scheme_tracer != SCALE_HIP_UD3 && g->flux_scheme_tracer != SCALE_HIP_CD4 &&
      g->flux_scheme_tracer != SCALE_HIP_UD3KOREN1993 && g->flux_scheme_tracer != SCALE_HIP_UD1 &&
      g->flux_scheme_tracer != SCALE_HIP_CD2 && g->flux_scheme_tracer != SCALE_HIP_UD5 &&
      g->flux_scheme_tracer != SCALE_HIP_CD6)
    return fail("the tracer flux scheme must be UD3 (3), CD4 (4), UD3KOREN1993 (31), UD1 (1), CD2 (2), UD5 (5) or CD6 (6)");
  if ((g->flux_scheme_tracer == SCALE_HIP_UD5 || g->flux_scheme_tracer == SCALE_HIP_CD6) && (H != 3 || KE - KS + 1 < 5))
    return fail("the UD5 and CD6 tracer schemes need IHALO = JHALO = 3 and KMAX >= 5");
  if (hipSetDevice(g->device) != hipSuccess) return fail("hipSetDevice failed");
  if (hipError_t e = noalloc_setup(g->device); e != hipSuccess) return hipfail(e, "hipGetSymbolAddress(scratch)");
  scale_hip_ctx* h = new scale_hip_ctx();
  Ctx& c = h->c;
  c.device = g->device;
  c.d.KA = g->KA; c.d.IA = g->IA; c.d.JA = g->JA;
  c.d.KS = KS; c.d.KE = KE; c.d.IS = IS; c.d.IE = IE; c.d.JS = JS; c.d.JE = JE;
  c.d.pl = (size_t)g->IA * g->JA;
  c.d.n3 = c.d.pl * g->KA;
  c.scheme = g->flux_scheme;
  c.scheme_tracer = g->flux_scheme_tracer;
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) { delete h; return fail("hipStreamCreate failed"); }
  if (hipDeviceGetAttribute(&c.ncu, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess || c.ncu <= 0) c.ncu = 256;
  const size_t n1 = 4 * (size_t)(c.d.KA + c.d.IA + c.d.JA);
  c.g1buf = dalloc(c, n1, nullptr);
  c.MAPF = dalloc(c, 8 * c.d.pl, nullptr);
  c.PHI = dalloc(c, c.d.n3, nullptr);
  bool ok = c.g1buf && c.MAPF && c.PHI;
  for (int q = 0; q < shp::NSCR && ok; ++q) ok = (c.scr[q] = dalloc(c, c.d.n3, nullptr)) != nullptr;
  if (!ok) { scale_hip_finalize(h); return fail("hipMalloc failed in scale_hip_init"); }
  double* p = c.g1buf;
  const int KA = c.d.KA, IA = c.d.IA, JA = c.d.JA;
  c.g1.CDZ = p; c.g1.FDZ = p + KA; c.g1.RCDZ = p + 2 * KA; c.g1.RFDZ = p + 3 * KA; p += 4 * KA;
  c.g1.CDX = p; c.g1.FDX = p + IA; c.g1.RCDX = p + 2 * IA; c.g1.RFDX = p + 3 * IA; p += 4 * IA;
  c.g1.CDY = p; c.g1.FDY = p + JA; c.g1.RCDY = p + 2 * JA; c.g1.RFDY = p + 3 * JA;
  // flat map factor until metrics are given
  std::vector<double> ones(8 * c.d.pl, 1.0);
  if (hipMemcpyAsync(c.MAPF, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice, c.stream) != hipSuccess ||
      hipStreamSynchronize(c.stream) != hipSuccess) {
    scale_hip_finalize(h);
    return fail("copying the initial map factors failed in scale_hip_init");
  }
  *out = h;
  return 0;
}

int scale_hip_finalize(scale_hip_ctx* h) {
  if (!h) return 0;
  Ctx& c = h->c;
  hipSetDevice(c.device);
  if (c.stream) hipStreamSynchronize(c.stream);
  auto F = [](void* p) { if (p) hipFree(p); };
  F(c.g1buf); F(c.GSQRT); F(c.J13G); F(c.J23G); F(c.MAPF); F(c.PHI);
  F(c.CNZ3); F(c.CNX3); F(c.CNY3); F(c.ndc); F(c.ndqc); F(c.stage);
  for (int q = 0; q < shp::NSCR; ++q) F(c.scr[q]);
  F(c.mfct);
  for (double* b : h->dbuf) F(b);
  shp::adj_free(c);
  if (c.model) {
    for (void* q : c.model->allocs) F(q);
    delete c.model;
  }
  shp::prof_destroy(c.prof);
  if (c.stream) hipStreamDestroy(c.stream);
  delete h;
  return 0;
}

int scale_hip_synchronize(scale_hip_ctx* h) {
  if (int r = check_ctx(h)) return r;
  CK(hipStreamSynchronize(h->c.stream));
  return 0;
}

int scale_hip_stream(scale_hip_ctx* h, void** stream) {
  if (int r = check_ctx(h)) return r;
  *stream = (void*)h->c.stream;
  return 0;
}

static int short_stage(bool heve, scale_hip_ctx* h,
    double* DENS_RK, double* MOMZ_RK, double* MOMX_RK, double* MOMY_RK, double* RHOT_RK,
    double* PROG_RK, double* mflx_hi, double* tflx_hi,
    const double* DENS0, const double* MOMZ0, const double* MOMX0, const double* MOMY0, const double* RHOT0,
    const double* DENS, const double* MOMZ, const double* MOMX, const double* MOMY, const double* RHOT,
    const double* DENS_t, const double* MOMZ_t, const double* MOMX_t, const double* MOMY_t, const double* RHOT_t,
    const double* PROG0, const double* PROG,
    const double* DPRES0, const double* RT2P, const double* CORIOLI,
    const double* num_diff, const double* wdamp_coef, const double* divdmp_coef, const double* DDIV,
    const int32_t* FLAG_FCT_MOMENTUM, const int32_t* FLAG_FCT_T, const int32_t* FLAG_FCT_ALONG_STREAM,
    const double* CDZ, const double* FDZ, const double* FDX, const double* FDY,
    const double* RCDZ, const double* RCDX, const double* RCDY, const double* RFDZ, const double* RFDX, const double* RFDY,
    const double* PHI, const double* GSQRT, const double* J13G, const double* J23G, const double* J33G, const double* MAPF,
    const double* REF_dens, const double* REF_rhot,
    const int32_t* BND_W, const int32_t* BND_E, const int32_t* BND_S, const int32_t* BND_N,
    const double* dtrk, const int32_t* last) {
  (void)PROG_RK; (void)PROG0; (void)PROG; (void)last;
  if (int r = check_ctx(h)) return r;
  if (!dtrk || !divdmp_coef || !J33G) return fail("null scalar argument");
  // HEVI ignores the FCT flags (HEVI.F90:224-226)
  const bool fct_mom = heve && FLAG_FCT_MOMENTUM && *FLAG_FCT_MOMENTUM;
  if ((BND_W && *BND_W) || (BND_E && *BND_E) || (BND_S && *BND_S) || (BND_N && *BND_N))
    return fail("only periodic lateral boundaries (BND_* = .false.) are supported");
  Ctx& c = h->c;
  if (int r = ensure_dropin(h)) return r;
  if (fct_mom && !c.mfct && !(c.mfct = dalloc(c, shp::MFCT_FIELDS * c.d.n3, nullptr)))
    return fail("hipMalloc failed for the momentum FCT work arrays");
  if (int r = set_grid1d(c, CDZ, FDZ, RCDZ, RFDZ, nullptr, FDX, RCDX, RFDX, nullptr, FDY, RCDY, RFDY)) return r;
  if (int r = set_metric3d(c, PHI, GSQRT, J13G, J23G, *J33G, MAPF)) return r;
  auto& B = h->dbuf;
  const double* s0[5] = {DENS0, MOMZ0, MOMX0, MOMY0, RHOT0};
  const double* s[5] = {DENS, MOMZ, MOMX, MOMY, RHOT};
  const double* t[5] = {DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t};
  double* o[5] = {DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK};
  for (int v = 0; v < 5; ++v) {
    if (int r = upload3(c, B[B_S0 + v], s0[v], 1)) return r;
    if (int r = upload3(c, B[B_S + v], s[v], 1)) return r;
    if (int r = upload3(c, B[B_T + v], t[v], 1)) return r;
    if (int r = upload3(c, B[B_OUT + v], o[v], 1)) return r;  // elements HEVI leaves alone keep their values
  }
  if (int r = upload3(c, B[B_DPRES0], DPRES0, 1)) return r;
  if (int r = upload3(c, B[B_RT2P], RT2P, 1)) return r;
  if (int r = upload3(c, B[B_REFD], REF_dens, 1)) return r;
  if (int r = upload3(c, B[B_REFR], REF_rhot, 1)) return r;
  if (int r = upload3(c, B[B_DDIV], DDIV, 1)) return r;
  if (int r = upload3(c, B[B_ND], num_diff, 15)) return r;
  if (int r = upload3(c, B[B_MF], mflx_hi, 3)) return r;
  if (int r = upload3(c, B[B_TF], tflx_hi, 3)) return r;
  if (int r = upload_raw(c, B[B_COR], CORIOLI, c.d.pl)) return r;
  if (int r = upload_raw(c, B[B_WD], wdamp_coef, (size_t)c.d.KA)) return r;
  const bool has_cor = !all_equal(CORIOLI, c.d.pl, 0.0);
  shp::StageIO io;
  io.ctx = &c;
  for (int v = 0; v < 5; ++v) { io.o[v] = B[B_OUT + v]; io.s0[v] = B[B_S0 + v]; io.s[v] = B[B_S + v]; io.t[v] = B[B_T + v]; }
  io.mflx_hi = B[B_MF]; io.tflx_hi = B[B_TF];
  io.DPRES0 = B[B_DPRES0]; io.RT2P = B[B_RT2P]; io.CORIOLI = B[B_COR]; io.num_diff = B[B_ND]; io.wdamp_coef = B[B_WD];
  io.DDIV = B[B_DDIV]; io.REF_dens = B[B_REFD]; io.REF_rhot = B[B_REFR];
  io.divdmp_coef = *divdmp_coef; io.dtrk = *dtrk; io.has_coriolis = has_cor;
  io.fct_momentum = fct_mom;
  io.fct_t = heve && FLAG_FCT_T && *FLAG_FCT_T;
  io.fct_along = !FLAG_FCT_ALONG_STREAM || *FLAG_FCT_ALONG_STREAM;
  CK(heve ? shp::heve_stage(io, c.stream) : shp::hevi_stage(io, c.stream));
  for (int v = 0; v < 5; ++v)
    if (int r = download3(c, o[v], B[B_OUT + v], 1)) return r;
  if (int r = download3(c, mflx_hi, B[B_MF], 3)) return r;
  if (int r = download3(c, tflx_hi, B[B_TF], 3)) return r;
  return 0;
}

int scale_hip_tstep_short_hevi(scale_hip_ctx* h,
    double* DENS_RK, double* MOMZ_RK, double* MOMX_RK, double* MOMY_RK, double* RHOT_RK,
    double* PROG_RK, double* mflx_hi, double* tflx_hi,
    const double* DENS0, const double* MOMZ0, const double* MOMX0, const double* MOMY0, const double* RHOT0,
    const double* DENS, const double* MOMZ, const double* MOMX, const double* MOMY, const double* RHOT,
    const double* DENS_t, const double* MOMZ_t, const double* MOMX_t, const double* MOMY_t, const double* RHOT_t,
    const double* PROG0, const double* PROG,
    const double* DPRES0, const double* RT2P, const double* CORIOLI,
    const double* num_diff, const double* wdamp_coef, const double* divdmp_coef, const double* DDIV,
    const int32_t* FLAG_FCT_MOMENTUM, const int32_t* FLAG_FCT_T, const int32_t* FLAG_FCT_ALONG_STREAM,
    const double* CDZ, const double* FDZ, const double* FDX, const double* FDY,
    const double* RCDZ, const double* RCDX, const double* RCDY, const double* RFDZ, const double* RFDX, const double* RFDY,
    const double* PHI, const double* GSQRT, const double* J13G, const double* J23G, const double* J33G, const double* MAPF,
    const double* REF_dens, const double* REF_rhot,
    const int32_t* BND_W, const int32_t* BND_E, const int32_t* BND_S, const int32_t* BND_N,
    const double* dtrk, const int32_t* last) {
  return short_stage(false, h, DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, PROG_RK, mflx_hi, tflx_hi, DENS0, MOMZ0,
      MOMX0, MOMY0, RHOT0, DENS, MOMZ, MOMX, MOMY, RHOT, DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, PROG0, PROG, DPRES0,
      RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV, FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM,
      CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY, PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot,
      BND_W, BND_E, BND_S, BND_N, dtrk, last);
}

int scale_hip_tstep_short_heve(scale_hip_ctx* h,
    double* DENS_RK, double* MOMZ_RK, double* MOMX_RK, double* MOMY_RK, double* RHOT_RK,
    double* PROG_RK, double* mflx_hi, double* tflx_hi,
    const double* DENS0, const double* MOMZ0, const double* MOMX0, const double* MOMY0, const double* RHOT0,
    const double* DENS, const double* MOMZ, const double* MOMX, const double* MOMY, const double* RHOT,
    const double* DENS_t, const double* MOMZ_t, const double* MOMX_t, const double* MOMY_t, const double* RHOT_t,
    const double* PROG0, const double* PROG,
    const double* DPRES0, const double* RT2P, const double* CORIOLI,
    const double* num_diff, const double* wdamp_coef, const double* divdmp_coef, const double* DDIV,
    const int32_t* FLAG_FCT_MOMENTUM, const int32_t* FLAG_FCT_T, const int32_t* FLAG_FCT_ALONG_STREAM,
    const double* CDZ, const double* FDZ, const double* FDX, const double* FDY,
    const double* RCDZ, const double* RCDX, const double* RCDY, const double* RFDZ, const double* RFDX, const double* RFDY,
    const double* PHI, const double* GSQRT, const double* J13G, const double* J23G, const double* J33G, const double* MAPF,
    const double* REF_dens, const double* REF_rhot,
    const int32_t* BND_W, const int32_t* BND_E, const int32_t* BND_S, const int32_t* BND_N,
    const double* dtrk, const int32_t* last) {
  return short_stage(true, h, DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, PROG_RK, mflx_hi, tflx_hi, DENS0, MOMZ0,
      MOMX0, MOMY0, RHOT0, DENS, MOMZ, MOMX, MOMY, RHOT, DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, PROG0, PROG, DPRES0,
      RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV, FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM,
      CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY, PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot,
      BND_W, BND_E, BND_S, BND_N, dtrk, last);
}

int scale_hip_tstep_short_hevi_adj(scale_hip_ctx* h,
    const double* DENS_RK_b, const double* MOMZ_RK_b, const double* MOMX_RK_b, const double* MOMY_RK_b, const double* RHOT_RK_b,
    double* DENS0_b, double* MOMZ0_b, double* MOMX0_b, double* MOMY0_b, double* RHOT0_b,
    double* DENS_b, double* MOMZ_b, double* MOMX_b, double* MOMY_b, double* RHOT_b,
    double* num_diff_b, double* DPRES0_b, double* RT2P_b, double* REF_rhot_b,
    const double* DENS0, const double* MOMZ0, const double* MOMX0, const double* MOMY0, const double* RHOT0,
    const double* DENS, const double* MOMZ, const double* MOMX, const double* MOMY, const double* RHOT,
    const double* DENS_t, const double* MOMZ_t, const double* MOMX_t, const double* MOMY_t, const double* RHOT_t,
    const double* DPRES0, const double* RT2P, const double* CORIOLI,
    const double* num_diff, const double* wdamp_coef,
    const double* CDZ, const double* FDZ, const double* FDX, const double* FDY,
    const double* RCDZ, const double* RCDX, const double* RCDY, const double* RFDZ, const double* RFDX, const double* RFDY,
    const double* PHI, const double* GSQRT, const double* J13G, const double* J23G, const double* J33G, const double* MAPF,
    const double* REF_dens, const double* REF_rhot, const double* dtrk) {
  if (int r = check_ctx(h)) return r;
  if (!dtrk || !J33G) return fail("null scalar argument");
  Ctx& c = h->c;
  if (c.scheme != shp::UD3 && c.scheme != shp::CD4) return fail("the adjoint supports the UD3 and CD4 dynamics schemes");
  if (c.d.IS != 2) return fail("the adjoint supports IHALO = 2 only");
  if (int r = ensure_dropin(h)) return r;
  if (int r = set_grid1d(c, CDZ, FDZ, RCDZ, RFDZ, nullptr, FDX, RCDX, RFDX, nullptr, FDY, RCDY, RFDY)) return r;
  if (int r = set_metric3d(c, PHI, GSQRT, J13G, J23G, *J33G, MAPF)) return r;
  CK(shp::adj_setup(c, 0, 0));
  auto& B = h->dbuf;
  const double* s0[5] = {DENS0, MOMZ0, MOMX0, MOMY0, RHOT0};
  const double* s[5] = {DENS, MOMZ, MOMX, MOMY, RHOT};
  const double* t[5] = {DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t};
  const double* ob[5] = {DENS_RK_b, MOMZ_RK_b, MOMX_RK_b, MOMY_RK_b, RHOT_RK_b};
  double* s0b[5] = {DENS0_b, MOMZ0_b, MOMX0_b, MOMY0_b, RHOT0_b};
  double* sb[5] = {DENS_b, MOMZ_b, MOMX_b, MOMY_b, RHOT_b};
  for (int v = 0; v < 5; ++v) {
    if (int r = upload3(c, B[B_S0 + v], s0[v], 1)) return r;
    if (int r = upload3(c, B[B_S + v], s[v], 1)) return r;
    if (int r = upload3(c, B[B_T + v], t[v], 1)) return r;
    if (int r = upload3(c, B[B_OUT + v], ob[v], 1)) return r;
  }
  if (int r = upload3(c, B[B_DPRES0], DPRES0, 1)) return r;
  if (int r = upload3(c, B[B_RT2P], RT2P, 1)) return r;
  if (int r = upload3(c, B[B_REFD], REF_dens, 1)) return r;
  if (int r = upload3(c, B[B_REFR], REF_rhot, 1)) return r;
  if (int r = upload3(c, B[B_ND], num_diff, 15)) return r;
  if (int r = upload_raw(c, B[B_COR], CORIOLI, c.d.pl)) return r;
  if (int r = upload_raw(c, B[B_WD], wdamp_coef, (size_t)c.d.KA)) return r;
  double *ds0b[5], *dsb[5], *dNDb, *dPb[3];
  shp::adj_stage_bufs(c, ds0b, dsb, &dNDb, dPb);
  const size_t bytes = c.d.n3 * sizeof(double);
  for (int v = 0; v < 5; ++v) {
    CK(hipMemsetAsync(ds0b[v], 0, bytes, c.stream));
    CK(hipMemsetAsync(dsb[v], 0, bytes, c.stream));
  }
  CK(hipMemsetAsync(dNDb, 0, 15 * bytes, c.stream));
  for (int q = 0; q < 3; ++q) CK(hipMemsetAsync(dPb[q], 0, bytes, c.stream));
  shp::StageAdjIO io{};
  io.ctx = &c;
  for (int v = 0; v < 5; ++v) {
    io.s0[v] = B[B_S0 + v]; io.s[v] = B[B_S + v]; io.t[v] = B[B_T + v]; io.ob[v] = B[B_OUT + v];
    io.s0b[v] = ds0b[v]; io.sb[v] = dsb[v];
  }
  io.DPRES0 = B[B_DPRES0]; io.RT2P = B[B_RT2P]; io.CORIOLI = B[B_COR]; io.num_diff = B[B_ND]; io.wdamp_coef = B[B_WD];
  io.REF_dens = B[B_REFD]; io.REF_rhot = B[B_REFR];
  io.dtrk = *dtrk; io.has_coriolis = !all_equal(CORIOLI, c.d.pl, 0.0);
  io.NDb = dNDb; io.DPRES0b = dPb[0]; io.RT2Pb = dPb[1]; io.REFrtb = dPb[2];
  CK(shp::hevi_stage_adj(c, io, c.stream));
  for (int v = 0; v < 5; ++v) {
    if (int r = download3(c, s0b[v], ds0b[v], 1)) return r;
    if (int r = download3(c, sb[v], dsb[v], 1)) return r;
  }
  if (int r = download3(c, num_diff_b, dNDb, 15)) return r;
  if (int r = download3(c, DPRES0_b, dPb[0], 1)) return r;
  if (int r = download3(c, RT2P_b, dPb[1], 1)) return r;
  return download3(c, REF_rhot_b, dPb[2], 1);
}

int scale_hip_tstep_tracer_heve(scale_hip_ctx* h,
    double* QTRCo, const double* QTRC, const double* QTRC0, const double* RHOQ_t,
    const double* DENS0, const double* DENS, const double* mflx_hi, const double* num_diff,
    const double* GSQRT, const double* MAPF, const double* CDZ,
    const double* RCDZ, const double* RCDX, const double* RCDY,
    const double* dtl, const int32_t* FLAG_FCT_TRACER, const int32_t* FLAG_FCT_ALONG_STREAM) {
  if (int r = check_ctx(h)) return r;
  if (!dtl || !FLAG_FCT_TRACER || !FLAG_FCT_ALONG_STREAM) return fail("null scalar argument");
  Ctx& c = h->c;
  const Dims& d = c.d;
  if (int r = ensure_dropin(h)) return r;
  if (int r = set_grid1d(c, CDZ, nullptr, RCDZ, nullptr, nullptr, nullptr, RCDX, nullptr, nullptr, nullptr, RCDY, nullptr)) return r;
  // metrics: GSQRT(KA,IA,JA,7) and the (IA,JA,2) map factor
  const bool flat = all_equal(GSQRT, 7 * d.n3, 1.0) && all_equal(MAPF, 2 * d.pl, 1.0);
  if (!flat) {
    std::vector<double> m4(8 * d.pl, 1.0);
    memcpy(m4.data(), MAPF, 2 * d.pl * sizeof(double));
    std::vector<double> zero(7 * d.n3, 0.0);
    if (int r = set_metric3d(c, nullptr, GSQRT, zero.data(), zero.data(), 1.0, m4.data())) return r;
  } else {
    c.terrain = false;
  }
  auto& B = h->dbuf;
  // reuse drop-in buffers: QTRCo, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx(3), num_diff(3)
  double* dQo = B[B_OUT]; double* dQ = B[B_S]; double* dQ0 = B[B_S + 1]; double* dRQ = B[B_S + 2];
  double* dD0 = B[B_S + 3]; double* dD = B[B_S + 4]; double* dMF = B[B_MF]; double* dND = B[B_ND];
  if (int r = upload3(c, dQo, QTRCo, 1)) return r;
  if (int r = upload3(c, dQ, QTRC, 1)) return r;
  if (int r = upload3(c, dQ0, QTRC0, 1)) return r;
  if (int r = upload3(c, dRQ, RHOQ_t, 1)) return r;
  if (int r = upload3(c, dD0, DENS0, 1)) return r;
  if (int r = upload3(c, dD, DENS, 1)) return r;
  if (int r = upload3(c, dMF, mflx_hi, 3)) return r;
  if (int r = upload3(c, dND, num_diff, 3)) return r;
  shp::TracerIO io;
  io.QTRCo = dQo; io.QTRC = dQ; io.QTRC0 = dQ0; io.RHOQ_t = dRQ; io.DENS0 = dD0; io.DENS = dD;
  io.mflx_hi = dMF; io.num_diff = dND; io.dtl = *dtl; io.flag_fct = *FLAG_FCT_TRACER; io.flag_along = *FLAG_FCT_ALONG_STREAM;
  CK(shp::tracer_stage(c, io, c.stream));
  return download3(c, QTRCo, dQo, 1);
}

int scale_hip_tstep_tracer_heve_adj(scale_hip_ctx* h,
    const double* QTRCo_b, double* QTRC_b, double* QTRC0_b, double* RHOQ_t_b, double* num_diff_b,
    const double* QTRC, const double* QTRC0,
    const double* DENS0, const double* DENS, const double* mflx_hi, const double* num_diff,
    const double* GSQRT, const double* MAPF, const double* CDZ,
    const double* RCDZ, const double* RCDX, const double* RCDY,
    const double* dtl, const int32_t* FLAG_FCT_TRACER, const int32_t* FLAG_FCT_ALONG_STREAM) {
  if (int r = check_ctx(h)) return r;
  if (!dtl || !FLAG_FCT_TRACER || !FLAG_FCT_ALONG_STREAM) return fail("null scalar argument");
  if (!QTRCo_b || !QTRC_b || !QTRC0_b || !RHOQ_t_b || !num_diff_b || !QTRC || !QTRC0 || !DENS0 || !DENS || !mflx_hi ||
      !num_diff || !GSQRT || !MAPF || !RCDZ || !RCDX || !RCDY)
    return fail("null array argument");
  Ctx& c = h->c;
  const Dims& d = c.d;
  if (c.scheme_tracer == shp::UD3KOREN)
    return fail("the tracer adjoint needs a face form linear in the tracer; UD3KOREN1993's limiter is not");
  if (int r = ensure_dropin(h)) return r;
  if (int r = set_grid1d(c, CDZ, nullptr, RCDZ, nullptr, nullptr, nullptr, RCDX, nullptr, nullptr, nullptr, RCDY, nullptr)) return r;
  const bool flat = all_equal(GSQRT, 7 * d.n3, 1.0) && all_equal(MAPF, 2 * d.pl, 1.0);
  if (!flat) {
    std::vector<double> m4(8 * d.pl, 1.0);
    memcpy(m4.data(), MAPF, 2 * d.pl * sizeof(double));
    std::vector<double> zero(7 * d.n3, 0.0);
    if (int r = set_metric3d(c, nullptr, GSQRT, zero.data(), zero.data(), 1.0, m4.data())) return r;
  } else {
    c.terrain = false;
  }
  auto& B = h->dbuf;
  // inputs: QTRCo_b, QTRC, QTRC0, DENS0, DENS, mflx(3), num_diff(3); outputs in
  // the tendency buffers and num_diff's components 3..5
  double* dOb = B[B_OUT]; double* dQ = B[B_S]; double* dQ0 = B[B_S + 1];
  double* dD0 = B[B_S + 3]; double* dD = B[B_S + 4]; double* dMF = B[B_MF]; double* dND = B[B_ND];
  double* dQb = B[B_T]; double* dQ0b = B[B_T + 1]; double* dRQb = B[B_T + 2]; double* dNDb = B[B_ND] + 3 * d.n3;
  if (int r = upload3(c, dOb, QTRCo_b, 1)) return r;
  if (int r = upload3(c, dQ, QTRC, 1)) return r;
  if (int r = upload3(c, dQ0, QTRC0, 1)) return r;
  if (int r = upload3(c, dD0, DENS0, 1)) return r;
  if (int r = upload3(c, dD, DENS, 1)) return r;
  if (int r = upload3(c, dMF, mflx_hi, 3)) return r;
  if (int r = upload3(c, dND, num_diff, 3)) return r;
  shp::TracerIO io{};
  io.QTRC = dQ; io.QTRC0 = dQ0; io.DENS0 = dD0; io.DENS = dD; io.mflx_hi = dMF; io.num_diff = dND;
  io.dtl = *dtl; io.flag_fct = *FLAG_FCT_TRACER; io.flag_along = *FLAG_FCT_ALONG_STREAM;
  shp::TracerAdjIO adj{dOb, dQb, dQ0b, dRQb, dNDb};
  CK(shp::tracer_stage_adj(c, io, adj, c.stream));
  if (int r = download3(c, QTRC_b, dQb, 1)) return r;
  if (int r = download3(c, QTRC0_b, dQ0b, 1)) return r;
  if (int r = download3(c, RHOQ_t_b, dRQb, 1)) return r;
  return download3(c, num_diff_b, dNDb, 3);
}

int scale_hip_tstep_tracer_heve_noalloc(scale_hip_ctx* h,
    double* QTRCo, const double* QTRC, const double* QTRC0, const double* RHOQ_t,
    const double* DENS0, const double* DENS, const double* mflx_hi, const double* num_diff,
    const double* GSQRT, const double* MAPF, const double* CDZ,
    const double* RCDZ, const double* RCDX, const double* RCDY,
    const double* dtl, const int32_t* FLAG_FCT_TRACER, const int32_t* FLAG_FCT_ALONG_STREAM) {
  (void)CDZ;  // unused by the stage, as in the reference (tracer_fvm_heve.F90:72-242)
  if (int r = check_ctx(h)) return r;
  if (!dtl || !FLAG_FCT_TRACER || !FLAG_FCT_ALONG_STREAM) return fail("null scalar argument");
  if (!QTRCo || !QTRC || !QTRC0 || !RHOQ_t || !DENS0 || !DENS || !mflx_hi || !num_diff || !GSQRT || !MAPF || !RCDZ ||
      !RCDX || !RCDY)
    return fail("null array argument");
  if (QTRCo == QTRC || QTRCo == QTRC0 || QTRCo == DENS0 || QTRCo == DENS)
    return fail("QTRCo must not alias an input of the no-allocation tracer stage");
  Ctx& c = h->c;
  const bool fct = *FLAG_FCT_TRACER != 0;
  if (fct && c.d.n3 > NOALLOC_MAX_N3)
    return fail("scale_hip_tstep_tracer_heve_noalloc: KA*IA*JA = " + std::to_string(c.d.n3) + " exceeds the maximum " +
                std::to_string(NOALLOC_MAX_N3) + " of its FCT work arrays");
  double* ptr = (c.device >= 0 && c.device < NOALLOC_MAX_DEV) ? g_noalloc_ptr[c.device] : nullptr;
  if (!ptr) return fail("scale_hip_tstep_tracer_heve_noalloc: no work array address for this device");
  CK(hipSetDevice(c.device));
  shp::TracerIO io;
  io.QTRCo = QTRCo; io.QTRC = QTRC; io.QTRC0 = QTRC0; io.RHOQ_t = RHOQ_t; io.DENS0 = DENS0; io.DENS = DENS;
  io.mflx_hi = mflx_hi; io.num_diff = num_diff; io.dtl = *dtl; io.flag_fct = fct; io.flag_along = *FLAG_FCT_ALONG_STREAM;
  // the named timers create events and grow host tables: off for this call
  shp::Prof* prof = c.prof;
  c.prof = nullptr;
  hipError_t e = shp::tracer_stage_ext(c, io, GSQRT, MAPF, RCDZ, RCDX, RCDY, ptr, c.stream);
  c.prof = prof;
  CK(e);
  return 0;
}

int scale_hip_numfilter_coef(scale_hip_ctx* h, double* num_diff,
    const double* DENS, const double* MOMZ, const double* MOMX, const double* MOMY, const double* RHOT,
    const double* CDZ, const double* CDX, const double* CDY, const double* FDZ, const double* FDX, const double* FDY,
    const double* DT, const double* REF_dens, const double* REF_pott,
    const double* ND_COEF, const int32_t* ND_ORDER, const double* ND_SFC_FACT, const int32_t* ND_USE_RS) {
  if (int r = check_ctx(h)) return r;
  if (!DT || !ND_COEF || !ND_ORDER || !ND_SFC_FACT || !ND_USE_RS) return fail("null scalar argument");
  if (*ND_ORDER != 1) return fail("ND_ORDER other than 1 is not supported");
  if (!*ND_USE_RS) return fail("ND_USE_RS = .false. is not supported");
  Ctx& c = h->c;
  if (int r = ensure_dropin(h)) return r;
  if (int r = set_grid1d(c, CDZ, FDZ, nullptr, nullptr, CDX, FDX, nullptr, nullptr, CDY, FDY, nullptr, nullptr)) return r;
  if (int r = set_numfilter_tables(c, *DT, *ND_COEF, *ND_SFC_FACT)) return r;
  auto& B = h->dbuf;
  const double* s[5] = {DENS, MOMZ, MOMX, MOMY, RHOT};
  for (int v = 0; v < 5; ++v)
    if (int r = upload3(c, B[B_S + v], s[v], 1)) return r;
  if (int r = upload3(c, B[B_REFD], REF_dens, 1)) return r;
  if (int r = upload3(c, B[B_REFR], REF_pott, 1)) return r;
  const double* cs[5] = {B[B_S], B[B_S + 1], B[B_S + 2], B[B_S + 3], B[B_S + 4]};
  CK(shp::numfilter(c, B[B_ND], cs, B[B_REFD], B[B_REFR], c.stream));
  return download3(c, num_diff, B[B_ND], 15);
}

int scale_hip_numfilter_coef_q(scale_hip_ctx* h, double* num_diff_q, const double* DENS, const double* QTRC,
    const int32_t* is_qv, const double* CDZ, const double* CDX, const double* CDY, const double* dt,
    const double* REF_qv, const int32_t* iq, const double* ND_COEF, const int32_t* ND_ORDER,
    const double* ND_SFC_FACT, const int32_t* ND_USE_RS) {
  (void)iq;
  if (int r = check_ctx(h)) return r;
  if (!is_qv || !dt || !ND_COEF || !ND_ORDER || !ND_SFC_FACT || !ND_USE_RS) return fail("null scalar argument");
  if (*ND_ORDER != 1) return fail("ND_ORDER other than 1 is not supported");
  if (*is_qv && !*ND_USE_RS) return fail("QV without ND_USE_RS (the 19-point smoothing) is not supported");
  Ctx& c = h->c;
  if (!c.CNZ3) return fail("the filter tables come from the grid: call scale_hip_numfilter_coef or create the model first");
  if (int r = ensure_dropin(h)) return r;
  if (int r = set_grid1d(c, CDZ, nullptr, nullptr, nullptr, CDX, nullptr, nullptr, nullptr, CDY, nullptr, nullptr, nullptr)) return r;
  if (int r = set_ndq_table(c, *dt, *ND_COEF, *ND_SFC_FACT)) return r;
  auto& B = h->dbuf;
  if (int r = upload3(c, B[B_S], DENS, 1)) return r;
  if (int r = upload3(c, B[B_S + 1], QTRC, 1)) return r;
  if (*is_qv) {
    if (int r = upload3(c, B[B_S + 2], REF_qv, 1)) return r;
  }
  CK(shp::numfilter_q(c, B[B_MF], B[B_S], B[B_S + 1], *is_qv ? B[B_S + 2] : nullptr, c.stream));
  return download3(c, num_diff_q, B[B_MF], 3);
}

// ---- device-resident model ---------------------------------------------------------
int scale_hip_set_metrics(scale_hip_ctx* h,
    const double* CDZ, const double* FDZ, const double* RCDZ, const double* RFDZ,
    const double* CDX, const double* FDX, const double* RCDX, const double* RFDX,
    const double* CDY, const double* FDY, const double* RCDY, const double* RFDY,
    const double* PHI, const double* GSQRT, const double* J13G, const double* J23G,
    const double* J33G, const double* MAPF) {
  if (int r = check_ctx(h)) return r;
  if (!J33G) return fail("null J33G");
  if (!CDZ || !FDZ || !RCDZ || !RFDZ || !CDX || !FDX || !RCDX || !RFDX || !CDY || !FDY || !RCDY || !RFDY)
    return fail("every 1-D grid array is required");
  if (int r = set_grid1d(h->c, CDZ, FDZ, RCDZ, RFDZ, CDX, FDX, RCDX, RFDX, CDY, FDY, RCDY, RFDY)) return r;
  return set_metric3d(h->c, PHI, GSQRT, J13G, J23G, *J33G, MAPF);
}

int scale_hip_set_terrain(scale_hip_ctx* h,
    const double* CDZ, const double* FDZ, const double* RCDZ, const double* RFDZ,
    const double* CDX, const double* FDX, const double* RCDX, const double* RFDX,
    const double* CDY, const double* FDY, const double* RCDY, const double* RFDY,
    const double* CZ, const double* FZ, const double* Zsfc, const double* MAPF) {
  if (int r = check_ctx(h)) return r;
  if (!CDZ || !FDZ || !RCDZ || !RFDZ || !CDX || !FDX || !RCDX || !RFDX || !CDY || !FDY || !RCDY || !RFDY || !CZ || !FZ ||
      !Zsfc || !MAPF)
    return fail("set_terrain: every array is required");
  Ctx& c = h->c;
  const Dims& d = c.d;
  if (int r = set_grid1d(c, CDZ, FDZ, RCDZ, RFDZ, CDX, FDX, RCDX, RFDX, CDY, FDY, RCDY, RFDY)) return r;
  if (!c.GSQRT) {
    c.GSQRT = dalloc(c, 7 * d.n3, nullptr);
    c.J13G = dalloc(c, 7 * d.n3, nullptr);
    c.J23G = dalloc(c, 7 * d.n3, nullptr);
    if (!c.GSQRT || !c.J13G || !c.J23G) return fail("hipMalloc of metric arrays failed");
  }
  // the surface and the reference heights go through the staging buffer
  const size_t nz = (size_t)d.pl + (size_t)d.KA + (size_t)d.KA + 1;
  if (int r = ensure_stage(c, nz)) return r;
  double* zs = c.stage;
  double* cz = zs + d.pl;
  double* fz = cz + d.KA;
  CK(hipMemcpyAsync(zs, Zsfc, d.pl * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(hipMemcpyAsync(cz, CZ, (size_t)d.KA * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CK(hipMemcpyAsync(fz, FZ, ((size_t)d.KA + 1) * sizeof(double), hipMemcpyHostToDevice, c.stream));
  const double Htop = FZ[d.KE + 1] - FZ[d.KS];  // FZ(KE) - FZ(KS-1) (REAL:410)
  CK(shp::terrain_metrics(c, zs, cz, fz, Htop, c.GSQRT, c.J13G, c.J23G, c.PHI, c.stream));
  if (int r = upload_raw(c, c.MAPF, MAPF, 8 * d.pl)) return r;
  CK(hipStreamSynchronize(c.stream));
  c.terrain = !(all_equal(Zsfc, d.pl, 0.0) && all_equal(MAPF, 8 * d.pl, 1.0));
  c.J33G = 1.0;
  c.metrics_set = true;
  return 0;
}

int scale_hip_get_metrics(scale_hip_ctx* h, double* GSQRT, double* J13G, double* J23G, double* PHI) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  if (!c.metrics_set) return fail("no metrics set");
  if (!c.GSQRT) return fail("the metrics are flat (GSQRT = 1, J13G = J23G = 0): nothing stored");
  if (GSQRT) { if (int r = download3(c, GSQRT, c.GSQRT, 7)) return r; }
  if (J13G) { if (int r = download3(c, J13G, c.J13G, 7)) return r; }
  if (J23G) { if (int r = download3(c, J23G, c.J23G, 7)) return r; }
  if (PHI) { if (int r = download3(c, PHI, c.PHI, 1)) return r; }
  return 0;
}

int scale_hip_model_create(scale_hip_ctx* h, const scale_hip_dyn_config* cfg, int32_t QA) {
  if (int r = check_ctx(h)) return r;
  if (!cfg) return fail("null config");
  Ctx& c = h->c;
  if (!c.metrics_set) return fail("scale_hip_set_metrics must be called first");
  if (c.model) return fail("model already created");
  if (cfg->nd_order != 1) return fail("nd_order other than 1 is not supported");
  if (cfg->divdmp_coef < 0.0) return fail("divdmp_coef must be >= 0");
  if (cfg->divdmp_coef > 0.0 && cfg->mixed_precision) return fail("divergence damping is not supported in mixed precision");
  if (!(cfg->dtl > 0.0) || !(cfg->dts > 0.0)) return fail("dtl and dts must be positive");
  if (QA < 0) return fail("QA must be >= 0");
  if (cfg->mixed_precision && c.d.IS != 2) return fail("mixed precision supports IHALO = 2 only");
  if (cfg->mixed_precision && c.terrain) return fail("mixed precision needs flat terrain (GSQRT = 1, J13G = J23G = 0, MAPF = 1)");
  if (cfg->mixed_precision && c.d.KA > shp::KA_DEV_MAX) return fail("mixed precision supports at most 512 levels (KA)");
  if (cfg->nd_coef_q < 0.0) return fail("nd_coef_q must be >= 0");
  if (cfg->short_type != 0 && cfg->short_type != 1) return fail("short_type must be 0 (FVM-HEVI) or 1 (FVM-HEVE)");
  if (cfg->short_type == 1 && cfg->mixed_precision) return fail("the HEVE short step runs in FP64 only");
  const int ts = cfg->tinteg_short == 0 ? SCALE_HIP_RK4 : cfg->tinteg_short;
  if (ts != SCALE_HIP_RK4 && ts != SCALE_HIP_RK3 && ts != SCALE_HIP_RK3WS2002)
    return fail("tinteg_short must be SCALE_HIP_RK4 (4), SCALE_HIP_RK3 (3) or SCALE_HIP_RK3WS2002 (32)");
  if (cfg->tinteg_tracer != SCALE_HIP_TRACER_RK3WS2002 && cfg->tinteg_tracer != SCALE_HIP_TRACER_EULER)
    return fail("tinteg_tracer must be SCALE_HIP_TRACER_RK3WS2002 (0) or SCALE_HIP_TRACER_EULER (1)");
  shp::Model* M = new shp::Model();
  c.model = M;
  const size_t n3 = c.d.n3;
  auto A = [&](size_t n) { return dalloc(c, n, &M->allocs); };
  M->QA = QA;
  M->DENS = A(n3); M->MOMZ = A(n3); M->MOMX = A(n3); M->MOMY = A(n3); M->RHOT = A(n3);
  M->QTRC = QA > 0 ? A(n3 * (size_t)QA) : nullptr;
  M->REF_dens = A(n3); M->REF_pott = A(n3); M->REF_pres = A(n3);
  M->CORIOLI = A(c.d.pl); M->wdamp_coef = A((size_t)c.d.KA);
  M->mflx_hi = A(3 * n3); M->tflx_hi = A(3 * n3);
  M->RT2P = A(n3); M->DPRES0 = A(n3); M->REF_rhot = A(n3); M->DENS00 = A(n3); M->mflx_av = A(3 * n3);
  M->num_diff = A(15 * n3); M->DDIV = A(n3);
  for (int v = 0; v < 5; ++v) { M->t[v] = A(n3); M->S0[v] = A(n3); }
  for (int r = 0; r < 4; ++r) {
    for (int v = 0; v < 5; ++v) M->R[r][v] = A(n3);
    M->mfR[r] = A(3 * n3); M->tfR[r] = A(3 * n3);
  }
  M->ndq = A(3 * n3); M->DENS_RK = A(n3); M->QRK1 = A(n3); M->QRK2 = A(n3);
  if (QA > 0 && cfg->tinteg_tracer == SCALE_HIP_TRACER_RK3WS2002) M->DENS_RK2 = A(n3);
  if (cfg->mixed_precision) {
    M->mixed = 1;
    auto F = [&](size_t n) { return (float*)dalloc(c, (n + 1) / 2, &M->allocs); };
    for (int v = 0; v < 5; ++v) { M->fS[v] = F(n3); M->ft[v] = F(n3); }
    for (int r = 0; r < 4; ++r) {
      for (int v = 0; v < 5; ++v) M->fR[r][v] = F(n3);
      M->fmfR[r] = F(3 * n3); M->ftfR[r] = F(3 * n3);
    }
    M->fnum_diff = F(15 * n3); M->fDDIV = F(n3); M->fmflx = F(3 * n3); M->ftflx = F(3 * n3);
    M->foff = A(2 * (size_t)c.d.KA);
  }
  for (void* p : M->allocs)
    if (!p) return fail("hipMalloc failed in scale_hip_model_create (out of device memory?)");
  if (M->mixed) {
    CK(hipMemsetAsync(M->fmflx, 0, 3 * n3 * sizeof(float), c.stream));
    CK(hipMemsetAsync(M->ftflx, 0, 3 * n3 * sizeof(float), c.stream));
    CK(hipMemsetAsync(M->fDDIV, 0, n3 * sizeof(float), c.stream));
    CK(hipMemsetAsync(M->foff, 0, 2 * (size_t)c.d.KA * sizeof(double), c.stream));
  }
  CK(hipMemsetAsync(M->mflx_hi, 0, 3 * n3 * sizeof(double), c.stream));
  CK(hipMemsetAsync(M->tflx_hi, 0, 3 * n3 * sizeof(double), c.stream));
  CK(hipMemsetAsync(M->ndq, 0, 3 * n3 * sizeof(double), c.stream));
  CK(hipMemsetAsync(M->DDIV, 0, n3 * sizeof(double), c.stream));
  CK(hipMemsetAsync(M->CORIOLI, 0, c.d.pl * sizeof(double), c.stream));
  CK(hipMemsetAsync(M->wdamp_coef, 0, (size_t)c.d.KA * sizeof(double), c.stream));
  M->dtl = cfg->dtl; M->dts = cfg->dts; M->nd_coef = cfg->nd_coef; M->nd_sfc_fact = cfg->nd_sfc_fact;
  M->divdmp_coef = cfg->divdmp_coef; M->fct = cfg->fct_tracer ? 1 : 0; M->fct_along = cfg->fct_along_stream ? 1 : 0;
  M->nd_coef_q = cfg->nd_coef_q;
  M->heve = cfg->short_type == 1 ? 1 : 0;
  M->tinteg_short = ts;
  M->tinteg_tracer = cfg->tinteg_tracer;
  const int nstep = (int)ceil((cfg->dtl - shp::EPS_RP) / cfg->dts);
  if (int r = set_numfilter_tables(c, cfg->dtl / nstep, cfg->nd_coef, cfg->nd_sfc_fact)) return r;
  if (cfg->nd_coef_q != 0.0)
    if (int r = set_ndq_table(c, cfg->dtl, cfg->nd_coef_q, cfg->nd_sfc_fact)) return r;
  CK(hipStreamSynchronize(c.stream));
  return 0;
}

static double* model_field(shp::Model* M, int32_t f, int* ncomp) {
  *ncomp = 1;
  switch (f) {
    case SCALE_HIP_DENS: return M->DENS;
    case SCALE_HIP_MOMZ: return M->MOMZ;
    case SCALE_HIP_MOMX: return M->MOMX;
    case SCALE_HIP_MOMY: return M->MOMY;
    case SCALE_HIP_RHOT: return M->RHOT;
    case SCALE_HIP_REF_DENS: return M->REF_dens;
    case SCALE_HIP_REF_POTT: return M->REF_pott;
    case SCALE_HIP_REF_PRES: return M->REF_pres;
    case SCALE_HIP_MFLX_HI: *ncomp = 3; return M->mflx_hi;
    case SCALE_HIP_TFLX_HI: *ncomp = 3; return M->tflx_hi;
    default: break;
  }
  return nullptr;
}

int scale_hip_model_set(scale_hip_ctx* h, int32_t f, const double* host) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  shp::Model* M = c.model;
  if (!M) return fail("model not created");
  if (!host) return fail("null array");
  const size_t n3 = c.d.n3;
  if (f == SCALE_HIP_CORIOLI) {
    if (int r = upload_raw(c, M->CORIOLI, host, c.d.pl)) return r;
    M->has_coriolis = !all_equal(host, c.d.pl, 0.0);
    CK(hipStreamSynchronize(c.stream));
    return 0;
  }
  if (f == SCALE_HIP_WDAMP_COEF) {
    if (int r = upload_raw(c, M->wdamp_coef, host, (size_t)c.d.KA)) return r;
    CK(hipStreamSynchronize(c.stream));
    return 0;
  }
  if (f >= SCALE_HIP_DENS_TP && f <= SCALE_HIP_RHOT_TP) {
    const int v = f - SCALE_HIP_DENS_TP;
    if (!M->tp[v]) { M->tp[v] = dalloc(c, n3, &M->allocs); if (!M->tp[v]) return fail("hipMalloc failed"); }
    return upload3(c, M->tp[v], host, 1);
  }
  if (f >= SCALE_HIP_RHOQ_TP && f < SCALE_HIP_RHOQ_TP + M->QA) {
    if (!M->RHOQ_tp) {
      M->RHOQ_tp = dalloc(c, n3 * (size_t)M->QA, &M->allocs);
      if (!M->RHOQ_tp) return fail("hipMalloc failed");
      CK(hipMemsetAsync(M->RHOQ_tp, 0, n3 * (size_t)M->QA * sizeof(double), c.stream));
    }
    return upload3(c, M->RHOQ_tp + n3 * (size_t)(f - SCALE_HIP_RHOQ_TP), host, 1);
  }
  if (f >= SCALE_HIP_QTRC && f < SCALE_HIP_QTRC + M->QA) return upload3(c, M->QTRC + n3 * (size_t)(f - SCALE_HIP_QTRC), host, 1);
  if (f == SCALE_HIP_MFLX_HI || f == SCALE_HIP_TFLX_HI) return fail("mflx_hi/tflx_hi are outputs");
  int nc;
  double* dst = model_field(M, f, &nc);
  if (!dst) return fail("unknown field id");
  int r = upload3(c, dst, host, nc);
  if (r == 0 && f >= SCALE_HIP_DENS && f <= SCALE_HIP_RHOT) CK(shp::model_state_to_f32(c));
  if (r == 0) CK(hipStreamSynchronize(c.stream));
  return r;
}

int scale_hip_model_get(scale_hip_ctx* h, int32_t f, double* host) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  shp::Model* M = c.model;
  if (!M) return fail("model not created");
  if (!host) return fail("null array");
  if (f >= SCALE_HIP_QTRC && f < SCALE_HIP_QTRC + M->QA) return download3(c, host, M->QTRC + c.d.n3 * (size_t)(f - SCALE_HIP_QTRC), 1);
  int nc;
  double* src = model_field(M, f, &nc);
  if (!src) return fail("unknown field id");
  return download3(c, host, src, nc);
}

int scale_hip_model_device_ptr(scale_hip_ctx* h, int32_t f, double** dptr) {
  if (int r = check_ctx(h)) return r;
  shp::Model* M = h->c.model;
  if (!M || !dptr) return fail("model not created");
  if (f >= SCALE_HIP_QTRC && f < SCALE_HIP_QTRC + M->QA) { *dptr = M->QTRC + h->c.d.n3 * (size_t)(f - SCALE_HIP_QTRC); return 0; }
  int nc;
  *dptr = model_field(M, f, &nc);
  return *dptr ? 0 : fail("unknown field id");
}

int scale_hip_model_large_step(scale_hip_ctx* h, int32_t n) {
  if (int r = check_ctx(h)) return r;
  if (!h->c.model) return fail("model not created");
  shp::prof_start(h->c, "model_large_step", h->c.stream);
  for (int s = 0; s < n; ++s) {
    CK(shp::model_state_to_f64(h->c, shp::F64_DENS | shp::F64_RHOT));  // the preparation reads them
    CK(shp::model_large_step(h->c));
  }
  CK(shp::model_state_to_f64(h->c));
  shp::prof_stop(h->c, h->c.stream);
  return 0;
}

int scale_hip_model_prepare(scale_hip_ctx* h) {
  if (int r = check_ctx(h)) return r;
  if (!h->c.model) return fail("model not created");
  CK(shp::model_state_to_f64(h->c));
  CK(shp::model_prepare(h->c));
  return 0;
}

int scale_hip_model_short_step(scale_hip_ctx* h, int32_t n) {
  if (int r = check_ctx(h)) return r;
  if (!h->c.model) return fail("model not created");
  shp::Model& M = *h->c.model;
  const int nstep = (int)ceil((M.dtl - shp::EPS_RP) / M.dts);
  const double dts = M.dtl / nstep;
  for (int s = 0; s < n; ++s) CK(shp::model_short_step_rk4(h->c, dts));
  CK(shp::model_state_to_f64(h->c));
  return 0;
}

// restart files: 8 bytes of magic, int32 KA, IA, JA, QA, then DENS, MOMZ,
// MOMX, MOMY, RHOT and QTRC(1..QA), each a FP64 (KA,IA,JA) array in Fortran
// order, halos included
static const char kRestartMagic[8] = {'S', 'C', 'L', 'H', 'I', 'P', 'R', '1'};

int scale_hip_model_restart_write(scale_hip_ctx* h, const char* path) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  shp::Model* M = c.model;
  if (!M) return fail("model not created");
  if (!path) return fail("null path");
  if (M->mixed) return fail("restart files hold the FP64 state: a model in mixed precision cannot write one");
  const size_t n3 = c.d.n3;
  std::vector<double> buf(n3);
  FILE* f = fopen(path, "wb");
  if (!f) return fail(std::string("cannot open ") + path + " for writing");
  const int32_t hdr[4] = {c.d.KA, c.d.IA, c.d.JA, M->QA};
  bool ok = fwrite(kRestartMagic, 1, 8, f) == 8 && fwrite(hdr, sizeof(int32_t), 4, f) == 4;
  double* fields[5] = {M->DENS, M->MOMZ, M->MOMX, M->MOMY, M->RHOT};
  for (int q = 0; q < 5 + M->QA && ok; ++q) {
    const double* src = q < 5 ? fields[q] : M->QTRC + n3 * (size_t)(q - 5);
    if (int r = download3(c, buf.data(), src, 1)) { fclose(f); return r; }
    ok = fwrite(buf.data(), sizeof(double), n3, f) == n3;
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? 0 : fail(std::string("writing ") + path + " failed");
}

int scale_hip_model_restart_read(scale_hip_ctx* h, const char* path) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  shp::Model* M = c.model;
  if (!M) return fail("model not created");
  if (!path) return fail("null path");
  if (M->mixed) return fail("restart files hold the FP64 state: a model in mixed precision cannot read one");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(std::string("cannot open ") + path);
  char magic[8];
  int32_t hdr[4];
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, kRestartMagic, 8) != 0 || fread(hdr, sizeof(int32_t), 4, f) != 4) {
    fclose(f);
    return fail(std::string(path) + " is not a restart file of this library");
  }
  if (hdr[0] != c.d.KA || hdr[1] != c.d.IA || hdr[2] != c.d.JA || hdr[3] != M->QA) {
    fclose(f);
    return fail(std::string(path) + ": grid (KA,IA,JA) = (" + std::to_string(hdr[0]) + "," + std::to_string(hdr[1]) + "," +
                std::to_string(hdr[2]) + ") with " + std::to_string(hdr[3]) + " tracers does not match the model's (" +
                std::to_string(c.d.KA) + "," + std::to_string(c.d.IA) + "," + std::to_string(c.d.JA) + ") with " +
                std::to_string(M->QA));
  }
  const size_t n3 = c.d.n3;
  std::vector<double> buf(n3);
  double* fields[5] = {M->DENS, M->MOMZ, M->MOMX, M->MOMY, M->RHOT};
  for (int q = 0; q < 5 + M->QA; ++q) {
    if (fread(buf.data(), sizeof(double), n3, f) != n3) {
      fclose(f);
      return fail(std::string(path) + " is truncated");
    }
    double* dst = q < 5 ? fields[q] : M->QTRC + n3 * (size_t)(q - 5);
    if (int r = upload3(c, dst, buf.data(), 1)) { fclose(f); return r; }
  }
  fclose(f);
  CK(hipStreamSynchronize(c.stream));
  return 0;
}

int scale_hip_model_mass(scale_hip_ctx* h, int32_t f, int32_t weight_dens, double* out) {
  if (int r = check_ctx(h)) return r;
  shp::Model* M = h->c.model;
  if (!M || !out) return fail("model not created");
  const double* src;
  if (f >= SCALE_HIP_QTRC && f < SCALE_HIP_QTRC + M->QA) src = M->QTRC + h->c.d.n3 * (size_t)(f - SCALE_HIP_QTRC);
  else { int nc; src = model_field(M, f, &nc); }
  if (!src) return fail("unknown field id");
  CK(shp::mass_sum(h->c, src, weight_dens ? M->DENS : nullptr, out));
  return 0;
}

int scale_hip_model_forward_checkpointed(scale_hip_ctx* h, int32_t n, int32_t every) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  if (!c.model) return fail("model not created");
  if (n < 1 || every < 1) return fail("n and every must be positive");
  hipError_t e = shp::model_forward_checkpointed(c, n, every);
  if (e == hipErrorNotSupported)
    return fail("the adjoint supports FP64, RK4 short steps, divdmp_coef = 0 and one short step per large step");
  CK(e);
  CK(hipStreamSynchronize(c.stream));
  return 0;
}

int scale_hip_model_adjoint(scale_hip_ctx* h, double* DENS_b, double* MOMZ_b, double* MOMX_b, double* MOMY_b, double* RHOT_b) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  if (!c.model) return fail("model not created");
  double* hb[5] = {DENS_b, MOMZ_b, MOMX_b, MOMY_b, RHOT_b};
  CK(shp::adj_setup(c, 0, 0));
  double* xb[5];
  for (int v = 0; v < 5; ++v) {
    xb[v] = shp::adj_io(c, v);
    if (int r = upload3(c, xb[v], hb[v], 1)) return r;
  }
  hipError_t e = shp::model_adjoint(c, xb);
  if (e == hipErrorNotReady) return fail("scale_hip_model_forward_checkpointed must run first");
  if (e == hipErrorNotSupported)
    return fail("the adjoint supports FP64, RK4 short steps, divdmp_coef = 0 and one short step per large step");
  CK(e);
  for (int v = 0; v < 5; ++v)
    if (int r = download3(c, hb[v], xb[v], 1)) return r;
  return 0;
}

// the member contexts of a batch: each with its model, all on one device
static int batch_members(scale_hip_ctx* const* members, int32_t n, std::vector<Ctx*>& cs) {
  if (!members || n <= 0) return fail("batch: members and n > 0 are required");
  for (int m = 0; m < n; ++m) {
    if (!members[m]) return fail("batch: null member context");
    Ctx& c = members[m]->c;
    if (!c.model) return fail("batch: every member needs a created model");
    if (c.device != members[0]->c.device) return fail("batch: the members must be on one device");
    for (Ctx* o : cs)
      if (o == &c) return fail("batch: a context appears twice");
    cs.push_back(&c);
  }
  CK(hipSetDevice(cs[0]->device));
  return 0;
}

int scale_hip_batch_forward_checkpointed(scale_hip_ctx* const* members, int32_t n, int32_t nsteps, int32_t every) {
  std::vector<Ctx*> cs;
  if (int r = batch_members(members, n, cs)) return r;
  if (nsteps < 1 || every < 1) return fail("nsteps and every must be positive");
  hipError_t e = shp::batch_forward_checkpointed(cs.data(), n, nsteps, every);
  if (e == hipErrorNotSupported)
    return fail("the adjoint supports FP64, RK4 short steps, divdmp_coef = 0 and one short step per large step");
  CK(e);
  for (Ctx* c : cs) CK(hipStreamSynchronize(c->stream));
  return 0;
}

int scale_hip_batch_adjoint(scale_hip_ctx* const* members, int32_t n, double* const* DENS_b, double* const* MOMZ_b,
                            double* const* MOMX_b, double* const* MOMY_b, double* const* RHOT_b) {
  std::vector<Ctx*> cs;
  if (int r = batch_members(members, n, cs)) return r;
  double* const* hb[5] = {DENS_b, MOMZ_b, MOMX_b, MOMY_b, RHOT_b};
  for (int v = 0; v < 5; ++v)
    if (!hb[v]) return fail("batch_adjoint: null array of member fields");
  std::vector<double*> xb(5 * (size_t)n);
  for (int m = 0; m < n; ++m) {
    Ctx& c = *cs[m];
    CK(shp::adj_setup(c, 0, 0));
    for (int v = 0; v < 5; ++v) {
      xb[5 * (size_t)m + v] = shp::adj_io(c, v);
      if (int r = upload3(c, xb[5 * (size_t)m + v], hb[v][m], 1)) return r;
    }
  }
  hipError_t e = shp::batch_adjoint(cs.data(), n, xb.data());
  if (e == hipErrorNotReady) return fail("scale_hip_batch_forward_checkpointed must run first");
  if (e == hipErrorNotSupported)
    return fail("the adjoint supports FP64, RK4 short steps, divdmp_coef = 0 and one short step per large step");
  CK(e);
  for (int m = 0; m < n; ++m)
    for (int v = 0; v < 5; ++v)
      if (int r = download3(*cs[m], hb[v][m], xb[5 * (size_t)m + v], 1)) return r;
  return 0;
}

int scale_hip_profile_enable(scale_hip_ctx* h, int32_t on) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  CK(hipStreamSynchronize(c.stream));
  if (on) {
    if (!c.prof) c.prof = shp::prof_create();
    shp::prof_collect(c);
    shp::prof_reset(c.prof);
    shp::prof_outer_only(c.prof, on == 2);
  } else if (c.prof) {
    shp::prof_collect(c);
    shp::prof_destroy(c.prof);
    c.prof = nullptr;
  }
  return 0;
}

int scale_hip_profile_query(scale_hip_ctx* h, const char* name, int64_t* count, double* total_ms) {
  if (int r = check_ctx(h)) return r;
  Ctx& c = h->c;
  if (!c.prof) return fail("profiling is not enabled");
  if (!name || !count || !total_ms) return fail("null argument");
  CK(hipStreamSynchronize(c.stream));
  shp::prof_collect(c);
  long n = 0;
  shp::prof_query(c.prof, name, &n, total_ms);
  *count = n;
  return 0;
}

int scale_hip_wdamp_setup(const scale_hip_grid* g, double* wdamp_coef, const double* wdamp_tau, const double* wdamp_height,
                          const double* FZ) {
  if (!g || !wdamp_coef || !wdamp_tau || !wdamp_height || !FZ) return fail("wdamp_setup: null argument");
  const int KA = g->KA, KS = g->KS - 1, KE = g->KE - 1;  // 0-based levels; FZ holds FZ(0:KA)
  if (KS < 1 || KE >= KA || KS > KE) return fail("wdamp_setup: inconsistent KS/KE/KA");
  const double* fz = FZ + 1;  // fz[k] = FZ(k+1), the top face of level k
  const double h = *wdamp_height;
  if (h < 0.0 || fz[KE] - h < shp::EPS_RP) {
    for (int k = 0; k < KA; ++k) wdamp_coef[k] = 0.0;
    return 0;
  }
  const double alpha = 1.0 / *wdamp_tau;
  const double PI = 3.14159265358979323846;
  for (int k = KS; k <= KE; ++k) {
    const double sw = 0.5 + copysign(0.5, fz[k] - h);
    wdamp_coef[k] = alpha * sw * 0.5 * (1.0 - cos(PI * (fz[k] - h) / (fz[KE] - h)));
  }
  for (int k = 0; k < KS; ++k) wdamp_coef[k] = wdamp_coef[KS];
  for (int k = KE + 1; k < KA; ++k) wdamp_coef[k] = wdamp_coef[KE];
  return 0;
}

int scale_hip_tune_set(scale_hip_ctx* h, int32_t knob, int32_t value) {
  if (int r = check_ctx(h)) return r;
  switch (knob) {
    case SCALE_HIP_TUNE_NF_TJ:
      if (value != 0 && value != 4 && value != 8 && value != 16) return fail("SCALE_HIP_TUNE_NF_TJ must be 0, 4, 8 or 16");
      h->c.nf_tj = value;
      return 0;
    case SCALE_HIP_TUNE_TRACER_TILE:
      if (value < 0 || value > 2) return fail("SCALE_HIP_TUNE_TRACER_TILE must be 0, 1 or 2");
      h->c.tracer_tile = value;
      return 0;
    case SCALE_HIP_TUNE_ADJ_TAPE:
      if (value < 0) return fail("SCALE_HIP_TUNE_ADJ_TAPE must be >= 0");
      h->c.adj_tape = value;
      return 0;
    default:
      return fail("unknown tuning knob " + std::to_string(knob));
  }
}

int scale_hip_tune_get(scale_hip_ctx* h, int32_t knob, int32_t* value) {
  if (int r = check_ctx(h)) return r;
  if (!value) return fail("null argument");
  switch (knob) {
    case SCALE_HIP_TUNE_NF_TJ: *value = h->c.nf_tj; return 0;
    case SCALE_HIP_TUNE_TRACER_TILE: *value = h->c.tracer_tile; return 0;
    case SCALE_HIP_TUNE_ADJ_TAPE: *value = h->c.adj_tape; return 0;
    default: return fail("unknown tuning knob " + std::to_string(knob));
  }
}

int scale_hip_hbm_copy_probe(scale_hip_ctx* h, int64_t n, int32_t reps, double* avg_ms) {
  if (int r = check_ctx(h)) return r;
  if (n <= 0 || reps <= 0 || !avg_ms) return fail("hbm_copy_probe: n > 0, reps > 0 and avg_ms are required");
  Ctx& c = h->c;
  double *src = nullptr, *dst = nullptr;
  CK(hipMalloc((void**)&src, (size_t)n * sizeof(double)));
  hipError_t e = hipMalloc((void**)&dst, (size_t)n * sizeof(double));
  if (e != hipSuccess) { hipFree(src); return hipfail(e, "hipMalloc"); }
  hipEvent_t a = nullptr, b = nullptr;
  float ms = 0.f;
  if ((e = hipMemsetAsync(src, 0, (size_t)n * sizeof(double), c.stream)) == hipSuccess &&
      (e = shp::hbm_copy(c, dst, src, (size_t)n, c.stream)) == hipSuccess &&  // warm-up
      (e = hipEventCreate(&a)) == hipSuccess && (e = hipEventCreate(&b)) == hipSuccess &&
      (e = hipEventRecord(a, c.stream)) == hipSuccess) {
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = shp::hbm_copy(c, dst, src, (size_t)n, c.stream);
    if (e == hipSuccess) e = hipEventRecord(b, c.stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  }
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  hipStreamSynchronize(c.stream);
  hipFree(src);
  hipFree(dst);
  if (e != hipSuccess) return hipfail(e, "hbm_copy_probe");
  *avg_ms = (double)ms / reps;
  return 0;
}

int scale_hip_hbm_probe(scale_hip_ctx* h, int32_t kind, int64_t n, int32_t reps, double* avg_ms) {
  if (int r = check_ctx(h)) return r;
  if (kind < 0 || kind > 4 || n <= 1 || reps <= 0 || !avg_ms)
    return fail("hbm_probe: kind 0..4, n > 1, reps > 0 and avg_ms are required");
  Ctx& c = h->c;
  double *src = nullptr, *dst = nullptr, *part = nullptr;
  CK(hipMalloc((void**)&src, (size_t)n * sizeof(double)));
  hipError_t e = hipMalloc((void**)&dst, (size_t)n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&part, 4096 * sizeof(double));
  hipEvent_t a = nullptr, b = nullptr;
  float ms = 0.f;
  if (e == hipSuccess && (e = hipMemsetAsync(src, 0, (size_t)n * sizeof(double), c.stream)) == hipSuccess &&
      (e = shp::hbm_probe(c, kind, dst, src, part, (size_t)n, c.stream)) == hipSuccess &&  // warm-up
      (e = hipEventCreate(&a)) == hipSuccess && (e = hipEventCreate(&b)) == hipSuccess &&
      (e = hipEventRecord(a, c.stream)) == hipSuccess) {
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = shp::hbm_probe(c, kind, dst, src, part, (size_t)n, c.stream);
    if (e == hipSuccess) e = hipEventRecord(b, c.stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  }
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  hipStreamSynchronize(c.stream);
  if (src) hipFree(src);
  if (dst) hipFree(dst);
  if (part) hipFree(part);
  if (e != hipSuccess) return hipfail(e, "hbm_probe");
  *avg_ms = (double)ms / reps;
  return 0;
}

}  // extern "C"

// ---- ensembles ---------------------------------------------------------------------
struct scale_hip_ensemble {
  shp::Ensemble E;
};

static int check_ens(scale_hip_ensemble* e) { return e ? 0 : fail("null ensemble"); }

// device copies of every member's state (the prognostics, their FP32
// encoding in mixed precision, the tracers), for the tuner to return to
struct EnsSaved {
  std::vector<std::pair<void*, std::pair<const void*, size_t>>> items;  // copy, (source, bytes)
  ~EnsSaved() { for (auto& it : items) (void)hipFree(it.first); }
  hipError_t add(const void* src, size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return e;
    items.push_back({p, {src, bytes}});
    return hipMemcpy(p, src, bytes, hipMemcpyDeviceToDevice);
  }
  hipError_t restore(hipStream_t s) {
    for (auto& it : items) {
      hipError_t e = hipMemcpyAsync(const_cast<void*>(it.second.first), it.first, it.second.second, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(s);
  }
};

extern "C" {

int scale_hip_ensemble_create(scale_hip_ctx* const* members, int32_t n, scale_hip_ensemble** out) {
  if (!members || n <= 0 || !out) return fail("ensemble_create: members, n > 0 and out are required");
  *out = nullptr;
  auto* h = new scale_hip_ensemble();
  shp::Ensemble& E = h->E;
  for (int m = 0; m < n; ++m) {
    if (!members[m]) { delete h; return fail("null member context"); }
    E.members.push_back(&members[m]->c);
  }
  E.device = E.members[0]->device;
  const char* why = nullptr;
  if (shp::ensemble_check(E, &why) != hipSuccess) { delete h; return fail(why ? why : "unsupported ensemble"); }
  if (hipSetDevice(E.device) != hipSuccess) { delete h; return fail("hipSetDevice failed"); }
  hipDeviceGetAttribute(&E.ncu, hipDeviceAttributeMultiprocessorCount, E.device);
  if (hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking) != hipSuccess) { delete h; return fail("hipStreamCreate failed"); }
  bool ok = hipEventCreateWithFlags(&E.ev_main, hipEventDisableTiming) == hipSuccess;
  for (int l = 0; l < E.nlanes && ok; ++l)
    ok = hipStreamCreateWithFlags(&E.lane[l], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&E.ev_lane[l], hipEventDisableTiming) == hipSuccess;
  hipError_t e = ok ? shp::ensemble_build(E) : hipErrorOutOfMemory;
  if (e != hipSuccess) {
    shp::ensemble_free(E);
    hipStreamDestroy(E.stream);
    delete h;
    return hipfail(e, "ensemble_build");
  }
  *out = h;
  return 0;
}

int scale_hip_ensemble_destroy(scale_hip_ensemble* h) {
  if (!h) return 0;
  shp::ensemble_free(h->E);
  if (h->E.stream) hipStreamDestroy(h->E.stream);
  delete h;
  return 0;
}

int scale_hip_ensemble_large_step(scale_hip_ensemble* h, int32_t n) {
  if (int r = check_ens(h)) return r;
  CK(hipSetDevice(h->E.device));
  CK(shp::ensemble_large_step(h->E, n));
  return 0;
}

int scale_hip_ensemble_synchronize(scale_hip_ensemble* h) {
  if (int r = check_ens(h)) return r;
  CK(hipStreamSynchronize(h->E.stream));
  return 0;
}

int scale_hip_ensemble_set_graph(scale_hip_ensemble* h, int32_t on) {
  if (int r = check_ens(h)) return r;
  shp::Ensemble& E = h->E;
  CK(hipStreamSynchronize(E.stream));
  if (E.exec) { (void)hipGraphExecDestroy(E.exec); E.exec = nullptr; }
  if (E.graph) { (void)hipGraphDestroy(E.graph); E.graph = nullptr; }
  E.use_graph = on != 0;
  return 0;
}

static int ens_set_shape(shp::Ensemble& E, int stage_bpc, int nf_bpc, int nf_tj) {
  if (stage_bpc < 0 || stage_bpc > 2 || nf_bpc < 0 || nf_bpc > 4 || (nf_tj != 4 && nf_tj != 8))
    return fail("ensemble shape: stage_bpc 0..2, nf_bpc 0..4 and nf_tj 4 or 8");
  CK(hipStreamSynchronize(E.stream));
  if (E.exec) { (void)hipGraphExecDestroy(E.exec); E.exec = nullptr; }
  if (E.graph) { (void)hipGraphDestroy(E.graph); E.graph = nullptr; }
  const bool rebuild = nf_tj != E.nf_tj;
  E.stage_bpc = stage_bpc; E.nf_bpc = nf_bpc; E.nf_tj = nf_tj;
  if (rebuild) CK(shp::ensemble_build(E));
  return 0;
}

int scale_hip_ensemble_set_shape(scale_hip_ensemble* h, int32_t stage_bpc, int32_t nf_bpc, int32_t nf_tj) {
  if (int r = check_ens(h)) return r;
  return ens_set_shape(h->E, stage_bpc, nf_bpc, nf_tj);
}

int scale_hip_ensemble_get_shape(scale_hip_ensemble* h, int32_t* stage_bpc, int32_t* nf_bpc, int32_t* nf_tj) {
  if (int r = check_ens(h)) return r;
  if (!stage_bpc || !nf_bpc || !nf_tj) return fail("null argument");
  *stage_bpc = h->E.stage_bpc; *nf_bpc = h->E.nf_bpc; *nf_tj = h->E.nf_tj;
  return 0;
}

int scale_hip_ensemble_tune(scale_hip_ensemble* h, int32_t steps, double* best_ms, char* report, int32_t len) {
  if (int r = check_ens(h)) return r;
  if (steps < 1 || !best_ms) return fail("ensemble_tune: steps >= 1 and best_ms are required");
  shp::Ensemble& E = h->E;
  CK(hipSetDevice(E.device));
  for (shp::Ctx* cp : E.members) CK(hipStreamSynchronize(cp->stream));
  EnsSaved saved;
  for (shp::Ctx* cp : E.members) {
    shp::Model& M = *cp->model;
    const size_t b = cp->d.n3 * sizeof(double);
    double* st[5] = {M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT};
    for (int v = 0; v < 5; ++v) CK(saved.add(st[v], b));
    if (M.QA > 0) CK(saved.add(M.QTRC, b * (size_t)M.QA));
    if (M.mixed) {
      for (int v = 0; v < 5; ++v) CK(saved.add(M.fS[v], cp->d.n3 * sizeof(float)));
      CK(saved.add(M.foff, 2 * (size_t)cp->d.KA * sizeof(double)));
    }
  }
  const bool graph = E.use_graph;
  E.use_graph = false;
  hipEvent_t a = nullptr, b = nullptr;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  std::string rep;
  double best = 1e300;
  int bs = E.stage_bpc, bn = E.nf_bpc, bt = E.nf_tj;
  const int stage_c[3] = {0, 1, 2}, nfb_c[2] = {0, 2}, tj_c[2] = {8, 4};
  int err = 0;
  for (int tj : tj_c) {
    for (int sb : stage_c) {
      for (int nb : nfb_c) {
        if ((err = ens_set_shape(E, sb, nb, tj))) break;
        hipError_t e = shp::ensemble_large_step(E, 1);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(a, E.stream);
        if (e == hipSuccess) e = shp::ensemble_large_step(E, steps);
        if (e == hipSuccess) e = hipEventRecord(b, E.stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess) e = saved.restore(E.stream);
        if (e != hipSuccess) { err = hipfail(e, "ensemble_tune"); break; }
        const double per = ms / steps;
        rep += "stage_bpc=" + std::to_string(sb) + " nf_bpc=" + std::to_string(nb) + " nf_tj=" + std::to_string(tj) +
               ": " + std::to_string(per) + " ms\n";
        if (per < best) { best = per; bs = sb; bn = nb; bt = tj; }
      }
      if (err) break;
    }
    if (err) break;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  E.use_graph = graph;
  if (err) return err;
  if (int r = ens_set_shape(E, bs, bn, bt)) return r;
  *best_ms = best;
  rep += "chosen: stage_bpc=" + std::to_string(bs) + " nf_bpc=" + std::to_string(bn) + " nf_tj=" + std::to_string(bt) + "\n";
  if (report && len > 0) {
    const size_t nc = rep.size() < (size_t)len - 1 ? rep.size() : (size_t)len - 1;
    memcpy(report, rep.data(), nc);
    report[nc] = 0;
  }
  return 0;
}

}  // extern "C"
