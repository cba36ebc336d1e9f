// numfilter.hip -- hyperdiffusion fluxes num_diff(KA,IA,JA,5,3) on gfx950.
//
// Reference: ATMOS_DYN_numfilter_coef with ND_ORDER = 1 and ND_USE_RS,
//   scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:385-985,
//   calc_diff3 :1652-1800, coefficient tables filter_setup :92-319.
// One thread per (k,i,j) over every level of the interior columns writes all
// 15 outputs; the differenced fields (DENS-REF_dens, w, u, v, theta-REF_pott)
// are recomputed from the prognostics at each stencil point instead of being
// materialised.  The periodic halo fill follows as a separate launch.
#include "common.h"
#include "shp_internal.h"

namespace shp {

struct NfArgs {
  Dims d;
  const double *DENS, *MOMZ, *MOMX, *MOMY, *RHOT, *REF_dens, *REF_pott;
  const double *CNZ3, *CNX3, *CNY3;  // (3,KA,2), (3,IA,2), (3,JA,2)
  const double *ndc;                 // nd_cdz(KA) nd_fdz(KA) nd_cdx(IA) nd_fdx(IA) nd_cdy(JA) nd_fdy(JA)
  double sfc;
  double* num_diff;
};

struct Nf {
  const NfArgs& a;
  const Dims& d;
  __device__ Nf(const NfArgs& a_) : a(a_), d(a_.d) {}
  __device__ __forceinline__ size_t ix(int k, int i, int j) const { return IX(d, k, i, j); }
  __device__ __forceinline__ double cz(int n, int k, int b) const { return a.CNZ3[(n - 1) + 3 * ((size_t)k + (size_t)d.KA * (b - 1))]; }
  __device__ __forceinline__ double cx(int n, int i, int b) const { return a.CNX3[(n - 1) + 3 * ((size_t)i + (size_t)d.IA * (b - 1))]; }
  __device__ __forceinline__ double cy(int n, int j, int b) const { return a.CNY3[(n - 1) + 3 * ((size_t)j + (size_t)d.JA * (b - 1))]; }
  __device__ __forceinline__ double D(int k, int i, int j) const { return a.DENS[ix(k, i, j)]; }

  // the five differenced fields (COMMON.F90:545, 631, 715, 798, 882)
  __device__ __forceinline__ double phi(int f, int k, int i, int j) const {
    size_t p = ix(k, i, j);
    switch (f) {
      case 0: return a.DENS[p] - a.REF_dens[p];
      case 1: return 2.0 * a.MOMZ[p] / (a.DENS[ix(k + 1, i, j)] + a.DENS[p]);
      case 2: return 2.0 * a.MOMX[p] / (a.DENS[ix(k, i + 1, j)] + a.DENS[p]);
      case 3: return 2.0 * a.MOMY[p] / (a.DENS[ix(k, i, j + 1)] + a.DENS[p]);
      default: return a.RHOT[p] / a.DENS[p] - a.REF_pott[p];
    }
  }
  // calc_diff3 ZDIR, KO = 0 (COMMON.F90:1668-1697); ghost levels are the
  // negated mirror images of interior ones
  __device__ double dz0(int f, int k, int i, int j) const {
    const int KS = d.KS, KE = d.KE;
    bool neg = false;
    if (k == KS - 1) { k = KS; neg = true; }
    else if (k == KS - 2) { k = KS + 1; neg = true; }
    else if (k == KE) { k = KE - 1; neg = true; }
    else if (k == KE + 1) { k = KE - 2; neg = true; }
    else if (k == KE + 2) return 0.0;
    double v;
    if (k == KS)
      v = (+cz(1, KS + 1, 1) * phi(f, KS + 2, i, j) - cz(2, KS + 1, 1) * phi(f, KS + 1, i, j)
           + cz(3, KS + 1, 1) * phi(f, KS, i, j) - cz(1, KS, 1) * phi(f, KS + 1, i, j));
    else if (k == KE - 1)
      v = (+cz(1, KE, 1) * phi(f, KE - 1, i, j) - cz(2, KE, 1) * phi(f, KE, i, j)
           + cz(3, KE, 1) * phi(f, KE - 1, i, j) - cz(1, KE - 1, 1) * phi(f, KE - 2, i, j));
    else
      v = (+cz(1, k + 1, 1) * phi(f, k + 2, i, j) - cz(2, k + 1, 1) * phi(f, k + 1, i, j)
           + cz(3, k + 1, 1) * phi(f, k, i, j) - cz(1, k, 1) * phi(f, k - 1, i, j));
    return neg ? -v : v;
  }
  // calc_diff3 ZDIR, KO = 1 (COMMON.F90:1699-1731)
  __device__ double dz1(int f, int k, int i, int j) const {
    const int KS = d.KS, KE = d.KE;
    bool neg = false;
    if (k == KS) { k = KS + 1; neg = true; }
    else if (k == KS - 1) { k = KS + 2; neg = true; }
    else if (k == KS - 2) { k = KS + 3; neg = true; }
    else if (k == KE + 1) { k = KE; neg = true; }
    else if (k == KE + 2) { k = KE - 1; neg = true; }
    double v;
    if (k == KS + 1)
      v = (+cz(1, KS + 1, 2) * phi(f, KS + 2, i, j) - cz(2, KS + 1, 2) * phi(f, KS + 1, i, j)
           + cz(3, KS + 1, 2) * phi(f, KS, i, j) - cz(1, KS, 2) * phi(f, KS + 1, i, j));
    else if (k == KE - 1)
      v = (-cz(2, KE - 1, 2) * phi(f, KE - 1, i, j) + cz(3, KE - 1, 2) * phi(f, KE - 2, i, j)
           - cz(1, KE - 2, 2) * phi(f, KE - 3, i, j));
    else if (k == KE)
      v = (+cz(1, KE, 2) * phi(f, KE - 1, i, j) + cz(3, KE, 2) * phi(f, KE - 1, i, j)
           - cz(1, KE - 1, 2) * phi(f, KE - 2, i, j));
    else
      v = (+cz(1, k, 2) * phi(f, k + 1, i, j) - cz(2, k, 2) * phi(f, k, i, j)
           + cz(3, k, 2) * phi(f, k - 1, i, j) - cz(1, k - 1, 2) * phi(f, k - 2, i, j));
    return neg ? -v : v;
  }
  // calc_diff3 XDIR (COMMON.F90:1733-1763)
  __device__ double dx(int f, int io, int k, int i, int j) const {
    if (io == 0)
      return (+cx(1, i + 1, 1) * phi(f, k, i + 2, j) - cx(2, i + 1, 1) * phi(f, k, i + 1, j)
              + cx(3, i + 1, 1) * phi(f, k, i, j) - cx(1, i, 1) * phi(f, k, i - 1, j));
    return (+cx(1, i, 2) * phi(f, k, i + 1, j) - cx(2, i, 2) * phi(f, k, i, j)
            + cx(3, i, 2) * phi(f, k, i - 1, j) - cx(1, i - 1, 2) * phi(f, k, i - 2, j));
  }
  // calc_diff3 YDIR (COMMON.F90:1770-1800)
  __device__ double dy(int f, int jo, int k, int i, int j) const {
    if (jo == 0)
      return (+cy(1, j + 1, 1) * phi(f, k, i, j + 2) - cy(2, j + 1, 1) * phi(f, k, i, j + 1)
              + cy(3, j + 1, 1) * phi(f, k, i, j) - cy(1, j, 1) * phi(f, k, i, j - 1));
    return (+cy(1, j, 2) * phi(f, k, i, j + 1) - cy(2, j, 2) * phi(f, k, i, j)
            + cy(3, j, 2) * phi(f, k, i, j - 1) - cy(1, j - 1, 2) * phi(f, k, i, j - 2));
  }
  __device__ __forceinline__ double sfc(int k, double v) const {
    if (k == d.KS) return v * a.sfc;
    if (k == d.KS + 1) return v * (1.0 + a.sfc) * 0.5;
    return v;
  }
};

__global__ void __launch_bounds__(256) numfilter_kernel(NfArgs a, unsigned ni_tiles, unsigned nk_tiles, unsigned nblocks) {
  const Dims& d = a.d;
  Nf n(a);
  const unsigned lb = xcd_block(blockIdx.x, nblocks);
  const unsigned it = lb % ni_tiles, kt = (lb / ni_tiles) % nk_tiles, jt = lb / (ni_tiles * nk_tiles);
  const int i = d.IS + (int)(it * 64 + threadIdx.x);
  const int k = (int)(kt * 4 + threadIdx.y);
  const int j = d.JS + (int)jt;
  if (i > d.IE || k >= d.KA || j > d.JE) return;
  const int KS = d.KS, KE = d.KE;
  const size_t p = IX(d, k, i, j);
  const double* nd_cdz = a.ndc;
  const double* nd_fdz = a.ndc + d.KA;
  const double* nd_cdx = a.ndc + 2 * d.KA;
  const double* nd_fdx = a.ndc + 2 * d.KA + d.IA;
  const double* nd_cdy = a.ndc + 2 * d.KA + 2 * d.IA;
  const double* nd_fdy = a.ndc + 2 * d.KA + 2 * d.IA + d.JA;
  auto out = [&](int v, int dir) -> double& { return a.num_diff[p + d.n3 * (size_t)(v + 5 * dir)]; };
  const bool kin = k >= KS && k <= KE;
  const bool kinm = k >= KS && k <= KE - 1;
  // density (COMMON.F90:552-621)
  out(I_DENS, ZDIR) = (k >= KS - 1 && k <= KE) ? n.dz0(0, k, i, j) * nd_cdz[k] : 0.0;
  out(I_DENS, XDIR) = kin ? n.sfc(k, n.dx(0, 0, k, i, j) * nd_cdx[i]) : 0.0;
  out(I_DENS, YDIR) = kin ? n.sfc(k, n.dy(0, 0, k, i, j) * nd_cdy[j]) : 0.0;
  // z-momentum (COMMON.F90:637-707)
  out(I_MOMZ, ZDIR) = (k >= KS + 1 && k <= KE - 1) ? n.dz1(1, k, i, j) * nd_fdz[k] * n.D(k, i, j) : 0.0;
  out(I_MOMZ, XDIR) = kinm ? n.sfc(k, n.dx(1, 0, k, i, j) * nd_cdx[i] * 0.25
                                         * (n.D(k + 1, i + 1, j) + n.D(k + 1, i, j) + n.D(k, i + 1, j) + n.D(k, i, j))) : 0.0;
  out(I_MOMZ, YDIR) = kinm ? n.sfc(k, n.dy(1, 0, k, i, j) * nd_cdy[j] * 0.25
                                         * (n.D(k + 1, i, j + 1) + n.D(k + 1, i, j) + n.D(k, i, j + 1) + n.D(k, i, j))) : 0.0;
  // x-momentum (COMMON.F90:721-790)
  out(I_MOMX, ZDIR) = kinm ? n.dz0(2, k, i, j) * nd_cdz[k] * 0.25
                               * (n.D(k + 1, i + 1, j) + n.D(k + 1, i, j) + n.D(k, i + 1, j) + n.D(k, i, j)) : 0.0;
  out(I_MOMX, XDIR) = kin ? n.sfc(k, n.dx(2, 1, k, i, j) * nd_fdx[i] * n.D(k, i, j)) : 0.0;
  out(I_MOMX, YDIR) = kin ? n.sfc(k, n.dy(2, 0, k, i, j) * nd_cdy[j] * 0.25
                                        * (n.D(k, i + 1, j + 1) + n.D(k, i + 1, j) + n.D(k, i, j + 1) + n.D(k, i, j))) : 0.0;
  // y-momentum (COMMON.F90:804-873)
  out(I_MOMY, ZDIR) = kinm ? n.dz0(3, k, i, j) * nd_cdz[k] * 0.25
                               * (n.D(k + 1, i, j + 1) + n.D(k + 1, i, j) + n.D(k, i, j + 1) + n.D(k, i, j)) : 0.0;
  out(I_MOMY, XDIR) = kin ? n.sfc(k, n.dx(3, 0, k, i, j) * nd_cdx[i] * 0.25
                                        * (n.D(k, i + 1, j + 1) + n.D(k, i, j + 1) + n.D(k, i + 1, j) + n.D(k, i, j))) : 0.0;
  out(I_MOMY, YDIR) = kin ? n.sfc(k, n.dy(3, 1, k, i, j) * nd_fdy[j] * n.D(k, i, j)) : 0.0;
  // rho*theta (COMMON.F90:888-962)
  double rz = 0.0;
  if (kinm) rz = n.dz0(4, k, i, j) * nd_cdz[k] * 0.5 * (n.D(k + 1, i, j) + n.D(k, i, j));
  else if (k == KS - 1) rz = n.dz0(4, KS - 1, i, j) * nd_cdz[KS - 1] * n.D(KS, i, j);
  else if (k == KE) rz = n.dz0(4, KE, i, j) * nd_cdz[KE] * n.D(KE, i, j);
  out(I_RHOT, ZDIR) = rz;
  out(I_RHOT, XDIR) = kin ? n.sfc(k, n.dx(4, 0, k, i, j) * nd_cdx[i] * 0.5 * (n.D(k, i + 1, j) + n.D(k, i, j))) : 0.0;
  out(I_RHOT, YDIR) = kin ? n.sfc(k, n.dy(4, 0, k, i, j) * nd_cdy[j] * 0.5 * (n.D(k, i, j + 1) + n.D(k, i, j))) : 0.0;
}

hipError_t numfilter(const Ctx& c, double* num_diff, const double* const st[5], const double* REF_dens,
                     const double* REF_pott, hipStream_t s) {
  NfArgs a;
  a.d = c.d;
  a.DENS = st[0]; a.MOMZ = st[1]; a.MOMX = st[2]; a.MOMY = st[3]; a.RHOT = st[4];
  a.REF_dens = REF_dens; a.REF_pott = REF_pott;
  a.CNZ3 = c.CNZ3; a.CNX3 = c.CNX3; a.CNY3 = c.CNY3; a.ndc = c.ndc; a.sfc = c.nd_sfc_fact;
  a.num_diff = num_diff;
  const Dims& d = c.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  const unsigned ni = (imax + 63) / 64, nk = (d.KA + 3) / 4;
  const unsigned nb = ni * nk * (unsigned)jmax;
  prof_start(c, "numfilter", s);
  hipLaunchKernelGGL(numfilter_kernel, dim3(nb), dim3(64, 4), 0, s, a, ni, nk, nb);
  prof_stop(c, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  double* arrs[15];
  for (int q = 0; q < 15; ++q) arrs[q] = num_diff + d.n3 * (size_t)q;
  return halo_fill(c, arrs, 15, s);
}

}  // namespace shp
