// metrics.hip -- terrain-following metrics generated on the device from the
// surface height, for a model on a terrain-following grid.
//
// Reference (scalelib/src/atmosphere/grid/):
//   scale_atmos_grid_cartesC_real.F90:391-563   Gal-Chen heights of the cell
//     centres and faces, z = (Htop - Zs) / Htop * zeta + Zs, at the x, u, v and
//     uv points (surface heights averaged to them), and PHI = GRAV * z
//   scale_atmos_grid_cartesC_metric.F90:200-345 GSQRT(7) and J13G/J23G(7) as
//     differences of those heights
// One thread per point (k, i, j) of the interior columns writes the 21
// components there; the periodic halo fill follows.  Every expression keeps
// the reference's grouping, so the fields equal a host computation of the
// same formulas bit for bit.
#include "common.h"
#include "shp_internal.h"

namespace shp {

struct TmArgs {
  Dims d;
  const double *Zsfc, *CZ, *FZ;  // Zsfc(IA,JA); CZ(KA); FZ(0:KA) as KA+1 values
  Grid1D g;
  double Htop;
  double *GS, *J13, *J23, *PHI;
};

// surface heights at the x, u, v and uv points (REAL:413-455, 478-545); the
// last column / row repeats the edge value
struct Zs4 {
  const TmArgs& a;
  __device__ explicit Zs4(const TmArgs& a_) : a(a_) {}
  __device__ double Z(int i, int j) const { return a.Zsfc[(size_t)i + (size_t)a.d.IA * j]; }
  __device__ double x(int i, int j) const { return Z(i, j); }
  __device__ double u(int i, int j) const { return i < a.d.IA - 1 ? (Z(i, j) + Z(i + 1, j)) * 0.5 : Z(a.d.IA - 1, j); }
  __device__ double v(int i, int j) const { return j < a.d.JA - 1 ? (Z(i, j) + Z(i, j + 1)) * 0.5 : Z(i, a.d.JA - 1); }
  __device__ double uv(int i, int j) const {
    const int IA = a.d.IA, JA = a.d.JA;
    if (i < IA - 1 && j < JA - 1) return (Z(i, j) + Z(i + 1, j) + Z(i, j + 1) + Z(i + 1, j + 1)) * 0.25;
    if (j < JA - 1) return (Z(IA - 1, j) + Z(IA - 1, j + 1)) * 0.5;
    if (i < IA - 1) return (Z(i, JA - 1) + Z(i + 1, JA - 1)) * 0.5;
    return Z(IA - 1, JA - 1);
  }
  // Gal-Chen height of a reference height zeta above a surface at Zs
  __device__ double h(double Zs, double zeta) const { return (a.Htop - Zs) / a.Htop * zeta + Zs; }
};

__global__ void __launch_bounds__(256) terrain_metrics_kernel(TmArgs a) {
  const Dims& d = a.d;
  const int i = d.IS + (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = d.JS + (int)blockIdx.z;
  if (i > d.IE || k >= d.KA || j > d.JE) return;
  const Zs4 z(a);
  const size_t p = IX(d, k, i, j), n3 = d.n3;
  const double* CZ = a.CZ;
  const double* FZ = a.FZ;
  const double* RCDZ = a.g.RCDZ;
  const double* RFDZ = a.g.RFDZ;
  const double RCDX = a.g.RCDX[i], RFDX = a.g.RFDX[i], RCDY = a.g.RCDY[j], RFDY = a.g.RFDY[j];
  // G^1/2 (METRIC:231-279): the cell height over the reference height
  a.GS[p + n3 * I_XYZ] = (z.h(z.x(i, j), FZ[k + 1]) - z.h(z.x(i, j), FZ[k])) * RCDZ[k];
  a.GS[p + n3 * I_UYZ] = (z.h(z.u(i, j), FZ[k + 1]) - z.h(z.u(i, j), FZ[k])) * RCDZ[k];
  a.GS[p + n3 * I_XVZ] = (z.h(z.v(i, j), FZ[k + 1]) - z.h(z.v(i, j), FZ[k])) * RCDZ[k];
  a.GS[p + n3 * I_UVZ] = (z.h(z.uv(i, j), FZ[k + 1]) - z.h(z.uv(i, j), FZ[k])) * RCDZ[k];
  // at the top level the W-point values repeat the level below (METRIC:270-272)
  const int kw = k < d.KA - 1 ? k : d.KA - 2;
  a.GS[p + n3 * I_XYW] = (z.h(z.x(i, j), CZ[kw + 1]) - z.h(z.x(i, j), CZ[kw])) * RFDZ[kw];
  a.GS[p + n3 * I_UYW] = (z.h(z.u(i, j), CZ[kw + 1]) - z.h(z.u(i, j), CZ[kw])) * RFDZ[kw];
  a.GS[p + n3 * I_XVW] = (z.h(z.v(i, j), CZ[kw + 1]) - z.h(z.v(i, j), CZ[kw])) * RFDZ[kw];
  // J13G, J23G (METRIC:297-323): minus the horizontal slope of the height surface
  const double cz = CZ[k], fz = FZ[k + 1];
  a.J13[p + n3 * I_XYZ] = -(z.h(z.u(i, j), cz) - z.h(z.u(i - 1, j), cz)) * RCDX;
  a.J13[p + n3 * I_XYW] = -(z.h(z.u(i, j), fz) - z.h(z.u(i - 1, j), fz)) * RCDX;
  a.J13[p + n3 * I_UYW] = -(z.h(z.x(i + 1, j), fz) - z.h(z.x(i, j), fz)) * RFDX;
  a.J13[p + n3 * I_XVW] = -(z.h(z.uv(i, j), fz) - z.h(z.uv(i - 1, j), fz)) * RCDX;
  a.J13[p + n3 * I_UYZ] = -(z.h(z.x(i + 1, j), cz) - z.h(z.x(i, j), cz)) * RFDX;
  a.J13[p + n3 * I_XVZ] = -(z.h(z.uv(i, j), cz) - z.h(z.uv(i - 1, j), cz)) * RCDX;
  a.J13[p + n3 * I_UVZ] = -(z.h(z.v(i + 1, j), cz) - z.h(z.v(i, j), cz)) * RFDX;
  a.J23[p + n3 * I_XYZ] = -(z.h(z.v(i, j), cz) - z.h(z.v(i, j - 1), cz)) * RCDY;
  a.J23[p + n3 * I_XYW] = -(z.h(z.v(i, j), fz) - z.h(z.v(i, j - 1), fz)) * RCDY;
  a.J23[p + n3 * I_XVW] = -(z.h(z.x(i, j + 1), fz) - z.h(z.x(i, j), fz)) * RFDY;
  a.J23[p + n3 * I_UYW] = -(z.h(z.uv(i, j), fz) - z.h(z.uv(i, j - 1), fz)) * RCDY;
  a.J23[p + n3 * I_XVZ] = -(z.h(z.x(i, j + 1), cz) - z.h(z.x(i, j), cz)) * RFDY;
  a.J23[p + n3 * I_UYZ] = -(z.h(z.uv(i, j), cz) - z.h(z.uv(i, j - 1), cz)) * RCDY;
  a.J23[p + n3 * I_UVZ] = -(z.h(z.u(i, j + 1), cz) - z.h(z.u(i, j), cz)) * RFDY;
}

// PHI = GRAV * z at every cell centre, halos included (REAL:549)
__global__ void __launch_bounds__(256) terrain_phi_kernel(TmArgs a) {
  const Dims& d = a.d;
  const int i = (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = (int)blockIdx.z;
  if (i >= d.IA || k >= d.KA || j >= d.JA) return;
  const Zs4 z(a);
  a.PHI[IX(d, k, i, j)] = GRAV * z.h(z.x(i, j), a.CZ[k]);
}

hipError_t terrain_metrics(const Ctx& c, const double* Zsfc, const double* CZ, const double* FZ, double Htop,
                           double* GSQRT, double* J13G, double* J23G, double* PHI, hipStream_t s) {
  const Dims& d = c.d;
  TmArgs a;
  a.d = d; a.Zsfc = Zsfc; a.CZ = CZ; a.FZ = FZ; a.g = c.g1; a.Htop = Htop;
  a.GS = GSQRT; a.J13 = J13G; a.J23 = J23G; a.PHI = PHI;
  const dim3 gi((d.IE - d.IS + 1 + 63) / 64, (d.KA + 3) / 4, d.JE - d.JS + 1);
  hipLaunchKernelGGL(terrain_metrics_kernel, gi, dim3(64, 4), 0, s, a);
  const dim3 ga((d.IA + 63) / 64, (d.KA + 3) / 4, d.JA);
  hipLaunchKernelGGL(terrain_phi_kernel, ga, dim3(64, 4), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  double* f[21];
  for (int q = 0; q < 7; ++q) {
    f[q] = GSQRT + d.n3 * (size_t)q;
    f[7 + q] = J13G + d.n3 * (size_t)q;
    f[14 + q] = J23G + d.n3 * (size_t)q;
  }
  return halo_fill(c, f, 21, s);
}

}  // namespace shp
