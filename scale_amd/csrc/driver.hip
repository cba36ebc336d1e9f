// driver.hip -- the kernels around the HEVI stage and the tracer step:
// periodic halo fill (one-rank COMM_vars8), layout transposes, the RK4 short
// step, the large-step preparation and ghost-level fills, the tracer RK3
// sequence and FP64 budgets.  The device-resident model (struct Model) keeps
// every field in HBM across steps.
//
// References (scalelib/src/...):
//   communication/scale_comm_cartesC.F90:345-397,1600-2043  (halo semantics)
//   atmos-rm/dynamics/scale_atmos_dyn_tinteg_short_rk4.F90:260-660
//   atmos-rm/dynamics/scale_atmos_dyn_tinteg_tracer_rk3.F90:92-208
//   atmos-rm/dynamics/scale_atmos_dyn_tstep_large_fvm_heve.F90:377-1127
#include <math.h>
#include <string.h>
#include "common.h"
#include "shp_internal.h"
#include "batch.h"
#include "model.h"

namespace shp {

// ---- periodic halo fill ----------------------------------------------------
template <typename T>
struct HaloArgs {
  Dims d;
  T* a[16];
  int n;
  int nW;  // halo points per level: west/east strips then south/north rows
};

// halo point t of a level (west/east strips of the interior rows, then the
// south/north rows) and the interior point it copies
__device__ __forceinline__ void halo_point(const Dims& d, int t, int& i, int& j, int& si, int& sj) {
  const int IMAX = d.IE - d.IS + 1, JMAX = d.JE - d.JS + 1;
  const int hw = d.IS;  // halo width in x
  const int hs = d.JS;  // halo width in y
  const int nstrip = 2 * hw * JMAX;  // west + east columns of interior rows
  if (t < nstrip) {
    const int r = t / (2 * hw), c = t % (2 * hw);
    j = d.JS + r;
    i = c < hw ? c : d.IE + 1 + (c - hw);
  } else {
    const int u = t - nstrip;
    const int r = u / d.IA, c = u % d.IA;
    i = c;
    j = r < hs ? r : d.JE + 1 + (r - hs);
  }
  si = i < d.IS ? i + IMAX : (i > d.IE ? i - IMAX : i);
  sj = j < d.JS ? j + JMAX : (j > d.JE ? j - JMAX : j);
}

template <typename T>
__global__ void halo_kernel(HaloArgs<T> h) {
  const Dims& d = h.d;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  const int f = blockIdx.z;
  if (t >= h.nW || k >= d.KA || f >= h.n) return;
  int i, j, si, sj;
  halo_point(d, t, i, j, si, sj);
  T* a = h.a[f];
  a[IX(d, k, i, j)] = a[IX(d, k, si, sj)];
}

// the halos of many fields of many grids in one launch: blockIdx.z is the
// field, each with its own grid (ensembles)
template <typename T>
struct HaloEnt {
  Dims d;
  T* a;
  int nW;
};
template <typename T>
__global__ void halo_batch_kernel(const HaloEnt<T>* __restrict__ tab) {
  const HaloEnt<T>& e = tab[blockIdx.z];
  const Dims& d = e.d;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (t >= e.nW || k >= d.KA) return;
  int i, j, si, sj;
  halo_point(d, t, i, j, si, sj);
  e.a[IX(d, k, i, j)] = e.a[IX(d, k, si, sj)];
}

template <typename T>
static hipError_t halo_fill_t(const Ctx& c, T* const* arrs, int n, hipStream_t s) {
  const Dims& d = c.d;
  const int JMAX = d.JE - d.JS + 1;
  const int nW = 2 * d.IS * JMAX + 2 * d.JS * d.IA;
  for (int base = 0; base < n; base += 16) {
    HaloArgs<T> h;
    h.d = d;
    h.n = n - base < 16 ? n - base : 16;
    for (int q = 0; q < h.n; ++q) h.a[q] = arrs[base + q];
    h.nW = nW;
    prof_start(c, "halo_fill", s);
    hipLaunchKernelGGL(halo_kernel<T>, dim3((nW + 255) / 256, d.KA, h.n), dim3(256), 0, s, h);
    prof_stop(c, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
hipError_t halo_fill(const Ctx& c, double* const* arrs, int n, hipStream_t s) { return halo_fill_t(c, arrs, n, s); }
hipError_t halo_fill(const Ctx& c, float* const* arrs, int n, hipStream_t s) { return halo_fill_t(c, arrs, n, s); }

// ---- layout transposes: Fortran (k,i,j) <-> device (i,j,k) -------------------
// 32x32 tiles over (k, i) for each (j, component), staged through LDS so both
// sides are read and written along their contiguous dimension.
__global__ void f2d_kernel(Dims d, double* dst, const double* src) {
  __shared__ double tile[32][33];
  const int kb = blockIdx.x * 32, ib = blockIdx.y * 32;
  const int jc = blockIdx.z;  // j + JA*comp
  const size_t base = (size_t)jc * d.KA * d.IA;  // Fortran (k,i) plane of this (j,comp)
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int i = ib + r, k = kb + threadIdx.x;
    if (i < d.IA && k < d.KA) tile[r][threadIdx.x] = src[base + (size_t)k + (size_t)d.KA * i];
  }
  __syncthreads();
  const int j = jc % d.JA, comp = jc / d.JA;
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int k = kb + r, i = ib + threadIdx.x;
    if (i < d.IA && k < d.KA) dst[(size_t)comp * d.n3 + IX(d, k, i, j)] = tile[threadIdx.x][r];
  }
}
__global__ void d2f_kernel(Dims d, double* dst, const double* src) {
  __shared__ double tile[32][33];
  const int kb = blockIdx.x * 32, ib = blockIdx.y * 32;
  const int jc = blockIdx.z;
  const int j = jc % d.JA, comp = jc / d.JA;
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int k = kb + r, i = ib + threadIdx.x;
    if (i < d.IA && k < d.KA) tile[threadIdx.x][r] = src[(size_t)comp * d.n3 + IX(d, k, i, j)];
  }
  __syncthreads();
  const size_t base = (size_t)jc * d.KA * d.IA;
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int i = ib + r, k = kb + threadIdx.x;
    if (i < d.IA && k < d.KA) dst[base + (size_t)k + (size_t)d.KA * i] = tile[r][threadIdx.x];
  }
}

hipError_t to_device_layout(const Ctx& c, double* dst, const double* src_f, int ncomp, hipStream_t s) {
  const Dims& d = c.d;
  dim3 grid((d.KA + 31) / 32, (d.IA + 31) / 32, (unsigned)(d.JA * ncomp));
  hipLaunchKernelGGL(f2d_kernel, grid, dim3(32, 8), 0, s, d, dst, src_f);
  return hipGetLastError();
}
hipError_t to_fortran_layout(const Ctx& c, double* dst_f, const double* src, int ncomp, hipStream_t s) {
  const Dims& d = c.d;
  dim3 grid((d.KA + 31) / 32, (d.IA + 31) / 32, (unsigned)(d.JA * ncomp));
  hipLaunchKernelGGL(d2f_kernel, grid, dim3(32, 8), 0, s, d, dst_f, src);
  return hipGetLastError();
}

// ---- RK4 combination + ghost levels (RK4.F90:562-660, LARGE.F90:908-927) -------
template <typename R>
struct CombArgs {
  Dims d;
  R* st[5];
  const R *r1[5], *r2[5], *r3[5], *r4[5], *s0[5];
  R *mflx, *tflx;
  const R *m1, *m2, *m3, *m4, *t1, *t2, *t3, *t4;
};

template <typename R>
__device__ __forceinline__ R comb(const CombArgs<R>& a, int v, size_t p) {
  return (a.r1[v][p] * R(2) + a.r2[v][p] * R(4) + a.r3[v][p] * R(2) + a.r4[v][p] - a.s0[v][p] * R(3)) / R(6);
}

// RK4 combination (tinteg_short_rk4.F90:560-660)
template <typename R>
__global__ void __launch_bounds__(256) rk4_combine(CombArgs<R> a) {
  // one thread per interior point; the threads of levels KS and KE also fill
  // the ghost levels below/above, so the state may be combined in place
  // (st aliasing s0: every thread reads its own point before writing it)
  const Dims& d = a.d;
  const int i = d.IS + blockIdx.x * 64 + threadIdx.x;
  const int k = d.KS + blockIdx.y * 4 + threadIdx.y;
  const int j = d.JS + blockIdx.z;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const size_t p = IX(d, k, i, j);
  const int KS = d.KS, KE = d.KE;
  R v4[4];
  const int vs[4] = {0, 2, 3, 4};
#pragma unroll
  for (int q = 0; q < 4; ++q) v4[q] = comb(a, vs[q], p);
  // MOMZ: combine on KS..KE-1, MOMZ(KE) keeps the last stage's value (0)
  const R w = (k <= KE - 1) ? comb(a, 1, p) : a.r4[1][p];
#pragma unroll
  for (int q = 0; q < 4; ++q) a.st[vs[q]][p] = v4[q];
  a.st[1][p] = w;
  // ghost levels: DENS, MOMX, MOMY, RHOT copy the combined KS / KE value; MOMZ is 0
  if (k == KS || k == KE) {
    const int lo = (k == KS) ? 0 : KE + 1, hi = (k == KS) ? KS - 1 : d.KA - 1;
    for (int kk = lo; kk <= hi; ++kk) {
      const size_t g = IX(d, kk, i, j);
#pragma unroll
      for (int q = 0; q < 4; ++q) a.st[vs[q]][g] = v4[q];
      a.st[1][g] = R(0);
    }
  }
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const size_t q = p + d.n3 * (size_t)n;
    a.mflx[q] = (a.m1[q] + a.m2[q] * R(2) + a.m3[q] * R(2) + a.m4[q]) / R(6);
    a.tflx[q] = (a.t1[q] + a.t2[q] * R(2) + a.t3[q] * R(2) + a.t4[q]) / R(6);
  }
}

// RK3 end of step (tinteg_short_rk3.F90:502-571): Heun's combination
// (3 r1 + 3 r3 - 2 phi_n)/4 of the state (MOMZ on KS..KE-1) and (f1 + 3 f3)/4
// of the fluxes, or for RK3WS2002 the third stage itself; the ghost levels
// as rk4_combine (r2 and r4 are unused)
template <typename R>
__global__ void __launch_bounds__(256) rk3_combine(CombArgs<R> a, int heun) {
  const Dims& d = a.d;
  const int i = d.IS + blockIdx.x * 64 + threadIdx.x;
  const int k = d.KS + blockIdx.y * 4 + threadIdx.y;
  const int j = d.JS + blockIdx.z;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const size_t p = IX(d, k, i, j);
  const int KS = d.KS, KE = d.KE;
  auto c3 = [&](int v) { return (a.r1[v][p] * R(3) + a.r3[v][p] * R(3) - a.s0[v][p] * R(2)) / R(4); };
  R v4[4];
  const int vs[4] = {0, 2, 3, 4};
#pragma unroll
  for (int q = 0; q < 4; ++q) v4[q] = heun ? c3(vs[q]) : a.r3[vs[q]][p];
  const R w = (heun && k <= KE - 1) ? c3(1) : a.r3[1][p];
#pragma unroll
  for (int q = 0; q < 4; ++q) a.st[vs[q]][p] = v4[q];
  a.st[1][p] = w;
  if (k == KS || k == KE) {
    const int lo = (k == KS) ? 0 : KE + 1, hi = (k == KS) ? KS - 1 : d.KA - 1;
    for (int kk = lo; kk <= hi; ++kk) {
      const size_t g = IX(d, kk, i, j);
#pragma unroll
      for (int q = 0; q < 4; ++q) a.st[vs[q]][g] = v4[q];
      a.st[1][g] = R(0);
    }
  }
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const size_t q = p + d.n3 * (size_t)n;
    a.mflx[q] = heun ? (a.m1[q] + a.m3[q] * R(3)) / R(4) : a.m3[q];
    a.tflx[q] = heun ? (a.t1[q] + a.t3[q] * R(3)) / R(4) : a.t3[q];
  }
}

// element-wise conversion between the FP64 and FP32 copies of a field
template <typename To, typename From>
__global__ void convert_kernel(To* dst, const From* src, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = To(src[q]);
}
template <typename To, typename From>
static hipError_t convert(To* dst, const From* src, size_t n, hipStream_t s) {
  hipLaunchKernelGGL((convert_kernel<To, From>), dim3(2048), dim3(256), 0, s, dst, src, n);
  return hipGetLastError();
}

// the same for a field stored in FP32 as the deviation from a level profile
// off(k): FP32 = x - off(k), FP64 = FP32 + off(k)
__global__ void to_dev32(float* dst, const double* src, const double* off, size_t n, size_t pl) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = float(src[q] - off[q / pl]);
}
__global__ void from_dev32(double* dst, const float* src, const double* off, size_t n, size_t pl) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = (double)src[q] + off[q / pl];
}
// the level profiles: DENS and RHOT of the column (IS, JS) of the FP64 state
__global__ void dev_profiles(Dims d, const double* DENS, const double* RHOT, double* off) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.KA) return;
  const size_t p = IX(d, k, d.IS, d.JS);
  off[k] = DENS[p];
  off[d.KA + k] = RHOT[p];
}

// ---- 3-D momentum divergence for divergence damping ----------------------------
// ATMOS_DYN_divergence (COMMON.F90:1437-1515): one thread per point of
// i = IS..IE+1, j = JS..JE+1, k = KS-1..KE+1, the levels KS and KE with their
// one-sided vertical terms
template <bool TERRAIN>
__global__ void divergence_kernel(Dims d, Grid1D g, Met<TERRAIN> m, double J33G, double* DDIV, const double* MOMZ,
                                  const double* MOMX, const double* MOMY) {
  const int i = d.IS + blockIdx.x * 64 + threadIdx.x;
  const int k = d.KS - 1 + blockIdx.y * 4 + threadIdx.y;
  const int j = d.JS + blockIdx.z;
  if (i > d.IE + 1 || k > d.KE + 1 || j > d.JE + 1) return;
  const int KS = d.KS, KE = d.KE;
  auto X = [&](int kk, int ii, int jj) { return MOMX[IX(d, kk, ii, jj)]; };
  auto Y = [&](int kk, int ii, int jj) { return MOMY[IX(d, kk, ii, jj)]; };
  auto p = [&](int kk, int ii, int jj) { return IX(d, kk, ii, jj); };
  const double hdiv = m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY)
                    * ((X(k, i, j) * m.gs(I_UYZ, p(k, i, j)) / m.mapf(i, j, 1, I_UY)
                        - X(k, i - 1, j) * m.gs(I_UYZ, p(k, i - 1, j)) / m.mapf(i - 1, j, 1, I_UY)) * g.RCDX[i]
                       + (Y(k, i, j) * m.gs(I_XVZ, p(k, i, j)) / m.mapf(i, j, 0, I_XV)
                          - Y(k, i, j - 1) * m.gs(I_XVZ, p(k, i, j - 1)) / m.mapf(i, j - 1, 0, I_XV)) * g.RCDY[j]);
  double v;
  if (k == KS) {
    v = J33G * (MOMZ[p(KS, i, j)]) * g.RCDZ[KS]
      + ((X(KS + 1, i, j) + X(KS + 1, i - 1, j)) * m.j13(I_XYW, p(KS + 1, i, j))
         - (X(KS - 1, i, j) + X(KS, i - 1, j)) * m.j13(I_XYW, p(KS, i, j))
         + (Y(KS + 1, i, j) + Y(KS + 1, i, j - 1)) * m.j23(I_XYW, p(KS + 1, i, j))
         - (Y(KS, i, j) + Y(KS, i, j - 1)) * m.j23(I_XYW, p(KS, i, j))) * g.RFDZ[KS]
      + hdiv;
  } else if (k == KE) {
    v = J33G * (-MOMZ[p(KE - 1, i, j)]) * g.RCDZ[KE]
      + ((X(KE, i, j) + X(KE, i - 1, j)) * m.j13(I_XYW, p(KE, i, j))
         - (X(KE - 1, i, j) + X(KE - 1, i - 1, j)) * m.j13(I_XYW, p(KE - 1, i, j))
         + (Y(KE, i, j) + Y(KE, i, j - 1)) * m.j23(I_XYW, p(KE, i, j))
         - (Y(KE - 1, i, j) + Y(KE - 1, i, j - 1)) * m.j23(I_XYW, p(KE - 1, i, j))) * g.RFDZ[KE]
      + hdiv;
  } else {
    v = J33G * (MOMZ[p(k, i, j)] - MOMZ[p(k - 1, i, j)]) * g.RCDZ[k]
      + ((X(k + 1, i, j) + X(k + 1, i - 1, j)) * m.j13(I_XYW, p(k + 1, i, j))
         - (X(k - 1, i, j) + X(k - 1, i - 1, j)) * m.j13(I_XYW, p(k - 1, i, j))
         + (Y(k + 1, i, j) + Y(k + 1, i, j - 1)) * m.j23(I_XYW, p(k + 1, i, j))
         - (Y(k - 1, i, j) + Y(k - 1, i, j - 1)) * m.j23(I_XYW, p(k - 1, i, j))) / (g.FDZ[k] + g.FDZ[k - 1])
      + hdiv;
  }
  DDIV[p(k, i, j)] = v;
}

hipError_t divergence(const Ctx& c, double* DDIV, const double* MOMZ, const double* MOMX, const double* MOMY, hipStream_t s) {
  const Dims& d = c.d;
  const dim3 grid((d.IE - d.IS + 2 + 63) / 64, (d.KE - d.KS + 3 + 3) / 4, d.JE - d.JS + 2);
  if (c.terrain) {
    Met<true> m;
    m.G = c.GSQRT; m.J13 = c.J13G; m.J23 = c.J23G; m.MAPF = c.MAPF; m.n3 = d.n3; m.pl = d.pl; m.IA = d.IA;
    hipLaunchKernelGGL(divergence_kernel<true>, grid, dim3(64, 4), 0, s, d, c.g1, m, c.J33G, DDIV, MOMZ, MOMX, MOMY);
  } else {
    Met<false> m;
    m.G = c.GSQRT; m.J13 = c.J13G; m.J23 = c.J23G; m.MAPF = c.MAPF; m.n3 = d.n3; m.pl = d.pl; m.IA = d.IA;
    hipLaunchKernelGGL(divergence_kernel<false>, grid, dim3(64, 4), 0, s, d, c.g1, m, c.J33G, DDIV, MOMZ, MOMX, MOMY);
  }
  return hipGetLastError();
}

// ---- large-step preparation (LARGE.F90:443-471), dry constants ------------------
// rhot(kk) is RHOT at level kk of the point's column
template <typename F>
__device__ __forceinline__ void large_prep_point(const Dims& d, int k, int i, int j, F rhot, const double* REF_pres,
                                                 double* RT2P, double* DPRES0, double* REF_rhot) {
  const int KS = d.KS, KE = d.KE;
  const double CVdry = CPDRY - RDRY;
  const double CPovCV = CPDRY / CVdry;
  const size_t p = IX(d, k, i, j);
  if (k >= KS && k <= KE) {
    const double R = rhot(k);
    const double PRES = P00 * pow(RDRY * R / P00, CPovCV);
    RT2P[p] = CPovCV * PRES / R;
    DPRES0[p] = PRES - REF_pres[p];
    REF_rhot[p] = R;
  } else if (k == KS - 1 || k == KE + 1) {
    // ghost levels from the adjacent-but-one level; recompute that PRES here
    const int kn = k == KS - 1 ? KS + 1 : KE - 1;
    const size_t q = IX(d, kn, i, j);
    const double PRES = P00 * pow(RDRY * rhot(kn) / P00, CPovCV);
    const double dp = PRES - REF_pres[q];
    DPRES0[p] = k == KS - 1 ? dp + (REF_pres[q] - REF_pres[p]) : dp + (REF_pres[q] - REF_pres[p]);
    RT2P[p] = 0.0;
    REF_rhot[p] = 0.0;
  } else {
    RT2P[p] = 0.0; DPRES0[p] = 0.0; REF_rhot[p] = 0.0;
  }
}
__global__ void large_prep(Dims d, const double* RHOT, const double* REF_pres, double* RT2P, double* DPRES0, double* REF_rhot) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y * 4 + threadIdx.y;
  const int j = blockIdx.z;
  if (i >= d.IA || k >= d.KA || j >= d.JA) return;
  large_prep_point(d, k, i, j, [&](int kk) { return RHOT[IX(d, kk, i, j)]; }, REF_pres, RT2P, DPRES0, REF_rhot);
}

// tendencies with DAMP_alpha = 0: X_t = X_tp on the active levels, 0 elsewhere (LARGE.F90:614-839)
__global__ void tend_prep(Dims d, double* t, const double* tp, int kmax_off) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y * 4 + threadIdx.y;
  const int j = blockIdx.z;
  if (i >= d.IA || k >= d.KA || j >= d.JA) return;
  const size_t p = IX(d, k, i, j);
  const bool act = k >= d.KS && k <= d.KE - kmax_off;
  t[p] = act && tp ? tp[p] : 0.0;
}

__global__ void axpy_interior(Dims d, double* y, const double* x, double scale_div, int mode) {
  // mode 0: y += x (interior, KS..KE); mode 1: y = x / scale_div (interior, KS..KE)
  const int i = d.IS + blockIdx.x * 64 + threadIdx.x;
  const int k = d.KS + blockIdx.y * 4 + threadIdx.y;
  const int j = d.JS + blockIdx.z;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const size_t p = IX(d, k, i, j);
  if (mode == 0) y[p] = y[p] + x[p];
  else y[p] = x[p] / scale_div;
}

// av += f on the interior, the three components of a mass flux in one launch
// (blockIdx.z = j + JMAX * component)
template <typename R>
__global__ void mflx_sum(Dims d, double* av, const R* f) {
  const int jmax = d.JE - d.JS + 1;
  const int i = d.IS + blockIdx.x * 64 + threadIdx.x;
  const int k = d.KS + blockIdx.y * 4 + threadIdx.y;
  const int j = d.JS + (int)(blockIdx.z % jmax);
  const size_t comp = blockIdx.z / jmax;
  if (i > d.IE || k > d.KE) return;
  const size_t p = IX(d, k, i, j) + d.n3 * comp;
  av[p] = av[p] + (double)f[p];
}
// mflx = av / nstep on the interior, three components in one launch
__global__ void mflx_mean(Dims d, double* mflx, const double* av, double nstep) {
  const int jmax = d.JE - d.JS + 1;
  const int i = d.IS + blockIdx.x * 64 + threadIdx.x;
  const int k = d.KS + blockIdx.y * 4 + threadIdx.y;
  const int j = d.JS + (int)(blockIdx.z % jmax);
  const size_t comp = blockIdx.z / jmax;
  if (i > d.IE || k > d.KE) return;
  const size_t p = IX(d, k, i, j) + d.n3 * comp;
  mflx[p] = av[p] / nstep;
}

// DENS_RK of the first and second tracer RK3 stages (tinteg_tracer_rk3.F90:
// 140-147, 168-175), both at once: the same for every tracer of the step
__global__ void dens_rk(Dims d, double* third, double* half, const double* D0, const double* D) {
  const int i = d.IS - 1 + blockIdx.x * 64 + threadIdx.x;
  const int k = d.KS + blockIdx.y * 4 + threadIdx.y;
  const int j = d.JS - 1 + blockIdx.z;
  if (i > d.IE + 1 || k > d.KE || j > d.JE + 1) return;
  const size_t p = IX(d, k, i, j);
  const double d0 = D0[p], d1 = D[p];
  third[p] = d0 + (d1 - d0) / 3.0;
  half[p] = d0 + (d1 - d0) * 0.5;
}

// FP64 budget: per-block partial sums of a*CDZ*CDX*CDY*GSQRT/(MAPF1*MAPF2), fixed order
__global__ void mass_partial(Dims d, const double* a, const double* b, const double* CDZ, const double* CDX, const double* CDY,
                             const double* GSQRT, const double* MAPF, int terrain, double* part) {
  __shared__ double sh[256];
  const int j = d.JS + blockIdx.x;
  double s = 0.0;
  const int imax = d.IE - d.IS + 1, kmax = d.KE - d.KS + 1;
  for (int t = threadIdx.x; t < imax * kmax; t += blockDim.x) {
    const int i = d.IS + t % imax, k = d.KS + t / imax;
    const size_t p = IX(d, k, i, j);
    double v = a[p];
    if (b) v = v * b[p];
    double w = CDZ[k] * CDX[i] * CDY[j];
    if (terrain) w = w * GSQRT[p] / (MAPF[(size_t)i + (size_t)d.IA * j] * MAPF[(size_t)i + (size_t)d.IA * j + d.pl]);
    s += v * w;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void mass_final(const double* part, int n, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < n; ++q) s += part[q];
    out[0] = s;
  }
}

hipError_t mass_sum(const Ctx& c, const double* a, const double* b, double* out_host) {
  const Dims& d = c.d;
  const int jmax = d.JE - d.JS + 1;
  double* part = c.scr[NSCR - 1];
  hipLaunchKernelGGL(mass_partial, dim3(jmax), dim3(256), 0, c.stream, d, a, b, c.g1.CDZ, c.g1.CDX, c.g1.CDY, c.GSQRT, c.MAPF,
                     (int)c.terrain, part);
  hipLaunchKernelGGL(mass_final, dim3(1), dim3(64), 0, c.stream, part, jmax, part + jmax);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(out_host, part + jmax, sizeof(double), hipMemcpyDeviceToHost, c.stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(c.stream);
}

// zero fill by a kernel (a step sequence may be captured in a graph, which
// takes no memset nodes here)
template <typename T>
__global__ void fill_zero_kernel(T* a, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) a[q] = T(0);
}
template <typename T>
static hipError_t fill_zero(T* a, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(1024), dim3(256), 0, s, a, n);
  return hipGetLastError();
}

// ---- launch helpers ------------------------------------------------------------
static inline dim3 grid_full(const Dims& d) { return dim3((d.IA + 63) / 64, (d.KA + 3) / 4, d.JA); }
static inline dim3 grid_int(const Dims& d) {
  return dim3((d.IE - d.IS + 1 + 63) / 64, (d.KE - d.KS + 1 + 3) / 4, d.JE - d.JS + 1);
}
static inline dim3 grid_int3(const Dims& d) {  // interior, three components
  return dim3((d.IE - d.IS + 1 + 63) / 64, (d.KE - d.KS + 1 + 3) / 4, 3 * (d.JE - d.JS + 1));
}
static inline dim3 grid_intk(const Dims& d) {  // interior columns, all levels
  return dim3((d.IE - d.IS + 1 + 63) / 64, (d.KA + 3) / 4, d.JE - d.JS + 1);
}

#define HIPCHK(x)                          \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) return e_;       \
  } while (0)

// ---- device-resident model ------------------------------------------------------
// the arrays one RK4 short step works on, in storage type R
template <typename R>
struct ShortBufs {
  R* st[5];
  R* (*Rk)[5];
  R** t;
  R* num_diff;
  R** mfR;
  R** tfR;
  R *mflx, *tflx;
  R* DDIV;
};

// the inputs and outputs of stage r of a short step of length dts
// the fourth RK4 stage combines the stages into the state itself (flat
// HEVI on the tile kernel: FP32 always, FP64 where the tiles fill the device
// or the stage is one of an ensemble's batched launches)
template <typename R>
static bool fused_rk4(const Ctx& c, bool batched) {
  const Model& M = *c.model;
  return M.tinteg_short == 4 && !c.terrain && !M.heve && (sizeof(R) == sizeof(float) || batched || hevi_flat_tile(c));
}

template <typename R>
static StageIOT<R> short_stage_io(Ctx& c, double dts, const ShortBufs<R>& B, int r, const StepRec* rec = nullptr,
                                  bool batched = false) {
  Model& M = *c.model;
  const Dims& d = c.d;
  R* const num_diff = rec ? (R*)rec->nd : B.num_diff;
  auto Rk = [&](int q, int v) -> R* { return (rec && q < 3) ? (R*)rec->R[q][v] : B.Rk[q][v]; };
  const bool rk4 = M.tinteg_short == 4;
  const double fact_dt1 = 1.0 / 3.0, fact_dt2 = M.tinteg_short == 32 ? 1.0 / 2.0 : 2.0 / 3.0;
  double dtrk[4] = {dts / 2.0, dts / 2.0, dts, dts};
  if (!rk4) { dtrk[0] = dts * fact_dt1; dtrk[1] = dts * fact_dt2; dtrk[2] = dts; }
  StageIOT<R> io;
  io.ctx = &c;
  for (int v = 0; v < 5; ++v) {
    io.o[v] = Rk(r, v);
    io.s0[v] = B.st[v];
    io.s[v] = r == 0 ? B.st[v] : Rk(r - 1, v);
    io.t[v] = B.t[v];
  }
  io.mflx_hi = B.mfR[r];
  io.tflx_hi = B.tfR[r];
  io.DPRES0 = M.DPRES0; io.RT2P = M.RT2P; io.CORIOLI = M.CORIOLI; io.num_diff = num_diff;
  io.wdamp_coef = M.wdamp_coef; io.DDIV = B.DDIV; io.REF_dens = M.REF_dens; io.REF_rhot = M.REF_rhot;
  io.divdmp_coef = M.divdmp_coef; io.dtrk = dtrk[r]; io.has_coriolis = M.has_coriolis;
  io.combine = (r == 3 && fused_rk4<R>(c, batched)) ? 1 : 0;
  for (int v = 0; v < 5; ++v) { io.c1[v] = Rk(0, v); io.c2[v] = Rk(1, v); io.c3[v] = Rk(2, v); io.st[v] = B.st[v]; }
  for (int n = 0; n < 3; ++n) { io.cm[n] = B.mfR[n]; io.ct[n] = B.tfR[n]; }
  io.mflx_av = B.mflx; io.tflx_av = B.tflx;
  if constexpr (sizeof(R) == sizeof(float)) { io.offD = M.foff; io.offR = M.foff + d.KA; }
  if (rec) { io.trPT = rec->tr[r][0]; io.trSt = rec->tr[r][1]; io.trCp = rec->tr[r][2]; }
  return io;
}

template <typename R>
static hipError_t short_step_t(Ctx& c, double dts, const ShortBufs<R>& B, const StepRec* rec = nullptr) {
  Model& M = *c.model;
  const Dims& d = c.d;
  hipStream_t s = c.stream;
  // rec (FP64 flat HEVI): num_diff, the stage outputs R1..R3 and the stage
  // traces go to the record itself, which the reverse sweep reads
  R* const num_diff = rec ? (R*)rec->nd : B.num_diff;
  auto Rk = [&](int r, int v) -> R* { return (rec && r < 3) ? (R*)rec->R[r][v] : B.Rk[r][v]; };
  // numerical filter on the state at the start of the short step (LARGE.F90:845-858)
  if (M.nd_coef == 0.0) {
    HIPCHK(fill_zero(num_diff, 15 * d.n3, s));
  } else {
    const R* cst[5] = {B.st[0], B.st[1], B.st[2], B.st[3], B.st[4]};
    if constexpr (sizeof(R) == sizeof(double)) {
      HIPCHK(numfilter(c, num_diff, cst, M.REF_dens, M.REF_pott, s));
    } else {
      HIPCHK(numfilter(c, num_diff, cst, M.REF_dens, M.REF_pott, s, M.foff, M.foff + d.KA));
    }
  }
  // divergence for the divergence damping (LARGE.F90:860-872)
  if constexpr (sizeof(R) == sizeof(double)) {
    if (M.divdmp_coef > 0.0) HIPCHK(divergence(c, B.DDIV, B.st[1], B.st[2], B.st[3], s));
  }
  // DENS0 = DENS, ... (RK4.F90:326-336): the stages only read the state, so
  // it serves as the start value directly and is overwritten by the combination.
  // RK3 (tinteg_short_rk3.F90:141-166, 355-490): dtrk = dt * fact_dt1, dt * fact_dt2, dt
  const bool rk4 = M.tinteg_short == 4;
  const int nst = rk4 ? 4 : 3;
  for (int r = 0; r < nst; ++r) {
    const StageIOT<R> io = short_stage_io(c, dts, B, r, rec);
    if constexpr (sizeof(R) == sizeof(double)) {
      HIPCHK(M.heve ? heve_stage(io, s) : hevi_stage(io, s));
    } else {
      HIPCHK(hevi_stage(io, s));
    }
    if (r < nst - 1) {
      R* const rr[5] = {Rk(r, 0), Rk(r, 1), Rk(r, 2), Rk(r, 3), Rk(r, 4)};
      HIPCHK(halo_fill(c, rr, 5, s));
    }
  }
  if (!rk4) {
    CombArgs<R> a;
    a.d = d;
    for (int v = 0; v < 5; ++v) {
      a.st[v] = B.st[v]; a.r1[v] = Rk(0, v); a.r2[v] = Rk(1, v); a.r3[v] = Rk(2, v); a.r4[v] = Rk(2, v);
      a.s0[v] = B.st[v];
    }
    a.mflx = B.mflx; a.tflx = B.tflx;
    a.m1 = B.mfR[0]; a.m2 = B.mfR[1]; a.m3 = B.mfR[2]; a.m4 = B.mfR[2];
    a.t1 = B.tfR[0]; a.t2 = B.tfR[1]; a.t3 = B.tfR[2]; a.t4 = B.tfR[2];
    prof_start(c, "rk3_combine", s);
    hipLaunchKernelGGL(rk3_combine<R>, grid_int(d), dim3(64, 4), 0, s, a, M.tinteg_short == 3 ? 1 : 0);
    prof_stop(c, s);
    HIPCHK(hipGetLastError());
    return halo_fill(c, B.st, 5, s);
  }
  if (fused_rk4<R>(c, false)) return halo_fill(c, B.st, 5, s);
  CombArgs<R> a;
  a.d = d;
  for (int v = 0; v < 5; ++v) {
    a.st[v] = B.st[v]; a.r1[v] = Rk(0, v); a.r2[v] = Rk(1, v); a.r3[v] = Rk(2, v); a.r4[v] = B.Rk[3][v];
    a.s0[v] = B.st[v];
  }
  a.mflx = B.mflx; a.tflx = B.tflx;
  a.m1 = B.mfR[0]; a.m2 = B.mfR[1]; a.m3 = B.mfR[2]; a.m4 = B.mfR[3];
  a.t1 = B.tfR[0]; a.t2 = B.tfR[1]; a.t3 = B.tfR[2]; a.t4 = B.tfR[3];
  prof_start(c, "rk4_combine", s);
  hipLaunchKernelGGL(rk4_combine<R>, grid_int(d), dim3(64, 4), 0, s, a);
  prof_stop(c, s);
  HIPCHK(hipGetLastError());
  return halo_fill(c, B.st, 5, s);
}

// RK4 combination of four stage outputs into out (with the ghost levels and
// the periodic halos), as the last stage of a short step leaves the state;
// the flux averages go to the scratch mf/tf (3 n3 each), read as every stage's
hipError_t rk4_combine_state(Ctx& c, double* const out[5], const double* const s0[5], double* const Rk[4][5], double* mf,
                             double* tf, hipStream_t s) {
  CombArgs<double> a;
  a.d = c.d;
  for (int v = 0; v < 5; ++v) {
    a.st[v] = out[v]; a.r1[v] = Rk[0][v]; a.r2[v] = Rk[1][v]; a.r3[v] = Rk[2][v]; a.r4[v] = Rk[3][v]; a.s0[v] = s0[v];
  }
  a.mflx = mf; a.tflx = tf;
  a.m1 = a.m2 = a.m3 = a.m4 = mf;
  a.t1 = a.t2 = a.t3 = a.t4 = tf;
  hipLaunchKernelGGL(rk4_combine<double>, grid_int(c.d), dim3(64, 4), 0, s, a);
  HIPCHK(hipGetLastError());
  return halo_fill(c, out, 5, s);
}

hipError_t model_short_step_rk4(Ctx& c, double dts, const StepRec* rec) {
  Model& M = *c.model;
  if (rec && (M.mixed || M.heve || M.divdmp_coef != 0.0 || M.tinteg_short != 4)) return hipErrorNotSupported;
  if (M.mixed) {
    ShortBufs<float> B{{M.fS[0], M.fS[1], M.fS[2], M.fS[3], M.fS[4]}, M.fR, M.ft, M.fnum_diff, M.fmfR, M.ftfR, M.fmflx,
                       M.ftflx, M.fDDIV};
    return short_step_t(c, dts, B);
  }
  ShortBufs<double> B{{M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT}, M.R, M.t, M.num_diff, M.mfR, M.tfR, M.mflx_hi,
                      M.tflx_hi, M.DDIV};
  return short_step_t(c, dts, B, rec);
}

// mixed precision: the FP32 state is the integrated one; these keep the FP64
// copies (read by model_get, the budgets and the large-step preparation) and
// the FP32 copies in step
hipError_t model_state_to_f64(Ctx& c, int parts) {
  Model& M = *c.model;
  if (!M.mixed) return hipSuccess;
  double* st[5] = {M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT};
  if (parts & F64_MOM)
    for (int v = 1; v < 4; ++v) HIPCHK(convert(st[v], (const float*)M.fS[v], c.d.n3, c.stream));
  if (parts & F64_DENS)
    hipLaunchKernelGGL(from_dev32, dim3(2048), dim3(256), 0, c.stream, M.DENS, M.fS[0], M.foff, c.d.n3, c.d.pl);
  if (parts & F64_RHOT)
    hipLaunchKernelGGL(from_dev32, dim3(2048), dim3(256), 0, c.stream, M.RHOT, M.fS[4], M.foff + c.d.KA, c.d.n3, c.d.pl);
  HIPCHK(hipGetLastError());
  if (!(parts & F64_FLUX)) return hipSuccess;
  HIPCHK(convert(M.mflx_hi, (const float*)M.fmflx, 3 * c.d.n3, c.stream));
  return convert(M.tflx_hi, (const float*)M.ftflx, 3 * c.d.n3, c.stream);
}
hipError_t model_state_to_f32(Ctx& c) {
  Model& M = *c.model;
  if (!M.mixed) return hipSuccess;
  const double* st[5] = {M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT};
  // DENS and RHOT as deviations from the profiles of the current state (set
  // anew here, with every FP32 field re-encoded from the FP64 state)
  hipLaunchKernelGGL(dev_profiles, dim3((c.d.KA + 63) / 64), dim3(64), 0, c.stream, c.d, M.DENS, M.RHOT, M.foff);
  HIPCHK(hipGetLastError());
  for (int v = 1; v < 4; ++v) HIPCHK(convert(M.fS[v], st[v], c.d.n3, c.stream));
  hipLaunchKernelGGL(to_dev32, dim3(2048), dim3(256), 0, c.stream, M.fS[0], M.DENS, M.foff, c.d.n3, c.d.pl);
  hipLaunchKernelGGL(to_dev32, dim3(2048), dim3(256), 0, c.stream, M.fS[4], M.RHOT, M.foff + c.d.KA, c.d.n3, c.d.pl);
  return hipGetLastError();
}

// RT2P, DPRES0, REF_rhot from RHOT alone (LARGE.F90:443-471), into the given arrays
hipError_t large_prep_into(Ctx& c, const double* RHOT, double* RT2P, double* DPRES0, double* REF_rhot, hipStream_t s) {
  hipLaunchKernelGGL(large_prep, grid_full(c.d), dim3(64, 4), 0, s, c.d, RHOT, c.model->REF_pres, RT2P, DPRES0, REF_rhot);
  return hipGetLastError();
}

// RT2P, DPRES0, REF_rhot, DENS00, tendencies (LARGE.F90:443-839)
// the tendencies (LARGE.F90:614-839): without physics tendencies they stay
// zero and are set once
static bool tend_needed(const Model& M) {
  const bool none = !M.tp[0] && !M.tp[1] && !M.tp[2] && !M.tp[3] && !M.tp[4];
  return !(none && M.tend_zero);
}
static hipError_t model_prepare_tend(Ctx& c) {
  Model& M = *c.model;
  const Dims& d = c.d;
  hipStream_t s = c.stream;
  const bool none = !M.tp[0] && !M.tp[1] && !M.tp[2] && !M.tp[3] && !M.tp[4];
  if (!tend_needed(M)) return hipSuccess;
  for (int v = 0; v < 5; ++v) {
    hipLaunchKernelGGL(tend_prep, grid_full(d), dim3(64, 4), 0, s, d, M.t[v], (const double*)M.tp[v], v == 1 ? 1 : 0);
    HIPCHK(hipGetLastError());
    if (M.mixed) HIPCHK(convert(M.ft[v], (const double*)M.t[v], d.n3, s));
  }
  M.tend_zero = none;
  return hipSuccess;
}
hipError_t model_prepare(Ctx& c) {
  Model& M = *c.model;
  const Dims& d = c.d;
  hipStream_t s = c.stream;
  const size_t bytes = d.n3 * sizeof(double);
  hipLaunchKernelGGL(large_prep, grid_full(d), dim3(64, 4), 0, s, d, M.RHOT, M.REF_pres, M.RT2P, M.DPRES0, M.REF_rhot);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(M.DENS00, M.DENS, bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(fill_zero(M.mflx_av, 3 * d.n3, s));
  return model_prepare_tend(c);
}

// the number of short steps of a large step (LARGE.F90:402-404)
static int model_nstep(const Model& M) { return (int)ceil((M.dtl - EPS_RP) / M.dts); }

// preparation of a large step (and the record of its pressure terms)
static hipError_t large_begin(Ctx& c, const StepRec* rec) {
  Model& M = *c.model;
  const size_t bytes = c.d.n3 * sizeof(double);
  HIPCHK(model_prepare(c));
  if (rec) {
    HIPCHK(hipMemcpyAsync(rec->RT2P, M.RT2P, bytes, hipMemcpyDeviceToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(rec->DPRES0, M.DPRES0, bytes, hipMemcpyDeviceToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(rec->REF_rhot, M.REF_rhot, bytes, hipMemcpyDeviceToDevice, c.stream));
  }
  return hipSuccess;
}

// after each short step: the sum of the mass fluxes over the short steps
// (LARGE.F90:960-970), the three components in one launch (in mixed
// precision from the FP32 fluxes, widened on the fly)
static hipError_t large_after_short(Ctx& c) {
  Model& M = *c.model;
  const Dims& d = c.d;
  if (M.mixed) {
    if (M.QA <= 0) return hipSuccess;  // only the tracers read the sum
    hipLaunchKernelGGL(mflx_sum<float>, grid_int3(d), dim3(64, 4), 0, c.stream, d, M.mflx_av, (const float*)M.fmflx);
  } else {
    hipLaunchKernelGGL(mflx_sum<double>, grid_int3(d), dim3(64, 4), 0, c.stream, d, M.mflx_av, (const double*)M.mflx_hi);
  }
  return hipGetLastError();
}

// the levels KS..KE of a tracer stage's output back into the tracer (its
// ghost levels keep their values, as the reference's QTRC does)
static hipError_t copy_levels(const Ctx& c, double* dst, const double* src, hipStream_t s) {
  const Dims& d = c.d;
  const size_t off = d.pl * (size_t)d.KS, n = d.pl * (size_t)(d.KE - d.KS + 1);
  return hipMemcpyAsync(dst + off, src + off, n * sizeof(double), hipMemcpyDeviceToDevice, s);
}

// the tracers of the large step with the mass flux averaged over its nstep
// short steps (LARGE.F90:1016-1127)
static hipError_t large_tracers(Ctx& c, int nstep) {
  Model& M = *c.model;
  const Dims& d = c.d;
  hipStream_t s = c.stream;
  const double dtl = M.dtl;
  if (M.QA <= 0) return hipSuccess;
  // mixed precision: the tracers advance in FP64 on the FP64 copy of the
  // end-of-step density and the FP64 sum of the mass fluxes
  if (M.mixed) HIPCHK(model_state_to_f64(c, F64_DENS));
  hipLaunchKernelGGL(mflx_mean, grid_int3(d), dim3(64, 4), 0, s, d, M.mflx_hi, (const double*)M.mflx_av, (double)nstep);
  HIPCHK(hipGetLastError());
  double* mf[3] = {M.mflx_hi, M.mflx_hi + d.n3, M.mflx_hi + 2 * d.n3};
  HIPCHK(halo_fill(c, mf, 3, s));
  if (M.tinteg_tracer != 1) {
    hipLaunchKernelGGL(dens_rk, dim3((d.IE - d.IS + 3 + 63) / 64, (d.KE - d.KS + 1 + 3) / 4, d.JE - d.JS + 3), dim3(64, 4), 0, s,
                       d, M.DENS_RK, M.DENS_RK2, (const double*)M.DENS00, (const double*)M.DENS);
    HIPCHK(hipGetLastError());
  }
  for (int iq = 0; iq < M.QA; ++iq) {
    double* q = M.QTRC + d.n3 * (size_t)iq;
    // zero sources and, with ND_COEF_Q = 0, zero num_diff_q go in as null
    // arrays: the stages add 0.0 without loading it
    const double* rq = M.RHOQ_tp ? M.RHOQ_tp + d.n3 * (size_t)iq : nullptr;
    // tracer filter from DENS00 (LARGE.F90:1049-1061)
    if (M.nd_coef_q != 0.0) HIPCHK(numfilter_q(c, M.ndq, M.DENS00, q, nullptr, s));
    TracerIO io;
    io.RHOQ_t = rq; io.DENS0 = M.DENS00; io.mflx_hi = M.mflx_hi; io.num_diff = M.nd_coef_q != 0.0 ? M.ndq : nullptr;
    io.flag_along = M.fct_along;
    if (M.tinteg_tracer == 1) {
      // EULER (tinteg_tracer_euler.F90:80-133): one stage of dtl with the
      // end-of-step density, FCT as configured; written to QRK1 then copied
      // back (the stage reads its input at neighbours)
      io.QTRCo = M.QRK1; io.QTRC = q; io.QTRC0 = q; io.DENS = M.DENS; io.dtl = dtl; io.flag_fct = M.fct;
      HIPCHK(tracer_stage(c, io, s));
      HIPCHK(copy_levels(c, q, M.QRK1, s));
      HIPCHK(halo_fill(c, &q, 1, s));
      continue;
    }
    // stage 1: dt/3, no FCT
    io.QTRCo = M.QRK1; io.QTRC = q; io.QTRC0 = q; io.DENS = M.DENS_RK; io.dtl = dtl / 3.0; io.flag_fct = 0;
    HIPCHK(tracer_stage(c, io, s));
    HIPCHK(halo_fill(c, &M.QRK1, 1, s));
    // stage 2: dt/2, no FCT
    io.QTRCo = M.QRK2; io.QTRC = M.QRK1; io.QTRC0 = q; io.DENS = M.DENS_RK2; io.dtl = dtl / 2.0; io.flag_fct = 0;
    HIPCHK(tracer_stage(c, io, s));
    HIPCHK(halo_fill(c, &M.QRK2, 1, s));
    // stage 3: dt, FCT as configured; written to QRK1 then copied back (QTRC0 is read at neighbours)
    io.QTRCo = M.QRK1; io.QTRC = M.QRK2; io.QTRC0 = q; io.DENS = M.DENS; io.dtl = dtl; io.flag_fct = M.fct;
    HIPCHK(tracer_stage(c, io, s));
    HIPCHK(copy_levels(c, q, M.QRK1, s));
    HIPCHK(halo_fill(c, &q, 1, s));
  }
  return hipSuccess;
}

hipError_t model_large_step(Ctx& c, const StepRec* rec) {
  Model& M = *c.model;
  const int nstep = model_nstep(M);
  const double dts = M.dtl / nstep;
  if (rec && nstep != 1) return hipErrorNotSupported;
  HIPCHK(large_begin(c, rec));
  for (int step = 0; step < nstep; ++step) {
    HIPCHK(model_short_step_rk4(c, dts, rec));
    HIPCHK(large_after_short(c));
  }
  return large_tracers(c, nstep);
}

// ---- ensembles ---------------------------------------------------------------------
// The supported configuration: every member a FP64 or every member a mixed
// precision HEVI model with RK4 short steps on flat terrain without divergence
// damping, one dynamics scheme and one number of short steps per large step.
hipError_t ensemble_check(const Ensemble& E, const char** why) {
  *why = nullptr;
  if (E.members.empty()) { *why = "an ensemble needs at least one member"; return hipErrorInvalidValue; }
  const Ctx& c0 = *E.members[0];
  for (const Ctx* cp : E.members) {
    const Ctx& c = *cp;
    const Model* M = c.model;
    if (!M) { *why = "every member needs a created model"; return hipErrorInvalidValue; }
    if (c.device != E.device) { *why = "the members must be on one device"; return hipErrorInvalidValue; }
    if (c.terrain || M->heve || M->tinteg_short != 4 || M->divdmp_coef != 0.0) {
      *why = "ensemble members must be HEVI with RK4 short steps on flat terrain without divergence damping";
      return hipErrorNotSupported;
    }
    if (M->mixed != c0.model->mixed || c.scheme != c0.scheme || model_nstep(*M) != model_nstep(*c0.model)) {
      *why = "ensemble members must share the precision, the dynamics scheme and the short steps per large step";
      return hipErrorNotSupported;
    }
  }
  return hipSuccess;
}

template <typename R>
static ShortBufs<R> model_bufs(Model& M);
template <>
ShortBufs<double> model_bufs<double>(Model& M) {
  return ShortBufs<double>{{M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT}, M.R, M.t, M.num_diff, M.mfR, M.tfR, M.mflx_hi,
                           M.tflx_hi, M.DDIV};
}
template <>
ShortBufs<float> model_bufs<float>(Model& M) {
  return ShortBufs<float>{{M.fS[0], M.fS[1], M.fS[2], M.fS[3], M.fS[4]}, M.fR, M.ft, M.fnum_diff, M.fmfR, M.ftfR,
                          M.fmflx, M.ftflx, M.fDDIV};
}

// a device table of halo entries for the given fields of every member
template <typename R>
static hipError_t halo_table(Ensemble::Tab& t, const std::vector<std::pair<const Ctx*, R*>>& fields) {
  std::vector<HaloEnt<R>> ents;
  int wmax = 0, kmax = 0;
  for (const auto& f : fields) {
    const Dims& d = f.first->d;
    HaloEnt<R> e;
    e.d = d; e.a = f.second; e.nW = 2 * d.IS * (d.JE - d.JS + 1) + 2 * d.JS * d.IA;
    wmax = e.nW > wmax ? e.nW : wmax;
    kmax = d.KA > kmax ? d.KA : kmax;
    ents.push_back(e);
  }
  if (t.dev) { (void)hipFree(t.dev); t.dev = nullptr; }
  HIPCHK(hipMalloc(&t.dev, ents.size() * sizeof(HaloEnt<R>)));
  HIPCHK(hipMemcpy(t.dev, ents.data(), ents.size() * sizeof(HaloEnt<R>), hipMemcpyHostToDevice));
  t.n = (int)ents.size(); t.w = wmax; t.k = kmax;
  return hipSuccess;
}
template <typename R>
static hipError_t halo_table_launch(const Ensemble::Tab& t, hipStream_t s) {
  hipLaunchKernelGGL(halo_batch_kernel<R>, dim3((t.w + 255) / 256, t.k, t.n), dim3(256), 0, s, (const HaloEnt<R>*)t.dev);
  return hipGetLastError();
}
static void tab_free(Ensemble::Tab& t) {
  if (t.dev) (void)hipFree(t.dev);
  t = Ensemble::Tab{};
}

// ---- ensemble tracers: point tables ---------------------------------------------
// the preparation of the tracer stages of one member: the mean mass flux
// (LARGE.F90:1016-1027) on the interior and the two RK3 stage densities on
// the interior and its one-cell ring, as mflx_mean and dens_rk do
struct TrPrepEnt {
  Dims d;
  double* mflx;
  const double* av;
  double nstep;
  double *rk1, *rk2;
  const double *D00, *D;
};
__global__ void __launch_bounds__(256) tr_prep_batch(const TrPrepEnt* __restrict__ tab, const unsigned* __restrict__ first,
                                                     int nmem) {
  const unsigned lt = blockIdx.x;
  const int m = batch_member(first, nmem, lt);
  const TrPrepEnt& e = tab[m];
  const Dims& d = e.d;
  const int ri = d.IE - d.IS + 3, rj = d.JE - d.JS + 3;
  const size_t t = (size_t)(lt - first[m]) * 256 + threadIdx.x;
  const size_t np = (size_t)ri * rj * (size_t)(d.KE - d.KS + 1);
  if (t >= np) return;
  const int i = d.IS - 1 + (int)(t % (size_t)ri);
  const int j = d.JS - 1 + (int)((t / (size_t)ri) % (size_t)rj);
  const int k = d.KS + (int)(t / ((size_t)ri * rj));
  const size_t p = IX(d, k, i, j);
  const double d0 = e.D00[p], d1 = e.D[p];
  e.rk1[p] = d0 + (d1 - d0) / 3.0;
  e.rk2[p] = d0 + (d1 - d0) * 0.5;
  if (i >= d.IS && i <= d.IE && j >= d.JS && j <= d.JE)
    for (int c = 0; c < 3; ++c) e.mflx[p + d.n3 * c] = e.av[p + d.n3 * c] / e.nstep;
}
// a stage's output back into the tracer, levels KS..KE (copy_levels)
struct CopyEnt {
  double* dst;
  const double* src;
  size_t n;
};
__global__ void __launch_bounds__(256) copy_batch(const CopyEnt* __restrict__ tab, const unsigned* __restrict__ first, int nmem) {
  const unsigned lt = blockIdx.x;
  const int m = batch_member(first, nmem, lt);
  const CopyEnt& e = tab[m];
  const size_t t0 = (size_t)(lt - first[m]) * 1024 + threadIdx.x;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const size_t t = t0 + 256 * (size_t)r;
    if (t < e.n) e.dst[t] = e.src[t];
  }
}
// a point table: the entries, then first[n+1] (blocks of each member)
template <typename E>
static hipError_t point_table(Ensemble::Tab& t, const std::vector<E>& ents, const std::vector<unsigned>& blocks) {
  std::vector<unsigned> first(ents.size() + 1, 0);
  for (size_t m = 0; m < ents.size(); ++m) first[m + 1] = first[m] + blocks[m];
  const size_t ne = ents.size() * sizeof(E), off = (ne + 255) / 256 * 256;
  tab_free(t);
  HIPCHK(hipMalloc(&t.dev, off + first.size() * sizeof(unsigned)));
  HIPCHK(hipMemcpy(t.dev, ents.data(), ne, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((char*)t.dev + off, first.data(), first.size() * sizeof(unsigned), hipMemcpyHostToDevice));
  t.n = (int)ents.size(); t.total = first.back();
  return hipSuccess;
}
template <typename E>
static const unsigned* point_first(const Ensemble::Tab& t) {
  return (const unsigned*)((const char*)t.dev + ((size_t)t.n * sizeof(E) + 255) / 256 * 256);
}

// the start of the large step of every member (large_begin without records,
// the tendencies aside): in mixed precision the FP64 DENS and RHOT from the
// FP32 state (from_dev32), then RT2P, DPRES0, REF_rhot (large_prep), DENS00 =
// DENS and the flux sum cleared; one thread per point of the whole array
struct BeginEnt {
  Dims d;
  const float *fD, *fR;  // mixed precision (else null)
  const double* off;     // their level profiles (DENS, then RHOT)
  double *DENS, *RHOT, *DENS00, *mflx_av;
  const double* REF_pres;
  double *RT2P, *DPRES0, *REF_rhot;
};
__global__ void __launch_bounds__(256) begin_batch(const BeginEnt* __restrict__ tab, const unsigned* __restrict__ first, int nmem) {
  const unsigned lt = blockIdx.x;
  const int m = batch_member(first, nmem, lt);
  const BeginEnt& e = tab[m];
  const Dims& d = e.d;
  const size_t q = (size_t)(lt - first[m]) * 256 + threadIdx.x;
  if (q >= d.n3) return;
  const int k = (int)(q / d.pl);
  const size_t r = q - (size_t)k * d.pl;
  const int j = (int)(r / (size_t)d.IA), i = (int)(r - (size_t)j * d.IA);
  double D;
  if (e.fD) {
    D = (double)e.fD[q] + e.off[k];
    e.DENS[q] = D;
    e.RHOT[q] = (double)e.fR[q] + e.off[d.KA + k];
  } else {
    D = e.DENS[q];
  }
  auto rhot = [&](int kk) {
    const size_t p = IX(d, kk, i, j);
    return e.fR ? (double)e.fR[p] + e.off[d.KA + kk] : e.RHOT[p];
  };
  large_prep_point(d, k, i, j, rhot, e.REF_pres, e.RT2P, e.DPRES0, e.REF_rhot);
  e.DENS00[q] = D;
  for (int c = 0; c < 3; ++c) e.mflx_av[q + d.n3 * c] = 0.0;
}
// the flux sum after each short step (large_after_short) of every member
struct SumEnt {
  Dims d;
  double* av;
  const float* f32;
  const double* f64;
};
__global__ void __launch_bounds__(256) sum_batch(const SumEnt* __restrict__ tab, const unsigned* __restrict__ first, int nmem) {
  const unsigned lt = blockIdx.x;
  const int m = batch_member(first, nmem, lt);
  const SumEnt& e = tab[m];
  const Dims& d = e.d;
  const int ni = d.IE - d.IS + 1, nj = d.JE - d.JS + 1;
  const size_t t = (size_t)(lt - first[m]) * 256 + threadIdx.x;
  if (t >= (size_t)ni * nj * (size_t)(d.KE - d.KS + 1)) return;
  const int i = d.IS + (int)(t % (size_t)ni);
  const int j = d.JS + (int)((t / (size_t)ni) % (size_t)nj);
  const int k = d.KS + (int)(t / ((size_t)ni * nj));
  const size_t p = IX(d, k, i, j);
  for (int c = 0; c < 3; ++c) {
    const size_t q = p + d.n3 * c;
    e.av[q] = e.av[q] + (e.f32 ? (double)e.f32[q] : e.f64[q]);
  }
}
static hipError_t ensemble_begin_build(Ensemble& E) {
  std::vector<BeginEnt> be;
  std::vector<unsigned> bb;
  std::vector<SumEnt> se;
  std::vector<unsigned> sb;
  for (Ctx* cp : E.members) {
    Model& M = *cp->model;
    const Dims& d = cp->d;
    BeginEnt b{};
    b.d = d;
    if (M.mixed) { b.fD = M.fS[0]; b.fR = M.fS[4]; b.off = M.foff; }
    b.DENS = M.DENS; b.RHOT = M.RHOT; b.DENS00 = M.DENS00; b.mflx_av = M.mflx_av;
    b.REF_pres = M.REF_pres; b.RT2P = M.RT2P; b.DPRES0 = M.DPRES0; b.REF_rhot = M.REF_rhot;
    be.push_back(b);
    bb.push_back((unsigned)((d.n3 + 255) / 256));
    if (M.mixed && M.QA <= 0) continue;  // only the tracers read the sum
    SumEnt e{d, M.mflx_av, M.mixed ? (const float*)M.fmflx : nullptr, M.mixed ? nullptr : (const double*)M.mflx_hi};
    se.push_back(e);
    sb.push_back((unsigned)(((size_t)(d.IE - d.IS + 1) * (d.JE - d.JS + 1) * (d.KE - d.KS + 1) + 255) / 256));
  }
  HIPCHK(point_table(E.begin, be, bb));
  if (se.empty()) tab_free(E.sum);
  else HIPCHK(point_table(E.sum, se, sb));
  return hipSuccess;
}

// the batched tracer tables, where every member qualifies (see Ensemble)
static hipError_t ensemble_tracers_build(Ensemble& E) {
  E.tr_batched = false;
  const Model& M0 = *E.members[0]->model;
  const int qa = M0.QA;
  if (qa <= 0) return hipSuccess;
  for (Ctx* cp : E.members) {
    const Model& M = *cp->model;
    const int st = cp->scheme_tracer;
    if (M.QA != qa || M.tinteg_tracer == 1 || M.nd_coef_q != 0.0 || M.fct != M0.fct || M.fct_along != M0.fct_along ||
        cp->terrain || st != E.members[0]->scheme_tracer || st == UD5 || st == CD6)
      return hipSuccess;  // the members' own tracer steps on the lanes
  }
  const int n = (int)E.members.size();
  std::vector<const Ctx*> cs;
  for (Ctx* cp : E.members) cs.push_back(cp);
  // the FCT stage on LDS tiles for the two-cell schemes (the tiles of all
  // members keep every CU busy even where one member's would not)
  const int st = E.members[0]->scheme_tracer;
  E.tr_tile = M0.fct && (st == UD3 || st == CD4 || st == UD3KOREN || st == UD1 || st == CD2);
  for (auto* v : {&E.tr_s1, &E.tr_s2, &E.tr_ph, &E.tr_s3})
    for (auto& b : *v) batch_free(b);
  for (auto* v : {&E.tr_q, &E.tr_copy})
    for (auto& t : *v) tab_free(t);
  E.tr_s1.assign((size_t)qa, BatchBuf{}); E.tr_s2.assign((size_t)qa, BatchBuf{});
  E.tr_ph.assign((size_t)qa, BatchBuf{}); E.tr_s3.assign((size_t)qa, BatchBuf{});
  E.tr_q.assign((size_t)qa, Ensemble::Tab{}); E.tr_copy.assign((size_t)qa, Ensemble::Tab{});
  for (int iq = 0; iq < qa; ++iq) {
    std::vector<TracerIO> s1, s2, s3;
    std::vector<CopyEnt> cp_ents;
    std::vector<unsigned> cp_blocks;
    std::vector<std::pair<const Ctx*, double*>> qf;
    for (Ctx* cp : E.members) {
      Model& M = *cp->model;
      const Dims& d = cp->d;
      double* q = M.QTRC + d.n3 * (size_t)iq;
      const double* rq = M.RHOQ_tp ? M.RHOQ_tp + d.n3 * (size_t)iq : nullptr;
      TracerIO io{};  // no tracer filter in the batched case: num_diff_q is zero (null)
      io.RHOQ_t = rq; io.DENS0 = M.DENS00; io.mflx_hi = M.mflx_hi; io.num_diff = nullptr; io.flag_along = M.fct_along;
      io.QTRCo = M.QRK1; io.QTRC = q; io.QTRC0 = q; io.DENS = M.DENS_RK; io.dtl = M.dtl / 3.0; io.flag_fct = 0;
      s1.push_back(io);
      io.QTRCo = M.QRK2; io.QTRC = M.QRK1; io.DENS = M.DENS_RK2; io.dtl = M.dtl / 2.0;
      s2.push_back(io);
      io.QTRCo = M.QRK1; io.QTRC = M.QRK2; io.DENS = M.DENS; io.dtl = M.dtl; io.flag_fct = M.fct;
      s3.push_back(io);
      const size_t off = d.pl * (size_t)d.KS, nn = d.pl * (size_t)(d.KE - d.KS + 1);
      cp_ents.push_back(CopyEnt{q + off, M.QRK1 + off, nn});
      cp_blocks.push_back((unsigned)((nn + 1023) / 1024));
      qf.push_back({cp, q});
    }
    HIPCHK(tracer_batch_build(cs.data(), s1.data(), n, false, E.tr_s1[(size_t)iq]));
    HIPCHK(tracer_batch_build(cs.data(), s2.data(), n, false, E.tr_s2[(size_t)iq]));
    if (E.tr_tile) {
      HIPCHK(trt_batch_build(cs.data(), s3.data(), n, E.tr_s3[(size_t)iq]));
    } else {
      HIPCHK(tracer_batch_build(cs.data(), s3.data(), n, true, E.tr_ph[(size_t)iq]));
      HIPCHK(tracer_batch_build(cs.data(), s3.data(), n, false, E.tr_s3[(size_t)iq]));
    }
    HIPCHK(point_table(E.tr_copy[(size_t)iq], cp_ents, cp_blocks));
    HIPCHK(halo_table<double>(E.tr_q[(size_t)iq], qf));
  }
  std::vector<std::pair<const Ctx*, double*>> mf, q1, q2, rr;
  std::vector<TrPrepEnt> prep;
  std::vector<unsigned> prep_blocks;
  for (Ctx* cp : E.members) {
    Model& M = *cp->model;
    const Dims& d = cp->d;
    for (int c = 0; c < 3; ++c) mf.push_back({cp, M.mflx_hi + d.n3 * (size_t)c});
    q1.push_back({cp, M.QRK1});
    q2.push_back({cp, M.QRK2});
    rr.push_back({cp, cp->scr[1]});
    rr.push_back({cp, cp->scr[2]});
    prep.push_back(TrPrepEnt{d, M.mflx_hi, M.mflx_av, (double)model_nstep(M), M.DENS_RK, M.DENS_RK2, M.DENS00, M.DENS});
    const size_t np = (size_t)(d.IE - d.IS + 3) * (size_t)(d.JE - d.JS + 3) * (size_t)(d.KE - d.KS + 1);
    prep_blocks.push_back((unsigned)((np + 255) / 256));
  }
  HIPCHK(halo_table<double>(E.tr_mf, mf));
  HIPCHK(halo_table<double>(E.tr_q1, q1));
  HIPCHK(halo_table<double>(E.tr_q2, q2));
  HIPCHK(halo_table<double>(E.tr_r, rr));
  HIPCHK(point_table(E.tr_prep, prep, prep_blocks));
  E.tr_qa = qa; E.tr_fct = M0.fct;
  E.tr_batched = true;
  return hipSuccess;
}

// the tracers of every member (large_tracers, batched): the preparation,
// then per tracer the three stages, each with the halo fill of its output
static hipError_t ensemble_tracers(Ensemble& E, hipStream_t s) {
  hipLaunchKernelGGL(tr_prep_batch, dim3(E.tr_prep.total), dim3(256), 0, s, (const TrPrepEnt*)E.tr_prep.dev,
                     point_first<TrPrepEnt>(E.tr_prep), E.tr_prep.n);
  HIPCHK(hipGetLastError());
  HIPCHK(halo_table_launch<double>(E.tr_mf, s));
  for (int iq = 0; iq < E.tr_qa; ++iq) {
    HIPCHK(tracer_batch_launch(E.tr_s1[(size_t)iq], 0, s));  // stage 1: dt/3, no FCT
    HIPCHK(halo_table_launch<double>(E.tr_q1, s));
    HIPCHK(tracer_batch_launch(E.tr_s2[(size_t)iq], 0, s));  // stage 2: dt/2, no FCT
    HIPCHK(halo_table_launch<double>(E.tr_q2, s));
    if (E.tr_fct && E.tr_tile) {  // stage 3: dt, FCT on tiles
      HIPCHK(trt_batch_launch(E.tr_s3[(size_t)iq], 0, s));
      HIPCHK(halo_table_launch<double>(E.tr_r, s));
      HIPCHK(trt_batch_launch(E.tr_s3[(size_t)iq], 1, s));
    } else if (E.tr_fct) {  // stage 3: dt, FCT
      HIPCHK(tracer_batch_launch(E.tr_ph[(size_t)iq], 1, s));
      HIPCHK(tracer_batch_launch(E.tr_s3[(size_t)iq], 2, s));
      HIPCHK(halo_table_launch<double>(E.tr_r, s));
      HIPCHK(tracer_batch_launch(E.tr_s3[(size_t)iq], 3, s));
    } else {
      HIPCHK(tracer_batch_launch(E.tr_s3[(size_t)iq], 0, s));
    }
    const Ensemble::Tab& t = E.tr_copy[(size_t)iq];
    hipLaunchKernelGGL(copy_batch, dim3(t.total), dim3(256), 0, s, (const CopyEnt*)t.dev, point_first<CopyEnt>(t), t.n);
    HIPCHK(hipGetLastError());
    HIPCHK(halo_table_launch<double>(E.tr_q[(size_t)iq], s));
  }
  return hipSuccess;
}

template <typename R>
static hipError_t ensemble_build_t(Ensemble& E) {
  const int n = (int)E.members.size();
  // halo tables: slot 0 the filter's 15 fields, 1..3 the stage outputs of
  // stages 1..3, 4 the state
  for (int slot = 0; slot < 5; ++slot) {
    std::vector<std::pair<const Ctx*, R*>> f;
    for (Ctx* cp : E.members) {
      const ShortBufs<R> B = model_bufs<R>(*cp->model);
      if (slot == 0) {
        for (int q = 0; q < 15; ++q) f.push_back({cp, B.num_diff + cp->d.n3 * (size_t)q});
      } else if (slot < 4) {
        for (int v = 0; v < 5; ++v) f.push_back({cp, B.Rk[slot - 1][v]});
      } else {
        for (int v = 0; v < 5; ++v) f.push_back({cp, B.st[v]});
      }
    }
    HIPCHK(halo_table<R>(E.halo[slot], f));
  }
  for (int r = 0; r < 4; ++r) {
    std::vector<StageIOT<R>> ios;
    for (Ctx* cp : E.members) {
      Model& M = *cp->model;
      const ShortBufs<R> B = model_bufs<R>(M);
      ios.push_back(short_stage_io(*cp, M.dtl / model_nstep(M), B, r, nullptr, true));
    }
    HIPCHK(hevi_batch_build(ios.data(), n, E.stage[r]));
  }
  std::vector<NfIn<R>> in;
  for (Ctx* cp : E.members) {
    Model& M = *cp->model;
    const ShortBufs<R> B = model_bufs<R>(M);
    NfIn<R> f;
    f.c = cp; f.num_diff = B.num_diff;
    for (int v = 0; v < 5; ++v) f.st[v] = B.st[v];
    f.REF_dens = M.REF_dens; f.REF_pott = M.REF_pott;
    f.offD = M.mixed ? M.foff : nullptr; f.offR = M.mixed ? M.foff + cp->d.KA : nullptr;
    in.push_back(f);
  }
  HIPCHK(numfilter_batch_build(in.data(), n, E.nf_tj, E.nf));
  HIPCHK(ensemble_begin_build(E));
  return ensemble_tracers_build(E);
}

hipError_t ensemble_build(Ensemble& E) {
  const char* why;
  HIPCHK(ensemble_check(E, &why));
  E.built = false;
  hipError_t e = E.members[0]->model->mixed ? ensemble_build_t<float>(E) : ensemble_build_t<double>(E);
  if (e == hipSuccess) E.built = true;
  return e;
}

// one large step of every member in the phase-major order: the batched
// launches on the ensemble's stream, each member's own launches on its lane
template <typename R>
static hipError_t ensemble_step_t(Ensemble& E) {
  hipStream_t s = E.stream;
  const int nstep = model_nstep(*E.members[0]->model);
  // the lanes wait for the ensemble stream, then the ensemble stream for them
  auto fork = [&]() -> hipError_t {
    HIPCHK(hipEventRecord(E.ev_main, s));
    for (int l = 0; l < E.nlanes; ++l) HIPCHK(hipStreamWaitEvent(E.lane[l], E.ev_main, 0));
    return hipSuccess;
  };
  auto join = [&]() -> hipError_t {
    for (int l = 0; l < E.nlanes; ++l) {
      HIPCHK(hipEventRecord(E.ev_lane[l], E.lane[l]));
      HIPCHK(hipStreamWaitEvent(s, E.ev_lane[l], 0));
    }
    return hipSuccess;
  };
  // the preparation of every member (reading the FP64 DENS and RHOT, refreshed
  // from the FP32 state in mixed precision) in one launch; the tendencies on
  // the lanes where they are (re)computed
  hipLaunchKernelGGL(begin_batch, dim3(E.begin.total), dim3(256), 0, s, (const BeginEnt*)E.begin.dev,
                     point_first<BeginEnt>(E.begin), E.begin.n);
  HIPCHK(hipGetLastError());
  bool tend = false;
  for (Ctx* cp : E.members) tend = tend || tend_needed(*cp->model);
  if (tend) {
    HIPCHK(fork());
    for (Ctx* cp : E.members) HIPCHK(model_prepare_tend(*cp));
    HIPCHK(join());
  }
  for (int step = 0; step < nstep; ++step) {
    // numerical filter (LARGE.F90:845-858) and its halos: one launch each
    // (a member with ND_COEF = 0 gets zeros)
    HIPCHK(numfilter_batch_launch(E.nf, E.nf_bpc, E.ncu, s));
    HIPCHK(halo_table_launch<R>(E.halo[0], s));
    for (Ctx* cp : E.members) {
      if (cp->model->nd_coef == 0.0) HIPCHK(fill_zero(model_bufs<R>(*cp->model).num_diff, 15 * cp->d.n3, s));
    }
    // four stages, the last with the RK4 combination into the state, each
    // with the halos of its outputs in one launch
    for (int r = 0; r < 4; ++r) {
      HIPCHK(hevi_batch_launch(E.stage[r], E.stage_bpc, E.ncu, s));
      HIPCHK(halo_table_launch<R>(E.halo[r + 1], s));
    }
    if (E.sum.dev) {  // the flux sums of every member
      hipLaunchKernelGGL(sum_batch, dim3(E.sum.total), dim3(256), 0, s, (const SumEnt*)E.sum.dev, point_first<SumEnt>(E.sum),
                         E.sum.n);
      HIPCHK(hipGetLastError());
    }
  }
  if (!E.tr_batched) {
    HIPCHK(fork());
    for (Ctx* cp : E.members) HIPCHK(large_tracers(*cp, nstep));
    return join();
  }
  if (E.members[0]->model->mixed) {  // the tracers advance on the FP64 copy of the density
    HIPCHK(fork());
    for (Ctx* cp : E.members) HIPCHK(model_state_to_f64(*cp, F64_DENS));
    HIPCHK(join());
  }
  return ensemble_tracers(E, s);
}

// every member's FP64 copies of its FP32 state and fluxes (mixed precision),
// as scale_hip_model_large_step leaves them
static hipError_t ensemble_finish(Ensemble& E) {
  hipStream_t s = E.stream;
  HIPCHK(hipEventRecord(E.ev_main, s));
  for (int l = 0; l < E.nlanes; ++l) HIPCHK(hipStreamWaitEvent(E.lane[l], E.ev_main, 0));
  for (Ctx* cp : E.members) HIPCHK(model_state_to_f64(*cp));
  for (int l = 0; l < E.nlanes; ++l) {
    HIPCHK(hipEventRecord(E.ev_lane[l], E.lane[l]));
    HIPCHK(hipStreamWaitEvent(s, E.ev_lane[l], 0));
  }
  return hipSuccess;
}

// the members' own streams and timers are set aside while the ensemble's
// lanes carry their work
struct MemberStreams {
  Ensemble& E;
  std::vector<hipStream_t> st;
  std::vector<Prof*> pr;
  explicit MemberStreams(Ensemble& e) : E(e) {
    for (size_t m = 0; m < E.members.size(); ++m) {
      Ctx* cp = E.members[m];
      st.push_back(cp->stream); pr.push_back(cp->prof);
      cp->stream = E.lane[m % (size_t)E.nlanes]; cp->prof = nullptr;
    }
  }
  ~MemberStreams() {
    for (size_t m = 0; m < E.members.size(); ++m) { E.members[m]->stream = st[m]; E.members[m]->prof = pr[m]; }
  }
};

static hipError_t ensemble_step_direct(Ensemble& E) {
  return E.members[0]->model->mixed ? ensemble_step_t<float>(E) : ensemble_step_t<double>(E);
}

hipError_t ensemble_large_step(Ensemble& E, int n) {
  if (!E.built) HIPCHK(ensemble_build(E));
  // the members' pending work on their own streams comes first
  for (Ctx* cp : E.members) HIPCHK(hipStreamSynchronize(cp->stream));
  MemberStreams guard(E);
  for (int q = 0; q < n; ++q) {
    if (!E.use_graph) {
      HIPCHK(ensemble_step_direct(E));
      continue;
    }
    if (!E.exec) {
      HIPCHK(hipStreamBeginCapture(E.stream, hipStreamCaptureModeRelaxed));
      const hipError_t e = ensemble_step_direct(E);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(E.stream, &g);
      if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return e; }
      HIPCHK(e2);
      E.graph = g;
      HIPCHK(hipGraphInstantiate(&E.exec, E.graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(E.exec, E.stream));
  }
  return E.members[0]->model->mixed ? ensemble_finish(E) : hipSuccess;
}

void ensemble_free(Ensemble& E) {
  if (E.stream) (void)hipStreamSynchronize(E.stream);
  for (int l = 0; l < Ensemble::MAX_LANES; ++l) {
    if (E.lane[l]) { (void)hipStreamSynchronize(E.lane[l]); (void)hipStreamDestroy(E.lane[l]); E.lane[l] = nullptr; }
    if (E.ev_lane[l]) { (void)hipEventDestroy(E.ev_lane[l]); E.ev_lane[l] = nullptr; }
  }
  if (E.ev_main) { (void)hipEventDestroy(E.ev_main); E.ev_main = nullptr; }
  if (E.exec) (void)hipGraphExecDestroy(E.exec);
  if (E.graph) (void)hipGraphDestroy(E.graph);
  E.exec = nullptr; E.graph = nullptr;
  for (auto& b : E.stage) batch_free(b);
  batch_free(E.nf);
  for (auto& t : E.halo) tab_free(t);
  for (auto* v : {&E.tr_s1, &E.tr_s2, &E.tr_ph, &E.tr_s3}) {
    for (auto& b : *v) batch_free(b);
    v->clear();
  }
  for (auto* t : {&E.tr_mf, &E.tr_q1, &E.tr_q2, &E.tr_r, &E.tr_prep, &E.begin, &E.sum}) tab_free(*t);
  for (auto* v : {&E.tr_q, &E.tr_copy}) {
    for (auto& t : *v) tab_free(t);
    v->clear();
  }
  E.tr_batched = false;
  E.built = false;
}

}  // namespace shp
