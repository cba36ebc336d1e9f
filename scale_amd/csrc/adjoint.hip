// adjoint.hip -- reverse-mode derivative (adjoint) of the HEVI stage, the
// numerical filter and the large step, with checkpointing.
//
// The reference has no adjoint.  This is the transpose of the Jacobian of the
// forward this library computes (ATMOS_DYN_Tstep_short_fvm_hevi,
// scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short_fvm_hevi.F90:128-1089;
// ATMOS_DYN_numfilter_coef, scale_atmos_dyn_common.F90:385-985; the RK4 short
// step, scale_atmos_dyn_tinteg_short_rk4.F90:300-660; and the large-step
// preparation, scale_atmos_dyn_tstep_large_fvm_heve.F90:443-471,908-927),
// derived statement by statement.  The upwind switch sign(vel) is held fixed
// (the derivative wherever no face velocity is exactly zero).
//
// Scope: flat or terrain-following grids (GSQRT, J13G, J23G, J33G and MAPF
// as the context holds them; the metrics are constants of the adjoint), UD3
// or CD4, FP64, divdmp_coef = 0, one short step per large step, periodic
// lateral boundaries.  Tracers are passive (they never feed back into the
// dynamics), so the adjoint of the dynamics ignores them.
//
// Structure: each loop nest of the forward becomes one launch over the same
// index range, which runs the loop body backwards: the adjoint of an output
// point is pushed to the inputs that point read, with FP64 global atomics
// where several points read the same input.  The vertical solve runs one
// thread per column (the transposed Thomas sweep).  Forward values the
// reverse needs (PT, St and the solved C of each stage) are recorded by
// hevi_stage_traced when the step's stages are recomputed.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include <stdlib.h>

#include "common.h"
#include "model.h"
#include "shp_internal.h"

namespace shp {

namespace {

__device__ __forceinline__ void aadd(double* p, double v) { unsafeAtomicAdd(p, v); }

// index box, inclusive bounds; one thread per (k, i, j), i fastest
struct Box {
  int k0, k1, i0, i1, j0, j1;
};

template <typename F>
__global__ void __launch_bounds__(256) for3_kernel(Box b, F f) {
  const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
  const size_t n = (size_t)ni * (size_t)nj * (size_t)nk;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = b.i0 + (int)(t % (size_t)ni);
  const size_t r = t / (size_t)ni;
  const int j = b.j0 + (int)(r % (size_t)nj);
  const int k = b.k0 + (int)(r / (size_t)nj);
  f(k, i, j);
}

template <typename F>
hipError_t for3(Box b, F f, hipStream_t s) {
  if (b.k1 < b.k0 || b.i1 < b.i0 || b.j1 < b.j0) return hipSuccess;
  const size_t n = (size_t)(b.i1 - b.i0 + 1) * (size_t)(b.j1 - b.j0 + 1) * (size_t)(b.k1 - b.k0 + 1);
  hipLaunchKernelGGL(for3_kernel<F>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, f);
  return hipGetLastError();
}

#define HIPCHK(x)                          \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) return e_;       \
  } while (0)

// d(face value)/d(qm1, q0, q1, q2) for a fixed sign of vel (kind 0 interior,
// 1 lowest face q0..q2, 2 highest face qm1..q1)
template <int S>
__device__ __forceinline__ void fv_weights(int kind, double vel, double w[4]) {
  w[0] = w[1] = w[2] = w[3] = 0.0;
  if constexpr (S == CD4) {
    if (kind == 0) { w[0] = F42; w[1] = F41; w[2] = F41; w[3] = F42; }
    else { w[1] = F2; w[2] = F2; }
  } else {
    const double sg = fsign(1.0, vel), sh = fsign(0.5, vel);
    if (kind == 0) {
      w[0] = F31 + sg * F31; w[1] = F32 + sg * F33; w[2] = F32 - sg * F33; w[3] = F31 - sg * F31;
    } else if (kind == 1) {
      w[1] = F2 * (0.5 + sh) + (2.0 / 6.0) * (0.5 - sh);
      w[2] = F2 * (0.5 + sh) + (5.0 / 6.0) * (0.5 - sh);
      w[3] = -(1.0 / 6.0) * (0.5 - sh);
    } else {
      w[0] = -(1.0 / 6.0) * (0.5 + sh);
      w[1] = (5.0 / 6.0) * (0.5 + sh) + F2 * (0.5 - sh);
      w[2] = (2.0 / 6.0) * (0.5 + sh) + F2 * (0.5 - sh);
    }
  }
}

// Face value of q at the face whose q0 is q[p] (stride st) and its adjoint:
// sink(m, scale * dfv/dq_m) for the points p + (m-1)*st.  Returns the face value.
template <int S, typename Sink>
__device__ __forceinline__ double face_adj_s(int kind, const double* q, size_t p, size_t st, double vel, double scale, Sink&& sink) {
  double w[4];
  fv_weights<S>(kind, vel, w);
  if (scale != 0.0) {
    if (kind != 1) sink(0, scale * w[0]);
    sink(1, scale * w[1]);
    sink(2, scale * w[2]);
    if (kind != 2) sink(3, scale * w[3]);
  }
  if (kind == 1) return fv_lo<S, double>(q[p], q[p + st], q[p + 2 * st], vel);
  if (kind == 2) return fv_hi<S, double>(q[p - st], q[p], q[p + st], vel);
  return fv_int<S, double>(q[p - st], q[p], q[p + st], q[p + 2 * st], vel);
}

// the same, adding to qb[p + (m-1)*st] with atomics
template <int S>
__device__ __forceinline__ double face_adj(int kind, const double* q, double* qb, size_t p, size_t st, double vel, double scale) {
  return face_adj_s<S>(kind, q, p, st, vel, scale, [&](int m, double v) { aadd(qb + p + (size_t)(m - 1) * st, v); });
}

// Faces of a box along z, y or x (DIR 0, 1, 2), L consecutive faces of a
// column (a y row, an x row) per thread, in order: the adjoint of a stencil
// along that direction (contributions to the points l-1..l+2 of the face's
// own line) collects in a register window and each point of the line is
// added once, without an atomic where every face that reaches it (faces
// q-2..q+1) is the thread's own.  f(k, i, j, win) adds win.a[m] for the point l-1+m of tgt, and
// win.b[m] for the point l-1+m of tgt2 (added with atomics: other lines may
// reach it too).  The x stencils keep one face per thread (xrow below): with
// a thread per x row the lanes of a wave read 64 rows apart, and the warm
// bubble's reverse sweep ran 32 % slower that way than with atomics.
struct ZWin {
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  double b[4] = {0.0, 0.0, 0.0, 0.0};
};

template <int DIR, typename F>
__global__ void __launch_bounds__(256) lseg_kernel(Dims d, Box b, int L, double* tgt, double* tgt2, F f) {
  const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
  // the fastest thread index and the other one across the lines (i and j,
  // i and k, j and k), and the line's own index (k, j, i)
  const int n1 = DIR == 2 ? nj : ni, n2 = DIR == 0 ? nj : nk, nl = DIR == 0 ? nk : (DIR == 1 ? nj : ni);
  const int nseg = (nl + L - 1) / L;
  const size_t n = (size_t)n1 * (size_t)n2 * (size_t)nseg;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int c1 = (int)(t % (size_t)n1);
  const size_t r = t / (size_t)n1;
  const int c2 = (int)(r % (size_t)n2);
  const int lb = DIR == 0 ? b.k0 : (DIR == 1 ? b.j0 : b.i0), le = DIR == 0 ? b.k1 : (DIR == 1 ? b.j1 : b.i1);
  const int l0 = lb + (int)(r / (size_t)n2) * L, l1 = min(le, l0 + L - 1);
  const int i = DIR == 2 ? 0 : b.i0 + c1;
  const int j = DIR == 2 ? b.j0 + c1 : (DIR == 0 ? b.j0 + c2 : 0);
  const int k = DIR == 0 ? 0 : b.k0 + c2;
  auto at = [&](int l) { return DIR == 0 ? IX(d, l, i, j) : (DIR == 1 ? IX(d, k, i, l) : IX(d, k, l, j)); };
  auto flush = [&](int l, double va, double vb, bool own) {
    if (va != 0.0) {
      double* q = tgt + at(l);
      if (own) *q += va;
      else aadd(q, va);
    }
    if (vb != 0.0) aadd(tgt2 + at(l), vb);
  };
  ZWin w;
  for (int l = l0; l <= l1; ++l) {
    if constexpr (DIR == 0) f(l, i, j, w);
    else if constexpr (DIR == 1) f(k, i, l, w);
    else f(k, l, j, w);
    flush(l - 1, w.a[0], w.b[0], l - 1 >= l0 + 2);  // no later face reaches l-1
    w.a[0] = w.a[1]; w.a[1] = w.a[2]; w.a[2] = w.a[3]; w.a[3] = 0.0;
    w.b[0] = w.b[1]; w.b[1] = w.b[2]; w.b[2] = w.b[3]; w.b[3] = 0.0;
  }
  flush(l1, w.a[0], w.b[0], false);
  flush(l1 + 1, w.a[1], w.b[1], false);
  flush(l1 + 2, w.a[2], w.b[2], false);
}

// Faces of a box along x, one face per thread as for3 maps them (i fastest):
// f(k, i, j, win) returns the face's contributions to the points p-1..p+2 of
// its x row (p = IX(k, i, j)) in win.v[t][0..3], for the target tg[t] (a field,
// or a field shifted to another row: Db + IA).  Neighbouring lanes of a wave
// hold neighbouring faces of a row, so each lane sums what its wave's faces
// give its own point (shuffles) and adds that with one atomic; contributions
// whose point no lane of the wave owns (row ends, wave ends) are added by
// their own lane.
template <int NT>
struct XWin {
  double v[NT][4];
};
template <int NT>
struct XTgt {
  double* p[NT];
};

template <int NT, typename F>
__global__ void __launch_bounds__(256) xrow_kernel(Dims d, Box b, XTgt<NT> tg, F f) {
  const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
  const size_t n = (size_t)ni * (size_t)nj * (size_t)nk;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  XWin<NT> w;
#pragma unroll
  for (int q = 0; q < NT; ++q) w.v[q][0] = w.v[q][1] = w.v[q][2] = w.v[q][3] = 0.0;
  long long p = -8;  // no point: never p-1 or p-2 of another lane's
  if (t < n) {
    const int i = b.i0 + (int)(t % (size_t)ni);
    const size_t r = t / (size_t)ni;
    const int j = b.j0 + (int)(r % (size_t)nj);
    const int k = b.k0 + (int)(r / (size_t)nj);
    p = (long long)IX(d, k, i, j);
    f(k, i, j, w);
  }
  // every lane of the wave takes part in the shuffles (a lane with no
  // neighbour in that direction reads its own value, which fails the test)
  const long long pu1 = __shfl_up(p, 1), pu2 = __shfl_up(p, 2), pd1 = __shfl_down(p, 1), pd2 = __shfl_down(p, 2);
  const bool u1 = pu1 == p - 1, u2 = u1 && pu2 == p - 2, d1 = pd1 == p + 1, d2 = d1 && pd2 == p + 2;
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    const double from_d1 = __shfl_down(w.v[q][0], 1), from_u1 = __shfl_up(w.v[q][2], 1), from_u2 = __shfl_up(w.v[q][3], 2);
    if (p < 0) continue;
    double* base = tg.p[q] + p;
    const double own = w.v[q][1] + (d1 ? from_d1 : 0.0) + (u1 ? from_u1 : 0.0) + (u2 ? from_u2 : 0.0);
    if (own != 0.0) aadd(base, own);
    if (!u1 && w.v[q][0] != 0.0) aadd(base - 1, w.v[q][0]);
    if (!d1 && w.v[q][2] != 0.0) aadd(base + 1, w.v[q][2]);
    if (!d2 && w.v[q][3] != 0.0) aadd(base + 2, w.v[q][3]);
  }
}

template <int NT, typename F>
hipError_t xrow(const Dims& d, Box b, XTgt<NT> tg, F f, hipStream_t s) {
  if (b.k1 < b.k0 || b.i1 < b.i0 || b.j1 < b.j0) return hipSuccess;
  const size_t n = (size_t)(b.i1 - b.i0 + 1) * (size_t)(b.j1 - b.j0 + 1) * (size_t)(b.k1 - b.k0 + 1);
  hipLaunchKernelGGL((xrow_kernel<NT, F>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, b, tg, f);
  return hipGetLastError();
}

// Points of a box along z or y (DIR 0, 1), L consecutive points of a column
// (a y row) per thread, i fastest across threads: f(k, i, j, win) adds its
// contributions to the points l-1..l+2 of the line in win.v[t][0..3] for the
// target tg[t] (a field, or a field shifted by whole levels, rows or
// columns); each point of the line is added once per thread and target, with
// an atomic (the shifted targets of other threads reach the same points),
// except, with OWN0, target 0's points that no other thread reaches (every
// face q-2..q+1 of point q the thread's own).
template <int DIR, int NT, bool OWN0, typename F>
__global__ void __launch_bounds__(256) lsegn_kernel(Dims d, Box b, int L, XTgt<NT> tg, F f) {
  const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
  const int n2 = DIR == 0 ? nj : nk, nl = DIR == 0 ? nk : nj;
  const int nseg = (nl + L - 1) / L;
  const size_t n = (size_t)ni * (size_t)n2 * (size_t)nseg;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = b.i0 + (int)(t % (size_t)ni);
  const size_t r = t / (size_t)ni;
  const int c2 = (int)(r % (size_t)n2);
  const int lb = DIR == 0 ? b.k0 : b.j0, le = DIR == 0 ? b.k1 : b.j1;
  const int l0 = lb + (int)(r / (size_t)n2) * L, l1 = min(le, l0 + L - 1);
  const int j = DIR == 0 ? b.j0 + c2 : 0, k = DIR == 0 ? 0 : b.k0 + c2;
  XWin<NT> w;
#pragma unroll
  for (int q = 0; q < NT; ++q) w.v[q][0] = w.v[q][1] = w.v[q][2] = w.v[q][3] = 0.0;
  auto flush = [&](int l, int m, bool own) {
    const size_t p = DIR == 0 ? IX(d, l, i, j) : IX(d, k, i, l);
#pragma unroll
    for (int q = 0; q < NT; ++q) {
      const double v = w.v[q][m];
      if (v == 0.0) continue;
      if (OWN0 && q == 0 && own) tg.p[0][p] += v;
      else aadd(tg.p[q] + p, v);
    }
  };
  for (int l = l0; l <= l1; ++l) {
    if constexpr (DIR == 0) f(l, i, j, w);
    else f(k, i, l, w);
    flush(l - 1, 0, l - 1 >= l0 + 2);  // no later face reaches l-1
#pragma unroll
    for (int q = 0; q < NT; ++q) {
      w.v[q][0] = w.v[q][1]; w.v[q][1] = w.v[q][2]; w.v[q][2] = w.v[q][3]; w.v[q][3] = 0.0;
    }
  }
  flush(l1, 0, false);
  flush(l1 + 1, 1, false);
  flush(l1 + 2, 2, false);
}

constexpr int ZSEG_L = 8;

template <int DIR, typename F>
hipError_t lseg(const Dims& d, Box b, double* tgt, double* tgt2, F f, hipStream_t s) {
  if (b.k1 < b.k0 || b.i1 < b.i0 || b.j1 < b.j0) return hipSuccess;
  const size_t ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
  const size_t nl = DIR == 0 ? nk : (DIR == 1 ? nj : ni);
  const size_t n = ni * nj * nk / nl * ((nl + ZSEG_L - 1) / ZSEG_L);
  hipLaunchKernelGGL((lseg_kernel<DIR, F>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, b, ZSEG_L, tgt, tgt2, f);
  return hipGetLastError();
}
template <int DIR, int NT, bool OWN0, typename F>
hipError_t lsegn(const Dims& d, Box b, XTgt<NT> tg, F f, hipStream_t s) {
  if (b.k1 < b.k0 || b.i1 < b.i0 || b.j1 < b.j0) return hipSuccess;
  const size_t nl = DIR == 0 ? b.k1 - b.k0 + 1 : b.j1 - b.j0 + 1, n2 = DIR == 0 ? b.j1 - b.j0 + 1 : b.k1 - b.k0 + 1;
  const size_t n = (size_t)(b.i1 - b.i0 + 1) * n2 * ((nl + ZSEG_L - 1) / ZSEG_L);
  hipLaunchKernelGGL((lsegn_kernel<DIR, NT, OWN0, F>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, b, ZSEG_L, tg,
                     f);
  return hipGetLastError();
}
template <typename F>
hipError_t zseg(const Dims& d, Box b, double* tgt, F f, hipStream_t s) { return lseg<0>(d, b, tgt, nullptr, f, s); }

__device__ __forceinline__ int zkind(int k, int KS, int KE) { return k == KS ? 1 : (k == KE - 1 ? 2 : 0); }

}  // namespace

// ---------------------------------------------------------------------------
// work arrays and checkpoints of one context
enum {
  W_O = 0,                 // 5 traced stage outputs
  W_PT = 5, W_ST, W_CP, W_POTT,
  W_POTTB, W_STB, W_SRB, W_SWB, W_DPRESB,
  W_MZ, W_MX, W_MY,
  W_CPB, W_PTB, W_AB, W_CB, W_CPV, W_DPV,
  W_FLD, W_W0, W_W1, W_W2, W_WB0, W_WB1, W_WB2, W_FB,  // W_WB0..W_FB adjacent: one fill clears them
  NW
};

struct AdjWork {
  int nsteps = 0, every = 0, recorded = 0, tape_limit = -1, share = 1;
  // the checkpointed forward ran its last `taped` segments with records: the
  // reverse sweep takes them without a recording forward
  int taped = 0;
  // tapes: one record per step (45 fields) and the step's state (5), for the
  // last `tseg` segments of the run, as many as the tape budget holds (at
  // least one segment, the work area of the recording forward of the others);
  // with no room for one segment's records each reversed step recomputes its
  // own instead
  int tseg = 0;
  std::vector<StepRec> recs;  // tseg x every records, segment by segment
  std::vector<double*> ckpt;  // states at steps 0, every, 2*every, ... (5 arrays each)
  std::vector<double*> seg;   // states of the taped steps (max(tseg, 1) x every x 5)
  double *xb[5] = {}, *S0b[5] = {}, *Rb[4][5] = {};
  double *NDb = nullptr, *Pb[3] = {};
  double *w[NW] = {};
  double *mf = nullptr, *tf = nullptr;  // traced stage fluxes (3 n3 each)
  double* io[5] = {};                   // adjoint arguments of the C ABI
  double* tr[4][3] = {};                // PT, St, C of the four stages of the step being reversed
  std::vector<void*> allocs;
};

double* adj_io(Ctx& c, int v) { return c.adj ? c.adj->io[v] : nullptr; }

void adj_stage_bufs(Ctx& c, double* s0b[5], double* sb[5], double** NDb, double* Pb[3]) {
  AdjWork& A = *c.adj;
  for (int v = 0; v < 5; ++v) { s0b[v] = A.S0b[v]; sb[v] = A.Rb[0][v]; }
  *NDb = A.NDb;
  for (int q = 0; q < 3; ++q) Pb[q] = A.Pb[q];
}

void adj_free(Ctx& c) {
  if (!c.adj) return;
  for (void* p : c.adj->allocs) (void)hipFree(p);
  delete c.adj;
  c.adj = nullptr;
}

static double* aalloc(AdjWork& A, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
  A.allocs.push_back(p);
  return (double*)p;
}

// allocate the stage/filter work (nsteps = 0) or, in addition, the
// checkpoint storage of a run of nsteps large steps saved every 'every' steps
hipError_t adj_setup(Ctx& c, int nsteps, int every, int share) {
  if (!c.adj) {
    c.adj = new AdjWork();
    AdjWork& A = *c.adj;
    const size_t n3 = c.d.n3;
    bool ok = true;
    // the work fields in one block (W order), and S0b, Rb, Pb in another,
    // so that each group is cleared by one fill
    double* wb = aalloc(A, NW * n3);
    double* bb = aalloc(A, 28 * n3);
    ok = wb && bb;
    for (int q = 0; q < NW && ok; ++q) A.w[q] = wb + n3 * q;
    for (int v = 0; v < 5 && ok; ++v) {
      A.S0b[v] = bb + n3 * v;
      for (int r = 0; r < 4; ++r) A.Rb[r][v] = bb + n3 * (5 + 5 * r + v);
      ok = (A.xb[v] = aalloc(A, n3)) && (A.io[v] = aalloc(A, n3));
    }
    for (int q = 0; q < 3 && ok; ++q) A.Pb[q] = bb + n3 * (25 + q);
    for (int r = 0; r < 4 && ok; ++r)
      for (int q = 0; q < 3 && ok; ++q) ok = (A.tr[r][q] = aalloc(A, n3)) != nullptr;
    ok = ok && (A.NDb = aalloc(A, 15 * n3)) && (A.mf = aalloc(A, 3 * n3)) && (A.tf = aalloc(A, 3 * n3));
    if (!ok) { adj_free(c); return hipErrorOutOfMemory; }
  }
  AdjWork& A = *c.adj;
  if (nsteps > 0 && (A.nsteps != nsteps || A.every != every || A.tape_limit != c.adj_tape || A.share != share)) {
    const size_t n3 = c.d.n3;
    const int nck = (nsteps + every - 1) / every;
    // drop the storage of a previous run (kept in allocs until adj_free) and size anew
    A.ckpt.clear(); A.seg.clear(); A.recs.clear(); A.tseg = 0;
    for (int q = 0; q < 5 * nck; ++q) {
      double* p = aalloc(A, n3);
      if (!p) return hipErrorOutOfMemory;
      A.ckpt.push_back(p);
    }
    // the tape budget: c.adj_tape segments, or (0) as many as this model's
    // share of half the free device memory holds; SCALE_HIP_ADJ_RECORDS=0
    // keeps no records at all (each reversed step recomputes its own)
    const char* env = getenv("SCALE_HIP_ADJ_RECORDS");
    const bool records = !env || atoi(env) != 0;
    const size_t per_seg = (size_t)every * 50 * n3 * sizeof(double);
    int want = c.adj_tape > 0 ? std::min(c.adj_tape, nck) : nck;
    if (c.adj_tape == 0) {
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess)
        want = (int)std::max<size_t>(1, std::min<size_t>(nck, fr / 2 / (size_t)std::max(share, 1) / per_seg));
    }
    if (!records) want = 0;
    for (int sg = 0; sg < want; ++sg) {
      std::vector<double*> sts;
      std::vector<StepRec> recs(every);
      bool room = true;
      for (int q = 0; q < 5 * every && room; ++q) {
        double* p = aalloc(A, n3);
        room = p != nullptr;
        if (room) sts.push_back(p);
      }
      for (int e = 0; e < every && room; ++e) {
        StepRec& r = recs[e];
        for (int q = 0; q < 3 && room; ++q)
          for (int v = 0; v < 5 && room; ++v) room = (r.R[q][v] = aalloc(A, n3)) != nullptr;
        room = room && (r.nd = aalloc(A, 15 * n3)) && (r.RT2P = aalloc(A, n3)) && (r.DPRES0 = aalloc(A, n3)) &&
               (r.REF_rhot = aalloc(A, n3));
        for (int q = 0; q < 4 && room; ++q)
          for (int t = 0; t < 3 && room; ++t) room = (r.tr[q][t] = aalloc(A, n3)) != nullptr;
      }
      if (!room) break;  // a failed allocation only means: no more tapes
      A.seg.insert(A.seg.end(), sts.begin(), sts.end());
      A.recs.insert(A.recs.end(), recs.begin(), recs.end());
      ++A.tseg;
    }
    (void)hipGetLastError();
    if (A.tseg == 0) {  // the per-step recomputation needs one segment's states
      A.seg.clear();
      for (int q = 0; q < 5 * every; ++q) {
        double* p = aalloc(A, n3);
        if (!p) return hipErrorOutOfMemory;
        A.seg.push_back(p);
      }
    }
    A.nsteps = nsteps; A.every = every; A.tape_limit = c.adj_tape; A.share = share;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Stage adjoint (reverse of HEVI.F90:128-1089).  TER = false is the flat
// grid: the metric accessors return the constants and the terrain-only loop
// nests (the J13/J23 fluxes) drop out.

namespace {
// F2H(k, pp) of a column with the GSQRT component cG (UD3.F90:75 and the
// F2H of HEVI.F90): CDZ(k+pp-1) G(k+pp-1) / (CDZ(k) G(k) + CDZ(k+1) G(k+1))
template <bool TER>
__device__ __forceinline__ double f2h(const Met<TER>& m, const Grid1D& g, const Dims& d, int cG, int k, int pp, int i, int j) {
  const size_t p = IX(d, k, i, j);
  return g.CDZ[k + pp - 1] * m.gs(cG, p + d.pl * (size_t)(pp - 1)) / (g.CDZ[k] * m.gs(cG, p) + g.CDZ[k + 1] * m.gs(cG, p + d.pl));
}

// the vertical mass flux before the solve at (k, i, j) (HEVI.F90:369-392)
template <bool TER>
__device__ __forceinline__ double mflx_z0(const Met<TER>& m, const Dims& d, const double* MX, const double* MY, const double* ND,
                                          int k, int i, int j) {
  const size_t p = IX(d, k, i, j), pl = d.pl, IA = d.IA;
  const double M0 = m.mapf(i, j, 0, I_XY), M1 = m.mapf(i, j, 1, I_XY);
  const double gz = m.gs(I_XYW, p) / (M0 * M1) * ND[p];
  if constexpr (TER)
    return m.j13(I_XYW, p) * 0.25 / M1 * (MX[p + pl] + MX[p + pl - 1] + MX[p] + MX[p - 1])
         + m.j23(I_XYW, p) * 0.25 / M0 * (MY[p + pl] + MY[p + pl - IA] + MY[p] + MY[p - IA]) + gz;
  else return gz;
}

// The adjoint of a flux face from the tendency adjoints of the two cells it
// separates, the transpose of each cell's difference of its faces
// (tendency(k) = -(F(k) - F(k-1)) R(k) ... / GSQRT(k)): a face gathers
// -u(lo) R(lo) + u(hi) R(hi) instead of each cell scattering to its faces.
// u is the cell's tendency adjoint (scale ob) over GSQRT, zero outside the
// cells [k0, k1] x [IS, IE] x [JS, JE]; the horizontal faces (and, with
// mm_all, the vertical ones) carry MAPF(1) MAPF(2) of component cM.
template <bool TER>
struct FaceAdj {
  Met<TER> m;
  Dims d;
  const double* ob;
  double scale;
  int cG, cM, k0, k1;
  bool mm_all;
  const double *RZ, *RX, *RY;
  __device__ __forceinline__ double cell(int k, int i, int j, bool horiz) const {
    if (k < k0 || k > k1 || i < d.IS || i > d.IE || j < d.JS || j > d.JE) return 0.0;
    const size_t p = IX(d, k, i, j);
    double u = scale * ob[p] / m.gs(cG, p);
    if (horiz || mm_all) u = u * m.mapf(i, j, 0, cM) * m.mapf(i, j, 1, cM);
    return u;
  }
  // the z face between levels k and k+1, the x face between i and i+1, the y face between j and j+1
  __device__ __forceinline__ double z(int k, int i, int j) const {
    return -cell(k, i, j, false) * RZ[k] + cell(k + 1, i, j, false) * RZ[k + 1];
  }
  __device__ __forceinline__ double x(int k, int i, int j) const {
    return -cell(k, i, j, true) * RX[i] + cell(k, i + 1, j, true) * RX[i + 1];
  }
  __device__ __forceinline__ double y(int k, int i, int j) const {
    return -cell(k, i, j, true) * RY[j] + cell(k, i, j + 1, true) * RY[j + 1];
  }
};
}  // namespace

template <int S, bool TER>
static hipError_t stage_adj_t(Ctx& c, const StageAdjIO& io, hipStream_t st) {
  const Dims d = c.d;
  const Grid1D g = c.g1;
  AdjWork& A = *c.adj;
  double* const* w = A.w;
  const size_t n3 = d.n3, pl = d.pl;
  const int IA = d.IA, KS = d.KS, KE = d.KE, IS = d.IS, IE = d.IE, JS = d.JS, JE = d.JE;
  const double dt = io.dtrk, J33G = c.J33G;
  const Met<TER> m{c.GSQRT, c.J13G, c.J23G, c.MAPF, n3, pl, IA};
  const double* PHI = c.PHI;

  // forward stage with PT, St and C recorded (unless the caller has them)
  if (!io.trPT) {
    StageIO f{};
    f.ctx = &c;
    for (int v = 0; v < 5; ++v) { f.o[v] = w[W_O + v]; f.s0[v] = io.s0[v]; f.s[v] = io.s[v]; f.t[v] = io.t[v]; }
    f.mflx_hi = A.mf; f.tflx_hi = A.tf;
    f.DPRES0 = io.DPRES0; f.RT2P = io.RT2P; f.CORIOLI = io.CORIOLI; f.num_diff = io.num_diff;
    f.wdamp_coef = io.wdamp_coef; f.DDIV = nullptr; f.REF_dens = io.REF_dens; f.REF_rhot = io.REF_rhot;
    f.divdmp_coef = 0.0; f.dtrk = dt; f.has_coriolis = io.has_coriolis; f.combine = 0;
    HIPCHK(hevi_stage_traced(f, w[W_PT], w[W_ST], w[W_CP], st));
  }
  HIPCHK(hipMemsetAsync(w[W_POTTB], 0, (W_MY - W_POTTB + 1) * n3 * sizeof(double), st));

  const double *D = io.s[0], *MZ = io.s[1], *MX = io.s[2], *MY = io.s[3], *RT = io.s[4];
  const double *ND = io.num_diff, *COR = io.CORIOLI, *RT2P = io.RT2P, *REFrt = io.REF_rhot, *wd = io.wdamp_coef;
  const double *PT = io.trPT ? io.trPT : w[W_PT], *St = io.trSt ? io.trSt : w[W_ST], *Cp = io.trCp ? io.trCp : w[W_CP];
  double *POTT = w[W_POTT];
  double *POTTb = w[W_POTTB], *Stb = w[W_STB], *Srb = w[W_SRB], *Swb = w[W_SWB], *DPRESb = w[W_DPRESB];
  double *mZb = w[W_MZ], *mXb = w[W_MX], *mYb = w[W_MY];
  double *Cpb = w[W_CPB], *PTb = w[W_PTB], *Cb = w[W_CB], *cpv = w[W_CPV], *dpv = w[W_DPV];
  double *Db = io.sb[0], *MZb = io.sb[1], *MXb = io.sb[2], *MYb = io.sb[3], *RTb = io.sb[4];
  double *D0b = io.s0b[0], *MZ0b = io.s0b[1], *MX0b = io.s0b[2], *MY0b = io.s0b[3], *RT0b = io.s0b[4];
  double *NDb = io.NDb, *DPRES0b = io.DPRES0b, *RT2Pb = io.RT2Pb, *REFrtb = io.REFrtb;
  const double *obD = io.ob[0], *obMZ = io.ob[1], *obMX = io.ob[2], *obMY = io.ob[3], *obRT = io.ob[4];
  const int cor = io.has_coriolis;
  auto nd = [n3](int v, int dir) { return n3 * (size_t)(v + 5 * dir); };
  const size_t oDZ = nd(I_DENS, ZDIR), oDX = nd(I_DENS, XDIR), oDY = nd(I_DENS, YDIR);
  const size_t oWZ = nd(I_MOMZ, ZDIR), oWX = nd(I_MOMZ, XDIR), oWY = nd(I_MOMZ, YDIR);
  const size_t oUZ = nd(I_MOMX, ZDIR), oUX = nd(I_MOMX, XDIR), oUY = nd(I_MOMX, YDIR);
  const size_t oVZ = nd(I_MOMY, ZDIR), oVX = nd(I_MOMY, XDIR), oVY = nd(I_MOMY, YDIR);
  const size_t oTZ = nd(I_RHOT, ZDIR), oTX = nd(I_RHOT, XDIR), oTY = nd(I_RHOT, YDIR);
  const Box inter{KS, KE, IS, IE, JS, JE};
  // the face adjoints of MOMY, MOMX, MOMZ (Sw), RHOT (St) and DENS (Sr)
  const FaceAdj<TER> fV{m, d, obMY, dt, I_XVZ, I_XV, KS, KE, false, g.RCDZ, g.RCDX, g.RFDY};
  const FaceAdj<TER> fU{m, d, obMX, dt, I_UYZ, I_UY, KS, KE, false, g.RCDZ, g.RFDX, g.RCDY};
  const FaceAdj<TER> fW{m, d, Swb, 1.0, I_XYW, I_XY, KS, KE - 1, false, g.RFDZ, g.RCDX, g.RCDY};
  const FaceAdj<TER> fT{m, d, Stb, 1.0, I_XYZ, I_XY, KS, KE, true, g.RCDZ, g.RCDX, g.RCDY};
  const FaceAdj<TER> fR{m, d, Srb, 1.0, I_XYZ, I_XY, KS, KE, true, g.RCDZ, g.RCDX, g.RCDY};

  // POTT = RHOT / DENS (HEVI.F90:588-598)
  HIPCHK(for3(Box{KS, KE, IS - 2, IE + 2, JS - 2, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    POTT[p] = RT[p] / D[p];
  }, st));

  // The J13/J23 fluxes of the horizontal momenta (UD3.F90:849-1039,
  // 1269-1459), terrain only: flux(k) = J vel / mp fv(val, J vel) with
  // vel = (f1 mom(k+1) + f2 mom(k)) / (f1 <D>(k+1) + f2 <D>(k)); mom is the
  // 4-point mean (quad) where it lives at the other horizontal staggering.
  // stag 0: the UYZ family (D averaged over i, i+1), 1: XVZ (over j, j+1).
  auto fluxj_uvz = [&](auto quadc, int stag, int jd, const double* val, double* valb, const FaceAdj<TER> fa) -> hipError_t {
    constexpr bool quad = decltype(quadc)::value;  // stag != jd
    const double* mom = jd == 0 ? MX : MY;
    double* momb = jd == 0 ? MXb : MYb;
    const ptrdiff_t sd = stag == 0 ? 1 : IA;
    const ptrdiff_t oa = stag == 0 ? 1 : -1, ob = stag == 0 ? -(ptrdiff_t)IA : (ptrdiff_t)IA;
    const int cG = stag == 0 ? I_UYZ : I_XVZ, cJ = stag == 0 ? I_UYW : I_XVW, cM = stag == 0 ? I_UY : I_XV;
    const int mn = jd == 0 ? 1 : 0;
    // z faces of a column in the segment's windows: the val adjoint (and,
    // with mom the same field, the mom adjoint of the own column), the
    // density adjoint of the column and its neighbour, and with quad the
    // mom adjoints of the four columns of the mean
    constexpr int NT = quad ? 7 : 3;
    XTgt<NT> tg;
    {
      double* all[7] = {valb, Db, Db + sd, momb, momb + oa, momb + ob, momb + oa + ob};
      for (int q = 0; q < NT; ++q) tg.p[q] = all[q];
    }
    return lsegn<0, NT, true>(d, Box{KS, KE - 1, IS, IE, JS, JE}, tg, [=] __device__(int k, int i, int j, XWin<NT>& zw) {
      const size_t p = IX(d, k, i, j);
      const double f1 = f2h(m, g, d, cG, k, 1, i, j), f2 = f2h(m, g, d, cG, k, 2, i, j);
      auto mm = [&](size_t q) {
        const double* r = mom + q;
        if constexpr (quad) return 0.25 * (r[0] + r[oa] + r[ob] + r[oa + ob]);
        else return r[0];
      };
      const double num = f1 * mm(p + pl) + f2 * mm(p);
      const double den = f1 * 0.5 * (D[p + pl] + D[p + pl + sd]) + f2 * 0.5 * (D[p] + D[p + sd]);
      const double vr = num / den;
      const double Jm = jd == 0 ? m.j13(cJ, p) : m.j23(cJ, p);
      const double vel = vr * Jm, mp = m.mapf(i, j, mn, cM);
      const double fb = fa.z(k, i, j);
      const double velb =
          fb * face_adj_s<S>(zkind(k, KS, KE), val, p, pl, vel, fb * vel / mp, [&](int q, double v) { zw.v[0][q] += v; }) / mp;
      const double vrb = velb * Jm;
      const double numb = vrb / den, denb = -vrb * vr / den;
      if constexpr (quad) {
#pragma unroll
        for (int q = 3; q < 7; ++q) { zw.v[q][2] += 0.25 * (f1 * numb); zw.v[q][1] += 0.25 * (f2 * numb); }
      } else {  // mom is val itself, on the own column
        zw.v[0][2] += f1 * numb; zw.v[0][1] += f2 * numb;
      }
      zw.v[1][2] += 0.5 * f1 * denb; zw.v[2][2] += 0.5 * f1 * denb;
      zw.v[1][1] += 0.5 * f2 * denb; zw.v[2][1] += 0.5 * f2 * denb;
    }, st);
  };

  // ---- momentum y (HEVI.F90:936-1047): the pressure-gradient, Coriolis and
  //      map-factor adjoints of a y row in windows (DPRES at levels k, k+1
  //      and k-1; MOMX at i and i-1) ----
  {
    constexpr int NY = TER ? 6 : 3;
    XTgt<NY> ty;
    {
      double* all[6] = {DPRESb, MXb, MXb - 1, Db, DPRESb + pl, DPRESb - pl};
      for (int q = 0; q < NY; ++q) ty.p[q] = all[q];
    }
    HIPCHK((lsegn<1, NY, false>)(d, inter, ty, [=] __device__(int k, int i, int j, XWin<NY>& yw) {
      const size_t p = IX(d, k, i, j);
      const double b = obMY[p];
      MY0b[p] += b;
      const double T = dt * b;
      const double u = T / m.gs(I_XVZ, p);
      // (the flux faces gather their adjoints from obMY: fV)
      // pg: the bracket times MAPF(i,j,2,I_XV)
      const double v = -u * m.mapf(i, j, 1, I_XV);
      yw.v[0][2] += v * g.RFDY[j] * m.gs(I_XYZ, p + IA); yw.v[0][1] += -v * g.RFDY[j] * m.gs(I_XYZ, p);
      if constexpr (TER) {
        const double wv = v * g.RCDZ[k] * 0.5;
        const double a1 = wv * m.j23(I_XVW, p), a2 = wv * m.j23(I_XVW, p - pl);
        const double f1k = f2h(m, g, d, I_XVZ, k, 1, i, j), f2k = f2h(m, g, d, I_XVZ, k, 2, i, j);
        const double f1m = f2h(m, g, d, I_XVZ, k - 1, 1, i, j), f2m = f2h(m, g, d, I_XVZ, k - 1, 2, i, j);
        yw.v[4][2] += a1 * f1k; yw.v[4][1] += a1 * f1k;
        yw.v[0][2] += a1 * f2k - a2 * f1m; yw.v[0][1] += a1 * f2k - a2 * f1m;
        yw.v[5][2] += -a2 * f2m; yw.v[5][1] += -a2 * f2m;
      }
      if (cor) {
        const double cc = -0.125 * (COR[(size_t)i + (size_t)IA * (j + 1)] + COR[(size_t)i + (size_t)IA * j]) * T;
        yw.v[1][2] += cc; yw.v[1][1] += cc; yw.v[2][2] += cc; yw.v[2][1] += cc;
      }
      if constexpr (TER) {
        // the map-factor term of cf, -c0 Xs (MOMY a - Xs/4 b) 2/(D(j+1) + D(j));
        // zero where the map factors are uniform
        const double aa = (1.0 / m.mapf(i, j, 1, I_UV) - 1.0 / m.mapf(i - 1, j, 1, I_UV)) * g.RCDX[i];
        const double bb = (1.0 / m.mapf(i, j + 1, 0, I_XY) - 1.0 / m.mapf(i, j, 0, I_XY)) * g.RFDY[j];
        if (aa != 0.0 || bb != 0.0) {
          const double Xs = MX[p] + MX[p - 1] + MX[p + IA] + MX[p - 1 + IA];
          const double c0 = 0.25 * m.mapf(i, j, 0, I_XV) * m.mapf(i, j, 1, I_XV);
          const double Dd = D[p + IA] + D[p], rD = 2.0 / Dd;
          const double xsb = -T * c0 * (MY[p] * aa - 0.5 * Xs * bb) * rD;
          yw.v[1][1] += xsb; yw.v[2][1] += xsb; yw.v[1][2] += xsb; yw.v[2][2] += xsb;
          MYb[p] += -T * c0 * Xs * aa * rD;  // the point's own: no other thread adds to it here
          const double cf2 = -c0 * Xs * (MY[p] * aa - 0.25 * Xs * bb) * rD;
          yw.v[3][2] += -T * cf2 / Dd; yw.v[3][1] += -T * cf2 / Dd;
        }
      }
    }, st));
  }
  // fluxZ_XVZ (UD3.F90:1173-1265), the MOMY adjoint of the column in a window
  HIPCHK(zseg(d, Box{KS, KE - 1, IS, IE, JS, JE}, MYb, [=] __device__(int k, int i, int j, ZWin& zw) {
    const size_t p = IX(d, k, i, j);
    const double f1 = f2h(m, g, d, I_XVW, k, 1, i, j), f2 = f2h(m, g, d, I_XVW, k, 2, i, j);
    const double den = f1 * 0.5 * (D[p + pl] + D[p + pl + IA]) + f2 * 0.5 * (D[p] + D[p + IA]);
    const double vel = 0.5 * (MZ[p] + MZ[p + IA]) / den;
    const double fb = fV.z(k, i, j);
    NDb[oVZ + p] += fb * m.gs(I_XVW, p);
    const double velb = J33G * fb * face_adj_s<S>(zkind(k, KS, KE), MY, p, pl, vel, J33G * fb * vel, [&](int q, double v) { zw.a[q] += v; });
    aadd(MZb + p, 0.5 * velb / den); aadd(MZb + p + IA, 0.5 * velb / den);
    const double denb = -velb * vel / den;
    aadd(Db + p + pl, 0.5 * f1 * denb); aadd(Db + p + pl + IA, 0.5 * f1 * denb);
    aadd(Db + p, 0.5 * f2 * denb); aadd(Db + p + IA, 0.5 * f2 * denb);
  }, st));
  if constexpr (TER) {
    HIPCHK(fluxj_uvz(std::true_type{}, 1, 0, MY, MYb, fV));  // fluxJ13_XVZ
    HIPCHK(fluxj_uvz(std::false_type{}, 1, 1, MY, MYb, fV));  // fluxJ23_XVZ
  }
  // fluxX_XVZ (UD3.F90:1463-1512)
  HIPCHK(xrow<3>(d, Box{KS, KE, IS - 1, IE, JS, JE}, XTgt<3>{{MYb, Db, Db + IA}}, [=] __device__(int k, int i, int j, XWin<3>& xw) {
    const size_t p = IX(d, k, i, j);
    const double den = 0.25 * (D[p] + D[p + 1] + D[p + IA] + D[p + 1 + IA]);
    const double vel = 0.5 * (MX[p] + MX[p + IA]) / den;
    const double fb = fV.x(k, i, j);
    const double G = m.gs(I_UVZ, p), gm = G / m.mapf(i, j, 1, I_UV);
    NDb[oVX + p] += fb * G;
    const double velb = fb * gm * face_adj_s<S>(0, MY, p, 1, vel, fb * gm * vel, [&](int c, double v) { xw.v[0][c] += v; });
    aadd(MXb + p, 0.5 * velb / den); aadd(MXb + p + IA, 0.5 * velb / den);
    const double denb = 0.25 * (-velb * vel / den);
    xw.v[1][1] += denb; xw.v[1][2] += denb; xw.v[2][1] += denb; xw.v[2][2] += denb;
  }, st));
  // fluxY_XVZ (UD3.F90:1536-1587): flux(k,i,j-1) at loop index j, here the
  // face jf = j-1; the MOMY and DENS adjoints of the y row in windows
  HIPCHK(lseg<1>(d, Box{KS, KE, IS, IE, JS - 1, JE}, MYb, Db, [=] __device__(int k, int i, int jf, ZWin& yw) {
    const int j = jf + 1;
    const size_t p = IX(d, k, i, j), q = p - IA;
    const double vel = 0.5 * (MY[p] + MY[q]) / D[p];
    const double fb = fV.y(k, i, jf);
    const double G = m.gs(I_XYZ, p), gm = G / m.mapf(i, j, 0, I_XY);
    NDb[oVY + p] += fb * G;
    const double velb = fb * gm * face_adj_s<S>(0, MY, q, IA, vel, fb * gm * vel, [&](int c, double v) { yw.a[c] += v; });
    yw.a[2] += 0.5 * velb / D[p]; yw.a[1] += 0.5 * velb / D[p];
    yw.b[2] += -velb * vel / D[p];
  }, st));

  // ---- momentum x (HEVI.F90:816-927): the pressure-gradient, Coriolis and
  //      map-factor adjoints of an x row summed across the wave (DPRES at
  //      levels k, k+1 and k-1; MOMY at j and j-1) ----
  {
    constexpr int NX = TER ? 6 : 3;
    XTgt<NX> tx;
    {
      double* all[6] = {DPRESb, MYb, MYb - IA, Db, DPRESb + pl, DPRESb - pl};
      for (int q = 0; q < NX; ++q) tx.p[q] = all[q];
    }
    HIPCHK(xrow<NX>(d, inter, tx, [=] __device__(int k, int i, int j, XWin<NX>& xw) {
      const size_t p = IX(d, k, i, j);
      const double b = obMX[p];
      MX0b[p] += b;
      const double T = dt * b;
      const double u = T / m.gs(I_UYZ, p);
      // (the flux faces gather their adjoints from obMX: fU)
      // pg: the bracket times MAPF(i,j,1,I_UY)
      const double v = -u * m.mapf(i, j, 0, I_UY);
      xw.v[0][2] += v * g.RFDX[i] * m.gs(I_XYZ, p + 1); xw.v[0][1] += -v * g.RFDX[i] * m.gs(I_XYZ, p);
      if constexpr (TER) {
        // both J13 terms take F2H at level k, as HEVI.F90 writes them
        const double wv = v * g.RCDZ[k] * 0.5;
        const double a1 = wv * m.j13(I_UYW, p), a2 = wv * m.j13(I_UYW, p - pl);
        const double f1k = f2h(m, g, d, I_UYZ, k, 1, i, j), f2k = f2h(m, g, d, I_UYZ, k, 2, i, j);
        xw.v[4][2] += a1 * f1k; xw.v[4][1] += a1 * f1k;
        xw.v[0][2] += a1 * f2k - a2 * f1k; xw.v[0][1] += a1 * f2k - a2 * f1k;
        xw.v[5][2] += -a2 * f2k; xw.v[5][1] += -a2 * f2k;
      }
      if (cor) {
        const double cc = 0.125 * (COR[(size_t)(i + 1) + (size_t)IA * j] + COR[(size_t)i + (size_t)IA * j]) * T;
        xw.v[1][2] += cc; xw.v[1][1] += cc; xw.v[2][2] += cc; xw.v[2][1] += cc;
      }
      if constexpr (TER) {
        // the map-factor term of cf, c0 Ys (Ys a - MOMX b) 2/(D(i+1) + D(i));
        // zero where the map factors are uniform
        const double aa = 0.25 * (1.0 / m.mapf(i + 1, j, 1, I_XY) - 1.0 / m.mapf(i, j, 1, I_XY)) * g.RFDX[i];
        const double bb = (1.0 / m.mapf(i, j, 0, I_UV) - 1.0 / m.mapf(i, j - 1, 0, I_UV)) * g.RCDY[j];
        if (aa != 0.0 || bb != 0.0) {
          const double Ys = MY[p] + MY[p - IA] + MY[p + 1] + MY[p + 1 - IA];
          const double c0 = 0.25 * m.mapf(i, j, 0, I_UY) * m.mapf(i, j, 1, I_UY);
          const double Dd = D[p + 1] + D[p], rD = 2.0 / Dd;
          const double ysb = T * c0 * (2.0 * Ys * aa - MX[p] * bb) * rD;
          xw.v[1][1] += ysb; xw.v[2][1] += ysb; xw.v[1][2] += ysb; xw.v[2][2] += ysb;
          MXb[p] += -T * c0 * Ys * bb * rD;  // the point's own: no other thread adds to it here
          const double cf2 = c0 * Ys * (Ys * aa - MX[p] * bb) * rD;
          xw.v[3][2] += -T * cf2 / Dd; xw.v[3][1] += -T * cf2 / Dd;
        }
      }
    }, st));
  }
  // fluxZ_UYZ (UD3.F90:753-845), the MOMX adjoint of the column in a window
  HIPCHK(zseg(d, Box{KS, KE - 1, IS, IE, JS, JE}, MXb, [=] __device__(int k, int i, int j, ZWin& zw) {
    const size_t p = IX(d, k, i, j);
    const double f1 = f2h(m, g, d, I_UYW, k, 1, i, j), f2 = f2h(m, g, d, I_UYW, k, 2, i, j);
    const double den = f1 * 0.5 * (D[p + pl] + D[p + pl + 1]) + f2 * 0.5 * (D[p] + D[p + 1]);
    const double vel = 0.5 * (MZ[p] + MZ[p + 1]) / den;
    const double fb = fU.z(k, i, j);
    NDb[oUZ + p] += fb * m.gs(I_UYW, p);
    const double velb = J33G * fb * face_adj_s<S>(zkind(k, KS, KE), MX, p, pl, vel, J33G * fb * vel, [&](int q, double v) { zw.a[q] += v; });
    aadd(MZb + p, 0.5 * velb / den); aadd(MZb + p + 1, 0.5 * velb / den);
    const double denb = -velb * vel / den;
    aadd(Db + p + pl, 0.5 * f1 * denb); aadd(Db + p + pl + 1, 0.5 * f1 * denb);
    aadd(Db + p, 0.5 * f2 * denb); aadd(Db + p + 1, 0.5 * f2 * denb);
  }, st));
  if constexpr (TER) {
    HIPCHK(fluxj_uvz(std::false_type{}, 0, 0, MX, MXb, fU));  // fluxJ13_UYZ
    HIPCHK(fluxj_uvz(std::true_type{}, 0, 1, MX, MXb, fU));  // fluxJ23_UYZ
  }
  // fluxX_UYZ (UD3.F90:1043-1097): flux(k,i-1,j) at loop index i, here the
  // face if = i-1 (q its point, p = q+1 the loop's)
  HIPCHK(xrow<2>(d, Box{KS, KE, IS - 1, IE, JS, JE}, XTgt<2>{{MXb, Db}}, [=] __device__(int k, int fi, int j, XWin<2>& xw) {
    const int i = fi + 1;
    const size_t q = IX(d, k, fi, j), p = q + 1;
    const double vel = 0.5 * (MX[p] + MX[q]) / D[p];
    const double fb = fU.x(k, fi, j);
    const double G = m.gs(I_XYZ, p), gm = G / m.mapf(i, j, 1, I_XY);
    NDb[oUX + p] += fb * G;
    const double velb = fb * gm * face_adj_s<S>(0, MX, q, 1, vel, fb * gm * vel, [&](int c, double v) { xw.v[0][c] += v; });
    xw.v[0][2] += 0.5 * velb / D[p]; xw.v[0][1] += 0.5 * velb / D[p];
    xw.v[1][2] += -velb * vel / D[p];
  }, st));
  // fluxY_UYZ (UD3.F90:1101-1150): the MOMX adjoint of the y row in a window
  HIPCHK(lseg<1>(d, Box{KS, KE, IS, IE, JS - 1, JE}, MXb, Db, [=] __device__(int k, int i, int j, ZWin& yw) {
    const size_t p = IX(d, k, i, j);
    const double den = 0.25 * (D[p] + D[p + 1] + D[p + IA] + D[p + 1 + IA]);
    const double vel = 0.5 * (MY[p] + MY[p + 1]) / den;
    const double fb = fU.y(k, i, j);
    const double G = m.gs(I_UVZ, p), gm = G / m.mapf(i, j, 0, I_UV);
    NDb[oUY + p] += fb * G;
    const double velb = fb * gm * face_adj_s<S>(0, MX, p, IA, vel, fb * gm * vel, [&](int c, double v) { yw.a[c] += v; });
    aadd(MYb + p, 0.5 * velb / den); aadd(MYb + p + 1, 0.5 * velb / den);
    const double denb = 0.25 * (-velb * vel / den);
    yw.b[1] += denb; yw.b[2] += denb;
    aadd(Db + p + 1, denb); aadd(Db + p + 1 + IA, denb);
  }, st));

  // ---- vertical implicit block (HEVI.F90:688-804) ----
  // (a) per level: adjoints of the back substitution's updates of MOMZ, DENS
  //     and RHOT (HEVI.F90:728-790), each level gathering what its own and
  //     the upper level's updates send to it
  HIPCHK(for3(inter, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    const double bd = obD[p], bt = obRT[p];
    D0b[p] += bd; Srb[p] += dt * bd;
    RT0b[p] += bt; Stb[p] += dt * bt;
    double cpb = 0.0, ptb = 0.0;
    if (k <= KE - 1) {
      const double b = obMZ[p];
      MZ0b[p] += b; MZb[p] -= b;
      const double r = dt * J33G * g.RCDZ[k] / m.gs(I_XYZ, p), r1 = dt * J33G * g.RCDZ[k + 1] / m.gs(I_XYZ, p + pl);
      const double bd1 = obD[p + pl], bt1 = obRT[p + pl];
      cpb = b - r * bd - r * PT[p] * bt + r1 * bd1 + r1 * PT[p] * bt1;
      ptb = -r * Cp[p] * bt + r1 * Cp[p] * bt1;
    }
    Cpb[p] = cpb; PTb[p] = ptb;
  }, st));
  // (b) the transposed Thomas solve, one thread per column:
  //     M(k,k-1) = F3(k), M(k,k) = F2(k), M(k,k+1) = F1(k)
  //     (the bands F1(k-1), F2(k), F3(k+1) of level k all read PT(k); A(k)
  //     is carried from the level below.  With few columns this runs about
  //     one wave per CU, so the loads go in chunks of ZCH levels, the next
  //     chunk's in flight during the current chunk's work)
  HIPCHK(for3(Box{0, 0, IS, IE, JS, JE}, [=] __device__(int, int i, int j) {
    constexpr int ZCH = 8;
    const size_t c0 = IX(d, 0, i, j);
    auto at = [&](int k) { return c0 + pl * (size_t)k; };
    auto Af = [&](int k, double rt) { return dt * dt * J33G * g.RCDZ[k] * rt * J33G / m.gs(I_XYZ, at(k)); };
    auto Bf = [&](int k) { return GRAV * (dt * dt) * J33G / (g.CDZ[k + 1] + g.CDZ[k]); };
    auto gw = [&](int k) { return m.gs(I_XYW, at(k)); };
    double ptk = PT[at(KS)];
    double Ak = Af(KS, RT2P[at(KS)]), Ak1 = Af(KS + 1, RT2P[at(KS + 1)]);
    double den = 1.0 + (ptk * g.RFDZ[KS] * (Ak1 + Ak)) / gw(KS);
    double cpk = (KS < KE - 1 ? -(ptk * g.RFDZ[KS + 1] * Ak1 - Bf(KS + 1)) / gw(KS + 1) : 0.0) / den;
    double dpk = Cpb[at(KS)] / den;
    cpv[at(KS)] = cpk; dpv[at(KS)] = dpk;
    // chunk of levels kc..kc+ZCH-1 (up to KE-1): PT(k), RT2P(k+1), Cpb(k)
    double pt[ZCH], rt[ZCH], cb[ZCH], npt[ZCH], nrt[ZCH], ncb[ZCH];
    auto load = [&](int kc, double* P, double* R, double* Cc) {
_Pragma("unroll")
      for (int c = 0; c < ZCH; ++c) {
        const int k = kc + c;
        const bool in = k <= KE - 1;
        P[c] = in ? PT[at(k)] : 0.0;
        R[c] = in ? RT2P[at(k + 1)] : 0.0;
        Cc[c] = in ? Cpb[at(k)] : 0.0;
      }
    };
    load(KS + 1, pt, rt, cb);
    for (int kc = KS + 1; kc <= KE - 1; kc += ZCH) {
      if (kc + ZCH <= KE - 1) load(kc + ZCH, npt, nrt, ncb);
_Pragma("unroll")
      for (int c = 0; c < ZCH; ++c) {
        const int k = kc + c;
        if (k > KE - 1) break;
        ptk = pt[c];
        Ak = Ak1;
        Ak1 = Af(k + 1, rt[c]);
        const double f1m = -(ptk * g.RFDZ[k - 1] * Ak + Bf(k - 1)) / gw(k - 1);
        den = (1.0 + (ptk * g.RFDZ[k] * (Ak1 + Ak)) / gw(k)) - f1m * cpk;
        cpk = (k < KE - 1 ? -(ptk * g.RFDZ[k + 1] * Ak1 - Bf(k + 1)) / gw(k + 1) : 0.0) / den;
        dpk = (cb[c] - f1m * dpk) / den;
        cpv[at(k)] = cpk; dpv[at(k)] = dpk;
      }
_Pragma("unroll")
      for (int c = 0; c < ZCH; ++c) { pt[c] = npt[c]; rt[c] = nrt[c]; cb[c] = ncb[c]; }
    }
    double cn = dpk;
    Cb[at(KE - 1)] = cn;
    // back substitution, chunks of levels kc..kc-ZCH+1 (down to KS)
    double dv[ZCH], cv[ZCH], ndv[ZCH], ncv[ZCH];
    auto loadb = [&](int kc, double* Dv, double* Cv) {
_Pragma("unroll")
      for (int c = 0; c < ZCH; ++c) {
        const int k = kc - c;
        const bool in = k >= KS;
        Dv[c] = in ? dpv[at(k)] : 0.0;
        Cv[c] = in ? cpv[at(k)] : 0.0;
      }
    };
    loadb(KE - 2, dv, cv);
    for (int kc = KE - 2; kc >= KS; kc -= ZCH) {
      if (kc - ZCH >= KS) loadb(kc - ZCH, ndv, ncv);
_Pragma("unroll")
      for (int c = 0; c < ZCH; ++c) {
        const int k = kc - c;
        if (k < KS) break;
        cn = dv[c] - cv[c] * cn;
        Cb[at(k)] = cn;
      }
_Pragma("unroll")
      for (int c = 0; c < ZCH; ++c) { dv[c] = ndv[c]; cv[c] = ncv[c]; }
    }
  }, st));
  // (c) per level: adjoints of the band coefficients (through PT and
  //     A = dt^2 J33G RCDZ RT2P J33G / GSQRT) and of the right-hand side
  //     C = MOMZ + dt (pg + Sw), level L gathering rows L-1, L and L+1
  HIPCHK(for3(inter, [=] __device__(int L, int i, int j) {
    const size_t p = IX(d, L, i, j);
    auto A_ = [&](int k) { const size_t q = IX(d, k, i, j); return dt * dt * J33G * g.RCDZ[k] * RT2P[q] * J33G / m.gs(I_XYZ, q); };
    auto gw = [&](int k) { return m.gs(I_XYW, IX(d, k, i, j)); };
    auto row = [&](int k) { return k >= KS && k <= KE - 1; };
    double ptb = 0.0, ab = 0.0;
    if (row(L)) {
      const double cb = Cb[p];
      const double F2b = -cb * Cp[p] / gw(L);
      ptb += F2b * g.RFDZ[L] * (A_(L + 1) + A_(L));
      ab += F2b * PT[p] * g.RFDZ[L];
      if (L >= KS + 1) ab -= (-cb * Cp[p - pl] / gw(L)) * PT[p - pl] * g.RFDZ[L];  // F3 of row L
    }
    if (row(L - 1)) {
      const double cbm = Cb[p - pl];
      ab += (-cbm * Cp[p - pl] / gw(L - 1)) * PT[p - pl] * g.RFDZ[L - 1];  // F2 of row L-1 through A(L)
      if (L - 1 <= KE - 2) {                                               // F1 of row L-1
        const double F1b = -cbm * Cp[p] / gw(L - 1);
        ptb -= F1b * g.RFDZ[L - 1] * A_(L);
        ab -= F1b * PT[p] * g.RFDZ[L - 1];
      }
    }
    if (row(L + 1) && L + 1 >= KS + 1) ptb -= (-Cb[p + pl] * Cp[p] / gw(L + 1)) * g.RFDZ[L + 1] * A_(L + 1);  // F3 of row L+1
    PTb[p] += ptb;
    double rt2pb = ab * dt * dt * J33G * g.RCDZ[L] * J33G / m.gs(I_XYZ, p);
    // right-hand side: pg = -(DPRES(k+1) + RT2P(k+1) dt St(k+1) - DPRES(k) - RT2P(k) dt St(k)) RFDZ(k) J33G / GSQRT
    //                       - GRAV (f1 (D(k+1) - REFd + dt Sr(k+1)) + f2 (D(k) - REFd + dt Sr(k)))
    double dpb = 0.0, stb = 0.0, db = 0.0, srb = 0.0;
    if (row(L)) {
      const double cb = Cb[p];
      MZb[p] += cb;
      Swb[p] += dt * cb;
      const double a = -dt * cb * g.RFDZ[L] * J33G / gw(L);
      dpb -= a; rt2pb -= a * dt * St[p]; stb -= a * RT2P[p] * dt;
      const double bb = -GRAV * dt * cb;
      const double f2 = f2h(m, g, d, I_XYZ, L, 2, i, j);
      db += bb * f2; srb += bb * f2 * dt;
    }
    if (row(L - 1)) {
      const double cbm = Cb[p - pl];
      const double a = -dt * cbm * g.RFDZ[L - 1] * J33G / gw(L - 1);
      dpb += a; rt2pb += a * dt * St[p]; stb += a * RT2P[p] * dt;
      const double bb = -GRAV * dt * cbm;
      const double f1 = f2h(m, g, d, I_XYZ, L - 1, 1, i, j);
      db += bb * f1; srb += bb * f1 * dt;
    }
    DPRESb[p] += dpb; RT2Pb[p] += rt2pb; Stb[p] += stb; Db[p] += db; Srb[p] += srb;
  }, st));
  // (d) PT = valueW_Z(MOMZ, POTT) (UD3.F90:98-153), the POTT adjoint of the column in a window
  HIPCHK(zseg(d, Box{KS, KE - 1, IS, IE, JS, JE}, POTTb, [=] __device__(int k, int i, int j, ZWin& zw) {
    const size_t p = IX(d, k, i, j);
    (void)face_adj_s<S>(zkind(k, KS, KE), POTT, p, pl, MZ[p], PTb[p], [&](int q, double v) { zw.a[q] += v; });
  }, st));

  // ---- St (HEVI.F90:634-657) and the theta fluxes (UD3.F90:157-354): each
  //      face gathers its adjoint from Stb (fT) and owns its mass-flux adjoint ----
  HIPCHK(zseg(d, Box{KS, KE - 1, IS, IE, JS, JE}, POTTb, [=] __device__(int k, int i, int j, ZWin& zw) {
    const size_t p = IX(d, k, i, j);
    const double fb = fT.z(k, i, j), vel = mflx_z0(m, d, MX, MY, ND + oDZ, k, i, j);
    NDb[oTZ + p] += fb * m.gs(I_XYW, p);
    mZb[p] += fb * face_adj_s<S>(zkind(k, KS, KE), POTT, p, pl, vel, fb * vel, [&](int q, double v) { zw.a[q] += v; });
  }, st));
  HIPCHK(xrow<1>(d, Box{KS, KE, IS - 1, IE, JS, JE}, XTgt<1>{{POTTb}}, [=] __device__(int k, int i, int j, XWin<1>& xw) {
    const size_t p = IX(d, k, i, j);
    const double G = m.gs(I_UYZ, p);
    const double fb = fT.x(k, i, j), vel = G / m.mapf(i, j, 1, I_UY) * (MX[p] + ND[oDX + p]);
    NDb[oTX + p] += fb * G;
    mXb[p] += fb * face_adj_s<S>(0, POTT, p, 1, vel, fb * vel, [&](int c, double v) { xw.v[0][c] += v; });
  }, st));
  HIPCHK(lseg<1>(d, Box{KS, KE, IS, IE, JS - 1, JE}, POTTb, nullptr, [=] __device__(int k, int i, int j, ZWin& yw) {
    const size_t p = IX(d, k, i, j);
    const double G = m.gs(I_XVZ, p);
    const double fb = fT.y(k, i, j), vel = G / m.mapf(i, j, 0, I_XV) * (MY[p] + ND[oDY + p]);
    NDb[oTY + p] += fb * G;
    mYb[p] += fb * face_adj_s<S>(0, POTT, p, IA, vel, fb * vel, [&](int c, double v) { yw.a[c] += v; });
  }, st));

  // ---- Sw (HEVI.F90:538-577) and the MOMZ fluxes (HEVI.F90:486-525): the
  //      faces gather their adjoints from Swb (fW); Sw's damping term ----
  HIPCHK(for3(Box{KS, KE - 1, IS, IE, JS, JE}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    MZ0b[p] += -wd[k] * Swb[p];
  }, st));
  // fluxZ_XYW (UD3.F90:358-452): face m at loop index m+1; the MOMZ adjoint
  // of the column (value and velocity) in a window
  HIPCHK(zseg(d, Box{KS, KE - 2, IS, IE, JS, JE}, MZb, [=] __device__(int f, int i, int j, ZWin& zw) {
    const size_t p = IX(d, f, i, j);
    const double vel = 0.5 * (MZ[p] + MZ[p + pl]) / D[p + pl];
    const double fb = fW.z(f, i, j);
    NDb[oWZ + p + pl] += fb * m.gs(I_XYZ, p + pl);
    const double velb = J33G * fb * face_adj_s<S>(f == KS ? 1 : 0, MZ, p, pl, vel, J33G * fb * vel, [&](int q, double v) { zw.a[q] += v; });
    zw.a[1] += 0.5 * velb / D[p + pl]; zw.a[2] += 0.5 * velb / D[p + pl];
    Db[p + pl] += -velb * vel / D[p + pl];
  }, st));
  // fluxJ13_XYW / fluxJ23_XYW (UD3.F90:457-593), terrain only: face f at
  // loop index f+1, the face KS with the mean velocity of levels KS, KS+1
  if constexpr (TER) {
    for (int jd = 0; jd < 2; ++jd) {
      const double* mom = jd == 0 ? MX : MY;
      double* momb = jd == 0 ? MXb : MYb;
      const size_t sd = jd == 0 ? 1 : (size_t)IA;  // mom averaged over (i, i-1) or (j, j-1)
      const int mn = jd == 0 ? 1 : 0;
      HIPCHK(zseg(d, Box{KS, KE - 2, IS, IE, JS, JE}, MZb, [=] __device__(int f, int i, int j, ZWin& zw) {
        const size_t p = IX(d, f, i, j), p1 = p + pl;
        const double Jm = jd == 0 ? m.j13(I_XYZ, p1) : m.j23(I_XYZ, p1);
        const double mp = m.mapf(i, j, mn, I_XY);
        const double t1 = 0.5 * (mom[p1] + mom[p1 - sd]) / D[p1];
        const double t0 = f == KS ? 0.5 * (mom[p] + mom[p - sd]) / D[p] : 0.0;
        const double vr = f == KS ? (t1 + t0) * 0.5 : t1;
        const double vel = vr * Jm;
        const double fb = fW.z(f, i, j);
        const double velb =
            fb * face_adj_s<S>(f == KS ? 1 : 0, MZ, p, pl, vel, fb * vel / mp, [&](int q, double v) { zw.a[q] += v; }) / mp;
        const double vrb = velb * Jm;
        const double t1b = f == KS ? 0.5 * vrb : vrb;
        aadd(momb + p1, 0.5 * t1b / D[p1]); aadd(momb + p1 - sd, 0.5 * t1b / D[p1]); Db[p1] += -t1b * t1 / D[p1];
        if (f == KS) {
          const double t0b = 0.5 * vrb;
          aadd(momb + p, 0.5 * t0b / D[p]); aadd(momb + p - sd, 0.5 * t0b / D[p]); Db[p] += -t0b * t0 / D[p];
        }
      }, st));
    }
  }
  // fluxX_XYW (UD3.F90:598-671)
  HIPCHK(xrow<3>(d, Box{KS, KE - 1, IS - 1, IE, JS, JE}, XTgt<3>{{MZb, Db, Db + pl}}, [=] __device__(int k, int i, int j, XWin<3>& xw) {
    const size_t p = IX(d, k, i, j);
    const double f1 = f2h(m, g, d, I_UYW, k, 1, i, j), f2 = f2h(m, g, d, I_UYW, k, 2, i, j);
    const double num = f1 * MX[p + pl] + f2 * MX[p];
    const double den = f1 * 0.5 * (D[p + pl] + D[p + pl + 1]) + f2 * 0.5 * (D[p] + D[p + 1]);
    const double vel = num / den;
    const double fb = fW.x(k, i, j);
    const double G = m.gs(I_UYW, p), gm = G / m.mapf(i, j, 1, I_UY);
    NDb[oWX + p] += fb * G;
    const double velb = fb * gm * face_adj_s<S>(0, MZ, p, 1, vel, fb * gm * vel, [&](int c, double v) { xw.v[0][c] += v; });
    const double numb = velb / den, denb = -velb * vel / den;
    aadd(MXb + p + pl, f1 * numb); aadd(MXb + p, f2 * numb);
    xw.v[2][1] += 0.5 * f1 * denb; xw.v[2][2] += 0.5 * f1 * denb;
    xw.v[1][1] += 0.5 * f2 * denb; xw.v[1][2] += 0.5 * f2 * denb;
  }, st));
  // fluxY_XYW (UD3.F90:675-748): the MOMZ adjoint of the y row in a window
  HIPCHK(lseg<1>(d, Box{KS, KE - 1, IS, IE, JS - 1, JE}, MZb, Db, [=] __device__(int k, int i, int j, ZWin& yw) {
    const size_t p = IX(d, k, i, j);
    const double f1 = f2h(m, g, d, I_XVW, k, 1, i, j), f2 = f2h(m, g, d, I_XVW, k, 2, i, j);
    const double num = f1 * MY[p + pl] + f2 * MY[p];
    const double den = f1 * 0.5 * (D[p + pl] + D[p + pl + IA]) + f2 * 0.5 * (D[p] + D[p + IA]);
    const double vel = num / den;
    const double fb = fW.y(k, i, j);
    const double G = m.gs(I_XVW, p), gm = G / m.mapf(i, j, 0, I_XV);
    NDb[oWY + p] += fb * G;
    const double velb = fb * gm * face_adj_s<S>(0, MZ, p, IA, vel, fb * gm * vel, [&](int c, double v) { yw.a[c] += v; });
    const double numb = velb / den, denb = -velb * vel / den;
    aadd(MYb + p + pl, f1 * numb); aadd(MYb + p, f2 * numb);
    aadd(Db + p + pl, 0.5 * f1 * denb); aadd(Db + p + pl + IA, 0.5 * f1 * denb);
    yw.b[1] += 0.5 * f2 * denb; yw.b[2] += 0.5 * f2 * denb;
  }, st));

  // ---- mass fluxes before the solve (HEVI.F90:369-438): the theta faces'
  //      adjoints plus those gathered from Sr (HEVI.F90:453-475, fR) ----
  //      Flat, each point adds to its own MOMX and MOMY adjoints only; over
  //      terrain the J13/J23 terms add to the neighbours and the level above
  //      too, summed in the column's windows.
  if constexpr (TER) {
    XTgt<4> tm{{MXb, MXb - 1, MYb, MYb - IA}};
    HIPCHK((lsegn<0, 4, false>)(d, Box{KS, KE, IS - 1, IE, JS - 1, JE}, tm, [=] __device__(int k, int i, int j, XWin<4>& zw) {
      const size_t p = IX(d, k, i, j);
      if (j >= JS) {
        if (k <= KE - 1) {
          const double b = mZb[p] + fR.z(k, i, j);
          const double M0 = m.mapf(i, j, 0, I_XY), M1 = m.mapf(i, j, 1, I_XY);
          NDb[oDZ + p] += b * (m.gs(I_XYW, p) / (M0 * M1));
          const double c13 = b * (m.j13(I_XYW, p) * 0.25 / M1), c23 = b * (m.j23(I_XYW, p) * 0.25 / M0);
          zw.v[0][2] += c13; zw.v[1][2] += c13; zw.v[0][1] += c13; zw.v[1][1] += c13;
          zw.v[2][2] += c23; zw.v[3][2] += c23; zw.v[2][1] += c23; zw.v[3][1] += c23;
        }
        const double gm = m.gs(I_UYZ, p) / m.mapf(i, j, 1, I_UY);
        const double bx = mXb[p] + fR.x(k, i, j);
        zw.v[0][1] += bx * gm; NDb[oDX + p] += bx * gm;
      }
      if (i >= IS) {
        const double gm = m.gs(I_XVZ, p) / m.mapf(i, j, 0, I_XV);
        const double by = mYb[p] + fR.y(k, i, j);
        zw.v[2][1] += by * gm; NDb[oDY + p] += by * gm;
      }
    }, st));
  } else {
    HIPCHK(for3(Box{KS, KE, IS - 1, IE, JS - 1, JE}, [=] __device__(int k, int i, int j) {
      const size_t p = IX(d, k, i, j);
      if (j >= JS) {
        if (k <= KE - 1) {
          const double b = mZb[p] + fR.z(k, i, j);
          const double M0 = m.mapf(i, j, 0, I_XY), M1 = m.mapf(i, j, 1, I_XY);
          NDb[oDZ + p] += b * (m.gs(I_XYW, p) / (M0 * M1));
        }
        const double gm = m.gs(I_UYZ, p) / m.mapf(i, j, 1, I_UY);
        const double bx = mXb[p] + fR.x(k, i, j);
        MXb[p] += bx * gm; NDb[oDX + p] += bx * gm;
      }
      if (i >= IS) {
        const double gm = m.gs(I_XVZ, p) / m.mapf(i, j, 0, I_XV);
        const double by = mYb[p] + fR.y(k, i, j);
        MYb[p] += by * gm; NDb[oDY + p] += by * gm;
      }
    }, st));
  }
  // ---- POTT (HEVI.F90:588-598) and DPRES (HEVI.F90:346-360) ----
  HIPCHK(for3(Box{KS, KE, IS - 2, IE + 2, JS - 2, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    RTb[p] += POTTb[p] / D[p];
    Db[p] += -POTTb[p] * POTT[p] / D[p];
    if (i >= IS && i <= IE + 1 && j >= JS && j <= JE + 1) {
      const double b = DPRESb[p];
      DPRES0b[p] += b;
      RT2Pb[p] += b * (RT[p] - REFrt[p]);
      RTb[p] += b * RT2P[p];
      REFrtb[p] += -b * RT2P[p];
      if constexpr (TER) {
        // the ghost levels, read by the terrain terms of the horizontal
        // pressure gradient: DPRES0 - DENS(KS) (PHI(KS-1) - PHI(KS+1)) and its KE mirror
        if (k == KS) {
          const double bg = DPRESb[p - pl];
          DPRES0b[p - pl] += bg;
          Db[p] += -bg * (PHI[p - pl] - PHI[p + pl]);
        }
        if (k == KE) {
          const double bg = DPRESb[p + pl];
          DPRES0b[p + pl] += bg;
          Db[p] += -bg * (PHI[p + pl] - PHI[p - pl]);
        }
      }
    }
  }, st));
  return hipSuccess;
}

hipError_t hevi_stage_adj(Ctx& c, const StageAdjIO& io, hipStream_t s) {
  HIPCHK(adj_setup(c, 0, 0));
  if (c.terrain) return c.scheme == UD3 ? stage_adj_t<UD3, true>(c, io, s) : stage_adj_t<CD4, true>(c, io, s);
  return c.scheme == UD3 ? stage_adj_t<UD3, false>(c, io, s) : stage_adj_t<CD4, false>(c, io, s);
}

// ---------------------------------------------------------------------------
// periodic halo fill (COMM_vars8), reverse: halo adjoints go to their sources
// up to 16 fields of n3 doubles, for one launch over all of them
struct Fields {
  double* p[16];
  int n;
};

// Transpose of the periodic halo fill for nf fields: the fill copies the x
// faces of the interior rows first, then the whole south/north rows, so the
// reverse folds the south/north halo rows (corners included) into their
// images, then the x halos of the interior rows, zeroing what it folded.
static hipError_t halo_adj_n(const Ctx& c, const Fields f, hipStream_t s) {
  const Dims d = c.d;
  const int IMAX = d.IE - d.IS + 1, JMAX = d.JE - d.JS + 1;
  const int JH = d.JS, IH = d.IS;  // halo widths (the halo rows below JS, columns left of IS)
  const int JN = d.JA - 1 - d.JE, IN = d.IA - 1 - d.IE;
  HIPCHK(for3(Box{0, d.KA - 1, 0, d.IA - 1, 0, JH + JN - 1}, [=] __device__(int k, int i, int jj) {
    const int j = jj < JH ? jj : d.JE + 1 + (jj - JH);
    const int jt = jj < JH ? j + JMAX : j - JMAX;
    for (int q = 0; q < f.n; ++q) {
      double* ab = f.p[q];
      const size_t p = IX(d, k, i, j);
      aadd(ab + IX(d, k, i, jt), ab[p]);
      ab[p] = 0.0;
    }
  }, s));
  return for3(Box{0, d.KA - 1, 0, IH + IN - 1, d.JS, d.JE}, [=] __device__(int k, int ii, int j) {
    const int i = ii < IH ? ii : d.IE + 1 + (ii - IH);
    const int it = ii < IH ? i + IMAX : i - IMAX;
    for (int q = 0; q < f.n; ++q) {
      double* ab = f.p[q];
      const size_t p = IX(d, k, i, j);
      aadd(ab + IX(d, k, it, j), ab[p]);
      ab[p] = 0.0;
    }
  }, s);
}

static hipError_t halo_adj5(const Ctx& c, double* const ab[5], hipStream_t s) {
  Fields f{};
  for (int v = 0; v < 5; ++v) f.p[v] = ab[v];
  f.n = 5;
  return halo_adj_n(c, f, s);
}

// ---------------------------------------------------------------------------
// numerical filter (COMMON.F90:385-985 with calc_diff3 :1652-1800), reverse

namespace {
struct NfTab {
  const double *CNZ3, *CNX3, *CNY3, *ndc;
  int KA, IA, JA;
  __device__ double cz(int n, int k, int b) const { return CNZ3[(n - 1) + 3 * ((size_t)k + (size_t)KA * (b - 1))]; }
  __device__ double cx(int n, int i, int b) const { return CNX3[(n - 1) + 3 * ((size_t)i + (size_t)IA * (b - 1))]; }
  __device__ double cy(int n, int j, int b) const { return CNY3[(n - 1) + 3 * ((size_t)j + (size_t)JA * (b - 1))]; }
  __device__ double cdz(int k) const { return ndc[k]; }
  __device__ double fdz(int k) const { return ndc[KA + k]; }
  __device__ double cdx(int i) const { return ndc[2 * KA + i]; }
  __device__ double fdx(int i) const { return ndc[2 * KA + IA + i]; }
  __device__ double cdy(int j) const { return ndc[2 * KA + 2 * IA + j]; }
  __device__ double fdy(int j) const { return ndc[2 * KA + 2 * IA + JA + j]; }
};

// the real rows of calc_diff3's z difference (ghost rows are mirrors)
__device__ __forceinline__ double dz_row(const NfTab& t, const Dims& d, const double* phi, size_t c0, int k, int KO) {
  const int KS = d.KS, KE = d.KE;
  const size_t pl = d.pl;
  auto P = [&](int kk) { return phi[c0 + pl * (size_t)kk]; };
  if (KO == 0) {
    if (k == KS) return t.cz(1, KS + 1, 1) * P(KS + 2) - t.cz(2, KS + 1, 1) * P(KS + 1) + t.cz(3, KS + 1, 1) * P(KS) - t.cz(1, KS, 1) * P(KS + 1);
    if (k == KE - 1) return t.cz(1, KE, 1) * P(KE - 1) - t.cz(2, KE, 1) * P(KE) + t.cz(3, KE, 1) * P(KE - 1) - t.cz(1, KE - 1, 1) * P(KE - 2);
    return t.cz(1, k + 1, 1) * P(k + 2) - t.cz(2, k + 1, 1) * P(k + 1) + t.cz(3, k + 1, 1) * P(k) - t.cz(1, k, 1) * P(k - 1);
  }
  if (k == KS + 1) return t.cz(1, KS + 1, 2) * P(KS + 2) - t.cz(2, KS + 1, 2) * P(KS + 1) + t.cz(3, KS + 1, 2) * P(KS) - t.cz(1, KS, 2) * P(KS + 1);
  if (k == KE - 1) return -t.cz(2, KE - 1, 2) * P(KE - 1) + t.cz(3, KE - 1, 2) * P(KE - 2) - t.cz(1, KE - 2, 2) * P(KE - 3);
  if (k == KE) return t.cz(1, KE, 2) * P(KE - 1) + t.cz(3, KE, 2) * P(KE - 1) - t.cz(1, KE - 1, 2) * P(KE - 2);
  return t.cz(1, k, 2) * P(k + 1) - t.cz(2, k, 2) * P(k) + t.cz(3, k, 2) * P(k - 1) - t.cz(1, k - 1, 2) * P(k - 2);
}
}  // namespace

// forward calc_diff3 at (k, i, j) of an interior column, levels KS-1..KE
static hipError_t diff3_fwd(const Ctx& c, const NfTab t, const double* phi, double* wZ, double* wX, double* wY, int KO, int IO, int JO,
                            hipStream_t s) {
  const Dims d = c.d;
  return for3(Box{d.KS - 1, d.KE, d.IS, d.IE, d.JS, d.JE}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j), c0 = IX(d, 0, i, j);
    const int KS = d.KS, KE = d.KE;
    double z;
    if (KO == 0) {
      if (k == KS - 1) z = -dz_row(t, d, phi, c0, KS, 0);
      else if (k == KE) z = -dz_row(t, d, phi, c0, KE - 1, 0);
      else z = dz_row(t, d, phi, c0, k, 0);
    } else {
      if (k == KS) z = -dz_row(t, d, phi, c0, KS + 1, 1);
      else if (k == KS - 1) z = -dz_row(t, d, phi, c0, KS + 2, 1);
      else z = dz_row(t, d, phi, c0, k, 1);
    }
    wZ[p] = z;
    if (k >= KS && k <= KE - KO) {
      const size_t IA = d.IA;
      wX[p] = IO == 0 ? t.cx(1, i + 1, 1) * phi[p + 2] - t.cx(2, i + 1, 1) * phi[p + 1] + t.cx(3, i + 1, 1) * phi[p] - t.cx(1, i, 1) * phi[p - 1]
                      : t.cx(1, i, 2) * phi[p + 1] - t.cx(2, i, 2) * phi[p] + t.cx(3, i, 2) * phi[p - 1] - t.cx(1, i - 1, 2) * phi[p - 2];
      wY[p] = JO == 0 ? t.cy(1, j + 1, 1) * phi[p + 2 * IA] - t.cy(2, j + 1, 1) * phi[p + IA] + t.cy(3, j + 1, 1) * phi[p] - t.cy(1, j, 1) * phi[p - IA]
                      : t.cy(1, j, 2) * phi[p + IA] - t.cy(2, j, 2) * phi[p] + t.cy(3, j, 2) * phi[p - IA] - t.cy(1, j - 1, 2) * phi[p - 2 * IA];
    } else {
      wX[p] = 0.0; wY[p] = 0.0;
    }
  }, s);
}

// reverse of calc_diff3: the w*b arrays are consumed, phib accumulated
static hipError_t diff3_adj(const Ctx& c, const NfTab t, double* dZb, const double* dXb, const double* dYb, double* phib, int KO, int IO,
                            int JO, hipStream_t s) {
  const Dims d = c.d;
  HIPCHK(for3(Box{0, 0, d.IS, d.IE, d.JS, d.JE}, [=] __device__(int, int i, int j) {
    const size_t c0 = IX(d, 0, i, j), pl = d.pl;
    const int KS = d.KS, KE = d.KE;
    double* db = dZb + c0;
    double* pb = phib + c0;
    auto B = [&](int k) -> double& { return db[pl * (size_t)k]; };
    auto Pb = [&](int k) -> double& { return pb[pl * (size_t)k]; };
    if (KO == 0) {
      B(KE - 2) -= B(KE + 1); B(KE - 1) -= B(KE); B(KE + 1) = 0.0; B(KE) = 0.0; B(KE + 2) = 0.0;
      { const double b = B(KE - 1);
        Pb(KE - 1) += t.cz(1, KE, 1) * b; Pb(KE) -= t.cz(2, KE, 1) * b; Pb(KE - 1) += t.cz(3, KE, 1) * b; Pb(KE - 2) -= t.cz(1, KE - 1, 1) * b; }
      B(KS + 1) -= B(KS - 2); B(KS) -= B(KS - 1); B(KS - 2) = 0.0; B(KS - 1) = 0.0;
      { const double b = B(KS);
        Pb(KS + 2) += t.cz(1, KS + 1, 1) * b; Pb(KS + 1) -= t.cz(2, KS + 1, 1) * b; Pb(KS) += t.cz(3, KS + 1, 1) * b; Pb(KS + 1) -= t.cz(1, KS, 1) * b; }
      // rows KS+1..KE-2 in a register window w0..w3 = Pb(k-1..k+2): each
      // level is loaded and stored once, its terms added in the rows' order
      if (KS + 1 <= KE - 2) {
        double w0 = Pb(KS), w1 = Pb(KS + 1), w2 = Pb(KS + 2), w3 = Pb(KS + 3);
        for (int k = KS + 1; k <= KE - 2; ++k) {
          const double b = B(k);
          w3 += t.cz(1, k + 1, 1) * b; w2 -= t.cz(2, k + 1, 1) * b; w1 += t.cz(3, k + 1, 1) * b; w0 -= t.cz(1, k, 1) * b;
          Pb(k - 1) = w0;
          w0 = w1; w1 = w2; w2 = w3; w3 = Pb(k + 3);
        }
        Pb(KE - 2) = w0; Pb(KE - 1) = w1; Pb(KE) = w2;
      }
    } else {
      B(KE - 1) -= B(KE + 2); B(KE) -= B(KE + 1); B(KE + 1) = 0.0; B(KE + 2) = 0.0;
      { const double b = B(KE);
        Pb(KE - 1) += t.cz(1, KE, 2) * b; Pb(KE - 1) += t.cz(3, KE, 2) * b; Pb(KE - 2) -= t.cz(1, KE - 1, 2) * b; }
      { const double b = B(KE - 1);
        Pb(KE - 1) -= t.cz(2, KE - 1, 2) * b; Pb(KE - 2) += t.cz(3, KE - 1, 2) * b; Pb(KE - 3) -= t.cz(1, KE - 2, 2) * b; }
      B(KS + 3) -= B(KS - 2); B(KS + 2) -= B(KS - 1); B(KS + 1) -= B(KS); B(KS - 2) = 0.0; B(KS - 1) = 0.0; B(KS) = 0.0;
      { const double b = B(KS + 1);
        Pb(KS + 2) += t.cz(1, KS + 1, 2) * b; Pb(KS + 1) -= t.cz(2, KS + 1, 2) * b; Pb(KS) += t.cz(3, KS + 1, 2) * b; Pb(KS + 1) -= t.cz(1, KS, 2) * b; }
      // rows KS+2..KE-2 in a register window w0..w3 = Pb(k-2..k+1)
      if (KS + 2 <= KE - 2) {
        double w0 = Pb(KS), w1 = Pb(KS + 1), w2 = Pb(KS + 2), w3 = Pb(KS + 3);
        for (int k = KS + 2; k <= KE - 2; ++k) {
          const double b = B(k);
          w3 += t.cz(1, k, 2) * b; w2 -= t.cz(2, k, 2) * b; w1 += t.cz(3, k, 2) * b; w0 -= t.cz(1, k - 1, 2) * b;
          Pb(k - 2) = w0;
          w0 = w1; w1 = w2; w2 = w3; w3 = Pb(k + 2);
        }
        Pb(KE - 3) = w0; Pb(KE - 2) = w1; Pb(KE - 1) = w2; Pb(KE) = w3;
      }
    }
  }, s));
  // the x and y differences, gathered: each point of phi sums the terms of
  // the (up to four) differences along x and along y that read it; the
  // differences exist on k = KS..KE-KO, i = IS..IE, j = JS..JE
  return for3(Box{d.KS, d.KE - KO, d.IS - 2, d.IE + 2, d.JS - 2, d.JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t q = IX(d, k, i, j), IA = d.IA;
    const bool rowx = j >= d.JS && j <= d.JE, rowy = i >= d.IS && i <= d.IE;
    auto inx = [&](int ii) { return rowx && ii >= d.IS && ii <= d.IE; };
    auto iny = [&](int jj) { return rowy && jj >= d.JS && jj <= d.JE; };
    double acc = 0.0;
    if (IO == 0) {  // dX(i) = cx(1,i+1,1) phi(i+2) - cx(2,i+1,1) phi(i+1) + cx(3,i+1,1) phi(i) - cx(1,i,1) phi(i-1)
      if (inx(i - 2)) acc += t.cx(1, i - 1, 1) * dXb[q - 2];
      if (inx(i - 1)) acc += -t.cx(2, i, 1) * dXb[q - 1];
      if (inx(i)) acc += t.cx(3, i + 1, 1) * dXb[q];
      if (inx(i + 1)) acc += -t.cx(1, i + 1, 1) * dXb[q + 1];
    } else {        // dX(i) = cx(1,i,2) phi(i+1) - cx(2,i,2) phi(i) + cx(3,i,2) phi(i-1) - cx(1,i-1,2) phi(i-2)
      if (inx(i - 1)) acc += t.cx(1, i - 1, 2) * dXb[q - 1];
      if (inx(i)) acc += -t.cx(2, i, 2) * dXb[q];
      if (inx(i + 1)) acc += t.cx(3, i + 1, 2) * dXb[q + 1];
      if (inx(i + 2)) acc += -t.cx(1, i + 1, 2) * dXb[q + 2];
    }
    if (JO == 0) {
      if (iny(j - 2)) acc += t.cy(1, j - 1, 1) * dYb[q - 2 * IA];
      if (iny(j - 1)) acc += -t.cy(2, j, 1) * dYb[q - IA];
      if (iny(j)) acc += t.cy(3, j + 1, 1) * dYb[q];
      if (iny(j + 1)) acc += -t.cy(1, j + 1, 1) * dYb[q + IA];
    } else {
      if (iny(j - 1)) acc += t.cy(1, j - 1, 2) * dYb[q - IA];
      if (iny(j)) acc += -t.cy(2, j, 2) * dYb[q];
      if (iny(j + 1)) acc += t.cy(3, j + 1, 2) * dYb[q + IA];
      if (iny(j + 2)) acc += -t.cy(1, j + 1, 2) * dYb[q + 2 * IA];
    }
    phib[q] += acc;
  }, s);
}

// num_diff_b (15 fields, consumed) -> adjoints of DENS..RHOT (accumulated)
hipError_t numfilter_adj(Ctx& c, double* NDb, const double* const st[5], const double* REF_dens, const double* REF_pott,
                         double* const stb[5], hipStream_t s) {
  HIPCHK(adj_setup(c, 0, 0));
  AdjWork& A = *c.adj;
  const Dims d = c.d;
  const size_t n3 = d.n3, pl = d.pl, IA = d.IA;
  const int KS = d.KS, KE = d.KE, IS = d.IS, IE = d.IE, JS = d.JS, JE = d.JE;
  const NfTab t{c.CNZ3, c.CNX3, c.CNY3, c.ndc, d.KA, d.IA, d.JA};
  const double sfc = c.nd_sfc_fact;
  double *fld = A.w[W_FLD], *fb = A.w[W_FB];
  double *wZ = A.w[W_W0], *wX = A.w[W_W1], *wY = A.w[W_W2], *wZb = A.w[W_WB0], *wXb = A.w[W_WB1], *wYb = A.w[W_WB2];
  const double *D = st[0], *MZ = st[1], *MX = st[2], *MY = st[3], *RT = st[4];
  double *Db = stb[0], *MZb = stb[1], *MXb = stb[2], *MYb = stb[3], *RTb = stb[4];
  auto NB = [=](int v, int dir) { return NDb + n3 * (size_t)(v + 5 * dir); };
  {
    Fields f{};
    for (int q = 0; q < 15; ++q) f.p[q] = NDb + n3 * (size_t)q;
    f.n = 15;
    HIPCHK(halo_adj_n(c, f, s));
  }
  // surface factor rows of the X and Y components
  HIPCHK(for3(Box{0, 9, IS, IE, JS, JE}, [=] __device__(int vd, int i, int j) {
    double* b = NDb + n3 * (size_t)((vd >> 1) + 5 * (1 + (vd & 1)));
    b[IX(d, KS, i, j)] *= sfc;
    b[IX(d, KS + 1, i, j)] *= (1.0 + sfc) * 0.5;
  }, s));
  auto zero_w = [&]() -> hipError_t {  // wZb, wXb, wYb and fb: adjacent work fields
    return hipMemsetAsync(wZb, 0, (W_FB - W_WB0 + 1) * n3 * sizeof(double), s);
  };

  // density (COMMON.F90:541-621)
  HIPCHK(zero_w());
  {
    const double *bZ = NB(I_DENS, ZDIR), *bX = NB(I_DENS, XDIR), *bY = NB(I_DENS, YDIR);
    HIPCHK(for3(Box{KS - 1, KE, IS, IE, JS, JE}, [=] __device__(int k, int i, int j) {
      const size_t p = IX(d, k, i, j);
      wZb[p] = bZ[p] * t.cdz(k);
      if (k >= KS) { wXb[p] = bX[p] * t.cdx(i); wYb[p] = bY[p] * t.cdy(j); }
    }, s));
  }
  HIPCHK(diff3_adj(c, t, wZb, wXb, wYb, fb, 0, 0, 0, s));
  HIPCHK(for3(Box{KS, KE, IS - 1, IE + 2, JS - 1, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    Db[p] += fb[p];
  }, s));

  // z-momentum (COMMON.F90:629-707)
  HIPCHK(for3(Box{KS, KE - 1, IS - 2, IE + 2, JS - 2, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    fld[p] = 2.0 * MZ[p] / (D[p + pl] + D[p]);
  }, s));
  HIPCHK(diff3_fwd(c, t, fld, wZ, wX, wY, 1, 0, 0, s));
  HIPCHK(zero_w());
  {
    const double *bZ = NB(I_MOMZ, ZDIR), *bX = NB(I_MOMZ, XDIR), *bY = NB(I_MOMZ, YDIR);
    HIPCHK(for3(Box{KS, KE - 1, IS, IE, JS, JE}, [=] __device__(int k, int i, int j) {
      const size_t p = IX(d, k, i, j);
      if (k >= KS + 1) {
        const double b = bZ[p];
        wZb[p] = b * t.fdz(k) * D[p];
        aadd(Db + p, b * wZ[p] * t.fdz(k));
      }
      double b = bX[p];
      wXb[p] = b * t.cdx(i) * 0.25 * (D[p + pl + 1] + D[p + pl] + D[p + 1] + D[p]);
      double cc = b * wX[p] * t.cdx(i) * 0.25;
      aadd(Db + p + pl + 1, cc); aadd(Db + p + pl, cc); aadd(Db + p + 1, cc); aadd(Db + p, cc);
      b = bY[p];
      wYb[p] = b * t.cdy(j) * 0.25 * (D[p + pl + IA] + D[p + pl] + D[p + IA] + D[p]);
      cc = b * wY[p] * t.cdy(j) * 0.25;
      aadd(Db + p + pl + IA, cc); aadd(Db + p + pl, cc); aadd(Db + p + IA, cc); aadd(Db + p, cc);
    }, s));
  }
  HIPCHK(diff3_adj(c, t, wZb, wXb, wYb, fb, 1, 0, 0, s));
  // fld = 2 MOMZ / (D(k+1) + D(k)) on k = KS..KE-1: MOMZ's adjoint at the
  // point, DENS's gathered from the two faces whose average reads it
  HIPCHK(for3(Box{KS, KE, IS - 2, IE + 2, JS - 2, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    auto dsb = [&](size_t r) { return -fb[r] * fld[r] / (D[r + pl] + D[r]); };
    double acc = 0.0;
    if (k <= KE - 1) {
      MZb[p] += fb[p] * 2.0 / (D[p + pl] + D[p]);
      acc += dsb(p);
    }
    if (k - 1 >= KS) acc += dsb(p - pl);
    Db[p] += acc;
  }, s));

  // x-momentum (COMMON.F90:713-790)
  HIPCHK(for3(Box{KS, KE, IS - 2, IE + 1, JS - 2, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    fld[p] = 2.0 * MX[p] / (D[p + 1] + D[p]);
  }, s));
  HIPCHK(diff3_fwd(c, t, fld, wZ, wX, wY, 0, 1, 0, s));
  HIPCHK(zero_w());
  {
    const double *bZ = NB(I_MOMX, ZDIR), *bX = NB(I_MOMX, XDIR), *bY = NB(I_MOMX, YDIR);
    HIPCHK(for3(Box{KS, KE, IS, IE, JS, JE}, [=] __device__(int k, int i, int j) {
      const size_t p = IX(d, k, i, j);
      if (k <= KE - 1) {
        const double b = bZ[p];
        wZb[p] = b * t.cdz(k) * 0.25 * (D[p + pl + 1] + D[p + pl] + D[p + 1] + D[p]);
        const double cc = b * wZ[p] * t.cdz(k) * 0.25;
        aadd(Db + p + pl + 1, cc); aadd(Db + p + pl, cc); aadd(Db + p + 1, cc); aadd(Db + p, cc);
      }
      double b = bX[p];
      wXb[p] = b * t.fdx(i) * D[p];
      aadd(Db + p, b * wX[p] * t.fdx(i));
      b = bY[p];
      wYb[p] = b * t.cdy(j) * 0.25 * (D[p + 1 + IA] + D[p + 1] + D[p + IA] + D[p]);
      const double cc = b * wY[p] * t.cdy(j) * 0.25;
      aadd(Db + p + 1 + IA, cc); aadd(Db + p + 1, cc); aadd(Db + p + IA, cc); aadd(Db + p, cc);
    }, s));
  }
  HIPCHK(diff3_adj(c, t, wZb, wXb, wYb, fb, 0, 1, 0, s));
  // fld = 2 MOMX / (D(i+1) + D(i)) on i = IS-2..IE+1, gathered as for MOMZ
  HIPCHK(for3(Box{KS, KE, IS - 2, IE + 2, JS - 2, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    auto dsb = [&](size_t r) { return -fb[r] * fld[r] / (D[r + 1] + D[r]); };
    double acc = 0.0;
    if (i <= IE + 1) {
      MXb[p] += fb[p] * 2.0 / (D[p + 1] + D[p]);
      acc += dsb(p);
    }
    if (i - 1 >= IS - 2) acc += dsb(p - 1);
    Db[p] += acc;
  }, s));

  // y-momentum (COMMON.F90:796-873)
  HIPCHK(for3(Box{KS, KE, IS - 2, IE + 2, JS - 2, JE + 1}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    fld[p] = 2.0 * MY[p] / (D[p + IA] + D[p]);
  }, s));
  HIPCHK(diff3_fwd(c, t, fld, wZ, wX, wY, 0, 0, 1, s));
  HIPCHK(zero_w());
  {
    const double *bZ = NB(I_MOMY, ZDIR), *bX = NB(I_MOMY, XDIR), *bY = NB(I_MOMY, YDIR);
    HIPCHK(for3(Box{KS, KE, IS, IE, JS, JE}, [=] __device__(int k, int i, int j) {
      const size_t p = IX(d, k, i, j);
      if (k <= KE - 1) {
        const double b = bZ[p];
        wZb[p] = b * t.cdz(k) * 0.25 * (D[p + pl + IA] + D[p + pl] + D[p + IA] + D[p]);
        const double cc = b * wZ[p] * t.cdz(k) * 0.25;
        aadd(Db + p + pl + IA, cc); aadd(Db + p + pl, cc); aadd(Db + p + IA, cc); aadd(Db + p, cc);
      }
      double b = bX[p];
      wXb[p] = b * t.cdx(i) * 0.25 * (D[p + 1 + IA] + D[p + IA] + D[p + 1] + D[p]);
      const double cc = b * wX[p] * t.cdx(i) * 0.25;
      aadd(Db + p + 1 + IA, cc); aadd(Db + p + IA, cc); aadd(Db + p + 1, cc); aadd(Db + p, cc);
      b = bY[p];
      wYb[p] = b * t.fdy(j) * D[p];
      aadd(Db + p, b * wY[p] * t.fdy(j));
    }, s));
  }
  HIPCHK(diff3_adj(c, t, wZb, wXb, wYb, fb, 0, 0, 1, s));
  // fld = 2 MOMY / (D(j+1) + D(j)) on j = JS-2..JE+1, gathered as for MOMZ
  HIPCHK(for3(Box{KS, KE, IS - 2, IE + 2, JS - 2, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    auto dsb = [&](size_t r) { return -fb[r] * fld[r] / (D[r + IA] + D[r]); };
    double acc = 0.0;
    if (j <= JE + 1) {
      MYb[p] += fb[p] * 2.0 / (D[p + IA] + D[p]);
      acc += dsb(p);
    }
    if (j - 1 >= JS - 2) acc += dsb(p - IA);
    Db[p] += acc;
  }, s));

  // rho*theta (COMMON.F90:880-962)
  HIPCHK(for3(Box{KS, KE, IS - 1, IE + 2, JS - 1, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    fld[p] = RT[p] / D[p] - REF_pott[p];
  }, s));
  HIPCHK(diff3_fwd(c, t, fld, wZ, wX, wY, 0, 0, 0, s));
  HIPCHK(zero_w());
  {
    const double *bZ = NB(I_RHOT, ZDIR), *bX = NB(I_RHOT, XDIR), *bY = NB(I_RHOT, YDIR);
    HIPCHK(for3(Box{KS - 1, KE, IS, IE, JS, JE}, [=] __device__(int k, int i, int j) {
      const size_t p = IX(d, k, i, j);
      const double b = bZ[p];
      if (k == KS - 1) {
        wZb[p] = b * t.cdz(k) * D[p + pl];
        aadd(Db + p + pl, b * wZ[p] * t.cdz(k));
        return;
      }
      if (k == KE) {
        wZb[p] = b * t.cdz(k) * D[p];
        aadd(Db + p, b * wZ[p] * t.cdz(k));
      } else {
        wZb[p] = b * t.cdz(k) * 0.5 * (D[p + pl] + D[p]);
        const double cc = b * wZ[p] * t.cdz(k) * 0.5;
        aadd(Db + p + pl, cc); aadd(Db + p, cc);
      }
      double bb = bX[p];
      wXb[p] = bb * t.cdx(i) * 0.5 * (D[p + 1] + D[p]);
      double cc = bb * wX[p] * t.cdx(i) * 0.5;
      aadd(Db + p + 1, cc); aadd(Db + p, cc);
      bb = bY[p];
      wYb[p] = bb * t.cdy(j) * 0.5 * (D[p + IA] + D[p]);
      cc = bb * wY[p] * t.cdy(j) * 0.5;
      aadd(Db + p + IA, cc); aadd(Db + p, cc);
    }, s));
  }
  HIPCHK(diff3_adj(c, t, wZb, wXb, wYb, fb, 0, 0, 0, s));
  HIPCHK(for3(Box{KS, KE, IS - 1, IE + 2, JS - 1, JE + 2}, [=] __device__(int k, int i, int j) {
    const size_t p = IX(d, k, i, j);
    RTb[p] += fb[p] / D[p];
    Db[p] += -fb[p] * RT[p] / (D[p] * D[p]);
  }, s));
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// large step (LARGE.F90:377-1127 with RK4.F90:300-660), reverse of one step
// from the state x at its start; xb: adjoint of the state after the step on
// entry, of x on return
// Forward of one large step from x with everything its reverse reads kept
// in rec; with xnext, the new state (RK4 combination, ghost levels, halos)
// (the tendencies M.t are the ones the large-step preparation left: they do
// not depend on the state)
static hipError_t step_fwd(Ctx& c, double* const x[5], const StepRec& rec, double* const xnext[5]) {
  Model& M = *c.model;
  AdjWork& A = *c.adj;
  const Dims d = c.d;
  hipStream_t s = c.stream;
  const size_t bytes = d.n3 * sizeof(double);
  // RT2P, DPRES0, REF_rhot of this step, from its RHOT, straight into the record
  HIPCHK(large_prep_into(c, x[4], rec.RT2P, rec.DPRES0, rec.REF_rhot, s));
  if (M.nd_coef == 0.0) {
    HIPCHK(hipMemsetAsync(rec.nd, 0, 15 * bytes, s));
  } else {
    const double* xs[5] = {x[0], x[1], x[2], x[3], x[4]};
    HIPCHK(numfilter(c, rec.nd, xs, M.REF_dens, M.REF_pott, s));
  }
  const double dtrk_f[4] = {M.dts / 2.0, M.dts / 2.0, M.dts, M.dts};
  double* R4[5];
  for (int v = 0; v < 5; ++v) R4[v] = A.w[W_O + v];
  for (int r = 0; r < 4; ++r) {
    StageIO f{};
    f.ctx = &c;
    for (int v = 0; v < 5; ++v) {
      f.o[v] = r < 3 ? rec.R[r][v] : R4[v];
      f.s0[v] = x[v];
      f.s[v] = r == 0 ? x[v] : rec.R[r - 1][v];
      f.t[v] = M.t[v];
    }
    f.mflx_hi = A.mf; f.tflx_hi = A.tf;
    f.DPRES0 = rec.DPRES0; f.RT2P = rec.RT2P; f.CORIOLI = M.CORIOLI; f.num_diff = rec.nd;
    f.wdamp_coef = M.wdamp_coef; f.DDIV = nullptr; f.REF_dens = M.REF_dens; f.REF_rhot = rec.REF_rhot;
    f.divdmp_coef = 0.0; f.dtrk = dtrk_f[r]; f.has_coriolis = M.has_coriolis; f.combine = 0;
    HIPCHK(hevi_stage_traced(f, rec.tr[r][0], rec.tr[r][1], rec.tr[r][2], s));
    if (r < 3) HIPCHK(halo_fill(c, rec.R[r], 5, s));
  }
  if (!xnext) return hipSuccess;
  double* Rk[4][5];
  for (int v = 0; v < 5; ++v) {
    for (int r = 0; r < 3; ++r) Rk[r][v] = rec.R[r][v];
    Rk[3][v] = R4[v];
  }
  const double* s0[5] = {x[0], x[1], x[2], x[3], x[4]};
  return rk4_combine_state(c, xnext, s0, Rk, A.mf, A.tf, s);
}

// Reverse of one large step: xb holds the adjoint of the state after the
// step on entry and of x on return
static hipError_t step_rev(Ctx& c, double* const x[5], double* const xb[5], const StepRec& rec) {
  Model& M = *c.model;
  AdjWork& A = *c.adj;
  const Dims d = c.d;
  hipStream_t s = c.stream;
  const size_t n3 = d.n3, bytes = n3 * sizeof(double);
  const int KS = d.KS, KE = d.KE;
  (void)n3;

  // halo fill and ghost levels of the new state (LARGE.F90:908-927)
  HIPCHK(halo_adj5(c, xb, s));
  {
    double* b0 = xb[0]; double* b1 = xb[1]; double* b2 = xb[2]; double* b3 = xb[3]; double* b4 = xb[4];
    HIPCHK(for3(Box{0, 0, d.IS, d.IE, d.JS, d.JE}, [=] __device__(int, int i, int j) {
      double* bs[5] = {b0, b1, b2, b3, b4};
      for (int v = 0; v < 5; ++v) {
        double* b = bs[v];
        for (int k = d.KA - 1; k >= KE + 1; --k) { if (v != 1) b[IX(d, KE, i, j)] += b[IX(d, k, i, j)]; b[IX(d, k, i, j)] = 0.0; }
        for (int k = 0; k <= KS - 1; ++k) { if (v != 1) b[IX(d, KS, i, j)] += b[IX(d, k, i, j)]; b[IX(d, k, i, j)] = 0.0; }
      }
    }, s));
  }
  // combination (RK4.F90:562-660)
  // S0b and Rb[0..2], one block (Rb[3] is only read, at the points written below)
  HIPCHK(hipMemsetAsync(A.S0b[0], 0, 20 * bytes, s));
  for (int v = 0; v < 5; ++v) {
    double *b = xb[v], *r1 = A.Rb[0][v], *r2 = A.Rb[1][v], *r3 = A.Rb[2][v], *r4 = A.Rb[3][v], *s0 = A.S0b[v];
    HIPCHK(for3(Box{KS, v == 1 ? KE - 1 : KE, d.IS, d.IE, d.JS, d.JE}, [=] __device__(int k, int i, int j) {
      const size_t p = IX(d, k, i, j);
      const double bb = b[p];
      r1[p] = bb * (2.0 / 6.0); r2[p] = bb * (4.0 / 6.0); r3[p] = bb * (2.0 / 6.0); r4[p] = bb * (1.0 / 6.0);
      s0[p] = -bb * (3.0 / 6.0);
    }, s));
  }
  HIPCHK(hipMemsetAsync(A.NDb, 0, 15 * bytes, s));
  HIPCHK(hipMemsetAsync(A.Pb[0], 0, 3 * bytes, s));
  // stages 4..1
  const double dtrk[4] = {M.dts / 2.0, M.dts / 2.0, M.dts, M.dts};
  for (int r = 3; r >= 0; --r) {
    if (r < 3)
      HIPCHK(halo_adj5(c, A.Rb[r], s));
    StageAdjIO io{};
    io.ctx = &c;
    for (int v = 0; v < 5; ++v) {
      io.s0[v] = x[v];
      io.s[v] = r == 0 ? x[v] : rec.R[r - 1][v];
      io.t[v] = M.t[v];
      io.ob[v] = A.Rb[r][v];
      io.s0b[v] = A.S0b[v];
      io.sb[v] = r == 0 ? A.S0b[v] : A.Rb[r - 1][v];
    }
    io.DPRES0 = rec.DPRES0; io.RT2P = rec.RT2P; io.CORIOLI = M.CORIOLI; io.num_diff = rec.nd;
    io.wdamp_coef = M.wdamp_coef; io.REF_dens = M.REF_dens; io.REF_rhot = rec.REF_rhot;
    io.dtrk = dtrk[r]; io.has_coriolis = M.has_coriolis;
    io.NDb = A.NDb; io.DPRES0b = A.Pb[0]; io.RT2Pb = A.Pb[1]; io.REFrtb = A.Pb[2];
    io.trPT = rec.tr[r][0]; io.trSt = rec.tr[r][1]; io.trCp = rec.tr[r][2];
    HIPCHK(hevi_stage_adj(c, io, s));
  }
  // RT2P, DPRES0, REF_rhot as functions of RHOT at the step start (LARGE.F90:443-471)
  {
    const double CPovCV = CPDRY / (CPDRY - RDRY);
    const double *DP0b = A.Pb[0], *R2Pb = A.Pb[1], *RRb = A.Pb[2], *R2P = rec.RT2P, *RT = x[4];
    double* RTb = A.S0b[4];
    double* DP0bw = A.Pb[0];
    HIPCHK(for3(Box{0, 0, 0, d.IA - 1, 0, d.JA - 1}, [=] __device__(int, int i, int j) {
      DP0bw[IX(d, KE - 1, i, j)] += DP0bw[IX(d, KE + 1, i, j)];
      DP0bw[IX(d, KS + 1, i, j)] += DP0bw[IX(d, KS - 1, i, j)];
    }, s));
    HIPCHK(for3(Box{KS, KE, 0, d.IA - 1, 0, d.JA - 1}, [=] __device__(int k, int i, int j) {
      const size_t p = IX(d, k, i, j);
      RTb[p] += R2Pb[p] * R2P[p] * (CPovCV - 1.0) / RT[p] + DP0b[p] * R2P[p] + RRb[p];
    }, s));
  }
  if (M.nd_coef != 0.0) {
    const double* xs[5] = {x[0], x[1], x[2], x[3], x[4]};
    HIPCHK(numfilter_adj(c, A.NDb, xs, M.REF_dens, M.REF_pott, A.S0b, s));
  }
  for (int v = 0; v < 5; ++v) HIPCHK(hipMemcpyAsync(xb[v], A.S0b[v], bytes, hipMemcpyDeviceToDevice, s));
  return hipSuccess;
}

// the single record the reverse uses when the segment's records do not fit:
// the model's own work arrays and the stage traces of AdjWork
static StepRec model_rec(Ctx& c) {
  Model& M = *c.model;
  AdjWork& A = *c.adj;
  StepRec r;
  for (int q = 0; q < 3; ++q)
    for (int v = 0; v < 5; ++v) r.R[q][v] = M.R[q][v];
  r.nd = M.num_diff; r.RT2P = M.RT2P; r.DPRES0 = M.DPRES0; r.REF_rhot = M.REF_rhot;
  for (int q = 0; q < 4; ++q)
    for (int t = 0; t < 3; ++t) r.tr[q][t] = A.tr[q][t];
  return r;
}

static hipError_t check_adjoint_model(const Ctx& c) {
  if (!c.model || c.d.IS != 2) return hipErrorNotSupported;
  const Model& M = *c.model;
  if (M.mixed || M.heve || M.divdmp_coef != 0.0 || M.tinteg_short != 4) return hipErrorNotSupported;
  if ((int)ceil((M.dtl - EPS_RP) / M.dts) != 1) return hipErrorNotSupported;
  return hipSuccess;
}

// The checkpointed forward in pieces: fwd_begin sets up the run,
// fwd_step(n) advances large step n (saving the checkpoint or the tape it
// owes), fwd_end marks the run recorded.  A forward of nsteps large steps
// keeps the state every 'every' steps; its last A.tseg segments also keep
// their step states and records (the tapes), as the reverse sweep's
// recording forward would.  share: the number of models that split the
// device memory for tapes.
static hipError_t fwd_begin(Ctx& c, int nsteps, int every, int share) {
  HIPCHK(check_adjoint_model(c));
  HIPCHK(adj_setup(c, nsteps, every, share));
  c.adj->recorded = 0;
  c.adj->taped = 0;
  return hipSuccess;
}

static hipError_t fwd_step(Ctx& c, int n) {
  AdjWork& A = *c.adj;
  Model& M = *c.model;
  const size_t bytes = c.d.n3 * sizeof(double);
  double* st[5] = {M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT};
  const int nseg = (A.nsteps + A.every - 1) / A.every;
  const int nfirst = (nseg - A.tseg) * A.every;  // first taped step
  if (n % A.every == 0)
    for (int v = 0; v < 5; ++v)
      HIPCHK(hipMemcpyAsync(A.ckpt[5 * (n / A.every) + v], st[v], bytes, hipMemcpyDeviceToDevice, c.stream));
  if (A.tseg > 0 && n >= nfirst) {
    for (int v = 0; v < 5; ++v)
      HIPCHK(hipMemcpyAsync(A.seg[5 * (n - nfirst) + v], st[v], bytes, hipMemcpyDeviceToDevice, c.stream));
    return model_large_step(c, &A.recs[n - nfirst]);
  }
  return model_large_step(c);
}

static void fwd_end(Ctx& c) {
  c.adj->recorded = 1;
  c.adj->taped = c.adj->tseg;
}

hipError_t model_forward_checkpointed(Ctx& c, int nsteps, int every) {
  HIPCHK(fwd_begin(c, nsteps, every, 1));
  prof_start(c, "model_forward_checkpointed", c.stream);
  for (int n = 0; n < nsteps; ++n) HIPCHK(fwd_step(c, n));
  prof_stop(c, c.stream);
  fwd_end(c);
  return hipSuccess;
}

// The reverse sweep in pieces: rev_begin keeps the final state and prepares
// the tendencies, rev_segment(sg) reverses segment sg (the segments run from
// the last to the first), rev_end restores the final state.
static hipError_t rev_begin(Ctx& c) {
  HIPCHK(check_adjoint_model(c));
  if (!c.adj || !c.adj->recorded) return hipErrorNotReady;
  AdjWork& A = *c.adj;
  Model& M = *c.model;
  const size_t bytes = c.d.n3 * sizeof(double);
  double* st[5] = {M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT};
  for (int v = 0; v < 5; ++v) HIPCHK(hipMemcpyAsync(A.xb[v], st[v], bytes, hipMemcpyDeviceToDevice, c.stream));
  return model_prepare(c);  // the tendencies M.t, which every step shares
}

static hipError_t rev_segment(Ctx& c, int sg, double* const xb[5]) {
  AdjWork& A = *c.adj;
  Model& M = *c.model;
  hipStream_t s = c.stream;
  const size_t bytes = c.d.n3 * sizeof(double);
  double* st[5] = {M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT};
  const int nseg = (A.nsteps + A.every - 1) / A.every;
  const bool recs = A.tseg > 0;
  const int n0 = sg * A.every, n1 = std::min(A.nsteps, n0 + A.every);
  // a taped segment reverses from its own slot of the tapes; the others
  // record into slot 0, free by then (its segment is reversed before them)
  const bool have = recs && sg >= nseg - A.taped;
  const int base = have ? (sg - (nseg - A.tseg)) * A.every : 0;
  auto seg = [&](int n, double** x) { for (int v = 0; v < 5; ++v) x[v] = A.seg[5 * (base + n - n0) + v]; };
  if (!have) {
    for (int v = 0; v < 5; ++v) HIPCHK(hipMemcpyAsync(A.seg[v], A.ckpt[5 * sg + v], bytes, hipMemcpyDeviceToDevice, s));
    // slot 0 no longer holds the tape of its own segment (the earliest
    // taped one): a later reverse sweep of the same run records it again
    if (recs) A.taped = std::min(A.taped, A.tseg - 1);
  }
  if (recs) {
    // one recording forward over the segment unless the checkpointed
    // forward kept its tape, then the reverse steps
    for (int n = n0; n < n1 && !have; ++n) {
      double *x[5], *xn[5];
      seg(n, x);
      if (n + 1 < n1) seg(n + 1, xn);
      HIPCHK(step_fwd(c, x, A.recs[base + n - n0], n + 1 < n1 ? xn : nullptr));
    }
    for (int n = n1 - 1; n >= n0; --n) {
      double* x[5];
      seg(n, x);
      HIPCHK(step_rev(c, x, xb, A.recs[base + n - n0]));
    }
    return hipSuccess;
  }
  // recompute the states of the segment, then each reversed step its own record
  const StepRec one = model_rec(c);
  for (int n = n0 + 1; n < n1; ++n) {
    for (int v = 0; v < 5; ++v) HIPCHK(hipMemcpyAsync(st[v], A.seg[5 * (n - 1 - n0) + v], bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(model_prepare(c));
    HIPCHK(model_short_step_rk4(c, M.dts));
    for (int v = 0; v < 5; ++v) HIPCHK(hipMemcpyAsync(A.seg[5 * (n - n0) + v], st[v], bytes, hipMemcpyDeviceToDevice, s));
  }
  for (int n = n1 - 1; n >= n0; --n) {
    double* x[5];
    seg(n, x);
    HIPCHK(step_fwd(c, x, one, nullptr));
    HIPCHK(step_rev(c, x, xb, one));
  }
  return hipSuccess;
}

static hipError_t rev_end(Ctx& c) {
  AdjWork& A = *c.adj;
  Model& M = *c.model;
  const size_t bytes = c.d.n3 * sizeof(double);
  double* st[5] = {M.DENS, M.MOMZ, M.MOMX, M.MOMY, M.RHOT};
  for (int v = 0; v < 5; ++v) HIPCHK(hipMemcpyAsync(st[v], A.xb[v], bytes, hipMemcpyDeviceToDevice, c.stream));
  return hipSuccess;
}

// xb (device, 5 fields): adjoint of the state after the recorded run on entry,
// of the state at its start on return.  The model state is left as the run
// left it.
hipError_t model_adjoint(Ctx& c, double* const xb[5]) {
  HIPCHK(rev_begin(c));
  const int nseg = (c.adj->nsteps + c.adj->every - 1) / c.adj->every;
  prof_start(c, "model_adjoint", c.stream);
  for (int sg = nseg - 1; sg >= 0; --sg) HIPCHK(rev_segment(c, sg, xb));
  prof_stop(c, c.stream);
  return rev_end(c);
}

// Batches of models (members of any grid size, flat or terrain), each on
// its own stream: the checkpointed forwards advance step by step, member
// after member, and the reverse sweeps segment by segment, so that the
// kernels of different members overlap on the device.  The members split
// the device memory for tapes.  xb[5 m + v]: field v of member m.
hipError_t batch_forward_checkpointed(Ctx* const* cs, int n, int nsteps, int every) {
  for (int m = 0; m < n; ++m) HIPCHK(check_adjoint_model(*cs[m]));
  for (int m = 0; m < n; ++m) HIPCHK(fwd_begin(*cs[m], nsteps, every, n));
  for (int k = 0; k < nsteps; ++k)
    for (int m = 0; m < n; ++m) HIPCHK(fwd_step(*cs[m], k));
  for (int m = 0; m < n; ++m) fwd_end(*cs[m]);
  return hipSuccess;
}

hipError_t batch_adjoint(Ctx* const* cs, int n, double* const* xb) {
  for (int m = 0; m < n; ++m) HIPCHK(rev_begin(*cs[m]));
  int nseg = 0;
  for (int m = 0; m < n; ++m) nseg = std::max(nseg, (cs[m]->adj->nsteps + cs[m]->adj->every - 1) / cs[m]->adj->every);
  for (int sg = nseg - 1; sg >= 0; --sg)
    for (int m = 0; m < n; ++m) {
      const AdjWork& A = *cs[m]->adj;
      const int off = nseg - (A.nsteps + A.every - 1) / A.every;  // runs of fewer segments end together
      if (sg - off >= 0) HIPCHK(rev_segment(*cs[m], sg - off, xb + 5 * (size_t)m));
    }
  for (int m = 0; m < n; ++m) HIPCHK(rev_end(*cs[m]));
  return hipSuccess;
}

}  // namespace shp
