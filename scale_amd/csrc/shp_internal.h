// shp_internal.h -- host-side context and launch entry points shared by the
// HIP translation units of libscalehip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "common.h"

namespace shp {

constexpr int NSCR = 8;  // scratch fields of KA*IA*JA doubles owned by the context
constexpr int MFCT_FIELDS = 10;  // velocity, low-order and anti-diffusive fluxes (3 each), phi_lo, R+, R-

struct Model;    // device-resident model state (driver.hip)
struct AdjWork;  // adjoint work arrays and checkpoints (adjoint.hip)
struct Prof;   // named kernel timers (profile.hip)

struct Ctx {
  Dims d;
  int device = 0;
  int ncu = 256;  // compute units of the device
  int scheme = UD3, scheme_tracer = UD3;
  bool terrain = false;          // general metrics (GSQRT/J13G/J23G/MAPF) in use
  bool metrics_set = false;
  hipStream_t stream = nullptr;
  // grid and metrics (device layout)
  double* g1buf = nullptr;       // 1-D arrays, see Grid1D
  Grid1D g1{};
  double *GSQRT = nullptr, *J13G = nullptr, *J23G = nullptr, *MAPF = nullptr, *PHI = nullptr;
  double J33G = 1.0;
  // numerical filter tables
  double *CNZ3 = nullptr, *CNX3 = nullptr, *CNY3 = nullptr, *ndc = nullptr;
  double nd_coef = 0.0, nd_dt = 0.0, nd_sfc_fact = 1.0;
  // tracer filter coefficients nd_coef_cd{z,x,y} for ND_COEF_Q and the large step
  double* ndqc = nullptr;
  double ndq_coef = -1.0, ndq_dt = -1.0, ndq_sfc = 1.0;
  // scratch
  double* scr[NSCR] = {};
  // work arrays of the HEVE momentum FCT (MFCT_FIELDS x n3), allocated on first use
  double* mfct = nullptr;
  // staging buffer for host <-> device layout transposes
  double* stage = nullptr;
  size_t stage_elems = 0;
  // device-resident model
  Model* model = nullptr;
  AdjWork* adj = nullptr;
  int adj_tape = 0;  // segments whose tapes the checkpointed forward keeps (0: half the free memory's worth)
  // named hipEvent timers, null when profiling is off
  Prof* prof = nullptr;
  // launch-shape knobs (scale_hip_tune_set); 0 = chosen from the grid
  int nf_tj = 0;  // numerical-filter tile height: 4, 8 or 16 rows
  int tracer_tile = 1;  // tracer stage on LDS tiles where supported (0: per-point kernels)
  // host copies of 1-D grid arrays (for the numerical-filter tables)
  std::vector<double> hCDZ, hFDZ, hCDX, hFDX, hCDY, hFDY;
};

// one HEVI stage, all pointers in device layout; R is the storage type of the
// prognostic, tendency, flux and num_diff arrays (float: mixed precision)
template <typename R>
struct StageIOT {
  const Ctx* ctx;
  R* o[5];                  // DENS_RK..RHOT_RK
  R *mflx_hi, *tflx_hi;     // (..,3)
  const R* s0[5];           // DENS0..RHOT0
  const R* s[5];            // DENS..RHOT
  const R* t[5];            // DENS_t..RHOT_t
  const double *DPRES0, *RT2P, *CORIOLI;
  const R* num_diff;
  const double* wdamp_coef;
  const R* DDIV;
  const double *REF_dens, *REF_rhot;
  double divdmp_coef, dtrk;
  int has_coriolis;
  // HEVE only: FLAG_FCT_MOMENTUM (needs Ctx::mfct), FLAG_FCT_T and FLAG_FCT_ALONG_STREAM
  int fct_momentum = 0, fct_t = 0, fct_along = 1;
  // FP32 stage: DENS and RHOT stored as deviations from these level profiles
  const double *offD = nullptr, *offR = nullptr;
  // FP64 flat-terrain HEVI only: PT kept in trPT (instead of the context's
  // scratch), St and the solved C recorded in trSt, trCp, for the adjoint
  double *trPT = nullptr, *trSt = nullptr, *trCp = nullptr;
  // last RK4 stage only (flat terrain): combine with the outputs of stages
  // 1-3 on the fly, writing the new state into st (which may alias s0) and the
  // averaged fluxes into mflx_av/tflx_av, instead of o / the Z-free parts of
  // mflx_hi, tflx_hi (tinteg_short_rk4.F90:560-660)
  int combine = 0;
  const R *c1[5], *c2[5], *c3[5];  // DENS..RHOT outputs of stages 1, 2, 3
  const R *cm[3], *ct[3];          // mflx_hi, tflx_hi (3 components, 3 n3 each) of stages 1, 2, 3
  R* st[5];
  R *mflx_av, *tflx_av;
};
using StageIO = StageIOT<double>;

// What the reverse of one large step reads from its forward: the stage
// outputs R1..R3, num_diff, RT2P/DPRES0/REF_rhot, and the PT, St and C of
// the four stages.
struct StepRec {
  double* R[3][5] = {};
  double* nd = nullptr;
  double *RT2P = nullptr, *DPRES0 = nullptr, *REF_rhot = nullptr;
  double* tr[4][3] = {};
};

struct TracerIO {
  double* QTRCo;
  const double *QTRC, *QTRC0, *RHOQ_t, *DENS0, *DENS, *mflx_hi, *num_diff;
  double dtl;
  int flag_fct, flag_along;
};

// ---- ensembles: one launch over the tiles of many members' grids ----
// Device tables of the members' argument structs with first[nmem+1] (first
// logical tile of each member) and ni[nmem] (tiles across), built once.
struct BatchBuf {
  void* dev = nullptr;
  int nmem = 0;
  unsigned total = 0;  // tiles over all members
  int scheme = UD3, tj = 8, fp32 = 0, cmb = 0, ka_max = 0;
};
void batch_free(BatchBuf& b);
// the HEVI stage of every member: flat terrain, no divergence damping, one
// scheme and one precision (cmb: the fused RK4 combination of the last stage)
hipError_t hevi_batch_build(const StageIOT<double>* ios, int n, BatchBuf& b);
hipError_t hevi_batch_build(const StageIOT<float>* ios, int n, BatchBuf& b);
// blocks_per_cu = 0: one block per tile; otherwise a persistent grid of
// blocks_per_cu blocks per CU walking the tiles
hipError_t hevi_batch_launch(const BatchBuf& b, int blocks_per_cu, int ncu, hipStream_t s);
// the flat FP64 stage on LDS tiles (false: the split launches, for grids
// too small to give every CU a tile)
bool hevi_flat_tile(const Ctx& c);
// the numerical filter of every member (then each member's halo fill)
template <typename R>
struct NfIn {
  const Ctx* c;
  R* num_diff;
  const R* st[5];
  const double *REF_dens, *REF_pott, *offD, *offR;
};
hipError_t numfilter_batch_build(const NfIn<double>* in, int n, int tj, BatchBuf& b);
hipError_t numfilter_batch_build(const NfIn<float>* in, int n, int tj, BatchBuf& b);
hipError_t numfilter_batch_launch(const BatchBuf& b, int blocks_per_cu, int ncu, hipStream_t s);

hipError_t hevi_stage(const StageIO& io, hipStream_t s);
// the explicit HEVE stage: same inputs and outputs, FP64, no fused combine
hipError_t heve_stage(const StageIO& io, hipStream_t s);
// FCT of the HEVE theta flux, run by heve_stage after the stage kernel (tracer.hip)
hipError_t heve_fct_theta(const StageIO& io, hipStream_t s);
// ATMOS_DYN_fct of one HEVE momentum component on stored low-order and
// anti-diffusive fluxes, then its MOM_RK update (tracer.hip)
struct MomFct {
  const double *vel, *D0, *D1;   // the velocity (with halos); DENS0, DENS_RK (with halos)
  int dk, di, dj;                // the staggering: densities averaged with (k+dk, i+di, j+dj)
  const double *lo, *anti, *mflx;  // 3 components each
  const double *RDZ, *RDX, *RDY, *G, *MAPF;
  double dt;
  bool along;
  double *phi_lo, *rjpls, *rjmns;  // scratch
  double* out;                     // MOM_RK
};
hipError_t fct_momentum_limit(const Ctx& c, const MomFct& f, hipStream_t s);
// mixed precision: flat terrain and divdmp_coef = 0 only
hipError_t hevi_stage(const StageIOT<float>& io, hipStream_t s);
// flat terrain, FP64, no divergence damping: the stage with PT, St and the
// solved C recorded (n3 each) for the adjoint
hipError_t hevi_stage_traced(const StageIO& io, double* PT, double* St, double* Cp, hipStream_t s);
// named timers around launches (no-ops unless c.prof is set): the counterpart
// of the reference's PROF_rapstart / PROF_rapend (scale_prof.F90:156-253)
void prof_start(const Ctx& c, const char* name, hipStream_t s);
void prof_stop(const Ctx& c, hipStream_t s);
void prof_collect(const Ctx& c);
Prof* prof_create();
void prof_destroy(Prof* p);
void prof_reset(Prof* p);
void prof_outer_only(Prof* p, bool on);  // time only the outermost of nested timers
int prof_query(Prof* p, const char* name, long* count, double* total_ms);
int prof_names(Prof* p, char* buf, int len);
// streaming copy of n doubles (timed as "hbm_copy")
hipError_t hbm_copy(const Ctx& c, double* dst, const double* src, size_t n, hipStream_t s);
// bandwidth probes: 0 copy, 1 non-temporal copy, 2 read-only, 3 write-only, 4 hipMemcpyAsync D2D
hipError_t hbm_probe(const Ctx& c, int kind, double* dst, const double* src, double* part, size_t n, hipStream_t s);
hipError_t numfilter(const Ctx& c, double* num_diff, const double* const st[5], const double* REF_dens,
                     const double* REF_pott, hipStream_t s);
// FP32: DENS and RHOT of st are deviations from the level profiles offD, offR
hipError_t numfilter(const Ctx& c, float* num_diff, const float* const st[5], const double* REF_dens,
                     const double* REF_pott, hipStream_t s, const double* offD, const double* offR);
hipError_t tracer_stage(const Ctx& c, const TracerIO& io, hipStream_t s);
// the LDS-tiled tracer stage (flat terrain; UD3, CD4, UD3KOREN1993, UD1, CD2)
// with R+/R- in the given fields; tracer_stage uses it where it applies
bool tracer_tile_supported(const Ctx& c);
hipError_t tracer_stage_tile(const Ctx& c, const TracerIO& io, double* rjpls, double* rjmns, hipStream_t s);
// the stage without FCT on the same tiles (flat, two-cell schemes)
hipError_t tracer_nofct_tile(const Ctx& c, const TracerIO& io, hipStream_t s);
// the same on caller-supplied device metrics (GSQRT(..,7), MAPF(IA,JA,2),
// RCDZ/RCDX/RCDY) and work arrays (3 n3 doubles: phi_lo, R+, R-)
hipError_t tracer_stage_ext(const Ctx& c, const TracerIO& io, const double* GSQRT, const double* MAPF_XY,
                            const double* RCDZ, const double* RCDX, const double* RCDY, double* work3, hipStream_t s);
// adjoint of one tracer stage (tracer_stage's inputs in io) with the FCT
// limiter frozen at its forward R+/R-: from QTRCo_b (read at the interior)
// overwrites the adjoints of QTRC, QTRC0, RHOQ_t and num_diff (3 planes) on
// the whole array.  Uses c.scr[0..7]; UD3KOREN1993 is rejected.
struct TracerAdjIO {
  const double* QTRCo_b;
  double *QTRC_b, *QTRC0_b, *RHOQ_t_b, *num_diff_b;
};
hipError_t tracer_stage_adj(const Ctx& c, const TracerIO& io, const TracerAdjIO& adj, hipStream_t s);
// one pass of the tracer stage of every member in one launch (flat terrain,
// one tracer scheme with two-cell faces, one FCT bound kind): the table of
// the members' stage arguments (ios: as tracer_stage takes them; the FCT
// work arrays are each member's c.scr[0..2]), for the phi_lo pass (philo) or
// for the others; kinds 0 the stage without FCT, 1 phi_lo, 2 R+/R-, 3 the
// limited update (the R+/R- halo fill between 2 and 3 is the caller's)
hipError_t tracer_batch_build(const Ctx* const* cs, const TracerIO* ios, int n, bool philo, BatchBuf& b);
hipError_t tracer_batch_launch(const BatchBuf& b, int kind, hipStream_t s);
// the same FCT passes on LDS tiles (tracer_tile.hip; two-cell schemes):
// kinds 0 R+/R- (phi_lo within), 1 the update
hipError_t trt_batch_build(const Ctx* const* cs, const TracerIO* ios, int n, BatchBuf& b);
hipError_t trt_batch_launch(const BatchBuf& b, int kind, hipStream_t s);
// ATMOS_DYN_numfilter_coef_q (ND_ORDER = 1) with the halo fill of its 3
// fields; REF_qv non-null for QV.  Needs the filter tables and c.ndqc.
hipError_t numfilter_q(const Ctx& c, double* num_diff_q, const double* DENS, const double* QTRC, const double* REF_qv,
                       hipStream_t s);

// Adjoint of one flat-terrain FP64 stage (adjoint.hip): given the stage's
// inputs and the adjoints ob of its five prognostic outputs (read at the
// points the stage writes), adds the adjoints of its inputs.
struct StageAdjIO {
  const Ctx* ctx;
  const double *s0[5], *s[5], *t[5];
  const double *DPRES0, *RT2P, *CORIOLI, *num_diff, *wdamp_coef, *REF_dens, *REF_rhot;
  double dtrk;
  int has_coriolis;
  const double* ob[5];
  double *s0b[5], *sb[5], *NDb, *DPRES0b, *RT2Pb, *REFrtb;
  // PT, St and the solved C of the forward stage (hevi_stage_traced); when
  // null the stage adjoint reruns the forward stage to record them
  const double *trPT = nullptr, *trSt = nullptr, *trCp = nullptr;
};
hipError_t hevi_stage_adj(Ctx& c, const StageAdjIO& io, hipStream_t s);
// num_diff adjoint (15 fields, read) -> adjoints of DENS..RHOT (added)
hipError_t numfilter_adj(Ctx& c, double* NDb, const double* const st[5], const double* REF_dens, const double* REF_pott,
                         double* const stb[5], hipStream_t s);
// adjoint work arrays (nsteps > 0: also checkpoints of a run of nsteps saved every 'every')
hipError_t adj_setup(Ctx& c, int nsteps, int every, int share = 1);
void adj_free(Ctx& c);
double* adj_io(Ctx& c, int v);  // five device fields for the adjoint arguments of the C ABI
// output buffers of a stage adjoint called through the C ABI (after adj_setup)
void adj_stage_bufs(Ctx& c, double* s0b[5], double* sb[5], double** NDb, double* Pb[3]);
// ATMOS_DYN_divergence on i = IS..IE+1, j = JS..JE+1, k = KS-1..KE+1
hipError_t divergence(const Ctx& c, double* DDIV, const double* MOMZ, const double* MOMX, const double* MOMY, hipStream_t s);
hipError_t halo_fill(const Ctx& c, double* const* arrs, int n, hipStream_t s);
// terrain-following metrics from the surface height Zsfc(IA,JA) (device) and
// the reference heights CZ(KA), FZ(0:KA) (device), Htop = FZ(KE) - FZ(KS-1):
// GSQRT, J13G, J23G (7 components each, with halos) and PHI (metrics.hip)
hipError_t terrain_metrics(const Ctx& c, const double* Zsfc, const double* CZ, const double* FZ, double Htop,
                           double* GSQRT, double* J13G, double* J23G, double* PHI, hipStream_t s);
hipError_t halo_fill(const Ctx& c, float* const* arrs, int n, hipStream_t s);
// Fortran (k fastest) <-> device (i fastest) layout; ncomp fields of KA*IA*JA
hipError_t to_device_layout(const Ctx& c, double* dst, const double* src_f, int ncomp, hipStream_t s);
hipError_t to_fortran_layout(const Ctx& c, double* dst_f, const double* src, int ncomp, hipStream_t s);
// 2-D (IA,JA,n) arrays are identical in both layouts

}  // namespace shp
