// model.h -- device-resident state of one integration (the arrays that
// ATMOS_DYN_Tstep_large_fvm_heve and its RK drivers keep between calls).
#pragma once
#include "shp_internal.h"

namespace shp {

struct Model {
  int QA = 0;
  // prognostics and tracers (device layout, valid halos)
  double *DENS = nullptr, *MOMZ = nullptr, *MOMX = nullptr, *MOMY = nullptr, *RHOT = nullptr;
  double* QTRC = nullptr;  // QA fields
  // reference state and forcing
  double *REF_dens = nullptr, *REF_pott = nullptr, *REF_pres = nullptr;
  double* tp[5] = {};        // physics tendencies (nullptr = zero)
  double* RHOQ_tp = nullptr; // QA fields or nullptr
  double* CORIOLI = nullptr; // (IA,JA)
  double* wdamp_coef = nullptr;
  int has_coriolis = 0;
  // outputs of the last large step
  double *mflx_hi = nullptr, *tflx_hi = nullptr;
  // large-step work arrays
  double *RT2P = nullptr, *DPRES0 = nullptr, *REF_rhot = nullptr, *DENS00 = nullptr, *mflx_av = nullptr;
  double *num_diff = nullptr, *DDIV = nullptr, *t[5] = {};
  double* S0[5] = {};
  double* R[4][5] = {};
  double *mfR[4] = {}, *tfR[4] = {};
  double *ndq = nullptr, *DENS_RK = nullptr, *QRK1 = nullptr, *QRK2 = nullptr;
  double* DENS_RK2 = nullptr;  // the second RK3 tracer stage's density (DENS_RK holds the first's)
  // mixed precision: FP32 state, stage, tendency, filter and flux arrays
  // (the FP64 arrays above keep the reference state, the pressure terms and
  // a copy of the state for the large-step preparation and the budgets)
  int mixed = 0;
  float* fS[5] = {};
  float* fR[4][5] = {};
  float* ft[5] = {};
  float *fnum_diff = nullptr, *fDDIV = nullptr, *fmflx = nullptr, *ftflx = nullptr;
  // the FP32 DENS and RHOT (state and stage values) are deviations from the
  // level profiles foff[0..KA) and foff[KA..2KA), so that their rounding
  // error scales with the deviation, not with the full value
  double* foff = nullptr;
  float *fmfR[4] = {}, *ftfR[4] = {};
  // configuration (PARAM_ATMOS_DYN subset)
  double dtl = 1.0, dts = 1.0, nd_coef = 1.0e-4, nd_sfc_fact = 1.0, divdmp_coef = 0.0, nd_coef_q = 0.0;
  int fct = 0, fct_along = 1;
  int heve = 0;  // ATMOS_DYN_TYPE = FVM-HEVE: explicit stages, separate RK4 combination
  // ATMOS_DYN_TINTEG_SHORT_TYPE: 4 = RK4, 3 = RK3 (Heun), 32 = RK3WS2002;
  // ATMOS_DYN_TINTEG_TRACER_TYPE: 0 = RK3WS2002, 1 = EULER
  int tinteg_short = 4, tinteg_tracer = 0;
  // the tendencies hold zeros with no physics tendencies set: not refilled
  bool tend_zero = false;
  // every allocation, freed together
  std::vector<void*> allocs;
};

// An ensemble: member contexts (each with its model, of any grid size) on
// one device, stepped together.  The filter and the four stages of a short
// step run as one launch each over the tiles of every member; the other
// launches stay per member, and the whole large step may be replayed from a
// captured hipGraph.  (The reference runs ensemble members as separate jobs
// of one MPI program: the bulk-job split of scale-rm/src/scale-rm.F90:180-184.)
struct Ensemble {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<Ctx*> members;
  int ncu = 256;
  // launch shapes: persistent blocks per CU of the batched stage and filter
  // (0: one block per tile) and the filter's tile rows
  int stage_bpc = 0, nf_bpc = 0, nf_tj = 8;
  BatchBuf stage[4], nf;
  // a device table of per-member entries: n entries, the widest w and
  // deepest k of them (halo tables), or total blocks (point tables)
  struct Tab {
    void* dev = nullptr;
    int n = 0, w = 0, k = 0;
    unsigned total = 0;
  };
  // halo tables: the filter's fields, the outputs of stages 1..3, the state
  Tab halo[5];
  // point tables of the large step's preparation and of the flux sum after
  // each short step, every member in one launch
  Tab begin, sum;
  // the tracers of every member in batched launches, where the members have
  // the same number of tracers, the RK3 tracer integration, no tracer filter,
  // one FCT setting and one two-cell tracer scheme: per tracer the tables of
  // stages 1 and 2, of the FCT phi_lo pass and of the last stage; halo tables
  // of the mass fluxes, QRK1, QRK2, R+/R- and each tracer; the point tables
  // of the preparation (flux mean, stage densities) and of each copy back
  bool tr_batched = false, tr_tile = false;
  int tr_qa = 0, tr_fct = 0;
  std::vector<BatchBuf> tr_s1, tr_s2, tr_ph, tr_s3;
  Tab tr_mf, tr_q1, tr_q2, tr_r, tr_prep;
  std::vector<Tab> tr_q, tr_copy;
  bool built = false;
  // lanes: streams the members' own launches spread over (member m on lane
  // m % nlanes), forked from and joined back to `stream` around each batched
  // launch, so that the small per-member kernels of different members overlap
  static constexpr int MAX_LANES = 8;
  int nlanes = 4;
  hipStream_t lane[MAX_LANES] = {};
  hipEvent_t ev_main = nullptr, ev_lane[MAX_LANES] = {};
  // the large step captured once and replayed
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};
hipError_t ensemble_check(const Ensemble& E, const char** why);
hipError_t ensemble_build(Ensemble& E);
hipError_t ensemble_large_step(Ensemble& E, int n);
void ensemble_free(Ensemble& E);

// rec (FP64 flat HEVI without divergence damping, one short step per large
// step): the step also keeps in rec what the reverse sweep reads
hipError_t model_short_step_rk4(Ctx& c, double dts, const StepRec* rec = nullptr);
hipError_t model_prepare(Ctx& c);
hipError_t large_prep_into(Ctx& c, const double* RHOT, double* RT2P, double* DPRES0, double* REF_rhot, hipStream_t s);
hipError_t model_large_step(Ctx& c, const StepRec* rec = nullptr);
hipError_t mass_sum(const Ctx& c, const double* a, const double* b, double* out_host);
// parts of the FP64 copies to refresh from the FP32 state (mixed precision)
enum { F64_DENS = 1, F64_RHOT = 2, F64_MOM = 4, F64_FLUX = 8, F64_ALL = 15 };
hipError_t model_state_to_f64(Ctx& c, int parts = F64_ALL);
hipError_t model_state_to_f32(Ctx& c);
// adjoint (adjoint.hip): flat terrain, FP64, divdmp_coef = 0, one short step per large step
hipError_t model_forward_checkpointed(Ctx& c, int nsteps, int every);
hipError_t rk4_combine_state(Ctx& c, double* const out[5], const double* const s0[5], double* const Rk[4][5], double* mf,
                             double* tf, hipStream_t s);
hipError_t model_adjoint(Ctx& c, double* const xb[5]);
// the same for n models at once, each on its own stream (xb[5 m + v])
hipError_t batch_forward_checkpointed(Ctx* const* cs, int n, int nsteps, int every);
hipError_t batch_adjoint(Ctx* const* cs, int n, double* const* xb);

}  // namespace shp
