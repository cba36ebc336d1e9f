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
  double *ndq = nullptr, *DENS_RK = nullptr, *QRK1 = nullptr, *QRK2 = nullptr, *zeros = nullptr;
  // configuration (PARAM_ATMOS_DYN subset)
  double dtl = 1.0, dts = 1.0, nd_coef = 1.0e-4, nd_sfc_fact = 1.0, divdmp_coef = 0.0;
  int fct = 0, fct_along = 1;
  // every allocation, freed together
  std::vector<void*> allocs;
};

hipError_t model_short_step_rk4(Ctx& c, double dts);
hipError_t model_prepare(Ctx& c);
hipError_t model_large_step(Ctx& c);
hipError_t mass_sum(const Ctx& c, const double* a, const double* b, double* out_host);

}  // namespace shp
