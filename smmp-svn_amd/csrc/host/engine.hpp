// Boundary between the host library and the HIP engine (device/engine.hip).
#pragma once
#include <string>
#include <vector>

#include "../../../include/smmp_mi355x.h"
#include "state.hpp"

namespace smmp {

struct Engine;  // device mirror of one context (opaque to the host code)

struct DeviceError {
  std::string msg;
  explicit DeviceError(std::string m) : msg(std::move(m)) {}
};
// End-of-sweep energy differs from the tracked one by more than 1e-6
// (metropolis.f:172-179); the results of the call are still written.
struct DriftError {
  std::string msg;
  explicit DriftError(std::string m) : msg(std::move(m)) {}
};

void engine_destroy(Engine* e);
void engine_invalidate(Engine* e);  // topology changed: rebuild device data on next use

// energy.f:12 for the context's current vlvr; writes eyel.. and coordinates
double engine_energy(Engine*& e, State& s);
// gradient.f:13-70 for nconf full variable vectors (nvt each): out_e[5 c ..]
// = eysm, eyel, eyvw, eyhb, eyvr, where eysm includes eysl with the analytic
// solvent as gradient.f:47 leaves it (eysl = eysm - the other four);
// out_g[nvt c ..] = dE/dv.  With itysol < 0 esolan's rotations draw from
// the context's twister, conformations in order.
void engine_gradient(Engine*& e, State& s, int nconf, const double* vlvr, double* out_e, double* out_g);
// Throws unless the context's model has derivatives here: vacuum (itysol 0;
// the reference stops for itysol > 0, gradient.f:59-62) and ECEPP (not
// FLEX).  The analytic solvent's device derivatives (k_esolan's gradient
// form, k_sol_project) are switched off: their first run faulted the GPU
// and no cause has been shown yet, so itysol < 0 is refused.
constexpr bool kSolventGradientOn = false;
inline void gradient_supported(const State& s) {
  if (s.itysol > 0) throw Error("gradient> Set itysol < 0");
  if (s.itysol < 0 && !kSolventGradientOn)
    throw Error("gradient> derivatives of the analytic solvent (opesol.f) are switched off on the device: set itysol "
                "to 0");
  if (s.itysol < 0 && s.ntlml > 1)
    throw Error("gradient> the analytic solvent's derivatives are provided for one chain (esolan.f numbers the atoms "
                "of a molecule from 1)");
  if (s.flex) throw Error("gradient> the FLEX force field (opeflx.f) is not provided");
}
// enysol.f:13 for the current coordinates
double engine_enysol(Engine*& e, State& s);
// metropolis.f:13 (context MT stream)
int engine_metropolis(Engine*& e, State& s, double* eol, double* acz, const smmp_weight& w);
// partem_s.f:13, serial-compatible single stream
int engine_partem_s(Engine*& e, State& s, int nrep, int nequi, int nswp, int switch_, const double* temps,
                    double* eol, int32_t* ipoi, double* traj, double* obs, const double* start, double* coor);
int engine_energy_batch(Engine*& e, std::vector<State*>& sys, const int32_t* conf_sys, int nconf,
                        const double* vlvr, int precision, double* out);
// Resident batch (smmp_batch_*): systems and conformations kept on the device.
struct EnergyBatch;
int engine_energy_batch_noalloc(std::vector<State*>& sys, const int32_t* conf_sys, int nconf, const double* vlvr,
                                int precision, double* out);
EnergyBatch* engine_batch_create(Engine*& e, std::vector<State*>& sys, const int32_t* conf_sys, int nconf);
void engine_batch_destroy(EnergyBatch* b);
void engine_batch_load(EnergyBatch* b, const double* vlvr, int64_t nvals);
void engine_batch_run(EnergyBatch* b, int precision);
int engine_batch_tune(EnergyBatch* b, int precision, int32_t* out);  // measured launch shapes; returns the classes
void engine_batch_read(EnergyBatch* b, double* out);
int64_t engine_batch_values(const EnergyBatch* b);  // doubles of vlvr the batch takes
int engine_sample(Engine*& e, State& s, const smmp_sampler_cfg& cfg, double* vlvr, double* eol, double* acc,
                  int32_t* ipoi, int32_t* hist, double* traj);
const char* engine_build_info();
void engine_profile(Engine*& e, int enable);
int engine_profile_read(Engine*& e, int kernel, double* ms, int64_t* launches, double* work);
int engine_profile_phases(Engine*& e, double* out);
int engine_sampler_info(Engine*& e, int32_t* out);
void engine_mlp_set(Engine*& e, const State& s, const float* theta);
void engine_mlp_get(Engine*& e, const State& s, float* theta);
double engine_mlp_grad(Engine*& e, const State& s, const double* vlvr, int nconf, const float* target, float* pred,
                       float* grad, int pairs);
void engine_mlp_fit(Engine*& e, const State& s, const double* vlvr, const float* target, int n, int epochs, int batch,
                    float lr, uint32_t seed, double* epoch_loss, int pairs);

// twister.f:104 grnd() on the context's stream (host side; the draws of the
// device samplers come from the same state copied to the device)
double host_grnd(State& s);

}  // namespace smmp
