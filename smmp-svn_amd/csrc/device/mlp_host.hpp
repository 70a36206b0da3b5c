// Host functions of the surrogate network (device/mlp.hip).  The network's
// parameters and Adam moments live with the context (MlpWork).
#pragma once
#include <hip/hip_runtime.h>

#include "../host/state.hpp"
#include "mlp_common.hpp"

namespace smmp {

struct MlpWork;  // device workspace of the training passes and the network
void mlp_work_destroy(MlpWork* w);
// Residue-window tables of the system of s on the device.
MlpResidues mlp_residues(MlpWork*& w, const State& s, hipStream_t st);
void mlp_set_theta(MlpWork*& w, hipStream_t st, const float* host);  // resets the Adam moments
void mlp_get_theta(MlpWork*& w, hipStream_t st, float* host);
const float* mlp_theta_dev(MlpWork*& w);
// Predictions of nconf host conformations; with targets returns the mean
// squared error and writes its gradient (host, may be null).
// pairs: the conformations come in pairs (x, x') with targets E(x') - E(x).
double mlp_grad(MlpWork*& w, const State& s, hipStream_t st, const double* vlvr, int nconf, const float* target,
                float* pred, float* grad, int pairs);
void mlp_fit(MlpWork*& w, const State& s, hipStream_t st, const double* vlvr, const float* target, int n, int epochs,
             int batch, float lr, uint32_t seed, double* epoch_loss, int pairs);

}  // namespace smmp
