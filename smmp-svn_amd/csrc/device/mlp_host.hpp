// Host functions of the surrogate network (device/mlp.hip).  The network's
// parameters and Adam moments live with the context (MlpWork).
#pragma once
#include <hip/hip_runtime.h>

#include "../host/state.hpp"
#include "mlp_common.hpp"

namespace smmp {

struct MlpWork;  // device workspace of the training passes and the network
void mlp_work_destroy(MlpWork* w);
// Residues of the system of s (all chains): the network's input covers them.
int mlp_residue_count(const State& s);
// Residue tables of the system of s on the device.
MlpResidues mlp_residues(MlpWork*& w, const State& s, hipStream_t st);
// Sets the network for nres residues (mlp_layout(nres).n values; resets the
// Adam moments) / reads it.
void mlp_set_theta(MlpWork*& w, hipStream_t st, const float* host, int nres);
void mlp_get_theta(MlpWork*& w, hipStream_t st, float* host, int nres);
// The network as the sampler's screen reads it (BF16 weight copies refreshed
// after training; enqueued on st).
MlpScreen mlp_screen(MlpWork*& w, const State& s, hipStream_t st);
// Predictions of nconf host conformations; with targets returns the mean
// squared error and writes its gradient (host, may be null).
// pairs: the conformations come in pairs (x, x') with targets E(x') - E(x).
double mlp_grad(MlpWork*& w, const State& s, hipStream_t st, const double* vlvr, int nconf, const float* target,
                float* pred, float* grad, int pairs);
void mlp_fit(MlpWork*& w, const State& s, hipStream_t st, const double* vlvr, const float* target, int n, int epochs,
             int batch, float lr, uint32_t seed, double* epoch_loss, int pairs);

}  // namespace smmp
