// minim.f on a context (device derivatives), single and batched.
#pragma once
#include <cstdint>

#include "state.hpp"

namespace smmp {

struct Engine;

constexpr int kGradEnergies = 5;  // eysm, eyel, eyvw, eyhb, eyvr per conformation

// minim(imin, maxit, eps) of minim.f:13-250 (ireg = 0): minimises from the
// context's variables; on return the context holds the best variables
// (wrapped into [-pi, pi]), the energies of energy() at the last point the
// minimiser evaluated and the derivatives there.  Returns the function
// calls used (maxit reached: the step limit).
int minim(Engine*& e, State& s, int imin, int maxit, double eps);

// nconf minimisations of the context's system from vlvr (nconf x nvt, full
// variable vectors; the fixed variables keep their values), side by side:
// each takes the same steps as minim would alone, and each round of calls
// is one device launch over all conformations still running.  On return
// vlvr holds each best point (free variables wrapped into [-pi, pi]),
// fout[c] its function value (eysm of gradient.f), nfun[c] its calls.
void minim_batch(Engine*& e, State& s, int imin, int nconf, double* vlvr, int maxit, double eps, double* fout,
                 int32_t* nfun);

}  // namespace smmp
