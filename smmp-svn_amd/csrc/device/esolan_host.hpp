// Host side of the analytic solvent term (itysol < 0, device/esolan.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "../host/devtopo.hpp"
#include "common.hpp"

namespace smmp {

struct EsolanWork;  // device workspace (grown on demand, kept by the engine)
void esolan_work_destroy(EsolanWork* w);

// Adds the analytic solvent energy of nconf conformations to their energy
// rows: d_out[c * out_stride + 6] = sum_i sigma_i A_i (esolan.f:1006, 1040)
// and the total d_out[c * out_stride + 0] = (eysm + eysl) + eysmi.
//   d_xyz: coordinates of conformation c at d_xyz + 3 * stride * c (x, y, z
//          per atom); conf_sys / d_conf_sys: its system (host / device copy).
//   draw(c): the next uniform number of the twister that conformation c's
//          energy draws from (esolan.f:199-200 draws on a degenerate north
//          point, in atom order, conformations in order).
//   d_gradan (may be null): also the Cartesian gradient of each
//          conformation's term, esolan.f's gradan (sum over atoms i of
//          sigma_i dA_i/dr_j) at d_gradan + 3 (stride c + j).
// Throws Error on coinciding atom centres (esolan.f:127-133 stops there) and
// on more than 191 neighbours of one atom.
void esolan_energies(EsolanWork*& W, hipStream_t st, const DevSys* d_systems, const std::vector<const HostSys*>& hsys,
                     const std::vector<int32_t>& conf_sys, const int32_t* d_conf_sys, int nconf, int stride,
                     const double* d_xyz, double* d_out, int out_stride, const std::function<double(int)>& draw,
                     double* d_gradan = nullptr);

// Kernel time of the last esolan_energies call (ms, both passes and the
// sums) and its work counters, summed over all conformations: atom pairs
// scanned for neighbours, circle pairs classified, arcs tested, circles
// tested against arc midpoints.
struct EsolanStats {
  double ms = 0.0;
  double work[4] = {0.0, 0.0, 0.0, 0.0};
  int passes = 0;
};
EsolanStats esolan_last_stats(EsolanWork* W);

}  // namespace smmp
