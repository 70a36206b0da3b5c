// Device-ready description of one system of one or more chains, derived
// from the host State after init_molecule: the chain-build schedule, the
// interaction set as symmetric 2-plane bit rows (pairs within a chain from
// the reference's lists, every pair between chains), the torsion terms, the
// solvent atoms and, per variable, the atoms a torsion change moves.
#pragma once
#include <cstdint>
#include <vector>

#include "state.hpp"

namespace smmp {

constexpr int kTypes = mxtyat;  // 18 ECEPP atom types

struct HostSys {
  int nat = 0, nvt = 0, nlev = 0, words = 0, nheavy = 0;
  // chains (0-based atoms): first and last atom, second and third backbone
  // atoms (bldmol fnd3ba), heavy-index range [h1, h2), global parameters
  int nml = 0;
  int16_t mol_a1[mxml] = {}, mol_a2[mxml] = {}, mol_i2[mxml] = {}, mol_i3[mxml] = {};
  int16_t mol_h1[mxml] = {}, mol_h2[mxml] = {};
  double gbpr[6 * mxml] = {};
  double boxsize = 0;
  // C-alpha atoms (c_alfa, contacts.f:65-97) of the first residue of the first
  // chain and the last residue of the last chain, -1 if none: rgyr's ee
  int16_t ca_first = -1, ca_last = -1;
  std::vector<int16_t> mol_of;      // per atom: its chain
  // atoms (0-based)
  std::vector<int16_t> parent;      // iowat-1, -1 for the first atom of a chain
  std::vector<int16_t> lev_atoms;   // atoms in build order (first atoms excluded)
  std::vector<int32_t> lev_off;     // nlev+1 offsets into lev_atoms
  std::vector<int8_t> type;         // ityat-1
  std::vector<double> q, bl, cba, sba, toat0, tref, rvdw, sigma;
  std::vector<int16_t> tvar;        // variable (0-based) whose value sets this torsion, -1 if fixed
  // variables (0-based over all variables of the chain)
  std::vector<int16_t> var_atom;    // iatvr-1
  std::vector<int16_t> var_base;    // iowat(iatvr)-1: axis end, not moved
  std::vector<int16_t> var_axis0;   // iowat(base)-1: axis start
  std::vector<int16_t> var_m1, var_m2;  // atoms a trial move rotates, [m1, m2]
  std::vector<int16_t> var_h1, var_h2;  // their heavy atoms: heavy indices [h1, h2)
  std::vector<double> var_sign;         // +1: by the torsion change, -1: by its opposite
  std::vector<double> var_e0, var_sg, var_rn;  // torsion potential (e0to/2, sign, period)
  std::vector<int32_t> var_is_torsion;
  // interaction set: bits[(i*words + w)*2 + plane], plane 0 = pair counted,
  // plane 1 = 1-4 pair (half repulsion); symmetric in (i, j)
  std::vector<uint64_t> bits;
  int64_t npairs = 0, n14 = 0, ninter = 0;  // ninter: pairs between chains
  std::vector<int16_t> heavy;       // atoms with rvdw > 0
  std::vector<int16_t> heavy_of;    // per atom: its index in 'heavy', or -1
  // parameters, [ti*18 + tj]: A (aij or ahb), A14 (a14 or ahb), C (cij or chb), hb flag
  std::vector<double> pA, pA14, pC;
  std::vector<int32_t> pHB;
  double conv = 0, plt = 0, slp = 0;
  int epsd = 0, itysol = 0, npnt = 0;
  std::vector<double> spoint;       // npnt*3, point-major (x,y,z)
};

// Builds the description of every chain of s.  Throws Error when a list
// pair occurs twice or a moving set is not contiguous.
HostSys build_host_sys(const State& s);

}  // namespace smmp
