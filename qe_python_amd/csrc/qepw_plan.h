// Host-side plan of the sphere-pruned local-potential FFT pipeline.
//
// Device G order ("stick-major"): the plane waves of one k-point are
// re-ordered so that every z-stick (fixed x = m0 mod n1, y = m1 mod n2) owns
// a contiguous run of rows.  Every Davidson operation is invariant under a
// common row permutation, so vectors, Hpsi, beta and the diagonal all live
// in this order on the device; ``perm[d]`` is the reference row of device
// row d (reference order: ``qepy_pw/basis.py:1516-1547``).
//
// Intermediate "stick buffer" layout per band: work[z * nst + s], i.e.
// (n3, nst) with sticks contiguous, so that the xy-plane stage reads and
// writes one contiguous run per plane and the z stages write runs of
// ``group`` sticks.
#pragma once

#include <vector>
#include "qepw_common.h"

namespace qepw {

struct StickTables {
  int nst = 0;          // active sticks
  int nxcols = 0;       // distinct x among active sticks
  int group = 1;        // sticks per workgroup in the z stages
  int ngroups = 0;
  int64_t npoints = 0;  // rows addressed (npw, or N for the full grid)
  int* d_off = nullptr;    // nst+1 row offsets of each stick
  int* d_gpos = nullptr;   // npoints: (stick * n3 + z) of each row
  int* d_plane = nullptr;  // nst: x * n2 + y of each stick
  int* d_xcols = nullptr;  // nxcols: active x values
  int* d_stickq = nullptr;  // nst: c * n2 + y, c = rank of the stick's x in xcols
  int* d_colmap = nullptr;  // n1: rank of x among active columns, or -1
  int* d_colstart = nullptr;  // nxcols+1: first stick of each active column (sticks
                              // are sorted by x, then y)
  int* d_colpos = nullptr;  // nxcols * n2: stick at (column c, y), or -1
  unsigned slot_mask = 0;   // team-FFT input slots holding an active x or stick y
};

struct Plan {
  int n1 = 0, n2 = 0, n3 = 0;
  int64_t npw = 0;
  LinePlan line[3];
  StickTables sphere;   // the plane-wave sphere of this k-point
  StickTables full;     // every stick of the grid (prepare_potential)
  double2* d_tw64[3] = {nullptr, nullptr, nullptr};
  float2* d_tw32[3] = {nullptr, nullptr, nullptr};
  int* d_perm = nullptr;
  std::vector<int> h_perm;
  int device = 0;
};

}  // namespace qepw
