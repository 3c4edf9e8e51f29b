// host.hpp -- host-side problem setup and direct-plugin factorization.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pdos.h"

namespace pdos_amd {

// Scaled problem as the ADMM iteration sees it (common.h:35-218).
struct Setup {
  int n = 0, m = 0, f = 0, l = 0;
  std::vector<int> soc_ptr;        // absolute row offsets of each SOC, size nsoc+1
  // A' rows = A columns (CSC of A), int32 indices, scaled values
  std::vector<int> cp, ci;
  std::vector<double> cx;
  // A rows (CSR of A) -- the reference's transpose At (cs.c:146-168)
  std::vector<int> rp, ri;
  std::vector<double> rx;
  std::vector<double> b, c, D, E;  // scaled b, c; D[m], E[n] (E includes x1e3)
  std::vector<double> E0;          // E before the x1e3 (1 without NORMALIZE)
  double lambda = 0, pscale = 0, dscale = 0;
  std::vector<double> x, s, y;     // scaled warm start (zeros if absent)
  // Setup on the device (build_setup_raw): cx, b, c, x, s, y hold the
  // caller's unscaled values, rx is empty, csr_src maps each entry of the
  // row form (rp, ri) to its CSC entry, and D, E, E0, lambda, pscale and
  // dscale are formed by the device (batch_setup.hip) and filled in after.
  bool dev = false;
  int normalize = 0;
  std::vector<int> csr_src;
  int nsoc() const { return (int)soc_ptr.size() - 1; }
  long nnz() const { return (long)cx.size(); }
};

// Fills `S` from the caller's data without writing to it.  Returns false and
// sets `err` on invalid input (NULL arrays, cone sizes that do not sum to m,
// indices out of range, sizes beyond int32).  `equil` selects the scaling:
// 0 common.h:35-218, 1 ruiz_equil_common.h:98-322.
bool build_setup(const Data* d, const Cone* k, Setup& S, std::string& err, int equil = 0);

// The same checks and the caller's values unscaled, with the row pattern of
// A and its map to the CSC entries (the stable transpose's order), for the
// scaling on the device (common.h:69-215 there, bit for bit).
bool build_setup_raw(const Data* d, const Cone* k, Setup& S, std::string& err);

// LDL' factorization of the permuted quasi-definite KKT [[-I, A'], [A, I]].
struct Factor {
  int N = 0;
  std::vector<int> P;              // P[k] = original index of pivot k
  std::vector<int> Lp, Li;         // L strictly lower, CSC (columns)
  std::vector<double> Lx, Ld;      // values, diagonal D
  std::vector<int> Rp, Ri;         // L strictly lower, CSR (rows) for the forward sweep
  std::vector<double> Rx;
  std::vector<int> fl_ptr, fl_rows;  // forward level sets (rows)
  std::vector<int> bl_ptr, bl_rows;  // backward level sets (columns)
  int nfl() const { return (int)fl_ptr.size() - 1; }
  int nbl() const { return (int)bl_ptr.size() - 1; }
};

// Builds the KKT of the scaled A, orders it with the approximate minimum
// degree ordering (or uses `perm` if non-empty), factors it (left-looking
// LDL', no pivoting: the KKT is quasi-definite), and derives level sets.
bool build_factor(const Setup& S, const std::vector<int>& perm, Factor& F, std::string& err);

// The fill-reducing ordering build_factor uses for S's KKT matrix.
bool kkt_order(const Setup& S, std::vector<int>& perm, std::string& err);

// Approximate minimum degree ordering (ordering.cpp) of the symmetric
// pattern whose upper (or lower, or full) triangle is the CSC (Kp, Ki) of
// order N; perm[k] = node eliminated k-th.
bool amd_order(int N, const std::vector<int>& Kp, const std::vector<int>& Ki, std::vector<int>& perm,
               std::string& err);

// amd_order through a per-process store of the orderings of recently seen
// patterns (the ordering depends on the pattern only).
bool amd_order_cached(int N, const std::vector<int>& Kp, const std::vector<int>& Ki, std::vector<int>& perm,
                      std::string& err);

}  // namespace pdos_amd
