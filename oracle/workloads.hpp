// Synthetic workloads of the oracle (test infrastructure only). See workloads.cpp.
#pragma once

#include <array>

#include "oracle.hpp"

namespace oracle {

// xoshiro256** seeded through splitmix64 from (seed, stream), so every platform draws the same
// numbers; uniform() is the top 53 bits scaled to [0, 1); normal() is Box-Muller's cosine
// branch.
struct Rng {
  std::array<uint64_t, 4> state;
  Rng(uint64_t seed, uint64_t stream);
  uint64_t next();
  double uniform();
  double normal();
  size_t below(size_t n);
};

struct Grid {
  size_t buses = 0;
  std::vector<double> bx, by;
  std::vector<size_t> from, to;  // from < to, branches sorted by (from, to)
  std::vector<double> x;         // reactance, pu
};

Grid make_grid(size_t n, uint64_t seed);
std::vector<bool> bridges(const Grid& g);

struct DcMatrices {
  Compressed bprime;  // CSR, (N-1) x (N-1): Bbus over buses 1..N-1
  Compressed bf;      // CSR, branches x (N-1): Bf over the same unknowns
  std::vector<size_t> branch_rows;
};
DcMatrices make_dc(const Grid& g, const std::vector<bool>* in_service);

struct AcJacobian {
  Compressed jacobian;  // CSC
  std::vector<size_t> pv, pq;
  std::vector<double> vm, va;
  std::vector<double> r, bc;  // branch resistance and line charging, pu
};
// `shift` > 0 moves the operating point as a Newton iterate would (every |V| and every non-slack
// angle by shift * U[-1, 1], drawn from `shift_seed`), keeping the branches and the PV/PQ split,
// so the Jacobian keeps its pattern
AcJacobian make_jacobian(const Grid& g, uint64_t seed, double shift = 0.0, uint64_t shift_seed = 0);

}  // namespace oracle
