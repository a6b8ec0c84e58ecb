// Test infrastructure only: a CPU FP64 restatement of PotentialFrameWork.jl's
// cell-list preparation, used as the checker for the HIP path. Nothing in the
// product library links or calls this code.
#pragma once
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

struct ArgumentError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DomainError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Julia `mod(x::Float64, y::Float64)` (rem-based, result has the sign of y).
double jl_mod(double x, double y);
int64_t jl_fld(int64_t a, int64_t b);
int64_t jl_imod(int64_t a, int64_t b);

struct CellImage {
  int cell;
  double shift[3];  // Cartesian basis * lattice
  bool self;
  int64_t lattice[3];
};

// Restates `Workspace` after `prepare!` (src/submodules/NeighbourList/storage.jl:32-58,130-241).
struct Prepared {
  int64_t n = 0;
  double c = 0;
  bool minimum_image = true;
  bool pb[3] = {false, false, false};
  double L[3] = {1, 1, 1};
  double basis[9];  // column-major 3x3
  double recip[9];
  double reach[3], extent[3], origin[3], sizecell[3];
  int64_t dims[3];
  bool brute = false;
  std::vector<double> x, y, z;  // sorted, wrapped coordinates
  std::vector<int64_t> p, inv;  // p[k] = original index of sorted atom k
  std::vector<std::array<int64_t, 3>> keys;
  std::vector<std::pair<int64_t, int64_t>> ranges;  // [first, last] inclusive
  std::vector<std::vector<CellImage>> images;
  std::vector<std::vector<int>> adjacency, footprints, phases;
};

// storage.jl:83-108
void periodic_basis(const double* H, const bool* pb, double* B);
// StaticArrays 3x3 inverse by cross products (used by storage.jl:148).
void inv3(const double* A, double* out);

void prepare(Prepared& w, int64_t n, const double* X, const double* H, const bool* pb,
             double cutoff, int bruteforce_threshold, bool schedule);

struct DirectedPair {
  int64_t i, j;
  int64_t S[3];
  double R[3];
};
// interface.jl:116-143: every directed image in original order.
void directed_pairs(const Prepared& w, const double* X, const double* H,
                    std::vector<DirectedPair>& out);

}  // namespace oracle
