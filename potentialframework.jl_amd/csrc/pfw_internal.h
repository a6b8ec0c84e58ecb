// Internal definitions shared by the HIP translation units of libpfw_mi355x.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pfw_mi355x.h"

namespace pfw {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PFW_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::pfw::Error(PFW_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device allocation owned by a scope (temporary buffers of one call).
template <typename T>
struct DeviceArray {
  T* p = nullptr;
  explicit DeviceArray(size_t n) {
    if (n) PFW_HIP(hipMalloc(&p, n * sizeof(T)));
  }
  ~DeviceArray() {
    if (p) (void)hipFree(p);
  }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
};

// Cell grid of one system. Static fields are set on the host by pfw_set_system;
// for open (non-periodic) axes origin/sizecell/used dims are recomputed on the
// device every step from the atoms' extent.
struct Grid {
  int cell0;          // first global cell of this configuration
  int ncells;         // cells of this configuration (product of layout dims)
  int natoms;         // atoms of this configuration
  int minimum_image;  // storage.jl:140
  int pb[3];
  int open_any;
  int dims[3];     // layout dims (open axes: capacity)
  int radius[3];   // stencil half-width per axis
  int nstencil;    // (2r+1) product
  double basis[9];  // column-major periodic basis (storage.jl:83-108), identity if minimum image
  double recip[9];  // its inverse
  double L[3];      // diag(H) for the minimum-image path
  double c, c2;
  double reach[3];
  double sizecell[3];  // periodic axes: extent/dims; open axes: device-computed
  double origin[3];    // open axes: device-computed minimum bin coordinate
  int used[3];         // open axes: device-computed used dims (<= dims)
};

// Device buffers for one system.
struct SystemBuffers {
  int64_t n = 0;
  int64_t cap = 0;
  double* X = nullptr;        // 3 x n original order (AoS)
  int32_t* species = nullptr;  // n original order
  double* xw = nullptr;       // wrapped coordinates, SoA [3][n], original order
  double* u = nullptr;        // bin coordinates, SoA [3][n], original order
  int32_t* cell_of = nullptr;  // n
  int32_t* count = nullptr;    // ncells
  int32_t* start = nullptr;    // ncells + 1
  int32_t* cursor = nullptr;   // ncells
  int32_t* blocksum = nullptr;  // scan scratch
  int32_t* perm = nullptr;     // sorted -> original
  double* xs = nullptr;        // sorted wrapped coordinates SoA [3][n]
  int32_t* sps = nullptr;      // sorted species
  uint64_t* bounds = nullptr;  // 6 ordered-int min/max of open axes per configuration
  Grid* grids = nullptr;       // device copy, one per configuration
  int32_t* atom_cfg = nullptr;  // n: configuration of each atom (original order)
  int32_t* cell_cfg = nullptr;  // ncells: configuration of each global cell
  int32_t* cfg_cell0 = nullptr;  // ncfg + 1: global cell offsets per configuration
};

struct PairModelDev {
  int kind = -1;
  int nz = 0;
  double cutoff = 0;
  std::vector<double> params;  // host copy
  double* d_params = nullptr;  // SMOOTH: nz*nz amplitudes; LJ: eps, sigma
};

constexpr int kStages = 6;  // binning, scan, scatter+sort, environments, interaction, reduction
constexpr size_t kMaxLds = 160 * 1024;  // LDS per compute unit (gfx950)

// One radial family of a resolved ACE model on the device, in both precisions.
struct AceFamilyDev {
  int nr = 0, nnodes = 0, nch = 0;
  bool uniform = true;  // nodes are linspace(-1, 1, nnodes)
  int* pq = nullptr;    // nch x 2 (p, q)
  int* env_kind = nullptr;
  int* env_pow = nullptr;  // nch x 2 integer envelope powers
  double* par64 = nullptr;  // nch x 12: a rin r0 yin ycut rcut | envelope x1 x2 s (or rc r0 -)
  float* par32 = nullptr;
  double* nodes64 = nullptr;
  float* nodes32 = nullptr;
  double* coef64 = nullptr;  // nch x nint x nr x 4
  float* coef32 = nullptr;
};

struct AceDev {
  bool ready = false;
  int nz = 0, lmax = 0, ndens = 0, nprod = 0, nbasis = 0, ny = 0;
  double cutoff = 0;
  std::vector<int32_t> c_rows, c_cols;  // host copy of the coupling
  std::vector<double> c_vals;
  double* reference = nullptr;
  AceFamilyDev radial, pair;
  int32_t* prod = nullptr;      // nprod x 4, -1 padded
  int nsteps = 0;               // product schedule: nsteps x 64 slots
  uint4* sched = nullptr;       // per slot: 4 x u16 factors, 4 x u16 accumulator targets
  int32_t* copyb = nullptr;     // ndens + 1: accumulator copies of each density
  int ntarget = 0;
  std::vector<int32_t> slot_prod;  // host: product of each slot, -1 = padding
  double* ws64 = nullptr;       // nz x slots readout weights in schedule order
  float* ws32 = nullptr;
  int32_t* dens = nullptr;      // ndens x 2
  int32_t* rlb = nullptr;       // lmax + 2: radial indices grouped by l
  int32_t* rl = nullptr;        // (radial index, first density of its m-block)
  int max_degree = 0;           // largest edge count of the last environment build
  double* w64 = nullptr;        // nz x nprod precontracted readout
  float* w32 = nullptr;
  double* wp64 = nullptr;       // nz x nq
  float* wp32 = nullptr;
  // environments and per-step buffers
  int64_t edge_cap = 0;       // allocated edge slots
  int ecap = 0;               // slots per site (0: not sized yet)
  int32_t* ecount = nullptr;  // n: edges of each site
  int32_t* enb = nullptr;     // n x ecap: neighbour (sorted index)
  double* eR = nullptr;       // 3 x slots
  int32_t* erev = nullptr;    // slot of the reverse edge
  int32_t* eent = nullptr;    // stencil entry of each edge (a site's edges are sorted by entry, neighbour)
  double* eg = nullptr;       // 3 x slots edge gradients
  int32_t* h_status = nullptr;  // pinned copy of status after each environment build
  double* siteE = nullptr;    // n (sorted order)
  void* Abuf = nullptr;       // n x ndens densities, then dE/dA (T of the step)
  double* siteW = nullptr;    // 9 x n
  double* rec = nullptr;      // 10 x n reduction records
  int32_t* status = nullptr;
};

}  // namespace pfw

struct pfw_ctx {
  int device = 0;
  int num_cu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  std::string err;
  // geometry: a batch of ncfg independent configurations (1 for a single system)
  int ncfg = 0;
  int ncells = 0;        // all configurations
  bool open_any = false;
  double cutoff = 0;
  std::vector<pfw::Grid> grids;       // host copies
  std::vector<int64_t> atom0;         // ncfg + 1 atom offsets (original order)
  std::vector<double> host_H;         // 9 per configuration
  std::vector<int> host_pb;           // 3 per configuration
  pfw::SystemBuffers sys;
  bool have_system = false;
  pfw::PairModelDev pair;
  pfw::AceDev ace;
  // outputs
  double* d_F = nullptr;        // 3 x n original order
  double* d_partials = nullptr;  // per-wave E + 6 W partials
  int64_t partial_cap = 0;
  double* d_red = nullptr;       // stage-1 reduction partials (64 x 16)
  double* d_cfg = nullptr;       // per-configuration sums (ncfg x 16)
  double* d_cfg_out = nullptr;   // per-configuration E, W[9], status (ncfg x 11)
  std::vector<double> h_cfg_out;
  double* d_out = nullptr;       // E, W[9], status
  double* h_out = nullptr;       // pinned copy of d_out
  uint32_t last_mask = 0;
  bool have_results = false;
  // timing
  bool timing = false;
  hipEvent_t ev[pfw::kStages + 1] = {};
  double stage_ms[pfw::kStages] = {};
};

namespace pfw {
// nl.hip
void nl_configure(pfw_ctx* ctx, double cutoff);
void batch_reduce(pfw_ctx* ctx, const double* records, int width, bool cells, int kind);
void nl_build(pfw_ctx* ctx);
void nl_export_pairs(pfw_ctx* ctx, int32_t* i, int32_t* j, int32_t* S, double* R, int64_t cap,
                     int64_t* count);
void nl_free(pfw_ctx* ctx);
void scan_counts(pfw_ctx* ctx, const int32_t* count, int m, int32_t* start, int32_t* blocksum,
                 int32_t* cursor);
// pair.hip
void pair_run(pfw_ctx* ctx, uint32_t mask, int precision);
void record_stage(pfw_ctx* ctx, int k);
void pair_adjoint(pfw_ctx* ctx, const double* Ebar, const double* Fbar, const double* Wbar, double* gA);
// ace.hip
void ace_set_model(pfw_ctx* ctx, const pfw_ace_model* m);
void ace_set_parameters(pfw_ctx* ctx, const double* WB, const double* Wpair);
void ace_run(pfw_ctx* ctx, uint32_t mask, int precision);
void ace_free(pfw_ctx* ctx);
void ace_free_edges(pfw_ctx* ctx);
void ace_site_energies(pfw_ctx* ctx, double* out);
void ace_adjoint(pfw_ctx* ctx, int precision, const double* Ebar, const double* Fbar, const double* Wbar, double* gWB,
                 double* gWpair);
}  // namespace pfw
