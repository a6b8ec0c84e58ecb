// Internal definitions shared by the HIP translation units of libpfw_mi355x.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
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
  double invL[3];   // 1 / L (the minimum-image rounding's fast path)
  double c, c2;
  double cs2;         // squared radius of the Verlet list (cutoff + skin, padded); c2 without a skin
  double reach[3];    // cell reach of the search radius (cutoff + skin)
  double sizecell[3];  // periodic axes: extent/dims; open axes: device-computed
  double origin[3];    // open axes: device-computed minimum bin coordinate
  int used[3];         // open axes: device-computed used dims (<= dims)
  // Verlet ghosts: the grid extended by the stencil radius on every periodic
  // axis; extended cells outside the real range hold images of real cells
  int ecell0;    // first extended cell of this configuration
  int edims[3];  // dims + 2 radius on periodic axes, dims on open axes
};

// Device buffers for one system.
struct SystemBuffers {
  int64_t n = 0;
  int64_t cap = 0;
  double* X = nullptr;        // 3 x n original order (AoS)
  int32_t* species = nullptr;  // n original order
  double* xw = nullptr;       // wrap records (x, y, z, species bits), [n][4], original order
  double* u = nullptr;        // bin coordinates, SoA [3][n], original order
  int32_t* cell_of = nullptr;  // n
  int32_t* count = nullptr;    // ncells
  int32_t* start = nullptr;    // ncells + 1
  int32_t* rank = nullptr;     // n: place of each atom among its cell's atoms (count-sort pass)
  int32_t* blocksum = nullptr;  // scan scratch
  int32_t* perm = nullptr;     // sorted -> original
  double* xs = nullptr;        // sorted wrapped coordinates SoA [3][n]
  int32_t* sps = nullptr;      // sorted species
  float4* xr = nullptr;        // sorted positions relative to their cell's first atom (FP32), species bits in w
  float4* cbox = nullptr;      // 2 per cell: min and max of its atoms' xr (x, y, z)
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

// Verlet list at cutoff + skin, reused while no atom has moved half the skin
// (the reference's GPUConfig(skin) connectivity cache, NeighbourList/gpu.jl:504-549).
// Periodic images are explicit ghost entries, so a list entry is a bare index
// into the extended position array xe: [0, n) the sorted atoms, [n, n + G) the
// ghosts in extended-cell order.
struct VerletDev {
  double skin = 0;              // 0: no list, fresh cells every step
  bool valid = false;
  uint64_t gen = 0;             // cell-build generation the list was made from
  int64_t ext_cells = 0;        // extended cells of all configurations
  int64_t n_ext = 0;            // n + ghosts
  int64_t ext_capacity = 0;     // allocated entries of xe / off / srcp
  int ecap = 0;                 // list slots per atom
  int64_t list_capacity = 0;    // allocated slots
  int parity = 0;               // which displacement word the next step accumulates into
  int64_t rebuilds = 0, reuses = 0;
  double last_max_disp = 0;     // largest displacement since the build, last step (Å)
  int32_t* ecfg0 = nullptr;     // ncfg + 1 extended-cell offsets
  int32_t* gcount = nullptr;    // ext_cells: ghosts of each extended cell
  int32_t* gstart = nullptr;    // ext_cells + 1
  int32_t* gblock = nullptr;
  int32_t* nbr = nullptr;       // ecap x n, column-major (slot-major): entry index into xe
  int32_t* ncount = nullptr;    // n: list length of each sorted atom
  double* off = nullptr;        // 4 x n_ext: build position minus caller position, species bits
  double* xe = nullptr;         // 4 x n_ext: current extended positions, species bits
  int32_t* srcp = nullptr;      // n_ext: original atom of each entry
  double* Xref = nullptr;       // 3 x n caller positions at the build (AoS, original order)
  unsigned long long* maxd2 = nullptr;  // 2 words: largest squared displacement (bits), double-buffered
  unsigned long long* h_maxd2 = nullptr;  // pinned copy
  int32_t* status = nullptr;    // 6: unused, largest degree, overflow, unused, list entries (uint64)
  int32_t* h_status = nullptr;  // pinned
  int64_t entries = 0;          // list entries of the last build
  double* part = nullptr;       // list kernel energy / virial records
  int64_t part_capacity = 0;
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
  int solid = 0;  // real solid harmonics (r^l Y_lm) instead of spherical
  // basis policy: species projection with orders order_lo..order_hi (proj), the
  // projected neighbour species as bits of sel, the neighbour species whose
  // pair term is kept as bits of pair_gate
  int proj = 0, order_lo = -1, order_hi = -1;
  uint32_t sel = 0, pair_gate = 0xffffffffu;
  double cutoff = 0;
  std::vector<int32_t> c_rows, c_cols;  // host copy of the coupling
  std::vector<double> c_vals;
  double* reference = nullptr;
  AceFamilyDev radial, pair;
  int32_t* prod = nullptr;      // nprod x 4, -1 padded
  std::vector<int32_t> h_prod;  // host copy of prod
  int nsteps = 0;               // product schedule: nsteps x 64 slots
  int order_step[5] = {0, 0, 0, 0, 0};  // steps [order_step[nu - 1], order_step[nu]) hold products of nu factors
  uint4* sched = nullptr;       // per slot: 4 x u16 factors, 4 x u16 accumulator targets
  int32_t* copyb = nullptr;     // ndens + 1: accumulator copies of each density
  int ntarget = 0;
  std::vector<int32_t> slot_prod;  // host: product of each slot, -1 = padding
  double* ws64 = nullptr;       // nz x slots readout weights in schedule order
  float* ws32 = nullptr;
  int32_t* dens = nullptr;      // ndens x 2
  int32_t* dperm = nullptr;     // densities ordered by chunks of radial indices
  int32_t* dchunk = nullptr;    // chunk offsets into dperm
  std::vector<int32_t> h_dens;  // host copy of dens: the density tiles are regrouped when rchunk changes
  int rchunk = 1;               // radial indices per chunk
  int32_t* rlb = nullptr;       // lmax + 2: radial indices grouped by l
  int32_t* rl = nullptr;        // (radial index, first density of its m-block)
  // matrix-core density tiles (FP32, unprojected models): groups of 16 tiles
  // of 4 radial indices x 4 harmonics, one group per MFMA; per group and lane
  // the radial row, the harmonic column and the density of each of the 4
  // accumulator registers (-1 none), [mgroups][6][64]
  int32_t* mtab = nullptr;
  int mgroups = 0;              // 0 when the model's tiles exceed the kernel's accumulators
  bool mdefault = false;        // tiles at least half useful: the built-in density contraction
  int mgchunk[4] = {0, 0, 0, 0};  // radial chunk of each group
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
  uint8_t* eseg = nullptr;    // n x 27: first slot of each stencil entry's edges in a site's row (radius-1 stencils)
  int32_t* eent = nullptr;    // stencil entry of each edge (a site's edges are sorted by entry, neighbour)
  double* eg = nullptr;       // 3 x slots edge gradients
  int32_t* h_status = nullptr;  // pinned copy of status after each environment build
  double* siteE = nullptr;    // n (sorted order)
  void* Abuf = nullptr;       // n x ndens densities, then dE/dA (T of the step)
  double* siteW = nullptr;    // 9 x n
  double* rec = nullptr;      // 10 x n reduction records
  int32_t* status = nullptr;
  // Verlet reuse (skin > 0): rows at cutoff + skin with the build's exact R;
  // every step compacts them into the rows above (edges inside the cutoff now)
  bool skin_valid = false;
  uint64_t skin_gen = 0;        // cell-build generation of the skin rows
  int ecap_s = 0;               // slots per site of the skin rows
  int64_t skin_cap = 0;         // allocated skin slots
  int skin_max_degree = 0;
  int compact_degree = 0;       // largest in-cutoff degree at the build
  int32_t* ecount_s = nullptr;  // n
  int32_t* enb_s = nullptr;     // n x ecap_s
  int32_t* eent_s = nullptr;
  int32_t* erev_s = nullptr;
  uint8_t* eseg_s = nullptr;    // n x 27
  double* eR_s = nullptr;       // 3 x slots, at the build
  int32_t* cslot = nullptr;     // skin slot -> compacted slot (-1: outside the cutoff now)
  int32_t* corig = nullptr;     // compacted slot -> skin slot
  double* Xref = nullptr;       // 3 x n caller positions at the build (AoS, original order)
  double* dX = nullptr;         // 3 x n displacement since the build, sorted SoA
  unsigned long long* maxd2 = nullptr;    // 2 words, as the pair list's
  unsigned long long* h_maxd2 = nullptr;  // pinned
  int parity = 0;
};

}  // namespace pfw

namespace pfw {
// Grow-only device buffers kept between calls (one per slot) for the large
// per-call arrays of the feature kernels; freed with the context.
struct GrowPool {
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  std::vector<Buf> slots;
  template <typename T>
  T* get(size_t slot, size_t count) {
    if (slots.size() <= slot) slots.resize(slot + 1);
    Buf& b = slots[slot];
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (b.bytes < bytes) {
      if (b.p) PFW_HIP(hipFree(b.p));
      b.p = nullptr;
      b.bytes = 0;
      PFW_HIP(hipMalloc(&b.p, bytes));
      b.bytes = bytes;
    }
    return (T*)b.p;
  }
  void release() {
    for (Buf& b : slots)
      if (b.p) (void)hipFree(b.p);
    slots.clear();
  }
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
  pfw::VerletDev verlet;
  pfw::GrowPool pool;
  uint64_t build_gen = 0;  // incremented by every cell build
  // outputs
  double* d_F = nullptr;        // 3 x n original order
  double* d_partials = nullptr;  // per-wave E + 6 W partials
  int64_t partial_cap = 0;
  double* d_red = nullptr;       // stage-1 reduction partials (64 x 16)
  double* d_cfg = nullptr;       // per-configuration sums (ncfg x 16)
  double* d_cfg_out = nullptr;   // per-configuration E, W[9], status (ncfg x 11)
  std::vector<double> h_cfg_out;
  double* d_Xstage = nullptr;    // 3 x n staging copy of caller device positions
  int32_t* d_flag = nullptr;     // validation flag of a device position update
  int32_t* h_flag = nullptr;     // pinned copy
  double* d_out = nullptr;       // E, W[9], status
  double* h_out = nullptr;       // pinned copy of d_out
  double* h_out_dev = nullptr;   // device address of h_out: the finish kernel writes it directly
  bool host_written = false;     // h_out already holds this step's outputs (no copy needed)
  uint32_t last_mask = 0;
  bool have_results = false;
  // pair-set audit (a diagnostic: per-atom digests of the pairs the
  // evaluating kernels accumulated)
  bool audit = false;
  int32_t* audit_count = nullptr;           // n, original order
  unsigned long long* audit_hash = nullptr;  // n, original order
  int64_t audit_cap = 0;
  bool audit_valid = false;                 // the digests belong to the last step
  // timing
  bool timing = false;
  hipEvent_t ev[pfw::kStages + 1] = {};
  double stage_ms[pfw::kStages] = {};
  // captured steps: the launches of one fresh (no-skin) pair or ACE step as an
  // instantiated HIP graph, replayed while only positions change
  bool graphs = false;
  hipGraphExec_t graph_exec = nullptr;
  int graph_kind = -1, graph_precision = -1;
  uint32_t graph_mask = 0;
  int64_t graph_captures = 0, graph_replays = 0;
  bool graph_warm = false;  // an eager step has run since the last reset (buffers sized)
  // launch-shape knobs (0 = the built-in choice): ACE sites per wavefront,
  // ACE radial indices per density tile, ACE product wavefronts per
  // workgroup, most lanes per atom in the pair kernel, ACE contractions (1
  // VALU, 2 densities on matrix cores, 3 densities and pullback)
  int tune[5] = {0, 0, 0, 0, 0};
};

namespace pfw {
// capi.hip: drop the captured step (every entry point that may change a
// buffer, a shape or a model calls it)
void graph_reset(pfw_ctx* ctx);
// Issue the launches of `issue` (stream work only: kernels and copies, no
// host synchronisation, no memset) on ctx->stream; with graphs on, through
// the graph captured for (kind, precision, mask), capturing it first.
void graph_issue(pfw_ctx* ctx, int kind, int precision, uint32_t mask, void (*issue)(pfw_ctx*, void*), void* arg);
// nl.hip
void nl_configure(pfw_ctx* ctx, double cutoff, double skin);
void batch_reduce(pfw_ctx* ctx, const double* records, int width, bool cells, int kind);
void nl_build(pfw_ctx* ctx);
// The buffers of one cell build (the context's, or views into a caller's
// scratch area).
struct BuildView {
  int64_t n;
  int ncells, ncfg;
  bool open_any;
  const double* X;
  double *xw, *u;
  int32_t *cell_of, *count, *rank;
  uint64_t* bounds;
  Grid* grids;
  const int32_t* atom_cfg;
  int32_t *start, *blocksum, *perm;
  double* xs;
  const int32_t* species;
  int32_t* sps;
  float4* xr;
  float4* cbox;
};
void nl_build_core(const BuildView& v, hipStream_t st, bool counts_zero, pfw_ctx* timing);
void fill_i32(int32_t* p, int64_t m, int32_t v, hipStream_t st);
// Host-side grid of one configuration (prepare!'s choices, storage.jl:130-193).
Grid nl_configure_one(const double* H, const int* pb, int64_t n, double c, double search, double limit);
// Directed pairs on the device, grouped by sorted atom (offsets off[0..n]);
// the destructor frees them.
struct DeviceExport {
  int64_t total = 0;
  int32_t* counts = nullptr;
  int64_t* off = nullptr;
  int32_t *i = nullptr, *j = nullptr, *S = nullptr;
  double* R = nullptr;
  void release();
  ~DeviceExport() { release(); }
};
void nl_export_device(pfw_ctx* ctx, DeviceExport& ex, bool with_S);
void nl_export_pairs(pfw_ctx* ctx, int32_t* i, int32_t* j, int32_t* S, double* R, int64_t cap,
                     int64_t* count);
void nl_free(pfw_ctx* ctx);
void scan_counts(pfw_ctx* ctx, const int32_t* count, int m, int32_t* start, int32_t* blocksum);
// E, W and the status word of one configuration from m records of `width`
// columns (kind 0: pair records, kind 1: ACE records) into d_out / d_cfg_out.
void reduce_outputs(pfw_ctx* ctx, const double* rec, int64_t m, int width, int kind, const int32_t* status);
// The same over explicit buffers: red holds 64 x 16 doubles, out and out2 11
// each (out3 may be NULL).
void reduce_core(const double* rec, int64_t m, int width, int kind, const int32_t* status, double* red, double* out,
                 double* out2, double* out3, hipStream_t st);
// d_out's 11 words into h_out on the stream, unless the finish kernel wrote them there.
void copy_outputs(pfw_ctx* ctx);
// pair.hip
void pair_run(pfw_ctx* ctx, uint32_t mask, int precision);
void pair_finish(pfw_ctx* ctx, uint32_t mask);
void record_stage(pfw_ctx* ctx, int k);
void pair_adjoint(pfw_ctx* ctx, const double* Ebar, const double* Fbar, const double* Wbar, double* gA);
void static_scratch_init(int device);
void pair_run_static(pfw_ctx* ctx, int kind, int nz, const double* params, double cutoff, int64_t n,
                     const double* X, const int32_t* species, const double* H, const uint8_t* pbc, uint32_t mask,
                     int precision, double* E, double* F, double* W);
// verlet.hip
void verlet_free(pfw_ctx* ctx);
void verlet_pair_run(pfw_ctx* ctx, uint32_t mask, int precision);
// ace.hip
void ace_set_model(pfw_ctx* ctx, const pfw_ace_model* m);
void ace_group_densities(pfw_ctx* ctx, int rchunk);
void ace_design(pfw_ctx* ctx, int precision, const double* basis_scale, double* out, int64_t ld);
void ace_set_parameters(pfw_ctx* ctx, const double* WB, const double* Wpair);
void ace_run(pfw_ctx* ctx, uint32_t mask, int precision);
void ace_free(pfw_ctx* ctx);
void ace_free_edges(pfw_ctx* ctx);
void ace_site_energies(pfw_ctx* ctx, double* out);
void ace_parameter_sets(pfw_ctx* ctx, int precision, int64_t nsets, const double* WB, const double* Wpair,
                        double* out);
void features_run(pfw_ctx* ctx, const pfw_feature_layer* layer, double* out);
void features_vjp_run(pfw_ctx* ctx, const pfw_feature_layer* layer, const double* cot, double* dx, double* dmixing,
                      double* dscale, double* dradial);
void ace_adjoint(pfw_ctx* ctx, int precision, const double* Ebar, const double* Fbar, const double* Wbar, double* gWB,
                 double* gWpair);
}  // namespace pfw
