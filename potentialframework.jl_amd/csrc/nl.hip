// Cell-list build on gfx950: FP64 wrap and binning, counting sort by cell
// (histogram, exclusive scan, scatter, per-cell ordering by original index),
// and the canonical directed-pair export used for parity.
//
// Reference semantics: src/submodules/NeighbourList/storage.jl:83-283 (wrap,
// reach, cell dims, image stencil with shift fld(key+off, dims)) and
// interface.jl:128-171 (directed pairs with shifts relative to the original
// coordinates). The grid may be coarser than the reference's (it is capped for
// huge or open systems) because the pair set does not depend on the grid, only
// on the exact membership test r^2 < c^2 evaluated below in FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pfw_device.h"

namespace pfw {

// ---------------------------------------------------------------- host setup

static void host_periodic_basis(const double* H, const int* pb, double* B) {
  double orth[3][3];
  int no = 0;
  for (int i = 0; i < 9; ++i) B[i] = 0;
  auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  for (int a = 0; a < 3; ++a) {
    if (!pb[a]) continue;
    double v[3] = {H[3 * a], H[3 * a + 1], H[3 * a + 2]};
    for (double t : v)
      if (!std::isfinite(t)) throw Error(PFW_INVALID_ARGUMENT, "periodic cell vectors must be finite");
    for (int k = 0; k < 3; ++k) B[3 * a + k] = v[k];
    for (int q = 0; q < no; ++q) {
      double s = dot(orth[q], v);
      for (int k = 0; k < 3; ++k) v[k] -= s * orth[q][k];
    }
    double nrm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(std::isfinite(nrm) && nrm > 0))
      throw Error(PFW_INVALID_ARGUMENT, "periodic cell vectors must be linearly independent");
    for (int k = 0; k < 3; ++k) orth[no][k] = v[k] / nrm;
    ++no;
  }
  for (int a = 0; a < 3; ++a) {
    if (pb[a]) continue;
    double best[3] = {0, 0, 0}, bestn = -1;
    for (int k = 0; k < 3; ++k) {
      double v[3] = {0, 0, 0};
      v[k] = 1;
      for (int q = 0; q < no; ++q) {
        double s = dot(orth[q], v);
        for (int m = 0; m < 3; ++m) v[m] -= s * orth[q][m];
      }
      double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      if (n2 > bestn) {
        bestn = n2;
        for (int m = 0; m < 3; ++m) best[m] = v[m];
      }
    }
    double nrm = std::sqrt(best[0] * best[0] + best[1] * best[1] + best[2] * best[2]);
    for (int m = 0; m < 3; ++m) {
      best[m] /= nrm;
      B[3 * a + m] = best[m];
      orth[no][m] = best[m];
    }
    ++no;
  }
}

// 3x3 inverse by cross products (the StaticArrays form used by storage.jl:148).
static void host_inv3(const double* A, double* out) {
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  double x0[3] = {A[0], A[1], A[2]}, x1[3] = {A[3], A[4], A[5]}, x2[3] = {A[6], A[7], A[8]};
  double y0[3], y1[3], y2[3];
  cross(x1, x2, y0);
  double d = x0[0] * y0[0] + x0[1] * y0[1] + x0[2] * y0[2];
  for (int k = 0; k < 3; ++k) {
    x0[k] /= d;
    y0[k] /= d;
  }
  cross(x2, x0, y1);
  cross(x0, x1, y2);
  for (int col = 0; col < 3; ++col) {
    out[3 * col + 0] = y0[col];
    out[3 * col + 1] = y1[col];
    out[3 * col + 2] = y2[col];
  }
}

template <typename T>
static void dmalloc(T** p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  PFW_HIP(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
}

void nl_free(pfw_ctx* ctx) {
  SystemBuffers& s = ctx->sys;
  void* ptrs[] = {s.X,     s.species, s.xw,     s.u,      s.cell_of,  s.count,    s.start,    s.rank,
                  s.blocksum, s.perm, s.xs,    s.sps,    s.bounds, s.grids, s.atom_cfg, s.cell_cfg, s.cfg_cell0, s.xr,
                  s.cbox};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  s = SystemBuffers();
}

// Grid of one configuration: mirrors prepare!'s choice of wrap path, basis,
// reach and dims (storage.jl:130-193), with the cell count capped by `limit`.
// The wrap path follows the cutoff c; cells are sized for the search radius
// (c plus the Verlet skin), so a stencil reaches every pair of the list.
Grid nl_configure_one(const double* H, const int* pb, int64_t n, double c, double search, double limit) {
  Grid g{};
  bool axisaligned = true;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (i != j && H[i + 3 * j] != 0) axisaligned = false;
  double L[3] = {H[0], H[4], H[8]};
  bool anyp = pb[0] || pb[1] || pb[2];
  bool mi = !anyp;
  if (anyp && axisaligned) {
    mi = true;
    for (int a = 0; a < 3; ++a)
      if (pb[a] && !(std::isfinite(L[a]) && L[a] > 2 * c)) mi = false;
  }
  g.minimum_image = mi;
  for (int a = 0; a < 3; ++a) {
    g.pb[a] = pb[a];
    g.L[a] = L[a];
    g.invL[a] = 1.0 / L[a];
  }
  if (mi) {
    for (int k = 0; k < 9; ++k) g.basis[k] = g.recip[k] = (k % 4 == 0) ? 1.0 : 0.0;
  } else {
    host_periodic_basis(H, pb, g.basis);
    host_inv3(g.basis, g.recip);
    for (double v : g.recip)
      if (!std::isfinite(v)) throw Error(PFW_INVALID_ARGUMENT, "periodic cell vectors must be linearly independent");
  }
  g.c = c;
  g.c2 = c * c;
  // list radius: an absolute pad keeps rounding of the build-time geometry
  // from dropping a pair that the displacement bound promises to keep
  g.cs2 = search > c ? (search + 1e-6) * (search + 1e-6) : g.c2;
  double extent[3];
  int nper = 0, nopen = 0;
  for (int a = 0; a < 3; ++a) {
    if (mi) {
      g.reach[a] = search;
      extent[a] = L[a];
    } else {
      double r0 = g.recip[a], r1 = g.recip[a + 3], r2 = g.recip[a + 6];
      g.reach[a] = search * std::sqrt(r0 * r0 + r1 * r1 + r2 * r2);
      extent[a] = 1.0;
    }
    if (pb[a])
      ++nper;
    else
      ++nopen;
  }
  // periodic axes: reach-sized cells (storage.jl:192), capped in count
  double dimsd[3] = {1, 1, 1};
  for (int a = 0; a < 3; ++a)
    if (pb[a]) dimsd[a] = std::max(1.0, std::floor(extent[a] / g.reach[a]));
  double prod = dimsd[0] * dimsd[1] * dimsd[2];
  if (prod > limit) {
    double f = std::pow(limit / prod, 1.0 / std::max(nper, 1));
    for (int a = 0; a < 3; ++a)
      if (pb[a]) dimsd[a] = std::max(1.0, std::floor(dimsd[a] * f));
  }
  double perprod = dimsd[0] * dimsd[1] * dimsd[2];
  int opencap = nopen ? (int)std::min(1024.0, std::max(1.0, std::ceil(std::pow(limit / perprod, 1.0 / nopen))))
                      : 1;
  for (int a = 0; a < 3; ++a) {
    if (pb[a]) {
      g.dims[a] = (int)dimsd[a];
      g.sizecell[a] = extent[a] / g.dims[a];
      // minimum image without a skin: L > 2c, so cells of at least c need one ring
      g.radius[a] = (mi && search == c) ? 1 : std::max(1, (int)std::ceil(g.reach[a] / g.sizecell[a]));
      g.used[a] = g.dims[a];
      g.origin[a] = 0;
    } else {
      g.dims[a] = opencap;
      g.radius[a] = 1;
      g.used[a] = opencap;
      g.sizecell[a] = g.reach[a];
      g.origin[a] = 0;
    }
  }
  g.open_any = nopen > 0;
  int64_t nc = (int64_t)g.dims[0] * g.dims[1] * g.dims[2];
  if (nc > (int64_t)1 << 30) throw Error(PFW_INVALID_ARGUMENT, "cell grid too large");
  g.ncells = (int)nc;
  g.natoms = (int)n;
  g.nstencil = (2 * g.radius[0] + 1) * (2 * g.radius[1] + 1) * (2 * g.radius[2] + 1);
  if (g.nstencil > 4096)
    throw Error(PFW_INVALID_ARGUMENT, "cutoff spans too many periodic images for the cell stencil");
  for (int a = 0; a < 3; ++a) g.edims[a] = pb[a] ? g.dims[a] + 2 * g.radius[a] : g.dims[a];
  return g;
}

// Grids of every configuration of the uploaded batch, packed one after the
// other in a single global cell index space.
void nl_configure(pfw_ctx* ctx, double c, double skin) {
  const int ncfg = ctx->ncfg;
  const int64_t n = ctx->sys.n;
  ctx->grids.resize(ncfg);
  int64_t total = 0, ext_total = 0;
  bool open_any = false;
  std::vector<int32_t> cell0(ncfg + 1, 0), atom_cfg(n);
  for (int k = 0; k < ncfg; ++k) {
    const int64_t nk = ctx->atom0[k + 1] - ctx->atom0[k];
    const double limit = ncfg == 1 ? std::max<double>(2.0 * (double)nk, 4096.0) : std::max<double>(2.0 * (double)nk, 64.0);
    Grid g = nl_configure_one(&ctx->host_H[9 * k], &ctx->host_pb[3 * k], nk, c, c + skin, limit);
    g.cell0 = (int)total;
    g.ecell0 = (int)ext_total;
    total += g.ncells;
    ext_total += (int64_t)g.edims[0] * g.edims[1] * g.edims[2];
    if (ext_total > ((int64_t)1 << 31) - 2) throw Error(PFW_INVALID_ARGUMENT, "extended cell grid too large");
    if (total > ((int64_t)1 << 31) - 2) throw Error(PFW_INVALID_ARGUMENT, "cell grid too large");
    cell0[k + 1] = (int32_t)total;
    open_any |= g.open_any != 0;
    ctx->grids[k] = g;
    for (int64_t i = ctx->atom0[k]; i < ctx->atom0[k + 1]; ++i) atom_cfg[i] = k;
  }
  ctx->ncells = (int)total;
  ctx->open_any = open_any;
  ctx->cutoff = c;
  ctx->verlet.ext_cells = ext_total;
  ctx->verlet.valid = false;
  std::vector<int32_t> cell_cfg(total);
  for (int k = 0; k < ncfg; ++k)
    for (int32_t q = cell0[k]; q < cell0[k + 1]; ++q) cell_cfg[q] = k;

  SystemBuffers& s = ctx->sys;
  dmalloc(&s.xw, 4 * n);  // wrap records: x, y, z, species bits
  dmalloc(&s.u, 3 * n);
  dmalloc(&s.cell_of, n);
  dmalloc(&s.perm, n);
  dmalloc(&s.xs, 3 * n);
  dmalloc(&s.sps, n);
  dmalloc(&s.xr, n);
  dmalloc(&s.cbox, 2 * total);
  dmalloc(&s.count, total);
  dmalloc(&s.start, total + 1);
  dmalloc(&s.rank, n);
  dmalloc(&s.blocksum, (total + 1023) / 1024 + 1);
  dmalloc(&s.bounds, 6 * (size_t)ncfg);
  dmalloc(&s.grids, ncfg);
  dmalloc(&s.atom_cfg, n);
  dmalloc(&s.cell_cfg, total);
  dmalloc(&s.cfg_cell0, ncfg + 1);
  hipStream_t st = ctx->stream;
  PFW_HIP(hipMemcpyAsync(s.grids, ctx->grids.data(), ncfg * sizeof(Grid), hipMemcpyHostToDevice, st));
  PFW_HIP(hipMemsetAsync(s.count, 0, sizeof(int32_t) * std::max<int64_t>(total, 1), st));  // later builds leave it zero
  if (n > 0) PFW_HIP(hipMemcpyAsync(s.atom_cfg, atom_cfg.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (total > 0)
    PFW_HIP(hipMemcpyAsync(s.cell_cfg, cell_cfg.data(), total * sizeof(int32_t), hipMemcpyHostToDevice, st));
  PFW_HIP(hipMemcpyAsync(s.cfg_cell0, cell0.data(), (ncfg + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
  PFW_HIP(hipStreamSynchronize(st));  // host staging vectors go out of scope
}

// ---------------------------------------------------------------- kernels

__device__ inline int bin_atom(const Grid& g, const double* u) {
  int k[3];
  for (int a = 0; a < 3; ++a) {
    double t = g.pb[a] ? u[a] / g.sizecell[a] : (u[a] - g.origin[a]) / g.sizecell[a];
    int lim = g.pb[a] ? g.dims[a] : g.used[a];
    double fl = floor(t);
    int v = fl < 0 ? 0 : (fl > (double)(lim - 1) ? lim - 1 : (int)fl);
    k[a] = v;
  }
  return cell_index(g, k);
}

// count[c] += 1 for every active lane, with one atomic per distinct cell of the
// wavefront (lanes of a wavefront often share a cell), all issued together.
// Returns the lane's rank among the atoms counted into its cell: the count
// before this wavefront's atomic plus the lane's place among the wavefront's
// lanes of that cell. The ranks of a cell are 0 .. count - 1, in an order that
// depends on scheduling; the cell sort fixes the order afterwards.
__device__ inline int wave_rank_count(int32_t* count, int c) {
  const int lane = (int)(threadIdx.x % kWave);
  const uint64_t below = (1ull << lane) - 1;
  uint64_t todo = __ballot(true);
  int leader = 0, before = 0, group = 0;
  while (todo) {  // group the active lanes by cell; the loop control is uniform
    const int first = __builtin_ctzll(todo);
    const int cl = __builtin_amdgcn_readlane(c, first);
    const uint64_t same = __ballot(c == cl) & todo;
    if (c == cl) {
      leader = first;
      before = __popcll(same & below);
      group = __popcll(same);
    }
    todo &= ~same;
  }
  int old = 0;
  if (lane == leader) old = atomicAdd(&count[c], group);
  return __shfl(old, leader, kWave) + before;
}

// Wrap atom i with its configuration's basis; when no configuration has an
// open axis also bin and count it. The wrapped position goes out as one
// 32-byte record (x, y, z, species bits), so that the cell sort gathers one
// record per atom.
__device__ inline void wrap_one(const Grid& g, int cfg, int64_t i, int64_t n, const double* __restrict__ X,
                                const int32_t* __restrict__ species, double* __restrict__ xw,
                                double* __restrict__ u, int32_t* __restrict__ cell_of, int32_t* __restrict__ count,
                                int32_t* __restrict__ rank, unsigned long long* __restrict__ bounds, int open_any) {
  double x[3] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]}, w[3], uu[3];
  wrap_atom(g, x, w, uu);
  double2* rec = (double2*)(xw + 4 * i);
  rec[0] = make_double2(w[0], w[1]);
  rec[1] = make_double2(w[2], __longlong_as_double((long long)species[i]));
  if (!open_any) {  // bin coordinates are needed later only for open axes
    const int c = bin_atom(g, uu);
    cell_of[i] = c;
    rank[i] = wave_rank_count(count, c);
    return;
  }
  for (int a = 0; a < 3; ++a) u[a * n + i] = uu[a];
  if (g.open_any) {
    for (int a = 0; a < 3; ++a) {
      if (g.pb[a]) continue;
      atomicMin(&bounds[6 * cfg + 2 * a], ordered_bits(uu[a]));
      atomicMax(&bounds[6 * cfg + 2 * a + 1], ordered_bits(uu[a]));
    }
  }
}

// single: one configuration, whose grid is then read with scalar loads (no
// dependence on a per-atom load).
__global__ void k_wrap(const Grid* __restrict__ grids, const int32_t* __restrict__ atom_cfg, int64_t n,
                       const double* __restrict__ X, const int32_t* __restrict__ species, double* __restrict__ xw,
                       double* __restrict__ u, int32_t* __restrict__ cell_of, int32_t* __restrict__ count,
                       int32_t* __restrict__ rank, unsigned long long* __restrict__ bounds, int open_any, int single) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (single) {
    wrap_one(grids[0], 0, i, n, X, species, xw, u, cell_of, count, rank, bounds, open_any);
  } else {
    const int cfg = atom_cfg[i];
    wrap_one(grids[cfg], cfg, i, n, X, species, xw, u, cell_of, count, rank, bounds, open_any);
  }
}

// Open axes: grid origin and cell size of every configuration from its atoms'
// extent (sparse-vacuum handling of storage.jl:187-197 with a bounded dense grid).
__global__ void k_open_grid(Grid* __restrict__ grids, const unsigned long long* __restrict__ bounds, int ncfg) {
  const int cfg = blockIdx.x * blockDim.x + threadIdx.x;
  if (cfg >= ncfg) return;
  Grid& g = grids[cfg];
  for (int a = 0; a < 3; ++a) {
    if (g.pb[a]) continue;
    if (g.natoms == 0) {
      g.origin[a] = 0;
      g.sizecell[a] = g.reach[a];
      g.used[a] = 1;
      continue;
    }
    const double lo = from_ordered_bits(bounds[6 * cfg + 2 * a]), hi = from_ordered_bits(bounds[6 * cfg + 2 * a + 1]);
    const double span = hi - lo;
    double sc = g.reach[a];
    const double nominal = floor(span / sc) + 1.0;
    if (nominal > (double)g.dims[a]) sc = fmax(sc, span / (double)g.dims[a] * (1.0 + 1e-12));
    const int used = (int)fmin((double)g.dims[a], floor(span / sc) + 1.0);
    g.origin[a] = lo;
    g.sizecell[a] = sc;
    g.used[a] = used < 1 ? 1 : used;
  }
}

__global__ void k_bin_open(const Grid* __restrict__ grids, const int32_t* __restrict__ atom_cfg, int64_t n,
                           const double* __restrict__ u, int32_t* __restrict__ cell_of, int32_t* __restrict__ count,
                           int32_t* __restrict__ rank) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Grid& g = grids[atom_cfg[i]];
  double uu[3] = {u[i], u[n + i], u[2 * n + i]};
  const int c = bin_atom(g, uu);
  cell_of[i] = c;
  rank[i] = wave_rank_count(count, c);
}

// Exclusive scan of cell counts: per-block totals, a one-block scan of the
// totals, then per-block rescans with offsets. 256 threads x 4 cells per block.
constexpr int kScanItems = 4;
constexpr int kScanThreads = 256;
constexpr int kScanBlock = kScanItems * kScanThreads;

__device__ inline int block_exclusive_scan(int v, int* lds, int* total) {
  // lds: blockDim.x / 64 + 1 ints. Wave scans, then one wave scans the wave totals.
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  const int incl = wave_incl_scan(v);
  if (lane == kWave - 1) lds[wid] = incl;
  __syncthreads();
  if (wid == 0) {
    const int t = lane < nw ? lds[lane] : 0;
    const int ti = wave_incl_scan(t);
    if (lane < nw) lds[lane] = ti - t;
    if (lane == nw - 1) lds[nw] = ti;
  }
  __syncthreads();
  const int woff = lds[wid];
  *total = lds[nw];
  __syncthreads();
  return woff + incl - v;
}

// One-workgroup exclusive scan of count[0, m) for m <= 1024 ROWS, read as
// rows of 1024 (coalesced, all in flight at once). All rows are scanned
// together: wave scans in registers, one pass over the wave totals of every
// row, one pass over the row totals. emit(k, s) receives the exclusive prefix
// s of every k < m (after every count has been read); returns the total.
constexpr int kScanOneThreads = 1024;
template <int ROWS, typename Emit>
__device__ inline int scan_rows(const int32_t* __restrict__ count, int m, Emit emit) {
  constexpr int NW = kScanOneThreads / kWave;
  __shared__ int wt[ROWS][NW];    // wave totals of each row
  __shared__ int woff[ROWS][NW];  // exclusive wave offsets within each row
  __shared__ int rt[ROWS];        // row totals
  __shared__ int roff[ROWS + 1];  // exclusive row offsets; roff[rows] = total
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  const int rows = (m + kScanOneThreads - 1) / kScanOneThreads;
  int v[ROWS], incl[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const int k = r * kScanOneThreads + tid;
    v[r] = (r < rows && k < m) ? count[k] : 0;
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    int x = v[r];
    if (r < rows) {  // uniform over the workgroup
      x = wave_incl_scan(x);
      if (lane == kWave - 1) wt[r][wid] = x;
    }
    incl[r] = x;
  }
  __syncthreads();
  if (tid < rows * NW) {
    const int r = tid / NW, w = tid % NW;
    int sacc = 0;
    for (int u = 0; u < w; ++u) sacc += wt[r][u];
    woff[r][w] = sacc;
    if (w == NW - 1) rt[r] = sacc + wt[r][w];
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int r = 0; r < rows; ++r) {
      roff[r] = acc;
      acc += rt[r];
    }
    roff[rows] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    const int k = r * kScanOneThreads + tid;
    if (r < rows && k < m) emit(k, roff[r] + woff[r][wid] + incl[r] - v[r]);
  }
  return roff[rows];
}

// The scan of up to kScanOneMax counts as one workgroup. With zero_count it
// also zeroes the counts themselves (ready for the next step's binning).
constexpr int kScanOneRows = 32;
constexpr int kScanOneMax = kScanOneThreads * kScanOneRows;
__global__ __launch_bounds__(1024) void k_scan_one(int32_t* __restrict__ count, int m, int32_t* __restrict__ start,
                                                   int64_t n, int zero_count) {
  const int total = scan_rows<kScanOneRows>(count, m, [&](int k, int s) {
    start[k] = s;
    if (zero_count) count[k] = 0;
  });
  if (threadIdx.x == 0) start[m] = n >= 0 ? (int32_t)n : total;
}

// Scan and count-sort placement in one launch, for grids of at most
// kFusedScanMax cells: every workgroup scans the whole count array (a few
// tens of KB, read from L2) into LDS, then places its 1024 atoms at
// start[c] + rank; workgroup 0 also writes start. This replaces the
// one-workgroup scan and the placement kernel (one launch and the scan's
// serial tail fewer). The counts are zeroed later, by the cell sort.
constexpr int kFusedScanRows = 15;
constexpr int kFusedScanMax = kScanOneThreads * kFusedScanRows;  // (m + 1) starts: at most 60 KB of LDS
__global__ __launch_bounds__(1024) void k_scan_scatter(const int32_t* __restrict__ count, int m,
                                                       int32_t* __restrict__ start, int64_t n,
                                                       const int32_t* __restrict__ cell_of,
                                                       const int32_t* __restrict__ rank, int32_t* __restrict__ perm) {
  extern __shared__ int st[];  // m + 1 exclusive starts
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = i < n;
  const int c = on ? cell_of[i] : 0, rk = on ? rank[i] : 0;  // in flight during the scan
  const bool first = blockIdx.x == 0;
  const int total = scan_rows<kFusedScanRows>(count, m, [&](int k, int s) {
    st[k] = s;
    if (first) start[k] = s;
  });
  if (threadIdx.x == 0) {
    st[m] = total;
    if (first) start[m] = (int32_t)n;
  }
  __syncthreads();
  if (on) {
    const int pos = st[c] + rk;
    if (pos < st[c + 1]) perm[pos] = (int32_t)i;  // ranks are 0 .. count - 1 by construction
  }
}

__global__ void k_scan_blocks(const int32_t* __restrict__ count, int ncells, int32_t* __restrict__ blocksum) {
  __shared__ int lds[kScanThreads / kWave + 1];
  int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  int s = 0;
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < ncells) s += count[base + k];
  int total;
  block_exclusive_scan(s, lds, &total);
  if (threadIdx.x == 0) blocksum[blockIdx.x] = total;
}

__global__ void k_scan_top(int32_t* __restrict__ blocksum, int nb) {
  __shared__ int lds[kScanThreads / kWave + 1];
  int carry = 0;
  for (int base = 0; base < nb; base += kScanThreads) {
    int i = base + threadIdx.x;
    int v = i < nb ? blocksum[i] : 0;
    int total;
    int ex = block_exclusive_scan(v, lds, &total);
    if (i < nb) blocksum[i] = carry + ex;
    carry += total;
    __syncthreads();
  }
}

__global__ void k_scan_final(const int32_t* __restrict__ count, int ncells, const int32_t* __restrict__ blocksum,
                             int32_t* __restrict__ start, int64_t n) {
  __shared__ int lds[kScanThreads / kWave + 1];
  int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  int v[kScanItems], s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < ncells ? count[base + k] : 0;
    s += v[k];
  }
  int total;
  int ex = block_exclusive_scan(s, lds, &total) + blocksum[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < ncells) {
      start[base + k] = ex;
    }
    ex += v[k];
  }
  // n < 0: the total is the scan's own sum (written by the block that holds the last item)
  if (n >= 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) start[ncells] = (int32_t)n;
  } else if (base <= ncells - 1 && ncells - 1 < base + kScanItems) {
    start[ncells] = ex;
  }
}

// Count-sort placement from the ranks of the counting pass (no atomics); the
// order inside a cell is fixed by the cell sort that follows.
__global__ void k_scatter(int64_t n, const int32_t* __restrict__ cell_of, const int32_t* __restrict__ start,
                          const int32_t* __restrict__ rank, int32_t* __restrict__ perm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of[i];
  const int pos = start[c] + rank[i];
  if (pos < start[c + 1]) perm[pos] = (int32_t)i;  // ranks are 0 .. count - 1 by construction
}

// Deterministic order inside each cell: ascending original index (the stable
// sortperm of storage.jl:199), then the gather into sorted SoA coordinates.
// One wavefront per pair of consecutive cells. When both hold at most 32 atoms
// (the common case) each half-wave sorts one of them with a register bitonic
// network over lane shuffles (of the smallest power of two that holds the
// larger cell), so one generation of wavefronts covers twice the cells.
// Otherwise the wavefront takes the two cells in turn: cells of up to 64 atoms
// by the same network over the whole wave, cells of up to kRankSort atoms by
// rank counting over an LDS copy (indices are distinct), larger cells (dense
// clusters only) by rank counting over a copy in `scratch` (n entries, free at
// this point). With `count` given, the cells' counts are zeroed (ready for the
// next binning; the fused scan does not zero them).
constexpr int kRankSort = 256;

// Bitonic sort of v within each aligned group of `size` lanes (size a power
// of two, at most 64): every group ends ascending. The last merge (k = size)
// is ascending in every group; taking its direction from lane bit k instead
// would leave the odd groups descending.
__device__ inline int bitonic_lanes(int v, int lane, int size) {
  for (int k = 2; k <= size; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int o = __shfl_xor(v, j, kWave);
      const bool up = k == size || (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      v = (lower == up) ? min(v, o) : max(v, o);
    }
  return v;
}

// Gather of one sorted atom (its sorted slot k, original index i) from its
// wrap record into the SoA arrays; xr relative to the cell's first atom
// (ox, oy, oz); extends the lane's box.
__device__ inline void gather_sorted(int64_t n, int k, int i, double ox, double oy, double oz,
                                     const double* __restrict__ xw, double* __restrict__ xs,
                                     int32_t* __restrict__ sps, float4* __restrict__ xr, float* lo, float* hi) {
  const double2* rec = (const double2*)(xw + 4 * (int64_t)i);
  const double2 a = rec[0], b = rec[1];
  const double x = a.x, y = a.y, z = b.x;
  const int sp = (int)__double_as_longlong(b.y);
  xs[k] = x;
  xs[n + k] = y;
  xs[2 * n + k] = z;
  sps[k] = sp;
  const float4 r = make_float4((float)(x - ox), (float)(y - oy), (float)(z - oz), __int_as_float(sp));
  xr[k] = r;
  lo[0] = fminf(lo[0], r.x);
  lo[1] = fminf(lo[1], r.y);
  lo[2] = fminf(lo[2], r.z);
  hi[0] = fmaxf(hi[0], r.x);
  hi[1] = fmaxf(hi[1], r.y);
  hi[2] = fmaxf(hi[2], r.z);
}

// One cell [b, e) by the whole wavefront.
__device__ inline void sort_gather_cell(int c, int b, int e, int64_t n, int lane, int* buf,
                                        int32_t* __restrict__ perm, const double* __restrict__ xw,
                                        double* __restrict__ xs,
                                        int32_t* __restrict__ sps, float4* __restrict__ xr,
                                        float4* __restrict__ cbox, int32_t* __restrict__ scratch) {
  const int cnt = e - b;
  if (cnt == 0) return;
  if (cnt <= kWave) {
    int v = lane < cnt ? perm[b + lane] : 0x7fffffff;
    v = bitonic_lanes(v, lane, cnt <= 16 ? 16 : (cnt <= 32 ? 32 : kWave));
    if (lane < cnt) perm[b + lane] = v;
  } else if (cnt <= kRankSort) {
    for (int k = lane; k < cnt; k += kWave) buf[k] = perm[b + k];
    __syncthreads();
    for (int k = lane; k < cnt; k += kWave) {
      const int v = buf[k];
      int rank = 0;
      for (int t = 0; t < cnt; ++t) rank += buf[t] < v;
      perm[b + rank] = v;
    }
    __syncthreads();
  } else {
    for (int k = b + lane; k < e; k += kWave) scratch[k] = perm[k];
    __syncthreads();  // one wavefront: orders its own global writes and reads
    for (int k = b + lane; k < e; k += kWave) {
      const int v = scratch[k];
      int rank = 0;
      for (int t = b; t < e; ++t) rank += scratch[t] < v;
      perm[b + rank] = v;
    }
    __syncthreads();
  }
  // xr: FP32 positions relative to the cell's first atom, species bits in w;
  // cbox: their box (the pair kernel's candidate cull)
  const int64_t i0 = perm[b];
  const double ox = xw[4 * i0], oy = xw[4 * i0 + 1], oz = xw[4 * i0 + 2];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = b + lane; k < e; k += kWave) gather_sorted(n, k, perm[k], ox, oy, oz, xw, xs, sps, xr, lo, hi);
  for (int off = kWave / 2; off > 0; off >>= 1)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, kWave));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, kWave));
    }
  if (lane == 0) {
    cbox[2 * (int64_t)c] = make_float4(lo[0], lo[1], lo[2], 0.f);
    cbox[2 * (int64_t)c + 1] = make_float4(hi[0], hi[1], hi[2], 0.f);
  }
}

__global__ __launch_bounds__(64) void k_cell_sort_gather(int ncells, int64_t n, const int32_t* __restrict__ start,
                                                         int32_t* __restrict__ perm, const double* __restrict__ xw,
                                                         double* __restrict__ xs, int32_t* __restrict__ sps,
                                                         float4* __restrict__ xr, float4* __restrict__ cbox,
                                                         int32_t* __restrict__ scratch, int32_t* __restrict__ count) {
  __shared__ int buf[kRankSort];
  const int lane = threadIdx.x;
  const int c0 = 2 * blockIdx.x;
  const bool two = c0 + 1 < ncells;
  const int b0 = start[c0], e0 = start[c0 + 1], e1 = two ? start[c0 + 2] : e0;
  if (count && lane < (two ? 2 : 1)) count[c0 + lane] = 0;
  const int n0 = e0 - b0, n1 = e1 - e0;
  if (n0 > 32 || n1 > 32) {
    sort_gather_cell(c0, b0, e0, n, lane, buf, perm, xw, xs, sps, xr, cbox, scratch);
    if (two) sort_gather_cell(c0 + 1, e0, e1, n, lane, buf, perm, xw, xs, sps, xr, cbox, scratch);
    return;
  }
  // half-wave h sorts cell c0 + h
  const int h = lane >> 5, hl = lane & 31;
  const int b = h ? e0 : b0, cnt = h ? n1 : n0;
  int v = hl < cnt ? perm[b + hl] : 0x7fffffff;
  v = bitonic_lanes(v, lane, max(n0, n1) <= 16 ? 16 : 32);  // shuffles stay inside each half
  const int64_t i0 = __shfl(v, lane & 32, kWave);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (hl < cnt) {
    perm[b + hl] = v;
    gather_sorted(n, b + hl, v, xw[4 * i0], xw[4 * i0 + 1], xw[4 * i0 + 2], xw, xs, sps, xr, lo, hi);
  }
  for (int off = 16; off > 0; off >>= 1)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, kWave));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, kWave));
    }
  if (hl == 0 && cnt > 0) {
    const int64_t c = c0 + h;
    cbox[2 * c] = make_float4(lo[0], lo[1], lo[2], 0.f);
    cbox[2 * c + 1] = make_float4(hi[0], hi[1], hi[2], 0.f);
  }
}

__global__ void k_colsum(const double* __restrict__ in, int64_t m, int width, double* __restrict__ out) {
  __shared__ double lds[16][4];
  const int64_t chunk = (m + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(m, lo + chunk);
  double s[16];
  for (int c = 0; c < width; ++c) s[c] = 0;
  for (int64_t r = lo + threadIdx.x; r < hi; r += blockDim.x)
    for (int c = 0; c < width; ++c) s[c] += in[r * width + c];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  for (int c = 0; c < width; ++c) {
    const double v = wave_sum(s[c]);
    if (lane == 0) lds[c][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < width) {
    const int c = threadIdx.x;
    out[(int64_t)blockIdx.x * width + c] = lds[c][0] + lds[c][1] + lds[c][2] + lds[c][3];
  }
}

// Final stage of the energy / virial reduction of one configuration: the
// stage-1 rows of k_colsum (nb x width) summed in fixed order, one lane per
// column, then E, W (3x3 column-major) and the status word written to
// out[0..10], out2[0..10] and, if given, out3[0..10] (host memory). kind 0: pair records (E, xx yy zz xy xz yz of
// sum d (x) g over directed pairs), halved; kind 1: ACE records (E, W[9]) with
// the status word of the environment build.
__global__ void k_finish(const double* __restrict__ part, int nb, int width, int kind,
                         const int32_t* __restrict__ status, double* __restrict__ out, double* __restrict__ out2,
                         double* __restrict__ out3) {
  __shared__ double rows[kRedBlocks * 16];
  __shared__ double t[16];
  for (int i = threadIdx.x; i < nb * width; i += blockDim.x) rows[i] = part[i];  // all loads in flight at once
  __syncthreads();
  if ((int)threadIdx.x < width) {
    double v = 0;
    for (int b = 0; b < nb; ++b) v += rows[b * width + threadIdx.x];
    t[threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double E, W[9];
  if (kind == 0) {
    E = 0.5 * t[0];
    W[0] = -0.5 * t[1];
    W[4] = -0.5 * t[2];
    W[8] = -0.5 * t[3];
    W[3] = W[1] = -0.5 * t[4];
    W[6] = W[2] = -0.5 * t[5];
    W[7] = W[5] = -0.5 * t[6];
  } else {
    E = t[0];
    for (int k = 0; k < 9; ++k) W[k] = t[1 + k];
  }
  bool bad = !isfinite(E) || (status && status[0] != 0);
  for (int k = 0; k < 9; ++k) bad |= !isfinite(W[k]);
  out[0] = out2[0] = E;
  for (int k = 0; k < 9; ++k) out[1 + k] = out2[1 + k] = W[k];
  out[10] = out2[10] = bad ? 1.0 : 0.0;
  if (out3) {
    out3[0] = E;
    for (int k = 0; k < 9; ++k) out3[1 + k] = W[k];
    out3[10] = bad ? 1.0 : 0.0;
  }
}

// E, W and the status word of one configuration: k_colsum, then k_finish.
// (A single launch whose last block finishes the sum needs a device-scope
// fence in every block, which writes back the L2 and was measured slower; one
// 1,024-thread workgroup doing both stages took 30 us against 9 us for the
// 256k pair step's 14,440 rows.)
void reduce_core(const double* rec, int64_t m, int width, int kind, const int32_t* status, double* red, double* out,
                 double* out2, double* out3, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3(kRedBlocks), dim3(256), 0, st, rec, m, width, red);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, st, red, kRedBlocks, width, kind, status, out, out2, out3);
  PFW_HIP(hipGetLastError());
}

void reduce_outputs(pfw_ctx* ctx, const double* rec, int64_t m, int width, int kind, const int32_t* status) {
  reduce_core(rec, m, width, kind, status, ctx->d_red, ctx->d_out, ctx->d_cfg_out, ctx->h_out_dev, ctx->stream);
  ctx->host_written = true;
}

AuditBufs audit_begin(pfw_ctx* ctx) {
  const int64_t n = ctx->sys.n;
  if (ctx->audit_cap < n || !ctx->audit_count) {
    if (ctx->audit_count) (void)hipFree(ctx->audit_count);
    if (ctx->audit_hash) (void)hipFree(ctx->audit_hash);
    ctx->audit_count = nullptr;
    ctx->audit_hash = nullptr;
    ctx->audit_cap = 0;
    PFW_HIP(hipMalloc(&ctx->audit_count, std::max<int64_t>(n, 1) * sizeof(int32_t)));
    PFW_HIP(hipMalloc(&ctx->audit_hash, std::max<int64_t>(n, 1) * sizeof(unsigned long long)));
    ctx->audit_cap = std::max<int64_t>(n, 1);
  }
  PFW_HIP(hipMemsetAsync(ctx->audit_count, 0, std::max<int64_t>(n, 1) * sizeof(int32_t), ctx->stream));
  PFW_HIP(hipMemsetAsync(ctx->audit_hash, 0, std::max<int64_t>(n, 1) * sizeof(unsigned long long), ctx->stream));
  ctx->audit_valid = true;
  return AuditBufs{ctx->audit_count, ctx->audit_hash, ctx->sys.X};
}

void copy_outputs(pfw_ctx* ctx) {
  if (!ctx->host_written)
    PFW_HIP(hipMemcpyAsync(ctx->h_out, ctx->d_out, 11 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ctx->host_written = false;
}

// Column sums of a (nblocks x width) table of per-workgroup partials, in
// fixed row order (deterministic).
__global__ void k_sum_partials(const double* __restrict__ partial, int nblocks, int width, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= width) return;
  double s = 0;
  for (int b = 0; b < nblocks; ++b) s += partial[(int64_t)b * width + k];
  out[k] = s;
}

static inline unsigned nblocks(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

// Exclusive scan of m counts into start[0..m] (start[m] = total).
// blocksum needs (m + 1023) / 1024 + 1 ints.
void scan_counts(pfw_ctx* ctx, const int32_t* count, int m, int32_t* start, int32_t* blocksum) {
  hipStream_t st = ctx->stream;
  if (m <= kScanOneMax) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanOneThreads), 0, st, const_cast<int32_t*>(count), m, start,
                       (int64_t)-1, 0);
    PFW_HIP(hipGetLastError());
    return;
  }
  int nb = (int)nblocks(m, kScanBlock);
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanThreads), 0, st, count, m, blocksum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, st, blocksum, nb);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanThreads), 0, st, count, m, blocksum, start, (int64_t)-1);
  PFW_HIP(hipGetLastError());
}

// Open-axis extent words: minimum (all ones) and maximum (zero) per axis.
__global__ void k_bounds_init(unsigned long long* __restrict__ bounds, int words) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < words) bounds[k] = (k % 2 == 0) ? ~0ull : 0ull;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t m, int32_t v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) p[k] = v;
}

void fill_i32(int32_t* p, int64_t m, int32_t v, hipStream_t st) {
  if (m > 0) hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(m, 256)), dim3(256), 0, st, p, m, v);
}

// The cell build over explicit buffers: wrap and count, (open axes: extent,
// grid, binning), scan, scatter and the per-cell gather. `counts_zero`: the
// count array is already zero (the previous build's one-workgroup scan leaves
// it so). Stage events are recorded on `timing` when it is given.
void nl_build_core(const BuildView& v, hipStream_t st, bool counts_zero, pfw_ctx* timing) {
  const int64_t n = v.n;
  const int ncells = v.ncells;
  auto stage = [&](int k) {
    if (timing) record_stage(timing, k);
  };
  stage(0);
  const bool one = ncells <= kScanOneMax;  // the one-workgroup scan leaves the counts zero
  if (!counts_zero) fill_i32(v.count, std::max(ncells, 1), 0, st);
  if (v.open_any)
    hipLaunchKernelGGL(k_bounds_init, dim3(nblocks(6 * (int64_t)v.ncfg, 64)), dim3(64), 0, st,
                       (unsigned long long*)v.bounds, 6 * v.ncfg);
  if (n > 0)
    hipLaunchKernelGGL(k_wrap, dim3(nblocks(n, 256)), dim3(256), 0, st, v.grids, v.atom_cfg, n, v.X, v.species, v.xw, v.u,
                       v.cell_of, v.count, v.rank, (unsigned long long*)v.bounds, (int)v.open_any,
                       (int)(v.ncfg == 1));
  if (v.open_any) {
    hipLaunchKernelGGL(k_open_grid, dim3(nblocks(v.ncfg, 64)), dim3(64), 0, st, v.grids,
                       (const unsigned long long*)v.bounds, v.ncfg);
    if (n > 0)
      hipLaunchKernelGGL(k_bin_open, dim3(nblocks(n, 256)), dim3(256), 0, st, v.grids, v.atom_cfg, n, v.u,
                         v.cell_of, v.count, v.rank);
  }
  PFW_HIP(hipGetLastError());
  stage(1);
  // counts are zeroed by the one-workgroup scan, or, after the fused scan, by the cell sort
  const bool fused = n > 0 && ncells <= kFusedScanMax;
  if (fused) {
    hipLaunchKernelGGL(k_scan_scatter, dim3(nblocks(n, kScanOneThreads)), dim3(kScanOneThreads),
                       (ncells + 1) * sizeof(int), st, v.count, ncells, v.start, n, v.cell_of, v.rank, v.perm);
  } else if (one) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanOneThreads), 0, st, v.count, ncells, v.start, n, 1);
  } else {
    int nb = (int)nblocks(ncells, kScanBlock);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanThreads), 0, st, v.count, ncells, v.blocksum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, st, v.blocksum, nb);
    hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanThreads), 0, st, v.count, ncells, v.blocksum, v.start, n);
  }
  PFW_HIP(hipGetLastError());
  stage(2);
  if (n > 0) {
    if (!fused)
      hipLaunchKernelGGL(k_scatter, dim3(nblocks(n, 256)), dim3(256), 0, st, n, v.cell_of, v.start, v.rank, v.perm);
    hipLaunchKernelGGL(k_cell_sort_gather, dim3(nblocks(ncells, 2)), dim3(kWave), 0, st, ncells, n, v.start, v.perm,
                       v.xw, v.xs, v.sps, v.xr, v.cbox, v.rank, fused ? v.count : nullptr);
  }
  PFW_HIP(hipGetLastError());
}

void nl_build(pfw_ctx* ctx) {
  ++ctx->build_gen;  // sorted indices change: a Verlet list made before is stale
  SystemBuffers& s = ctx->sys;
  BuildView v{s.n,    ctx->ncells, ctx->ncfg, ctx->open_any, s.X,     s.xw,      s.u,    s.cell_of, s.count,
              s.rank, s.bounds,    s.grids,   s.atom_cfg,    s.start, s.blocksum, s.perm, s.xs,      s.species,
              s.sps,  s.xr,      s.cbox};
  nl_build_core(v, ctx->stream, ctx->ncells <= kScanOneMax, ctx);
}

// ---------------------------------------------------------------- per-configuration sums

// Column sums of each configuration's records: rows [lo, hi) with lo/hi the
// configuration's global cell offsets, or (atoms != nullptr) the sorted-atom
// ranges start[cell offsets].
__global__ void k_segsum(const double* __restrict__ in, int width, const int32_t* __restrict__ cfg_cell0,
                         const int32_t* __restrict__ start, double* __restrict__ out) {
  __shared__ double lds[16][4];
  const int cfg = blockIdx.x;
  int lo = cfg_cell0[cfg], hi = cfg_cell0[cfg + 1];
  if (start) {
    lo = start[lo];
    hi = start[hi];
  }
  double s[16];
  for (int c = 0; c < width; ++c) s[c] = 0;
  for (int r = lo + threadIdx.x; r < hi; r += blockDim.x)
    for (int c = 0; c < width; ++c) s[c] += in[(int64_t)r * width + c];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  for (int c = 0; c < width; ++c) {
    const double v = wave_sum(s[c]);
    if (lane == 0) lds[c][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < width) {
    const int c = threadIdx.x;
    out[16 * (int64_t)cfg + c] = lds[c][0] + lds[c][1] + lds[c][2] + lds[c][3];
  }
}

// Per-configuration E, W (3x3 column-major) and status from the column sums.
// kind 0 (pair): columns E, xx yy zz xy xz yz of d (x) g over directed pairs
// (E and W carry the 1/2 of the full list); kind 1 (ACE): E, W[9].
__global__ void k_batch_finish(const double* __restrict__ sums, int ncfg, int kind, const int32_t* __restrict__ status,
                               double* __restrict__ out) {
  const int cfg = blockIdx.x * blockDim.x + threadIdx.x;
  if (cfg >= ncfg) return;
  const double* t = sums + 16 * (int64_t)cfg;
  double* o = out + 11 * (int64_t)cfg;
  if (kind == 0) {
    o[0] = 0.5 * t[0];
    o[1] = -0.5 * t[1];
    o[5] = -0.5 * t[2];
    o[9] = -0.5 * t[3];
    o[4] = o[2] = -0.5 * t[4];
    o[7] = o[3] = -0.5 * t[5];
    o[8] = o[6] = -0.5 * t[6];
  } else {
    for (int c = 0; c < 10; ++c) o[c] = t[c];
  }
  bool bad = status && status[0] != 0;
  for (int c = 0; c < 10; ++c) bad |= !isfinite(o[c]);
  o[10] = bad ? 1.0 : 0.0;
}

__global__ void k_total(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x >= 11) return;
  double t = 0;
  for (int b = 0; b < nb; ++b) t += part[11 * b + threadIdx.x];
  out[threadIdx.x] = t;
}

// records: (cells ? ncells : n) x width; kind 0 pair, 1 ACE. Writes per-configuration
// results to d_cfg_out and the batch totals (status = number of failed configurations) to d_out.
void batch_reduce(pfw_ctx* ctx, const double* records, int width, bool cells, int kind) {
  ctx->host_written = false;
  hipStream_t st = ctx->stream;
  const int ncfg = ctx->ncfg;
  hipLaunchKernelGGL(k_segsum, dim3(ncfg), dim3(256), 0, st, records, width, ctx->sys.cfg_cell0,
                     cells ? (const int32_t*)nullptr : (const int32_t*)ctx->sys.start, ctx->d_cfg);
  hipLaunchKernelGGL(k_batch_finish, dim3(nblocks(ncfg, 64)), dim3(64), 0, st, ctx->d_cfg, ncfg, kind,
                     kind == 1 ? (const int32_t*)ctx->ace.status : (const int32_t*)nullptr, ctx->d_cfg_out);
  hipLaunchKernelGGL(k_colsum, dim3(kRedBlocks), dim3(256), 0, st, ctx->d_cfg_out, (int64_t)ncfg, 11, ctx->d_red);
  hipLaunchKernelGGL(k_total, dim3(1), dim3(64), 0, st, ctx->d_red, kRedBlocks, ctx->d_out);
  PFW_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- pair export

__device__ inline void winding(const Grid& g, const double* X, int i, int* w) {
  double x[3] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]};
  if (g.minimum_image) {
    for (int a = 0; a < 3; ++a) w[a] = g.pb[a] ? (int)floor(x[a] / g.L[a]) : 0;
  } else {
    double f[3];
    exact_matvec(g.recip, x, f);
    for (int a = 0; a < 3; ++a) w[a] = g.pb[a] ? (int)floor(f[a]) : 0;
  }
}

// One thread per sorted atom: walk the image stencil and emit every directed
// pair strictly inside the cutoff (interface.jl:144-171). Pass 0 counts, pass 1 writes.
__global__ void k_export(const Grid* __restrict__ grids, const int32_t* __restrict__ atom_cfg, int64_t n,
                         const double* __restrict__ X,
                         const double* __restrict__ xs, const int32_t* __restrict__ start,
                         const int32_t* __restrict__ cell_of, const int32_t* __restrict__ perm,
                         int32_t* __restrict__ counts, const int64_t* __restrict__ offsets, int pass,
                         int32_t* __restrict__ oi, int32_t* __restrict__ oj, int32_t* __restrict__ oS,
                         double* __restrict__ oR) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int i = perm[k];
  const Grid& g = grids[atom_cfg[i]];
  int cell = cell_of[i];
  int key[3];
  cell_key(g, cell, key);
  double xi = xs[k], yi = xs[n + k], zi = xs[2 * n + k];
  int wi[3];
  if (pass) winding(g, X, i, wi);
  int64_t pos = pass ? offsets[k] : 0;
  int cnt = 0;
  for (int e = 0; e < g.nstencil; ++e) {
    int lat[3];
    bool zero;
    int nc = stencil_entry(g, key, e, lat, &zero);
    if (nc < 0) continue;
    double sh[3];
    exact_shift(g, lat, sh);
    for (int m = start[nc]; m < start[nc + 1]; ++m) {
      if (zero && m == k) continue;
      double d[3];
      int l[3];
      if (!exact_member(g, xi, yi, zi, xs[m], xs[n + m], xs[2 * n + m], lat, sh, d, l)) continue;
      if (pass) {
        int j = perm[m];
        int wj[3];
        winding(g, X, j, wj);
        oi[pos] = i;
        oj[pos] = j;
        if (oS)
          for (int a = 0; a < 3; ++a) oS[3 * pos + a] = l[a] + wi[a] - wj[a];
        if (oR)
          for (int a = 0; a < 3; ++a) oR[3 * pos + a] = d[a];
        ++pos;
      }
      ++cnt;
    }
  }
  if (!pass) counts[k] = cnt;
}

// The directed pairs of the current system on the device, grouped by sorted
// atom: pairs of sorted atom k at [off[k], off[k + 1]) (fresh cell build).
// Every array is allocated here (DeviceExport frees them); S may be skipped.
void nl_export_device(pfw_ctx* ctx, DeviceExport& ex, bool with_S) {
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  nl_build(ctx);
  ex.release();
  ex.total = 0;
  if (n == 0) return;
  int32_t* d_counts = nullptr;
  PFW_HIP(hipMalloc(&d_counts, n * sizeof(int32_t)));
  ex.counts = d_counts;
  PFW_HIP(hipMalloc(&ex.off, (n + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(k_export, dim3(nblocks(n, 128)), dim3(128), 0, st, s.grids, s.atom_cfg, n, s.X, s.xs, s.start,
                     s.cell_of, s.perm, d_counts, (const int64_t*)nullptr, 0, nullptr, nullptr, nullptr, nullptr);
  PFW_HIP(hipGetLastError());
  std::vector<int32_t> counts(n);
  PFW_HIP(hipMemcpyAsync(counts.data(), d_counts, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> off(n + 1);
  int64_t total = 0;
  for (int64_t k = 0; k < n; ++k) {
    off[k] = total;
    total += counts[k];
  }
  off[n] = total;
  ex.total = total;
  PFW_HIP(hipMemcpyAsync(ex.off, off.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  PFW_HIP(hipMalloc(&ex.i, std::max<int64_t>(total, 1) * sizeof(int32_t)));
  PFW_HIP(hipMalloc(&ex.j, std::max<int64_t>(total, 1) * sizeof(int32_t)));
  if (with_S) PFW_HIP(hipMalloc(&ex.S, std::max<int64_t>(total, 1) * 3 * sizeof(int32_t)));
  PFW_HIP(hipMalloc(&ex.R, std::max<int64_t>(total, 1) * 3 * sizeof(double)));
  hipLaunchKernelGGL(k_export, dim3(nblocks(n, 128)), dim3(128), 0, st, s.grids, s.atom_cfg, n, s.X, s.xs, s.start,
                     s.cell_of, s.perm, d_counts, (const int64_t*)ex.off, 1, ex.i, ex.j, ex.S, ex.R);
  PFW_HIP(hipGetLastError());
  PFW_HIP(hipStreamSynchronize(st));  // the staging vector goes out of scope
}

void DeviceExport::release() {
  void* p[] = {counts, off, i, j, S, R};
  for (void* q : p)
    if (q) (void)hipFree(q);
  counts = i = j = S = nullptr;
  off = nullptr;
  R = nullptr;
}

void nl_export_pairs(pfw_ctx* ctx, int32_t* hi, int32_t* hj, int32_t* hS, double* hR, int64_t cap,
                     int64_t* count) {
  hipStream_t st = ctx->stream;
  DeviceExport ex;
  nl_export_device(ctx, ex, true);
  const int64_t total = ex.total;
  *count = total;
  if (total > cap) throw Error(PFW_CAPACITY, "pair export capacity too small");
  if (total > 0) {
    if (hi) PFW_HIP(hipMemcpyAsync(hi, ex.i, total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (hj) PFW_HIP(hipMemcpyAsync(hj, ex.j, total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (hS) PFW_HIP(hipMemcpyAsync(hS, ex.S, total * 3 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (hR) PFW_HIP(hipMemcpyAsync(hR, ex.R, total * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  PFW_HIP(hipStreamSynchronize(st));
}

}  // namespace pfw
