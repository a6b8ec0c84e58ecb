// Verlet-list reuse for the pair models on gfx950: a list at cutoff + skin,
// built from the cell grid with explicit ghost images, kept while no atom has
// moved half the skin since the build.
//
// Reference: the skin path of the GPU pair workspace,
// src/submodules/NeighbourList/gpu.jl:404-549 (gpu_displacements!,
// gpu_connectivity!, gpu_buffered_central!, gpu_skin!) and docs/gpu_pairs.md:116-131:
// rebuild when 2 max|x - x_ref| >= skin (1 - 16 eps) over raw caller
// coordinates (a wrapped jump therefore rebuilds; no winding is inferred),
// reference positions fixed between rebuilds, every interaction at the
// physical cutoff from current positions. The reference restricts this path
// to orthorhombic cells longer than 2 (cutoff + skin) and stores bare
// neighbour indices with minimum-image displacements; here triclinic and
// multi-image cells are covered by ghosts instead:
//
//  * the grid is extended by the stencil radius on every periodic axis; an
//    extended cell outside the real range is an image of a real cell with
//    lattice shift fld(key, dims), and its atoms are ghost entries
//    xe[n + g] = x_src + H s;
//  * a list entry is a bare int32 index into xe = [sorted atoms | ghosts], so
//    the interaction loop reads one index and one 32-byte position per pair
//    and computes d = xe_j - xe_i with no shift arithmetic;
//  * every entry keeps off = (build position) - (caller position at the
//    build); a step sets xe = off + X_now for atoms and ghosts alike, which is
//    x_j - x_i + H S of the current raw coordinates with the build's shifts S.
//
// The list holds every image (j, S) with |d| < cutoff + skin at the build, so
// with max displacement m and 2 m < skin every pair within the cutoff now is
// listed; membership itself is r^2 < c^2 of the current d in FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pfw_device.h"

namespace pfw {

static inline unsigned nblk(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

// Extended cell `local` of grid g: true for a ghost cell, with its real
// source cell and lattice shift.
__device__ inline bool ext_source(const Grid& g, int local, int* src, int* lat) {
  int ek[3];
  ek[0] = local % g.edims[0];
  const int t = local / g.edims[0];
  ek[1] = t % g.edims[1];
  ek[2] = t / g.edims[1];
  bool ghost = false;
  int sk[3];
  for (int a = 0; a < 3; ++a) {
    const int v = ek[a] - (g.pb[a] ? g.radius[a] : 0);
    if (g.pb[a]) {
      sk[a] = jl_imod(v, g.dims[a]);
      lat[a] = jl_fld(v, g.dims[a]);
      ghost |= lat[a] != 0;
    } else {
      sk[a] = v;
      lat[a] = 0;
    }
  }
  *src = cell_index(g, sk);
  return ghost;
}

// The configuration owning extended cell e: the last k with ecfg0[k] <= e.
__device__ inline int ext_config(const int32_t* ecfg0, int ncfg, int64_t e) {
  int lo = 0, hi = ncfg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ecfg0[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ void k_ext_count(int64_t ext_cells, const int32_t* __restrict__ ecfg0, int ncfg,
                            const Grid* __restrict__ grids, const int32_t* __restrict__ start,
                            int32_t* __restrict__ gcount) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ext_cells) return;
  const Grid& g = grids[ext_config(ecfg0, ncfg, e)];
  int src, lat[3];
  gcount[e] = ext_source(g, (int)(e - g.ecell0), &src, lat) ? start[src + 1] - start[src] : 0;
}

// Sorted atoms: build position xs minus the caller's position, species in the
// fourth word.
__global__ void k_real_fill(int64_t n, const double* __restrict__ xs, const int32_t* __restrict__ sps,
                            const int32_t* __restrict__ perm, const double* __restrict__ X, double* __restrict__ off,
                            int32_t* __restrict__ srcp) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int p = perm[k];
  off[4 * k] = xs[k] - X[3 * (int64_t)p];
  off[4 * k + 1] = xs[n + k] - X[3 * (int64_t)p + 1];
  off[4 * k + 2] = xs[2 * n + k] - X[3 * (int64_t)p + 2];
  off[4 * k + 3] = __longlong_as_double((long long)sps[k]);
  srcp[k] = p;
}

// Ghosts: one workgroup per extended cell; the atoms of a ghost cell are the
// source cell's atoms shifted by H s.
__global__ void k_ghost_fill(int64_t ext_cells, const int32_t* __restrict__ ecfg0, int ncfg,
                             const Grid* __restrict__ grids, const int32_t* __restrict__ start,
                             const int32_t* __restrict__ gstart, int64_t n, const double* __restrict__ xs,
                             const int32_t* __restrict__ sps, const int32_t* __restrict__ perm,
                             const double* __restrict__ X, double* __restrict__ off, int32_t* __restrict__ srcp) {
  const int64_t e = blockIdx.x;
  const Grid& g = grids[ext_config(ecfg0, ncfg, e)];
  int src, lat[3];
  if (!ext_source(g, (int)(e - g.ecell0), &src, lat)) return;
  double sh[3];
  exact_shift(g, lat, sh);
  const int b = start[src], cnt = start[src + 1] - b;
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int64_t m = b + t, dst = n + gstart[e] + t;
    const int p = perm[m];
    off[4 * dst] = (xs[m] + sh[0]) - X[3 * (int64_t)p];
    off[4 * dst + 1] = (xs[n + m] + sh[1]) - X[3 * (int64_t)p + 1];
    off[4 * dst + 2] = (xs[2 * n + m] + sh[2]) - X[3 * (int64_t)p + 2];
    off[4 * dst + 3] = __longlong_as_double((long long)sps[m]);
    srcp[dst] = p;
  }
}

// List build: every (i, image of j) with |x_j + H s - x_i|^2 < cs2, written
// slot-major (slot q of sorted atom i at q n + i) so that the interaction
// kernel's lanes read consecutive words. Pass 0 only counts. One wavefront per
// cell: the stencil entries are staged in LDS with the xe index of their first
// atom (the real cell, or the ghost cell of that image); lanes hold 64
// candidates at a time and test them against each centre of the cell in turn,
// a ballot compacting the hits (so each row comes out in stencil order).
// status[1]: largest degree; status[2]: some row exceeded ecap.
__global__ __launch_bounds__(64) void k_verlet_build(const Grid* __restrict__ grids,
                                                     const int32_t* __restrict__ cell_cfg, int64_t n,
                                                     const double* __restrict__ xs,
                                                     const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ gstart, int pass, int ecap,
                                                     int32_t* __restrict__ ncount, int32_t* __restrict__ nbr,
                                                     int32_t* __restrict__ status) {
  __shared__ int s_pref[kWave + 1], s_start[kWave], s_base[kWave];
  __shared__ double s_sh[kWave][3];
  __shared__ float s_ax[kWave], s_ay[kWave], s_az[kWave];  // the centres, relative to the cell's first atom
  __shared__ int s_zero;
  const int cell = blockIdx.x;
  const int lane = threadIdx.x;
  const int b = start[cell], cnt = start[cell + 1] - b;
  if (cnt == 0) return;
  const Grid& g = grids[cell_cfg[cell]];
  int key[3];
  cell_key(g, cell, key);
  const double rx = xs[b], ry = xs[n + b], rz = xs[2 * n + b];
  const double cs2 = g.cs2;
  const float c2hi = (float)(cs2 * (1.0 + 1e-4));  // FP32 prefilter ahead of the FP64 test
  const int w0 = 2 * g.radius[0] + 1, w1 = 2 * g.radius[1] + 1;
  for (int ag = 0; ag < cnt; ag += kWave) {
    const int na = min(kWave, cnt - ag);
    const int64_t mine = b + ag + lane;
    int pos = 0;  // the centre's next slot (lane of the same index)
    if (lane < na) {
      s_ax[lane] = (float)(xs[mine] - rx);
      s_ay[lane] = (float)(xs[n + mine] - ry);
      s_az[lane] = (float)(xs[2 * n + mine] - rz);
    }
    for (int e0 = 0; e0 < g.nstencil; e0 += kWave) {
      const int e = e0 + lane;
      int c = 0;
      if (lane == 0) s_zero = -1;
      __syncthreads();
      if (e < g.nstencil) {
        int lat[3];
        bool zero;
        const int nc = stencil_entry(g, key, e, lat, &zero);
        if (nc >= 0) {
          s_start[lane] = start[nc];
          c = start[nc + 1] - start[nc];
          double sh[3];
          exact_shift(g, lat, sh);
          for (int q = 0; q < 3; ++q) s_sh[lane][q] = sh[q];
          if (lat[0] == 0 && lat[1] == 0 && lat[2] == 0) {
            s_base[lane] = start[nc];  // a real cell: the sorted atoms themselves
          } else {                     // the ghost cell of this image, in the extended grid
            const int o[3] = {e % w0 - g.radius[0], (e / w0) % w1 - g.radius[1], e / (w0 * w1) - g.radius[2]};
            int loc = 0;
            for (int a = 2; a >= 0; --a) loc = loc * g.edims[a] + key[a] + o[a] + (g.pb[a] ? g.radius[a] : 0);
            s_base[lane] = (int)n + gstart[g.ecell0 + loc];
          }
          if (zero) s_zero = lane;
        }
      }
      const int incl = wave_incl_scan(c);
      s_pref[lane + 1] = incl;
      if (lane == 0) s_pref[0] = 0;
      const int total = __shfl(incl, kWave - 1, kWave);
      __syncthreads();
      const int ez = s_zero;
      for (int t0 = 0; t0 < total; t0 += kWave) {
        const int t = t0 + lane;
        const bool cand = t < total;
        int m = 0, lo = 0, entry = -1;
        double xj = 0, yj = 0, zj = 0;
        float fx = 1e30f, fy = 1e30f, fz = 1e30f;
        if (cand) {
          int hi = kWave;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_pref[mid] <= t)
              lo = mid;
            else
              hi = mid;
          }
          m = s_start[lo] + (t - s_pref[lo]);
          entry = s_base[lo] + (t - s_pref[lo]);
          xj = xs[m] + s_sh[lo][0];
          yj = xs[n + m] + s_sh[lo][1];
          zj = xs[2 * n + m] + s_sh[lo][2];
          fx = (float)(xj - rx);
          fy = (float)(yj - ry);
          fz = (float)(zj - rz);
        }
        for (int a = 0; a < na; ++a) {
          const int64_t k = b + ag + a;
          bool hit = false;
          const float ux = fx - s_ax[a], uy = fy - s_ay[a], uz = fz - s_az[a];
          if (ux * ux + uy * uy + uz * uz < c2hi && !(lo == ez && m == (int)k)) {
            const double dx = xj - xs[k], dy = yj - xs[n + k], dz = zj - xs[2 * n + k];
            hit = dx * dx + dy * dy + dz * dz < cs2;
          }
          const uint64_t mask = __ballot(hit);
          const int base = __shfl(pos, a, kWave);  // outside the branch: the source lane must be active
          const int slot = base + __popcll(mask & ((1ull << lane) - 1));
          if (pass && hit && slot < ecap) nbr[(int64_t)slot * n + k] = entry;
          if (lane == a) pos += __popcll(mask);
        }
      }
      __syncthreads();
    }
    if (lane < na) ncount[mine] = min(pos, ecap);
    int dmax = lane < na ? pos : 0;  // one atomic per wavefront (same-address atomics serialise)
    int dsum = lane < na ? min(pos, ecap) : 0;
    for (int off = kWave / 2; off > 0; off >>= 1) {
      dmax = max(dmax, __shfl_xor(dmax, off, kWave));
      dsum += __shfl_xor(dsum, off, kWave);
    }
    if (lane == 0) {
      atomicMax(status + 1, dmax);
      if (dmax > ecap) atomicOr(status + 2, 1);  // capacity too small: the host grows it and builds again
      if (pass) atomicAdd(reinterpret_cast<unsigned long long*>(status + 4), (unsigned long long)dsum);
    }
  }
}

// Per step: xe = off + X for every atom and ghost entry, and the largest
// squared displacement of the caller's positions since the build, as the
// bits of a nonnegative double (monotone as unsigned) in words[parity]; the
// other word is cleared for the next step.
__global__ __launch_bounds__(256) void k_verlet_refresh(int64_t n, int64_t n_ext, const double4* __restrict__ off,
                                                        const int32_t* __restrict__ srcp,
                                                        const double* __restrict__ X,
                                                        const double* __restrict__ Xref, double4* __restrict__ xe,
                                                        unsigned long long* __restrict__ words, int parity) {
  __shared__ double wmax[256 / kWave];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_ext) {
    const int64_t p = srcp[t];
    const double4 o = off[t];
    xe[t] = make_double4(o.x + X[3 * p], o.y + X[3 * p + 1], o.z + X[3 * p + 2], o.w);
  } else if (t == n_ext) {  // the sentinel: beyond every cutoff, squares finite in FP32
    xe[t] = make_double4(1e15, 1e15, 1e15, __longlong_as_double(0ll));
  }
  double m2 = 0;
  if (t < n) {
    const double dx = X[3 * t] - Xref[3 * t], dy = X[3 * t + 1] - Xref[3 * t + 1], dz = X[3 * t + 2] - Xref[3 * t + 2];
    m2 = dx * dx + dy * dy + dz * dz;
  }
  for (int s = kWave / 2; s > 0; s >>= 1) m2 = fmax(m2, __shfl_xor(m2, s, kWave));
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) wmax[wid] = m2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
    for (int w = 1; w < (int)(blockDim.x / kWave); ++w) m = fmax(m, wmax[w]);
    if (m > 0) atomicMax(&words[parity], (unsigned long long)__double_as_longlong(m));
    if (blockIdx.x == 0) words[1 - parity] = 0;
  }
}

// Pair energy / force / virial from the list. A wavefront takes 64 / P
// consecutive sorted atoms; the P lanes of an atom take every P-th slot of its
// row, so a row of ~80 entries is ~80 / P dependent steps and the slot-major
// layout keeps each step's index loads coalesced. The body is branch-free over
// U slots at a time (unused slots read the far sentinel entry xe[n_ext]), so
// U index loads and U position gathers are in flight per lane. d and r^2 in
// FP64 and membership r^2 < c^2 in FP64; the model in T with T partial sums,
// the sums in FP64, reduced over the P lanes of an atom in fixed order.
// Records per wavefront (E, xx yy zz xy xz yz of d (x) g) or, for a batch, per atom.
constexpr int kListParts = 4;   // lanes per atom
constexpr int kListUnroll = 4;  // slots per lane per step; the row capacity is a multiple of P U
// Membership: r^2 < c^2 of the current FP64 displacement; a pair whose r^2
// lies within a relative 1e-10 of c^2 is re-decided by exact_image_member from
// the caller's positions, as a fresh cell build decides it, so the pair set
// is that of a fresh build bit for bit (X, grids, atom_cfg and srcp are read
// only there). AUDIT: also add every accumulated pair to the per-atom
// pair-set digest.
template <typename T, int KIND, bool AUDIT>
__global__ __launch_bounds__(64) void k_pair_list(int64_t n, int sentinel, const int32_t* __restrict__ ncount,
                                                  const int32_t* __restrict__ nbr, const double4* __restrict__ xe,
                                                  const int32_t* __restrict__ perm, const double* __restrict__ params,
                                                  int nz, double c2, LJConsts lj, uint32_t mask,
                                                  double* __restrict__ F, double* __restrict__ rec, int per_atom,
                                                  AuditBufs au, const Grid* __restrict__ grids,
                                                  const int32_t* __restrict__ atom_cfg,
                                                  const int32_t* __restrict__ srcp, const double* __restrict__ X) {
  constexpr int P = kListParts, U = kListUnroll, NA = kWave / P;
  extern __shared__ __align__(16) unsigned char dyn[];  // SmoothPair: nz x nz amplitudes, then -6 A / c^2
  T* sA = (T*)dyn;
  T* sA6 = sA + nz * nz;
  const int lane = threadIdx.x;
  if (KIND == PFW_PAIR_SMOOTH) {
    for (int k = lane; k < nz * nz; k += kWave) {
      const double a = params[k];
      sA[k] = (T)a;
      sA6[k] = (T)(-6.0 * a / c2);
    }
  }
  __syncthreads();
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int part = lane / NA;
  const int64_t i = (int64_t)blk * NA + lane % NA;
  const bool live = i < n;
  const int64_t ii = live ? i : 0;
  const T invc2 = (T)(1.0 / c2);
  const T c4e = (T)lj.c4e, c24e = (T)lj.c24e, sig2 = (T)lj.sig2, eshift = (T)lj.eshift;
  const double4 xi = xe[ii];
  const int cnt = live ? ncount[i] : 0;
  const int zi = (int)__double_as_longlong(xi.w);
  int kmax = cnt;
  for (int s = kWave / 2; s > 0; s >>= 1) kmax = max(kmax, __shfl_xor(kmax, s, kWave));
  T er = 0, gxr = 0, gyr = 0, gzr = 0, w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0;
  // the next step's indices are loaded while this step's positions arrive and
  // its pairs are computed; the list has one spare step of slots past the row
  // capacity (a multiple of P U), so the prefetch after the last step stays in bounds
  int jn[U];
  const int64_t stride = (int64_t)P * U * n;
  const int32_t* row = nbr + (int64_t)part * U * n + ii;
#pragma unroll
  for (int u = 0; u < U; ++u) jn[u] = row[(int64_t)u * n];
  for (int k0 = part * U; k0 < kmax; k0 += P * U) {
    int j[U];
#pragma unroll
    for (int u = 0; u < U; ++u) j[u] = k0 + u < cnt ? jn[u] : sentinel;
    double4 xj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xj[u] = xe[j[u]];
    row += stride;
#pragma unroll
    for (int u = 0; u < U; ++u) jn[u] = row[(int64_t)u * n];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double ddx = xj[u].x - xi.x, ddy = xj[u].y - xi.y, ddz = xj[u].z - xi.z;
      const double rr = ddx * ddx + ddy * ddy + ddz * ddz;
      bool hit = rr < c2;
      if (fabs(rr - c2) <= c2 * 1e-10) {  // rare: decide as a fresh build would
        const double d[3] = {ddx, ddy, ddz};
        const int oi = perm[i], oj = srcp[j[u]];
        const Grid& g = grids[atom_cfg[oi]];
        int S[3];
        audit_shift(g, d, X, oi, oj, S);
        hit = exact_image_member(g, X, oi, oj, S);
      }
      if constexpr (AUDIT) {
        if (hit) {
          const double d[3] = {ddx, ddy, ddz};
          const int oi = perm[i];
          audit_add(au, grids[atom_cfg[oi]], d, oi, srcp[j[u]]);
        }
      }
      const int zj = (int)__double_as_longlong(xj[u].w);
      const T dx = (T)ddx, dy = (T)ddy, dz = (T)ddz, r2 = (T)rr;
      T e, coef;
      if (KIND == PFW_PAIR_SMOOTH) {
        const T a = sA[zi * nz + zj], a6 = sA6[zi * nz + zj];
        const T qq = hit ? (T)1 - r2 * invc2 : (T)0;
        const T q2 = qq * qq;
        e = a * q2 * qq;
        coef = a6 * q2;
      } else {
        T inv;
        if constexpr (sizeof(T) == 4)
          inv = __builtin_amdgcn_rcpf(r2);
        else
          inv = (T)1 / r2;
        const T s2 = sig2 * inv;
        const T s6 = s2 * s2 * s2;
        e = hit ? c4e * (s6 * s6 - s6) - eshift : (T)0;
        coef = hit ? c24e * (s6 - (T)2 * s6 * s6) * inv : (T)0;
      }
      const T gx = coef * dx, gy = coef * dy, gz = coef * dz;
      er += e;
      gxr += gx;
      gyr += gy;
      gzr += gz;
      w0 += dx * gx;
      w1 += dy * gy;
      w2 += dz * gz;
      w3 += dx * gy;
      w4 += dx * gz;
      w5 += dy * gz;
    }
  }
  double v[10] = {(double)er, (double)gxr, (double)gyr, (double)gzr, (double)w0,
                  (double)w1, (double)w2,  (double)w3,  (double)w4,  (double)w5};
  for (int s = NA; s < kWave; s <<= 1)  // the P lanes of an atom, fixed order
#pragma unroll
    for (int c = 0; c < 10; ++c) v[c] += __shfl_xor(v[c], s, kWave);
  if (live && part == 0 && (mask & PFW_FORCES)) {
    const int64_t o = perm[i];
    F[3 * o] = v[1];
    F[3 * o + 1] = v[2];
    F[3 * o + 2] = v[3];
  }
  double r[7] = {v[0], v[4], v[5], v[6], v[7], v[8], v[9]};
  if (per_atom) {
    if (live && part == 0)
      for (int c = 0; c < 7; ++c) rec[7 * i + c] = r[c];
  } else {
    if (!live || part != 0)
      for (int c = 0; c < 7; ++c) r[c] = 0;
    for (int c = 0; c < 7; ++c) r[c] = wave_sum(r[c]);
    if (lane == 0)
      for (int c = 0; c < 7; ++c) rec[7 * (int64_t)blk + c] = r[c];
  }
}

// ---------------------------------------------------------------- host

template <typename P>
static void dfree(P*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

void verlet_free(pfw_ctx* ctx) {
  VerletDev& V = ctx->verlet;
  dfree(V.ecfg0);
  dfree(V.gcount);
  dfree(V.gstart);
  dfree(V.gblock);
  dfree(V.nbr);
  dfree(V.ncount);
  dfree(V.off);
  dfree(V.xe);
  dfree(V.srcp);
  dfree(V.Xref);
  dfree(V.maxd2);
  dfree(V.status);
  dfree(V.part);
  if (V.h_maxd2) (void)hipHostFree(V.h_maxd2);
  if (V.h_status) (void)hipHostFree(V.h_status);
  V.h_maxd2 = nullptr;
  V.h_status = nullptr;
  V.valid = false;
  V.ecap = 0;
  V.n_ext = V.ext_capacity = V.list_capacity = V.part_capacity = 0;
  V.parity = 0;
}

template <typename P>
static void dalloc(P*& p, size_t count) {
  dfree(p);
  PFW_HIP(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(*p)));
}

// Buffers whose size follows the system and grid (allocated at the first build
// after the system or the skin changed).
static void verlet_setup(pfw_ctx* ctx) {
  VerletDev& V = ctx->verlet;
  const int64_t n = ctx->sys.n;
  const int64_t ec = V.ext_cells;
  dalloc(V.ecfg0, ctx->ncfg + 1);
  dalloc(V.gcount, ec);
  dalloc(V.gstart, ec + 1);
  dalloc(V.gblock, (ec + 1023) / 1024 + 1);
  dalloc(V.ncount, n);
  dalloc(V.Xref, 3 * n);
  dalloc(V.maxd2, 2);
  dalloc(V.status, 6);
  dalloc(V.part, 7 * n);
  V.part_capacity = 7 * n;
  if (!V.h_maxd2) PFW_HIP(hipHostMalloc((void**)&V.h_maxd2, 2 * sizeof(unsigned long long), hipHostMallocDefault));
  if (!V.h_status) PFW_HIP(hipHostMalloc((void**)&V.h_status, 6 * sizeof(int32_t), hipHostMallocDefault));
  std::vector<int32_t> e0(ctx->ncfg + 1);
  for (int k = 0; k < ctx->ncfg; ++k) e0[k] = ctx->grids[k].ecell0;
  e0[ctx->ncfg] = (int32_t)ec;
  hipStream_t st = ctx->stream;
  PFW_HIP(hipMemcpyAsync(V.ecfg0, e0.data(), e0.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  PFW_HIP(hipMemsetAsync(V.maxd2, 0, 2 * sizeof(unsigned long long), st));
  PFW_HIP(hipStreamSynchronize(st));  // the staging vector goes out of scope
  V.parity = 0;
}

static void build_pass(pfw_ctx* ctx, int pass) {
  VerletDev& V = ctx->verlet;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  PFW_HIP(hipMemsetAsync(V.status, 0, 6 * sizeof(int32_t), st));
  hipLaunchKernelGGL(k_verlet_build, dim3(ctx->ncells), dim3(kWave), 0, st, s.grids, s.cell_cfg, s.n, s.xs, s.start,
                     V.gstart, pass, pass ? V.ecap : 0, V.ncount, V.nbr, V.status);
  PFW_HIP(hipGetLastError());
  PFW_HIP(hipMemcpyAsync(V.h_status, V.status, 6 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipStreamSynchronize(st));
  if (pass) V.entries = (int64_t)(*reinterpret_cast<const unsigned long long*>(V.h_status + 4));
}

static void grow_list(pfw_ctx* ctx, int degree) {
  VerletDev& V = ctx->verlet;
  const int64_t n = ctx->sys.n;
  const int step = kListParts * kListUnroll;  // the interaction kernel reads whole steps of slots
  V.ecap = std::max(V.ecap, (degree + std::max(8, degree / 8) + step - 1) / step * step);  // headroom for the next builds
  if ((int64_t)V.ecap * n > (int64_t)1 << 33) throw Error(PFW_INVALID_ARGUMENT, "Verlet list too large");
  const int64_t slots = (int64_t)(V.ecap + step) * n;  // + the interaction kernel's prefetch step
  if (slots > V.list_capacity) {
    dalloc(V.nbr, (size_t)slots);
    V.list_capacity = slots;
  }
}

// Fresh cells at cutoff + skin, ghost entries and the list (gpu_build_cells! +
// gpu_connectivity!, gpu.jl:497-541).
static void verlet_rebuild(pfw_ctx* ctx) {
  VerletDev& V = ctx->verlet;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  if (!V.gcount) verlet_setup(ctx);
  nl_build(ctx);
  record_stage(ctx, 3);
  const int64_t ec = V.ext_cells;
  hipLaunchKernelGGL(k_ext_count, dim3(nblk(ec, 256)), dim3(256), 0, st, ec, V.ecfg0, ctx->ncfg, s.grids, s.start,
                     V.gcount);
  PFW_HIP(hipGetLastError());
  scan_counts(ctx, V.gcount, (int)ec, V.gstart, V.gblock);
  int32_t ghosts = 0;
  PFW_HIP(hipMemcpyAsync(&ghosts, V.gstart + ec, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipStreamSynchronize(st));
  if (ghosts < 0 || n + (int64_t)ghosts > ((int64_t)1 << 31) - 1)
    throw Error(PFW_INVALID_ARGUMENT, "too many ghost images for the Verlet list");
  V.n_ext = n + ghosts;
  if (V.n_ext + 1 > V.ext_capacity) {  // + the far sentinel entry xe[n_ext]
    const int64_t cap = V.n_ext + 1 + V.n_ext / 8;
    dalloc(V.off, 4 * cap);
    dalloc(V.xe, 4 * cap);
    dalloc(V.srcp, cap);
    V.ext_capacity = cap;
  }
  hipLaunchKernelGGL(k_real_fill, dim3(nblk(n, 256)), dim3(256), 0, st, n, s.xs, s.sps, s.perm, s.X, V.off, V.srcp);
  if (ghosts > 0)
    hipLaunchKernelGGL(k_ghost_fill, dim3((unsigned)ec), dim3(kWave), 0, st, ec, V.ecfg0, ctx->ncfg, s.grids,
                       s.start, V.gstart, n, s.xs, s.sps, s.perm, s.X, V.off, V.srcp);
  PFW_HIP(hipGetLastError());
  PFW_HIP(hipMemcpyAsync(V.Xref, s.X, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (V.ecap == 0) {  // first build: size the rows
    build_pass(ctx, 0);
    grow_list(ctx, V.h_status[1]);
  }
  build_pass(ctx, 1);
  if (V.h_status[2]) {  // a row outgrew the capacity since the last build
    grow_list(ctx, V.h_status[1]);
    build_pass(ctx, 1);
    if (V.h_status[2]) throw Error(PFW_DEVICE, "Verlet list build overflowed after resizing");
  }
  V.valid = true;
  V.gen = ctx->build_gen;
  ++V.rebuilds;
}

template <typename T, int KIND>
static void launch_list(pfw_ctx* ctx, uint32_t mask, bool per_atom) {
  VerletDev& V = ctx->verlet;
  const int64_t n = ctx->sys.n;
  const size_t dyn = KIND == PFW_PAIR_SMOOTH ? 2 * sizeof(T) * ctx->pair.nz * ctx->pair.nz : 0;
  if (ctx->audit) {
    const AuditBufs au = audit_begin(ctx);
    hipLaunchKernelGGL((k_pair_list<T, KIND, true>), dim3(nblk(n, kWave / kListParts)), dim3(kWave), dyn,
                       ctx->stream, n, (int)V.n_ext, V.ncount, V.nbr, (const double4*)V.xe, ctx->sys.perm,
                       ctx->pair.d_params, ctx->pair.nz, ctx->pair.cutoff * ctx->pair.cutoff, lj_consts(ctx->pair),
                       mask, ctx->d_F, V.part, (int)per_atom, au, ctx->sys.grids, ctx->sys.atom_cfg, V.srcp,
                       ctx->sys.X);
  } else {
    hipLaunchKernelGGL((k_pair_list<T, KIND, false>), dim3(nblk(n, kWave / kListParts)), dim3(kWave), dyn,
                       ctx->stream, n, (int)V.n_ext, V.ncount, V.nbr, (const double4*)V.xe, ctx->sys.perm,
                       ctx->pair.d_params, ctx->pair.nz, ctx->pair.cutoff * ctx->pair.cutoff, lj_consts(ctx->pair),
                       mask, ctx->d_F, V.part, (int)per_atom, AuditBufs{}, ctx->sys.grids, ctx->sys.atom_cfg,
                       V.srcp, ctx->sys.X);
  }
}

// One step on the list: refresh of the extended positions with the
// displacement check, the interaction, the FP64 reductions and the output
// copies. Evaluated speculatively on the current list; when the displacement
// read back with the outputs shows that the list may miss a pair, the list is
// rebuilt and the step repeated (the reference instead reads the displacement
// before evaluating, gpu.jl:516-520).
void verlet_pair_run(pfw_ctx* ctx, uint32_t mask, int precision) {
  ctx->host_written = false;
  VerletDev& V = ctx->verlet;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  bool fresh = false;
  if (!V.valid || V.gen != ctx->build_gen) {
    verlet_rebuild(ctx);
    fresh = true;
  } else {
    for (int k = 0; k < 4; ++k) record_stage(ctx, k);
  }
  const bool per_atom = ctx->ncfg > 1;
  const bool fp32 = precision == PFW_MIXED;
  for (;;) {
    hipLaunchKernelGGL(k_verlet_refresh, dim3(nblk(V.n_ext + 1, 256)), dim3(256), 0, st, n, V.n_ext,
                       (const double4*)V.off, V.srcp, s.X, V.Xref, (double4*)V.xe, V.maxd2, V.parity);
    PFW_HIP(hipGetLastError());
    record_stage(ctx, 4);
    if (ctx->pair.kind == PFW_PAIR_SMOOTH) {
      if (fp32)
        launch_list<float, PFW_PAIR_SMOOTH>(ctx, mask, per_atom);
      else
        launch_list<double, PFW_PAIR_SMOOTH>(ctx, mask, per_atom);
    } else {
      if (fp32)
        launch_list<float, PFW_PAIR_LJ>(ctx, mask, per_atom);
      else
        launch_list<double, PFW_PAIR_LJ>(ctx, mask, per_atom);
    }
    PFW_HIP(hipGetLastError());
    record_stage(ctx, 5);
    if (per_atom) {
      batch_reduce(ctx, V.part, 7, false, 0);
    } else {
      reduce_outputs(ctx, V.part, (int64_t)nblk(n, kWave / kListParts), 7, 0, nullptr);
    }
    PFW_HIP(hipGetLastError());
    copy_outputs(ctx);
    PFW_HIP(hipMemcpyAsync(V.h_maxd2, V.maxd2 + V.parity, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    record_stage(ctx, 6);
    PFW_HIP(hipStreamSynchronize(st));
    V.parity ^= 1;
    const double m2 = __builtin_bit_cast(double, (unsigned long long)V.h_maxd2[0]);
    V.last_max_disp = std::sqrt(m2);
    // gpu_skin!'s test (gpu.jl:519): rebuild once 2 max|dx| >= skin (1 - 16 eps)
    const bool stale = !(2.0 * V.last_max_disp < V.skin * (1.0 - 16.0 * 2.220446049250313e-16));
    if (fresh || !stale) break;
    verlet_rebuild(ctx);
    fresh = true;
  }
  if (!fresh) ++V.reuses;
}

}  // namespace pfw
