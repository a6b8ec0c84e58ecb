// Pair energy / force / virial on gfx950 (SmoothPair and energy-shifted
// Lennard-Jones), full directed list without atomics.
//
// Reference: the half-list executor src/submodules/NeighbourList/execution.jl:173-217
// with the callbacks src/submodules/PairPotentials/models.jl:153-167 and
// lennard_jones.jl:5-17; the GPU boundary src/submodules/NeighbourList/gpu.jl:236-402.
// Every directed pair (i -> j, S) contributes g = (dE/dr / r) d to F_i, half its
// energy to E and -1/2 d (x) g to W, so each unordered pair image is counted once.
//
// Layout: one 64-lane wavefront per cell. The neighbour cells' atoms are staged
// in LDS as coordinates relative to the cell's first atom (computed in FP64,
// stored in T). Lanes are (atom, slice) pairs: a cell of n_i atoms uses
// S = min(4, 64 / n_i) slices, slice s testing one contiguous range of each
// LDS chunk (so its candidate reads take immediate offsets). Hits are
// compacted per lane into a 64-bit mask before the model body runs. Membership
// inside a thin band around the cutoff is re-decided with the exact FP64
// formula of the oracle, so the pair set is bit-exact in both precisions.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "pfw_device.h"

namespace pfw {

constexpr int kEnt = 32;     // stencil entries per entry chunk (the 27 of radius 1 in one)
constexpr int kSlices = 4;   // most lanes per atom
// Candidates per LDS chunk. Slice s takes the chunk's candidates s, s + S,
// s + 2 S, ... (interleaved, so that every slice sees every stencil cell and
// the slices' hit counts stay alike), stored contiguously: L entries, L a
// multiple of 8 (the test unroll), starting L + 1 entries after the previous
// slice so that the slices' concurrent reads fall in different LDS banks. The
// slices' tails past the chunk hold far sentinels.
template <typename T>
constexpr int kCandOf = 256;
template <typename T>
constexpr int kCandPadOf = kCandOf<T> + 9 * kSlices;
static_assert(kCandOf<float> % 64 == 0 && kCandOf<double> % 64 == 0, "the fill walks a chunk in whole waves");
// Candidate-to-entry table over a whole entry chunk (larger neighbourhoods
// search the entry offsets instead).
constexpr int kEofMax = 768;

// Enumeration order of a neighbourhood's candidates. The entry-major index t
// (stencil entry, then the cell's sorted atoms) is visited at position u with
// t = u q mod total, q a prime that does not divide total (so the map is a
// bijection). Consecutive positions then jump across the whole neighbourhood,
// so every chunk and every 64-candidate block of a slice holds a spatially
// even sample of it: the lanes' hit counts per block stay alike and the
// divergent hit loop runs fewer rounds (on the B2 256k workload 30.5 instead
// of 40.8 hit iterations per wavefront, counted on the host).
__device__ inline uint32_t scramble_mul(int total) {
  if (total % 37) return 37;
  if (total % 41) return 41;
  if (total % 43) return 43;
  if (total % 47) return 47;
  return 1;
}
// t = u q mod total for 0 <= u < total
__device__ inline int scramble_of(uint32_t u, uint32_t q, int total) {
  if (total < (1 << 25)) return (int)((u * q) % (uint32_t)total);  // u q < 2^31
  return (int)(((uint64_t)u * q) % (uint64_t)total);
}

template <typename T>
struct Band;
template <>
struct Band<float> {
  static constexpr double rel = 1e-4;
};
template <>
struct Band<double> {
  static constexpr double rel = 1e-11;
};

// A staged candidate: coordinates relative to the cell's first atom and, in
// FP32, its squared norm (pre-shrunk for the prefilter), read with one LDS
// instruction.
template <typename T>
struct alignas(4 * sizeof(T)) Cand {
  T x, y, z, w;
};
// FP64: no squared norm (the difference form is used; LDS is scarcer)
template <>
struct Cand<double> {
  double x, y, z;
};

// Relative slack of the |xi|^2 + |xj|^2 - 2 xi.xj prefilter: it exceeds the
// rounding error of that form (a few ulp of |xi|^2 + |xj|^2), so with both
// squares scaled by (1 - slack) the prefilter never drops a pair the exact
// difference form keeps; extra candidates are rejected in the pair loop.
template <typename T>
struct DotSlack;
template <>
struct DotSlack<float> {
  static constexpr float value = 16 * 1.1920929e-7f;
};
template <>
struct DotSlack<double> {
  static constexpr double value = 16 * 2.220446049250313e-16;
};

// The test of a candidate against the lane's atom, as r^2 - c_hi^2 (negative
// for a candidate of the band-widened sphere): FP32 by the dot form (one FMA
// chain, with -c_hi^2 folded into the lane's constant), FP64 by differences.
__device__ inline float prefilter_d(const Cand<float>& c, float, float, float, float m2x, float m2y, float m2z,
                                    float base, float) {
  return m2x * c.x + (m2y * c.y + (m2z * c.z + (base + c.w)));
}
__device__ inline double prefilter_d(const Cand<double>& c, double xi, double yi, double zi, double, double, double,
                                     double, double c2hi) {
  const double dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
  return (dx * dx + dy * dy + dz * dz) - c2hi;
}
// The sign bit of a test value (its high word for FP64).
__device__ inline uint32_t sign_word(float v) { return __float_as_uint(v); }
__device__ inline uint32_t sign_word(double v) { return (uint32_t)__double2hiint(v); }

// An entry's Cartesian image shift: the staged FP64 copy, or (FP32 kernel)
// formed again from its integer lattice shift.
template <typename T>
__device__ inline void entry_shift(const Grid& g, const int* lat, const double* staged, double* sh) {
  if constexpr (sizeof(T) == 8) {
    for (int a = 0; a < 3; ++a) sh[a] = staged[a];
  } else {
    exact_shift(g, lat, sh);
  }
}

// Per entry chunk: candidate count, entry of the zero offset (or -1), entries
// in the chunk, candidate-to-entry table in use, cull in use, enumeration
// multiplier and its step per 64 positions.
struct EntryInfo {
  int total, ez, ne;
  bool table, cull;
  uint32_t sq, tstep;
};

template <typename T>
struct Pair2;
template <>
struct Pair2<float> {
  using type = float2;
};
template <>
struct Pair2<double> {
  using type = double2;
};

// Five wavefronts per SIMD for the FP32 kernel (at most 96 registers): the
// candidate fills and LDS round trips of one wavefront hide behind the tests
// of the others.
// AUDIT: also add every accumulated pair to the per-atom pair-set digest.
template <typename T, int KIND, bool AUDIT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : 4))) void k_pair(const Grid* __restrict__ grids, const int32_t* __restrict__ cell_cfg,
                                             int64_t n, const int32_t* __restrict__ start,
                                             const double* __restrict__ xs, const float4* __restrict__ xr,
                                             const float4* __restrict__ cbox,
                                             const int32_t* __restrict__ sps,
                                             const int32_t* __restrict__ perm, const double* __restrict__ params,
                                             int nz, LJConsts lj, uint32_t mask, double* __restrict__ F,
                                             double* __restrict__ partials, int ncells, AuditBufs au,
                                             int max_slices) {
  // candidates as (x, y, z, (1 - slack) |x|^2), then far sentinels up to the end
  // of the last slice's range; after the candidate loops of an atom group the
  // same bytes hold the per-slice force sums
  constexpr int kCand = kCandOf<T>;
  constexpr size_t kCoordBytes = kCandPadOf<T> * sizeof(Cand<T>), kRedBytes = kWave * 3 * sizeof(double);
  __shared__ __align__(16) unsigned char cbuf[kCoordBytes > kRedBytes ? kCoordBytes : kRedBytes];
  Cand<T>* cand = (Cand<T>*)cbuf;
  double(*red)[3] = (double(*)[3])cbuf;
  __shared__ int cm[kCandPadOf<T>];  // per candidate position: (stencil entry of the chunk) << 8 | species
  __shared__ int e_pref[kEnt + 1], e_start[kEnt], e_lat[kEnt][3];
  __shared__ uint8_t e_of[kEofMax];  // stencil entry of each entry-major candidate index of the entry chunk
  __shared__ int selfpos[kWave];     // enumeration position of each home atom's self pair (atom group), or -1
  // FP64: each entry's Cartesian image shift (FP32 forms it again in the rare
  // exact-membership path, which keeps the LDS footprint at five wavefronts per SIMD)
  __shared__ double e_sh[sizeof(T) == 8 ? kEnt : 1][3];
  // FP32: each entry's cell origin (its first atom) plus image shift, relative
  // to this cell's first atom; a candidate is its cell-relative position plus this
  __shared__ __align__(16) float e_off[sizeof(T) == 4 ? kEnt : 1][4];
  extern __shared__ __align__(16) unsigned char dyn[];  // SmoothPair: per (z_i, z_j) the pair (A, -6 A / c^2)
  using T2 = typename Pair2<T>::type;
  const T2* sAA = (const T2*)dyn;
  __shared__ int e_zero;

  const int lane = threadIdx.x;
  const int cell = xcd_remap(blockIdx.x, gridDim.x);
  const Grid& g = grids[cell_cfg ? cell_cfg[cell] : 0];  // no table for one configuration: no dependent load
  const T c2 = (T)g.c2;
  const T c2lo = (T)(g.c2 * (1.0 - Band<T>::rel)), c2hi = (T)(g.c2 * (1.0 + Band<T>::rel));
  const T invc2 = (T)(1.0 / g.c2);
  if (KIND == PFW_PAIR_SMOOTH) {
    for (int k = lane; k < nz * nz; k += kWave) {
      const double a = params[k];
      ((T2*)dyn)[k] = T2{(T)a, (T)(-6.0 * a / g.c2)};
    }
  }
  Cand<T> far;
  if constexpr (sizeof(T) == 4)
    far = Cand<T>{(T)0, (T)0, (T)0, (T)1e30};
  else
    far = Cand<T>{(T)1e30, (T)1e30, (T)1e30};
  const T c4e = (T)lj.c4e, c24e = (T)lj.c24e, sig2 = (T)lj.sig2, eshift = (T)lj.eshift;
  double eacc = 0;
  T w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0;
  __syncthreads();

  const int b = start[cell], ncnt = start[cell + 1] - b;
  if (ncnt > 0) {
    int key[3];
    cell_key(g, cell, key);
    const double rx = xs[b], ry = xs[n + b], rz = xs[2 * n + b];
    // One chunk of up to kEnt stencil entries: each entry's cell range, lattice
    // and offsets in LDS, the entry-major prefix of their candidates, the
    // candidate-to-entry table and the enumeration multiplier. A radius-1
    // stencil is one chunk, set up once per cell before the atom groups.
    auto setup_entries = [&](int e0) -> EntryInfo {
      int cntE = 0;
      const int e = e0 + lane;
      if (lane == 0) e_zero = -1;
      __syncthreads();
      if (lane < kEnt && e < g.nstencil) {
        int lat[3];
        bool zero;
        int nc = stencil_entry(g, key, e, lat, &zero);
        if (nc >= 0) {
          int s0 = start[nc];
          cntE = start[nc + 1] - s0;
          e_start[lane] = s0;
          double sh[3];
          exact_shift(g, lat, sh);
          for (int a = 0; a < 3; ++a) {
            if constexpr (sizeof(T) == 8) e_sh[lane][a] = sh[a];
            e_lat[lane][a] = lat[a];
          }
          if constexpr (sizeof(T) == 4) {
            e_off[lane][0] = (float)((xs[s0] + sh[0]) - rx);
            e_off[lane][1] = (float)((xs[n + s0] + sh[1]) - ry);
            e_off[lane][2] = (float)((xs[2 * n + s0] + sh[2]) - rz);
          }
          if (zero) e_zero = lane;
        }
      }
      const int incl = wave_incl_scan(cntE);  // lanes past kEnt add nothing
      if (lane < kEnt) e_pref[lane + 1] = incl;
      if (lane == 0) e_pref[0] = 0;
      EntryInfo r;
      r.total = __builtin_amdgcn_readfirstlane(__shfl(incl, kWave - 1, kWave));
      r.table = r.total <= kEofMax;
      if (r.table && lane < kEnt)  // each entry marks its candidates
        for (int t = incl - cntE; t < incl; ++t) e_of[t] = (uint8_t)lane;
      r.sq = __builtin_amdgcn_readfirstlane(scramble_mul(r.total));
      r.tstep = __builtin_amdgcn_readfirstlane(r.total > kWave ? (uint32_t)scramble_of((uint32_t)kWave, r.sq, r.total)
                                                                 : 0u);
      // culling records a candidate's index in its cell in 19 bits of its metadata
      r.cull = r.total < (1 << 19);
      r.ne = min(kEnt, g.nstencil - e0);  // entries of this chunk
      __syncthreads();
      r.ez = __builtin_amdgcn_readfirstlane(e_zero);
      return r;
    };
    const bool one_chunk = g.nstencil <= kEnt;
    EntryInfo ei{};
    if (one_chunk) ei = setup_entries(0);
    for (int ib = 0; ib < ncnt; ib += kWave) {
      const int nb = min(kWave, ncnt - ib);
      const int S = min(max_slices, kWave / nb);  // max_slices in 1..kSlices
      const int il = lane % nb, split = lane / nb;
      const bool on = split < S;
      const int is = b + ib + il;
      // the FP64 coordinates are read again in the rare exact-membership path
      // (fewer live registers, more wavefronts per SIMD)
      T xi, yi, zi;
      if constexpr (sizeof(T) == 4) {  // the cell-relative copy: the same values
        const float4 v = xr[is];
        xi = v.x;
        yi = v.y;
        zi = v.z;
      } else {
        xi = (T)(xs[is] - rx);
        yi = (T)(xs[n + is] - ry);
        zi = (T)(xs[2 * n + is] - rz);
      }
      // prefilter r^2 = |xi|^2 + |xj|^2 - 2 xi.xj, both squares shrunk by the
      // slack that bounds its rounding error, so it never misses a pair; the
      // test value is that minus c_hi^2
      const T m2x = (T)-2 * xi, m2y = (T)-2 * yi, m2z = (T)-2 * zi;
      const T base = (xi * xi + yi * yi + zi * zi) * ((T)1 - DotSlack<T>::value) - c2hi;
      const int zsp = sps[is];
      // the cell's box (its atoms' FP32 cell-relative positions) and the cull
      // radius: those coordinates differ from the exact ones by about 1e-6 A,
      // far inside the 1e-3 relative margin
      const float4 blo = cbox[2 * (int64_t)cell], bhi = cbox[2 * (int64_t)cell + 1];
      const T bx0 = (T)blo.x, bx1 = (T)bhi.x, by0 = (T)blo.y, by1 = (T)bhi.y, bz0 = (T)blo.z, bz1 = (T)bhi.z;
      const T cullc2 = (T)(g.c2 * (1.0 + 1e-3));
      double fx = 0, fy = 0, fz = 0;
      for (int e0 = 0; e0 < g.nstencil; e0 += kEnt) {
        if (!one_chunk) ei = setup_entries(e0);
        const int total = ei.total, ez = ei.ez, ne = ei.ne;
        const bool table = ei.table, cull = ei.cull;
        const uint32_t sq = ei.sq, tstep = ei.tstep;
        selfpos[lane] = -1;
        __syncthreads();
        // sorted index of a staged candidate from its metadata, entry el and u =
        // chunk start + slot (without the cull the slot is its enumeration position)
        auto cand_index = [&](int meta, int el, int u) -> int {
          return e_start[el] + (cull ? (meta >> 13) : scramble_of((uint32_t)u, sq, total) - e_pref[el]);
        };
        for (int t0 = 0; t0 < total; t0 += kCand) {
          const int tn = min(kCand, total - t0);
          // slice s at positions [s (Lcap + 1), s (Lcap + 1) + Lcap): room for the
          // whole chunk, of which only the candidates that pass the cull are kept
          const int Lcap = ((tn + S - 1) / S + 7) / 8 * 8;
          const int stride = Lcap + 1;
          const uint32_t invS = (65536u + S - 1) / S;  // t / S = (t invS) >> 16 for t < 4096
          // entry-major index of this lane's fill position t0 + t
          uint32_t tq = lane < tn ? (uint32_t)scramble_of((uint32_t)(t0 + lane), sq, total) : 0;
          int kept = 0;  // candidates of this chunk kept so far (the same in every lane)
          for (int p0 = 0; p0 < tn; p0 += kWave) {
            const int t = p0 + lane;
            const bool act = t < tn;
            const int tt = (int)tq;  // the candidate at position t0 + t
            tq += tstep;
            if (tq >= (uint32_t)total) tq -= (uint32_t)total;
            int lo = 0, m = 0;
            T x = 0, y = 0, z = 0;
            int spj = 0;
            if (act) {
              if (table) {
                lo = e_of[tt];
              } else {  // the last entry starting at or before tt (empty entries share its offset)
                int lo2 = 0, hi2 = ne;
                while (hi2 - lo2 > 1) {
                  const int mid = (lo2 + hi2) >> 1;
                  if (e_pref[mid] <= tt)
                    lo2 = mid;
                  else
                    hi2 = mid;
                }
                lo = lo2;
              }
              m = e_start[lo] + (tt - e_pref[lo]);
              if constexpr (sizeof(T) == 4) {
                const float4 v = xr[m];
                const float4 o = *(const float4*)e_off[lo];
                x = v.x + o.x;
                y = v.y + o.y;
                z = v.z + o.z;
                spj = __float_as_int(v.w);
              } else {
                x = (T)((xs[m] + e_sh[lo][0]) - rx);
                y = (T)((xs[n + m] + e_sh[lo][1]) - ry);
                z = (T)((xs[2 * n + m] + e_sh[lo][2]) - rz);
                spj = sps[m];
              }
            }
            // cull: a candidate farther than the cutoff from the box around the
            // atom group cannot pair with any of its atoms
            const T gx = fmax(fmax(bx0 - x, x - bx1), (T)0), gy = fmax(fmax(by0 - y, y - by1), (T)0),
                    gz = fmax(fmax(bz0 - z, z - bz1), (T)0);
            const bool keep = act && (!cull || gx * gx + gy * gy + gz * gz < cullc2);
            const uint64_t ball = __ballot(keep);
            if (keep) {
              const int slot =
                  kept + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ball >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ball, 0));
              if (lo == ez) {  // a home-cell atom: record where its self pair sits
                const int j = m - b - ib;
                if (j >= 0 && j < nb) selfpos[j] = t0 + slot;
              }
              const int q = (int)(((uint32_t)slot * invS) >> 16);
              const int pos = (slot - q * S) * stride + q;
              // culled chunks record each candidate's index in its cell for the
              // exact-membership path; otherwise slot = t and the order gives it
              const int meta = (cull ? (tt - e_pref[lo]) << 13 : 0) | (lo << 8) | spj;
              if constexpr (sizeof(T) == 4)
                cand[pos] = Cand<T>{x, y, z, (x * x + y * y + z * z) * ((T)1 - DotSlack<T>::value)};
              else
                cand[pos] = Cand<T>{x, y, z};
              cm[pos] = meta;
            }
            kept += __builtin_popcountll(ball);
          }
          const int L = ((kept + S - 1) / S + 7) / 8 * 8;  // <= Lcap
          // the gaps and the slices' tails past the kept candidates hold far sentinels
          for (int k = max(0, kept / S - 1) + lane; k <= L; k += kWave)
            for (int sl = 0; sl < S; ++sl)
              if (k == L || k * S + sl >= kept) cand[sl * stride + k] = far;
          __syncthreads();
          if (on) {
            // cnt (8 .. 32, a multiple of 8) tests of candidates p[0 .. cnt): bit
            // 31 - q flags candidate q; the bits are shifted in (h = 2 h + flag) so
            // that no per-test mask or shift is needed
            // the sign bit of each test value is shifted in with one alignbit
            auto round = [&](const Cand<T>* p, int cnt) -> uint32_t {
              uint32_t h = 0;
              for (int q0 = 0; q0 < cnt; q0 += 8) {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                  h = __builtin_amdgcn_alignbit(
                      h, sign_word(prefilter_d(p[q0 + q], xi, yi, zi, m2x, m2y, m2z, base, c2hi)), 31);
              }
              return h << (32 - cnt);
            };
            // the lane's own atom among this chunk's candidates: its place in this
            // lane's slice, or -1
            const int us = selfpos[il];
            const int tself = us >= 0 ? us - t0 : -1;
            int kself = -1;
            if (tself >= 0 && tself < kept) {
              const int q = (int)(((uint32_t)tself * invS) >> 16);
              if (tself - q * S == split) kself = q;
            }
            const Cand<T>* mine = cand + split * stride;
            const int* cmine = cm + split * stride;
            for (int r0 = 0; r0 < L; r0 += 64) {
              // two rounds, then their hits: bit b of `hits` flags the slice's candidate r0 + 63 - b
              const int n1 = min(32, L - r0), n2 = min(32, L - r0 - 32);
              uint64_t hits = (uint64_t)round(mine + r0, n1) << 32;
              if (n2 > 0) hits |= round(mine + r0 + 32, n2);
              if (kself >= r0 && kself < r0 + 64) hits &= ~(1ull << (63 - (kself - r0)));
              // this pair of rounds' sums (at most 64 pairs) in T, then added to the FP64 totals
              T er = 0, fxr = 0, fyr = 0, fzr = 0;
              while (hits) {
                const int k = r0 + 63 - __builtin_ctzll(hits);  // within the slice
                hits &= hits - 1;
                const Cand<T> c = mine[k];
                const int meta = cmine[k];
                const T dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
                const T r2 = dx * dx + dy * dy + dz * dz;
                const int el = (meta >> 8) & 31, zj = meta & 255;
                if (!(r2 < c2lo)) {
                  const int m = cand_index(meta, el, t0 + k * S + split);
                  double sh[3], d[3];
                  int img[3];
                  entry_shift<T>(g, e_lat[el], e_sh[sizeof(T) == 8 ? el : 0], sh);
                  if (!exact_member(g, xs[is], xs[n + is], xs[2 * n + is], xs[m], xs[n + m], xs[2 * n + m], e_lat[el],
                                    sh, d, img))
                    continue;
                }
                if constexpr (AUDIT) {
                  const int m = cand_index(meta, el, t0 + k * S + split);
                  double sh[3], d[3];
                  entry_shift<T>(g, e_lat[el], e_sh[sizeof(T) == 8 ? el : 0], sh);
                  for (int a = 0; a < 3; ++a) d[a] = (xs[a * n + m] + sh[a]) - xs[a * n + is];
                  audit_add(au, g, d, perm[is], perm[m]);
                }
                T e, coef;
                if (KIND == PFW_PAIR_SMOOTH) {
                  const T2 aa = sAA[zsp * nz + zj];
                  const T a = aa.x, a6 = aa.y;
                  const T qq = (T)1 - r2 * invc2;
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
                  e = c4e * (s6 * s6 - s6) - eshift;
                  coef = c24e * (s6 - (T)2 * s6 * s6) * inv;
                }
                const T gx = coef * dx, gy = coef * dy, gz = coef * dz;
                er += e;
                fxr += gx;
                fyr += gy;
                fzr += gz;
                w0 += dx * gx;
                w1 += dy * gy;
                w2 += dz * gz;
                w3 += dx * gy;
                w4 += dx * gz;
                w5 += dy * gz;
              }
              eacc += (double)er;
              fx += (double)fxr;
              fy += (double)fyr;
              fz += (double)fzr;
            }
          }
          __syncthreads();
        }
      }
      if (mask & PFW_FORCES) {
        red[lane][0] = fx;
        red[lane][1] = fy;
        red[lane][2] = fz;
        __syncthreads();
        if (on && split == 0) {
          double sx = 0, sy = 0, sz = 0;
          for (int s = 0; s < S; ++s) {
            sx += red[il + s * nb][0];
            sy += red[il + s * nb][1];
            sz += red[il + s * nb][2];
          }
          const int64_t o = perm[is];
          F[3 * o] = sx;
          F[3 * o + 1] = sy;
          F[3 * o + 2] = sz;
        }
        __syncthreads();
      }
    }
  }
  double v[7] = {eacc, (double)w0, (double)w1, (double)w2, (double)w3, (double)w4, (double)w5};
  for (int k = 0; k < 7; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
    for (int k = 0; k < 7; ++k) partials[7 * (int64_t)cell + k] = v[k];
}

void record_stage(pfw_ctx* ctx, int k) {
  if (ctx->timing) PFW_HIP(hipEventRecord(ctx->ev[k], ctx->stream));
}

static LJConsts lj_consts_of(double eps, double sig, double c) {
  LJConsts lj;
  const double rc = sig / c, qc2 = rc * rc, qc = qc2 * qc2 * qc2;
  lj.c4e = 4.0 * eps;
  lj.c24e = 24.0 * eps;
  lj.sig2 = sig * sig;
  lj.eshift = 4.0 * eps * (qc * qc - qc);
  return lj;
}

LJConsts lj_consts(const PairModelDev& m) {
  if (m.kind == PFW_PAIR_LJ) return lj_consts_of(m.params[0], m.params[1], m.cutoff);
  return LJConsts{0, 0, 0, 0};
}

template <typename T, int KIND>
static void launch_pair(pfw_ctx* ctx, uint32_t mask) {
  SystemBuffers& s = ctx->sys;
  const LJConsts lj = lj_consts(ctx->pair);
  const size_t dyn = KIND == PFW_PAIR_SMOOTH ? 2 * sizeof(T) * ctx->pair.nz * ctx->pair.nz : 0;
  const int slices = ctx->tune[3] > 0 ? std::min(ctx->tune[3], kSlices) : kSlices;
  const int32_t* cfgs = ctx->ncfg == 1 ? nullptr : s.cell_cfg;  // one configuration: grid 0 for every cell
  if (ctx->audit) {
    const AuditBufs au = audit_begin(ctx);
    hipLaunchKernelGGL((k_pair<T, KIND, true>), dim3(ctx->ncells), dim3(kWave), dyn, ctx->stream, s.grids,
                       cfgs, s.n, s.start, s.xs, s.xr, s.cbox, s.sps, s.perm, ctx->pair.d_params, ctx->pair.nz, lj,
                       mask, ctx->d_F, ctx->d_partials, ctx->ncells, au, slices);
  } else {
    hipLaunchKernelGGL((k_pair<T, KIND, false>), dim3(ctx->ncells), dim3(kWave), dyn, ctx->stream, s.grids,
                       cfgs, s.n, s.start, s.xs, s.xr, s.cbox, s.sps, s.perm, ctx->pair.d_params, ctx->pair.nz, lj,
                       mask, ctx->d_F, ctx->d_partials, ctx->ncells, AuditBufs{}, slices);
  }
}

// The pair kernel and the reductions of a fresh step (stream work only).
static void pair_launches(pfw_ctx* ctx, uint32_t mask, int precision) {
  const bool fp32 = precision == PFW_MIXED;
  if (ctx->pair.kind == PFW_PAIR_SMOOTH) {
    if (fp32)
      launch_pair<float, PFW_PAIR_SMOOTH>(ctx, mask);
    else
      launch_pair<double, PFW_PAIR_SMOOTH>(ctx, mask);
  } else {
    if (fp32)
      launch_pair<float, PFW_PAIR_LJ>(ctx, mask);
    else
      launch_pair<double, PFW_PAIR_LJ>(ctx, mask);
  }
  PFW_HIP(hipGetLastError());
  record_stage(ctx, 5);
  if (ctx->ncfg == 1) {
    reduce_outputs(ctx, ctx->d_partials, ctx->ncells, 7, 0, nullptr);
  } else {
    batch_reduce(ctx, ctx->d_partials, 7, true, 0);
  }
  PFW_HIP(hipGetLastError());
}

void pair_run(pfw_ctx* ctx, uint32_t mask, int precision) {
  ctx->host_written = false;
  ctx->audit_valid = false;
  if (ctx->pair.kind < 0) throw Error(PFW_STALE, "no pair model set");
  if (ctx->pair.cutoff != ctx->cutoff)
    throw Error(PFW_STALE, "pair model cutoff differs from the system's cell-list cutoff; set the system again");
  if (ctx->pair.kind == PFW_PAIR_SMOOTH && ctx->pair.nz > kMaxZ)
    throw Error(PFW_INVALID_ARGUMENT, "too many species for the pair table");
  if (ctx->verlet.skin > 0 && mask != 0 && ctx->sys.n > 0) {
    verlet_pair_run(ctx, mask, precision);  // ends with the outputs copied and the stream synchronised
    pair_finish(ctx, mask);
    return;
  }
  hipStream_t st = ctx->stream;
  if (ctx->graphs && ctx->graph_warm && !ctx->timing && !ctx->audit && mask != 0 && ctx->sys.n > 0) {
    struct Launch {
      uint32_t mask;
      int precision;
    } arg{mask, precision};
    graph_issue(ctx, 0, precision, mask, [](pfw_ctx* c, void* p) {
      const Launch& a = *(const Launch*)p;
      nl_build(c);
      pair_launches(c, a.mask, a.precision);
      copy_outputs(c);
    }, &arg);
    PFW_HIP(hipStreamSynchronize(st));
    pair_finish(ctx, mask);
    return;
  }
  nl_build(ctx);
  record_stage(ctx, 3);
  record_stage(ctx, 4);
  if (mask != 0 && ctx->sys.n > 0) {
    pair_launches(ctx, mask, precision);
  } else {
    if (ctx->audit) audit_begin(ctx);  // nothing accumulated: empty digests
    record_stage(ctx, 5);
    PFW_HIP(hipMemsetAsync(ctx->d_out, 0, 11 * sizeof(double), st));
    PFW_HIP(hipMemsetAsync(ctx->d_cfg_out, 0, 11 * (size_t)ctx->ncfg * sizeof(double), st));
    if (ctx->sys.n > 0) PFW_HIP(hipMemsetAsync(ctx->d_F, 0, 3 * ctx->sys.n * sizeof(double), st));
  }
  copy_outputs(ctx);
  record_stage(ctx, 6);
  PFW_HIP(hipStreamSynchronize(st));
  pair_finish(ctx, mask);
  ctx->graph_warm = true;
}

// After the step's outputs reached h_out: per-configuration rows, stage
// times, and the status check (DomainError on a singular or nonfinite result).
void pair_finish(pfw_ctx* ctx, uint32_t mask) {
  if (ctx->ncfg > 1) {
    ctx->h_cfg_out.resize(11 * (size_t)ctx->ncfg);
    PFW_HIP(hipMemcpy(ctx->h_cfg_out.data(), ctx->d_cfg_out, ctx->h_cfg_out.size() * sizeof(double),
                      hipMemcpyDeviceToHost));
  } else {
    ctx->h_cfg_out.assign(ctx->h_out, ctx->h_out + 11);
  }
  if (ctx->timing)
    for (int k = 0; k < kStages; ++k) {
      float ms = 0;
      PFW_HIP(hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]));
      ctx->stage_ms[k] = ms;
    }
  ctx->last_mask = mask;
  ctx->have_results = true;
  if (ctx->h_out[10] != 0.0)
    throw Error(PFW_DOMAIN, "singular origin or nonfinite pair callback output");
}

// ---------------------------------------------------------------- allocation-free pair step
//
// pfw_pair_evaluate_noalloc: one complete pair step (cell build, k_pair, FP64
// reductions) whose every intermediate buffer lies in `scratch`, a static
// device array of fixed size, so the call allocates nothing. Each device
// holds its own copy of the array, at an address taken once per device with
// hipGetSymbolAddress when a context is created there (static_scratch_init);
// host code uses that address, never the array's name. All calls on one
// device share the array: calls that overlap on a device race.
constexpr int64_t kStaticAtoms = 262144;  // atoms of one system
constexpr int64_t kStaticCells = 262144;  // cells of its grid
constexpr size_t kStaticAlign = 256;

constexpr size_t static_round(size_t b) { return (b + kStaticAlign - 1) / kStaticAlign * kStaticAlign; }

// Byte offsets of the sub-buffers, in order.
struct StaticLayout {
  size_t grid, params, flag, out, out2, red, atom_cfg, count, start, blocksum, bounds, cell_of, rank, perm,
      sps, xw, u, xs, xr, cbox, partials, total;
};
constexpr StaticLayout static_layout() {
  StaticLayout L{};
  size_t o = 0;
  auto take = [&o](size_t bytes) {
    const size_t at = o;
    o += static_round(bytes);
    return at;
  };
  L.grid = take(sizeof(Grid));
  L.params = take(sizeof(double) * kMaxZ * kMaxZ);
  L.flag = take(sizeof(int32_t) * 4);
  L.out = take(sizeof(double) * 16);
  L.out2 = take(sizeof(double) * 16);
  L.red = take(sizeof(double) * 64 * 16);
  L.atom_cfg = take(sizeof(int32_t) * kStaticAtoms);
  L.count = take(sizeof(int32_t) * kStaticCells);
  L.start = take(sizeof(int32_t) * (kStaticCells + 1));
  L.blocksum = take(sizeof(int32_t) * ((kStaticCells + 1023) / 1024 + 1));
  L.bounds = take(sizeof(uint64_t) * 6);
  L.cell_of = take(sizeof(int32_t) * kStaticAtoms);
  L.rank = take(sizeof(int32_t) * kStaticAtoms);
  L.perm = take(sizeof(int32_t) * kStaticAtoms);
  L.sps = take(sizeof(int32_t) * kStaticAtoms);
  L.xw = take(sizeof(double) * 4 * kStaticAtoms);
  L.u = take(sizeof(double) * 3 * kStaticAtoms);
  L.xs = take(sizeof(double) * 3 * kStaticAtoms);
  L.xr = take(sizeof(float) * 4 * kStaticAtoms);
  L.cbox = take(sizeof(float) * 8 * kStaticCells);
  L.partials = take(sizeof(double) * 7 * kStaticCells);
  L.total = o;
  return L;
}
constexpr StaticLayout kStatic = static_layout();  // about 56 MB (55,587,840 bytes)

static __device__ __align__(256) unsigned char scratch[kStatic.total];

constexpr int kMaxDevices = 64;
static unsigned char* g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mutex;

// Called by pfw_ctx_create with `device` current: the device address of
// that device's copy of `scratch`, once per device.
void static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(PFW_INVALID_ARGUMENT, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch[device]) return;
  void* p = nullptr;
  PFW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(PFW_DEVICE, "no device address for the static scratch area");
  g_scratch[device] = (unsigned char*)p;
}

// Nonfinite positions or species outside 0..nz-1 set flag[0].
__global__ void k_static_check(int64_t n, const double* __restrict__ X, const int32_t* __restrict__ sp, int nz,
                               int32_t* __restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  bool bad = !(isfinite(X[3 * k]) && isfinite(X[3 * k + 1]) && isfinite(X[3 * k + 2]));
  bad |= sp[k] < 0 || sp[k] >= nz;
  if (bad) atomicOr(flag, 1);
}

__global__ void k_static_outputs(const double* __restrict__ out, double* __restrict__ E, double* __restrict__ W,
                                 int32_t* __restrict__ flag) {
  const int k = threadIdx.x;
  if (k == 0 && E) E[0] = out[0];
  if (k < 9 && W) W[k] = out[1 + k];
  if (k == 0) flag[1] = out[10] != 0.0 ? 1 : 0;
}

__global__ void k_static_zero(double* __restrict__ out) {
  if (threadIdx.x < 11) out[threadIdx.x] = 0.0;
}

template <typename T, int KIND>
static void launch_pair_static(const Grid* grid, const int32_t* cell_cfg, int64_t n, const int32_t* start,
                               const double* xs, const float4* xr, const float4* cbox, const int32_t* sps,
                               const int32_t* perm,
                               const double* params, int nz, LJConsts lj, uint32_t mask, double* F, double* partials,
                               int ncells, hipStream_t st) {
  const size_t dyn = KIND == PFW_PAIR_SMOOTH ? 2 * sizeof(T) * nz * nz : 0;
  hipLaunchKernelGGL((k_pair<T, KIND, false>), dim3(ncells), dim3(kWave), dyn, st, grid, cell_cfg, n, start, xs, xr, cbox,
                     sps, perm, params, nz, lj, mask, F, partials, ncells, AuditBufs{}, kSlices);
}

void pair_run_static(pfw_ctx* ctx, int kind, int nz, const double* params, double cutoff, int64_t n,
                     const double* X, const int32_t* species, const double* H, const uint8_t* pbc, uint32_t mask,
                     int precision, double* E, double* F, double* W) {
  unsigned char* base = ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
  if (!base) throw Error(PFW_DEVICE, "static scratch area not initialised on this device");
  if (kind != PFW_PAIR_SMOOTH && kind != PFW_PAIR_LJ) throw Error(PFW_INVALID_ARGUMENT, "unknown pair model kind");
  if (precision != PFW_FP64 && precision != PFW_MIXED) throw Error(PFW_INVALID_ARGUMENT, "unknown precision mode");
  if (mask > 7) throw Error(PFW_INVALID_ARGUMENT, "unknown request bits");
  if (n < 0) throw Error(PFW_DIMENSION, "negative atom count");
  if (n > kStaticAtoms)
    throw Error(PFW_CAPACITY, "the allocation-free pair step holds at most 262,144 atoms");
  if (!H || !pbc || !params || (n > 0 && (!X || !species)))
    throw Error(PFW_INVALID_ARGUMENT, "null pointer argument");
  if ((mask & PFW_FORCES) && n > 0 && !F) throw Error(PFW_INVALID_ARGUMENT, "forces requested without an output");
  if (!(std::isfinite(cutoff) && cutoff > 0)) throw Error(PFW_INVALID_ARGUMENT, "cutoff must be finite and positive (Å)");
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(H[k])) throw Error(PFW_INVALID_ARGUMENT, "cell must be finite");
  LJConsts lj{0, 0, 0, 0};
  if (kind == PFW_PAIR_SMOOTH) {
    if (nz < 1 || nz > kMaxZ) throw Error(PFW_INVALID_ARGUMENT, "species count outside 1..16");
    for (int k = 0; k < nz * nz; ++k)
      if (!std::isfinite(params[k])) throw Error(PFW_INVALID_ARGUMENT, "amplitudes must be finite");
  } else {
    if (!(std::isfinite(params[0]) && std::isfinite(params[1]) && params[1] > 0))
      throw Error(PFW_INVALID_ARGUMENT, "Lennard-Jones epsilon and sigma must be finite, sigma positive");
    lj = lj_consts_of(params[0], params[1], cutoff);
    nz = kMaxZ;  // species are not read by the model; any channel below 16 is accepted
  }
  int pb[3];
  for (int a = 0; a < 3; ++a) {
    if (pbc[a] > 1) throw Error(PFW_INVALID_ARGUMENT, "pbc flags must be 0 or 1");
    pb[a] = pbc[a];
  }
  Grid g = nl_configure_one(H, pb, n, cutoff, cutoff, std::max<double>(2.0 * (double)n, 4096.0));
  g.cell0 = 0;
  g.ecell0 = 0;
  if (g.ncells > kStaticCells)
    throw Error(PFW_CAPACITY, "the cell grid exceeds the 262,144 cells of the allocation-free pair step");
  const int ncells = g.ncells;
  hipStream_t st = ctx->stream;
  const StaticLayout& L = kStatic;
  Grid* d_grid = (Grid*)(base + L.grid);
  double* d_params = (double*)(base + L.params);
  int32_t* d_flag = (int32_t*)(base + L.flag);
  double* d_out = (double*)(base + L.out);
  // host copies of the small inputs; the stream is synchronised before return
  PFW_HIP(hipMemcpyAsync(d_grid, &g, sizeof(Grid), hipMemcpyHostToDevice, st));
  if (kind == PFW_PAIR_SMOOTH)
    PFW_HIP(hipMemcpyAsync(d_params, params, sizeof(double) * nz * nz, hipMemcpyHostToDevice, st));
  fill_i32(d_flag, 4, 0, st);
  int32_t* h_flag = ctx->h_flag;  // page-locked, from the context's creation
  if (n > 0) {
    hipLaunchKernelGGL(k_static_check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, X, species, nz, d_flag);
    PFW_HIP(hipGetLastError());
    PFW_HIP(hipMemcpyAsync(h_flag, d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PFW_HIP(hipStreamSynchronize(st));
    if (h_flag[0]) throw Error(PFW_INVALID_ARGUMENT, "positions must be finite and species channels in range");
  }
  fill_i32((int32_t*)(base + L.atom_cfg), n, 0, st);
  BuildView v{n,
              ncells,
              1,
              g.open_any != 0,
              X,
              (double*)(base + L.xw),
              (double*)(base + L.u),
              (int32_t*)(base + L.cell_of),
              (int32_t*)(base + L.count),
              (int32_t*)(base + L.rank),
              (uint64_t*)(base + L.bounds),
              d_grid,
              (const int32_t*)(base + L.atom_cfg),
              (int32_t*)(base + L.start),
              (int32_t*)(base + L.blocksum),
              (int32_t*)(base + L.perm),
              (double*)(base + L.xs),
              species,
              (int32_t*)(base + L.sps),
              (float4*)(base + L.xr),
              (float4*)(base + L.cbox)};
  nl_build_core(v, st, false, nullptr);
  if (mask != 0 && n > 0) {
    const bool fp32 = precision == PFW_MIXED;
    double* part = (double*)(base + L.partials);
#define PFW_STATIC_LAUNCH(T, K)                                                                                  \
  launch_pair_static<T, K>(d_grid, (const int32_t*)nullptr, n, v.start, v.xs, v.xr, v.cbox, v.sps,              \
                           v.perm, d_params, nz, lj, mask, F, part, ncells, st)
    if (kind == PFW_PAIR_SMOOTH) {
      if (fp32)
        PFW_STATIC_LAUNCH(float, PFW_PAIR_SMOOTH);
      else
        PFW_STATIC_LAUNCH(double, PFW_PAIR_SMOOTH);
    } else {
      if (fp32)
        PFW_STATIC_LAUNCH(float, PFW_PAIR_LJ);
      else
        PFW_STATIC_LAUNCH(double, PFW_PAIR_LJ);
    }
#undef PFW_STATIC_LAUNCH
    PFW_HIP(hipGetLastError());
    reduce_core(part, ncells, 7, 0, nullptr, (double*)(base + L.red), d_out, (double*)(base + L.out2), nullptr, st);
  } else {
    hipLaunchKernelGGL(k_static_zero, dim3(1), dim3(64), 0, st, d_out);
  }
  hipLaunchKernelGGL(k_static_outputs, dim3(1), dim3(64), 0, st, d_out, E, W, d_flag);
  PFW_HIP(hipGetLastError());
  PFW_HIP(hipMemcpyAsync(h_flag, d_flag + 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipStreamSynchronize(st));
  if (h_flag[0]) throw Error(PFW_DOMAIN, "singular origin or nonfinite pair callback output");
}

// Parameter adjoint of SmoothPair: dL/dA (nz x nz, the entry each directed
// pair reads) for dL/dE per configuration, dL/dF (3 x N original order) and
// optional dL/dW. E, F and W are linear in A (parameters.jl:126-138), so each
// directed pair i -> j at d = x_j + S - x_i with q = 1 - r^2/c^2 adds
//   1/2 Ebar q^3 - (6 q^2 / c^2) (Fbar_i . d - 1/2 d^T Wbar d)
// to dL/dA[z_i, z_j] (the pair_vjp of ext/PotentialFrameWorkForwardDiffExt.jl:27-54).
// FP64, one thread per sorted atom enumerating the same exact pair set as the
// force kernel; per-thread sums per neighbour species, then a fixed-order
// block reduction (deterministic).
__global__ __launch_bounds__(256) void k_pair_adjoint(const Grid* __restrict__ grids,
                                                      const int32_t* __restrict__ atom_cfg, int64_t n,
                                                      const double* __restrict__ xs,
                                                      const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ cell_of,
                                                      const int32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ sps, int nz,
                                                      const double* __restrict__ Ebar,
                                                      const double* __restrict__ Fbar,
                                                      const double* __restrict__ Wbar,
                                                      double* __restrict__ partial) {
  __shared__ double tab[256][kMaxZ + 1];
  __shared__ int zs[256];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[kMaxZ];
#pragma unroll
  for (int z = 0; z < kMaxZ; ++z) acc[z] = 0;
  int zi = -1;
  if (k < n) {
    const int oi = perm[k];
    const int cfg = atom_cfg[oi];
    const Grid& g = grids[cfg];
    zi = sps[k];
    int key[3];
    cell_key(g, cell_of[oi], key);
    const double xi = xs[k], yi = xs[n + k], zz = xs[2 * n + k];
    const double eb = 0.5 * Ebar[cfg], f0 = Fbar[3 * oi], f1 = Fbar[3 * oi + 1], f2 = Fbar[3 * oi + 2];
    double wb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) wb[q] = Wbar ? 0.5 * Wbar[9 * cfg + q] : 0.0;
    const double invc2 = 1.0 / g.c2, s6 = -6.0 * invc2;
    for (int e = 0; e < g.nstencil; ++e) {
      int lat[3];
      bool zero;
      const int nc = stencil_entry(g, key, e, lat, &zero);
      if (nc < 0) continue;
      double sh[3];
      exact_shift(g, lat, sh);
      for (int m = start[nc]; m < start[nc + 1]; ++m) {
        if (zero && m == k) continue;
        double d[3];
        int img[3];
        if (!exact_member(g, xi, yi, zz, xs[m], xs[n + m], xs[2 * n + m], lat, sh, d, img)) continue;
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double qq = 1.0 - r2 * invc2, q2 = qq * qq;
        const double fd = f0 * d[0] + f1 * d[1] + f2 * d[2];
        const double dwd = d[0] * (wb[0] * d[0] + wb[3] * d[1] + wb[6] * d[2]) +
                           d[1] * (wb[1] * d[0] + wb[4] * d[1] + wb[7] * d[2]) +
                           d[2] * (wb[2] * d[0] + wb[5] * d[1] + wb[8] * d[2]);
        const double t = eb * q2 * qq + s6 * q2 * (fd - dwd);
        const int zj = sps[m];
#pragma unroll
        for (int z = 0; z < kMaxZ; ++z) acc[z] += z == zj ? t : 0.0;
      }
    }
  }
  for (int z = 0; z < kMaxZ; ++z) tab[threadIdx.x][z] = acc[z];
  zs[threadIdx.x] = zi;
  __syncthreads();
  const int nn = nz * nz;
  for (int idx = threadIdx.x; idx < nn; idx += blockDim.x) {
    const int a = idx / nz, b = idx % nz;
    double v = 0;
    for (int t = 0; t < (int)blockDim.x; ++t)
      if (zs[t] == a) v += tab[t][b];
    partial[(int64_t)blockIdx.x * nn + idx] = v;
  }
}

void pair_adjoint(pfw_ctx* ctx, const double* Ebar, const double* Fbar, const double* Wbar, double* gA) {
  if (ctx->pair.kind < 0) throw Error(PFW_STALE, "no pair model set");
  if (ctx->pair.kind != PFW_PAIR_SMOOTH)
    throw Error(PFW_INVALID_ARGUMENT, "parameter adjoints are implemented for SmoothPair amplitudes");
  if (ctx->pair.cutoff != ctx->cutoff)
    throw Error(PFW_STALE, "pair model cutoff differs from the system's cell-list cutoff; set the system again");
  const int nz = ctx->pair.nz;
  if (nz > kMaxZ) throw Error(PFW_INVALID_ARGUMENT, "too many species for the pair table");
  SystemBuffers& s = ctx->sys;
  const int64_t n = s.n;
  const int nn = nz * nz;
  for (int k = 0; k < nn; ++k) gA[k] = 0;
  if (n == 0) return;
  for (int c = 0; c < ctx->ncfg; ++c)
    if (!std::isfinite(Ebar[c])) throw Error(PFW_INVALID_ARGUMENT, "nonfinite energy adjoint");
  for (int64_t k = 0; k < 3 * n; ++k)
    if (!std::isfinite(Fbar[k])) throw Error(PFW_INVALID_ARGUMENT, "nonfinite force adjoint");
  if (Wbar)
    for (int64_t k = 0; k < 9 * (int64_t)ctx->ncfg; ++k)
      if (!std::isfinite(Wbar[k])) throw Error(PFW_INVALID_ARGUMENT, "nonfinite virial adjoint");
  nl_build(ctx);
  hipStream_t st = ctx->stream;
  const int nb = (int)((n + 255) / 256);
  DeviceArray<double> dE(ctx->ncfg), dF(3 * n), dW(Wbar ? 9 * (size_t)ctx->ncfg : 0), part((size_t)nb * nn),
      sum(nn);
  PFW_HIP(hipMemcpyAsync(dE.p, Ebar, ctx->ncfg * sizeof(double), hipMemcpyHostToDevice, st));
  PFW_HIP(hipMemcpyAsync(dF.p, Fbar, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
  if (Wbar) PFW_HIP(hipMemcpyAsync(dW.p, Wbar, 9 * (size_t)ctx->ncfg * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_pair_adjoint, dim3(nb), dim3(256), 0, st, s.grids, s.atom_cfg, n, s.xs, s.start, s.cell_of,
                     s.perm, s.sps, nz, dE.p, dF.p, dW.p, part.p);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, part.p, nb, nn, sum.p);
  PFW_HIP(hipGetLastError());
  PFW_HIP(hipMemcpyAsync(gA, sum.p, nn * sizeof(double), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipStreamSynchronize(st));
  bool bad = false;
  for (int k = 0; k < nn; ++k) bad |= !std::isfinite(gA[k]);
  if (bad) throw Error(PFW_DOMAIN, "nonfinite pair parameter gradient");
}

}  // namespace pfw
