// Device-side helpers shared by the cell-list, pair and ACE kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pfw_internal.h"

namespace pfw {

constexpr int kWave = 64;
constexpr int kMaxZ = 16;  // species table size limit of the pair kernels

// Energy-shifted Lennard-Jones constants (lennard_jones.jl:5-17): 4 eps, 24 eps,
// sigma^2 and the energy at the cutoff.
struct LJConsts {
  double c4e, c24e, sig2, eshift;
};
LJConsts lj_consts(const PairModelDev& m);

// Block index -> work item, giving each of the 8 XCDs (dispatched round-robin)
// a contiguous range so that neighbouring items share an L2.
__device__ inline int xcd_remap(int b, int nb) {
  int q = nb / 8, r = nb % 8;
  int x = b % 8, idx = b / 8;
  return x < r ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// Julia `mod(x::Float64, y::Float64)`: rem-based, result carries the sign of y.
__device__ __host__ inline double jl_mod(double x, double y) {
  double r = fmod(x, y);
  if (r == 0) return copysign(0.0, y);
  if ((r > 0) != (y > 0)) return r + y;
  return r;
}
__device__ __host__ inline int jl_fld(int a, int b) {
  int q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
  return q;
}
__device__ __host__ inline int jl_imod(int a, int b) { return a - jl_fld(a, b) * b; }

// Exact FP64 geometry: every product and sum is separately rounded (no FMA
// contraction), in the association order of the reference's StaticArrays
// matrix-vector products, so membership decisions are reproducible bit for bit.
__device__ inline void exact_matvec(const double* M, const double* v, double* o) {
#pragma clang fp contract(off)
  for (int a = 0; a < 3; ++a) o[a] = M[a] * v[0] + M[a + 3] * v[1] + M[a + 6] * v[2];
}
__device__ inline void exact_shift(const Grid& g, const int* s, double* o) {
#pragma clang fp contract(off)
  if (g.minimum_image) {  // orthorhombic: image shift s_a L_a (the basis is the identity here)
    for (int a = 0; a < 3; ++a) o[a] = g.pb[a] ? (double)s[a] * g.L[a] : 0.0;
    return;
  }
  double v[3] = {(double)s[0], (double)s[1], (double)s[2]};
  for (int a = 0; a < 3; ++a) o[a] = g.basis[a] * v[0] + g.basis[a + 3] * v[1] + g.basis[a + 6] * v[2];
}
// Exact membership of the directed pair (sorted i -> sorted j) through the
// stencil entry with integer lattice `lat` and Cartesian shift `sh`
// (geometry.jl:5-19): d = x_j - x_i + H s on the image path; on the
// minimum-image path d = raw - round(raw/L) L, and the entry must carry that
// same lattice so that two stencil entries aliasing one cell (dims == 2)
// cannot both claim the pair. Writes d and the lattice of the image in `img`.
__device__ inline bool exact_member(const Grid& g, double xi, double yi, double zi, double xj, double yj,
                                    double zj, const int* lat, const double* sh, double* d, int* img) {
#pragma clang fp contract(off)
  double raw[3] = {xj - xi, yj - yi, zj - zi};
  if (g.minimum_image) {
    for (int a = 0; a < 3; ++a) {
      if (g.pb[a]) {
        // m = rint(raw / L). raw (1/L) differs from raw / L by a few ulp, so it
        // decides the rounding unless it lies within 1e-9 of a half-integer;
        // there the division decides
        double m = rint(raw[a] * g.invL[a]);
        if (fabs(fabs(raw[a] * g.invL[a] - m) - 0.5) < 1e-9) m = rint(raw[a] / g.L[a]);
        d[a] = raw[a] - m * g.L[a];
        img[a] = -(int)m;
        if (img[a] != lat[a]) return false;
      } else {
        d[a] = raw[a];
        img[a] = 0;
      }
    }
  } else {
    for (int a = 0; a < 3; ++a) {
      d[a] = raw[a] + sh[a];
      img[a] = lat[a];
    }
  }
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < g.c2;
}

// d of a directed pair that exact_member has accepted through the stencil
// entry with lattice `lat` and shift `sh`, by the same operations (on the
// minimum-image path an accepted pair has m = -lat).
__device__ inline void member_disp(const Grid& g, double xi, double yi, double zi, double xj, double yj, double zj,
                                   const int* lat, const double* sh, double* d) {
#pragma clang fp contract(off)
  const double raw[3] = {xj - xi, yj - yi, zj - zi};
  for (int a = 0; a < 3; ++a) {
    if (!g.minimum_image) {
      d[a] = raw[a] + sh[a];
    } else if (g.pb[a]) {
      const double m = -(double)lat[a];
      d[a] = raw[a] - m * g.L[a];
    } else {
      d[a] = raw[a];
    }
  }
}

// Global cell index <-> key of the configuration's grid.
__device__ inline void cell_key(const Grid& g, int cell, int* k) {
  const int local = cell - g.cell0;
  k[0] = local % g.dims[0];
  const int t = local / g.dims[0];
  k[1] = t % g.dims[1];
  k[2] = t / g.dims[1];
}
__device__ inline int cell_index(const Grid& g, const int* k) {
  return g.cell0 + k[0] + g.dims[0] * (k[1] + g.dims[1] * k[2]);
}

// Stencil entry e of the cell with key k: global neighbour cell (or -1 when it lies
// outside an open axis), integer lattice shift and whether it is the zero
// offset. Periodic axes wrap with shift fld(k+o, dims) (storage.jl:265-283).
__device__ inline int stencil_entry(const Grid& g, const int* k, int e, int* lat, bool* zero) {
  int o[3];
  if (g.radius[0] == 1 && g.radius[1] == 1 && g.radius[2] == 1) {  // 3 x 3 x 3: constant divisors
    o[0] = e % 3 - 1;
    o[1] = (e / 3) % 3 - 1;
    o[2] = e / 9 - 1;
  } else {
    const int w0 = 2 * g.radius[0] + 1, w1 = 2 * g.radius[1] + 1;
    o[0] = e % w0 - g.radius[0];
    o[1] = (e / w0) % w1 - g.radius[1];
    o[2] = e / (w0 * w1) - g.radius[2];
  }
  int nk[3];
  *zero = (o[0] == 0 && o[1] == 0 && o[2] == 0);
  for (int a = 0; a < 3; ++a) {
    int v = k[a] + o[a];
    if (g.pb[a]) {
      const int d = g.dims[a];
      if (v >= 0 && v < d) {  // fld(v, d) = 0 without a division
        nk[a] = v;
        lat[a] = 0;
      } else if (v < 0 && v >= -d) {
        nk[a] = v + d;
        lat[a] = -1;
      } else if (v >= d && v < 2 * d) {
        nk[a] = v - d;
        lat[a] = 1;
      } else {
        nk[a] = jl_imod(v, d);
        lat[a] = jl_fld(v, d);
      }
    } else {
      if (v < 0 || v >= g.used[a]) return -1;
      nk[a] = v;
      lat[a] = 0;
    }
  }
  return cell_index(g, nk);
}

// Julia mod(x, y) for y > 0 with a fast path for x already in [0, y): mod
// returns x itself there, except -0.0, which becomes +0.0.
__device__ inline double mod_pos(double x, double y) {
  if (x >= 0 && x < y) return x == 0 ? 0.0 : x;
  return jl_mod(x, y);
}

__device__ inline void wrap_atom(const Grid& g, const double* x, double* w, double* u) {
  if (g.minimum_image) {
    for (int a = 0; a < 3; ++a) {
      w[a] = g.pb[a] ? mod_pos(x[a], g.L[a]) : x[a];
      u[a] = w[a];
    }
  } else {
    double f[3];
    exact_matvec(g.recip, x, f);
    for (int a = 0; a < 3; ++a)
      if (g.pb[a]) f[a] = mod_pos(f[a], 1.0);
    exact_matvec(g.basis, f, w);
    for (int a = 0; a < 3; ++a) u[a] = f[a];
  }
}

// Winding of caller position x (original coordinates): the lattice vector
// that wrap_atom takes off, fld of the fractional coordinate on periodic axes.
__device__ inline void winding_of(const Grid& g, const double* x, int* w) {
  if (g.minimum_image) {
    for (int a = 0; a < 3; ++a) w[a] = g.pb[a] ? (int)floor(x[a] / g.L[a]) : 0;
  } else {
    double f[3];
    exact_matvec(g.recip, x, f);
    for (int a = 0; a < 3; ++a) w[a] = g.pb[a] ? (int)floor(f[a]) : 0;
  }
}

// Exact membership of the image (oj, S) of centre oi (original indices, S
// relative to the caller's coordinates X) decided as a fresh cell build
// decides it: both positions wrapped now, the image lattice in the wrapped
// frame l = S - w_i + w_j, then exact_member. Used to re-decide pairs found
// by other arithmetic (list reuse) inside a thin band around the cutoff.
__device__ inline bool exact_image_member(const Grid& g, const double* __restrict__ X, int oi, int oj, const int* S) {
  double xi[3], xj[3], wi_[3], wj_[3], ui[3], uj[3];
  int wi[3], wj[3], lat[3];
  for (int a = 0; a < 3; ++a) {
    xi[a] = X[3 * (int64_t)oi + a];
    xj[a] = X[3 * (int64_t)oj + a];
  }
  wrap_atom(g, xi, wi_, ui);
  wrap_atom(g, xj, wj_, uj);
  winding_of(g, xi, wi);
  winding_of(g, xj, wj);
  for (int a = 0; a < 3; ++a) lat[a] = S[a] - wi[a] + wj[a];
  double sh[3], d[3];
  int img[3];
  exact_shift(g, lat, sh);
  return exact_member(g, wi_[0], wi_[1], wi_[2], wj_[0], wj_[1], wj_[2], lat, sh, d, img);
}

// ---- pair-set audit: an order-independent digest of the directed pairs an
// evaluating kernel actually accumulated, per centre atom (original order).
// A pair is identified by the neighbour's original index j and its integer
// image shift S relative to the caller's original coordinates, R = x_j - x_i +
// H S (the convention of the directed snapshot, interface.jl:116-143). The
// digest is the count and the wrapping sum of a 64-bit mix of (j, S).
__device__ __host__ inline unsigned long long audit_mix(int j, const int* S) {
  unsigned long long z = (unsigned long long)(uint32_t)j;
  z |= (unsigned long long)((S[0] + 1024) & 2047) << 32;
  z |= (unsigned long long)((S[1] + 1024) & 2047) << 43;
  z |= (unsigned long long)((S[2] + 1024) & 2047) << 54;
  // splitmix64 finaliser (Steele, Lea and Flood, OOPSLA 2014)
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
// The image shift of a pair from the FP64 displacement d the kernel used and
// the caller's positions X (3 x N, original order): H S = d - (x_j - x_i) up
// to rounding, so S is that difference in lattice coordinates, rounded.
__device__ inline void audit_shift(const Grid& g, const double* d, const double* __restrict__ X, int oi, int oj,
                                   int* S) {
  double t[3];
  for (int a = 0; a < 3; ++a) t[a] = d[a] - (X[3 * (int64_t)oj + a] - X[3 * (int64_t)oi + a]);
  if (g.minimum_image) {
    for (int a = 0; a < 3; ++a) S[a] = g.pb[a] ? (int)rint(t[a] / g.L[a]) : 0;
  } else {
    for (int a = 0; a < 3; ++a) {
      const double f = g.recip[a] * t[0] + g.recip[a + 3] * t[1] + g.recip[a + 6] * t[2];
      S[a] = g.pb[a] ? (int)rint(f) : 0;
    }
  }
}
struct AuditBufs {
  int32_t* count;             // n, original order
  unsigned long long* hash;   // n, original order
  const double* X;            // the caller's positions, 3 x N original order
};
__device__ inline void audit_add(const AuditBufs& au, const Grid& g, const double* d, int oi, int oj) {
  int S[3];
  audit_shift(g, d, au.X, oi, oj, S);
  atomicAdd(au.count + oi, 1);
  atomicAdd(au.hash + oi, audit_mix(oj, S));
}

// Cleared digest buffers for the next evaluating kernel (allocated on first use).
AuditBufs audit_begin(pfw_ctx* ctx);

__device__ inline bool is_finite(double x) { return isfinite(x); }

// Order-preserving map of doubles onto unsigned 64-bit integers (for atomic min/max).
__device__ inline unsigned long long ordered_bits(double x) {
  unsigned long long b = __double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double from_ordered_bits(unsigned long long b) {
  return __longlong_as_double((b >> 63) ? (b & 0x7fffffffffffffffull) : ~b);
}

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Inclusive prefix sum over the 64 lanes of a wavefront with DPP moves (no LDS
// round trips): Hillis-Steele within each row of 16 lanes (row_shr 1, 2, 4, 8;
// lanes whose source lies before the row start add 0), then row 0's total into
// rows 1 and 3 (row_bcast:15) and rows 0 + 1's total into rows 2 and 3
// (row_bcast:31). Every lane of the wavefront must be active.
__device__ inline int wave_incl_scan(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
  return x;
}

// Stage 1 of a deterministic two-stage FP64 column reduction: block b sums
// columns of records [b*chunk, (b+1)*chunk) of an (m x width) row-major table
// into out[b*width + c] (width <= 16, blockDim 256).
constexpr int kRedBlocks = 64;
__global__ void k_colsum(const double* __restrict__ in, int64_t m, int width, double* __restrict__ out);
__global__ void k_sum_partials(const double* __restrict__ partial, int nblocks, int width, double* __restrict__ out);
__global__ void k_finish(const double* __restrict__ part, int nb, int width, int kind,
                         const int32_t* __restrict__ status, double* __restrict__ out, double* __restrict__ out2,
                         double* __restrict__ out3);

}  // namespace pfw
