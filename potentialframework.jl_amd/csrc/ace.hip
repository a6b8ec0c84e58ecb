// ACE site pipeline on gfx950: directed environments from the cell list,
// Agnesi-spline radials, real L2 spherical harmonics with Cartesian gradients,
// atomic-base densities, product basis, precontracted linear readout and the
// force pullback. Neighbour forces are gathered over reverse edges, so no
// global atomics are used anywhere.
//
// Reference: site_data / site_reverse (src/submodules/ACE/serial.jl:51-101),
// NI.evaluate_site! (src/submodules/ACE/cpu.jl:162-237), the radial spline
// and envelopes (src/submodules/ACE/resolved.jl:156-209), the Y_lm recurrence
// (src/submodules/ACE/angular_blocks.jl:25-81), unrestricted_products!
// (src/submodules/ACE/projection.jl:39-56), reduce_site!
// (src/submodules/NeighbourList/execution.jl:105-122); the KA kernels being
// replaced are src/submodules/ACE/gpu.jl:99-270.
//
// Layout: one 64-lane wavefront per site (sorted order). Phase A: lanes are
// edges; each lane evaluates its radial row and Y_lm into LDS (row stride 65
// to spread banks), then lanes are densities and sum R Y over the chunk.
// Phase B: lanes are products for the energy and densities for dE/dA (a
// gather over a per-density CSR, so no LDS atomics). Phase C: lanes are edges
// again; Y and grad Y live in registers (compile-time lmax), the radial rows
// are recomputed, and g = r^ sum_r R'_r T_r + sum_m U_m grad Y_m.
#include <algorithm>
#include <cstring>
#include <array>
#include <map>
#include <cmath>
#include <vector>

#include "pfw_device.h"

namespace pfw {

constexpr int kStride = kWave + 1;
constexpr int kRadialChunk = 20;  // radial indices per density-kernel tile
// the built-in density contraction: the matrix cores where the model's tiles
// fit (FP32, unprojected) and at least half of the tile entries are
// densities. 32k-atom BCC, lmax 4 (69 % useful): 127.9 -> 124.0 us per
// launch; the May30 model (40 % useful) ran slower on the matrix cores
// (interaction 1.17 -> 1.22 ms per step)
constexpr bool kMfmaPullDefault = false;  // the built-in pullback contraction: VALU

template <typename T>
struct FamArgs {
  int nr, nnodes, nint, uniform;
  const int* pq;
  const int* env_kind;
  const int* env_pow;
  const T* par;    // nch x 12
  const T* nodes;  // nch x nnodes
  const T* coef;   // nch x nint x nr x 4
};

template <typename T>
struct AceArgs {
  int64_t n;
  int nz, ndens, nprod, ny;
  const int32_t* perm;
  const int32_t* sps;
  const int32_t* ecount;  // edges of each site (sorted centre); site i's edges start at i * ecap
  int ecap;
  const int32_t* enb;
  const double* eR;
  const double* reference;
  FamArgs<T> rad, pair;
  const int32_t* dens;
  const int32_t* dperm;   // densities ordered by radial index
  const int32_t* dchunk;  // offsets into dperm of each chunk of rchunk radial indices
  int rchunk;
  const int32_t* rlb;     // L + 2: radial indices grouped by l
  const int32_t* rl;      // (radial index, first density of its m-block)
  const int32_t* prod;
  const T* w;   // nz x nprod
  const T* wp;  // nz x nq
  int nsteps;          // product schedule: nsteps x 64 slots
  const uint4* sched;  // per slot: 4 x u16 factor indices (0xffff = none), 4 x u16 targets
  const T* ws;         // nz x slots: readout weight of each slot's product
  const int32_t* copyb;  // ndens + 1: accumulator copies of each density
  int ntarget;
  T* Abuf;              // n x ndens densities, then dE/dA in place
  int grads;
  double* siteE;
  double* siteW;
  T* eg;  // 3 x E edge gradients in the arithmetic precision
  int solid;  // angular basis: 1 real solid harmonics r^l Y_lm, 0 spherical Y_lm
  // basis policy (interface.jl:115-125): with proj, each site has densities U
  // (neighbours outside the projected species) and V (inside), stored as
  // [U | V] (dstride = 2 ndens), and a product keeps only the orders
  // olo..ohi of prod (U + t V); the pair term of a neighbour of species z is
  // kept when bit z of pair_gate is set
  int proj, dstride, olo, ohi;
  uint32_t sel, pair_gate;
  int ostep[5];  // product schedule: steps [ostep[nu - 1], ostep[nu]) have nu factors
};

template <typename T>
__device__ inline T ipow(T x, int n) {
  T v = (T)1;
  for (int k = 0; k < n; ++k) v *= x;
  return v;
}

__host__ __device__ inline size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

// Agnesi coordinate x(r), dx/dr, envelope and its r-derivative, spline
// interval and local coordinate (resolved.jl:156-209).
template <typename T>
struct RadialPoint {
  int k;
  T u, e, de;
  T ed;  // e dx/dr / (interval length): the envelope times du/dr
};

template <typename T>
__device__ inline RadialPoint<T> radial_point(const FamArgs<T>& f, int ch, T r) {
  const T* par = f.par + 12 * ch;
  const int p = f.pq[2 * ch], q = f.pq[2 * ch + 1];
  const T a = par[0], rin = par[1], r0 = par[2], yin = par[3], ycut = par[4];
  const T s = (r - rin) / (r0 - rin);
  const int h = q - p;
  const T sq = ipow(s, q), sh = ipow(s, h);
  const T den = (T)1 + sh;
  const T y = r <= rin ? (T)1 : (T)1 / ((T)1 + a * sq / den);
  T x = (T)-1 + (T)2 * (y - yin) / (ycut - yin);
  x = x < (T)-1 ? (T)-1 : (x > (T)1 ? (T)1 : x);
  T dx = 0;
  if (!(r <= rin || x <= (T)-1 || x >= (T)1)) {
    const T u = sq / den;
    const T du = ((T)q * ipow(s, q - 1) / den - (h == 0 ? (T)0 : (T)h * ipow(s, q + h - 1) / (den * den))) / (r0 - rin);
    const T t = (T)1 + a * u;
    dx = (T)-2 * a * du / (t * t * (ycut - yin));
  }
  RadialPoint<T> o;
  const int kind = f.env_kind[ch];
  const int e0 = f.env_pow[2 * ch], e1 = f.env_pow[2 * ch + 1];
  if (kind == PFW_ENV_POLY2SX) {
    const T x1 = par[6], x2 = par[7], sc = par[8];
    if (x1 < x && x < x2) {
      o.e = sc * ipow(x - x1, e0) * ipow(x2 - x, e1);
      o.de = sc * ((T)e0 * ipow(x - x1, e0 - 1) * ipow(x2 - x, e1) - (T)e1 * ipow(x - x1, e0) * ipow(x2 - x, e1 - 1)) * dx;
    } else {
      o.e = 0;
      o.de = 0;
    }
  } else if (kind == PFW_ENV_POLY1SR) {
    const T rc = par[6];
    if (r >= rc) {
      o.e = o.de = 0;
    } else {
      const T ss = r / rc;
      const T sp = ipow(ss, e0);
      o.e = ((T)1 / sp - (T)1) * ((T)1 - ss);
      o.de = (-(T)e0 / (sp * ss) * ((T)1 - ss) - ((T)1 / sp - (T)1)) / rc;
    }
  } else {
    const T rc = par[6], rr0 = par[7];
    if (r > rc) {
      o.e = o.de = 0;
    } else {
      const T ss = r / rr0, sc = rc / rr0;
      o.e = (T)1 / ipow(ss, e0) - (T)1 / ipow(sc, e0) + (T)e0 / ipow(sc, e0 + 1) * (ss - sc);
      o.de = (T)e0 / rr0 * ((T)1 / ipow(sc, e0 + 1) - (T)1 / ipow(ss, e0 + 1));
    }
  }
  const T* nodes = f.nodes + (int64_t)f.nnodes * ch;
  int k;
  if (f.uniform) {
    k = (int)((x + (T)1) * (T)0.5 * (T)(f.nnodes - 1));
    k = k < 0 ? 0 : (k > f.nint - 1 ? f.nint - 1 : k);
    if (k > 0 && x < nodes[k]) --k;
    if (k < f.nint - 1 && x >= nodes[k + 1]) ++k;
  } else {
    int lo = 0, hi = f.nnodes;  // largest index with nodes[idx] <= x
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (nodes[mid] <= x)
        lo = mid;
      else
        hi = mid;
    }
    k = lo > f.nint - 1 ? f.nint - 1 : lo;
  }
  const T step = nodes[k + 1] - nodes[k];
  const T invstep = (T)1 / step;
  o.k = k;
  o.u = (x - nodes[k]) * invstep;
  o.ed = o.e * dx * invstep;
  return o;
}

template <typename T>
__device__ inline void radial_value(const FamArgs<T>& f, int ch, const RadialPoint<T>& P, int n, T& v, T& d) {
  const T* c = f.coef + (((int64_t)ch * f.nint + P.k) * f.nr + n) * 4;
  const T c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
  // Horner for the cubic and its u-derivative, sharing p1 = c3 u + c2:
  // 3 c3 u^2 + 2 c2 u + c1 = (c3 u + 2 p1) u + c1
  const T p1 = c3 * P.u + c2;
  const T poly = (p1 * P.u + c1) * P.u + c0;
  const T dpoly = (c3 * P.u + (p1 + p1)) * P.u + c1;
  v = P.e * poly;
  d = P.de * poly + P.ed * dpoly;
}

template <typename T>
struct Jet {
  T v, g0, g1, g2;
};
template <typename T>
__device__ inline Jet<T> jm(const Jet<T>& a, const Jet<T>& b) {
  return {a.v * b.v, a.g0 * b.v + a.v * b.g0, a.g1 * b.v + a.v * b.g1, a.g2 * b.v + a.v * b.g2};
}
template <typename T>
__device__ inline Jet<T> js(T s, const Jet<T>& a) {
  return {s * a.v, s * a.g0, s * a.g1, s * a.g2};
}
template <typename T>
__device__ inline Jet<T> jadd(const Jet<T>& a, const Jet<T>& b) {
  return {a.v + b.v, a.g0 + b.g0, a.g1 + b.g1, a.g2 + b.g2};
}
template <typename T>
__device__ inline Jet<T> jsub(const Jet<T>& a, const Jet<T>& b) {
  return {a.v - b.v, a.g0 - b.g0, a.g1 - b.g1, a.g2 - b.g2};
}
// s a - t b
template <typename T>
__device__ inline Jet<T> jlin(T s, const Jet<T>& a, T t, const Jet<T>& b) {
  return {s * a.v - t * b.v, s * a.g0 - t * b.g0, s * a.g1 - t * b.g1, s * a.g2 - t * b.g2};
}

// Real L2 Y_lm (no Condon-Shortley phase, index l*l+l+m) of the direction of
// (dx, dy, dz) and their Cartesian gradients (angular_blocks.jl:36-81); with
// `solid` the same recurrence on the raw coordinates, the real solid harmonics
// r^l Y_lm (angular_blocks.jl:47-48). Each harmonic is handed to
// emit(index, jet) as soon as it is formed, so a caller that contracts them
// need not hold all values and gradients at once.
// Fl holds F[l][m] = sqrt((2l+1)/2pi) * prod(-1/sqrt((l+m')(l+1-m'))).
template <typename T, int L, typename Emit>
__device__ inline void ylm_emit(T dx, T dy, T dz, const T* Fl, int solid, Emit emit) {
  const T r = sqrt(dx * dx + dy * dy + dz * dz);
  const T ri = (T)1 / r;
  const T ux = dx * ri, uy = dy * ri, uz = dz * ri;
  Jet<T> x{ux, ((T)1 - ux * ux) * ri, -ux * uy * ri, -ux * uz * ri};
  Jet<T> y{uy, -uy * ux * ri, ((T)1 - uy * uy) * ri, -uy * uz * ri};
  Jet<T> z{uz, -uz * ux * ri, -uz * uy * ri, ((T)1 - uz * uz) * ri};
  if (solid) {
    x = {dx, (T)1, (T)0, (T)0};
    y = {dy, (T)0, (T)1, (T)0};
    z = {dz, (T)0, (T)0, (T)1};
  }
  const Jet<T> r2 = jadd(jadd(jm(x, x), jm(y, y)), jm(z, z));
  // m outer, l inner: the recurrences Q[l][l] = -(2l-1) Q[l-1][l-1],
  // Q[l][l-1] = (2l-1) z Q[l-1][l-1] and, for m <= l-2,
  // Q[l][m] = ((2l-1) z Q[l-1][m] - (l+m-1) r^2 Q[l-2][m]) / (l-m), with
  // cos/sin(m phi) advanced with m; every value is formed by the same
  // operations as in l-outer order, but only a few jets are live at a time
  Jet<T> c{(T)1, 0, 0, 0}, sgn{(T)0, 0, 0, 0}, diag{(T)1, 0, 0, 0};
#pragma unroll
  for (int m = 0; m <= L; ++m) {
    if (m > 0) {
      const Jet<T> c1 = jsub(jm(c, x), jm(sgn, y));
      sgn = jadd(jm(sgn, x), jm(c, y));
      c = c1;
      diag = js((T) - (2 * m - 1), diag);
    }
    const Jet<T> cm = m == 0 ? Jet<T>{(T)0.70710678118654752440, 0, 0, 0} : c;
    Jet<T> q2 = diag, q1 = diag;  // Q[l-2][m], Q[l-1][m]
#pragma unroll
    for (int l = m; l <= L; ++l) {
      Jet<T> q;
      if (l == m)
        q = diag;
      else if (l == m + 1)
        q = js((T)(2 * l - 1), jm(z, q1));
      else  // the constants (2l-1)/(l-m) and (l+m-1)/(l-m) folded: one multiply and one FMA per component
        q = jlin((T)(2 * l - 1) / (T)(l - m), jm(z, q1), (T)(l + m - 1) / (T)(l - m), jm(r2, q2));
      const T F = Fl[l * (L + 1) + m];
      emit(l * l + l + m, js(F, jm(q, cm)));
      if (m > 0) emit(l * l + l - m, js(F, jm(q, sgn)));
      q2 = q1;
      q1 = q;
    }
  }
}

// All (L+1)^2 values Y and gradients G (3 per harmonic).
template <typename T, int L>
__device__ inline void ylm(T dx, T dy, T dz, const T* Fl, T* Y, T* G, int solid) {
  ylm_emit<T, L>(dx, dy, dz, Fl, solid, [&](int i, const Jet<T>& j) {
    Y[i] = j.v;
    G[3 * i] = j.g0;
    G[3 * i + 1] = j.g1;
    G[3 * i + 2] = j.g2;
  });
}

// The site pipeline runs as three kernels:
//   k_ace_dens   per-edge radials and Y_lm, densities A (site x ndens) and the
//                pair-radial energy of each site;
//   k_ace_prod   products, readout and dE/dA with the model tables in LDS,
//                dE/dA written over A in place;
//   k_ace_pull   per-edge force pullback and site virials from dE/dA.
// k_ace_dens and k_ace_pull give each wavefront SPW sites (SPW = 1 or 2); each
// site owns LPS = 64 / SPW consecutive lanes, one edge per lane, looping over
// chunks of LPS edges.

// Fl[l][m] = sqrt((2l+1)/2pi) * prod(-1/sqrt((l+m')(l+1-m'))) (angular_blocks.jl:25-34).
template <typename T, int L>
__device__ inline void fill_fl(T* Fl) {
  constexpr int NY = (L + 1) * (L + 1);
  const int lane = threadIdx.x;
  if (lane < NY) {
    int l = 0;
    while ((l + 1) * (l + 1) <= lane) ++l;
    const int m = lane - l * l;
    if (m <= l) {
      double F = sqrt((2.0 * l + 1) / (2 * 3.14159265358979323846));
      for (int mm = 1; mm <= m; ++mm) F = -F / sqrt((double)(l + mm) * (l + 1 - mm));
      Fl[l * (L + 1) + m] = (T)F;
    }
  }
}

template <typename T, int L, int SPW>
__global__ __launch_bounds__(64) void k_ace_dens(AceArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NY = (L + 1) * (L + 1);
  constexpr int LPS = kWave / SPW;
  T* Fl = (T*)smem;                      // (L+1)^2
  T* A = Fl + NY;                        // SPW x dstride
  T* Rb = A + SPW * a.dstride;           // rchunk x kStride: one chunk of radial indices
  T* Yb = Rb + a.rchunk * kStride;       // NY x kStride (with proj: the U part)
  T* Yv = Yb + NY * kStride;             // with proj: NY x kStride, the V part
  const int lane = threadIdx.x;
  const int sl = lane / LPS, ll = lane % LPS;
  const int64_t site = (int64_t)blockIdx.x * SPW + sl;
  const bool valid = site < a.n;
  fill_fl<T, L>(Fl);
  for (int k = lane; k < SPW * a.dstride; k += kWave) A[k] = 0;
  T* As = A + sl * a.dstride;
  const int iz = valid ? a.sps[site] : 0;
  const int64_t e0 = valid ? site * a.ecap : 0;
  const int ne = valid ? a.ecount[site] : 0;
  int nemax = ne;
  for (int off = LPS; off < kWave; off <<= 1) nemax = max(nemax, __shfl_xor(nemax, off, kWave));
  const int nq = a.pair.nr;
  const int col0 = sl * LPS;
  double epair = 0;
  __syncthreads();
  // A[k] = sum_e R[e, r_k] Y[e, m_k], the radial indices taken in chunks of
  // rchunk (densities in dperm order, grouped by chunk) so that the radial
  // tile stays small for models with many radial functions
  for (int c0 = 0; c0 < nemax; c0 += LPS) {
    const int e = c0 + ll;
    const bool on = e < ne;
    int ch = 0;
    RadialPoint<T> P{};
    if (on) {
      const double* R = a.eR + 3 * (int64_t)(e0 + e);
      const T dx = (T)R[0], dy = (T)R[1], dz = (T)R[2];
      const T r = (T)sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      const int jz = a.sps[a.enb[e0 + e]];
      ch = iz + a.nz * jz;
      P = radial_point(a.rad, ch, r);
      if (nq && ((a.pair_gate >> jz) & 1u)) {
        const RadialPoint<T> PQ = radial_point(a.pair, ch, r);
        T s = 0;
        for (int n = 0; n < nq; ++n) {
          T v, d;
          radial_value(a.pair, ch, PQ, n, v, d);
          s += v * a.wp[iz * nq + n];
        }
        epair += (double)s;
      }
      T Y[NY], G[3 * NY];
      ylm<T, L>(dx, dy, dz, Fl, Y, G, a.solid);
      if (a.proj) {
        const bool in_v = (a.sel >> jz) & 1u;
#pragma unroll
        for (int m = 0; m < NY; ++m) {
          Yb[m * kStride + lane] = in_v ? (T)0 : Y[m];
          Yv[m * kStride + lane] = in_v ? Y[m] : (T)0;
        }
      } else {
#pragma unroll
        for (int m = 0; m < NY; ++m) Yb[m * kStride + lane] = Y[m];
      }
    }
      const int cn = min(LPS, ne - c0);
    const bool single = a.rchunk >= a.rad.nr;
    for (int r0 = 0, q = 0; r0 < a.rad.nr; r0 += a.rchunk, ++q) {
      if (on) {
        const int r1 = min(a.rad.nr, r0 + a.rchunk);
#pragma unroll 8
        for (int n = r0; n < r1; ++n) {
          T v, d;
          radial_value(a.rad, ch, P, n, v, d);
          Rb[(n - r0) * kStride + lane] = v;
        }
      }
      __syncthreads();
      for (int i = a.dchunk[q] + ll; i < a.dchunk[q + 1]; i += LPS) {
        const int k = single ? i : a.dperm[i];  // one chunk: dperm is the identity
        const T* rr = Rb + (a.dens[2 * k] - r0) * kStride + col0;
        const T* yy = Yb + a.dens[2 * k + 1] * kStride + col0;
        T s = 0;
#pragma unroll 8
        for (int t = 0; t < cn; ++t) s += rr[t] * yy[t];
        As[k] += s;
        if (a.proj) {
          const T* yv = Yv + a.dens[2 * k + 1] * kStride + col0;
          T sv = 0;
#pragma unroll 8
          for (int t = 0; t < cn; ++t) sv += rr[t] * yv[t];
          As[a.ndens + k] += sv;
        }
      }
      __syncthreads();
    }
  }
  if (valid) {
    for (int k = ll; k < a.dstride; k += LPS) a.Abuf[site * a.dstride + k] = As[k];
    for (int off = LPS / 2; off > 0; off >>= 1) epair += __shfl_xor(epair, off, kWave);
    if (ll == 0) a.siteE[site] = epair;
  }
}

// The density contraction on the matrix cores (FP32, unprojected models).
// The densities of one l form the block A_l = sum_e R_l(e)^T Y_l(e) of radial
// indices of degree l by the 2l+1 harmonics of l. The host cuts each block
// (per radial chunk) into tiles of 4 radial indices x 4 harmonics and packs
// 16 tiles into a group: one v_mfma_f32_4x4x1_16b_f32 per group and edge
// adds the 16 outer products of the edge's radial rows and harmonic columns
// to the group's accumulators. Lane 4b+i supplies row i and column i of tile
// b; after the last edge, register i of lane 4b+j holds the tile's entry
// (i, j). Partial tiles read a zero row or column. Staging (radials and Y_lm
// per edge in LDS) is k_ace_dens's; the accumulators stay in registers over
// every chunk and are written to Abuf once.
constexpr int kMfmaGroups = 4;  // tile groups a site keeps in registers (4 x 4 VGPRs per site)
typedef float mfma_v4 __attribute__((ext_vector_type(4)));
struct MfmaTiles {
  const int32_t* tab;  // [groups][6][64]: row, column, densities of registers 0..3
  int groups;
  int chunk[kMfmaGroups];
};

template <int L, int SPW>
__global__ __launch_bounds__(64) void k_ace_dens_mfma(AceArgs<float> a, MfmaTiles t) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NY = (L + 1) * (L + 1);
  constexpr int LPS = kWave / SPW;
  float* Fl = (float*)smem;                     // (L+1)^2
  float* Rb = Fl + NY;                          // (rchunk + 1) x kStride: radial tile, then a zero row
  float* Yb = Rb + (a.rchunk + 1) * kStride;    // (NY + 1) x kStride: harmonics, then a zero row
  const int lane = threadIdx.x;
  const int sl = lane / LPS, ll = lane % LPS;
  const int64_t site = (int64_t)blockIdx.x * SPW + sl;
  const bool valid = site < a.n;
  fill_fl<float, L>(Fl);
  for (int k = lane; k < kStride; k += kWave) {
    Rb[a.rchunk * kStride + k] = 0;
    Yb[NY * kStride + k] = 0;
  }
  int row[kMfmaGroups], col[kMfmaGroups];
  mfma_v4 acc[SPW][kMfmaGroups];
#pragma unroll
  for (int g = 0; g < kMfmaGroups; ++g) {
    row[g] = g < t.groups ? t.tab[(6 * g) * kWave + lane] : a.rchunk;
    col[g] = g < t.groups ? t.tab[(6 * g + 1) * kWave + lane] : NY;
#pragma unroll
    for (int s = 0; s < SPW; ++s) acc[s][g] = mfma_v4{0, 0, 0, 0};
  }
  const int iz = valid ? a.sps[site] : 0;
  const int64_t e0 = valid ? site * a.ecap : 0;
  const int ne = valid ? a.ecount[site] : 0;
  int nemax = ne;
  for (int off = LPS; off < kWave; off <<= 1) nemax = max(nemax, __shfl_xor(nemax, off, kWave));
  const int nq = a.pair.nr;
  double epair = 0;
  __syncthreads();
  for (int c0 = 0; c0 < nemax; c0 += LPS) {
    const int e = c0 + ll;
    const bool on = e < ne;
    int ch = 0;
    RadialPoint<float> P{};
    if (on) {
      const double* R = a.eR + 3 * (int64_t)(e0 + e);
      const float dx = (float)R[0], dy = (float)R[1], dz = (float)R[2];
      const float r = (float)sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      const int jz = a.sps[a.enb[e0 + e]];
      ch = iz + a.nz * jz;
      P = radial_point(a.rad, ch, r);
      if (nq && ((a.pair_gate >> jz) & 1u)) {
        const RadialPoint<float> PQ = radial_point(a.pair, ch, r);
        float sp = 0;
        for (int n = 0; n < nq; ++n) {
          float v, d;
          radial_value(a.pair, ch, PQ, n, v, d);
          sp += v * a.wp[iz * nq + n];
        }
        epair += (double)sp;
      }
      float Y[NY], G[3 * NY];
      ylm<float, L>(dx, dy, dz, Fl, Y, G, a.solid);
#pragma unroll
      for (int m = 0; m < NY; ++m) Yb[m * kStride + lane] = Y[m];
    }
    // edges of each half in this chunk (wave-uniform)
    int cns[SPW];
    int emax = 0;
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      cns[s] = __builtin_amdgcn_readfirstlane(max(0, min(LPS, __shfl(ne, s * LPS, kWave) - c0)));
      emax = max(emax, cns[s]);
    }
    for (int r0 = 0, q = 0; r0 < a.rad.nr; r0 += a.rchunk, ++q) {
      if (on) {
        const int r1 = min(a.rad.nr, r0 + a.rchunk);
#pragma unroll 8
        for (int n = r0; n < r1; ++n) {
          float v, d;
          radial_value(a.rad, ch, P, n, v, d);
          Rb[(n - r0) * kStride + lane] = v;
        }
      }
      __syncthreads();
#pragma unroll
      for (int g = 0; g < kMfmaGroups; ++g) {
        if (g < t.groups && t.chunk[g] == q) {
          const float* rr = Rb + row[g] * kStride;
          const float* yy = Yb + col[g] * kStride;
          // four edges at a time, the operand loads ahead of the MFMAs; a
          // half past its last edge multiplies zeros (the stale columns
          // beyond it may hold anything, so both operands are replaced).
          // Columns stay below (s + 1) LPS since LPS is a multiple of 4.
          for (int ee = 0; ee < emax; ee += 4) {
            float x[SPW][4], y[SPW][4];
#pragma unroll
            for (int s = 0; s < SPW; ++s)
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                x[s][u] = rr[s * LPS + ee + u];
                y[s][u] = yy[s * LPS + ee + u];
              }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
              for (int s = 0; s < SPW; ++s) {
                const bool in = ee + u < cns[s];
                acc[s][g] = __builtin_amdgcn_mfma_f32_4x4x1f32(in ? x[s][u] : 0.0f, in ? y[s][u] : 0.0f, acc[s][g], 0, 0,
                                                               0);
              }
          }
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int s = 0; s < SPW; ++s) {
    const int64_t ss = (int64_t)blockIdx.x * SPW + s;
    if (ss < a.n) {
#pragma unroll
      for (int g = 0; g < kMfmaGroups; ++g) {
        if (g < t.groups) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int k = t.tab[(6 * g + 2 + i) * kWave + lane];
            if (k >= 0) a.Abuf[ss * a.dstride + k] = acc[s][g][i];
          }
        }
      }
    }
  }
  for (int off = LPS / 2; off > 0; off >>= 1) epair += __shfl_xor(epair, off, kWave);
  if (valid && ll == 0) a.siteE[site] = epair;
}

// Products P_k = prod A[f], E = ref + pair + sum_k w_k P_k, and dE/dA
// (projection.jl:39-56; cpu.jl:214-236), forward and reverse in one pass.
// The products run in a host-built schedule: steps of 64 lanes, each lane one
// product with its factors in slots 0..3 and, per slot, a target accumulator
// for that factor's partial derivative. A density owns several accumulators
// (copies) so that within one step and slot no two lanes share a target: the
// partials are plain read-modify-writes in LDS, race-free without atomics, and
// each accumulator sums in schedule order (deterministic). The schedule also
// keeps gathers of A and the accumulator updates in distinct LDS banks as far
// as it can. dE/dA[a] is the sum of a's copies. The schedule, copy offsets and
// the readout w (in schedule order, per species) are staged in LDS once per
// workgroup; each wavefront then walks sites with a grid stride, one site at a
// time, with no workgroup barrier inside the walk.
// GT: the schedule, weights and copy offsets are read from global memory (L2)
// instead of LDS, for models whose tables do not fit beside the per-wavefront
// accumulators.
template <typename T, bool GT>
__global__ __launch_bounds__(1024) void k_ace_prod(AceArgs<T> a, int grads) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr unsigned kNone = 0xffff;
  const int nwaves = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int nslot = a.nsteps * kWave;
  const T* ws;
  const uint4* sched;
  const int* cb;
  unsigned char* pw = smem;
  if constexpr (GT) {
    ws = a.ws;
    sched = a.sched;
    cb = a.copyb;
  } else {
    T* lws = (T*)smem;                                                   // nz x nslot
    uint4* lsched = (uint4*)(smem + align16(sizeof(T) * a.nz * nslot));  // nslot: factors and targets (u16)
    int* lcb = (int*)(lsched + nslot);                                   // ndens + 1 copy offsets
    pw = smem + align16((size_t)((unsigned char*)(lcb + a.ndens + 1) - smem));
    for (int i = threadIdx.x; i < a.nz * nslot; i += blockDim.x) lws[i] = a.ws[i];
    for (int i = threadIdx.x; i < nslot; i += blockDim.x) lsched[i] = a.sched[i];
    for (int i = threadIdx.x; i <= a.ndens; i += blockDim.x) lcb[i] = a.copyb[i];
    __syncthreads();
    ws = lws;
    sched = lsched;
    cb = lcb;
  }
  // accumulators: ntarget + 1 (the last one takes the padding slots' updates);
  // with proj the dE/dV accumulators follow the dE/dU ones
  const int nacc1 = a.ntarget + 1;
  const int nacc = a.proj ? 2 * nacc1 : nacc1;
  const size_t per_wave = align16(sizeof(T) * a.dstride) + align16(sizeof(T) * nacc);
  T* A = (T*)(pw + wave * per_wave);
  T* acc = (T*)((unsigned char*)A + align16(sizeof(T) * a.dstride));
  // kept orders as weights of the coefficients of t^0 .. t^4
  T om[6];
  for (int k = 0; k < 6; ++k) om[k] = (k <= 4 && a.olo <= k && k <= a.ohi) ? (T)1 : (T)0;
  for (int64_t site = (int64_t)blockIdx.x * nwaves + wave; site < a.n; site += (int64_t)gridDim.x * nwaves) {
    T* Ag = a.Abuf + site * a.dstride;
    for (int k = lane; k < a.dstride; k += kWave) A[k] = Ag[k];
    if (grads)
      for (int k = lane; k < nacc; k += kWave) acc[k] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int iz = a.sps[site];
    const T* wz = ws + iz * nslot;
    double eprod = 0;
    if (a.proj) {
      for (int st = 0; st < a.nsteps; ++st) {
        const int slot = st * kWave + lane;
        const uint4 f = sched[slot];
        const T w = wz[slot];
        // the truncated product of projection.jl:7-38: the coefficients c_k of
        // prod_s (U_s + t V_s) over the (at most four) factors; P = sum of the
        // kept orders; d/dU_s and d/dV_s from the coefficients o_k of the
        // product of the other factors: dP/dU_s = sum om_k o_k,
        // dP/dV_s = sum om_{k+1} o_k (unused factors are 1 + 0 t)
        const unsigned fs[4] = {f.x & 0xffff, f.x >> 16, f.y & 0xffff, f.y >> 16};
        T u[4], v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          u[q] = fs[q] != kNone ? A[fs[q]] : (T)1;
          v[q] = fs[q] != kNone ? A[a.ndens + fs[q]] : (T)0;
        }
        T c[5] = {(T)1, (T)0, (T)0, (T)0, (T)0};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
          for (int k = 4; k > 0; --k) c[k] = u[q] * c[k] + v[q] * c[k - 1];
          c[0] = u[q] * c[0];
        }
        T pv = 0;
#pragma unroll
        for (int k = 0; k <= 4; ++k) pv += om[k] * c[k];
        eprod += (double)(w * pv);
        if (grads) {
          const unsigned ts[4] = {f.z & 0xffff, f.z >> 16, f.w & 0xffff, f.w >> 16};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (fs[q] == kNone) continue;
            T o[4] = {(T)1, (T)0, (T)0, (T)0};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              if (r == q) continue;
#pragma unroll
              for (int k = 3; k > 0; --k) o[k] = u[r] * o[k] + v[r] * o[k - 1];
              o[0] = u[r] * o[0];
            }
            T du = 0, dv = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              du += om[k] * o[k];
              dv += om[k + 1] * o[k];
            }
            acc[ts[q]] += w * du;
            acc[nacc1 + ts[q]] += w * dv;
          }
        }
      }
    } else {
      // one order class at a time: every slot of a step has nu factors
#pragma unroll
      for (int nu = 1; nu <= 4; ++nu)
        for (int st = a.ostep[nu - 1]; st < a.ostep[nu]; ++st) {
          const int slot = st * kWave + lane;
          const uint4 f = sched[slot];
          const T w = wz[slot];
          const T x0 = A[f.x & 0xffff];
          if (nu == 1) {
            eprod += (double)(w * x0);
            if (grads) acc[f.z & 0xffff] += w;
          } else if (nu == 2) {
            const T x1 = A[f.x >> 16];
            eprod += (double)(w * x0 * x1);
            if (grads) {
              acc[f.z & 0xffff] += w * x1;
              acc[f.z >> 16] += w * x0;
            }
          } else if (nu == 3) {
            const T x1 = A[f.x >> 16], x2 = A[f.y & 0xffff];
            const T w2 = w * x2;
            eprod += (double)(w2 * x0 * x1);
            if (grads) {
              acc[f.z & 0xffff] += w2 * x1;
              acc[f.z >> 16] += w2 * x0;
              acc[f.w & 0xffff] += w * x0 * x1;
            }
          } else {
            const T x1 = A[f.x >> 16], x2 = A[f.y & 0xffff], x3 = A[f.y >> 16];
            const T p01 = x0 * x1, p23 = x2 * x3;
            eprod += (double)(w * p01 * p23);
            if (grads) {
              const T w23 = w * p23, w01 = w * p01;
              acc[f.z & 0xffff] += w23 * x1;
              acc[f.z >> 16] += w23 * x0;
              acc[f.w & 0xffff] += w01 * x3;
              acc[f.w >> 16] += w01 * x2;
            }
          }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (grads)
      for (int k = lane; k < a.dstride; k += kWave) {
        const int kd = k < a.ndens ? k : k - a.ndens;
        const T* ac = k < a.ndens ? acc : acc + nacc1;
        T d = 0;
        for (int c = cb[kd]; c < cb[kd + 1]; ++c) d += ac[c];
        Ag[k] = d;  // dE/dA (with proj: dE/dU, then dE/dV) over A
      }
    const double es = wave_sum(eprod);
    if (lane == 0) a.siteE[site] = a.reference[iz] + a.siteE[site] + es;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Force pullback g_e = r^ sum_r R'_r T_r + sum_m U_m grad Y_m with
// T_r = sum_m dE/dA[r, m] Y_m and U_m = sum_r dE/dA[r, m] R_r (serial.jl:85-101;
// cpu.jl:226-236), plus the pair radials; edge gradients and site virials.
// FP32 with lmax <= 4 fits 128 VGPRs without spilling: four wavefronts per SIMD
// instead of three (larger lmax and FP64 keep the compiler's choice)
template <typename T, int L, int SPW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && L <= 4 ? 4 : 1))) void k_ace_pull(
    AceArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NY = (L + 1) * (L + 1);
  constexpr int LPS = kWave / SPW;
  T* Fl = (T*)smem;                      // (L+1)^2
  T* dA = Fl + NY;                       // SPW x dstride
  const int lane = threadIdx.x;
  const int sl = lane / LPS, ll = lane % LPS;
  const int64_t site = (int64_t)blockIdx.x * SPW + sl;
  const bool valid = site < a.n;
  fill_fl<T, L>(Fl);
  T* dAs = dA + sl * a.dstride;
  if (valid)
    for (int k = ll; k < a.dstride; k += LPS) dAs[k] = a.Abuf[site * a.dstride + k];
  const int iz = valid ? a.sps[site] : 0;
  const int64_t e0 = valid ? site * a.ecap : 0;
  const int ne = valid ? a.ecount[site] : 0;
  int nemax = ne;
  for (int off = LPS; off < kWave; off <<= 1) nemax = max(nemax, __shfl_xor(nemax, off, kWave));
  const int nq = a.pair.nr;
  __syncthreads();
  double wv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int c0 = 0; c0 < nemax; c0 += LPS) {
    const int e = c0 + ll;
    if (e >= ne) continue;
    const double* R = a.eR + 3 * (int64_t)(e0 + e);
    const T dx = (T)R[0], dy = (T)R[1], dz = (T)R[2];
    const T r = (T)sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    const T ri = (T)1 / r;
    const int jz = a.sps[a.enb[e0 + e]];
    const int ch = iz + a.nz * jz;
    // with proj, the edge reads dE/dV when its neighbour is projected, else dE/dU
    const T* dsel = dAs + ((a.proj && ((a.sel >> jz) & 1u)) ? a.ndens : 0);
    // U_m = sum_r dE/dA[r, m] R_r and V_m = sum_r dE/dA[r, m] R'_r per harmonic;
    // then g = (sum_m V_m Y_m) r^ + sum_m U_m grad Y_m, the harmonics contracted
    // as they are formed (no arrays of values and gradients held)
    T U[NY], V[NY];
#pragma unroll
    for (int m = 0; m < NY; ++m) U[m] = V[m] = 0;
    T gs = 0;
    const RadialPoint<T> P = radial_point(a.rad, ch, r);
#pragma unroll
    for (int l = 0; l <= L; ++l) {
#pragma unroll 4
      for (int q = a.rlb[l]; q < a.rlb[l + 1]; ++q) {
        const int n = a.rl[2 * q], a0 = a.rl[2 * q + 1];
        T v, d;
        radial_value(a.rad, ch, P, n, v, d);
#pragma unroll
        for (int mm = 0; mm <= 2 * l; ++mm) {
          const T c = dsel[a0 + mm];
          U[l * l + mm] += c * v;
          V[l * l + mm] += c * d;
        }
      }
    }
    if (nq && ((a.pair_gate >> jz) & 1u)) {
      const RadialPoint<T> PQ = radial_point(a.pair, ch, r);
      for (int n = 0; n < nq; ++n) {
        T v, d;
        radial_value(a.pair, ch, PQ, n, v, d);
        gs += a.wp[iz * nq + n] * d;
      }
    }
    T h0 = 0, h1 = 0, h2 = 0;
    ylm_emit<T, L>(dx, dy, dz, Fl, a.solid, [&](int i, const Jet<T>& j) {
      gs += V[i] * j.v;
      h0 += U[i] * j.g0;
      h1 += U[i] * j.g1;
      h2 += U[i] * j.g2;
    });
    const T g0 = gs * dx * ri + h0, g1 = gs * dy * ri + h1, g2 = gs * dz * ri + h2;
    T* go = a.eg + 3 * (int64_t)(e0 + e);
    go[0] = g0;
    go[1] = g1;
    go[2] = g2;
    const double gg[3] = {(double)g0, (double)g1, (double)g2};
    for (int bb = 0; bb < 3; ++bb)
      for (int aa = 0; aa < 3; ++aa) wv[aa + 3 * bb] -= R[aa] * gg[bb];
  }
  for (int k = 0; k < 9; ++k) {
    double v = wv[k];
    for (int off = LPS / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (valid && ll == 0) a.siteW[9 * site + k] = v;
  }
}

// The pullback's contractions U_m = sum_r dE/dA[r, m] R_r and V_m = sum_r
// dE/dA[r, m] R'_r on the matrix cores (FP32, unprojected models). For each
// l, each group of 4 harmonics m0..m0+3 and each radial index r of degree l,
// one v_mfma_f32_4x4x1_16b_f32 takes A = the site's dE/dA[r, m0 + i] from
// lane 4b+i and B = R_r of the edge of lane 4b+j, so register i of every lane
// accumulates U_{m0+i} of the lane's own edge (no transposes); a second MFMA
// with B = R'_r gives V. The four lanes of a block share a site (LPS is a
// multiple of 4), so they share the weights. Lanes past their site's edges
// supply zero radials. The rest (pair radials, Y_lm jets, gradients and
// virials) is k_ace_pull's.
__host__ __device__ constexpr int mgroups_below(int l) { return l <= 0 ? 0 : mgroups_below(l - 1) + (2 * (l - 1) + 1 + 3) / 4; }

template <int L, int SPW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(L <= 4 ? 3 : 1))) void k_ace_pull_mfma(AceArgs<float> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NY = (L + 1) * (L + 1);
  constexpr int LPS = kWave / SPW;
  constexpr int NG = mgroups_below(L + 1);
  float* Fl = (float*)smem;   // (L+1)^2
  float* dA = Fl + NY;        // SPW x dstride
  const int lane = threadIdx.x;
  const int sl = lane / LPS, ll = lane % LPS;
  const int64_t site = (int64_t)blockIdx.x * SPW + sl;
  const bool valid = site < a.n;
  fill_fl<float, L>(Fl);
  float* dAs = dA + sl * a.dstride;
  if (valid)
    for (int k = ll; k < a.dstride; k += LPS) dAs[k] = a.Abuf[site * a.dstride + k];
  const int iz = valid ? a.sps[site] : 0;
  const int64_t e0 = valid ? site * a.ecap : 0;
  const int ne = valid ? a.ecount[site] : 0;
  int nemax = ne;
  for (int off = LPS; off < kWave; off <<= 1) nemax = max(nemax, __shfl_xor(nemax, off, kWave));
  const int nq = a.pair.nr;
  const int li = lane & 3;
  __syncthreads();
  double wv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int c0 = 0; c0 < nemax; c0 += LPS) {
    const int e = c0 + ll;
    const bool on = e < ne;
    const double* R = a.eR + 3 * (int64_t)(e0 + (on ? e : 0));
    float dx = 0, dy = 0, dz = 0, r = 1;
    int jz = 0, ch = 0;
    RadialPoint<float> P{};
    if (on) {
      dx = (float)R[0];
      dy = (float)R[1];
      dz = (float)R[2];
      r = (float)sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      jz = a.sps[a.enb[e0 + e]];
      ch = iz + a.nz * jz;
      P = radial_point(a.rad, ch, r);
    }
    mfma_v4 aU[NG], aV[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) aU[g] = aV[g] = mfma_v4{0, 0, 0, 0};
    // four radial indices at a time (unrolled by hand: the compiler does not
    // unroll loops around the MFMA builtin), their coefficient loads ahead of
    // the MFMAs; a lane past its edges has a zero envelope (P = {}), so its
    // radials are zero, and an index past the degree's last repeats the last
    // with zero weights
#pragma unroll
    for (int l = 0; l <= L; ++l) {
      const int q1 = a.rlb[l + 1];
      for (int q = a.rlb[l]; q < q1; q += 4) {
        float v[4], d[4];
        int a0s[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int qq = min(q + u, q1 - 1);
          a0s[u] = a.rl[2 * qq + 1];
          radial_value(a.rad, ch, P, a.rl[2 * qq], v[u], d[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool ok = q + u < q1;
#pragma unroll
          for (int mg = 0; mg < (2 * l + 1 + 3) / 4; ++mg) {
            const float w = ok && 4 * mg + li <= 2 * l ? dAs[a0s[u] + 4 * mg + li] : 0.0f;
            aU[mgroups_below(l) + mg] = __builtin_amdgcn_mfma_f32_4x4x1f32(w, v[u], aU[mgroups_below(l) + mg], 0, 0, 0);
            aV[mgroups_below(l) + mg] = __builtin_amdgcn_mfma_f32_4x4x1f32(w, d[u], aV[mgroups_below(l) + mg], 0, 0, 0);
          }
        }
      }
    }
    if (!on) continue;
    float U[NY], V[NY];
#pragma unroll
    for (int l = 0; l <= L; ++l)
#pragma unroll
      for (int mi = 0; mi <= 2 * l; ++mi) {
        U[l * l + mi] = aU[mgroups_below(l) + mi / 4][mi % 4];
        V[l * l + mi] = aV[mgroups_below(l) + mi / 4][mi % 4];
      }
    float gs = 0;
    if (nq && ((a.pair_gate >> jz) & 1u)) {
      const RadialPoint<float> PQ = radial_point(a.pair, ch, r);
      for (int n = 0; n < nq; ++n) {
        float v, d;
        radial_value(a.pair, ch, PQ, n, v, d);
        gs += a.wp[iz * nq + n] * d;
      }
    }
    const float ri = 1.0f / r;
    float h0 = 0, h1 = 0, h2 = 0;
    ylm_emit<float, L>(dx, dy, dz, Fl, a.solid, [&](int i, const Jet<float>& j) {
      gs += V[i] * j.v;
      h0 += U[i] * j.g0;
      h1 += U[i] * j.g1;
      h2 += U[i] * j.g2;
    });
    const float g0 = gs * dx * ri + h0, g1 = gs * dy * ri + h1, g2 = gs * dz * ri + h2;
    float* go = a.eg + 3 * (int64_t)(e0 + e);
    go[0] = g0;
    go[1] = g1;
    go[2] = g2;
    const double gg[3] = {(double)g0, (double)g1, (double)g2};
    for (int bb = 0; bb < 3; ++bb)
      for (int aa = 0; aa < 3; ++aa) wv[aa + 3 * bb] -= R[aa] * gg[bb];
  }
  for (int k = 0; k < 9; ++k) {
    double v = wv[k];
    for (int off = LPS / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if (valid && ll == 0) a.siteW[9 * site + k] = v;
  }
}

// Directed environments (interface.jl:116-143 + NeighbourList/environments.jl:42-95) in
// fixed-capacity rows: site i's edges occupy slots [i ecap, i ecap + ecount[i]).
// pass 0 only counts (to size ecap); pass 1 writes neighbour, stencil entry and
// exact R, flags status[2] when a site has more than ecap edges, and records
// the largest degree in status[1].
// One workgroup of kEdgeWaves wavefronts per cell. The stencil entries are
// staged in LDS, and the cell's candidates (in stencil-entry, cell order) are
// taken 64 at a time: chunk q of a round by wavefront q mod kEdgeWaves, which
// stages it and tests it against each centre of the cell in turn (the FP32
// prefilter, then the exact FP64 test), keeping every (chunk, centre) hit
// ballot in LDS. A prefix over the round's chunks, in chunk order, then gives
// each centre's first slot in every chunk, and each wavefront writes its
// chunks' edges compacted by the ballots: every centre's edges come out sorted
// by (entry, neighbour) without atomics. The staging (entry search and
// neighbour loads) is thereby spread over the wavefronts; splitting the
// centres instead, each wavefront staging every candidate, was measured twice
// as slow as one wavefront per cell.
// k_ace_edges's status: words 0..3 as the host reads them (0 failure flags,
// 1 largest degree, 2 capacity overflow, 3 largest in-cutoff degree of skin
// rows), then kStatusStripes stripes of (1, 2, 3) that the kernel's atomics
// go to and fold_status() folds into words 1..3 on the host
constexpr int kStatusStripes = 32;
constexpr int kStatusWords = 4 + 3 * kStatusStripes;
constexpr int kEdgeWaves = 4;
constexpr int kEdgeRound = 1;                         // chunks of one wavefront held at once
constexpr int kEdgeChunks = kEdgeWaves * kEdgeRound;  // chunks of one round
constexpr int kSegEntries = 27;  // the radius-1 stencil, for the direct reverse index
__global__ __launch_bounds__(kWave * kEdgeWaves) __attribute__((amdgpu_waves_per_eu(8))) void k_ace_edges(const Grid* __restrict__ grids, const int32_t* __restrict__ cell_cfg,
                                                  int64_t n, const double* __restrict__ xs,
                                                  const int32_t* __restrict__ start, int pass, int ecap,
                                                  int32_t* __restrict__ ecount, int32_t* __restrict__ enb,
                                                  int32_t* __restrict__ eent, double* __restrict__ eR,
                                                  int32_t* __restrict__ status, int skin_rows,
                                                  int32_t* __restrict__ erank, uint8_t* __restrict__ eseg,
                                                  int single) {
  __shared__ int s_pref[kWave + 1], s_start[kWave], s_lat[kWave][3];
  __shared__ double s_sh[kWave][3];
  __shared__ float s_ax[kWave], s_ay[kWave], s_az[kWave];  // the centres, relative to the cell's first atom
  __shared__ double s_cx[kWave], s_cy[kWave], s_cz[kWave];  // and exactly
  __shared__ int s_zero;
  __shared__ unsigned long long s_mask[kEdgeChunks][kWave];  // hit ballot of each (chunk of the round, centre)
  __shared__ int s_off[kEdgeChunks][kWave];                  // each centre's first slot in each chunk
  __shared__ int s_pos[kWave], s_cut[kWave];                  // each centre's hits so far (and inside the cutoff)
  const int cell = blockIdx.x;
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int b = start[cell], cnt = start[cell + 1] - b;
  if (cnt == 0) return;
  const Grid& g = grids[single ? 0 : cell_cfg[cell]];  // one configuration: no dependent load
  int key[3];
  cell_key(g, cell, key);
  const double rx = xs[b], ry = xs[n + b], rz = xs[2 * n + b];
  // FP32 prefilter: pairs outside a thin band above the cutoff never reach the
  // exact FP64 test (which decides, and gives R, for the rest)
  // skin rows: every stencil image inside the list radius, R = (x_j - x_i) + H s
  // (so that the reverse edge's R is exactly -R on every path)
  const double r2max = skin_rows ? g.cs2 : g.c2;
  const float c2hi = (float)(r2max * (1.0 + 1e-4));
  // direct reverse index (k_ace_rev): each centre's segment starts, when the
  // stencil has radius 1 and a row fits 8-bit slots, and each edge's rank
  // among this cell's centres that reach its neighbour through the same
  // entry, when the cell is also one group of centres
  const bool seg_ok = pass && eseg && g.nstencil == kSegEntries && ecap <= 255;
  const bool direct = seg_ok && cnt <= kWave;
  const uint64_t below = (1ull << lane) - 1;
  for (int ag = 0; ag < cnt; ag += kWave) {
    const int na = min(kWave, cnt - ag);
    const int64_t mine = b + ag + lane;
    if (wid == 0) {
      if (lane < na) {
        s_cx[lane] = xs[mine];
        s_cy[lane] = xs[n + mine];
        s_cz[lane] = xs[2 * n + mine];
        s_ax[lane] = (float)(xs[mine] - rx);
        s_ay[lane] = (float)(xs[n + mine] - ry);
        s_az[lane] = (float)(xs[2 * n + mine] - rz);
      }
      s_pos[lane] = 0;
      s_cut[lane] = 0;
    }
    for (int e0 = 0; e0 < g.nstencil; e0 += kWave) {
      const int e = e0 + lane;
      int c = 0;
      if (threadIdx.x == 0) s_zero = -1;
      __syncthreads();
      if (wid == 0 && e < g.nstencil) {  // wavefront 0 stages the stencil entries
        int lat[3];
        bool zero;
        const int nc = stencil_entry(g, key, e, lat, &zero);
        if (nc >= 0) {
          s_start[lane] = start[nc];
          c = start[nc + 1] - start[nc];
          double sh[3];
          exact_shift(g, lat, sh);
          for (int q = 0; q < 3; ++q) {
            s_lat[lane][q] = lat[q];
            s_sh[lane][q] = sh[q];
          }
          if (zero) s_zero = lane;
        }
      }
      if (wid == 0) {
        const int incl = wave_incl_scan(c);
        s_pref[lane + 1] = incl;
        if (lane == 0) s_pref[0] = 0;
      }
      __syncthreads();
      const int total = s_pref[kWave];
      const int ez = s_zero;
      const int nchunks = (total + kWave - 1) / kWave;
      for (int r0 = 0; r0 < nchunks; r0 += kEdgeChunks) {
        const int nq = min(kEdgeChunks, nchunks - r0);
        // this wavefront's chunks q = wid + kEdgeWaves j of the round: stage
        // each (kept in registers for the writes) and ballot its hits per centre
        int cm[kEdgeRound], clo[kEdgeRound];
        double cxj[kEdgeRound], cyj[kEdgeRound], czj[kEdgeRound];
        bool ccand[kEdgeRound];
#pragma unroll
        for (int j = 0; j < kEdgeRound; ++j) {
          const int q = wid + kEdgeWaves * j;
          const int t = (r0 + q) * kWave + lane;
          const bool cand = q < nq && t < total;
          int m = 0, lo = 0;
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
            xj = xs[m];
            yj = xs[n + m];
            zj = xs[2 * n + m];
            fx = (float)((xj + s_sh[lo][0]) - rx);
            fy = (float)((yj + s_sh[lo][1]) - ry);
            fz = (float)((zj + s_sh[lo][2]) - rz);
          }
          ccand[j] = cand;
          cm[j] = m;
          clo[j] = lo;
          cxj[j] = xj;
          cyj[j] = yj;
          czj[j] = zj;
          if (q < nq) {  // uniform over the wavefront
            for (int a = 0; a < na; ++a) {
              bool hit = false;
              const float ux = fx - s_ax[a], uy = fy - s_ay[a], uz = fz - s_az[a];
              if (ux * ux + uy * uy + uz * uz < c2hi && !(lo == ez && m == (int)(b + ag + a))) {
                if (skin_rows) {
#pragma clang fp contract(off)
                  const double d0 = (xj - s_cx[a]) + s_sh[lo][0];
                  const double d1 = (yj - s_cy[a]) + s_sh[lo][1];
                  const double d2 = (zj - s_cz[a]) + s_sh[lo][2];
                  hit = d0 * d0 + d1 * d1 + d2 * d2 < r2max;
                } else {
                  double d[3];
                  int img[3];
                  hit = exact_member(g, s_cx[a], s_cy[a], s_cz[a], xj, yj, zj, s_lat[lo], s_sh[lo], d, img);
                }
              }
              const uint64_t mask = __ballot(hit);
              if (lane == 0) s_mask[q][a] = mask;
            }
          }
        }
        __syncthreads();
        if (wid == 0 && lane < na) {  // each centre's first slot in each chunk, in chunk order
          int acc = s_pos[lane];
          for (int q = 0; q < nq; ++q) {
            s_off[q][lane] = acc;
            acc += __popcll(s_mask[q][lane]);
          }
          s_pos[lane] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kEdgeRound; ++j) {
          const int q = wid + kEdgeWaves * j;
          if (q >= nq) continue;  // uniform over the wavefront
          const int m = cm[j], lo = clo[j];
          const bool seg_first = ccand[j] && (r0 + q) * kWave + lane == s_pref[lo];  // the first candidate of its entry
          int hitc = 0;  // centres so far that reach this candidate (ascending sorted index)
          for (int a = 0; a < na; ++a) {
            const int64_t k = b + ag + a;
            const uint64_t mask = s_mask[q][a];
            const bool hit = (mask >> lane) & 1ull;
            const int slot = s_off[q][a] + __popcll(mask & below);
            double d[3] = {0, 0, 0};
            if (hit) {  // R again, by the same operations as the test
              if (skin_rows) {
#pragma clang fp contract(off)
                d[0] = (cxj[j] - s_cx[a]) + s_sh[lo][0];
                d[1] = (cyj[j] - s_cy[a]) + s_sh[lo][1];
                d[2] = (czj[j] - s_cz[a]) + s_sh[lo][2];
              } else {
                member_disp(g, s_cx[a], s_cy[a], s_cz[a], cxj[j], cyj[j], czj[j], s_lat[lo], s_sh[lo], d);
              }
            }
            if (skin_rows) {  // of which inside the cutoff itself (a launch-shape hint)
              const int hc = __popcll(__ballot(hit && d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < g.c2));
              if (lane == 0 && hc) atomicAdd(&s_cut[a], hc);
            }
            if (pass && hit && slot < ecap) {
              const int64_t qq = k * ecap + slot;
              enb[qq] = m;
              eent[qq] = e0 + lo;
              eR[3 * qq] = d[0];
              eR[3 * qq + 1] = d[1];
              eR[3 * qq + 2] = d[2];
              if (erank) erank[qq] = direct ? hitc : -1;
            }
            if (seg_ok && seg_first && slot < ecap) eseg[k * kSegEntries + lo] = (uint8_t)slot;
            hitc += hit ? 1 : 0;
          }
        }
        __syncthreads();  // the ballots and offsets are rewritten by the next round
      }
    }
    if (wid == 0) {
      const int deg = lane < na ? s_pos[lane] : 0, cut = lane < na ? s_cut[lane] : 0;
      if (lane < na) ecount[mine] = min(deg, ecap);
      // one atomic per workgroup and word (same-address atomics serialise
      // across the device). Skipping the atomic when a coherent read already
      // showed as large a value under-reported the maximum in the Verlet list
      // build (an extra build pass), so every workgroup issues it.
      int dmax = deg, cmax = cut;
      for (int off = kWave / 2; off > 0; off >>= 1) {
        dmax = max(dmax, __shfl_xor(dmax, off, kWave));
        cmax = max(cmax, __shfl_xor(cmax, off, kWave));
      }
      if (lane == 0) {
        // into one of kStatusStripes stripes (the host folds them): same-address
        // device atomics serialize, and every workgroup issues these
        int32_t* sw = status + 4 + 3 * (blockIdx.x % kStatusStripes);
        atomicMax(sw, dmax);
        if (dmax > ecap) atomicOr(sw + 1, 1);  // capacity too small: the host repeats the step
        if (skin_rows) atomicMax(sw + 2, cmax);
      }
    }
    __syncthreads();  // the centres' staging and counters are rewritten by the next group
  }
}

// rev[e] = the edge (j -> i, -R) of edge e = (i -> j, R); R negates exactly.
// The reverse of a pair found through stencil offset o is found through -o,
// the mirrored entry nstencil - 1 - entry, and every centre's edges are sorted
// by (entry, neighbour), so it is located by binary search in j's list (and
// its displacement checked to be exactly -R). One thread per edge slot; every
// used slot's erev is written (-1 when not found, which also sets the status
// word).
__global__ void k_ace_rev(int64_t nslots, int ecap, const Grid* __restrict__ grids,
                          const int32_t* __restrict__ atom_cfg, const int32_t* __restrict__ perm,
                          const int32_t* __restrict__ ecount, const int32_t* __restrict__ enb,
                          const int32_t* __restrict__ eent, const double* __restrict__ eR,
                          int32_t* __restrict__ erev, int32_t* __restrict__ status, const uint8_t* __restrict__ eseg,
                          int nstencil) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nslots) return;
  const int i = (int)(e / ecap);
  if ((int)(e - (int64_t)i * ecap) >= ecount[i]) return;  // an unused slot
  const int j = enb[e];
  // nstencil > 0: every configuration's stencil size
  const int want = (nstencil > 0 ? nstencil : grids[atom_cfg[perm[i]]].nstencil) - 1 - eent[e];
  const int rank = erev[e];  // from k_ace_edges: the direct index, or -1
  int64_t a = (int64_t)j * ecap, b = a + ecount[j];  // first f in [a, b) with (eent, enb)[f] >= (want, i)
  const int64_t end = b;
  if (rank >= 0) {
    // j's segment of entry `want` lists the centres of i's cell that reach j in
    // ascending order; i is the rank-th of them. The displacement needs no
    // check here: both directions form it as fl(fl(x_j - x_i) + H s) with the
    // shift of the mirrored entry exactly -H s (or with minimum-image
    // rounding, odd in its argument), so R(rev) = -R bitwise.
    a += (int64_t)eseg[(int64_t)j * kSegEntries + want] + rank;
    const int found = (a < end && eent[a] == want && enb[a] == i) ? (int)a : -1;
    erev[e] = found;
    if (found < 0) atomicOr(status, 1);
    return;
  } else {
    while (a < b) {
      const int64_t mid = (a + b) >> 1;
      const int em = eent[mid];
      if (em < want || (em == want && enb[mid] < i))
        a = mid + 1;
      else
        b = mid;
    }
  }
  int found = -1;
  if (a < end && eent[a] == want && enb[a] == i && eR[3 * a] == -eR[3 * e] && eR[3 * a + 1] == -eR[3 * e + 1] &&
      eR[3 * a + 2] == -eR[3 * e + 2])
    found = (int)a;
  erev[e] = found;
  if (found < 0) atomicOr(status, 1);
}

// F_k = sum over k's out-edges of (g_e - g_rev(e)), written in original order.
// One wavefront per atom, lanes over its edges, fixed-order wave reduction.
template <typename G>
__global__ __launch_bounds__(64) void k_ace_force(int64_t n, int ecap, const int32_t* __restrict__ ecount,
                                                  const int32_t* __restrict__ erev, const G* __restrict__ eg,
                                                  const int32_t* __restrict__ perm, double* __restrict__ F) {
  const int64_t k = blockIdx.x;
  const int lane = threadIdx.x;
  double f0 = 0, f1 = 0, f2 = 0;
  for (int64_t e = k * ecap + lane; e < k * ecap + ecount[k]; e += kWave) {
    const int r = erev[e];  // -1 only when the status word already reports the failure
    f0 += (double)eg[3 * e];
    f1 += (double)eg[3 * e + 1];
    f2 += (double)eg[3 * e + 2];
    if (r >= 0) {
      f0 -= (double)eg[3 * r];
      f1 -= (double)eg[3 * r + 1];
      f2 -= (double)eg[3 * r + 2];
    }
  }
  f0 = wave_sum(f0);
  f1 = wave_sum(f1);
  f2 = wave_sum(f2);
  if (lane == 0) {
    const int64_t o = perm[k];
    F[3 * o] = f0;
    F[3 * o + 1] = f1;
    F[3 * o + 2] = f2;
  }
}

// ---- Verlet reuse of the environments (skin > 0; the reference's GPUExecutor
// with GPUConfig(skin), src/submodules/ACE/gpu.jl:301-343 over gpu_skin!,
// src/submodules/NeighbourList/gpu.jl:510-549)

// Displacement of every sorted atom since the build (SoA) and the largest
// squared one, as the bits of a nonnegative double in words[parity]; the other
// word is cleared for the next step.
__global__ __launch_bounds__(256) void k_ace_delta(int64_t n, const int32_t* __restrict__ perm,
                                                   const double* __restrict__ X, const double* __restrict__ Xref,
                                                   double* __restrict__ dX, unsigned long long* __restrict__ words,
                                                   int parity) {
  __shared__ double wmax[256 / kWave];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double m2 = 0;
  if (k < n) {
    const int64_t p = perm[k];
    const double dx = X[3 * p] - Xref[3 * p], dy = X[3 * p + 1] - Xref[3 * p + 1], dz = X[3 * p + 2] - Xref[3 * p + 2];
    dX[k] = dx;
    dX[n + k] = dy;
    dX[2 * n + k] = dz;
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

// The edges of site i inside the cutoff now, R = R_build + (dx_j - dx_i), in
// skin-row order (a ballot compaction, one wavefront per site). The reverse
// edge's R is exactly -R (R_build negates exactly, and so does the
// displacement difference), so an edge and its reverse are kept together.
// An edge whose r^2 lies within a relative 1e-10 of c^2 is re-decided by
// exact_image_member from the caller's positions X, as a fresh build decides
// it (a symmetric test), so the kept set is a fresh build's bit for bit.
__global__ __launch_bounds__(64) void k_ace_compact(int64_t n, int ecap, const int32_t* __restrict__ ecount_s,
                                                    const int32_t* __restrict__ enb_s,
                                                    const double* __restrict__ eR_s, const double* __restrict__ dX,
                                                    double c2, int32_t* __restrict__ ecount,
                                                    int32_t* __restrict__ enb, double* __restrict__ eR,
                                                    int32_t* __restrict__ cslot, int32_t* __restrict__ corig,
                                                    const int32_t* __restrict__ perm, const Grid* __restrict__ grids,
                                                    const int32_t* __restrict__ atom_cfg,
                                                    const double* __restrict__ X) {
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t base = i * ecap;
  const int cnt = ecount_s[i];
  const double xi = dX[i], yi = dX[n + i], zi = dX[2 * n + i];
  int pos = 0;
  for (int q0 = 0; q0 < cnt; q0 += kWave) {
    const int q = q0 + lane;
    bool act = false;
    int j = 0;
    double r[3] = {0, 0, 0};
    if (q < cnt) {
      const int64_t e = base + q;
      j = enb_s[e];
      r[0] = eR_s[3 * e] + (dX[j] - xi);
      r[1] = eR_s[3 * e + 1] + (dX[n + j] - yi);
      r[2] = eR_s[3 * e + 2] + (dX[2 * n + j] - zi);
      const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      act = r2 < c2;
      if (fabs(r2 - c2) <= c2 * 1e-10) {
        const int oi = perm[i], oj = perm[j];
        const Grid& g = grids[atom_cfg[oi]];
        int S[3];
        audit_shift(g, r, X, oi, oj, S);
        act = exact_image_member(g, X, oi, oj, S);
      }
    }
    const uint64_t mask = __ballot(act);
    const int c = pos + __popcll(mask & ((1ull << lane) - 1));
    if (q < cnt) cslot[base + q] = act ? (int32_t)(base + c) : -1;
    if (act) {
      const int64_t o = base + c;
      enb[o] = j;
      eR[3 * o] = r[0];
      eR[3 * o + 1] = r[1];
      eR[3 * o + 2] = r[2];
      corig[o] = (int32_t)(base + q);
    }
    pos += __popcll(mask);
  }
  if (lane == 0) ecount[i] = pos;
}

// Reverse edges of the compacted rows from the skin rows' reverse index.
__global__ void k_ace_revmap(int64_t slots, int ecap, const int32_t* __restrict__ ecount,
                             const int32_t* __restrict__ corig, const int32_t* __restrict__ erev_s,
                             const int32_t* __restrict__ cslot, int32_t* __restrict__ erev,
                             int32_t* __restrict__ status) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slots) return;
  const int64_t i = e / ecap;
  if (e - i * ecap >= ecount[i]) return;
  const int rs = erev_s[corig[e]];
  const int r = rs >= 0 ? cslot[rs] : -1;
  erev[e] = r;
  if (r < 0) atomicOr(status, 1);
}

// Pair-set digest of the rows the site kernels evaluated: every used slot's
// (neighbour, image) with its exact FP64 R.
__global__ void k_ace_audit(int64_t slots, int ecap, const int32_t* __restrict__ ecount,
                            const int32_t* __restrict__ enb, const double* __restrict__ eR,
                            const int32_t* __restrict__ perm, const Grid* __restrict__ grids,
                            const int32_t* __restrict__ atom_cfg, AuditBufs au) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slots) return;
  const int64_t i = e / ecap;
  if (e - i * ecap >= ecount[i]) return;
  const int oi = perm[i];
  const double d[3] = {eR[3 * e], eR[3 * e + 1], eR[3 * e + 2]};
  audit_add(au, grids[atom_cfg[oi]], d, oi, perm[enb[e]]);
}

// siteE and siteW packed as 10-column records for the two-stage reduction.
__global__ void k_ace_pack(int64_t n, const double* __restrict__ siteE, const double* __restrict__ siteW, int grads,
                           double* __restrict__ rec) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  rec[10 * k] = siteE[k];
  for (int c = 0; c < 9; ++c) rec[10 * k + 1 + c] = grads ? siteW[9 * k + c] : 0.0;
}

__global__ void k_ace_scatter_sites(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ siteE,
                                    double* __restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[perm[k]] = siteE[k];
}

// ---------------------------------------------------------------- adjoints
//
// Gradient of a loss L(E_c, F) with respect to the precontracted readout
// w_k(z) and the pair weights Wpair(n, z), given Ebar_c = dL/dE_c and
// Fbar = dL/dF and Wbar_c = dL/dW_c. With g_e = dE_i/dR_e for edge
// e = (i -> j), F_i += g_e, F_j -= g_e and W_ab -= R_a g_b, so the force and
// virial part is sum_e g_e . delta_e with delta_e = Fbar_i - Fbar_j - Wbar^T R_e:
//   dL/dw_k(z)     = sum_{i: z_i = z} [Ebar P_k(A_i) + sum_a dP_k/dA_a Adot_a]
//   dL/dWpair(n,z) = sum_{i: z_i = z} sum_e [Ebar Q_n(r_e) + Q'_n(r_e) r^_e . delta_e]
// with Adot_a = sum_e [R'_ra (r^ . delta_e) Y_ma + R_ra (grad Y_ma . delta_e)],
// the JVP of the site basis along the edge directions (the reference's
// ace_vjp_impl / scalar_site_product, ext/PotentialFrameWorkForwardDiffExt.jl:63-109,
// and site_parameter_gradient, src/submodules/ACE/serial.jl:103-121).
// Environments (radials, Y_lm) are recomputed, nothing per edge is stored.
// Each workgroup walks sites with a grid stride and accumulates its partial
// gradient in LDS; a second kernel sums the workgroup partials in FP64.
template <typename T, int L>
__global__ __launch_bounds__(64) void k_ace_adjoint(AceArgs<T> a, const int32_t* __restrict__ atom_cfg,
                                                    const double* __restrict__ Ebar, const double* __restrict__ Fbar,
                                                    const double* __restrict__ Wbar, double* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NY = (L + 1) * (L + 1);
  const int nq = a.pair.nr;
  const int gsize = a.nz * (a.nprod + nq);
  double* G = (double*)smem;             // nz x (nprod + nq) partial gradient, FP64
  T* Fl = (T*)(G + gsize);               // (L+1)^2
  T* A = Fl + NY;                        // ndens
  T* Ad = A + a.ndens;                   // ndens
  T* Rb = Ad + a.ndens;                  // nr x kStride
  T* Rdb = Rb + a.rad.nr * kStride;      // nr x kStride
  T* Yb = Rdb + a.rad.nr * kStride;      // NY x kStride
  T* Ydb = Yb + NY * kStride;            // NY x kStride
  const int lane = threadIdx.x;
  if (lane < NY) {
    int l = 0;
    while ((l + 1) * (l + 1) <= lane) ++l;
    const int m = lane - l * l;
    if (m <= l) {
      double F = sqrt((2.0 * l + 1) / (2 * 3.14159265358979323846));
      for (int mm = 1; mm <= m; ++mm) F = -F / sqrt((double)(l + mm) * (l + 1 - mm));
      Fl[l * (L + 1) + m] = (T)F;
    }
  }
  for (int k = lane; k < gsize; k += kWave) G[k] = 0;
  __syncthreads();
  for (int64_t site = blockIdx.x; site < a.n; site += gridDim.x) {
    const int iz = a.sps[site];
    const int oi = a.perm[site];
    const int cfg = atom_cfg[oi];
    const T eb = (T)Ebar[cfg];
    double wb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) wb[k] = Wbar ? Wbar[9 * cfg + k] : 0.0;
    const double fi0 = Fbar[3 * oi], fi1 = Fbar[3 * oi + 1], fi2 = Fbar[3 * oi + 2];
    const int64_t e0 = site * a.ecap;
    const int ne = a.ecount[site];
    for (int k = lane; k < a.ndens; k += kWave) A[k] = Ad[k] = 0;
    __syncthreads();
    for (int c0 = 0; c0 < ne; c0 += kWave) {
      const int e = c0 + lane;
      const bool act = e < ne;
      T dx = 1, dy = 0, dz = 0, r = 1, de0 = 0, de1 = 0, de2 = 0;
      int ch = 0;
      if (act) {
        const double* R = a.eR + 3 * (int64_t)(e0 + e);
        const int j = a.enb[e0 + e];
        const int oj = a.perm[j];
        dx = (T)R[0];
        dy = (T)R[1];
        dz = (T)R[2];
        r = (T)sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        de0 = (T)(fi0 - Fbar[3 * oj] - (wb[0] * R[0] + wb[1] * R[1] + wb[2] * R[2]));
        de1 = (T)(fi1 - Fbar[3 * oj + 1] - (wb[3] * R[0] + wb[4] * R[1] + wb[5] * R[2]));
        de2 = (T)(fi2 - Fbar[3 * oj + 2] - (wb[6] * R[0] + wb[7] * R[1] + wb[8] * R[2]));
        ch = iz + a.nz * a.sps[j];
      }
      const T sdot = (dx * de0 + dy * de1 + dz * de2) / r;  // r^ . delta
      if (act) {
        const RadialPoint<T> P = radial_point(a.rad, ch, r);
#pragma unroll 8
        for (int n = 0; n < a.rad.nr; ++n) {
          T v, d;
          radial_value(a.rad, ch, P, n, v, d);
          Rb[n * kStride + lane] = v;
          Rdb[n * kStride + lane] = d * sdot;
        }
        T Y[NY], Gy[3 * NY];
        ylm<T, L>(dx, dy, dz, Fl, Y, Gy, a.solid);
#pragma unroll
        for (int m = 0; m < NY; ++m) {
          Yb[m * kStride + lane] = Y[m];
          Ydb[m * kStride + lane] = Gy[3 * m] * de0 + Gy[3 * m + 1] * de1 + Gy[3 * m + 2] * de2;
        }
      }
      if (nq) {  // pair weights: wave sums over the chunk's edges
        RadialPoint<T> PQ;
        if (act) PQ = radial_point(a.pair, ch, r);
        for (int n = 0; n < nq; ++n) {
          T v = 0, d = 0;
          if (act) radial_value(a.pair, ch, PQ, n, v, d);
          const T t = wave_sum(act ? eb * v + d * sdot : (T)0);
          if (lane == 0) G[iz * (a.nprod + nq) + a.nprod + n] += (double)t;
        }
      }
      __syncthreads();
      const int cn = min(kWave, ne - c0);
      for (int k = lane; k < a.ndens; k += kWave) {
        const int rn = a.dens[2 * k], mm = a.dens[2 * k + 1];
        const T* rr = Rb + rn * kStride;
        const T* rd = Rdb + rn * kStride;
        const T* yy = Yb + mm * kStride;
        const T* yd = Ydb + mm * kStride;
        T s = 0, sd = 0;
#pragma unroll 8
        for (int t = 0; t < cn; ++t) {
          s += rr[t] * yy[t];
          sd += rd[t] * yy[t] + rr[t] * yd[t];
        }
        A[k] += s;
        Ad[k] += sd;
      }
      __syncthreads();
    }
    // products: value and tangent; each lane owns its products' gradient entries
    double* Gz = G + iz * (a.nprod + nq);
    for (int p = lane; p < a.nprod; p += kWave) {
      const int* f = a.prod + 4 * p;
      T v = 1, dv = 0;
      for (int t = 0; t < 4 && f[t] >= 0; ++t) {
        const T x = A[f[t]], dx = Ad[f[t]];
        dv = dv * x + v * dx;
        v *= x;
      }
      Gz[p] += (double)(eb * v + dv);
    }
    __syncthreads();
  }
  for (int k = lane; k < gsize; k += kWave) partial[(int64_t)blockIdx.x * gsize + k] = G[k];
}

// ---------------------------------------------------------------- host

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <typename T>
static T* upload(const std::vector<T>& v, hipStream_t st) {
  T* d = nullptr;
  PFW_HIP(hipMalloc((void**)&d, std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (!v.empty()) PFW_HIP(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return d;
}

static void free_family(AceFamilyDev& f) {
  dfree(f.pq);
  dfree(f.env_kind);
  dfree(f.env_pow);
  dfree(f.par64);
  dfree(f.par32);
  dfree(f.nodes64);
  dfree(f.nodes32);
  dfree(f.coef64);
  dfree(f.coef32);
  f = AceFamilyDev();
}

void ace_free_edges(pfw_ctx* ctx) {
  AceDev& a = ctx->ace;
  if (a.h_status) (void)hipHostFree(a.h_status);
  a.h_status = nullptr;
  dfree(a.ecount);
  dfree(a.enb);
  dfree(a.eR);
  dfree(a.erev);
  dfree(a.eseg);
  dfree(a.eent);
  dfree(a.eg);
  dfree(a.siteE);
  dfree(a.Abuf);
  dfree(a.siteW);
  dfree(a.rec);
  dfree(a.status);
  a.edge_cap = 0;
  a.ecap = 0;
  dfree(a.corig);
  dfree(a.ecount_s);
  dfree(a.enb_s);
  dfree(a.eent_s);
  dfree(a.erev_s);
  dfree(a.eseg_s);
  dfree(a.eR_s);
  dfree(a.cslot);
  dfree(a.Xref);
  dfree(a.dX);
  dfree(a.maxd2);
  if (a.h_maxd2) (void)hipHostFree(a.h_maxd2);
  a.h_maxd2 = nullptr;
  a.skin_valid = false;
  a.compact_degree = 0;
  a.ecap_s = 0;
  a.skin_cap = 0;
}

void ace_free(pfw_ctx* ctx) {
  AceDev& a = ctx->ace;
  ace_free_edges(ctx);
  free_family(a.radial);
  free_family(a.pair);
  dfree(a.reference);
  dfree(a.prod);
  dfree(a.sched);
  dfree(a.copyb);
  dfree(a.ws64);
  dfree(a.ws32);
  dfree(a.dens);
  dfree(a.rlb);
  dfree(a.dperm);
  dfree(a.dchunk);
  dfree(a.mtab);
  a.mgroups = 0;
  dfree(a.rl);
  dfree(a.w64);
  dfree(a.w32);
  dfree(a.wp64);
  dfree(a.wp32);
  a = AceDev();
}

static void load_family(const pfw_ace_radial& src, int nz, AceFamilyDev& f, double& cutoff, hipStream_t st) {
  f.nr = src.nr;
  f.nnodes = src.nnodes;
  f.nch = nz * nz;
  if (src.nr < 0) throw Error(PFW_INVALID_ARGUMENT, "negative radial count");
  if (src.nr == 0) return;
  if (src.nnodes < 2 || !src.transform || !src.envelope_kind || !src.envelope || !src.nodes || !src.coefficients)
    throw Error(PFW_INVALID_ARGUMENT, "incomplete radial family");
  const int nch = f.nch, nn = src.nnodes, nint = nn - 1;
  std::vector<int> pq(2 * nch), kind(nch), epow(2 * nch);
  std::vector<double> par(12 * nch, 0.0);
  std::vector<double> nodes(src.nodes, src.nodes + (size_t)nch * nn);
  std::vector<double> coef(src.coefficients, src.coefficients + (size_t)nch * 4 * src.nr * nint);
  for (int c = 0; c < nch; ++c) {
    const double* t = src.transform + 8 * c;
    int p = (int)t[0], q = (int)t[1];
    if (!(p > 0 && q >= p && t[2] > 0 && t[7] > t[4] && t[4] > t[3]))
      throw Error(PFW_INVALID_ARGUMENT, "invalid Agnesi transform");
    pq[2 * c] = p;
    pq[2 * c + 1] = q;
    for (int k = 0; k < 5; ++k) par[12 * c + k] = t[2 + k];  // a rin r0 yin ycut
    par[12 * c + 5] = t[7];                                // rcut
    cutoff = std::max(cutoff, t[7]);
    const double* e = src.envelope + 5 * c;
    kind[c] = src.envelope_kind[c];
    if (kind[c] == PFW_ENV_POLY2SX) {
      par[12 * c + 6] = e[0];
      par[12 * c + 7] = e[1];
      par[12 * c + 8] = e[4];
      epow[2 * c] = (int)e[2];
      epow[2 * c + 1] = (int)e[3];
    } else if (kind[c] == PFW_ENV_POLY1SR) {
      par[12 * c + 6] = e[0];
      epow[2 * c] = (int)e[1];
    } else if (kind[c] == PFW_ENV_ACE1_POLY1SR) {
      par[12 * c + 6] = e[0];
      par[12 * c + 7] = e[1];
      epow[2 * c] = (int)e[2];
    } else {
      throw Error(PFW_INVALID_ARGUMENT, "unsupported radial envelope");
    }
    const double* nd = src.nodes + (size_t)c * nn;
    if (!(nd[0] == -1.0 && nd[nn - 1] == 1.0)) throw Error(PFW_INVALID_ARGUMENT, "invalid spline nodes");
    for (int k = 0; k + 1 < nn; ++k)
      if (!(nd[k + 1] > nd[k])) throw Error(PFW_INVALID_ARGUMENT, "invalid spline nodes");
    for (int k = 0; k < nn; ++k)
      if (std::fabs(nd[k] - (-1.0 + 2.0 * k / (nn - 1))) > 1e-12) f.uniform = false;
  }
  for (double v : coef)
    if (!std::isfinite(v)) throw Error(PFW_INVALID_ARGUMENT, "nonfinite spline coefficients");
  std::vector<float> par32(par.begin(), par.end()), nodes32(nodes.begin(), nodes.end()), coef32(coef.begin(), coef.end());
  f.pq = upload(pq, st);
  f.env_kind = upload(kind, st);
  f.env_pow = upload(epow, st);
  f.par64 = upload(par, st);
  f.par32 = upload(par32, st);
  f.nodes64 = upload(nodes, st);
  f.nodes32 = upload(nodes32, st);
  f.coef64 = upload(coef, st);
  f.coef32 = upload(coef32, st);
}

// Schedule of the product basis for k_ace_prod. Steps of 64 lanes; each lane
// takes one product (padding only where nothing fits) with its factors
// permuted into slots 0..3 and, per slot, an accumulator copy of that factor's
// density that no other lane of the step uses in the same slot (a density a
// has copies[a] = ceil(occurrences / ideal steps) accumulators). Among the
// next candidates in product order a lane takes the first whose gathers of A
// (same index or a free bank) and accumulator updates (a free bank) are
// conflict-free within its 32-lane half, else the one with fewest conflicts.
struct ProductSlot {
  int32_t k = -1;
  int32_t f[4] = {-1, -1, -1, -1};
  int32_t t[4] = {-1, -1, -1, -1};
};

// Products are scheduled one order class at a time (all products of one
// step have the same number of factors, so the kernel runs a body without
// per-lane factor checks); steps [order_step[nu - 1], order_step[nu]) hold the
// products of nu factors. Padding slots read density 0 and update the spare
// accumulator ntarget (never read), with a zero readout weight.
static std::vector<ProductSlot> schedule_class(const std::vector<std::vector<int>>& factors,
                                               std::vector<int> remaining, const std::vector<int32_t>& copyb,
                                               int ndens, int ntarget, int nu);

static std::vector<ProductSlot> schedule_products(const std::vector<std::vector<int>>& factors, int ndens,
                                                  std::vector<int32_t>& copyb, int* order_step) {
  const int nprod = (int)factors.size();
  const int ideal = std::max(1, (nprod + kWave - 1) / kWave);
  std::vector<int> occ(ndens, 0);
  for (const auto& f : factors)
    for (int d : f) ++occ[d];
  copyb.assign(ndens + 1, 0);
  // at least two accumulator copies per density, so that a step can usually
  // pick one whose LDS bank is free in its half-wave (at most 64 are chosen from)
  for (int d = 0; d < ndens; ++d) copyb[d + 1] = copyb[d] + std::max(2, (occ[d] + ideal - 1) / ideal);
  std::vector<ProductSlot> out;
  order_step[0] = 0;
  for (int nu = 1; nu <= 4; ++nu) {
    std::vector<int> cls;
    for (int k = 0; k < nprod; ++k)
      if ((int)factors[k].size() == nu) cls.push_back(k);
    if (!cls.empty()) {
      std::vector<ProductSlot> part = schedule_class(factors, cls, copyb, ndens, copyb[ndens], nu);
      out.insert(out.end(), part.begin(), part.end());
    }
    order_step[nu] = (int)(out.size() / kWave);
  }
  return out;
}

static std::vector<ProductSlot> schedule_class(const std::vector<std::vector<int>>& factors,
                                               std::vector<int> remaining, const std::vector<int32_t>& copyb,
                                               int ndens, int ntarget, int nu_class) {
  std::vector<ProductSlot> out;
  constexpr size_t kWindow = 512;
  // per factor slot and density: the accumulator copies this step has taken
  std::vector<std::vector<uint64_t>> used(4, std::vector<uint64_t>(ndens));
  while (!remaining.empty()) {
    for (auto& c : used) std::fill(c.begin(), c.end(), 0ull);
    int abank[2][4][32];
    for (auto& h : abank)
      for (auto& t : h)
        for (int& v : t) v = -1;
    uint32_t tbank[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    std::vector<char> taken(remaining.size(), 0);
    for (int lane = 0; lane < kWave; ++lane) {
      const int half = lane / 32;
      ProductSlot best;
      int best_cost = 1 << 30;
      size_t best_r = remaining.size(), seen = 0;
      for (size_t r = 0; r < remaining.size() && seen < kWindow && best_cost > 0; ++r) {
        if (taken[r]) continue;
        ++seen;
        const std::vector<int>& f = factors[remaining[r]];
        const int nu = (int)f.size();
        int perm[4] = {0, 1, 2, 3};
        do {
          int cost = 0;
          bool fits = true;
          ProductSlot cand;
          cand.k = remaining[r];
          for (int t = 0; t < nu && fits; ++t) {
            const int d = f[perm[t]];
            // a copy of d not yet taken in this step and slot, preferably in a
            // bank no lane of this half-wave writes in this slot
            const int nc = std::min(64, copyb[d + 1] - copyb[d]);
            int tgt = -1, tcost = 2;
            for (int c = 0; c < nc && tcost > 0; ++c) {
              if ((used[t][d] >> c) & 1ull) continue;
              const int cand_t = copyb[d] + c;
              const int cc = ((tbank[half][t] >> (cand_t & 31)) & 1u) ? 1 : 0;
              if (cc < tcost) {
                tcost = cc;
                tgt = cand_t;
              }
            }
            if (tgt < 0) {
              fits = false;
              break;
            }
            const int ab = abank[half][t][d & 31];
            cost += (ab >= 0 && ab != d) ? 1 : 0;
            cost += tcost;
            cand.f[t] = d;
            cand.t[t] = tgt;
          }
          if (fits && cost < best_cost) {
            best_cost = cost;
            best = cand;
            best_r = r;
            if (cost == 0) break;
          }
        } while (std::next_permutation(perm, perm + nu));
      }
      if (best_r < remaining.size()) {
        for (int t = 0; t < 4 && best.f[t] >= 0; ++t) {
          const int d = best.f[t];
          used[t][d] |= 1ull << (best.t[t] - copyb[d]);
          if (abank[half][t][d & 31] < 0) abank[half][t][d & 31] = d;
          tbank[half][t] |= 1u << (best.t[t] & 31);
        }
        taken[best_r] = 1;
      } else {  // padding: density 0 into the spare accumulator
        for (int t = 0; t < nu_class; ++t) {
          best.f[t] = 0;
          best.t[t] = ntarget;
        }
      }
      out.push_back(best);
    }
    std::vector<int> rest;
    for (size_t r = 0; r < remaining.size(); ++r)
      if (!taken[r]) rest.push_back(remaining[r]);
    if (rest.size() == remaining.size()) throw Error(PFW_INVALID_ARGUMENT, "product schedule failed");
    remaining.swap(rest);
  }
  return out;
}

// Device readout weights of (WB, Wpair): the precontracted product weights
// w (nz x nprod), the same in product-schedule order ws (nz x slots) and the
// pair weights wp (nz x nq).
static void readout_weights(const AceDev& a, const double* WB, const double* Wpair, std::vector<double>& w,
                            std::vector<double>& ws, std::vector<double>& wp) {
  const int nz = a.nz, nq = a.pair.nr;
  w.assign((size_t)nz * a.nprod, 0.0);  // z-major: w[z * nprod + k]
  for (size_t t = 0; t < a.c_vals.size(); ++t)
    for (int z = 0; z < nz; ++z) {
      double b = WB[(size_t)a.c_rows[t] + (size_t)a.nbasis * z];
      if (!std::isfinite(b)) throw Error(PFW_INVALID_ARGUMENT, "nonfinite readout parameters");
      w[(size_t)z * a.nprod + a.c_cols[t]] += a.c_vals[t] * b;
    }
  wp.assign((size_t)nz * std::max(nq, 1), 0.0);
  for (int z = 0; z < nz; ++z)
    for (int q = 0; q < nq; ++q) {
      double v = Wpair ? Wpair[q + (size_t)nq * z] : 0.0;
      if (!std::isfinite(v)) throw Error(PFW_INVALID_ARGUMENT, "nonfinite pair parameters");
      wp[(size_t)z * nq + q] = v;
    }
  const size_t nslot = a.slot_prod.size();  // readout weights in schedule order
  ws.assign((size_t)nz * std::max<size_t>(nslot, 1), 0.0);
  for (int z = 0; z < nz; ++z)
    for (size_t q = 0; q < nslot; ++q)
      if (a.slot_prod[q] >= 0) ws[z * nslot + q] = w[(size_t)z * a.nprod + a.slot_prod[q]];
}

void ace_set_parameters(pfw_ctx* ctx, const double* WB, const double* Wpair) {
  AceDev& a = ctx->ace;
  if (!a.ready && a.nprod == 0) throw Error(PFW_STALE, "no ACE model set");
  std::vector<double> w, ws, wp;
  readout_weights(a, WB, Wpair, w, ws, wp);
  std::vector<float> w32(w.begin(), w.end()), wp32(wp.begin(), wp.end()), ws32(ws.begin(), ws.end());
  dfree(a.ws64);
  dfree(a.ws32);
  a.ws64 = upload(ws, ctx->stream);
  a.ws32 = upload(ws32, ctx->stream);
  dfree(a.w64);
  dfree(a.w32);
  dfree(a.wp64);
  dfree(a.wp32);
  hipStream_t st = ctx->stream;
  a.w64 = upload(w, st);
  a.w32 = upload(w32, st);
  a.wp64 = upload(wp, st);
  a.wp32 = upload(wp32, st);
  PFW_HIP(hipStreamSynchronize(st));
}

// Densities grouped by chunks of `rchunk` radial indices for the density
// kernel's radial tile (0: the built-in 40, capped at the radial count).
// The matrix-core density tiles of k_ace_dens_mfma for the current radial
// chunk: per chunk and l, the chunk's radial indices of degree l in groups of
// 4 (rows) by the 2l+1 harmonics in groups of 4 (columns); 16 tiles to a
// group, a new group for each chunk. mgroups = 0 when the tiles need more
// than kMfmaGroups groups or the model is species-projected.
static void ace_mfma_tiles(pfw_ctx* ctx) {
  AceDev& a = ctx->ace;
  dfree(a.mtab);
  a.mgroups = 0;
  a.mdefault = false;
  const int nd = (int)(a.h_dens.size() / 2);
  if (a.proj || nd == 0) return;
  const int Lk = a.lmax <= 2 ? 2 : a.lmax <= 4 ? 4 : 6;  // the kernel's harmonic rows (dispatch_sites)
  const int zcol = (Lk + 1) * (Lk + 1), zrow = a.rchunk;
  // first density and degree of each radial index (m-blocks are contiguous and complete)
  std::vector<int> first(a.radial.nr, -1), deg(a.radial.nr, -1);
  for (int k = 0; k < nd; ++k) {
    const int r = a.h_dens[2 * k], h = a.h_dens[2 * k + 1];
    int l = 0;
    while ((l + 1) * (l + 1) <= h) ++l;
    if (h == l * l) first[r] = k;
    deg[r] = l;
  }
  std::vector<int32_t> tab;
  int groups = 0, in_group = 16, tiles = 0;
  int gch[kMfmaGroups] = {0, 0, 0, 0};
  for (int r0 = 0, q = 0; r0 < a.radial.nr; r0 += a.rchunk, ++q) {
    in_group = 16;  // every chunk starts a group
    for (int l = 0; l <= a.lmax; ++l) {
      std::vector<int> rs;
      for (int r = r0; r < std::min(a.radial.nr, r0 + a.rchunk); ++r)
        if (first[r] >= 0 && deg[r] == l) rs.push_back(r);
      for (size_t rg = 0; rg < rs.size(); rg += 4)
        for (int mg = 0; mg < 2 * l + 1; mg += 4) {
          if (in_group == 16) {
            if (groups == kMfmaGroups) return;
            gch[groups++] = q;
            in_group = 0;
            const size_t base = tab.size();
            tab.resize(base + 6 * kWave, -1);
            for (int x = 0; x < kWave; ++x) {
              tab[base + x] = zrow;
              tab[base + kWave + x] = zcol;
            }
          }
          int32_t* T = tab.data() + tab.size() - 6 * kWave;
          const int b = in_group++;
          ++tiles;
          for (int i = 0; i < 4; ++i) {
            const bool ri = rg + i < rs.size(), ci = mg + i <= 2 * l;
            T[4 * b + i] = ri ? rs[rg + i] - r0 : zrow;
            T[kWave + 4 * b + i] = ci ? l * l + mg + i : zcol;
            for (int j = 0; j < 4; ++j)  // register i of lane 4b + j: entry (i, j)
              T[(2 + i) * kWave + 4 * b + j] = ri && mg + j <= 2 * l ? first[rs[rg + i]] + mg + j : -1;
          }
        }
    }
  }
  a.mtab = upload(tab, ctx->stream);
  a.mgroups = groups;
  a.mdefault = 2 * nd >= 16 * tiles;
  for (int g = 0; g < kMfmaGroups; ++g) a.mgchunk[g] = gch[g];
}

void ace_group_densities(pfw_ctx* ctx, int rchunk) {
  AceDev& a = ctx->ace;
  a.rchunk = std::max(1, std::min(a.radial.nr, rchunk > 0 ? rchunk : kRadialChunk));
  std::vector<int32_t> dperm, dchunk(1, 0);
  const int nd = (int)(a.h_dens.size() / 2);
  for (int r0 = 0; r0 < std::max(a.radial.nr, 1); r0 += a.rchunk) {
    for (int k = 0; k < nd; ++k)
      if (a.h_dens[2 * k] >= r0 && a.h_dens[2 * k] < r0 + a.rchunk) dperm.push_back(k);
    dchunk.push_back((int32_t)dperm.size());
  }
  dfree(a.dperm);
  dfree(a.dchunk);
  a.dperm = upload(dperm, ctx->stream);
  a.dchunk = upload(dchunk, ctx->stream);
  ace_mfma_tiles(ctx);
}

void ace_set_model(pfw_ctx* ctx, const pfw_ace_model* m) {
  ace_free(ctx);
  AceDev& a = ctx->ace;
  hipStream_t st = ctx->stream;
  if (m->nz <= 0 || m->lmax < 0 || m->ndens < 0 || m->nprod < 0 || m->nbasis < 0)
    throw Error(PFW_INVALID_ARGUMENT, "invalid ACE model dimensions");
  if (m->lmax > 6) throw Error(PFW_INVALID_ARGUMENT, "lmax above 6 is not supported by the device kernels");
  a.nz = m->nz;
  a.lmax = m->lmax;
  if (m->angular != PFW_ANGULAR_SPHERICAL && m->angular != PFW_ANGULAR_SOLID)
    throw Error(PFW_INVALID_ARGUMENT, "unknown angular basis");
  a.solid = m->angular == PFW_ANGULAR_SOLID ? 1 : 0;
  // basis policy: species projection with restricted orders (interface.jl:115-125)
  a.proj = m->order_lo >= 0 ? 1 : 0;
  a.order_lo = m->order_lo;
  a.order_hi = m->order_hi;
  a.sel = 0;
  a.pair_gate = 0xffffffffu;
  if (a.proj) {
    if (m->order_hi < m->order_lo) throw Error(PFW_INVALID_ARGUMENT, "orders must be a nonnegative inclusive range");
    if (m->nz > 32) throw Error(PFW_INVALID_ARGUMENT, "species projection supports at most 32 species channels");
    for (int z = 0; z < m->nz; ++z)
      if (m->selected && m->selected[z]) a.sel |= 1u << z;
    // order_allowed(z in species ? 1 : 0) of pair_allowed (interface.jl:123)
    a.pair_gate = 0;
    for (int z = 0; z < m->nz; ++z) {
      const int k = ((a.sel >> z) & 1u) ? 1 : 0;
      if (m->order_lo <= k && k <= m->order_hi) a.pair_gate |= 1u << z;
    }
  } else if (m->order_hi >= 0) {
    throw Error(PFW_INVALID_ARGUMENT, "orders must be a nonnegative inclusive range");
  }
  a.ndens = m->ndens;
  a.nprod = m->nprod;
  a.nbasis = m->nbasis;
  a.ny = (m->lmax + 1) * (m->lmax + 1);
  double cutoff = 0;
  load_family(m->radial, m->nz, a.radial, cutoff, st);
  load_family(m->pair, m->nz, a.pair, cutoff, st);
  if (m->radial.nr <= 0) throw Error(PFW_INVALID_ARGUMENT, "ACE model needs radial functions");
  a.cutoff = cutoff;
  std::vector<double> ref(m->reference, m->reference + m->nz);
  for (double v : ref)
    if (!std::isfinite(v)) throw Error(PFW_INVALID_ARGUMENT, "nonfinite reference energies");
  a.reference = upload(ref, st);
  // densities: each radial index must own one contiguous, complete m-block
  // (r, l*l .. l*l+2l), as the resolved schema's (n, l) channels do.
  std::vector<int32_t> dens(m->density, m->density + 2 * (size_t)m->ndens);
  std::vector<int> first(a.radial.nr, -1), lof(a.radial.nr, -1), cnt(a.radial.nr, 0);
  for (int k = 0; k < m->ndens; ++k) {
    int r = dens[2 * k], h = dens[2 * k + 1];
    if (r < 0 || r >= a.radial.nr || h < 0 || h >= a.ny) throw Error(PFW_INVALID_ARGUMENT, "invalid density indices");
    int l = 0;
    while ((l + 1) * (l + 1) <= h) ++l;
    if (first[r] < 0) {
      first[r] = k;
      lof[r] = l;
    }
    if (lof[r] != l || k != first[r] + cnt[r] || h != l * l + cnt[r])
      throw Error(PFW_INVALID_ARGUMENT, "densities of a radial channel must form one contiguous complete m-block");
    ++cnt[r];
  }
  std::vector<int32_t> rlb(a.lmax + 2, 0), rl;
  for (int l = 0; l <= a.lmax; ++l) {
    for (int r = 0; r < a.radial.nr; ++r)
      if (first[r] >= 0 && lof[r] == l) {
        if (cnt[r] != 2 * l + 1)
          throw Error(PFW_INVALID_ARGUMENT, "densities of a radial channel must form one contiguous complete m-block");
        rl.push_back(r);
        rl.push_back(first[r]);
      }
    rlb[l + 1] = (int32_t)(rl.size() / 2);
  }
  a.dens = upload(dens, st);
  a.h_dens = dens;
  ace_group_densities(ctx, ctx->tune[1]);
  a.rlb = upload(rlb, st);
  a.rl = upload(rl, st);
  // products (factor lists, -1 padded) and their conflict-free schedule
  if (m->ndens >= 0xffff) throw Error(PFW_INVALID_ARGUMENT, "more than 65534 densities are not supported on the device");
  std::vector<int32_t> prod(4 * (size_t)m->nprod, -1);
  std::vector<std::vector<int>> factors(m->nprod);
  for (int k = 0; k < m->nprod; ++k) {
    int b = m->product_ptr[k], nu = m->product_ptr[k + 1] - b;
    if (nu <= 0 || nu > 4) throw Error(PFW_INVALID_ARGUMENT, "product orders must be 1..4 on the device");
    for (int t = 0; t < nu; ++t) {
      int d = m->product_idx[b + t];
      if (d < 0 || d >= m->ndens) throw Error(PFW_INVALID_ARGUMENT, "invalid product indices");
      prod[4 * (size_t)k + t] = d;
      factors[k].push_back(d);
    }
  }
  a.prod = upload(prod, st);
  a.h_prod = prod;
  std::vector<int32_t> copyb;
  std::vector<ProductSlot> slots = schedule_products(factors, m->ndens, copyb, a.order_step);
  a.nsteps = (int)(slots.size() / kWave);
  a.ntarget = copyb[m->ndens];
  if (a.ntarget >= 0xffff) throw Error(PFW_INVALID_ARGUMENT, "product schedule needs too many accumulators");
  a.slot_prod.resize(slots.size());
  std::vector<uint4> sched(slots.size());
  auto u16 = [](int32_t v) { return (uint32_t)(v < 0 ? 0xffff : v); };
  for (size_t q = 0; q < slots.size(); ++q) {
    const ProductSlot& e = slots[q];
    a.slot_prod[q] = e.k;
    sched[q] = make_uint4(u16(e.f[0]) | (u16(e.f[1]) << 16), u16(e.f[2]) | (u16(e.f[3]) << 16),
                          u16(e.t[0]) | (u16(e.t[1]) << 16), u16(e.t[2]) | (u16(e.t[3]) << 16));
  }
  a.sched = upload(sched, st);
  a.copyb = upload(copyb, st);
  a.c_rows.assign(m->coupling_rows, m->coupling_rows + m->coupling_nnz);
  a.c_cols.assign(m->coupling_cols, m->coupling_cols + m->coupling_nnz);
  a.c_vals.assign(m->coupling_vals, m->coupling_vals + m->coupling_nnz);
  for (int64_t t = 0; t < m->coupling_nnz; ++t)
    if (a.c_rows[t] < 0 || a.c_rows[t] >= m->nbasis || a.c_cols[t] < 0 || a.c_cols[t] >= m->nprod ||
        !std::isfinite(a.c_vals[t]))
      throw Error(PFW_INVALID_ARGUMENT, "invalid coupling entries");
  ace_set_parameters(ctx, m->WB, m->Wpair);
  a.ready = true;
}

template <typename T>
static FamArgs<T> fam_args(const AceFamilyDev& f) {
  FamArgs<T> o;
  o.nr = f.nr;
  o.nnodes = f.nnodes;
  o.nint = f.nnodes - 1;
  o.uniform = f.uniform;
  o.pq = f.pq;
  o.env_kind = f.env_kind;
  o.env_pow = f.env_pow;
  if constexpr (sizeof(T) == 8) {
    o.par = f.par64;
    o.nodes = f.nodes64;
    o.coef = f.coef64;
  } else {
    o.par = f.par32;
    o.nodes = f.nodes32;
    o.coef = f.coef32;
  }
  return o;
}

static inline unsigned nblk(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

template <typename T>
static AceArgs<T> make_args(pfw_ctx* ctx, int grads) {
  AceDev& ad = ctx->ace;
  SystemBuffers& s = ctx->sys;
  AceArgs<T> a;
  a.n = s.n;
  a.nz = ad.nz;
  a.ndens = ad.ndens;
  a.nprod = ad.nprod;
  a.ny = ad.ny;
  a.perm = s.perm;
  a.sps = s.sps;
  a.ecount = ad.ecount;
  a.ecap = ad.ecap;
  a.enb = ad.enb;
  a.eR = ad.eR;
  a.reference = ad.reference;
  a.rad = fam_args<T>(ad.radial);
  a.pair = fam_args<T>(ad.pair);
  a.dens = ad.dens;
  a.dperm = ad.dperm;
  a.dchunk = ad.dchunk;
  a.rchunk = ad.rchunk;
  a.rlb = ad.rlb;
  a.rl = ad.rl;
  a.prod = ad.prod;
  if constexpr (sizeof(T) == 8) {
    a.w = ad.w64;
    a.wp = ad.wp64;
    a.ws = ad.ws64;
  } else {
    a.w = ad.w32;
    a.wp = ad.wp32;
    a.ws = ad.ws32;
  }
  a.grads = grads;
  a.siteE = ad.siteE;
  a.siteW = ad.siteW;
  a.eg = (T*)ad.eg;
  a.nsteps = ad.nsteps;
  a.sched = ad.sched;
  a.copyb = ad.copyb;
  a.ntarget = ad.ntarget;
  a.Abuf = (T*)ad.Abuf;
  a.solid = ad.solid;
  a.proj = ad.proj;
  a.dstride = ad.proj ? 2 * ad.ndens : ad.ndens;
  a.olo = ad.order_lo;
  a.ohi = ad.order_hi;
  a.sel = ad.sel;
  a.pair_gate = ad.pair_gate;
  for (int k = 0; k < 5; ++k) a.ostep[k] = ad.order_step[k];
  return a;
}

template <typename K>
static void allow_lds(K kernel, size_t lds, const char* what) {
  if (lds > kMaxLds) throw Error(PFW_INVALID_ARGUMENT, std::string("ACE model too large for the LDS tiles of ") + what);
  if (lds > 64 * 1024)
    PFW_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

// LDS bytes of k_ace_prod's tables (the kernel's layout) and per wavefront.
template <typename T>
static void prod_lds(const AceDev& ad, size_t& tables, size_t& per_wave) {
  const size_t nslot = (size_t)ad.nsteps * kWave;
  tables = align16(align16(sizeof(T) * ad.nz * nslot) + 16 * nslot + 4 * (size_t)(ad.ndens + 1));
  const int k = ad.proj ? 2 : 1;
  per_wave = align16(sizeof(T) * k * ad.ndens) + align16(sizeof(T) * k * (ad.ntarget + 1));
}

template <typename T, int L, int SPW>
static void launch_sites(pfw_ctx* ctx, int grads) {
  AceDev& ad = ctx->ace;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  AceArgs<T> a = make_args<T>(ctx, grads);
  constexpr int NY = (L + 1) * (L + 1);
  const unsigned blocks = (unsigned)((s.n + SPW - 1) / SPW);
  const size_t dstride = ad.proj ? 2 * (size_t)ad.ndens : (size_t)ad.ndens;
  const size_t lds_dens =
      sizeof(T) * ((size_t)NY + SPW * dstride + (size_t)(ad.rchunk + (ad.proj ? 2 : 1) * NY) * kStride);
  bool mfma = false;
  if constexpr (sizeof(T) == 4) {
    mfma = ad.mgroups > 0 && (ctx->tune[4] >= 2 || (ctx->tune[4] == 0 && ad.mdefault));
    if (mfma) {
      MfmaTiles t{ad.mtab, ad.mgroups, {ad.mgchunk[0], ad.mgchunk[1], ad.mgchunk[2], ad.mgchunk[3]}};
      const size_t lds_m = sizeof(float) * ((size_t)NY + (size_t)(ad.rchunk + 1 + NY + 1) * kStride);
      allow_lds(k_ace_dens_mfma<L, SPW>, lds_m, "the density kernel");
      hipLaunchKernelGGL((k_ace_dens_mfma<L, SPW>), dim3(blocks), dim3(kWave), lds_m, st, a, t);
    }
  }
  if (!mfma) {
    allow_lds(k_ace_dens<T, L, SPW>, lds_dens, "the density kernel");
    hipLaunchKernelGGL((k_ace_dens<T, L, SPW>), dim3(blocks), dim3(kWave), lds_dens, st, a);
  }
  PFW_HIP(hipGetLastError());
  size_t tables, per_wave;
  prod_lds<T>(ad, tables, per_wave);
  // tables in LDS when at least 4 wavefronts' accumulators fit beside them
  const bool global_tables = tables + 4 * per_wave > kMaxLds;
  if (global_tables) tables = 0;
  if (tables + per_wave > kMaxLds) throw Error(PFW_INVALID_ARGUMENT, "ACE product accumulators too large for LDS");
  int waves = (int)std::min<size_t>(16, (kMaxLds - tables) / per_wave);
  if (ctx->tune[2] > 0) waves = std::min(waves, ctx->tune[2]);
  const size_t lds_prod = tables + waves * per_wave;
  const int64_t want = (s.n + waves - 1) / waves;
  const unsigned pblocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, ctx->num_cu));
  if (global_tables) {
    allow_lds(k_ace_prod<T, true>, lds_prod, "the product kernel");
    hipLaunchKernelGGL((k_ace_prod<T, true>), dim3(pblocks), dim3(waves * kWave), lds_prod, st, a, grads);
  } else {
    allow_lds(k_ace_prod<T, false>, lds_prod, "the product kernel");
    hipLaunchKernelGGL((k_ace_prod<T, false>), dim3(pblocks), dim3(waves * kWave), lds_prod, st, a, grads);
  }
  PFW_HIP(hipGetLastError());
  if (grads) {
    const size_t lds_pull = sizeof(T) * ((size_t)NY + SPW * dstride);
    bool mpull = false;
    if constexpr (sizeof(T) == 4) {
      mpull = !ad.proj && (ctx->tune[4] == 3 || (ctx->tune[4] == 0 && kMfmaPullDefault));
      if (mpull) {
        allow_lds(k_ace_pull_mfma<L, SPW>, lds_pull, "the pullback kernel");
        hipLaunchKernelGGL((k_ace_pull_mfma<L, SPW>), dim3(blocks), dim3(kWave), lds_pull, st, a);
      }
    }
    if (!mpull) {
      allow_lds(k_ace_pull<T, L, SPW>, lds_pull, "the pullback kernel");
      hipLaunchKernelGGL((k_ace_pull<T, L, SPW>), dim3(blocks), dim3(kWave), lds_pull, st, a);
    }
    PFW_HIP(hipGetLastError());
  }
}

template <typename T, int L>
static void launch_sites_spw(pfw_ctx* ctx, int grads) {
  // two sites per wavefront unless some site has more than 48 edges: a site
  // above 32 takes a second pass in its half-wavefront, which costs less than
  // halving the sites per wavefront for every site
  const int spw = ctx->tune[0] > 0 ? ctx->tune[0] : ctx->ace.max_degree <= 48 ? 2 : 1;
  if (spw == 2)
    launch_sites<T, L, 2>(ctx, grads);
  else
    launch_sites<T, L, 1>(ctx, grads);
}

template <typename T>
static void dispatch_sites(pfw_ctx* ctx, int grads) {
  switch (ctx->ace.lmax <= 2 ? 2 : ctx->ace.lmax <= 4 ? 4 : 6) {
    case 2:
      launch_sites_spw<T, 2>(ctx, grads);
      break;
    case 4:
      launch_sites_spw<T, 4>(ctx, grads);
      break;
    default:
      launch_sites_spw<T, 6>(ctx, grads);
  }
}

// Directed environments of every site from the cell list built by nl_build,
// in rows of ecap slots; with `rev` also the reverse-edge index. The first
// build counts once to size ecap; later builds write directly and report a
// capacity overflow in status[2] (the caller repeats the step with the degree
// found in status[1]).
static void ace_rows_alloc(pfw_ctx* ctx, int ecap);

// The stencil size shared by every configuration's grid, or 0.
static int uniform_stencil(const pfw_ctx* ctx) {
  int ns = -1;
  for (const Grid& g : ctx->grids) {
    if (ns >= 0 && g.nstencil != ns) return 0;
    ns = g.nstencil;
  }
  return ns > 0 ? ns : 0;
}

// Words 1..3 of k_ace_edges's status from its stripes (host copy).
static void fold_status(int32_t* h) {
  for (int k = 0; k < kStatusStripes; ++k) {
    const int32_t* w = h + 4 + 3 * k;
    h[1] = std::max(h[1], w[0]);
    h[2] |= w[1];
    h[3] = std::max(h[3], w[2]);
  }
}

static void ace_environment(pfw_ctx* ctx, int rev) {
  AceDev& ad = ctx->ace;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  ace_rows_alloc(ctx, std::max(ad.ecap, 1));
  if (ad.ecap == 0) {  // size the rows: one counting pass
    PFW_HIP(hipMemsetAsync(ad.status, 0, kStatusWords * sizeof(int32_t), st));
    hipLaunchKernelGGL(k_ace_edges, dim3((unsigned)ctx->ncells), dim3(kWave * kEdgeWaves), 0, st, s.grids, s.cell_cfg, n, s.xs,
                       s.start, 0, 0, ad.ecount, (int32_t*)nullptr, (int32_t*)nullptr, (double*)nullptr, ad.status, 0,
                       (int32_t*)nullptr, (uint8_t*)nullptr, (int)(ctx->ncfg == 1));
    PFW_HIP(hipGetLastError());
    PFW_HIP(hipMemcpyAsync(ad.h_status, ad.status, kStatusWords * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PFW_HIP(hipStreamSynchronize(st));
    fold_status(ad.h_status);
    ad.max_degree = ad.h_status[1];
    ad.ecap = std::max(1, ad.max_degree + std::max(4, ad.max_degree / 8));
  }
  ace_rows_alloc(ctx, ad.ecap);
  const int64_t slots = n * (int64_t)ad.ecap;
  fill_i32(ad.status, kStatusWords, 0, st);  // a fill kernel: this sequence may be captured in a graph
  hipLaunchKernelGGL(k_ace_edges, dim3((unsigned)ctx->ncells), dim3(kWave * kEdgeWaves), 0, st, s.grids, s.cell_cfg, n, s.xs,
                     s.start, 1, ad.ecap, ad.ecount, ad.enb, ad.eent, ad.eR, ad.status, 0, rev ? ad.erev : nullptr,
                     rev ? ad.eseg : nullptr, (int)(ctx->ncfg == 1));
  PFW_HIP(hipGetLastError());
  if (rev)
    hipLaunchKernelGGL(k_ace_rev, dim3(nblk(slots, 256)), dim3(256), 0, st, slots, ad.ecap, s.grids, s.atom_cfg,
                       s.perm, ad.ecount, ad.enb, ad.eent, ad.eR, ad.erev, ad.status, (const uint8_t*)ad.eseg,
                       uniform_stencil(ctx));
  PFW_HIP(hipMemcpyAsync(ad.h_status, ad.status, kStatusWords * sizeof(int32_t), hipMemcpyDeviceToHost, st));
}

// After the step's synchronisation: true when some site had more edges than
// its row holds; ecap then grows to the degree found, and the step is repeated.
static bool ace_overflowed(AceDev& ad) {
  fold_status(ad.h_status);
  ad.max_degree = ad.h_status[1];
  if (!ad.h_status[2]) return false;
  ad.ecap = ad.max_degree + std::max(4, ad.max_degree / 8);
  return true;
}

// Sites, force gather and reductions over the current rows (ecount, enb, eR,
// erev), ending with the output copies and the stream synchronisation.
static void ace_evaluate_rows(pfw_ctx* ctx, int grads, int precision) {
  AceDev& ad = ctx->ace;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  if (precision == PFW_MIXED)
    dispatch_sites<float>(ctx, grads);
  else
    dispatch_sites<double>(ctx, grads);
  record_stage(ctx, 5);
  if (grads && precision == PFW_MIXED)
    hipLaunchKernelGGL(k_ace_force<float>, dim3((unsigned)n), dim3(kWave), 0, st, n, ad.ecap, ad.ecount, ad.erev,
                       (const float*)ad.eg, s.perm, ctx->d_F);
  else if (grads)
    hipLaunchKernelGGL(k_ace_force<double>, dim3((unsigned)n), dim3(kWave), 0, st, n, ad.ecap, ad.ecount, ad.erev,
                       (const double*)ad.eg, s.perm, ctx->d_F);
  hipLaunchKernelGGL(k_ace_pack, dim3(nblk(n, 256)), dim3(256), 0, st, n, ad.siteE, ad.siteW, grads, ad.rec);
  if (ctx->ncfg == 1) {
    reduce_outputs(ctx, ad.rec, n, 10, 1, ad.status);
  } else {
    batch_reduce(ctx, ad.rec, 10, false, 1);
  }
  PFW_HIP(hipGetLastError());
}

// Buffers of the per-step rows, sized for n sites of `ecap` slots.
static void ace_rows_alloc(pfw_ctx* ctx, int ecap) {
  AceDev& ad = ctx->ace;
  const int64_t n = ctx->sys.n;
  if (!ad.ecount) {
    PFW_HIP(hipMalloc(&ad.ecount, n * sizeof(int32_t)));
    PFW_HIP(hipMalloc(&ad.siteE, n * sizeof(double)));
    PFW_HIP(hipMalloc(&ad.Abuf, std::max<size_t>(1, (size_t)n * ad.ndens * (ad.proj ? 2 : 1)) * sizeof(double)));
    PFW_HIP(hipMalloc(&ad.siteW, 9 * n * sizeof(double)));
    PFW_HIP(hipMalloc(&ad.rec, 10 * n * sizeof(double)));
    PFW_HIP(hipMalloc(&ad.status, kStatusWords * sizeof(int32_t)));
    PFW_HIP(hipHostMalloc(&ad.h_status, kStatusWords * sizeof(int32_t), hipHostMallocDefault));
    PFW_HIP(hipMalloc(&ad.eseg, std::max<int64_t>(n, 1) * kSegEntries));
  }
  const int64_t slots = n * (int64_t)ecap;
  if (slots > ad.edge_cap || !ad.enb) {
    dfree(ad.enb);
    dfree(ad.eR);
    dfree(ad.erev);
    dfree(ad.eent);
    dfree(ad.eg);
    dfree(ad.corig);
    PFW_HIP(hipMalloc(&ad.enb, slots * sizeof(int32_t)));
    PFW_HIP(hipMalloc(&ad.eR, 3 * slots * sizeof(double)));
    PFW_HIP(hipMalloc(&ad.erev, slots * sizeof(int32_t)));
    PFW_HIP(hipMalloc(&ad.eent, slots * sizeof(int32_t)));
    PFW_HIP(hipMalloc(&ad.eg, 3 * slots * sizeof(double)));
    PFW_HIP(hipMalloc(&ad.corig, slots * sizeof(int32_t)));
    ad.edge_cap = slots;
  }
}

// Skin rows at cutoff + skin from a fresh cell build, with their reverse
// index (gpu_build_cells! + gpu_connectivity!, src/submodules/NeighbourList/gpu.jl:497-541).
static void ace_skin_rebuild(pfw_ctx* ctx) {
  AceDev& ad = ctx->ace;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  nl_build(ctx);
  record_stage(ctx, 3);
  ace_rows_alloc(ctx, std::max(ad.ecap, 1));
  if (!ad.ecount_s) {
    PFW_HIP(hipMalloc(&ad.ecount_s, n * sizeof(int32_t)));
    PFW_HIP(hipMalloc(&ad.eseg_s, std::max<int64_t>(n, 1) * kSegEntries));
    PFW_HIP(hipMalloc(&ad.Xref, 3 * n * sizeof(double)));
    PFW_HIP(hipMalloc(&ad.dX, 3 * n * sizeof(double)));
    PFW_HIP(hipMalloc(&ad.maxd2, 2 * sizeof(unsigned long long)));
    PFW_HIP(hipHostMalloc(&ad.h_maxd2, 2 * sizeof(unsigned long long), hipHostMallocDefault));
    PFW_HIP(hipMemsetAsync(ad.maxd2, 0, 2 * sizeof(unsigned long long), st));
    ad.parity = 0;
  }
  auto build = [&](int pass) {
    PFW_HIP(hipMemsetAsync(ad.status, 0, kStatusWords * sizeof(int32_t), st));
    hipLaunchKernelGGL(k_ace_edges, dim3((unsigned)ctx->ncells), dim3(kWave * kEdgeWaves), 0, st, s.grids, s.cell_cfg, n, s.xs,
                       s.start, pass, pass ? ad.ecap_s : 0, ad.ecount_s, ad.enb_s, ad.eent_s, ad.eR_s, ad.status, 1,
                       pass ? ad.erev_s : nullptr, pass ? ad.eseg_s : nullptr, (int)(ctx->ncfg == 1));
    PFW_HIP(hipGetLastError());
    if (pass)
      hipLaunchKernelGGL(k_ace_rev, dim3(nblk(n * (int64_t)ad.ecap_s, 256)), dim3(256), 0, st, n * (int64_t)ad.ecap_s,
                         ad.ecap_s, s.grids, s.atom_cfg, s.perm, ad.ecount_s, ad.enb_s, ad.eent_s, ad.eR_s, ad.erev_s,
                         ad.status, (const uint8_t*)ad.eseg_s, uniform_stencil(ctx));
    PFW_HIP(hipGetLastError());
    PFW_HIP(hipMemcpyAsync(ad.h_status, ad.status, kStatusWords * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PFW_HIP(hipStreamSynchronize(st));
    fold_status(ad.h_status);
  };
  auto size_rows = [&](int degree) {
    ad.ecap_s = std::max(ad.ecap_s, degree + std::max(4, degree / 8));
    const int64_t slots = n * (int64_t)ad.ecap_s;
    if (slots > ad.skin_cap) {
      dfree(ad.enb_s);
      dfree(ad.eent_s);
      dfree(ad.erev_s);
      dfree(ad.eR_s);
      dfree(ad.cslot);
      PFW_HIP(hipMalloc(&ad.enb_s, slots * sizeof(int32_t)));
      PFW_HIP(hipMalloc(&ad.eent_s, slots * sizeof(int32_t)));
      PFW_HIP(hipMalloc(&ad.erev_s, slots * sizeof(int32_t)));
      PFW_HIP(hipMalloc(&ad.eR_s, 3 * slots * sizeof(double)));
      PFW_HIP(hipMalloc(&ad.cslot, slots * sizeof(int32_t)));
      ad.skin_cap = slots;
    }
  };
  if (ad.ecap_s == 0) {  // first build: size the rows
    build(0);
    size_rows(ad.h_status[1]);
  }
  build(1);
  if (ad.h_status[2]) {  // a row outgrew its capacity since the last build
    size_rows(ad.h_status[1]);
    build(1);
  }
  if (ad.h_status[2] || ad.h_status[0]) throw Error(PFW_DEVICE, "ACE skin rows could not be built consistently");
  ad.skin_max_degree = ad.h_status[1];
  ad.compact_degree = ad.h_status[3];
  PFW_HIP(hipMemcpyAsync(ad.Xref, s.X, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  ace_rows_alloc(ctx, ad.ecap_s);
  ad.skin_valid = true;
  ad.skin_gen = ctx->build_gen;
  ++ctx->verlet.rebuilds;
}

// One ACE step on the skin rows: displacements, compaction to the edges
// inside the cutoff now, then the usual sites, force gather and reductions.
// As for the pair list, the step is evaluated on the current rows and
// repeated after a rebuild when the displacement read back with the outputs
// shows that the rows may miss an edge.
static void ace_skin_step(pfw_ctx* ctx, int grads, int precision) {
  AceDev& ad = ctx->ace;
  VerletDev& V = ctx->verlet;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  bool fresh = false;
  if (!ad.skin_valid || ad.skin_gen != ctx->build_gen || !ad.ecount_s) {
    ace_skin_rebuild(ctx);
    fresh = true;
  } else {
    for (int k = 0; k < 4; ++k) record_stage(ctx, k);
  }
  for (;;) {
    ad.ecap = ad.ecap_s;
    // sites per wavefront follow the largest in-cutoff degree at the build (any
    // degree is correct; a smaller one only packs two sites per wavefront)
    ad.max_degree = ad.compact_degree;
    PFW_HIP(hipMemsetAsync(ad.status, 0, kStatusWords * sizeof(int32_t), st));
    hipLaunchKernelGGL(k_ace_delta, dim3(nblk(n, 256)), dim3(256), 0, st, n, s.perm, s.X, ad.Xref, ad.dX, ad.maxd2,
                       ad.parity);
    hipLaunchKernelGGL(k_ace_compact, dim3((unsigned)n), dim3(kWave), 0, st, n, ad.ecap, ad.ecount_s, ad.enb_s,
                       ad.eR_s, ad.dX, ad.cutoff * ad.cutoff, ad.ecount, ad.enb, ad.eR, ad.cslot, ad.corig,
                       s.perm, s.grids, s.atom_cfg, s.X);
    if (grads)
      hipLaunchKernelGGL(k_ace_revmap, dim3(nblk(n * (int64_t)ad.ecap, 256)), dim3(256), 0, st, n * (int64_t)ad.ecap,
                         ad.ecap, ad.ecount, ad.corig, ad.erev_s, ad.cslot, ad.erev, ad.status);
    PFW_HIP(hipGetLastError());
    record_stage(ctx, 4);
    ace_evaluate_rows(ctx, grads, precision);
    copy_outputs(ctx);
    PFW_HIP(hipMemcpyAsync(ad.h_maxd2, ad.maxd2 + ad.parity, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    record_stage(ctx, 6);
    PFW_HIP(hipStreamSynchronize(st));
    ad.parity ^= 1;
    const double m2 = __builtin_bit_cast(double, (unsigned long long)ad.h_maxd2[0]);
    V.last_max_disp = std::sqrt(m2);
    const bool stale = !(2.0 * V.last_max_disp < V.skin * (1.0 - 16.0 * 2.220446049250313e-16));
    if (fresh || !stale) break;
    ace_skin_rebuild(ctx);
    fresh = true;
  }
  if (!fresh) ++V.reuses;
}

void ace_run(pfw_ctx* ctx, uint32_t mask, int precision) {
  ctx->host_written = false;
  ctx->audit_valid = false;
  AceDev& ad = ctx->ace;
  if (!ad.ready) throw Error(PFW_STALE, "no ACE model set");
  if (ad.cutoff != ctx->cutoff)
    throw Error(PFW_STALE, "the system's cell list was built for another cutoff than the ACE search cutoff");
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  const int grads = (mask & (PFW_FORCES | PFW_VIRIAL)) ? 1 : 0;
  if (ctx->verlet.skin > 0 && n > 0 && mask != 0) {
    ace_skin_step(ctx, grads, precision);
  } else if (ctx->graphs && ctx->graph_warm && !ctx->timing && !ctx->audit && n > 0 && mask != 0 && ad.ecap > 0) {
    struct Launch {
      int grads, precision;
    } arg{grads, precision};
    graph_issue(ctx, 1, precision, mask, [](pfw_ctx* c, void* p) {
      const Launch& a = *(const Launch*)p;
      nl_build(c);
      ace_environment(c, a.grads);
      ace_evaluate_rows(c, a.grads, a.precision);
      copy_outputs(c);
    }, &arg);
    PFW_HIP(hipStreamSynchronize(st));
    if (ace_overflowed(ad)) {  // rows outgrew their capacity: drop the graph, repeat eagerly with wider rows
      graph_reset(ctx);
      ace_run(ctx, mask, precision);
      return;
    }
  } else {
    nl_build(ctx);
    record_stage(ctx, 3);
    if (n == 0 || mask == 0) {
      record_stage(ctx, 4);
      record_stage(ctx, 5);
      PFW_HIP(hipMemsetAsync(ctx->d_out, 0, 11 * sizeof(double), st));
      PFW_HIP(hipMemsetAsync(ctx->d_cfg_out, 0, 11 * (size_t)ctx->ncfg * sizeof(double), st));
      if (n > 0) PFW_HIP(hipMemsetAsync(ctx->d_F, 0, 3 * n * sizeof(double), st));
    } else {
      ace_environment(ctx, grads);
      record_stage(ctx, 4);
      ace_evaluate_rows(ctx, grads, precision);
    }
    copy_outputs(ctx);
    record_stage(ctx, 6);
    PFW_HIP(hipStreamSynchronize(st));
    if (n > 0 && mask != 0 && ace_overflowed(ad)) {  // an environment outgrew its row: once more, wider
      ace_run(ctx, mask, precision);
      return;
    }
  }
  if (ctx->audit) {
    const AuditBufs au = audit_begin(ctx);
    if (n > 0 && mask != 0) {
      const int64_t slots = n * (int64_t)ad.ecap;
      hipLaunchKernelGGL(k_ace_audit, dim3(nblk(slots, 256)), dim3(256), 0, st, slots, ad.ecap, ad.ecount, ad.enb,
                         ad.eR, s.perm, s.grids, s.atom_cfg, au);
      PFW_HIP(hipGetLastError());
    }
    PFW_HIP(hipStreamSynchronize(st));
  }
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
  if (ctx->h_out[10] != 0.0) throw Error(PFW_DOMAIN, "nonfinite ACE output (coincident atoms or singular envelope)");
  ctx->graph_warm = true;
}

// ---------------------------------------------------------------- equivariant features

// Per-atom equivariant features (ace_features, src/submodules/ACE/features.jl:141-178)
// of a layer with smooth exponential radials (radials.jl:193-206). One
// wavefront per atom: its edges, 64 at a time, put their radials, harmonics
// and neighbour species in LDS; each lane then sums, for its densities, the
// chunk's edges in order (species-resolved channel 0 = any species); then the
// products, the coupling rows and the mixed outputs, lanes over entries.
// One coupling nonzero: its value and the density indices of its product's
// factors, 16 bits each, unused factors pointing at the constant 1 that
// follows the densities (products of at most four factors).
struct FeatEntry {
  double val;
  uint64_t f;
};

struct FeatArgs {
  int64_t n;
  const int64_t* off;  // n + 1 edge offsets per sorted atom
  const int32_t* nbr;  // neighbour (original index) of each edge
  const double* R;     // 3 per edge: x_j + H s - x_i
  const int32_t* perm;
  const int32_t* species;  // original order, 0..nz-1
  int nz, nr, ndens, nrows, nblocks;
  int64_t nnz;
  const int32_t *dz, *dn, *dh;
  const int32_t* rptr;      // coupling row offsets (nrows + 1)
  const FeatEntry* ent;     // coupling entries in row order, each with its product's factors
  const int32_t* erow;      // row of each entry
  const int64_t* boff;  // per block: row offset, mixing offset, output offset, dim, nbasis, channels
  const double* mix;
  double rc, rate, amplitude;
  const int32_t* powers;
  const double* scale;  // nr x nz x nz (radial, centre, neighbour), column-major
  double* out;
  int32_t* status;  // bits: 1 two atoms coincide, 2 an edge without its reverse, 4 a nonfinite cotangent
};

// The densities A of one atom (edges [e0, e1), centre species zi) into LDS;
// Ysh (NY x 64), Rsh (nr x 64) and Zsh (64) are the chunk staging.
template <int L>
__device__ inline void feat_densities(const FeatArgs& a, int64_t e0, int64_t e1, int zi, const double* Fl,
                                      double* Ysh, double* Rsh, int* Zsh, double* A) {
  const int lane = threadIdx.x;
  for (int64_t c0 = e0; c0 < e1; c0 += kWave) {
    const int64_t e = c0 + lane;
    int zj = -1;
    if (e < e1) {
      const double dx = a.R[3 * e], dy = a.R[3 * e + 1], dz = a.R[3 * e + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r > 0) {
        zj = a.species[a.nbr[e]];
        const double x = r / a.rc, q = fmax(1.0 - x, 0.0), q2 = q * q;
        const double base = a.amplitude * q2 * q2 * exp(-a.rate * r);
        for (int nn = 0; nn < a.nr; ++nn) {
          double xn = 1;
          for (int p = 0; p < a.powers[nn]; ++p) xn *= x;
          Rsh[nn * 64 + lane] = base * xn * a.scale[nn + a.nr * (zi + a.nz * zj)];
        }
        ylm_emit<double, L>(dx, dy, dz, Fl, 0, [&](int h, const Jet<double>& jv) { Ysh[h * 64 + lane] = jv.v; });
      } else {
        atomicOr(a.status, 1);  // coincident atoms: undefined harmonics
      }
    }
    Zsh[lane] = zj;
    __syncthreads();
    const int ne = (int)min((int64_t)kWave, e1 - c0);
    for (int t = lane; t < a.ndens; t += kWave) {
      const int zc = a.dz[t], rn = a.dn[t], h = a.dh[t];
      double sum = 0;
      for (int u = 0; u < ne; ++u) {
        const int zu = Zsh[u];
        if (zu >= 0 && (zc == 0 || zc == zu + 1)) sum += Rsh[rn * 64 + u] * Ysh[h * 64 + u];
      }
      A[t] += sum;
    }
    __syncthreads();
  }
}

// Coupled rows B = C P of one atom, each product P_k = prod A formed where its
// coupling entry is read (A[ndens] = 1 pads the factors).
__device__ inline void feat_rows(const FeatArgs& a, const double* A, double* B) {
  const int lane = threadIdx.x;
  for (int t = lane; t < a.nrows; t += kWave) {
    double v = 0;
#pragma unroll 4
    for (int q = a.rptr[t]; q < a.rptr[t + 1]; ++q) {
      const FeatEntry e = a.ent[q];
      const double p = A[e.f & 0xffff] * A[(e.f >> 16) & 0xffff] * A[(e.f >> 32) & 0xffff] * A[e.f >> 48];
      v += e.val * p;
    }
    B[t] = v;
  }
  __syncthreads();
}

template <int L>
__global__ __launch_bounds__(64) void k_features(FeatArgs a) {
  constexpr int NY = (L + 1) * (L + 1);
  extern __shared__ __align__(16) unsigned char fsm[];
  double* Fl = (double*)fsm;   // (L+1)^2
  double* Ysh = Fl + NY;       // NY x 64
  double* Rsh = Ysh + NY * 64;  // nr x 64
  double* A = Rsh + a.nr * 64;  // ndens + 1 (the last is 1)
  double* B = A + a.ndens + 1;  // nrows
  int* Zsh = (int*)(B + a.nrows);  // 64
  const int lane = threadIdx.x;
  const int64_t k = blockIdx.x;
  const int i = a.perm[k], zi = a.species[i];
  fill_fl<double, L>(Fl);
  for (int t = lane; t <= a.ndens; t += kWave) A[t] = t < a.ndens ? 0.0 : 1.0;
  __syncthreads();
  feat_densities<L>(a, a.off[k], a.off[k + 1], zi, Fl, Ysh, Rsh, Zsh, A);
  feat_rows(a, A, B);
  for (int b = 0; b < a.nblocks; ++b) {
    const int64_t* o = a.boff + 6 * b;
    const int dim = (int)o[3], nb = (int)o[4], ch = (int)o[5];
    const int nout = dim * ch;
    int P = 1;  // lanes per output (a power of two), summing every P-th basis function
    while (P < 8 && nout * P * 2 <= kWave) P *= 2;
    for (int t0 = 0; t0 < nout * P; t0 += kWave) {
      const int t = t0 + lane, oi = t / P, part = t % P;
      double v = 0;
      if (oi < nout) {
        const int m = oi % dim, c = oi / dim;
        const double* W = a.mix + o[1] + (int64_t)nb * (c + (int64_t)ch * zi);
#pragma unroll 4
        for (int beta = part; beta < nb; beta += P) v += B[o[0] + beta * dim + m] * W[beta];
      }
      for (int off = 1; off < P; off <<= 1) v += __shfl_xor(v, off, kWave);  // fixed tree over the P lanes
      if (oi < nout && part == 0) {
        const int m = oi % dim, c = oi / dim;
        a.out[o[2] + m + (int64_t)dim * (c + (int64_t)ch * i)] = v;
      }
    }
  }
}

// The pullback of the features (feature_vjp / feature_local_reverse,
// features.jl:204-255) for one atom per wavefront: the forward densities and
// rows again; the mixing partial B^T G; dB = G W^T over the rows; then dA by
// the product rule, lanes over the coupling entries, each adding
// val dB[row] dP/dA_f to its own column of a per-lane table that is summed
// over lanes in a fixed order; then per edge, lane per edge,
// g = sum_a dA_a (R'_n Y_m r^ + R_n grad Y_m), the densities of each harmonic
// visited as ylm_emit forms it. Per atom it writes one parameter row (mixing
// partials, radial scale partials per neighbour species, rate and amplitude
// partials); g goes to the edge.
struct FeatVjpArgs {
  const double* cot;  // the cotangents, laid out as the forward output
  const int32_t *hptr, *hidx;  // harmonic -> densities
  double* gE;    // 3 per edge
  double* prow;  // n x wp, original atom order
  int wp, so;    // row width; offset of the scale partials (after the mixing partials)
  int gper;      // cotangents per atom (sum over blocks of dim x channels)
};

constexpr int kLaneStride = 65;  // per-lane dA table: column stride (odd: no bank conflicts in the sums)

template <int L>
__global__ __launch_bounds__(64) void k_features_vjp(FeatArgs a, FeatVjpArgs v) {
  constexpr int NY = (L + 1) * (L + 1);
  extern __shared__ __align__(16) unsigned char fsm[];
  double* Fl = (double*)fsm;          // (L+1)^2
  double* Ysh = Fl + NY;              // NY x 64
  double* Rsh = Ysh + NY * 64;        // nr x 64: scaled radials
  double* Ush = Rsh + a.nr * 64;      // nr x 64: radials at unit amplitude and scale
  double* Tsh = Ush + a.nr * 64;      // nr x 64: sum of dA Y per radial
  double* A = Tsh + a.nr * 64;        // ndens + 1 (the last is 1)
  double* dA = A + a.ndens + 1;       // ndens
  double* B = dA + a.ndens;           // nrows, then dB
  double* dAl = B + a.nrows;          // (ndens + 1) x kLaneStride
  double* Gsc = dAl + (a.ndens + 1) * kLaneStride;  // nr x nz: scale partials (radial, neighbour species)
  double* Gl = Gsc + a.nr * a.nz;     // the atom's cotangents (sum over blocks of dim x channels)
  int* Zsh = (int*)(Gl + v.gper);     // 64
  const int lane = threadIdx.x;
  const int64_t k = blockIdx.x;
  const int i = a.perm[k], zi = a.species[i];
  const int64_t e0 = a.off[k], e1 = a.off[k + 1];
  fill_fl<double, L>(Fl);
  for (int t = lane; t <= a.ndens; t += kWave) A[t] = t < a.ndens ? 0.0 : 1.0;
  for (int t = lane; t < a.nr * a.nz; t += kWave) Gsc[t] = 0;
  for (int t = 0; t <= a.ndens; ++t) dAl[t * kLaneStride + lane] = 0;
  for (int b = 0, g0 = 0; b < a.nblocks; ++b) {
    const int64_t* o = a.boff + 6 * b;
    const int w = (int)(o[3] * o[5]);
    for (int t = lane; t < w; t += kWave) {
      const double c = v.cot[o[2] + (int64_t)w * i + t];
      if (!isfinite(c)) atomicOr(a.status, 4);  // nonfinite cotangent
      Gl[g0 + t] = c;
    }
    g0 += w;
  }
  __syncthreads();
  feat_densities<L>(a, e0, e1, zi, Fl, Ysh, Rsh, Zsh, A);
  feat_rows(a, A, B);
  double* row = v.prow + (int64_t)i * v.wp;
  int64_t mo = 0;
  for (int b = 0, g0 = 0; b < a.nblocks; ++b) {  // mixing partials B^T G
    const int64_t* o = a.boff + 6 * b;
    const int dim = (int)o[3], nb = (int)o[4], ch = (int)o[5];
    const double* G = Gl + g0;
    g0 += dim * ch;
    for (int t = lane; t < nb * ch; t += kWave) {
      const int beta = t % nb, c = t / nb;
      double s = 0;
      for (int m = 0; m < dim; ++m) s += B[o[0] + beta * dim + m] * G[m + dim * c];
      row[mo + t] = s;
    }
    mo += (int64_t)nb * ch;
  }
  __syncthreads();
  for (int b = 0, g0 = 0; b < a.nblocks; ++b) {  // dB = G W^T, over B
    const int64_t* o = a.boff + 6 * b;
    const int dim = (int)o[3], nb = (int)o[4], ch = (int)o[5];
    const double* G = Gl + g0;
    g0 += dim * ch;
    const double* W = a.mix + o[1] + (int64_t)nb * ch * zi;
    for (int t = lane; t < dim * nb; t += kWave) {
      const int m = t % dim, beta = t / dim;
      double s = 0;
      for (int c = 0; c < ch; ++c) s += G[m + dim * c] * W[beta + nb * c];
      B[o[0] + t] = s;
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int64_t q = lane; q < a.nnz; q += kWave) {  // dA by the product rule, per lane
    const FeatEntry e = a.ent[q];
    const double c = e.val * B[a.erow[q]];
    const int f0 = (int)(e.f & 0xffff), f1 = (int)((e.f >> 16) & 0xffff), f2 = (int)((e.f >> 32) & 0xffff),
              f3 = (int)(e.f >> 48);
    const double a0 = A[f0], a1 = A[f1], a2 = A[f2], a3 = A[f3];
    const double p01 = a0 * a1, p23 = a2 * a3;
    dAl[f0 * kLaneStride + lane] += c * (a1 * p23);
    dAl[f1 * kLaneStride + lane] += c * (a0 * p23);
    dAl[f2 * kLaneStride + lane] += c * (p01 * a3);
    dAl[f3 * kLaneStride + lane] += c * (p01 * a2);
  }
  __syncthreads();
  for (int t = lane; t < a.ndens; t += kWave) {
    double s = 0;
    for (int u = 0; u < kWave; ++u) s += dAl[t * kLaneStride + u];
    dA[t] = s;
  }
  __syncthreads();
  double grate = 0, gamp = 0;
  for (int64_t c0 = e0; c0 < e1; c0 += kWave) {
    const int64_t e = c0 + lane;
    int zj = -1;
    for (int nn = 0; nn < a.nr; ++nn) Tsh[nn * 64 + lane] = 0;
    if (e < e1) {
      const double dx = a.R[3 * e], dy = a.R[3 * e + 1], dz = a.R[3 * e + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      double gx = 0, gy = 0, gz = 0;
      if (r > 0) {
        zj = a.species[a.nbr[e]];
        const double* sc = a.scale + a.nr * (zi + a.nz * zj);
        const double x = r / a.rc, q = fmax(1.0 - x, 0.0), q2 = q * q;
        const double ex = exp(-a.rate * r), base = q2 * q2 * ex;
        for (int nn = 0; nn < a.nr; ++nn) {
          double xn = 1;
          for (int p = 0; p < a.powers[nn]; ++p) xn *= x;
          Ush[nn * 64 + lane] = base * xn;
          Rsh[nn * 64 + lane] = a.amplitude * base * xn * sc[nn];
        }
        ylm_emit<double, L>(dx, dy, dz, Fl, 0, [&](int h, const Jet<double>& jv) {
          for (int u = v.hptr[h]; u < v.hptr[h + 1]; ++u) {
            const int t = v.hidx[u], zc = a.dz[t];
            if (zc == 0 || zc == zj + 1) {
              const int rn = a.dn[t];
              const double w = dA[t];
              Tsh[rn * 64 + lane] += w * jv.v;
              const double cf = w * Rsh[rn * 64 + lane];
              gx += cf * jv.g0;
              gy += cf * jv.g1;
              gz += cf * jv.g2;
            }
          }
        });
        const double ri = 1.0 / r, ux = dx * ri, uy = dy * ri, uz = dz * ri;
        double gr = 0;
        for (int nn = 0; nn < a.nr; ++nn) {
          const int pw = a.powers[nn];
          double xn = 1, xm = 0;  // x^p and p x^(p-1)
          for (int p = 0; p < pw; ++p) {
            xm = xm * x + xn;
            xn *= x;
          }
          const double T = Tsh[nn * 64 + lane], U = Ush[nn * 64 + lane];
          const double dU = ex * ((xm * q - 4.0 * xn) * q2 * q / a.rc - a.rate * xn * q2 * q2);
          gr += a.amplitude * sc[nn] * dU * T;
          grate -= sc[nn] * T * r * a.amplitude * U;
          gamp += sc[nn] * T * U;
        }
        gx += gr * ux;
        gy += gr * uy;
        gz += gr * uz;
      }
      v.gE[3 * e] = gx;
      v.gE[3 * e + 1] = gy;
      v.gE[3 * e + 2] = gz;
    }
    Zsh[lane] = zj;
    __syncthreads();
    const int ne = (int)min((int64_t)kWave, e1 - c0);
    for (int t = lane; t < a.nr * a.nz; t += kWave) {  // scale partials by neighbour species
      const int nn = t % a.nr, z = t / a.nr;
      double s = 0;
      for (int u = 0; u < ne; ++u)
        if (Zsh[u] == z) s += Tsh[nn * 64 + u] * a.amplitude * Ush[nn * 64 + u];
      Gsc[t] += s;
    }
    __syncthreads();
  }
  for (int t = lane; t < a.nr * a.nz; t += kWave) row[v.so + t] = Gsc[t];
  grate = wave_sum(grate);
  gamp = wave_sum(gamp);
  if (lane == 0) {
    row[v.so + a.nr * a.nz] = grate;
    row[v.so + a.nr * a.nz + 1] = gamp;
  }
}

// Position gradient of each atom (original order) from the edge vectors g:
// dx_i = sum over its edges e of (g of the reverse edge - g_e), the reverse
// of (i -> j, S) being (j -> i, -S) in j's edges; one wavefront per sorted
// atom, lane per edge, fixed summation order. An edge whose reverse is absent
// (a pair on the cutoff sphere to rounding) adds its g to `orphan` atomically
// and sets status bit 2.
__global__ __launch_bounds__(64) void k_features_dx(int64_t n, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ nbr, const int32_t* __restrict__ S,
                                                    const int32_t* __restrict__ perm, const int32_t* __restrict__ inv,
                                                    const double* __restrict__ gE, double* __restrict__ dx,
                                                    double* __restrict__ orphan, int32_t* __restrict__ status) {
  const int64_t k = blockIdx.x;
  const int lane = threadIdx.x;
  const int i = perm[k];
  double sx = 0, sy = 0, sz = 0;
  for (int64_t e = off[k] + lane; e < off[k + 1]; e += kWave) {
    sx -= gE[3 * e];
    sy -= gE[3 * e + 1];
    sz -= gE[3 * e + 2];
    const int j = nbr[e];
    const int s0 = -S[3 * e], s1 = -S[3 * e + 1], s2 = -S[3 * e + 2];
    const int kj = inv[j];
    bool found = false;
    for (int64_t f = off[kj]; f < off[kj + 1]; ++f)
      if (nbr[f] == i && S[3 * f] == s0 && S[3 * f + 1] == s1 && S[3 * f + 2] == s2) {
        sx += gE[3 * f];
        sy += gE[3 * f + 1];
        sz += gE[3 * f + 2];
        found = true;
        break;
      }
    if (!found) {
      atomicAdd(&orphan[3 * (int64_t)j], gE[3 * e]);
      atomicAdd(&orphan[3 * (int64_t)j + 1], gE[3 * e + 1]);
      atomicAdd(&orphan[3 * (int64_t)j + 2], gE[3 * e + 2]);
      atomicOr(status, 2);
    }
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  sz = wave_sum(sz);
  if (lane == 0) {
    dx[3 * (int64_t)i] = sx;
    dx[3 * (int64_t)i + 1] = sy;
    dx[3 * (int64_t)i + 2] = sz;
  }
}

__global__ void k_inverse_perm(int64_t n, const int32_t* __restrict__ perm, int32_t* __restrict__ inv) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) inv[perm[k]] = (int32_t)k;
}

// Per-species column sums of the per-atom parameter rows: chunk c of
// kRowChunk atoms gives part[(c nz + z) wp + col]; k_species_total sums the
// chunks in order. Both fixed order.
constexpr int kRowChunk = 128;
__global__ void k_species_rows(int64_t n, int wp, int nz, const double* __restrict__ prow,
                               const int32_t* __restrict__ species, double* __restrict__ part) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= wp) return;
  const int64_t c = blockIdx.y, a0 = c * kRowChunk, a1 = min(n, a0 + kRowChunk);
  constexpr int kZ = 8;  // species per pass over the rows
  for (int z0 = 0; z0 < nz; z0 += kZ) {
    double s[kZ];
#pragma unroll
    for (int u = 0; u < kZ; ++u) s[u] = 0;
#pragma unroll 4
    for (int64_t t = a0; t < a1; ++t) {
      const int z = species[t] - z0;
      const double x = prow[t * wp + col];
#pragma unroll
      for (int u = 0; u < kZ; ++u)
        if (z == u) s[u] += x;
    }
    for (int u = 0; u < kZ && z0 + u < nz; ++u) part[(c * nz + z0 + u) * wp + col] = s[u];
  }
}

__global__ void k_species_total(int64_t nchunks, int64_t width, const double* __restrict__ part,
                                double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= width) return;
  double s = 0;
#pragma unroll 8
  for (int64_t c = 0; c < nchunks; ++c) s += part[c * width + t];
  out[t] = s;
}

template <typename T>
static T* upload_raw(const T* p, size_t count, std::vector<void*>& owned) {
  T* d = nullptr;
  PFW_HIP(hipMalloc((void**)&d, std::max<size_t>(count, 1) * sizeof(T)));
  owned.push_back(d);
  if (count) PFW_HIP(hipMemcpy(d, p, count * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

namespace {
enum { kPoolOut, kPoolCot, kPoolEdge, kPoolRows, kPoolPart };  // context pool slots of the feature calls

// A validated feature layer with its tables on the device and the current
// system's pairs exported (when it has atoms).
struct FeatPlan {
  FeatArgs a{};
  DeviceExport ex;
  std::vector<void*> owned;
  std::vector<int64_t> boff;
  int64_t n = 0, nrows = 0, nmix = 0, nout = 0;
  int Lk = 2;
  size_t nyk = 9;
  ~FeatPlan() {
    for (void* q : owned) (void)hipFree(q);
  }
  template <typename T>
  T* up(const T* p, size_t count) {
    return upload_raw(p, count, owned);
  }
  template <typename T>
  T* alloc(size_t count) {
    T* d = nullptr;
    PFW_HIP(hipMalloc((void**)&d, std::max<size_t>(count, 1) * sizeof(T)));
    owned.push_back(d);
    return d;
  }
};

void feat_prepare(pfw_ctx* ctx, const pfw_feature_layer* l, FeatPlan& p, bool with_S) {
  if (!l) throw Error(PFW_INVALID_ARGUMENT, "null feature layer");
  if (l->radial_kind != PFW_RADIAL_SMOOTH_EXPONENTIAL)
    throw Error(PFW_INVALID_ARGUMENT, "unsupported feature radial family");
  if (l->nz <= 0 || l->nr <= 0 || l->lmax < 0 || l->ndens < 0 || l->nprod < 0 || l->nblocks <= 0)
    throw Error(PFW_INVALID_ARGUMENT, "invalid feature layer dimensions");
  if (l->lmax > 6) throw Error(PFW_INVALID_ARGUMENT, "lmax above 6 is not supported by the device kernels");
  if (!(std::isfinite(l->cutoff) && l->cutoff > 0 && std::isfinite(l->rate) && std::isfinite(l->amplitude)))
    throw Error(PFW_INVALID_ARGUMENT, "nonfinite feature radial parameters");
  if (l->cutoff != ctx->cutoff)
    throw Error(PFW_STALE, "the system's cell list was built for another cutoff than the feature layer's");
  const int ny = (l->lmax + 1) * (l->lmax + 1);
  for (int t = 0; t < l->ndens; ++t)
    if (l->dens_z[t] < 0 || l->dens_z[t] > l->nz || l->dens_n[t] < 0 || l->dens_n[t] >= l->nr || l->dens_h[t] < 0 ||
        l->dens_h[t] >= ny)
      throw Error(PFW_INVALID_ARGUMENT, "invalid feature density entry");
  if (l->prod_ptr[0] != 0) throw Error(PFW_INVALID_ARGUMENT, "invalid feature product offsets");
  if (l->ndens > 0xfffe) throw Error(PFW_INVALID_ARGUMENT, "more than 65534 feature densities");
  for (int t = 0; t < l->nprod; ++t) {
    if (l->prod_ptr[t + 1] <= l->prod_ptr[t]) throw Error(PFW_INVALID_ARGUMENT, "empty feature product");
    if (l->prod_ptr[t + 1] - l->prod_ptr[t] > 4)
      throw Error(PFW_INVALID_ARGUMENT, "feature products of more than four factors are not supported by the device kernels");
    for (int q = l->prod_ptr[t]; q < l->prod_ptr[t + 1]; ++q)
      if (l->prod_idx[q] < 0 || l->prod_idx[q] >= l->ndens) throw Error(PFW_INVALID_ARGUMENT, "invalid product factor");
  }
  p.n = ctx->sys.n;
  p.boff.assign(6 * (size_t)l->nblocks, 0);
  for (int b = 0; b < l->nblocks; ++b) {
    const int dim = 2 * l->block_L[b] + 1, nb = l->block_nbasis[b], ch = l->block_channels[b];
    if (l->block_L[b] < 0 || nb <= 0 || ch <= 0) throw Error(PFW_INVALID_ARGUMENT, "invalid feature block");
    int64_t* o = p.boff.data() + 6 * b;
    o[0] = p.nrows;
    o[1] = p.nmix;
    o[2] = p.nout;
    o[3] = dim;
    o[4] = nb;
    o[5] = ch;
    p.nrows += (int64_t)dim * nb;
    p.nmix += (int64_t)nb * ch * l->nz;
    p.nout += (int64_t)dim * ch * p.n;
  }
  if (l->row_ptr[0] != 0) throw Error(PFW_INVALID_ARGUMENT, "invalid coupling row offsets");
  for (int64_t r = 0; r < p.nrows; ++r) {
    if (l->row_ptr[r + 1] < l->row_ptr[r]) throw Error(PFW_INVALID_ARGUMENT, "invalid coupling row offsets");
    for (int q = l->row_ptr[r]; q < l->row_ptr[r + 1]; ++q)
      if (l->col[q] < 0 || l->col[q] >= l->nprod || !std::isfinite(l->val[q]))
        throw Error(PFW_INVALID_ARGUMENT, "invalid coupling entry");
  }
  for (int t = 0; t < l->nr; ++t)
    if (l->powers[t] < 0) throw Error(PFW_INVALID_ARGUMENT, "negative radial power");
  p.Lk = l->lmax <= 2 ? 2 : l->lmax <= 4 ? 4 : 6;  // the kernels' harmonic degree
  p.nyk = (size_t)(p.Lk + 1) * (p.Lk + 1);
  if (p.n > 0) {
    std::vector<int32_t> sp(p.n);
    PFW_HIP(hipMemcpy(sp.data(), ctx->sys.species, p.n * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t t = 0; t < p.n; ++t)
      if (sp[t] < 0 || sp[t] >= l->nz) throw Error(PFW_INVALID_ARGUMENT, "species channel outside the feature layer");
  }
  if (p.n == 0) return;
  nl_export_device(ctx, p.ex, with_S);
  const int64_t nval = l->row_ptr[p.nrows];
  FeatArgs& a = p.a;
  a.n = p.n;
  a.off = p.ex.off;
  a.nbr = p.ex.j;
  a.R = p.ex.R;
  a.perm = ctx->sys.perm;
  a.species = ctx->sys.species;
  a.nz = l->nz;
  a.nr = l->nr;
  a.ndens = l->ndens;
  a.nrows = (int)p.nrows;
  a.nnz = nval;
  a.nblocks = l->nblocks;
  a.dz = p.up(l->dens_z, l->ndens);
  a.dn = p.up(l->dens_n, l->ndens);
  a.dh = p.up(l->dens_h, l->ndens);
  a.rptr = p.up(l->row_ptr, (size_t)p.nrows + 1);
  {  // coupling entries with their products' factors, in row order
    std::vector<FeatEntry> ent(nval);
    std::vector<int32_t> erow(nval);
    for (int64_t r = 0; r < p.nrows; ++r)
      for (int q = l->row_ptr[r]; q < l->row_ptr[r + 1]; ++q) {
        const int k = l->col[q];
        uint64_t f = 0;
        for (int u = 0; u < 4; ++u) {
          const int at = l->prod_ptr[k] + u;
          const uint64_t d = at < l->prod_ptr[k + 1] ? (uint64_t)l->prod_idx[at] : (uint64_t)l->ndens;
          f |= d << (16 * u);
        }
        ent[q] = FeatEntry{l->val[q], f};
        erow[q] = (int32_t)r;
      }
    a.ent = p.up(ent.data(), ent.size());
    a.erow = p.up(erow.data(), erow.size());
  }
  a.boff = p.up(p.boff.data(), p.boff.size());
  a.mix = p.up(l->mixing, (size_t)p.nmix);
  a.rc = l->radial_cutoff;
  a.rate = l->rate;
  a.amplitude = l->amplitude;
  a.powers = p.up(l->powers, (size_t)l->nr);
  a.scale = p.up(l->scale, (size_t)l->nr * l->nz * l->nz);
  a.status = p.alloc<int32_t>(1);
  PFW_HIP(hipMemsetAsync(a.status, 0, sizeof(int32_t), ctx->stream));
}

template <template <int> class K, typename... Args>
void launch_lk(int Lk, size_t lds, int64_t n, hipStream_t st, Args... args) {
  switch (Lk) {
    case 2:
      allow_lds(K<2>::fn, lds, "the feature kernels");
      hipLaunchKernelGGL(K<2>::fn, dim3((unsigned)n), dim3(kWave), lds, st, args...);
      break;
    case 4:
      allow_lds(K<4>::fn, lds, "the feature kernels");
      hipLaunchKernelGGL(K<4>::fn, dim3((unsigned)n), dim3(kWave), lds, st, args...);
      break;
    default:
      allow_lds(K<6>::fn, lds, "the feature kernels");
      hipLaunchKernelGGL(K<6>::fn, dim3((unsigned)n), dim3(kWave), lds, st, args...);
  }
  PFW_HIP(hipGetLastError());
}
template <int L>
struct FeatFwd {
  static constexpr auto fn = k_features<L>;
};
template <int L>
struct FeatVjp {
  static constexpr auto fn = k_features_vjp<L>;
};
}  // namespace

void features_run(pfw_ctx* ctx, const pfw_feature_layer* l, double* out) {
  if (!out) throw Error(PFW_INVALID_ARGUMENT, "null feature output");
  FeatPlan p;
  feat_prepare(ctx, l, p, false);
  const size_t lds =
      sizeof(double) * (p.nyk + (p.nyk + l->nr) * 64 + l->ndens + 1 + p.nrows) + 64 * sizeof(int);
  if (lds > kMaxLds) throw Error(PFW_INVALID_ARGUMENT, "feature tables too large for LDS");
  if (p.n == 0) return;  // no outputs (nout = 0); otherwise the copy below writes all of out
  double* dout = ctx->pool.get<double>(kPoolOut, p.nout);
  p.a.out = dout;
  hipStream_t st = ctx->stream;
  launch_lk<FeatFwd>(p.Lk, lds, p.n, st, p.a);
  int32_t h_status = 0;
  PFW_HIP(hipMemcpyAsync(&h_status, p.a.status, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipMemcpyAsync(out, dout, p.nout * sizeof(double), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipStreamSynchronize(st));
  if (h_status) throw Error(PFW_DOMAIN, "coincident physical atoms have undefined harmonics");
}

void features_vjp_run(pfw_ctx* ctx, const pfw_feature_layer* l, const double* cot, double* dx, double* dmixing,
                      double* dscale, double* dradial) {
  if (!cot || !dx || !dmixing || !dscale || !dradial) throw Error(PFW_INVALID_ARGUMENT, "null feature pullback buffer");
  FeatPlan p;
  feat_prepare(ctx, l, p, true);
  const int nz = l->nz, nr = l->nr;
  const int64_t nsc = (int64_t)nr * nz * nz;
  const size_t lds = sizeof(double) * (p.nyk + (p.nyk + 3 * (size_t)nr) * 64 + 2 * (size_t)l->ndens + 1 + p.nrows +
                                       (size_t)(l->ndens + 1) * kLaneStride + (size_t)nr * nz + p.nout / std::max<int64_t>(p.n, 1)) +
                     64 * sizeof(int);
  if (lds > kMaxLds) throw Error(PFW_INVALID_ARGUMENT, "feature tables too large for LDS");
  std::fill(dx, dx + 3 * p.n, 0.0);
  std::fill(dmixing, dmixing + p.nmix, 0.0);
  std::fill(dscale, dscale + nsc, 0.0);
  dradial[0] = dradial[1] = 0;
  if (p.n == 0) return;
  const int ndens = l->ndens;
  std::vector<int32_t> hptr(p.nyk + 1, 0), hidx(ndens);  // harmonic -> densities
  for (int t = 0; t < ndens; ++t) ++hptr[l->dens_h[t] + 1];
  for (size_t h = 0; h < p.nyk; ++h) hptr[h + 1] += hptr[h];
  {
    std::vector<int32_t> fill(hptr.begin(), hptr.end() - 1);
    for (int t = 0; t < ndens; ++t) hidx[fill[l->dens_h[t]]++] = t;
  }
  FeatVjpArgs v;
  v.cot = ctx->pool.get<double>(kPoolCot, p.nout);
  PFW_HIP(hipMemcpyAsync((void*)v.cot, cot, p.nout * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  v.hptr = p.up(hptr.data(), hptr.size());
  v.hidx = p.up(hidx.data(), hidx.size());
  const int64_t total = p.ex.total;
  v.gE = ctx->pool.get<double>(kPoolEdge, 3 * (size_t)total);
  v.so = (int)(p.nmix / nz);
  v.gper = (int)(p.nout / p.n);
  v.wp = v.so + nr * nz + 2;
  v.prow = ctx->pool.get<double>(kPoolRows, (size_t)p.n * v.wp);
  hipStream_t st = ctx->stream;
  launch_lk<FeatVjp>(p.Lk, lds, p.n, st, p.a, v);
  int32_t* inv = p.alloc<int32_t>(p.n);
  hipLaunchKernelGGL(k_inverse_perm, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, st, p.n, (const int32_t*)ctx->sys.perm,
                     inv);
  double* d_dx = p.alloc<double>(3 * (size_t)p.n);
  double* orphan = p.alloc<double>(3 * (size_t)p.n);
  PFW_HIP(hipMemsetAsync(orphan, 0, 3 * p.n * sizeof(double), st));
  hipLaunchKernelGGL(k_features_dx, dim3((unsigned)p.n), dim3(kWave), 0, st, p.n, (const int64_t*)p.ex.off,
                     (const int32_t*)p.ex.j, (const int32_t*)p.ex.S, (const int32_t*)ctx->sys.perm,
                     (const int32_t*)inv, (const double*)v.gE, d_dx, orphan, p.a.status);
  PFW_HIP(hipGetLastError());
  const int64_t nchunks = (p.n + kRowChunk - 1) / kRowChunk, width = (int64_t)nz * v.wp;
  double* part = ctx->pool.get<double>(kPoolPart, (size_t)nchunks * width);
  double* dsum = p.alloc<double>((size_t)width);
  hipLaunchKernelGGL(k_species_rows, dim3((unsigned)((v.wp + 63) / 64), (unsigned)nchunks), dim3(64), 0, st, p.n,
                     v.wp, nz, (const double*)v.prow, (const int32_t*)ctx->sys.species, part);
  hipLaunchKernelGGL(k_species_total, dim3((unsigned)((width + 255) / 256)), dim3(256), 0, st, nchunks, width,
                     (const double*)part, dsum);
  PFW_HIP(hipGetLastError());
  int32_t h_status = 0;
  std::vector<double> hsum(width), horph(3 * p.n);
  PFW_HIP(hipMemcpyAsync(&h_status, p.a.status, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipMemcpyAsync(dx, d_dx, 3 * p.n * sizeof(double), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipMemcpyAsync(hsum.data(), dsum, width * sizeof(double), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipMemcpyAsync(horph.data(), orphan, 3 * p.n * sizeof(double), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipStreamSynchronize(st));
  if (h_status & 4) throw Error(PFW_INVALID_ARGUMENT, "nonfinite feature cotangent");
  if (h_status & 1) throw Error(PFW_DOMAIN, "coincident physical atoms have undefined harmonics");
  if (h_status & 2)
    for (int64_t t = 0; t < 3 * p.n; ++t) dx[t] += horph[t];
  // species rows -> mixing (basis, channels, species per block), scale (radial, centre, neighbour), rate, amplitude
  for (int z = 0; z < nz; ++z) {
    const double* s = hsum.data() + (int64_t)z * v.wp;
    int64_t mo = 0;
    for (int b = 0; b < l->nblocks; ++b) {
      const int64_t* o = p.boff.data() + 6 * b;
      const int64_t w = o[4] * o[5];
      for (int64_t t = 0; t < w; ++t) dmixing[o[1] + w * z + t] = s[mo + t];
      mo += w;
    }
    for (int zj = 0; zj < nz; ++zj)
      for (int nn = 0; nn < nr; ++nn) dscale[nn + (int64_t)nr * (z + (int64_t)nz * zj)] = s[v.so + nn + nr * zj];
    dradial[0] += s[v.so + nr * nz];
    dradial[1] += s[v.so + nr * nz + 1];
  }
}

// Force rows of one output column: atom i (original order) of configuration c
// has its rows at 10 c + 1 + 3 i + axis (each configuration's block holds its
// energy, 3 n_c force rows and 9 virial rows).
__global__ void k_design_forces(int64_t n, const int32_t* __restrict__ atom_cfg, const double* __restrict__ F,
                                double* __restrict__ col) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * n) return;
  col[10 * (int64_t)atom_cfg[k / 3] + 1 + k] = F[k];
}

// Energy and virial rows of one output column from the per-configuration
// rows (E, W[9], status) of the reductions; a nonfinite status sets `bad`.
__global__ void k_design_scalars(int ncfg, const int64_t* __restrict__ atom0, const double* __restrict__ cfg_out,
                                 double* __restrict__ col, int32_t* __restrict__ bad) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncfg) return;
  const double* o = cfg_out + 11 * (int64_t)c;
  col[10 * (int64_t)c + 3 * atom0[c]] = o[0];
  for (int k = 0; k < 9; ++k) col[10 * (int64_t)c + 3 * atom0[c + 1] + 1 + k] = o[1 + k];
  if (o[10] != 0.0) atomicOr(bad, 1);
}

// Device buffers owned by one call (freed on every exit path).
struct DeviceScratch {
  std::vector<void*> p;
  template <typename T>
  T* get(size_t count) {
    void* q = nullptr;
    PFW_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    p.push_back(q);
    return (T*)q;
  }
  ~DeviceScratch() {
    for (void* q : p) (void)hipFree(q);
  }
};

// Outputs of nsets parameter sets on one environment build: set k's outputs
// form column k of `out` (rows x nsets, column-major, rows = 10 ncfg + 3 n):
// for every configuration in order its energy, vec(F) (3 x n_c, original atom
// order) and vec(W) (3 x 3, column-major). E, F and W are linear in the
// readout, so with unit parameter sets the columns, less the column of the
// zero set, are the linear design of design_matrix_impl
// (src/submodules/ACE/fitting.jl:18-63), and the zero set's column is its
// baseline. Each set runs the site kernels, the force gather and the
// reductions of a step; the parameters set before the call stay in force.
void ace_parameter_sets(pfw_ctx* ctx, int precision, int64_t nsets, const double* WB, const double* Wpair,
                        double* out) {
  ctx->host_written = false;
  AceDev& ad = ctx->ace;
  if (!ad.ready) throw Error(PFW_STALE, "no ACE model set");
  if (ad.cutoff != ctx->cutoff)
    throw Error(PFW_STALE, "the system's cell list was built for another cutoff than the ACE search cutoff");
  const int nz = ad.nz, nb = ad.nbasis, nq = ad.pair.nr;
  if (nsets < 0) throw Error(PFW_INVALID_ARGUMENT, "negative number of parameter sets");
  if (nsets == 0) return;
  if (!WB || !out || (nq > 0 && !Wpair)) throw Error(PFW_INVALID_ARGUMENT, "null parameter sets or output");
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n, rows = 10 * (int64_t)ctx->ncfg + 3 * n;
  if (rows > 0 && nsets > ((int64_t)1 << 40) / rows) throw Error(PFW_INVALID_ARGUMENT, "design output too large");
  for (;;) {  // environments with the reverse index, as in a step
    nl_build(ctx);
    if (n == 0) break;
    ace_environment(ctx, 1);
    PFW_HIP(hipStreamSynchronize(st));
    if (!ace_overflowed(ad)) break;
  }
  // every set's device weights at once; the kernels are pointed at one set after the other
  std::vector<double> w, ws, wp, W, WS, WP;
  for (int64_t k = 0; k < nsets; ++k) {
    readout_weights(ad, WB + k * (int64_t)nb * nz, nq > 0 ? Wpair + k * (int64_t)nq * nz : nullptr, w, ws, wp);
    W.insert(W.end(), w.begin(), w.end());
    WS.insert(WS.end(), ws.begin(), ws.end());
    WP.insert(WP.end(), wp.begin(), wp.end());
  }
  const size_t nw = w.size(), nws = ws.size(), nwp = wp.size();
  std::vector<float> W32(W.begin(), W.end()), WS32(WS.begin(), WS.end()), WP32(WP.begin(), WP.end());
  DeviceScratch dev;
  double* W64d = dev.get<double>(W.size());
  double* WS64d = dev.get<double>(WS.size());
  double* WP64d = dev.get<double>(WP.size());
  float* W32d = dev.get<float>(W32.size());
  float* WS32d = dev.get<float>(WS32.size());
  float* WP32d = dev.get<float>(WP32.size());
  PFW_HIP(hipMemcpy(W64d, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice));
  PFW_HIP(hipMemcpy(WS64d, WS.data(), WS.size() * sizeof(double), hipMemcpyHostToDevice));
  PFW_HIP(hipMemcpy(WP64d, WP.data(), WP.size() * sizeof(double), hipMemcpyHostToDevice));
  PFW_HIP(hipMemcpy(W32d, W32.data(), W32.size() * sizeof(float), hipMemcpyHostToDevice));
  PFW_HIP(hipMemcpy(WS32d, WS32.data(), WS32.size() * sizeof(float), hipMemcpyHostToDevice));
  PFW_HIP(hipMemcpy(WP32d, WP32.data(), WP32.size() * sizeof(float), hipMemcpyHostToDevice));
  double* dout = dev.get<double>((size_t)rows * nsets);
  int64_t* atom0 = dev.get<int64_t>(ctx->atom0.size());
  int32_t* bad = dev.get<int32_t>(1);
  PFW_HIP(hipMemcpy(atom0, ctx->atom0.data(), ctx->atom0.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  PFW_HIP(hipMemset(bad, 0, sizeof(int32_t)));
  double* keep64[3] = {ad.w64, ad.ws64, ad.wp64};
  float* keep32[3] = {ad.w32, ad.ws32, ad.wp32};
  auto restore = [&] {
    ad.w64 = keep64[0];
    ad.ws64 = keep64[1];
    ad.wp64 = keep64[2];
    ad.w32 = keep32[0];
    ad.ws32 = keep32[1];
    ad.wp32 = keep32[2];
  };
  try {
    for (int64_t k = 0; k < nsets; ++k) {
      ad.w64 = W64d + k * nw;
      ad.ws64 = WS64d + k * nws;
      ad.wp64 = WP64d + k * nwp;
      ad.w32 = W32d + k * nw;
      ad.ws32 = WS32d + k * nws;
      ad.wp32 = WP32d + k * nwp;
      double* col = dout + k * rows;
      if (n > 0) {
        ace_evaluate_rows(ctx, 1, precision);
        hipLaunchKernelGGL(k_design_forces, dim3(nblk(3 * n, 256)), dim3(256), 0, st, n, s.atom_cfg, ctx->d_F, col);
      } else {
        PFW_HIP(hipMemsetAsync(ctx->d_cfg_out, 0, 11 * (size_t)ctx->ncfg * sizeof(double), st));
      }
      hipLaunchKernelGGL(k_design_scalars, dim3(nblk(ctx->ncfg, 64)), dim3(64), 0, st, ctx->ncfg, atom0,
                         ctx->d_cfg_out, col, bad);
      PFW_HIP(hipGetLastError());
    }
    PFW_HIP(hipStreamSynchronize(st));
  } catch (...) {
    restore();
    ctx->host_written = false;
    throw;
  }
  restore();
  ctx->host_written = false;
  ctx->have_results = false;  // d_out / d_F hold the last set's outputs, not a step's
  int32_t h_bad = 0;
  PFW_HIP(hipMemcpy(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost));
  if (h_bad) throw Error(PFW_DOMAIN, "nonfinite ACE output (coincident atoms or singular envelope)");
  PFW_HIP(hipMemcpy(out, dout, (size_t)rows * nsets * sizeof(double), hipMemcpyDeviceToHost));
}

// ---- the linear design by per-site Jacobians (design_matrix_impl,
// src/submodules/ACE/fitting.jl:18-63, forms each site's Jacobian once). For
// the unprojected basis, the column of basis function k and centre species z
// holds the energy, forces and virial of B_k summed over the sites of species
// z. Per site: A (from k_ace_dens), G = dB/dA at the model's nonzero (k, d)
// entries (each a sum of product-rule terms), B; per edge e and basis
// function k, g_e^k = sum_d G_kd dA_d/dR_e with dA_d/dR_e = R'_r Y_m r^ +
// R_r grad Y_m (r, m of density d). The forces gather g over each atom's
// edges and their reverses, as the step's force gather does, into the column
// block of the source site's species.

constexpr int kDesignThreads = 256;  // four wavefronts per site

template <typename T>
struct DesignArgs {
  int nb, nG;
  const int32_t* row;   // nb + 1: entries of each basis row
  const int32_t* d;     // nG: density of each entry
  const int32_t* term;  // nG + 1: terms of each entry
  const int32_t* f;     // 3 per term: the other factors (-1: none)
  const T* c;           // per term: the coupling value (times the basis scale)
  const int32_t* brow;  // nb + 1: products of each basis row
  const int32_t* bprod;
  const T* bc;
  T* Bsite;  // n x nb
  T* gbuf;   // (n ecap) x nb x 3
};

template <typename T>
struct V3 {
  T x, y, z;
};

// Byte offsets of k_design_site's LDS regions (CE edges per chunk).
struct DesignLds {
  size_t fl, A, G, gd, yv, yg, rv, rd, eg, rp, ech, D, total;
};
template <typename T>
__host__ __device__ inline DesignLds design_lds(int ny, int ndens, int nG, int nr, int ce) {
  DesignLds o;
  size_t b = 0;
  o.fl = b;
  b = align16(b + sizeof(T) * ny);
  o.A = b;
  b = align16(b + sizeof(T) * ndens);
  o.G = b;
  b = align16(b + sizeof(T) * nG);
  o.gd = b;
  b = align16(b + sizeof(uint16_t) * nG);
  o.yv = b;
  b = align16(b + sizeof(T) * ce * ny);
  o.yg = b;
  b = align16(b + sizeof(T) * ce * 3 * ny);
  o.rv = b;
  b = align16(b + sizeof(T) * ce * nr);
  o.rd = b;
  b = align16(b + sizeof(T) * ce * nr);
  o.eg = b;
  b = align16(b + sizeof(T) * ce * 4);
  o.rp = b;
  b = align16(b + sizeof(RadialPoint<T>) * ce);
  o.ech = b;
  b = align16(b + sizeof(int) * ce);
  o.D = b;
  b = align16(b + sizeof(V3<T>) * ce * ndens);
  o.total = b;
  return o;
}

template <typename T, int L, int CE>
__global__ __launch_bounds__(kDesignThreads) void k_design_site(AceArgs<T> a, DesignArgs<T> t) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NY = (L + 1) * (L + 1);
  const int nr = a.rad.nr;
  const DesignLds o = design_lds<T>(NY, a.ndens, t.nG, nr, CE);
  T* Fl = (T*)(smem + o.fl);
  T* A = (T*)(smem + o.A);
  T* G = (T*)(smem + o.G);
  uint16_t* Gd = (uint16_t*)(smem + o.gd);
  T* Yv = (T*)(smem + o.yv);
  T* Yg = (T*)(smem + o.yg);
  T* Rv = (T*)(smem + o.rv);
  T* Rd = (T*)(smem + o.rd);
  T* Eg = (T*)(smem + o.eg);
  RadialPoint<T>* Rp = (RadialPoint<T>*)(smem + o.rp);
  int* Ech = (int*)(smem + o.ech);
  V3<T>* D = (V3<T>*)(smem + o.D);
  const int lane = threadIdx.x;
  const int64_t i = blockIdx.x;
  fill_fl<T, L>(Fl);
  for (int k = lane; k < a.ndens; k += kDesignThreads) A[k] = a.Abuf[i * a.dstride + k];
  __syncthreads();
  for (int g = lane; g < t.nG; g += kDesignThreads) {
    T s = 0;
    for (int u = t.term[g]; u < t.term[g + 1]; ++u) {
      T v = t.c[u];
      for (int q = 0; q < 3; ++q) {
        const int f = t.f[3 * u + q];
        if (f >= 0) v *= A[f];
      }
      s += v;
    }
    G[g] = s;
    Gd[g] = (uint16_t)t.d[g];
  }
  for (int k = lane; k < t.nb; k += kDesignThreads) {
    T s = 0;
    for (int u = t.brow[k]; u < t.brow[k + 1]; ++u) {
      const int p = t.bprod[u];
      T v = t.bc[u];
      for (int q = 0; q < 4; ++q) {
        const int f = a.prod[4 * p + q];
        if (f >= 0) v *= A[f];
      }
      s += v;
    }
    t.Bsite[i * t.nb + k] = s;
  }
  __syncthreads();
  const int iz = a.sps[i];
  const int ne = a.ecount[i];
  const int64_t e0 = i * (int64_t)a.ecap;
  for (int c0 = 0; c0 < ne; c0 += CE) {
    const int ce = min(CE, ne - c0);
    if (lane < ce) {  // the edge's geometry, radial point and harmonics with gradients
      const double* R = a.eR + 3 * (e0 + c0 + lane);
      const T dx = (T)R[0], dy = (T)R[1], dz = (T)R[2];
      const T r = (T)sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      const T ri = (T)1 / r;
      Eg[4 * lane] = dx * ri;
      Eg[4 * lane + 1] = dy * ri;
      Eg[4 * lane + 2] = dz * ri;
      const int ch = iz + a.nz * a.sps[a.enb[e0 + c0 + lane]];
      Ech[lane] = ch;
      Rp[lane] = radial_point(a.rad, ch, r);
      T* yv = Yv + lane * NY;
      T* yg = Yg + lane * 3 * NY;
      ylm_emit<T, L>(dx, dy, dz, Fl, a.solid, [&](int m, const Jet<T>& j) {
        yv[m] = j.v;
        yg[3 * m] = j.g0;
        yg[3 * m + 1] = j.g1;
        yg[3 * m + 2] = j.g2;
      });
    }
    __syncthreads();
    for (int u = lane; u < ce * nr; u += kDesignThreads) {
      const int e = u / nr, n = u - e * nr;
      T v, d;
      radial_value(a.rad, Ech[e], Rp[e], n, v, d);
      Rv[e * nr + n] = v;
      Rd[e * nr + n] = d;
    }
    __syncthreads();
    for (int u = lane; u < ce * a.ndens; u += kDesignThreads) {
      const int e = u / a.ndens, k = u - e * a.ndens;
      const int rr = a.dens[2 * k], mm = a.dens[2 * k + 1];
      const T s = Rd[e * nr + rr] * Yv[e * NY + mm], v = Rv[e * nr + rr];
      const T* yg = Yg + (e * NY + mm) * 3;
      D[e * a.ndens + k] = V3<T>{s * Eg[4 * e] + v * yg[0], s * Eg[4 * e + 1] + v * yg[1], s * Eg[4 * e + 2] + v * yg[2]};
    }
    __syncthreads();
    for (int k = lane; k < t.nb; k += kDesignThreads) {
      T acc[CE][3];
#pragma unroll
      for (int e = 0; e < CE; ++e) acc[e][0] = acc[e][1] = acc[e][2] = 0;
      for (int g = t.row[k]; g < t.row[k + 1]; ++g) {
        const T gv = G[g];
        const int d = Gd[g];
#pragma unroll
        for (int e = 0; e < CE; ++e) {  // edges past ce read in-bounds stale tiles; their sums are not written
          const V3<T> q = D[e * a.ndens + d];
          acc[e][0] += gv * q.x;
          acc[e][1] += gv * q.y;
          acc[e][2] += gv * q.z;
        }
      }
#pragma unroll
      for (int e = 0; e < CE; ++e)
        if (e < ce) {
          T* out = t.gbuf + ((e0 + c0 + e) * t.nb + k) * 3;
          out[0] = acc[e][0];
          out[1] = acc[e][1];
          out[2] = acc[e][2];
        }
    }
    __syncthreads();
  }
}

// Force rows of every atom: one workgroup per (sorted) atom, threads over basis
// functions; F_a^{k, z_a} += sum_out g_e^k and, for each out-edge's reverse
// (an edge of neighbour site j), F_a^{k, z_j} -= g_rev^k, in edge order
// (deterministic). Also the atom's own-site virial -sum_out R_e (x) g_e^k.
template <typename T>
__global__ __launch_bounds__(kDesignThreads) void k_design_gather(AceArgs<T> a, DesignArgs<T> t, const int32_t* __restrict__ erev,
                                                      const int32_t* __restrict__ atom_cfg, double* __restrict__ Wpart,
                                                      double* __restrict__ out, int64_t ld) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* acc = (double*)smem;  // nz x 3 per thread: acc[(thread * nz + z) * 3 + c]
  const int lane = threadIdx.x;
  const int64_t i = blockIdx.x;
  const int nz = a.nz, nb = t.nb;
  const int zi = a.sps[i];
  const int ne = a.ecount[i];
  const int64_t e0 = i * (int64_t)a.ecap;
  const int64_t o = a.perm[i];
  const int64_t row0 = 10 * (int64_t)atom_cfg[o] + 1 + 3 * o;
  double* my = acc + (size_t)lane * nz * 3;
  for (int k = lane; k < nb; k += kDesignThreads) {
    for (int q = 0; q < 3 * nz; ++q) my[q] = 0;
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double f0 = 0, f1 = 0, f2 = 0;
    for (int e = 0; e < ne; ++e) {
      const int64_t s = e0 + e;
      const T* g = t.gbuf + (s * nb + k) * 3;
      const double g0 = (double)g[0], g1 = (double)g[1], g2 = (double)g[2];
      f0 += g0;
      f1 += g1;
      f2 += g2;
      const double* R = a.eR + 3 * s;
      for (int c = 0; c < 3; ++c) {
        w[c] -= R[c] * g0;
        w[c + 3] -= R[c] * g1;
        w[c + 6] -= R[c] * g2;
      }
      const int r = erev[s];
      if (r >= 0) {
        const int zj = a.sps[a.enb[s]];
        const T* h = t.gbuf + ((int64_t)r * nb + k) * 3;
        my[3 * zj] -= (double)h[0];
        my[3 * zj + 1] -= (double)h[1];
        my[3 * zj + 2] -= (double)h[2];
      }
    }
    my[3 * zi] += f0;
    my[3 * zi + 1] += f1;
    my[3 * zi + 2] += f2;
    for (int z = 0; z < nz; ++z)
      for (int c = 0; c < 3; ++c) out[(row0 + c) * ld + k + (int64_t)nb * z] = my[3 * z + c];
    for (int c = 0; c < 9; ++c) Wpart[(i * nb + k) * 9 + c] = w[c];
  }
}

// Energy and virial rows of every configuration: a block of ng x 64 threads per
// (configuration, 64 basis functions); thread group q sums the sites q, q + ng,
// ... of the configuration (sorted order) per centre species, and the groups'
// sums are added in group order (deterministic).
template <typename T>
__global__ __launch_bounds__(256) void k_design_rows(int nb, int nz, const int32_t* __restrict__ cfg_cell0,
                                                     const int32_t* __restrict__ start, const int64_t* __restrict__ atom0,
                                                     const int32_t* __restrict__ sps, const T* __restrict__ Bsite,
                                                     const double* __restrict__ Wpart, double* __restrict__ out,
                                                     int64_t ld) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* acc = (double*)smem;  // ng groups x 64 lanes x nz x 10
  const int c = blockIdx.x;
  const int lane = threadIdx.x % 64, grp = threadIdx.x / 64, ng = blockDim.x / 64;
  const int k = blockIdx.y * 64 + lane;
  double* my = acc + ((size_t)grp * 64 + lane) * nz * 10;
  for (int q = 0; q < nz * 10; ++q) my[q] = 0;
  const int lo = start[cfg_cell0[c]], hi = start[cfg_cell0[c + 1]];
  if (k < nb)
    for (int i = lo + grp; i < hi; i += ng) {
      double* a = my + 10 * sps[i];
      a[0] += (double)Bsite[(int64_t)i * nb + k];
      const double* w = Wpart + ((int64_t)i * nb + k) * 9;
      for (int q = 0; q < 9; ++q) a[1 + q] += w[q];
    }
  __syncthreads();
  if (grp != 0 || k >= nb) return;
  const int64_t erow = 10 * (int64_t)c + 3 * atom0[c], wrow = 10 * (int64_t)c + 3 * atom0[c + 1] + 1;
  for (int z = 0; z < nz; ++z) {
    double v[10];
    for (int q = 0; q < 10; ++q) {
      v[q] = 0;
      for (int g = 0; g < ng; ++g) v[q] += acc[((size_t)g * 64 + lane) * nz * 10 + 10 * z + q];
    }
    out[erow * ld + k + (int64_t)nb * z] = v[0];
    for (int q = 0; q < 9; ++q) out[(wrow + q) * ld + k + (int64_t)nb * z] = v[1 + q];
  }
}

constexpr size_t kDesignLds = 40 * 1024;

template <typename T, int L, int CE>
static void launch_design_site(pfw_ctx* ctx, const DesignArgs<T>& t, size_t lds) {
  AceArgs<T> a = make_args<T>(ctx, 0);
  if (lds > 64 * 1024)
    PFW_HIP(hipFuncSetAttribute((const void*)k_design_site<T, L, CE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  hipLaunchKernelGGL((k_design_site<T, L, CE>), dim3((unsigned)ctx->sys.n), dim3(kDesignThreads), lds, ctx->stream, a,
                     t);
  PFW_HIP(hipGetLastError());
}

// The densities alone (k_ace_dens, one site per wavefront).
template <typename T, int L>
static void launch_dens_only(pfw_ctx* ctx) {
  AceDev& ad = ctx->ace;
  AceArgs<T> a = make_args<T>(ctx, 0);
  constexpr int NY = (L + 1) * (L + 1);
  const size_t lds = sizeof(T) * ((size_t)NY + (size_t)ad.ndens + (size_t)(ad.rchunk + NY) * kStride);
  allow_lds(k_ace_dens<T, L, 1>, lds, "the density kernel");
  hipLaunchKernelGGL((k_ace_dens<T, L, 1>), dim3((unsigned)ctx->sys.n), dim3(kWave), lds, ctx->stream, a);
  PFW_HIP(hipGetLastError());
}

template <typename T, int L>
static void design_run(pfw_ctx* ctx, const std::vector<int32_t>& row, const std::vector<int32_t>& dd,
                       const std::vector<int32_t>& term, const std::vector<int32_t>& ff, const std::vector<double>& cc,
                       const std::vector<int32_t>& brow, const std::vector<int32_t>& bprod,
                       const std::vector<double>& bc, double* out, int64_t ld) {
  AceDev& ad = ctx->ace;
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int nb = ad.nbasis, nz = ad.nz;
  const int64_t n = s.n, ncols = (int64_t)nb * nz, rows = 10 * (int64_t)ctx->ncfg + 3 * n;
  constexpr int NY = (L + 1) * (L + 1);
  const int nG = (int)dd.size();
  // the largest edge chunk whose tiles leave room for several workgroups per CU
  int ce = 0;
  size_t lds = 0;
  for (int c : {8, 4, 2}) {
    lds = design_lds<T>(NY, ad.ndens, nG, ad.radial.nr, c).total;
    ce = c;
    if (lds <= kDesignLds) break;
  }
  if (lds > kMaxLds) throw Error(PFW_INVALID_ARGUMENT, "ACE model too large for the design kernel's LDS tiles");
  const double gbytes = (double)n * ad.ecap * nb * 3 * sizeof(T);
  if (gbytes > 64e9) throw Error(PFW_CAPACITY, "design: the per-edge basis gradients exceed 64 GB; use fewer configurations");
  DeviceScratch dev;
  auto up32 = [&](const std::vector<int32_t>& v) {
    int32_t* p = dev.get<int32_t>(v.size());
    if (!v.empty()) PFW_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    return p;
  };
  auto upT = [&](const std::vector<double>& v) {
    std::vector<T> w(v.begin(), v.end());
    T* p = dev.get<T>(w.size());
    if (!w.empty()) PFW_HIP(hipMemcpyAsync(p, w.data(), w.size() * sizeof(T), hipMemcpyHostToDevice, st));
    PFW_HIP(hipStreamSynchronize(st));  // w leaves scope
    return p;
  };
  DesignArgs<T> t;
  t.nb = nb;
  t.nG = nG;
  t.row = up32(row);
  t.d = up32(dd);
  t.term = up32(term);
  t.f = up32(ff);
  t.c = upT(cc);
  t.brow = up32(brow);
  t.bprod = up32(bprod);
  t.bc = upT(bc);
  t.Bsite = dev.get<T>((size_t)n * nb);
  t.gbuf = dev.get<T>((size_t)n * ad.ecap * nb * 3);
  double* Wpart = dev.get<double>((size_t)n * nb * 9);
  double* dout = dev.get<double>((size_t)rows * ld);
  int64_t* atom0 = dev.get<int64_t>(ctx->atom0.size());
  PFW_HIP(hipMemcpyAsync(atom0, ctx->atom0.data(), ctx->atom0.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
  launch_dens_only<T, L>(ctx);
  if (ce == 8)
    launch_design_site<T, L, 8>(ctx, t, lds);
  else if (ce == 4)
    launch_design_site<T, L, 4>(ctx, t, lds);
  else
    launch_design_site<T, L, 2>(ctx, t, lds);
  AceArgs<T> a = make_args<T>(ctx, 0);
  const size_t glds = sizeof(double) * kDesignThreads * 3 * nz;
  if (glds > 64 * 1024) throw Error(PFW_INVALID_ARGUMENT, "too many species for the design gather");
  hipLaunchKernelGGL(k_design_gather<T>, dim3((unsigned)n), dim3(kDesignThreads), glds, st, a, t, (const int32_t*)ad.erev,
                     (const int32_t*)s.atom_cfg, Wpart, dout, ld);
  PFW_HIP(hipGetLastError());
  const size_t glds_rows = sizeof(double) * 64 * 10 * nz;  // per group of 64 threads
  const int ng = (int)std::min<size_t>(4, (64 * 1024) / glds_rows);
  if (ng < 1) throw Error(PFW_INVALID_ARGUMENT, "too many species for the design rows");
  hipLaunchKernelGGL(k_design_rows<T>, dim3((unsigned)ctx->ncfg, (unsigned)((nb + 63) / 64)), dim3(64 * ng),
                     ng * glds_rows, st, nb, nz,
                     (const int32_t*)s.cfg_cell0, (const int32_t*)s.start, (const int64_t*)atom0, (const int32_t*)s.sps,
                     (const T*)t.Bsite, (const double*)Wpart, dout, ld);
  PFW_HIP(hipGetLastError());
  // whole rows of the caller's width in one copy (the columns past ncols are the caller's to fill)
  PFW_HIP(hipMemcpyAsync(out, dout, (size_t)rows * ld * sizeof(double), hipMemcpyDeviceToHost, st));
  PFW_HIP(hipStreamSynchronize(st));
}

// The basis columns of the linear design (rows as ace_parameter_sets', row-major
// rows x nb nz, column k + nb z), by per-site Jacobians in one pass.
void ace_design(pfw_ctx* ctx, int precision, const double* basis_scale, double* out, int64_t ld) {
  ctx->host_written = false;
  ctx->have_results = false;
  AceDev& ad = ctx->ace;
  if (!ad.ready) throw Error(PFW_STALE, "no ACE model set");
  if (ad.cutoff != ctx->cutoff)
    throw Error(PFW_STALE, "the system's cell list was built for another cutoff than the ACE search cutoff");
  if (ad.proj) throw Error(PFW_INVALID_ARGUMENT, "species-projected bases: use pfw_ace_parameter_sets");
  if (precision != PFW_FP64 && precision != PFW_MIXED) throw Error(PFW_INVALID_ARGUMENT, "unknown precision");
  if (!out) throw Error(PFW_INVALID_ARGUMENT, "null output");
  const int nb = ad.nbasis;
  SystemBuffers& s = ctx->sys;
  const int64_t n = s.n, ncols = (int64_t)nb * ad.nz, rows = 10 * (int64_t)ctx->ncfg + 3 * n;
  if (ld < ncols) throw Error(PFW_INVALID_ARGUMENT, "design row stride below nbasis x nz");
  if (n == 0 || ncols == 0) {
    for (int64_t r = 0; r < rows; ++r) std::memset(out + r * ld, 0, sizeof(double) * (size_t)ncols);
    return;
  }
  // host tables: the entries (k, d) of dB/dA with their product-rule terms,
  // and the products of each basis row, scaled by the basis scale (the mask)
  std::vector<double> scale(nb, 1.0);
  if (basis_scale)
    for (int k = 0; k < nb; ++k) scale[k] = basis_scale[k];
  std::map<std::pair<int, int>, std::vector<std::array<double, 4>>> ent;  // (k, d) -> (c, f0, f1, f2)
  std::vector<std::vector<std::pair<int, double>>> bp(nb);
  for (size_t u = 0; u < ad.c_vals.size(); ++u) {
    const int k = ad.c_rows[u], p = ad.c_cols[u];
    const double c = ad.c_vals[u] * scale[k];
    bp[k].push_back({p, c});
    int fac[4], nu = 0;
    for (int q = 0; q < 4; ++q)
      if (ad.h_prod[4 * (size_t)p + q] >= 0) fac[nu++] = ad.h_prod[4 * (size_t)p + q];
    for (int tt = 0; tt < nu; ++tt) {
      std::array<double, 4> term = {c, -1, -1, -1};
      for (int q = 0, w = 1; q < nu; ++q)
        if (q != tt) term[w++] = fac[q];
      ent[{k, fac[tt]}].push_back(term);
    }
  }
  std::vector<int32_t> row(nb + 1, 0), dd, term(1, 0), ff, brow(nb + 1, 0), bprod;
  std::vector<double> cc, bc;
  for (const auto& kv : ent) {
    row[kv.first.first + 1]++;
    dd.push_back(kv.first.second);
    for (const auto& tm : kv.second) {
      cc.push_back(tm[0]);
      for (int q = 1; q < 4; ++q) ff.push_back((int32_t)tm[q]);
    }
    term.push_back((int32_t)cc.size());
  }
  for (int k = 0; k < nb; ++k) row[k + 1] += row[k];
  for (int k = 0; k < nb; ++k) {
    for (const auto& pc : bp[k]) {
      bprod.push_back(pc.first);
      bc.push_back(pc.second);
    }
    brow[k + 1] = (int32_t)bprod.size();
  }
  for (;;) {  // environments with the reverse index, as in a step
    nl_build(ctx);
    ace_environment(ctx, 1);
    PFW_HIP(hipStreamSynchronize(ctx->stream));
    if (!ace_overflowed(ad)) break;
  }
  const int L = ad.lmax <= 2 ? 2 : ad.lmax <= 4 ? 4 : 6;
#define PFW_DESIGN(T, LL) design_run<T, LL>(ctx, row, dd, term, ff, cc, brow, bprod, bc, out, ld)
  if (precision == PFW_MIXED) {
    if (L == 2)
      PFW_DESIGN(float, 2);
    else if (L == 4)
      PFW_DESIGN(float, 4);
    else
      PFW_DESIGN(float, 6);
  } else {
    if (L == 2)
      PFW_DESIGN(double, 2);
    else if (L == 4)
      PFW_DESIGN(double, 4);
    else
      PFW_DESIGN(double, 6);
  }
#undef PFW_DESIGN
}

template <typename T, int L>
static void launch_adjoint(pfw_ctx* ctx, const double* Ebar, const double* Fbar, const double* Wbar, double* part,
                           int nb) {
  AceDev& ad = ctx->ace;
  AceArgs<T> a = make_args<T>(ctx, 0);
  constexpr int NY = (L + 1) * (L + 1);
  const size_t gsize = (size_t)ad.nz * (ad.nprod + ad.pair.nr);
  const size_t lds = sizeof(double) * gsize +
                     sizeof(T) * ((size_t)NY + 2 * (size_t)ad.ndens + 2 * (size_t)(ad.radial.nr + NY) * kStride);
  if (lds > 160 * 1024) throw Error(PFW_INVALID_ARGUMENT, "ACE model too large for the adjoint kernel's LDS tiles");
  if (lds > 64 * 1024)
    PFW_HIP(hipFuncSetAttribute((const void*)k_ace_adjoint<T, L>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_ace_adjoint<T, L>), dim3(nb), dim3(kWave), lds, ctx->stream, a,
                     (const int32_t*)ctx->sys.atom_cfg, Ebar, Fbar, Wbar, part);
  PFW_HIP(hipGetLastError());
}

// dL/dWB (nbasis x nz) and dL/dWpair (nq x nz) of a loss with the given
// dL/dE per configuration, dL/dF (3 x N, original order) and, optionally,
// dL/dW (3 x 3 column-major per configuration).
void ace_adjoint(pfw_ctx* ctx, int precision, const double* Ebar, const double* Fbar, const double* Wbar, double* gWB,
                 double* gWpair) {
  AceDev& ad = ctx->ace;
  if (!ad.ready) throw Error(PFW_STALE, "no ACE model set");
  if (ad.cutoff != ctx->cutoff)
    throw Error(PFW_STALE, "the system's cell list was built for another cutoff than the ACE search cutoff");
  if (ad.proj)
    throw Error(PFW_INVALID_ARGUMENT,
                "parameter adjoints of species-projected, order-restricted ACE bases are not implemented on the device");
  SystemBuffers& s = ctx->sys;
  hipStream_t st = ctx->stream;
  const int64_t n = s.n;
  const int nq = ad.pair.nr;
  const size_t gsize = (size_t)ad.nz * (ad.nprod + nq);
  std::vector<double> g(gsize, 0.0);
  if (n > 0) {
    for (int c = 0; c < ctx->ncfg; ++c)
      if (!std::isfinite(Ebar[c])) throw Error(PFW_INVALID_ARGUMENT, "nonfinite energy adjoint");
    for (int64_t k = 0; k < 3 * n; ++k)
      if (!std::isfinite(Fbar[k])) throw Error(PFW_INVALID_ARGUMENT, "nonfinite force adjoint");
    if (Wbar)
      for (int64_t k = 0; k < 9 * (int64_t)ctx->ncfg; ++k)
        if (!std::isfinite(Wbar[k])) throw Error(PFW_INVALID_ARGUMENT, "nonfinite virial adjoint");
    nl_build(ctx);
    ace_environment(ctx, 0);
    const int nb = (int)std::min<int64_t>(n, 1024);
    DeviceArray<double> dE(ctx->ncfg), dF(3 * n), dW(Wbar ? 9 * (size_t)ctx->ncfg : 0), part((size_t)nb * gsize),
        sum(gsize);
    PFW_HIP(hipMemcpyAsync(dE.p, Ebar, ctx->ncfg * sizeof(double), hipMemcpyHostToDevice, st));
    PFW_HIP(hipMemcpyAsync(dF.p, Fbar, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
    if (Wbar) PFW_HIP(hipMemcpyAsync(dW.p, Wbar, 9 * (size_t)ctx->ncfg * sizeof(double), hipMemcpyHostToDevice, st));
    const int L = ad.lmax <= 2 ? 2 : ad.lmax <= 4 ? 4 : 6;
    auto launch = precision == PFW_MIXED
                      ? (L == 2 ? launch_adjoint<float, 2> : L == 4 ? launch_adjoint<float, 4> : launch_adjoint<float, 6>)
                      : (L == 2 ? launch_adjoint<double, 2>
                                : L == 4 ? launch_adjoint<double, 4> : launch_adjoint<double, 6>);
    launch(ctx, dE.p, dF.p, dW.p, part.p, nb);
    hipLaunchKernelGGL(k_sum_partials, dim3(nblk((int64_t)gsize, 256)), dim3(256), 0, st, part.p, nb, (int)gsize,
                       sum.p);
    PFW_HIP(hipGetLastError());
    PFW_HIP(hipMemcpyAsync(g.data(), sum.p, gsize * sizeof(double), hipMemcpyDeviceToHost, st));
    PFW_HIP(hipStreamSynchronize(st));
    if (ace_overflowed(ad)) {  // an environment outgrew its row: once more, wider
      ace_adjoint(ctx, precision, Ebar, Fbar, Wbar, gWB, gWpair);
      return;
    }
  }
  // dL/dWB[b, z] = sum_k C[b, k] dL/dw_k(z): the transpose of the precontraction
  if (gWB) {
    for (size_t q = 0; q < (size_t)ad.nbasis * ad.nz; ++q) gWB[q] = 0;
    for (size_t t = 0; t < ad.c_vals.size(); ++t)
      for (int z = 0; z < ad.nz; ++z)
        gWB[(size_t)ad.c_rows[t] + (size_t)ad.nbasis * z] +=
            ad.c_vals[t] * g[(size_t)z * (ad.nprod + nq) + ad.c_cols[t]];
  }
  if (gWpair)
    for (int z = 0; z < ad.nz; ++z)
      for (int q = 0; q < nq; ++q) gWpair[q + (size_t)nq * z] = g[(size_t)z * (ad.nprod + nq) + ad.nprod + q];
  bool bad = false;
  for (double v : g) bad |= !std::isfinite(v);
  if (bad) throw Error(PFW_DOMAIN, "nonfinite ACE parameter gradient (coincident atoms or singular envelope)");
}

void ace_site_energies(pfw_ctx* ctx, double* out) {
  SystemBuffers& s = ctx->sys;
  if (!ctx->ace.siteE || s.n == 0) return;
  double* d = nullptr;
  PFW_HIP(hipMalloc(&d, s.n * sizeof(double)));
  hipLaunchKernelGGL(k_ace_scatter_sites, dim3(nblk(s.n, 256)), dim3(256), 0, ctx->stream, s.n, s.perm,
                     ctx->ace.siteE, d);
  PFW_HIP(hipMemcpyAsync(out, d, s.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  PFW_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(d);
}

}  // namespace pfw
