// Test infrastructure only (the oracle). CPU FP64 restatement of the ACE site
// pipeline of PotentialFrameWork.jl for resolved (schema v1) models, with the
// basis policy's species projection and order restriction:
//   radial_ed / transformed / envelope(_derivative)  src/submodules/ACE/resolved.jl:156-209
//   BlockAngular recurrence with Cartesian jets      src/submodules/ACE/angular_blocks.jl:25-81
//   site_data / site_reverse (U/V species split)     src/submodules/ACE/serial.jl:51-101
//   unrestricted_products!                           src/submodules/ACE/projection.jl:39-56
//   product_projection! (truncated (U + tV) products) src/submodules/ACE/projection.jl:7-38
//   order_allowed / pair_allowed / reference_value   src/submodules/ACE/interface.jl:122-124
//   reduce_site!                                     src/submodules/NeighbourList/execution.jl:105-122
// Environments are the directed snapshot of the neighbour-list oracle grouped
// by neighbour species in traversal order (src/submodules/NeighbourList/environments.jl:42-102).
#include <algorithm>
#include <cstdint>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "nl_oracle.h"

namespace oracle {

enum { ENV_POLY2SX = 0, ENV_POLY1SR = 1, ENV_ACE1_POLY1SR = 2 };

struct Family {
  int nr;       // radial functions per channel
  int nnodes;   // spline nodes per channel
  const double* tf;    // nch x 8: p q a rin r0 yin ycut rcut
  const int* env_kind;  // nch
  const double* env;   // nch x 5
  const double* nodes;  // nch x nnodes
  const double* coef;   // nch x [4 x nr x (nnodes-1)] (power fastest, Julia layout)
};

}  // namespace oracle

extern "C" struct OracleAceModel {
  int nz, lmax, ndens, nprod;
  const double* reference;  // nz
  oracle::Family radial, pair;
  const int* dens_n;    // ndens (0-based radial index)
  const int* dens_m;    // ndens (0-based harmonic index l*l+l+m)
  const int* prod_ptr;  // nprod + 1
  const int* prod_idx;  // 0-based density indices
  const double* weights;  // nprod x nz, column-major: C^T (WB .* mask)
  const double* wpair;    // nq x nz column-major
  int solid;              // 1: real solid harmonics r^l Y_lm ("real_solid"), 0: spherical
  // basis policy: order_lo < 0 means no order restriction (every product is
  // the full product of its densities); otherwise orders order_lo:order_hi of
  // the species projection, with selected[z] != 0 for the projected neighbour
  // species channels (project_species, interface.jl:115-124)
  const int* selected;    // nz flags, or NULL
  int order_lo, order_hi;
  int exclude;            // test aid: neighbour channel whose edges are dropped (-1: none)
};

namespace oracle {

static inline double ipow(double x, int n) {
  double v = 1;
  for (int k = 0; k < n; ++k) v *= x;
  return v;
}

// resolved.jl:156-160
static double transformed(const double* t, double r) {
  int p = (int)t[0], q = (int)t[1];
  double a = t[2], rin = t[3], r0 = t[4], yin = t[5], ycut = t[6];
  double s = (r - rin) / (r0 - rin);
  double y = r <= rin ? 1.0 : 1.0 / (1 + a * ipow(s, q) / (1 + ipow(s, q - p)));
  double x = -1 + 2 * (y - yin) / (ycut - yin);
  return std::min(std::max(x, -1.0), 1.0);
}

// resolved.jl:161-179, 193-209
static void envelope(int kind, const double* e, double r, double x, double dx, double* val, double* der) {
  if (kind == ENV_POLY2SX) {
    int p = (int)e[2], q = (int)e[3];
    if (e[0] < x && x < e[1]) {
      *val = e[4] * ipow(x - e[0], p) * ipow(e[1] - x, q);
      *der = e[4] * (p * ipow(x - e[0], p - 1) * ipow(e[1] - x, q) - q * ipow(x - e[0], p) * ipow(e[1] - x, q - 1)) * dx;
    } else {
      *val = 0;
      *der = 0;
    }
  } else if (kind == ENV_POLY1SR) {
    double rc = e[0];
    int p = (int)e[1];
    if (r >= rc) {
      *val = *der = 0;
      return;
    }
    if (!(r > 0)) throw DomainError("singular imported pair envelope at coincident atoms");
    double s = r / rc;
    *val = (1.0 / ipow(s, p) - 1) * (1 - s);
    *der = (-p / ipow(s, p + 1) * (1 - s) - (1.0 / ipow(s, p) - 1)) / rc;
  } else {
    double rc = e[0], r0 = e[1];
    int p = (int)e[2];
    if (r > rc) {
      *val = *der = 0;
      return;
    }
    if (!(r > 0)) throw DomainError("singular imported pair envelope at coincident atoms");
    double s = r / r0, sc = rc / r0;
    *val = 1.0 / ipow(s, p) - 1.0 / ipow(sc, p) + p / ipow(sc, p + 1) * (s - sc);
    *der = p / r0 * (1.0 / ipow(rc / r0, p + 1) - 1.0 / ipow(r / r0, p + 1));
  }
}

// radial_ed (serial.jl:19-42) over radial_values / envelope_derivative
// (resolved.jl:184-209) for channel ch of family f.
static void radial_ed(const Family& f, int ch, double r, double* v, double* d) {
  const double* t = f.tf + 8 * ch;
  int p = (int)t[0], q = (int)t[1];
  double a = t[2], rin = t[3], r0 = t[4], yin = t[5], ycut = t[6];
  double x = transformed(t, r);
  double s = (r - rin) / (r0 - rin);
  int h = q - p;
  double dx;
  if (r <= rin || x <= -1 || x >= 1) {
    dx = 0;
  } else {
    double den = 1 + ipow(s, h);
    double u = ipow(s, q) / den;
    double du = (q * ipow(s, q - 1) / den - (h == 0 ? 0.0 : h * ipow(s, q + h - 1) / (den * den))) / (r0 - rin);
    dx = -2 * a * du / ((1 + a * u) * (1 + a * u) * (ycut - yin));
  }
  double e, de;
  envelope(f.env_kind[ch], f.env + 5 * ch, r, x, dx, &e, &de);
  const double* nodes = f.nodes + (size_t)f.nnodes * ch;
  int k = (int)(std::upper_bound(nodes, nodes + f.nnodes, x) - nodes) - 1;
  k = std::min(std::max(k, 0), f.nnodes - 2);
  double step = nodes[k + 1] - nodes[k];
  double uu = (x - nodes[k]) / step;
  const double* c = f.coef + (size_t)ch * 4 * f.nr * (f.nnodes - 1) + (size_t)4 * f.nr * k;
  for (int n = 0; n < f.nr; ++n) {
    const double* cn = c + 4 * n;
    double val = ((cn[3] * uu + cn[2]) * uu + cn[1]) * uu + cn[0];
    double dv = (3 * cn[3] * uu * uu + 2 * cn[2] * uu + cn[1]) / step;
    v[n] = e * val;
    d[n] = de * val + e * dv * dx;
  }
}

struct Jet {
  double v, g[3];
};
static inline Jet jmul(const Jet& a, const Jet& b) {
  Jet o;
  o.v = a.v * b.v;
  for (int k = 0; k < 3; ++k) o.g[k] = a.g[k] * b.v + a.v * b.g[k];
  return o;
}
static inline Jet jadd(const Jet& a, const Jet& b) {
  Jet o{a.v + b.v, {a.g[0] + b.g[0], a.g[1] + b.g[1], a.g[2] + b.g[2]}};
  return o;
}
static inline Jet jsub(const Jet& a, const Jet& b) {
  Jet o{a.v - b.v, {a.g[0] - b.g[0], a.g[1] - b.g[1], a.g[2] - b.g[2]}};
  return o;
}
static inline Jet jscale(double s, const Jet& a) {
  Jet o{s * a.v, {s * a.g[0], s * a.g[1], s * a.g[2]}};
  return o;
}

// angular_blocks.jl:36-81 for one displacement R: real L2 Y_lm (no
// Condon-Shortley phase) in l*l+l+m order and their Cartesian gradients. With
// `solid` the same recurrence runs on the raw coordinates (angular_blocks.jl:47-48),
// giving the real solid harmonics r^l Y_lm.
static void angular(int L, const double* R, double* Y, double* dY, int solid = 0) {
  double r = std::sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  if (!(r > 0)) throw DomainError("physical zero directions are unsupported");
  double u[3] = {R[0] / r, R[1] / r, R[2] / r};
  Jet j[3];
  for (int a = 0; a < 3; ++a) {
    j[a].v = solid ? R[a] : u[a];
    for (int c = 0; c < 3; ++c)
      j[a].g[c] = solid ? (a == c ? 1.0 : 0.0) : ((a == c ? 1.0 : 0.0) - u[a] * u[c]) / r;
  }
  Jet x = j[0], y = j[1], z = j[2];
  Jet r2 = jadd(jadd(jmul(x, x), jmul(y, y)), jmul(z, z));
  std::vector<Jet> cs(L + 1), sn(L + 1), Q((L + 1) * (L + 1));
  Jet one{1, {0, 0, 0}}, zero{0, {0, 0, 0}};
  cs[0] = one;
  sn[0] = zero;
  for (int m = 1; m <= L; ++m) {
    cs[m] = jsub(jmul(cs[m - 1], x), jmul(sn[m - 1], y));
    sn[m] = jadd(jmul(sn[m - 1], x), jmul(cs[m - 1], y));
  }
  cs[0] = jscale(1.0 / std::sqrt(2.0), one);
  auto q = [&](int l, int m) -> Jet& { return Q[l * (L + 1) + m]; };
  q(0, 0) = one;
  for (int l = 0; l <= L; ++l) {
    if (l > 0) {
      q(l, l) = jscale(-(2.0 * l - 1), q(l - 1, l - 1));
      q(l, l - 1) = jscale(2.0 * l - 1, jmul(z, q(l - 1, l - 1)));
      for (int m = 0; m <= l - 2; ++m)
        q(l, m) = jscale(1.0 / (l - m), jsub(jscale(2.0 * l - 1, jmul(z, q(l - 1, m))),
                                              jscale((double)(l + m - 1), jmul(r2, q(l - 2, m)))));
    }
    double F = std::sqrt((2.0 * l + 1) / (2 * 3.14159265358979323846));
    for (int m = 0; m <= l; ++m) {
      if (m > 0) F = -F / std::sqrt((double)(l + m) * (l + 1 - m));
      Jet pos = jscale(F, jmul(q(l, m), cs[m]));
      Jet neg = jscale(F, jmul(q(l, m), sn[m]));
      int ip = l * l + l + m, im = l * l + l - m;
      Y[ip] = pos.v;
      for (int c = 0; c < 3; ++c) dY[3 * ip + c] = pos.g[c];
      if (m > 0) {
        Y[im] = neg.v;
        for (int c = 0; c < 3; ++c) dY[3 * im + c] = neg.g[c];
      }
    }
  }
}

struct Edge {
  int64_t j;
  int zj;
  double R[3];
};

struct SiteOut {
  double energy;
  std::vector<double> g;  // 3 per edge
};

// The truncated product projection of projection.jl:7-38 for one product of
// nu factors: P = sum over k in [lo, hi] of the coefficient of t^k in
// prod_s (U[a_s] + t V[a_s]), i.e. of the terms taking k factors from V.
// fwd[k][t] holds that coefficient for the first t factors; the reverse sweep
// carries the adjoint weight down the same recurrence, adding w dP/dU and
// w dP/dV into dU and dV. Returns P.
static double projected_product(const int* idx, int nu, const double* U, const double* V, int lo, int hi, double w,
                                double* dU, double* dV) {
  const int K = std::min(nu, hi);  // coefficients above the highest kept order never matter
  std::vector<double> fwd((K + 1) * (nu + 1), 0.0), rev((K + 1) * (nu + 1), 0.0);
  auto F = [&](int k, int t) -> double& { return fwd[k * (nu + 1) + t]; };
  auto G = [&](int k, int t) -> double& { return rev[k * (nu + 1) + t]; };
  F(0, 0) = 1;
  for (int t = 1; t <= nu; ++t) {
    const int a = idx[t - 1];
    for (int k = 0; k <= std::min(t, K); ++k) F(k, t) = U[a] * F(k, t - 1) + (k > 0 ? V[a] * F(k - 1, t - 1) : 0.0);
  }
  double value = 0;
  for (int k = std::max(lo, 0); k <= K; ++k) {
    value += F(k, nu);
    G(k, nu) = w;
  }
  for (int t = nu; t >= 1; --t) {
    const int a = idx[t - 1];
    for (int k = 0; k <= std::min(t, K); ++k) {
      const double g = G(k, t);
      dU[a] += g * F(k, t - 1);
      G(k, t - 1) += g * U[a];
      if (k > 0) {
        dV[a] += g * F(k - 1, t - 1);
        G(k - 1, t - 1) += g * V[a];
      }
    }
  }
  return value;
}

// Precision study: 0 keeps R, R', Y and grad Y in FP64; 1, 2 and 3 round each
// of them to bfloat16, binary16 or binary32 (nearest even) before the density
// contraction and the pullback, as an MFMA with those input types would see
// them; products, sums and everything else stay FP64.
static int g_input_rounding = 0;

static double round_to(double v, int mode) {
  if (mode == 3) return (double)(float)v;
  if (mode == 2) {  // binary16: 11 significant bits, exponent range ample for these values
    if (v == 0 || !std::isfinite(v)) return v;
    int ex;
    std::frexp(v, &ex);
    const double q = std::ldexp(1.0, ex - 11);
    return std::nearbyint(v / q) * q;
  }
  if (mode == 1) {  // bfloat16: FP32 with the low 16 mantissa bits rounded off (nearest even)
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
  }
  return v;
}

// site_data / site_reverse (serial.jl:51-101) with cpu.jl:214-236 for the
// readout and gradients.
static void site(const OracleAceModel& m, int iz, const std::vector<Edge>& edges, bool grads, SiteOut& out) {
  int ne = (int)edges.size(), nr = m.radial.nr, nq = m.pair.nr, ny = (m.lmax + 1) * (m.lmax + 1);
  std::vector<double> rs(ne), R(ne * nr), dR(ne * nr), Q(ne * nq), dQ(ne * nq), Y(ne * ny), dY(ne * ny * 3);
  for (int e = 0; e < ne; ++e) {
    const double* d = edges[e].R;
    rs[e] = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(rs[e] > 0)) throw DomainError("coincident physical atoms have undefined imported harmonics/pair envelope");
    int ch = iz + m.nz * edges[e].zj;
    radial_ed(m.radial, ch, rs[e], &R[e * nr], &dR[e * nr]);
    if (nq) radial_ed(m.pair, ch, rs[e], &Q[e * nq], &dQ[e * nq]);
    angular(m.lmax, d, &Y[e * ny], &dY[e * ny * 3], m.solid);
    if (g_input_rounding) {
      const int md = g_input_rounding;
      for (int k = 0; k < nr; ++k) {
        R[e * nr + k] = round_to(R[e * nr + k], md);
        dR[e * nr + k] = round_to(dR[e * nr + k], md);
      }
      for (int k = 0; k < ny; ++k) Y[e * ny + k] = round_to(Y[e * ny + k], md);
      for (int k = 0; k < 3 * ny; ++k) dY[e * ny * 3 + k] = round_to(dY[e * ny * 3 + k], md);
    }
  }
  // restricted orders: densities split by neighbour species into U (others)
  // and V (projected species); the pair term of an edge counts as order 1
  // when its neighbour is projected, else 0; the reference energy as order 0
  const bool proj = m.order_lo >= 0;
  auto order_allowed = [&](int k) { return !proj || (m.order_lo <= k && k <= m.order_hi); };
  auto selected = [&](int z) { return proj && m.selected && m.selected[z] != 0; };
  std::vector<double> A(m.ndens, 0.0), Vd(m.ndens, 0.0);
  for (int a = 0; a < m.ndens; ++a)
    for (int e = 0; e < ne; ++e) {
      const double v = R[e * nr + m.dens_n[a]] * Y[e * ny + m.dens_m[a]];
      if (selected(edges[e].zj))
        Vd[a] += v;
      else
        A[a] += v;
    }
  const double* w = m.weights + (size_t)m.nprod * iz;
  double energy = order_allowed(0) ? m.reference[iz] : 0.0;
  std::vector<double> dA(m.ndens, 0.0), dV(m.ndens, 0.0);
  for (int k = 0; k < m.nprod; ++k) {
    int b = m.prod_ptr[k], nu = m.prod_ptr[k + 1] - b;
    if (proj) {
      energy += w[k] * projected_product(m.prod_idx + b, nu, A.data(), Vd.data(), m.order_lo, m.order_hi, w[k],
                                         dA.data(), dV.data());
      continue;
    }
    double pre[8];
    pre[0] = 1;
    for (int t = 0; t < nu; ++t) pre[t + 1] = pre[t] * A[m.prod_idx[b + t]];
    energy += w[k] * pre[nu];
    double adj = w[k];
    for (int t = nu - 1; t >= 0; --t) {
      int a = m.prod_idx[b + t];
      dA[a] += adj * pre[t];
      adj *= A[a];
    }
  }
  const double* wp = m.wpair + (size_t)nq * iz;
  for (int e = 0; e < ne; ++e)
    if (order_allowed(selected(edges[e].zj) ? 1 : 0))
      for (int n = 0; n < nq; ++n) energy += Q[e * nq + n] * wp[n];
  out.energy = energy;
  if (!grads) return;
  out.g.assign(3 * ne, 0.0);
  for (int e = 0; e < ne; ++e) {
    double dir[3] = {edges[e].R[0] / rs[e], edges[e].R[1] / rs[e], edges[e].R[2] / rs[e]};
    double g[3] = {0, 0, 0};
    const bool sel = selected(edges[e].zj);
    const std::vector<double>& adj = sel ? dV : dA;
    for (int a = 0; a < m.ndens; ++a) {
      int n = m.dens_n[a], mm = m.dens_m[a];
      for (int c = 0; c < 3; ++c)
        g[c] += adj[a] * (dR[e * nr + n] * Y[e * ny + mm] * dir[c] + R[e * nr + n] * dY[(e * ny + mm) * 3 + c]);
    }
    double sq = 0;
    if (order_allowed(sel ? 1 : 0))
      for (int n = 0; n < nq; ++n) sq += wp[n] * dQ[e * nq + n];
    for (int c = 0; c < 3; ++c) out.g[3 * e + c] = g[c] + sq * dir[c];
  }
}

}  // namespace oracle

using namespace oracle;

// Sets the precision-study rounding of site() (0 none, 1 bfloat16, 2 binary16,
// 3 binary32); not thread-safe against a running evaluation.
extern "C" void oracle_ace_set_input_rounding(int mode) { g_input_rounding = (mode >= 0 && mode <= 3) ? mode : 0; }

static int ace_fail(char* err, int errlen, const char* msg, int code) {
  if (err && errlen > 0) {
    std::strncpy(err, msg, errlen - 1);
    err[errlen - 1] = 0;
  }
  return code;
}

extern "C" {

// Greedy site colouring of color_environment! (NeighbourList/environments.jl:78-95):
// a site's footprint is itself and every neighbour it writes; each site takes
// the smallest colour not yet used by a site whose footprint shares an atom,
// so the sites of one colour (phase) write disjoint force entries.
static std::vector<std::vector<int64_t>> colour_sites(int64_t n, const std::vector<std::vector<Edge>>& env) {
  std::vector<std::vector<int>> incidence(n);  // colours touching each atom
  std::vector<std::vector<int64_t>> phases;
  std::vector<int64_t> fp;
  std::vector<char> forbidden;
  for (int64_t i = 0; i < n; ++i) {
    fp.assign(1, i);
    for (const Edge& e : env[i]) fp.push_back(e.j);
    std::sort(fp.begin(), fp.end());
    fp.erase(std::unique(fp.begin(), fp.end()), fp.end());
    forbidden.assign(phases.size() + 1, 0);
    for (int64_t a : fp)
      for (int c : incidence[a]) forbidden[c] = 1;
    int c = 0;
    while (forbidden[c]) ++c;
    if (c == (int)phases.size()) phases.emplace_back();
    phases[c].push_back(i);
    for (int64_t a : fp) incidence[a].push_back(c);
  }
  return phases;
}

// E, F (3 x n original order), W (3 x 3 column-major), site energies (n) of a
// resolved ACE model; cutoff is the model's search cutoff. threads > 1 splits
// sites over threads: with coloured = 0 each thread keeps private force and
// virial buffers; with coloured = 1 the sites run phase by phase over the
// greedy colouring (execute_sites!, NeighbourList/execution.jl:153-167), the
// threads of a phase taking every threads-th site and writing one shared
// force array. Returns 0, 1 (ArgumentError) or 3 (DomainError).
int oracle_ace_evaluate_scheduled(int64_t n, const double* X, const int32_t* species, const double* H,
                                  const uint8_t* pbc, double cutoff, const OracleAceModel* m, int threads,
                                  int coloured, unsigned mask, double* E, double* F, double* W, double* siteE,
                                  char* err, int errlen);

int oracle_ace_evaluate(int64_t n, const double* X, const int32_t* species, const double* H, const uint8_t* pbc,
                        double cutoff, const OracleAceModel* m, int threads, unsigned mask, double* E, double* F,
                        double* W, double* siteE, char* err, int errlen) {
  return oracle_ace_evaluate_scheduled(n, X, species, H, pbc, cutoff, m, threads, 0, mask, E, F, W, siteE, err,
                                       errlen);
}

int oracle_ace_evaluate_scheduled(int64_t n, const double* X, const int32_t* species, const double* H,
                                  const uint8_t* pbc, double cutoff, const OracleAceModel* m, int threads,
                                  int coloured, unsigned mask, double* E, double* F, double* W, double* siteE,
                                  char* err, int errlen) {
  try {
    bool pb[3] = {pbc[0] != 0, pbc[1] != 0, pbc[2] != 0};
    Prepared w;
    prepare(w, n, X, H, pb, cutoff, 32, false);
    std::vector<DirectedPair> pairs;
    directed_pairs(w, X, H, pairs);
    for (int64_t i = 0; i < n; ++i)
      if (species[i] < 0 || species[i] >= m->nz) throw ArgumentError("invalid species channel");
    // per-centre environments in traversal order, grouped by neighbour species (stable)
    std::vector<std::vector<Edge>> env(n);
    for (auto& p : pairs)
      if (species[p.j] != m->exclude) env[p.i].push_back(Edge{p.j, species[p.j], {p.R[0], p.R[1], p.R[2]}});
    for (auto& e : env)
      std::stable_sort(e.begin(), e.end(), [](const Edge& a, const Edge& b) { return a.zj < b.zj; });
    bool grads = (mask & 6) != 0;
    int T = std::max(1, threads);
    if (coloured && T > 1) {
      const std::vector<std::vector<int64_t>> phases = colour_sites(n, env);
      std::vector<double> Fs(grads ? 3 * n : 0, 0.0);
      std::vector<std::array<double, 9>> Wb(T);
      for (auto& w : Wb) w.fill(0.0);
      std::vector<double> es(n, 0.0);
      std::vector<std::string> errs(T);
      for (const auto& phase : phases) {
        auto work = [&](int t) {
          try {
            SiteOut so;
            for (size_t k = t; k < phase.size(); k += T) {
              const int64_t i = phase[k];
              site(*m, species[i], env[i], grads, so);
              es[i] = so.energy;
              if (!grads) continue;
              for (size_t e = 0; e < env[i].size(); ++e) {
                const double* g = &so.g[3 * e];
                const int64_t j = env[i][e].j;
                for (int a = 0; a < 3; ++a) {
                  Fs[3 * i + a] += g[a];
                  Fs[3 * j + a] -= g[a];
                  for (int b = 0; b < 3; ++b) Wb[t][a + 3 * b] -= env[i][e].R[a] * g[b];
                }
              }
            }
          } catch (const std::exception& e) {
            errs[t] = e.what();
          }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < T; ++t) pool.emplace_back(work, t);
        for (auto& th : pool) th.join();
        for (auto& s_ : errs)
          if (!s_.empty()) throw DomainError(s_);
      }
      double e = 0;
      for (double v : es) e += v;
      if (E) *E = (mask & 1) ? e : 0.0;
      if (siteE)
        for (int64_t i = 0; i < n; ++i) siteE[i] = es[i];
      if (F)
        for (int64_t k = 0; k < 3 * n; ++k) F[k] = (mask & 2) ? Fs[k] : 0.0;
      if (W)
        for (int k = 0; k < 9; ++k) {
          double s_ = 0;
          if (mask & 4)
            for (int t = 0; t < T; ++t) s_ += Wb[t][k];
          W[k] = s_;
        }
      return 0;
    }
    std::vector<std::vector<double>> Fb(T, std::vector<double>(grads ? 3 * n : 0, 0.0));
    std::vector<std::array<double, 9>> Wb(T);
    std::vector<double> es(n, 0.0);
    std::vector<std::string> errs(T);
    auto work = [&](int t) {
      try {
        for (auto& v : Wb[t]) v = 0;
        SiteOut so;
        for (int64_t i = t; i < n; i += T) {
          site(*m, species[i], env[i], grads, so);
          es[i] = so.energy;
          if (!grads) continue;
          for (size_t e = 0; e < env[i].size(); ++e) {
            const double* g = &so.g[3 * e];
            int64_t j = env[i][e].j;
            for (int a = 0; a < 3; ++a) {
              Fb[t][3 * i + a] += g[a];
              Fb[t][3 * j + a] -= g[a];
              for (int b = 0; b < 3; ++b) Wb[t][a + 3 * b] -= env[i][e].R[a] * g[b];
            }
          }
        }
      } catch (const std::exception& e) {
        errs[t] = e.what();
      }
    };
    if (T == 1) {
      work(0);
    } else {
      std::vector<std::thread> pool;
      for (int t = 0; t < T; ++t) pool.emplace_back(work, t);
      for (auto& th : pool) th.join();
    }
    for (auto& s : errs)
      if (!s.empty()) throw DomainError(s);
    double e = 0;
    for (double v : es) e += v;
    if (E) *E = (mask & 1) ? e : 0.0;
    if (siteE)
      for (int64_t i = 0; i < n; ++i) siteE[i] = es[i];
    if (F) {
      for (int64_t k = 0; k < 3 * n; ++k) {
        double s = 0;
        if (mask & 2)
          for (int t = 0; t < T; ++t) s += Fb[t][k];
        F[k] = s;
      }
    }
    if (W)
      for (int k = 0; k < 9; ++k) {
        double s = 0;
        if (mask & 4)
          for (int t = 0; t < T; ++t) s += Wb[t][k];
        W[k] = s;
      }
    return 0;
  } catch (const ArgumentError& e) {
    return ace_fail(err, errlen, e.what(), 1);
  } catch (const DomainError& e) {
    return ace_fail(err, errlen, e.what(), 3);
  }
}

}  // extern "C"

// The projected product of one product (P, with w dP/dU and w dP/dV added into
// dU and dV) for tests against a brute-force enumeration.
extern "C" double oracle_projected_product(const int* idx, int nu, const double* U, const double* V, int lo, int hi,
                                           double w, double* dU, double* dV) {
  return oracle::projected_product(idx, nu, U, V, lo, hi, w, dU, dV);
}

extern "C" int oracle_ylm(int L, int solid, const double* R, double* Y, double* dY) {
  try {
    oracle::angular(L, R, Y, dY, solid);
    return 0;
  } catch (const std::exception&) {
    return 3;
  }
}
