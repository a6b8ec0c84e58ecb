// Test infrastructure only (the oracle). CPU FP64 restatement of the
// NeighbourList preparation of PotentialFrameWork.jl. Built with
// -ffp-contract=off so every product and sum rounds separately, as Julia does.
#include "nl_oracle.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <set>

namespace oracle {

double jl_mod(double x, double y) {
  if (std::isinf(y) && std::isfinite(x)) return x;
  double r = std::fmod(x, y);
  if (r == 0) return std::copysign(0.0, y);
  if ((r > 0) != (y > 0)) return r + y;
  return r;
}
int64_t jl_fld(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
  return q;
}
int64_t jl_imod(int64_t a, int64_t b) { return a - jl_fld(a, b) * b; }

static inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// storage.jl:83-108
void periodic_basis(const double* H, const bool* pb, double* B) {
  std::vector<std::array<double, 3>> orth;
  for (int i = 0; i < 9; ++i) B[i] = 0;
  for (int a = 0; a < 3; ++a) {
    if (!pb[a]) continue;
    double v[3] = {H[3 * a], H[3 * a + 1], H[3 * a + 2]};
    for (double t : v)
      if (!std::isfinite(t)) throw ArgumentError("periodic cell vectors must be finite");
    for (int k = 0; k < 3; ++k) B[3 * a + k] = v[k];
    for (auto& q : orth) {
      double s = dot3(q.data(), v);
      for (int k = 0; k < 3; ++k) v[k] -= s * q[k];
    }
    double nrm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(std::isfinite(nrm) && nrm > 0))
      throw ArgumentError("periodic cell vectors must be linearly independent");
    orth.push_back({v[0] / nrm, v[1] / nrm, v[2] / nrm});
  }
  for (int a = 0; a < 3; ++a) {
    if (pb[a]) continue;
    std::array<double, 3> best{};
    double bestn = -1;
    for (int k = 0; k < 3; ++k) {
      double v[3] = {0, 0, 0};
      v[k] = 1;
      for (auto& q : orth) {
        double s = dot3(q.data(), v);
        for (int m = 0; m < 3; ++m) v[m] -= s * q[m];
      }
      double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      if (n2 > bestn) {  // first maximum wins, as argmax
        bestn = n2;
        best = {v[0], v[1], v[2]};
      }
    }
    double nrm = std::sqrt(best[0] * best[0] + best[1] * best[1] + best[2] * best[2]);
    for (int m = 0; m < 3; ++m) best[m] /= nrm;
    for (int m = 0; m < 3; ++m) B[3 * a + m] = best[m];
    orth.push_back(best);
  }
}

static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

void inv3(const double* A, double* out) {
  double x0[3] = {A[0], A[1], A[2]}, x1[3] = {A[3], A[4], A[5]}, x2[3] = {A[6], A[7], A[8]};
  double y0[3], y1[3], y2[3];
  cross3(x1, x2, y0);
  double d = dot3(x0, y0);
  for (int k = 0; k < 3; ++k) {
    x0[k] /= d;
    y0[k] /= d;
  }
  cross3(x2, x0, y1);
  cross3(x0, x1, y2);
  // rows of the inverse are y0, y1, y2 (column-major storage)
  for (int col = 0; col < 3; ++col) {
    out[3 * col + 0] = y0[col];
    out[3 * col + 1] = y1[col];
    out[3 * col + 2] = y2[col];
  }
}

static inline void matvec(const double* M, const double* v, double* o) {
  for (int a = 0; a < 3; ++a) o[a] = M[a] * v[0] + M[a + 3] * v[1] + M[a + 6] * v[2];
}

static void pair_footprint(Prepared& w, int i, const std::vector<CellImage>& entries) {
  std::set<int> nb;
  for (auto& e : entries) nb.insert(e.cell);
  w.adjacency.emplace_back(nb.begin(), nb.end());
  nb.insert(i);
  w.footprints.emplace_back(nb.begin(), nb.end());
}

// storage.jl:244-261
static void minimum_topology(Prepared& w, const std::map<std::array<int64_t, 3>, int>& lookup,
                             bool query_only) {
  for (size_t i = 0; i < w.keys.size(); ++i) {
    auto& entries = w.images[i];
    if (w.brute) {
      entries.push_back(CellImage{(int)i, {0, 0, 0}, false, {0, 0, 0}});
    } else {
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            int64_t off[3] = {dx, dy, dz};
            std::array<int64_t, 3> other;
            for (int a = 0; a < 3; ++a)
              other[a] = w.pb[a] ? jl_imod(w.keys[i][a] + off[a], w.dims[a]) : w.keys[i][a] + off[a];
            auto it = lookup.find(other);
            int j = it == lookup.end() ? -1 : it->second;
            if (j >= (int)i) entries.push_back(CellImage{j, {0, 0, 0}, false, {0, 0, 0}});
          }
    }
    std::stable_sort(entries.begin(), entries.end(),
                     [](const CellImage& a, const CellImage& b) { return a.cell < b.cell; });
    entries.erase(std::unique(entries.begin(), entries.end(),
                              [](const CellImage& a, const CellImage& b) { return a.cell == b.cell; }),
                  entries.end());
    if (!query_only) pair_footprint(w, (int)i, entries);
  }
}

static bool lex_positive(const int64_t* s) {
  for (int a = 0; a < 3; ++a) {
    if (s[a] > 0) return true;
    if (s[a] < 0) return false;
  }
  return false;
}

// storage.jl:263-283
static void image_topology(Prepared& w, const std::map<std::array<int64_t, 3>, int>& lookup,
                           bool query_only) {
  int64_t radius[3];
  for (int a = 0; a < 3; ++a)
    radius[a] = w.brute ? (w.pb[a] ? (int64_t)std::ceil(w.reach[a]) : 0)
                        : std::max<int64_t>(1, (int64_t)std::ceil(w.reach[a] / w.sizecell[a]));
  for (size_t i = 0; i < w.keys.size(); ++i) {
    auto& entries = w.images[i];
    for (int64_t z = -radius[2]; z <= radius[2]; ++z)
      for (int64_t y = -radius[1]; y <= radius[1]; ++y)
        for (int64_t x = -radius[0]; x <= radius[0]; ++x) {
          int64_t off[3] = {x, y, z};
          std::array<int64_t, 3> other;
          for (int a = 0; a < 3; ++a)
            other[a] = w.pb[a] ? jl_imod(w.keys[i][a] + off[a], w.dims[a]) : w.keys[i][a] + off[a];
          int j;
          if (w.brute) {
            j = (int)i;
          } else {
            auto it = lookup.find(other);
            j = it == lookup.end() ? -1 : it->second;
          }
          if (j < (int)i) continue;
          CellImage e;
          e.cell = j;
          for (int a = 0; a < 3; ++a)
            e.lattice[a] = w.pb[a] ? (w.brute ? off[a] : jl_fld(w.keys[i][a] + off[a], w.dims[a])) : 0;
          double s[3] = {(double)e.lattice[0], (double)e.lattice[1], (double)e.lattice[2]};
          matvec(w.basis, s, e.shift);
          e.self = lex_positive(e.lattice);
          entries.push_back(e);
        }
    if (!query_only) pair_footprint(w, (int)i, entries);
  }
}

// storage.jl:291-306, greedy colouring of cell write footprints.
static void schedule(Prepared& w) {
  std::vector<std::set<int>> incidence(w.keys.size());
  w.phases.clear();
  for (size_t task = 0; task < w.keys.size(); ++task) {
    std::set<int> forbidden;
    for (int cell : w.footprints[task]) forbidden.insert(incidence[cell].begin(), incidence[cell].end());
    int color = 0;
    while (forbidden.count(color)) ++color;
    if (color >= (int)w.phases.size()) w.phases.emplace_back();
    w.phases[color].push_back((int)task);
    for (int cell : w.footprints[task]) incidence[cell].insert(color);
  }
}

// storage.jl:130-241
void prepare(Prepared& w, int64_t n, const double* X, const double* H, const bool* pbin,
             double cutoff, int bruteforce_threshold, bool do_schedule) {
  w = Prepared();
  double c = cutoff;
  if (!(std::isfinite(c) && c > 0)) throw ArgumentError("cutoff must be finite and positive (Å)");
  bool pb[3] = {pbin[0], pbin[1], pbin[2]};
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
  w.n = n;
  w.c = c;
  w.minimum_image = mi;
  for (int a = 0; a < 3; ++a) {
    w.pb[a] = pb[a];
    w.L[a] = L[a];
  }
  if (mi) {
    for (int k = 0; k < 9; ++k) w.basis[k] = w.recip[k] = (k % 4 == 0) ? 1.0 : 0.0;
  } else {
    periodic_basis(H, pb, w.basis);
    inv3(w.basis, w.recip);
    for (double v : w.recip)
      if (!std::isfinite(v)) throw ArgumentError("periodic cell vectors must be linearly independent");
  }
  std::vector<double> coords[3], frac[3];
  for (int a = 0; a < 3; ++a) coords[a].resize(n);
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      double v = X[3 * i + a];
      if (!std::isfinite(v)) throw ArgumentError("positions must be finite");
      coords[a][i] = (mi && pb[a]) ? jl_mod(v, L[a]) : v;
    }
  if (!mi) {
    for (int a = 0; a < 3; ++a) frac[a].resize(n);
    for (int64_t i = 0; i < n; ++i) {
      double v[3] = {coords[0][i], coords[1][i], coords[2][i]}, f[3], r[3];
      matvec(w.recip, v, f);
      for (int a = 0; a < 3; ++a)
        if (pb[a]) f[a] = jl_mod(f[a], 1.0);
      matvec(w.basis, f, r);
      for (int a = 0; a < 3; ++a) {
        coords[a][i] = r[a];
        frac[a][i] = f[a];
      }
    }
  }
  w.brute = n <= bruteforce_threshold;
  std::vector<double>* bin = mi ? coords : frac;
  for (int a = 0; a < 3; ++a) {
    if (mi) {
      w.reach[a] = c;
      w.extent[a] = L[a];
    } else {
      double r0 = w.recip[a], r1 = w.recip[a + 3], r2 = w.recip[a + 6];
      w.reach[a] = c * std::sqrt(r0 * r0 + r1 * r1 + r2 * r2);
      w.extent[a] = 1.0;
    }
    if (pb[a] || n == 0) {
      w.origin[a] = 0;
    } else {
      w.origin[a] = *std::min_element(bin[a].begin(), bin[a].end());
    }
    w.dims[a] = pb[a] ? std::max<int64_t>(1, (int64_t)std::floor(w.extent[a] / w.reach[a])) : 1;
    w.sizecell[a] = pb[a] ? w.extent[a] / (double)w.dims[a] : w.reach[a];
  }
  std::vector<std::array<int64_t, 3>> atomkeys(n);
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      if (w.brute) {
        atomkeys[i][a] = 0;
      } else if (pb[a]) {
        int64_t k = (int64_t)std::floor(bin[a][i] / w.sizecell[a]);
        atomkeys[i][a] = std::min(std::max<int64_t>(k, 0), w.dims[a] - 1);
      } else {
        atomkeys[i][a] = (int64_t)std::floor((bin[a][i] - w.origin[a]) / w.sizecell[a]);
      }
    }
  w.p.resize(n);
  std::iota(w.p.begin(), w.p.end(), 0);
  std::stable_sort(w.p.begin(), w.p.end(),
                   [&](int64_t a, int64_t b) { return atomkeys[a] < atomkeys[b]; });
  w.inv.resize(n);
  w.x.resize(n);
  w.y.resize(n);
  w.z.resize(n);
  for (int64_t k = 0; k < n; ++k) {
    int64_t i = w.p[k];
    w.inv[i] = k;
    w.x[k] = coords[0][i];
    w.y[k] = coords[1][i];
    w.z[k] = coords[2][i];
    if (w.keys.empty() || atomkeys[i] != w.keys.back()) {
      w.keys.push_back(atomkeys[i]);
      w.ranges.push_back({k, k});
    } else {
      w.ranges.back().second = k;
    }
  }
  w.images.assign(w.keys.size(), {});
  std::map<std::array<int64_t, 3>, int> lookup;
  for (size_t i = 0; i < w.keys.size(); ++i) lookup[w.keys[i]] = (int)i;
  if (mi)
    minimum_topology(w, lookup, !do_schedule);
  else
    image_topology(w, lookup, !do_schedule);
  if (do_schedule) schedule(w);
}

// image_windings! (interface.jl:100-114)
static void windings(const Prepared& w, const double* X, const double* H,
                     std::vector<std::array<int64_t, 3>>& out) {
  double recip[9];
  if (!w.minimum_image) {
    double B[9];
    periodic_basis(H, w.pb, B);
    inv3(B, recip);
  }
  out.resize(w.n);
  for (int64_t i = 0; i < w.n; ++i) {
    double r[3] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]}, f[3];
    if (w.minimum_image) {
      f[0] = r[0];
      f[1] = r[1];
      f[2] = r[2];
    } else {
      matvec(recip, r, f);
    }
    for (int a = 0; a < 3; ++a)
      out[i][a] = !w.pb[a] ? 0 : (int64_t)std::floor(w.minimum_image ? f[a] / w.L[a] : f[a]);
  }
}

// geometry.jl:5-19 for one lane.
static inline bool lane_geometry(const Prepared& w, int64_t i, int64_t j, const CellImage& im,
                                 double* d) {
  double raw[3] = {w.x[j] - w.x[i], w.y[j] - w.y[i], w.z[j] - w.z[i]};
  for (int a = 0; a < 3; ++a) {
    if (w.minimum_image)
      d[a] = w.pb[a] ? raw[a] - std::nearbyint(raw[a] / w.L[a]) * w.L[a] : raw[a];
    else
      d[a] = raw[a] + im.shift[a];
  }
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < w.c * w.c;
}

void directed_pairs(const Prepared& w, const double* X, const double* H,
                    std::vector<DirectedPair>& out) {
  std::vector<std::array<int64_t, 3>> wind;
  windings(w, X, H, wind);
  out.clear();
  for (size_t task = 0; task < w.ranges.size(); ++task)
    for (int64_t i = w.ranges[task].first; i <= w.ranges[task].second; ++i)
      for (const CellImage& im : w.images[task]) {
        auto r = w.ranges[im.cell];
        for (int64_t j = std::max(r.first, i + (im.self ? 0 : 1)); j <= r.second; ++j) {
          double d[3];
          if (!lane_geometry(w, i, j, im, d)) continue;
          int64_t a = w.p[i], b = w.p[j];
          int64_t lat[3];
          if (w.minimum_image) {
            double raw[3] = {w.x[j] - w.x[i], w.y[j] - w.y[i], w.z[j] - w.z[i]};
            for (int ax = 0; ax < 3; ++ax)
              lat[ax] = w.pb[ax] ? -(int64_t)std::nearbyint(raw[ax] / w.L[ax]) : 0;
          } else {
            for (int ax = 0; ax < 3; ++ax) lat[ax] = im.lattice[ax];
          }
          DirectedPair f, g;
          f.i = a;
          f.j = b;
          g.i = b;
          g.j = a;
          for (int ax = 0; ax < 3; ++ax) {
            f.S[ax] = lat[ax] + wind[a][ax] - wind[b][ax];
            g.S[ax] = -f.S[ax];
            f.R[ax] = d[ax];
            g.R[ax] = -d[ax];
          }
          out.push_back(f);
          out.push_back(g);
        }
      }
}

}  // namespace oracle
