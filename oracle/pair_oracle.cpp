// Test infrastructure only (the oracle). CPU FP64 restatement of the
// PotentialFrameWork.jl pair executor (half list, phase-coloured cell tasks)
// and of the SmoothPair / energy-shifted Lennard-Jones callbacks. Exposed to
// the tests and to the CPU baseline through a small C interface.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

#include "nl_oracle.h"

namespace oracle {

enum { KIND_SMOOTH = 0, KIND_LJ = 1 };
enum { WANT_E = 1, WANT_F = 2, WANT_V = 4 };

struct PairModel {
  int kind;
  int nz;
  const double* params;  // SMOOTH: nz*nz amplitudes (column-major); LJ: epsilon, sigma
  double radius;
};

// models.jl:153-167 (PreparedSmooth) and lennard_jones.jl:5-17 + models.jl:128-136
// (PreparedRadial). Returns energy and dE/dr for distance r.
static inline void pair_callback(const PairModel& m, int zi, int zj, double r, double& e,
                                 double& dedr) {
  if (m.kind == KIND_SMOOTH) {
    double a = m.params[zi + m.nz * zj];
    double invc2 = 1.0 / (m.radius * m.radius);
    double q = 1 - r * r * invc2;
    e = a * q * q * q;
    dedr = -6 * a * q * q * invc2 * r;
  } else {
    double eps = m.params[0], s = m.params[1], c = m.radius;
    if (!(r > 0)) throw DomainError("singular Lennard-Jones origin at coincident atoms");
    double ratio = s / r, ratio2 = ratio * ratio, q = ratio2 * ratio2 * ratio2;
    double rc = s / c, qc2 = rc * rc, qc = qc2 * qc2 * qc2;
    e = 4.0 * eps * (q * q - q - qc * qc + qc);
    dedr = 24.0 * eps * (q - 2.0 * q * q) / r;
  }
}

struct Outputs {
  std::vector<double> force;  // 3 x n, sorted order
  std::vector<double> energies, virials;
};

// execution.jl:173-217 with W = 1.
static void cell_task(const Prepared& w, const std::vector<int>& sp, const PairModel& m,
                      unsigned mask, size_t task, Outputs& o) {
  double e = 0, vir[9] = {0};
  for (int64_t i = w.ranges[task].first; i <= w.ranges[task].second; ++i) {
    double central[3] = {0, 0, 0};
    for (const CellImage& im : w.images[task]) {
      auto r = w.ranges[im.cell];
      for (int64_t j = std::max(r.first, i + (im.self ? 0 : 1)); j <= r.second; ++j) {
        double d[3];
        double raw[3] = {w.x[j] - w.x[i], w.y[j] - w.y[i], w.z[j] - w.z[i]};
        for (int a = 0; a < 3; ++a) {
          if (w.minimum_image)
            d[a] = w.pb[a] ? raw[a] - std::nearbyint(raw[a] / w.L[a]) * w.L[a] : raw[a];
          else
            d[a] = raw[a] + im.shift[a];
        }
        if (!(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < w.c * w.c)) continue;
        double rr = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double en, dedr;
        pair_callback(m, sp[w.p[i]], sp[w.p[j]], rr, en, dedr);
        if (mask & WANT_E) e += en;
        if (mask & (WANT_F | WANT_V)) {
          double coef = dedr / (rr > 0 ? rr : 1.0);
          double g[3] = {coef * d[0], coef * d[1], coef * d[2]};
          if (mask & WANT_F) {
            for (int a = 0; a < 3; ++a) {
              central[a] += g[a];
              o.force[3 * j + a] -= g[a];
            }
          }
          if (mask & WANT_V)
            for (int b = 0; b < 3; ++b)
              for (int a = 0; a < 3; ++a) vir[a + 3 * b] -= d[a] * g[b];
        }
      }
    }
    if (mask & WANT_F)
      for (int a = 0; a < 3; ++a) o.force[3 * i + a] += central[a];
  }
  o.energies[task] = e;
  for (int k = 0; k < 9; ++k) o.virials[9 * task + k] = vir[k];
}

class Barrier {
 public:
  explicit Barrier(int n) : n_(n) {}
  void wait() {
    std::unique_lock<std::mutex> lk(mu_);
    int gen = gen_;
    if (++count_ == n_) {
      count_ = 0;
      ++gen_;
      cv_.notify_all();
    } else {
      cv_.wait(lk, [&] { return gen != gen_; });
    }
  }

 private:
  std::mutex mu_;
  std::condition_variable cv_;
  int n_, count_ = 0, gen_ = 0;
};

// execution.jl:223-238: phases are barriers, tasks inside a phase run concurrently.
static void execute(const Prepared& w, const std::vector<int>& sp, const PairModel& m,
                    unsigned mask, int threads, Outputs& o) {
  size_t ntask = w.ranges.size();
  o.force.assign(3 * w.n, 0.0);
  o.energies.assign(ntask, 0.0);
  o.virials.assign(9 * ntask, 0.0);
  if (mask == 0) return;
  if (threads <= 1) {
    for (size_t t = 0; t < ntask; ++t) cell_task(w, sp, m, mask, t, o);
    return;
  }
  Barrier bar(threads);
  std::vector<std::atomic<size_t>> next(w.phases.size());
  for (auto& a : next) a = 0;
  std::mutex err_mu;
  std::string err;
  auto worker = [&]() {
    for (size_t ph = 0; ph < w.phases.size(); ++ph) {
      const auto& tasks = w.phases[ph];
      for (;;) {
        size_t k = next[ph].fetch_add(1);
        if (k >= tasks.size()) break;
        try {
          cell_task(w, sp, m, mask, tasks[k], o);
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> lk(err_mu);
          err = e.what();
        }
      }
      bar.wait();
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  if (!err.empty()) throw DomainError(err);
}

}  // namespace oracle

using namespace oracle;

static int fail(char* err, int errlen, const char* msg, int code) {
  if (err && errlen > 0) {
    std::strncpy(err, msg, errlen - 1);
    err[errlen - 1] = 0;
  }
  return code;
}

extern "C" {

// Returns 0 on success, 1 ArgumentError, 3 DomainError.
int oracle_pair_evaluate(int64_t n, const double* X, const int32_t* species, const double* H,
                         const uint8_t* pbc, int kind, int nz, const double* params,
                         double cutoff, int bf_threshold, int threads, unsigned mask, double* E,
                         double* F, double* W, char* err, int errlen) {
  try {
    bool pb[3] = {pbc[0] != 0, pbc[1] != 0, pbc[2] != 0};
    Prepared w;
    prepare(w, n, X, H, pb, cutoff, bf_threshold, true);
    std::vector<int> sp(species, species + n);
    for (int s : sp)
      if (s < 0 || s >= nz) throw ArgumentError("potential has no parameters for species");
    PairModel m{kind, nz, params, cutoff};
    Outputs o;
    execute(w, sp, m, mask, threads, o);
    double e = 0;
    for (double v : o.energies) e += v;
    if (E) *E = e;
    if (W) {
      for (int k = 0; k < 9; ++k) W[k] = 0;
      for (size_t t = 0; t < o.energies.size(); ++t)
        for (int k = 0; k < 9; ++k) W[k] += o.virials[9 * t + k];
    }
    if (F)
      for (int64_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) F[3 * w.p[k] + a] = o.force[3 * k + a];
    return 0;
  } catch (const ArgumentError& e) {
    return fail(err, errlen, e.what(), 1);
  } catch (const DomainError& e) {
    return fail(err, errlen, e.what(), 3);
  }
}

struct NLHandle {
  Prepared w;
  std::vector<DirectedPair> pairs;
};

void* oracle_nl_build(int64_t n, const double* X, const double* H, const uint8_t* pbc,
                      double cutoff, int bf_threshold, char* err, int errlen) {
  try {
    auto* h = new NLHandle;
    bool pb[3] = {pbc[0] != 0, pbc[1] != 0, pbc[2] != 0};
    prepare(h->w, n, X, H, pb, cutoff, bf_threshold, false);
    directed_pairs(h->w, X, H, h->pairs);
    return h;
  } catch (const std::exception& e) {
    fail(err, errlen, e.what(), 1);
    return nullptr;
  }
}
int64_t oracle_nl_count(void* h) { return (int64_t)static_cast<NLHandle*>(h)->pairs.size(); }
void oracle_nl_copy(void* hv, int64_t* i, int64_t* j, int64_t* S, double* R) {
  auto* h = static_cast<NLHandle*>(hv);
  for (size_t k = 0; k < h->pairs.size(); ++k) {
    const auto& p = h->pairs[k];
    i[k] = p.i;
    j[k] = p.j;
    for (int a = 0; a < 3; ++a) {
      S[3 * k + a] = p.S[a];
      R[3 * k + a] = p.R[a];
    }
  }
}
void oracle_nl_info(void* hv, int* minimum_image, int* brute, int64_t* dims, int64_t* ncells,
                    double* basis, double* recip) {
  auto* h = static_cast<NLHandle*>(hv);
  *minimum_image = h->w.minimum_image;
  *brute = h->w.brute;
  for (int a = 0; a < 3; ++a) dims[a] = h->w.dims[a];
  *ncells = (int64_t)h->w.keys.size();
  for (int k = 0; k < 9; ++k) {
    basis[k] = h->w.basis[k];
    recip[k] = h->w.recip[k];
  }
}
void oracle_nl_free(void* h) { delete static_cast<NLHandle*>(h); }

}  // extern "C"
