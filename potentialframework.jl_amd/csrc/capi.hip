// extern "C" entry points of libpfw_mi355x (declared in include/pfw_mi355x.h).
// Host-side validation mirrors the reference's error contract:
// ArgumentError for invalid geometry/parameters (storage.jl:87-94,134,162;
// models.jl:34-43), DimensionMismatch for shapes, DomainError for nonfinite
// model output (gpu.jl:390-395).
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "pfw_device.h"

using pfw::Error;

namespace {

// Every entry point runs through here. Unless KEEP (steps, position updates
// and read-outs), the captured step graph is dropped first: any other call may
// replace a buffer, a shape or a model the graph's launches refer to.
template <bool KEEP = false, typename F>
int run(pfw_ctx* ctx, F&& f) {
  if (!ctx) return PFW_INVALID_ARGUMENT;
  try {
    PFW_HIP(hipSetDevice(ctx->device));
    if (!KEEP) pfw::graph_reset(ctx);
    f();
    ctx->err.clear();
    return PFW_OK;
  } catch (const Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return PFW_DEVICE;
  }
}

void check_positions(int64_t n, const double* X) {
  if (n > 0 && !X) throw Error(PFW_INVALID_ARGUMENT, "positions pointer is null");
  for (int64_t k = 0; k < 3 * n; ++k)
    if (!std::isfinite(X[k])) throw Error(PFW_INVALID_ARGUMENT, "positions must be finite");
}

template <typename F>
int run_keep(pfw_ctx* ctx, F&& f) {
  return run<true>(ctx, std::forward<F>(f));
}

// Copy of caller device positions into the staging buffer, flagging any
// nonfinite coordinate.
__global__ void k_copy_positions(int64_t m, const double* __restrict__ src, double* __restrict__ dst,
                                 int32_t* __restrict__ bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const double v = src[k];
  dst[k] = v;
  if (!isfinite(v)) atomicOr(bad, 1);
}

}  // namespace

namespace pfw {

void graph_reset(pfw_ctx* ctx) {
  if (ctx->graph_exec) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipGraphExecDestroy(ctx->graph_exec);
  }
  ctx->graph_exec = nullptr;
  ctx->graph_kind = -1;
  ctx->graph_warm = false;
}

void graph_issue(pfw_ctx* ctx, int kind, int precision, uint32_t mask, void (*issue)(pfw_ctx*, void*), void* arg) {
  if (ctx->graph_exec && ctx->graph_kind == kind && ctx->graph_precision == precision && ctx->graph_mask == mask) {
    PFW_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
    ++ctx->build_gen;  // the replay rebuilds the cell list
    ++ctx->graph_replays;
    return;
  }
  if (ctx->graph_exec) {  // another step kind, precision or mask: replace the graph
    PFW_HIP(hipStreamSynchronize(ctx->stream));
    (void)hipGraphExecDestroy(ctx->graph_exec);
    ctx->graph_exec = nullptr;
    ctx->graph_kind = -1;
  }
  hipGraph_t g = nullptr;
  PFW_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  try {
    issue(ctx, arg);
  } catch (...) {
    (void)hipStreamEndCapture(ctx->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  PFW_HIP(hipStreamEndCapture(ctx->stream, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  PFW_HIP(e);
  ctx->graph_exec = exec;
  ctx->graph_kind = kind;
  ctx->graph_precision = precision;
  ctx->graph_mask = mask;
  ++ctx->graph_captures;
  PFW_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
}

}  // namespace pfw

static void realloc_doubles(double*& p, size_t count) {
  if (p) (void)hipFree(p);
  p = nullptr;
  PFW_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double)));
}

// Cell grids for the cutoff and the current skin, with the per-cell partials.
static void configure_grid(pfw_ctx* ctx, double cutoff) {
  pfw::nl_configure(ctx, cutoff, ctx->verlet.skin);
  realloc_doubles(ctx->d_partials, (size_t)ctx->ncells * 7);
}

extern "C" {

int pfw_ctx_create(int device, pfw_ctx** out) {
  if (!out) return PFW_INVALID_ARGUMENT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PFW_DEVICE;
  if (device < 0 || device >= ndev) return PFW_INVALID_ARGUMENT;
  auto* ctx = new pfw_ctx;
  ctx->device = device;
  int rc = run(ctx, [&] {
    PFW_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    PFW_HIP(hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, device));
    PFW_HIP(hipMalloc(&ctx->d_out, 16 * sizeof(double)));
    PFW_HIP(hipMalloc(&ctx->d_red, 64 * 16 * sizeof(double)));
    PFW_HIP(hipHostMalloc(&ctx->h_out, 16 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    PFW_HIP(hipHostGetDevicePointer((void**)&ctx->h_out_dev, ctx->h_out, 0));
    PFW_HIP(hipMalloc(&ctx->d_flag, sizeof(int32_t)));
    PFW_HIP(hipHostMalloc(&ctx->h_flag, sizeof(int32_t), hipHostMallocDefault));
    for (auto& e : ctx->ev) PFW_HIP(hipEventCreate(&e));
    pfw::static_scratch_init(device);
  });
  if (rc != PFW_OK) {
    pfw_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return PFW_OK;
}

void pfw_ctx_destroy(pfw_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  pfw::graph_reset(ctx);
  pfw::nl_free(ctx);
  pfw::ace_free(ctx);
  pfw::verlet_free(ctx);
  ctx->pool.release();
  if (ctx->d_F) (void)hipFree(ctx->d_F);
  if (ctx->d_partials) (void)hipFree(ctx->d_partials);
  if (ctx->d_out) (void)hipFree(ctx->d_out);
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  if (ctx->d_cfg) (void)hipFree(ctx->d_cfg);
  if (ctx->d_cfg_out) (void)hipFree(ctx->d_cfg_out);
  if (ctx->h_out) (void)hipHostFree(ctx->h_out);
  if (ctx->d_flag) (void)hipFree(ctx->d_flag);
  if (ctx->h_flag) (void)hipHostFree(ctx->h_flag);
  if (ctx->d_Xstage) (void)hipFree(ctx->d_Xstage);
  if (ctx->audit_count) (void)hipFree(ctx->audit_count);
  if (ctx->audit_hash) (void)hipFree(ctx->audit_hash);
  if (ctx->pair.d_params) (void)hipFree(ctx->pair.d_params);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* pfw_last_error(const pfw_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pfw_set_batch(pfw_ctx* ctx, int32_t ncfg, const int64_t* offsets, const double* X, const int32_t* species,
                  const double* H, const uint8_t* pbc, double cutoff) {
  return run(ctx, [&] {
    if (ncfg <= 0 || !offsets) throw Error(PFW_DIMENSION, "a batch needs at least one configuration");
    if (offsets[0] != 0) throw Error(PFW_DIMENSION, "atom offsets must start at 0");
    for (int k = 0; k < ncfg; ++k)
      if (offsets[k + 1] < offsets[k]) throw Error(PFW_DIMENSION, "atom offsets must be nondecreasing");
    const int64_t n = offsets[ncfg];
    if (n > ((int64_t)1 << 31) - 1) throw Error(PFW_DIMENSION, "atom count out of range");
    if (!H || !pbc || (n > 0 && !species)) throw Error(PFW_INVALID_ARGUMENT, "null geometry pointer");
    if (!(std::isfinite(cutoff) && cutoff > 0)) throw Error(PFW_INVALID_ARGUMENT, "cutoff must be finite and positive (Å)");
    check_positions(n, X);
    for (int64_t k = 0; k < n; ++k)
      if (species[k] < 0) throw Error(PFW_INVALID_ARGUMENT, "species channels must be nonnegative");
    ctx->have_system = false;
    ctx->have_results = false;
    ctx->audit_valid = false;
    pfw::nl_free(ctx);
    pfw::ace_free_edges(ctx);
    pfw::verlet_free(ctx);
    if (ctx->d_Xstage) (void)hipFree(ctx->d_Xstage);
    ctx->d_Xstage = nullptr;
    ctx->ncfg = ncfg;
    ctx->atom0.assign(offsets, offsets + ncfg + 1);
    // open cell vectors may be infinite or zero (storage.jl:80-82); periodic ones are checked by the basis
    ctx->host_H.assign(H, H + 9 * (size_t)ncfg);
    ctx->host_pb.resize(3 * (size_t)ncfg);
    for (size_t k = 0; k < ctx->host_pb.size(); ++k) ctx->host_pb[k] = pbc[k] ? 1 : 0;
    ctx->sys.n = n;
    PFW_HIP(hipMalloc(&ctx->sys.X, std::max<int64_t>(3 * n, 1) * sizeof(double)));
    PFW_HIP(hipMalloc(&ctx->sys.species, std::max<int64_t>(n, 1) * sizeof(int32_t)));
    if (n > 0) {
      PFW_HIP(hipMemcpyAsync(ctx->sys.X, X, 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      PFW_HIP(hipMemcpyAsync(ctx->sys.species, species, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    }
    configure_grid(ctx, cutoff);
    realloc_doubles(ctx->d_F, 3 * (size_t)n);
    realloc_doubles(ctx->d_cfg, (size_t)ncfg * 16);
    realloc_doubles(ctx->d_cfg_out, (size_t)ncfg * 11);
    PFW_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_system = true;
  });
}

int pfw_set_system(pfw_ctx* ctx, int64_t n, const double* X, const int32_t* species, const double* H,
                   const uint8_t* pbc, double cutoff) {
  if (n < 0 || n > ((int64_t)1 << 31) - 1) {
    if (ctx) ctx->err = "atom count out of range";
    return ctx ? PFW_DIMENSION : PFW_INVALID_ARGUMENT;
  }
  const int64_t offsets[2] = {0, n};
  return pfw_set_batch(ctx, 1, offsets, X, species, H, pbc, cutoff);
}

int pfw_update_positions_device(pfw_ctx* ctx, const double* X) {
  return run_keep(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    const int64_t n = ctx->sys.n;
    if (n > 0 && !X) throw Error(PFW_INVALID_ARGUMENT, "positions pointer is null");
    if (n == 0) return;
    hipStream_t st = ctx->stream;
    if (!ctx->d_Xstage) PFW_HIP(hipMalloc(&ctx->d_Xstage, 3 * n * sizeof(double)));
    PFW_HIP(hipMemsetAsync(ctx->d_flag, 0, sizeof(int32_t), st));
    hipLaunchKernelGGL(k_copy_positions, dim3((unsigned)((3 * n + 255) / 256)), dim3(256), 0, st, 3 * n, X,
                       ctx->d_Xstage, ctx->d_flag);
    PFW_HIP(hipGetLastError());
    PFW_HIP(hipMemcpyAsync(ctx->h_flag, ctx->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PFW_HIP(hipStreamSynchronize(st));
    if (*ctx->h_flag) throw Error(PFW_INVALID_ARGUMENT, "positions must be finite");
    if (ctx->graph_exec)  // a captured step reads the positions buffer it was captured with
      PFW_HIP(hipMemcpyAsync(ctx->sys.X, ctx->d_Xstage, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
    else
      std::swap(ctx->sys.X, ctx->d_Xstage);  // the previous positions stay untouched on failure
    ctx->have_results = false;
  });
}

int pfw_update_positions(pfw_ctx* ctx, const double* X) {
  return run_keep(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    check_positions(ctx->sys.n, X);
    if (ctx->sys.n > 0)
      PFW_HIP(hipMemcpyAsync(ctx->sys.X, X, 3 * ctx->sys.n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    PFW_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_results = false;
  });
}

int pfw_set_pair_model(pfw_ctx* ctx, int kind, int nz, const double* params, double cutoff) {
  return run(ctx, [&] {
    if (!params) throw Error(PFW_INVALID_ARGUMENT, "null parameters");
    if (!(std::isfinite(cutoff) && cutoff > 0)) throw Error(PFW_INVALID_ARGUMENT, "cutoff must be finite and positive");
    std::vector<double> p;
    if (kind == PFW_PAIR_SMOOTH) {
      if (nz <= 0) throw Error(PFW_INVALID_ARGUMENT, "SmoothPair needs at least one species");
      p.assign(params, params + (size_t)nz * nz);
      for (int a = 0; a < nz; ++a)
        for (int b = 0; b < nz; ++b) {
          double v = p[a + nz * b];
          if (!std::isfinite(v) || v != p[b + nz * a])
            throw Error(PFW_INVALID_ARGUMENT, "amplitude table must be finite and symmetric");
        }
    } else if (kind == PFW_PAIR_LJ) {
      p.assign(params, params + 2);
      if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && p[0] > 0 && p[1] > 0))
        throw Error(PFW_INVALID_ARGUMENT, "LJ parameters must be positive and finite");
      nz = 1;
    } else {
      throw Error(PFW_INVALID_ARGUMENT, "unknown pair model kind");
    }
    if (ctx->pair.d_params) (void)hipFree(ctx->pair.d_params);
    ctx->pair.d_params = nullptr;
    PFW_HIP(hipMalloc(&ctx->pair.d_params, p.size() * sizeof(double)));
    PFW_HIP(hipMemcpyAsync(ctx->pair.d_params, p.data(), p.size() * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream));
    PFW_HIP(hipStreamSynchronize(ctx->stream));
    ctx->pair.kind = kind;
    ctx->pair.nz = nz;
    ctx->pair.cutoff = cutoff;
    ctx->pair.params = p;
  });
}

int pfw_pair_step(pfw_ctx* ctx, uint32_t mask, int precision) {
  return run_keep(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (precision != PFW_FP64 && precision != PFW_MIXED) throw Error(PFW_INVALID_ARGUMENT, "unknown precision");
    if (mask & ~7u) throw Error(PFW_INVALID_ARGUMENT, "unknown request bits");
    ctx->have_results = false;
    pfw::pair_run(ctx, mask, precision);
  });
}

int pfw_read_results(pfw_ctx* ctx, double* E, double* F, double* W) {
  return run_keep(ctx, [&] {
    if (!ctx->have_results) throw Error(PFW_STALE, "no results: run a step first");
    if (E) *E = (ctx->last_mask & PFW_ENERGY) ? ctx->h_out[0] : 0.0;
    if (W)
      for (int k = 0; k < 9; ++k) W[k] = (ctx->last_mask & PFW_VIRIAL) ? ctx->h_out[1 + k] : 0.0;
    if (F && ctx->sys.n > 0) {
      if (ctx->last_mask & PFW_FORCES) {
        PFW_HIP(hipMemcpyAsync(F, ctx->d_F, 3 * ctx->sys.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        PFW_HIP(hipStreamSynchronize(ctx->stream));
      } else {
        std::memset(F, 0, 3 * ctx->sys.n * sizeof(double));
      }
    }
  });
}

int pfw_read_batch_results(pfw_ctx* ctx, double* E, double* F, double* W) {
  return run_keep(ctx, [&] {
    if (!ctx->have_results) throw Error(PFW_STALE, "no results: run a step first");
    const int ncfg = ctx->ncfg;
    for (int k = 0; k < ncfg; ++k) {
      const double* o = ctx->h_cfg_out.data() + 11 * (size_t)k;
      if (E) E[k] = (ctx->last_mask & PFW_ENERGY) ? o[0] : 0.0;
      if (W)
        for (int c = 0; c < 9; ++c) W[9 * (size_t)k + c] = (ctx->last_mask & PFW_VIRIAL) ? o[1 + c] : 0.0;
    }
    if (F && ctx->sys.n > 0) {
      if (ctx->last_mask & PFW_FORCES) {
        PFW_HIP(hipMemcpyAsync(F, ctx->d_F, 3 * ctx->sys.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        PFW_HIP(hipStreamSynchronize(ctx->stream));
      } else {
        std::memset(F, 0, 3 * ctx->sys.n * sizeof(double));
      }
    }
  });
}

int pfw_pair_evaluate(pfw_ctx* ctx, uint32_t mask, int precision, double* E, double* F, double* W) {
  int rc = pfw_pair_step(ctx, mask, precision);
  if (rc != PFW_OK) return rc;
  return pfw_read_results(ctx, E, F, W);
}

int pfw_export_pairs(pfw_ctx* ctx, int32_t* i, int32_t* j, int32_t* S, double* R, int64_t cap, int64_t* count) {
  return run_keep(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (!count) throw Error(PFW_INVALID_ARGUMENT, "null count");
    pfw::nl_export_pairs(ctx, i, j, S, R, cap, count);
  });
}

int pfw_grid_info(pfw_ctx* ctx, int64_t* ncells, int32_t* dims, int32_t* minimum_image) {
  return run_keep(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (ncells) *ncells = ctx->ncells;
    if (dims)
      for (int a = 0; a < 3; ++a) dims[a] = ctx->grids[0].dims[a];
    if (minimum_image) *minimum_image = ctx->grids[0].minimum_image;
  });
}

int pfw_set_skin(pfw_ctx* ctx, double skin) {
  return run(ctx, [&] {
    if (!(std::isfinite(skin) && skin >= 0)) throw Error(PFW_INVALID_ARGUMENT, "skin must be finite and nonnegative (Å)");
    if (skin == ctx->verlet.skin) return;
    pfw::verlet_free(ctx);
    ctx->ace.skin_valid = false;
    ctx->verlet.skin = skin;
    if (ctx->have_system) {
      ctx->have_system = false;
      configure_grid(ctx, ctx->cutoff);  // cells follow cutoff + skin
      PFW_HIP(hipStreamSynchronize(ctx->stream));
      ctx->have_system = true;
    }
    ctx->have_results = false;
  });
}

int pfw_verlet_info(pfw_ctx* ctx, int64_t* info, double* max_displacement) {
  return run_keep(ctx, [&] {
    const pfw::VerletDev& V = ctx->verlet;
    if (info) {
      info[0] = V.rebuilds;
      info[1] = V.reuses;
      info[2] = V.valid ? V.n_ext - ctx->sys.n : 0;
      info[3] = V.ecap;
      info[4] = V.valid ? V.entries : 0;
    }
    if (max_displacement) *max_displacement = V.last_max_disp;
  });
}

int pfw_set_timing(pfw_ctx* ctx, int enabled) {
  return run(ctx, [&] { ctx->timing = enabled != 0; });
}

int pfw_get_timings(pfw_ctx* ctx, double* ms, int n) {
  if (!ctx || !ms) return 0;
  int m = std::min(n, pfw::kStages);
  for (int k = 0; k < m; ++k) ms[k] = ctx->stage_ms[k];
  return m;
}

int pfw_pair_evaluate_noalloc(pfw_ctx* ctx, int kind, int nz, const double* params, double cutoff, int64_t n,
                              const double* X_device, const int32_t* species_device, const double* H,
                              const uint8_t* pbc, uint32_t request_mask, int precision, double* E_device,
                              double* F_device, double* W_device) {
  return run(ctx, [&] {
    pfw::pair_run_static(ctx, kind, nz, params, cutoff, n, X_device, species_device, H, pbc, request_mask, precision,
                         E_device, F_device, W_device);
  });
}

int pfw_set_audit(pfw_ctx* ctx, int enabled) {
  return run(ctx, [&] {
    ctx->audit = enabled != 0;
    ctx->audit_valid = false;
  });
}

int pfw_read_audit(pfw_ctx* ctx, int32_t* count, uint64_t* hash) {
  return run_keep(ctx, [&] {
    if (!ctx->audit || !ctx->audit_valid) throw Error(PFW_STALE, "no audited pair or ACE step since the audit was enabled");
    const int64_t n = ctx->sys.n;
    if (n == 0) return;
    if (count) PFW_HIP(hipMemcpy(count, ctx->audit_count, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (hash) PFW_HIP(hipMemcpy(hash, ctx->audit_hash, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
  });
}

int pfw_set_ace_model(pfw_ctx* ctx, const pfw_ace_model* model) {
  return run(ctx, [&] {
    if (!model) throw Error(PFW_INVALID_ARGUMENT, "null model");
    pfw::ace_set_model(ctx, model);
  });
}

int pfw_set_ace_parameters(pfw_ctx* ctx, const double* WB, const double* Wpair) {
  return run(ctx, [&] {
    if (!ctx->ace.ready) throw Error(PFW_STALE, "no ACE model set");
    if (!WB) throw Error(PFW_INVALID_ARGUMENT, "null readout parameters");
    pfw::ace_set_parameters(ctx, WB, Wpair);
  });
}

int pfw_ace_step(pfw_ctx* ctx, uint32_t mask, int precision) {
  return run_keep(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (precision != PFW_FP64 && precision != PFW_MIXED) throw Error(PFW_INVALID_ARGUMENT, "unknown precision");
    if (mask & ~7u) throw Error(PFW_INVALID_ARGUMENT, "unknown request bits");
    ctx->have_results = false;
    pfw::ace_run(ctx, mask, precision);
  });
}

int pfw_ace_evaluate(pfw_ctx* ctx, uint32_t mask, int precision, double* E, double* F, double* W) {
  int rc = pfw_ace_step(ctx, mask, precision);
  if (rc != PFW_OK) return rc;
  return pfw_read_results(ctx, E, F, W);
}

int pfw_ace_site_energies(pfw_ctx* ctx, double* out) {
  return run_keep(ctx, [&] {
    if (!ctx->have_results) throw Error(PFW_STALE, "no results: run an ACE step first");
    if (!out) throw Error(PFW_INVALID_ARGUMENT, "null output");
    pfw::ace_site_energies(ctx, out);
  });
}

int pfw_ace_adjoint(pfw_ctx* ctx, int precision, const double* dE, const double* dF, const double* dW, double* gWB,
                    double* gWpair) {
  return run(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (!ctx->ace.ready) throw Error(PFW_STALE, "no ACE model set");
    if (precision != PFW_FP64 && precision != PFW_MIXED) throw Error(PFW_INVALID_ARGUMENT, "unknown precision");
    if (!dE || !dF || !gWB) throw Error(PFW_INVALID_ARGUMENT, "null adjoint argument");
    pfw::ace_adjoint(ctx, precision, dE, dF, dW, gWB, gWpair);
  });
}

int pfw_ace_design(pfw_ctx* ctx, int precision, const double* basis_scale, double* out, int64_t ld) {
  return run(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    pfw::ace_design(ctx, precision, basis_scale, out, ld);
  });
}

int pfw_ace_parameter_sets(pfw_ctx* ctx, int precision, int64_t nsets, const double* WB, const double* Wpair,
                           double* out) {
  return run(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (precision != PFW_FP64 && precision != PFW_MIXED) throw Error(PFW_INVALID_ARGUMENT, "unknown precision");
    ctx->have_results = false;
    pfw::ace_parameter_sets(ctx, precision, nsets, WB, Wpair, out);
  });
}

int pfw_features(pfw_ctx* ctx, const pfw_feature_layer* layer, double* out) {
  return run(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (ctx->ncfg != 1) throw Error(PFW_INVALID_ARGUMENT, "features are evaluated for one configuration");
    ctx->have_results = false;
    pfw::features_run(ctx, layer, out);
  });
}

int pfw_features_vjp(pfw_ctx* ctx, const pfw_feature_layer* layer, const double* cotangents, double* dx,
                     double* dmixing, double* dscale, double* dradial) {
  return run(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (ctx->ncfg != 1) throw Error(PFW_INVALID_ARGUMENT, "features are evaluated for one configuration");
    ctx->have_results = false;
    pfw::features_vjp_run(ctx, layer, cotangents, dx, dmixing, dscale, dradial);
  });
}

int pfw_pair_adjoint(pfw_ctx* ctx, const double* dE, const double* dF, const double* dW, double* gA) {
  return run(ctx, [&] {
    if (!ctx->have_system) throw Error(PFW_STALE, "no system set");
    if (!dE || !dF || !gA) throw Error(PFW_INVALID_ARGUMENT, "null adjoint argument");
    pfw::pair_adjoint(ctx, dE, dF, dW, gA);
  });
}

// Pinned host blocks for result arrays: a process-wide cache keyed by size
// (a request takes the smallest cached size of at least its own and at most
// twice it, the most recently returned block of that size first, else a new
// hipHostMalloc block); returned blocks are kept up to kHostCacheBytes in
// total, beyond that freed. Blocks still cached at exit are left to the
// process teardown.
namespace {
constexpr size_t kHostCacheBytes = size_t(256) << 20;
struct HostPool {
  std::mutex mu;
  std::map<size_t, std::vector<void*>> free;  // size -> blocks, most recent last
  std::map<void*, size_t> size_of;            // every live or cached block
  size_t cached = 0;
};
HostPool& host_pool() {
  static HostPool* p = new HostPool;  // never destroyed: outlives every caller
  return *p;
}
}  // namespace

void* pfw_host_alloc(size_t bytes) {
  if (bytes == 0) return nullptr;
  const size_t want = (bytes + 65535) & ~size_t(65535);
  HostPool& P = host_pool();
  {
    std::lock_guard<std::mutex> g(P.mu);
    auto it = P.free.lower_bound(want);
    if (it != P.free.end() && it->first <= 2 * want) {
      void* q = it->second.back();
      P.cached -= it->first;
      it->second.pop_back();
      if (it->second.empty()) P.free.erase(it);
      return q;
    }
  }
  void* q = nullptr;
  if (hipHostMalloc(&q, want, hipHostMallocDefault) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> g(P.mu);
  P.size_of[q] = want;
  return q;
}

void pfw_host_free(void* p) {
  if (!p) return;
  HostPool& P = host_pool();
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.size_of.find(p);
  if (it == P.size_of.end()) return;  // not a block of this pool
  if (P.cached + it->second > kHostCacheBytes) {
    (void)hipHostFree(p);
    P.size_of.erase(it);
    return;
  }
  P.cached += it->second;
  P.free[it->second].push_back(p);
}

int pfw_set_graphs(pfw_ctx* ctx, int enabled) {
  return run(ctx, [&] { ctx->graphs = enabled != 0; });
}

int pfw_graph_info(pfw_ctx* ctx, int64_t* out) {
  return run_keep(ctx, [&] {
    if (!out) throw Error(PFW_INVALID_ARGUMENT, "null output");
    out[0] = ctx->graph_captures;
    out[1] = ctx->graph_replays;
    out[2] = ctx->graph_exec ? 1 : 0;
  });
}

int pfw_set_tuning(pfw_ctx* ctx, const int32_t* knobs, int nknobs) {
  return run(ctx, [&] {
    if (!knobs || nknobs < 0 || nknobs > 5) throw Error(PFW_INVALID_ARGUMENT, "knobs: at most 5 values");
    int t[5] = {ctx->tune[0], ctx->tune[1], ctx->tune[2], ctx->tune[3], ctx->tune[4]};
    for (int k = 0; k < nknobs; ++k) t[k] = knobs[k];
    if (t[0] < 0 || t[0] > 2) throw Error(PFW_INVALID_ARGUMENT, "ACE sites per wavefront must be 0 (built-in), 1 or 2");
    if (t[1] < 0 || t[1] > 4096) throw Error(PFW_INVALID_ARGUMENT, "ACE radial chunk must be 0 (built-in) or 1..4096");
    if (t[2] < 0 || t[2] > 16) throw Error(PFW_INVALID_ARGUMENT, "ACE product wavefronts must be 0 (built-in) or 1..16");
    if (t[3] < 0 || t[3] > 4) throw Error(PFW_INVALID_ARGUMENT, "pair lanes per atom must be 0 (built-in) or 1..4");
    if (t[4] < 0 || t[4] > 3)
      throw Error(PFW_INVALID_ARGUMENT,
                  "ACE contractions must be 0 (built-in), 1 (VALU), 2 (densities on matrix cores) or 3 (densities and "
                  "pullback on matrix cores)");
    PFW_HIP(hipStreamSynchronize(ctx->stream));
    const bool regroup = t[1] != ctx->tune[1];
    for (int k = 0; k < 5; ++k) ctx->tune[k] = t[k];
    if (regroup && ctx->ace.ready) pfw::ace_group_densities(ctx, t[1]);
  });
}

int pfw_get_tuning(pfw_ctx* ctx, int32_t* knobs, int nknobs) {
  return run_keep(ctx, [&] {
    if (!knobs || nknobs < 0 || nknobs > 5) throw Error(PFW_INVALID_ARGUMENT, "knobs: at most 5 values");
    for (int k = 0; k < nknobs; ++k) knobs[k] = ctx->tune[k];
  });
}

int pfw_synchronize(pfw_ctx* ctx) {
  return run_keep(ctx, [&] { PFW_HIP(hipStreamSynchronize(ctx->stream)); });
}

}  // extern "C"
