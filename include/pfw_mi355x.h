/*
 * pfw_mi355x.h — C ABI of the MI355X (gfx950) force-evaluation hot path of
 * PotentialFrameWork.jl: triclinic multi-image cell lists, SmoothPair and
 * Lennard-Jones energy/force/virial, and the ACE site pipeline.
 *
 * Conventions (the reference's, src/submodules/PairPotentials/evaluation.jl and
 * README.md "conventions"): lengths in Å, energies in eV, forces in eV/Å, virial
 * in eV (configurational, W = -sum d (x) g). Positions are 3 x N column-major
 * doubles in ORIGINAL atom order; the cell H is 3 x 3 column-major with the
 * lattice vectors in COLUMNS; pbc is three 0/1 bytes. Forces come back in
 * original order. Every entry point returns a pfw_status; pfw_last_error()
 * gives the message of the last failure on that context.
 *
 * Ownership and threading: the caller owns every host buffer; the library owns
 * device memory. A context is NOT safe for concurrent calls (the reference's
 * Workspace rule, README.md:217-219). Results are written before return.
 */
#ifndef PFW_MI355X_H
#define PFW_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PFW_OK = 0,
  PFW_INVALID_ARGUMENT = 1, /* Julia ArgumentError */
  PFW_DIMENSION = 2,        /* Julia DimensionMismatch */
  PFW_DOMAIN = 3,           /* Julia DomainError: singular or nonfinite model output */
  PFW_STALE = 4,            /* Julia ErrorException: stale topology / model */
  PFW_CAPACITY = 5,         /* a caller-supplied capacity is too small */
  PFW_DEVICE = 6            /* HIP runtime failure */
} pfw_status;

/* Request mask, as AtomsCalculators Energy()/Forces()/Virial() tuples
 * (src/submodules/NeighbourList/execution.jl:4-12). Any subset, including 0. */
enum { PFW_ENERGY = 1, PFW_FORCES = 2, PFW_VIRIAL = 4 };

/* Arithmetic of the pair and ACE kernels. FP64: plain double everywhere.
 * MIXED: float per-pair / per-edge arithmetic, double force, energy and virial
 * accumulation; neighbour membership is still decided bit-exactly in double. */
enum { PFW_FP64 = 0, PFW_MIXED = 1 };

/* Pair model kinds. */
enum { PFW_PAIR_SMOOTH = 0, PFW_PAIR_LJ = 1 };

typedef struct pfw_ctx pfw_ctx;

/* Replaces constructing NeighbourList.Workspace / GPUWorkspace
 * (src/submodules/NeighbourList/storage.jl:59-65, gpu.jl:82-116). */
int pfw_ctx_create(int device, pfw_ctx** out);
void pfw_ctx_destroy(pfw_ctx* ctx);
const char* pfw_last_error(const pfw_ctx* ctx);

/* Geometry upload and validation: the extraction half of prepare!
 * (src/submodules/NeighbourList/storage.jl:130-185, periodic_basis :83-108).
 * species are channel ids 0..nz-1, already mapped from ChemicalSpecies by the
 * host (bind_species, src/submodules/PairPotentials/models.jl:50-58). The cell
 * grid is fixed here for the given cutoff; positions may later be replaced
 * with pfw_update_positions without changing the cell. */
int pfw_set_system(pfw_ctx* ctx, int64_t n, const double* X, const int32_t* species,
                   const double* H, const uint8_t* pbc, double cutoff);
/* A batch of ncfg independent configurations packed into one cell list
 * (the fitting-set case; the reference evaluates configurations one by one).
 * atom_offsets (ncfg + 1, first 0) delimit each configuration's atoms in X /
 * species; H holds ncfg column-major 3 x 3 cells, pbc 3 bytes per
 * configuration. Every later step covers all configurations with one launch
 * per stage; pfw_read_batch_results returns per-configuration outputs.
 * pfw_set_system is the ncfg = 1 case. No counterpart entry point exists in
 * the reference; it replaces looping evaluate! over a fitting set
 * (src/submodules/ACE/fitting.jl:18-63). */
int pfw_set_batch(pfw_ctx* ctx, int32_t ncfg, const int64_t* atom_offsets, const double* X,
                  const int32_t* species, const double* H, const uint8_t* pbc, double cutoff);
/* Per-configuration outputs of the last step: E (ncfg), F (3 x N, original
 * order), W (9 x ncfg, column-major per configuration); NULL skips. */
int pfw_read_batch_results(pfw_ctx* ctx, double* E, double* F, double* W);

/* New positions for the same atoms and cell (host buffer). */
int pfw_update_positions(pfw_ctx* ctx, const double* X);
/* The same from device memory (3 x N doubles on the context's device, e.g. an
 * integrator's array): copied on the device and checked for finiteness, the
 * call synchronises; on PFW_INVALID_ARGUMENT the previous positions are kept.
 * Replaces copyto!(executor.positions, ...) before execute_gpu!
 * (src/submodules/ACE/gpu.jl:329-343; NeighbourList/gpu.jl:363-372). */
int pfw_update_positions_device(pfw_ctx* ctx, const double* X_device);

/* Pair model (NI.prepare_potential for SmoothPair / RadialPotential{LennardJonesRadial},
 * src/submodules/PairPotentials/models.jl:145-176, lennard_jones.jl:5-22;
 * GPU certification prepare_gpu_potential, gpu.jl:15-36). params:
 *   PFW_PAIR_SMOOTH: nz*nz symmetric amplitude table, column-major (eV)
 *   PFW_PAIR_LJ:     {epsilon (eV), sigma (Å)}; nz ignored.
 * The cutoff must equal the one given to pfw_set_system. */
int pfw_set_pair_model(pfw_ctx* ctx, int kind, int nz, const double* params, double cutoff);

/* One complete step on the device: fresh cell build, pair E/F/V, FP64
 * reductions and the nonfinite status check; outputs stay on the device.
 * Replaces gpu_execute! (src/submodules/NeighbourList/gpu.jl:363-402) and the
 * CPU execute! (src/submodules/NeighbourList/execution.jl:223-238). */
int pfw_pair_step(pfw_ctx* ctx, uint32_t request_mask, int precision);
/* Copy the last step's outputs to the host (any pointer may be NULL).
 * E: 1 double; F: 3 x N column-major, original order; W: 3 x 3 column-major. */
int pfw_read_results(pfw_ctx* ctx, double* E, double* F, double* W);
/* pfw_pair_step + pfw_read_results: the evaluate! boundary
 * (src/submodules/PairPotentials/evaluation.jl:31-52; gpu.jl:45-55). */
int pfw_pair_evaluate(pfw_ctx* ctx, uint32_t request_mask, int precision, double* E,
                      double* F, double* W);

/* Canonical directed neighbour set of the current system, in original order:
 * for every directed image (i -> j, S) with R = x_j - x_i + H S strictly inside
 * the cutoff, excluding (i, i, 0). Replaces neighbour_list / materialize /
 * visit_directed_images! (src/submodules/NeighbourList/interface.jl:47-143).
 * i, j are 0-based; S is 3 ints per pair; R is 3 doubles per pair (either may
 * be NULL). If cap is too small, *count receives the size and PFW_CAPACITY is
 * returned. Pairs are ordered by (i, then traversal). */
int pfw_export_pairs(pfw_ctx* ctx, int32_t* i, int32_t* j, int32_t* S, double* R, int64_t cap,
                     int64_t* count);

/* Verlet skin (Å, default 0): the GPUConfig(skin=...) option of the
 * reference's GPU pair workspace (src/submodules/NeighbourList/gpu.jl:61-72,
 * 504-549; docs/gpu_pairs.md:116-131). With skin > 0, pfw_pair_step keeps a
 * neighbour list at cutoff + skin and rebuilds it only when some atom has moved
 * at least skin/2 (times 1 - 16 eps) from its position at the last build, in
 * the caller's raw coordinates (a wrapped jump rebuilds), or when the system,
 * skin or cell build changed; every interaction uses the current positions
 * and the model cutoff. The cell grid is sized for cutoff + skin. Unlike the
 * reference, triclinic and multi-image cells are allowed (periodic images are
 * explicit ghost entries). skin = 0 returns to fresh cells every step. */
int pfw_set_skin(pfw_ctx* ctx, double skin);
/* Verlet state, info[5]: info[0] rebuilds, info[1] reused steps (the
 * reference's w.rebuilds counter, gpu.jl:100,508), info[2] ghost entries,
 * info[3] list slots per atom, info[4] list entries (pairs within cutoff +
 * skin at the last build, each direction counted); max_displacement: largest displacement since the build seen
 * by the last step (Å). Either pointer may be NULL. */
int pfw_verlet_info(pfw_ctx* ctx, int64_t* info, double* max_displacement);

/* Diagnostics: number of cells and the cell grid (dims[3]) of the current
 * system, and whether the minimum-image path is active (storage.jl:140). */
int pfw_grid_info(pfw_ctx* ctx, int64_t* ncells, int32_t* dims, int32_t* minimum_image);

/* Per-stage device times of the last step in ms, measured with HIP events on
 * the context's stream when timing is enabled (the reference's w.timings
 * dictionary, storage.jl:185-236). stages: 0 binning, 1 scan, 2 scatter+sort,
 * 3 environments (ACE edge lists), 4 interaction (the pair or ACE site
 * kernel), 5 reduction (force gather, FP64 sums, status read). Returns the
 * number written. */
int pfw_set_timing(pfw_ctx* ctx, int enabled);
int pfw_get_timings(pfw_ctx* ctx, double* ms, int n);

/* Allocation-free pair step, for callers that cannot allocate device memory
 * during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). No counterpart in the reference: it is
 * pfw_set_system + pfw_set_pair_model + pfw_pair_evaluate in one call whose
 * intermediate buffers (cell grid, sorted positions, partial sums) live in a
 * static device area of about 56 MB that each device holds from the first
 * context created on it until the process ends; nothing is allocated or
 * freed. Inputs X_device (3 x n doubles, column-major, original order) and
 * species_device (n channels) and outputs E_device (1), F_device (3 x n,
 * original order) and W_device (3 x 3, column-major) are device pointers on
 * the context's device (an output may be NULL when its quantity is not
 * requested); H, pbc and params are host values as for pfw_set_system and
 * pfw_set_pair_model. Limits: n <= 262,144 atoms and a cell grid of at most
 * 262,144 cells (else PFW_CAPACITY). Calls that overlap on one device share
 * the static area and race; the call synchronises the context's stream
 * before it returns. */
int pfw_pair_evaluate_noalloc(pfw_ctx* ctx, int kind, int nz, const double* params, double cutoff, int64_t n,
                              const double* X_device, const int32_t* species_device, const double* H,
                              const uint8_t* pbc, uint32_t request_mask, int precision, double* E_device,
                              double* F_device, double* W_device);

/* Pair-set audit, a diagnostic for parity tests (no counterpart in the
 * reference, whose CPU executor and GPU kernels read one neighbour set). While
 * enabled, every pair step (fresh cells or Verlet list) and ACE step records,
 * per centre atom i in original order, the number of directed pairs (j, S)
 * the evaluating kernel actually accumulated (count) and the wrapping 64-bit
 * sum of a mix of each pair's identity (hash): j the neighbour's original
 * index, S the integer image shift relative to the caller's original
 * coordinates, R = x_j - x_i + H S, as in the directed snapshot of
 * visit_directed_images! (src/submodules/NeighbourList/interface.jl:116-143).
 * The mix is splitmix64's finaliser of j | (S+1024 & 2047) << 32 | ... << 43 |
 * ... << 54 after adding 0x9e3779b97f4a7c15, so a caller can digest a
 * reference pair set the same way and compare bit for bit. Audited steps are
 * slower (one atomic per pair). pfw_read_audit copies the last pair or ACE
 * step's digests (n each; either pointer may be NULL); PFW_STALE when no step
 * ran since the audit was enabled or the system changed. */
int pfw_set_audit(pfw_ctx* ctx, int enabled);
int pfw_read_audit(pfw_ctx* ctx, int32_t* count, uint64_t* hash);

/* Captured steps. With graphs enabled, the launches of a fresh pair or ACE
 * step without skin (cell build, the interaction kernels, the FP64
 * reductions) are captured once as a HIP graph and replayed while only the
 * positions change (pfw_update_positions / pfw_update_positions_device); any
 * other entry point that may change a buffer, shape or model drops the graph,
 * and the step after that runs eagerly before the next capture. Timing and
 * audit runs are never captured. Results are those of the eager step. The
 * reference runs its steps as KernelAbstractions launches with no capture
 * (src/submodules/NeighbourList/gpu.jl:236-402, ACE/gpu.jl:301-330); this is
 * a launch-mode switch beside the GPUConfig knobs of gpu.jl:61-73.
 * pfw_graph_info: out[0] captures, out[1] replays, out[2] 1 while a graph is held. */
int pfw_set_graphs(pfw_ctx* ctx, int enabled);
int pfw_graph_info(pfw_ctx* ctx, int64_t* out);

/* Launch-shape knobs, the tunable counterparts of GPUConfig's strategy,
 * workgroup and tile (src/submodules/NeighbourList/gpu.jl:61-73) and the ACE
 * strategy (src/submodules/ACE/gpu.jl:301-303). knobs[0]: ACE sites per
 * wavefront (0 built-in, 1 or 2); knobs[1]: ACE radial indices per density
 * tile (0 built-in = 20, or 1..4096); knobs[2]: ACE product wavefronts per
 * workgroup (0 built-in, 1..16); knobs[3]: most lanes per atom in the pair
 * kernel (0 built-in = 4, 1..4); knobs[4]: ACE contractions (0 built-in,
 * 1 VALU, 2 densities on the matrix cores, 3 densities and pullback on the
 * matrix cores; the matrix cores serve FP32 steps of unprojected models, the
 * densities only when their 4 x 4 tiles fit four groups of 16; the built-in
 * takes them for the densities when at least half of the tile entries are
 * densities). Every setting gives
 * the same pair sets and results to rounding. nknobs <= 5 values are read or
 * written from the first; PFW_INVALID_ARGUMENT for a value outside its
 * range. */
int pfw_set_tuning(pfw_ctx* ctx, const int32_t* knobs, int nknobs);
int pfw_get_tuning(pfw_ctx* ctx, int32_t* knobs, int nknobs);

/* Synchronise the context's stream. */
int pfw_synchronize(pfw_ctx* ctx);

/* Page-locked host memory for result arrays, from a process-wide cache (at
 * most 256 MiB of returned blocks are kept; beyond that they are freed): the
 * library's device-to-host copies into such a block run as one DMA, where a
 * pageable destination is copied through the runtime's staging buffers. No
 * counterpart in the reference (Julia owns its arrays); for host bindings that
 * want owning result arrays filled at DMA speed. pfw_host_alloc returns NULL
 * when no block can be had (bytes == 0 included); pfw_host_free returns a block
 * to the cache (NULL and foreign pointers are ignored). Thread-safe. */
void* pfw_host_alloc(size_t bytes);
void pfw_host_free(void* p);

/* ---------------------------------------------------------------- ACE
 * Resolved ACE model, schema v1 (src/submodules/ACE/resolved.jl:18-142):
 * piecewise-cubic Agnesi radial splines per (centre, neighbour) species
 * channel (index centre + nz*neighbour), real L2 spherical (or solid)
 * harmonics in l*l+l+m order, density (radial index, harmonic index) pairs, product
 * multisets of density indices (order <= 4), sparse coupling C (nbasis x
 * nprod), readout WB (nbasis x nz), Wpair (nq x nz), and the basis policy's
 * species projection (the caller applies the readout mask and the pair and
 * reference flags to WB, Wpair and the reference energies). */
enum { PFW_ENV_POLY2SX = 0, PFW_ENV_POLY1SR = 1, PFW_ENV_ACE1_POLY1SR = 2 };
/* Angular basis (resolved.jl:109-110; angular_blocks.jl:47-48). */
enum { PFW_ANGULAR_SPHERICAL = 0, PFW_ANGULAR_SOLID = 1 };

typedef struct {
  int32_t nr;                    /* radial functions per channel */
  int32_t nnodes;                /* spline nodes per channel (>= 2, first -1, last 1) */
  const double* transform;       /* nz*nz x 8: p q a rin r0 yin ycut rcut */
  const int32_t* envelope_kind;  /* nz*nz: PFW_ENV_* */
  const double* envelope;        /* nz*nz x 5: (x1 x2 p1 p2 s) | (rcut p) | (rcut r0 p) */
  const double* nodes;           /* nz*nz x nnodes */
  const double* coefficients;    /* nz*nz x (4 x nr x (nnodes-1)), power fastest */
} pfw_ace_radial;

typedef struct {
  int32_t nz, lmax, ndens, nprod, nbasis;
  const double* reference;       /* nz reference energies (eV) */
  pfw_ace_radial radial, pair;   /* pair.nr may be 0 */
  const int32_t* density;        /* ndens x 2: (radial index, harmonic index), 0-based */
  const int32_t* product_ptr;    /* nprod + 1 */
  const int32_t* product_idx;    /* 0-based density indices */
  int64_t coupling_nnz;
  const int32_t* coupling_rows;  /* 0-based basis row */
  const int32_t* coupling_cols;  /* 0-based product column */
  const double* coupling_vals;
  const double* WB;              /* nbasis x nz, column-major */
  const double* Wpair;           /* pair.nr x nz, column-major */
  int32_t angular;               /* PFW_ANGULAR_*: "real_spherical" or "real_solid" (r^l Y_lm) */
  /* Basis policy of project_species (src/submodules/ACE/interface.jl:115-125):
   * order_lo = order_hi = -1 for no order restriction; otherwise every product
   * keeps the orders order_lo..order_hi of prod(U + t V), U and V the densities
   * of neighbours outside and inside the projected species (selected: nz flags,
   * NULL = none; at most 32 channels), and a pair term counts as order 1 for a
   * projected neighbour, 0 otherwise (pair_allowed). The reference energies
   * (order 0) are the caller's to zero (reference_value). */
  const int32_t* selected;
  int32_t order_lo, order_hi;
} pfw_ace_model;

/* Device tables of a resolved model: adapt_model / pack_model
 * (src/submodules/ACE/adaptation.jl:37-98). The search cutoff is the largest
 * radial cutoff (src/submodules/ACE/cpu.jl:316); the system must have been set
 * with that cutoff. */
int pfw_set_ace_model(pfw_ctx* ctx, const pfw_ace_model* model);
/* New readout parameters; refreshes the precontraction w = C^T WB
 * (NI.update_prepared!, src/submodules/ACE/cpu.jl:330-335; ace_weights!, gpu.jl:172-181). */
int pfw_set_ace_parameters(pfw_ctx* ctx, const double* WB, const double* Wpair);
/* One ACE step on the device: fresh cell list and directed environments,
 * site densities, products, readout and force pullback, force gather over
 * reverse edges (no atomics), FP64 reductions and the status check. Replaces
 * execute_gpu! (src/submodules/ACE/gpu.jl:329-391) and the CPU
 * execute_prepared! (src/submodules/ACE/cpu.jl:253-275). */
int pfw_ace_step(pfw_ctx* ctx, uint32_t request_mask, int precision);
/* pfw_ace_step + outputs (prepared_result, src/submodules/ACE/cpu.jl:279-281). */
int pfw_ace_evaluate(pfw_ctx* ctx, uint32_t request_mask, int precision, double* E, double* F, double* W);
/* Per-site energies of the last ACE step, original order (n doubles). */
int pfw_ace_site_energies(pfw_ctx* ctx, double* out);

/* ---- parameter adjoints (reverse mode over the current system or batch) ----
 * For cotangents dE (one per configuration), dF (3 x N column-major, original
 * order) and dW (3 x 3 column-major per configuration, or NULL for zero), the
 * gradient of L = sum_c dE_c E_c + <dF, F> + sum_c <dW_c, W_c> with respect to
 * the model parameters. Environments are rebuilt and recomputed on the device;
 * nothing from a previous step is reused.
 *
 * ACE: gWB (nbasis x nz, column-major) and gWpair (pair.nr x nz; may be NULL)
 * of the readout WB and pair weights. Replaces the ace_evaluate rrule pullback
 * (ext/PotentialFrameWorkChainRulesCoreExt.jl:54-76) with ace_vjp_impl /
 * scalar_site_product (ext/PotentialFrameWorkForwardDiffExt.jl:63-113) and
 * site_parameter_gradient (src/submodules/ACE/serial.jl:103-121). precision:
 * PFW_FP64, or PFW_MIXED (FP32 site arithmetic, FP64 accumulation). */
int pfw_ace_adjoint(pfw_ctx* ctx, int precision, const double* dE, const double* dF, const double* dW,
                    double* gWB, double* gWpair);
/* SmoothPair: gA (nz x nz) of the amplitude table entry each directed pair
 * reads, FP64. The gradient with respect to the independent upper-triangle
 * amplitudes is gA + gA^T off the diagonal and gA on it. Replaces the
 * pair_evaluate rrule pullback (ext/PotentialFrameWorkChainRulesCoreExt.jl:26-52,
 * pair_vjp in ext/PotentialFrameWorkForwardDiffExt.jl:27-54) and
 * energy_parameter_gradient(::SmoothPair) (src/submodules/PairPotentials/parameters.jl:126-138). */
int pfw_pair_adjoint(pfw_ctx* ctx, const double* dE, const double* dF, const double* dW, double* gA);

/* ---- linear fitting design (design_matrix_impl, src/submodules/ACE/fitting.jl:18-63) ----
 * The outputs of nsets ACE parameter sets over the current system or batch,
 * on one environment build. Set k is WB + k nbasis nz (nbasis x nz,
 * column-major) and Wpair + k nq nz (pair.nr x nz; may be NULL when the
 * model has no pair functions). Column k of out (rows x nsets, column-major,
 * rows = 10 ncfg + 3 N) holds, for every configuration in order, its energy,
 * its forces vec(F) (3 x n, original order) and its virial vec(W) (3 x 3,
 * column-major): the row layout of design_matrix_impl. The outputs are linear
 * in the parameters, so unit sets give the design's columns plus the
 * baseline (the zero set's column). The parameters set with
 * pfw_set_ace_parameters stay in force; the last step's results are cleared.
 * precision: PFW_FP64 or PFW_MIXED. */
int pfw_ace_parameter_sets(pfw_ctx* ctx, int precision, int64_t nsets, const double* WB, const double* Wpair,
                           double* out);

/* The basis columns of the linear design in one pass, by per-site Jacobians
 * (design_matrix_impl builds each site's Jacobian once, src/submodules/ACE/fitting.jl:18-63):
 * out (rows x ld, row-major, ld >= nbasis nz: the first nbasis nz columns of a
 * possibly wider matrix; the rest of each row is overwritten with unspecified
 * values for the caller to fill; rows as for pfw_ace_parameter_sets: per
 * configuration E, vec(F), vec(W)) gets, in column k + nbasis z, the energy,
 * forces and virial of basis function k summed over the sites of centre
 * species z, times basis_scale[k] (the readout mask; NULL = all ones). These
 * are the columns of vec(WB) of the design, without the baseline (reference
 * energies and the pair term; pfw_ace_parameter_sets gives the baseline and
 * the Wpair columns). Species-projected bases return PFW_INVALID_ARGUMENT
 * (their columns come from pfw_ace_parameter_sets). precision: PFW_FP64, or
 * PFW_MIXED (FP32 site arithmetic, FP64 sums). */
int pfw_ace_design(pfw_ctx* ctx, int precision, const double* basis_scale, double* out, int64_t ld);

/* ---- equivariant features (ace_features, src/submodules/ACE/features.jl:141-178) ---- */
enum { PFW_RADIAL_SMOOTH_EXPONENTIAL = 1 };
/* Numerical tables of an ACEFeatureLayer (features.jl:9-19) and its
 * parameters. Indices are 0-based. density: neighbour species channel
 * dens_z (0 = every species, 1..nz = that species), radial dens_n and
 * harmonic dens_h (l*l+l+m); products: factors prod_idx[prod_ptr[k] ..
 * prod_ptr[k+1]); coupling: the rows of every block in turn ((2L+1) x nbasis
 * rows per block, m fastest) as CSR (row_ptr, col = product, val); mixing:
 * per block (nbasis x channels x nz), column-major, blocks in turn. Radials:
 * the smooth exponential family (radials.jl:193-206), amplitude (r/rc)^p
 * (1 - r/rc)^4 exp(-rate r) with rc = radial_cutoff and one power per radial,
 * times scale (nr x nz x nz: radial, centre, neighbour species, column-major). */
typedef struct {
  int32_t nz, nr, lmax, ndens, nprod, nblocks;
  const int32_t *dens_z, *dens_n, *dens_h;
  const int32_t *prod_ptr, *prod_idx;
  const int32_t *block_L, *block_channels, *block_nbasis;
  const int32_t *row_ptr, *col;
  const double* val;
  const double* mixing;
  double cutoff;        /* the layer's neighbour cutoff; the system must be set with it */
  double radial_cutoff; /* the radial family's cutoff (<= cutoff) */
  int32_t radial_kind;  /* PFW_RADIAL_SMOOTH_EXPONENTIAL */
  const int32_t* powers;
  double rate, amplitude;
  const double* scale;
} pfw_feature_layer;
/* Features of every atom of the current system (one configuration): for each
 * block in turn a (2L+1) x channels x N array, column-major, original atom
 * order, FP64. Replaces ace_features (features.jl:164-178) over a fresh
 * cell list of the layer's cutoff. Coincident atoms give PFW_DOMAIN. */
int pfw_features(pfw_ctx* ctx, const pfw_feature_layer* layer, double* out);
/* Pullback of pfw_features for cotangents laid out as its output: replaces
 * feature_vjp (features.jl:204-225) with feature_local_reverse (:227-255).
 * dx: 3 x N (column-major, original order), the gradient of
 * sum(cotangents . features) with respect to the positions; dmixing: laid
 * out as layer->mixing; dscale: as layer->scale; dradial[2]: the rate and
 * amplitude gradients. Coincident atoms give PFW_DOMAIN; nonfinite
 * cotangents PFW_INVALID_ARGUMENT. */
int pfw_features_vjp(pfw_ctx* ctx, const pfw_feature_layer* layer, const double* cotangents, double* dx,
                     double* dmixing, double* dscale, double* dradial);

#ifdef __cplusplus
}
#endif
#endif /* PFW_MI355X_H */
