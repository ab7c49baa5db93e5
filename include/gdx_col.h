/* gdx_col.h -- C ABI of the MI355X swarm collision operator.
 *
 * Drop-in boundary for GameDev's swarm collision operator (bijiaqing/GameDev). Each entry point names
 * the reference interface it replaces. All state pointers are device pointers on the current HIP
 * device; sizes are element counts. Every function returns a status code (GDX_COL_OK on success) and
 * keeps a message for gdx_col_last_error(). Calls on one handle must not overlap.
 *
 * One library is built per compile-time model directory (gamedev_amd/mod/<model>/const_defs.h), the
 * analogue of the reference's mod/<model>/const_defs.cuh (Makefile:129-151, README.md:201-321):
 * N_K, the coagulation kernel, the disk and controller constants are constexpr in the library.
 */
#ifndef GDX_COL_H
#define GDX_COL_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: 1-6 are the chain errors of col_chain_run (src/swarm/col_chain_run.cu:258,264,
 *      280,299,327,372) thrown as "local collision chain error N" (inc/swarm/_col_chain.cuh:403-405) */
#define GDX_COL_OK                 0
#define GDX_COL_ERR_CHAIN_RATE     1   /* invalid pair rate (NaN or negative) */
#define GDX_COL_ERR_CHAIN_TOTAL    2   /* invalid total rate */
#define GDX_COL_ERR_CHAIN_WAIT     3   /* invalid waiting time */
#define GDX_COL_ERR_CHAIN_CLOCK    4   /* event time does not advance */
#define GDX_COL_ERR_CHAIN_PARTNER  5   /* no positive partner slot */
#define GDX_COL_ERR_CHAIN_MASS     6   /* mass conservation or invalid outcome */
#define GDX_COL_ERR_CONTINUATION   7   /* continuation limit (_col_chain.cuh:381-384) */
#define GDX_COL_ERR_RATE_MOMENTS   8   /* invalid change-based rate moments (_col_bound.cuh:22-33, 111) */
#define GDX_COL_ERR_AUDIT          9   /* invalid audit state / no active mass (_col_bound.cuh:136-146; _col_chain.cuh:434) */
#define GDX_COL_ERR_SCHEDULE      10   /* scheduler horizon or time resolution (_col_sched.cuh:30-49, 135-143) */
#define GDX_COL_ERR_MORTON_OVERFLOW 11 /* Morton stack or node-index overflow (swarm_runtime.cu:441-444; morton_index.cuh:215) */
#define GDX_COL_ERR_MORTON_ARGS   12   /* invalid Morton arguments (morton_index.cuh:172-176) */
#define GDX_COL_ERR_ALLOC         13   /* device allocation failure */
#define GDX_COL_ERR_NONFINITE     14   /* non-finite particle state (swarm_runtime.cu:361-411) */
#define GDX_COL_ERR_ARG           15   /* invalid argument */
#define GDX_COL_ERR_STATE         16   /* call order: geometry or state missing */
#define GDX_COL_ERR_HIP           17   /* HIP runtime failure (inc/gpu.cuh:103-120 exits instead) */

/* ---- flags ---- */
#define GDX_COL_DEFER_WRITEBACK 1      /* keep size/number resident between operator calls; gdx_col_store_state writes back */
#define GDX_COL_PROFILE         2      /* time every operator kernel with HIP events on the launch stream */
#define GDX_COL_DIAGNOSTICS     4      /* the reference's COL_DIAGNOSTICS: controller archive, per-operator log, event categories */

/* neighbour search of the geometry epoch (COLLISION_SEARCH=morton|kdtree) */
#define GDX_COL_SEARCH_MORTON   0      /* adaptive Morton hierarchy, rows in (d^2, code) order */
#define GDX_COL_SEARCH_KDTREE   1      /* left-balanced k-d tree, rows in heap order (_col_cache.cuh:15-71) */

/* kernels timed under GDX_COL_PROFILE */
#define GDX_COL_KERNEL_PUBLISH   0
#define GDX_COL_KERNEL_RATE      1
#define GDX_COL_KERNEL_SCREEN    2
#define GDX_COL_KERNEL_CHAIN     3
#define GDX_COL_KERNEL_AUDIT     4
#define GDX_COL_KERNEL_MAJORANT  5   /* learned: features and network per due owner */
#define GDX_COL_KERNEL_CANDIDATES 6  /* learned: exact K-sum and first event of each thinning candidate */
#define GDX_COL_KERNEL_STALE     7   /* learned: stale marking of the owners whose inputs changed */
#define GDX_COL_KERNEL_PREFIX    8   /* per-group queue prefix of a chain phase */
#define GDX_COL_KERNELS          9

typedef struct gdx_col_ctx *gdx_col_handle;

/* compile-time constants of the model directory this library was built for */
typedef struct {
    char model[32];
    int n_k, dim, coag_kernel, n_x, n_y, n_z;
    int local_groups, col_bin_s, morton_leaf_target, morton_max_level;
    double x_min, x_max, y_min, y_max, z_min, z_max;
    double h_search, v_frag, col_bath_max, col_bath_eps, col_bath_alpha, init_smin, rho_0;
} gdx_col_model;

typedef struct {
    int n;                          /* representatives: the runtime N_P (const_defs.cuh:39) */
    double total_dust_mass;         /* lambda_0 = N / N_K / total_dust_mass (_col_chain.cuh:187) */
    unsigned long long seed;        /* random-number key; replaces rngstate_init (src/swarm/rngstate_init.cu:11-18) */
    int flags;                      /* GDX_COL_DEFER_WRITEBACK */
    int search;                     /* GDX_COL_SEARCH_*: the build's COLLISION_SEARCH (Makefile:129-151) */
} gdx_col_config;

typedef struct {
    int waves;                      /* refresh waves: the reference's count_col increments (_col_chain.cuh:480) */
    int baths;                      /* audited group baths */
    long long owner_updates;        /* owners advanced over all waves */
    long long active_owners;        /* active representatives */
    long long events;               /* accepted collision events */
    long long chain_owners;         /* owners that entered the event chain */
    long long pair_evals;           /* pair-rate evaluations */
    double dt_col;                  /* smallest bath duration of the last wave (_col_chain.cuh:481-482) */
    double clock_dyn;               /* operator clock at return: the horizon */
    long long max_chain_events;     /* most events of one owner in one bath (the longest serial chain) */
    /* learned majorant (gdx_col_use_majorant; zero for the plain method) */
    long long thinned;              /* owner baths screened against the majorant instead of the exact rate */
    long long candidates;           /* thinning candidates inside the bath */
    long long accepted;             /* candidates accepted as events */
    long long exact_checks;         /* owner baths whose exact rate was compared with the majorant */
    long long violations;           /* of those, exact rate above the majorant */
    long long rated_owners;         /* owner baths rated with the exact K-sum by rate passes */
    long long candidate_pairs;      /* of pair_evals, those of the thinning-candidate pass */
    long long bound_evals;          /* owner bounds evaluated by the network (kept bounds excluded) */
} gdx_col_stats;

/* library and model */
int gdx_col_model_info(gdx_col_model *out);
const char *gdx_col_last_error(gdx_col_handle h);

/* construction: replaces the device allocations of main (src/swarm/swarm_runtime.cu:121-184) and
 * local_workspace (inc/swarm/_col_chain.cuh:36-96) */
int gdx_col_create(const gdx_col_config *cfg, gdx_col_handle *out);
int gdx_col_destroy(gdx_col_handle h);

/* state binding: particle-order device arrays. SoA fields (positions are azimuth x, spherical radius y,
 * polar angle z as in swarm.position, const_defs.cuh:194-203) or the reference's 64-byte swarm records */
int gdx_col_bind_soa(gdx_col_handle h, const double *x, const double *y, const double *z,
                     double *size, double *numr);
int gdx_col_bind_swarm(gdx_col_handle h, void *swarm_records);
int gdx_col_set_stream(gdx_col_handle h, void *hip_stream);

/* geometry epoch: col_site_init + morton_ghost_index::build + col_cache_get + col_space_bin +
 * col_dep_graph (src/swarm/swarm_runtime.cu:384-448; _col_chain.cuh:161-180). Also loads the state. */
int gdx_col_geometry(gdx_col_handle h);
/* finiteness screen of a reused geometry epoch (colstate_flag, src/swarm/colstate_flag.cu:13-29) */
int gdx_col_check_state(gdx_col_handle h);
/* gather bound size/number into the operator layout, or write them back */
int gdx_col_load_state(gdx_col_handle h);
int gdx_col_store_state(gdx_col_handle h);

/* Morton refinement threshold of the next geometry epochs (default: the model's MORTON_LEAF_TARGET,
 * morton_index::build's leaf_target argument, morton_index.cuh:168-229). Neighbour rows, slot order and
 * every operator result do not depend on it; only the search time does. */
int gdx_col_set_leaf_target(gdx_col_handle h, int leaf_target);
/* kernel timing of later operator calls on or off (the GDX_COL_PROFILE flag of the configuration) */
int gdx_col_set_profile(gdx_col_handle h, int on);
/* autotuner (no reference counterpart; the reference fixes MORTON_LEAF_TARGET per model): times the
 * geometry epoch of the bound state for each candidate leaf target (best of `repeats`, host wall clock
 * around a synchronised epoch), keeps the fastest, and leaves a valid epoch built with it. A deferred
 * state is written back first. ms (may be NULL) gets `count` times. */
int gdx_col_autotune(gdx_col_handle h, const int *candidates, int count, int repeats, int *best, double *ms);
/* top-K selection of the Morton search, no reference counterpart (the reference sorts each 64-record tile
 * with a block-wide bitonic merge, morton_index.cuh:353-455): 0 inserts every eligible record into the
 * query lane's sorted register list; 8, 16 or 32 collects eligible records per lane in LDS and merges them
 * that many at a time by register sorting networks. Rows are identical for every choice; only the epoch's
 * time differs. The autotuner times the epoch at each candidate tile (at the current leaf target) and
 * keeps the fastest. */
int gdx_col_set_topk_tile(gdx_col_handle h, int tile);
int gdx_col_autotune_tile(gdx_col_handle h, const int *candidates, int count, int repeats, int *best, double *ms);
/* block shape of the Morton search, no reference counterpart (the reference launches one block per
 * 64-query tile): 1, 2 or 4 wavefronts per block of the top-K kernel, each wavefront with its own LDS
 * stack and tile. Rows are identical for every shape; the autotuner times the epoch at each candidate (at
 * the current leaf target and tile) and keeps the fastest. */
int gdx_col_set_topk_block(gdx_col_handle h, int waves);
int gdx_col_autotune_block(gdx_col_handle h, const int *candidates, int count, int repeats, int *best, double *ms);

/* one operator call over `duration`: evolve_local_collisions (inc/swarm/_col_chain.cuh:108-539) */
int gdx_col_operator(gdx_col_handle h, double duration, gdx_col_stats *out);
/* one operator call on each of `count` independent swarms (realizations, each with its own handle,
 * geometry, state, controller and random key): the swarms' due groups share every kernel launch and
 * host synchronisation; results equal separate gdx_col_operator calls. No reference counterpart (the
 * reference runs one swarm per process); for ensembles of small realizations. Launches go to the
 * first handle's stream; out holds `count` entries. */
int gdx_col_operator_batch(const gdx_col_handle *handles, int count, double duration, gdx_col_stats *out);
/* no reference counterpart: `steps` operator calls of `duration` on every swarm of an ensemble of
 * independent realizations. The swarms split into `streams` groups (<= 0: 4; at most 8), dealt in turn
 * in order of the refresh waves of their last call (before any call: of their active-owner count), so
 * the heaviest land on different groups. Each group runs its batched calls on its own HIP stream from its
 * own host thread and advances through the steps without waiting for the other groups, so that one
 * realization's long sequence of refresh waves overlaps the others' work. Results equal separate
 * gdx_col_operator calls; the handles' own streams are synchronised first and not used. out (may be NULL)
 * holds steps x count entries, step-major (out[s * count + i]). */
int gdx_col_operator_ensemble(const gdx_col_handle *handles, int count, double duration, int steps, int streams,
                              gdx_col_stats *out);
/* controller memory reset at output boundaries (src/swarm/swarm_runtime.cu:491-496); also resets the
 * adaptive scale of an attached majorant's bound */
int gdx_col_reset_controller(gdx_col_handle h);
/* test access, no reference counterpart: > 0 makes every bath of later operator calls last
 * min(duration, horizon) instead of the controller's choice (audits still run); 0 restores the
 * controller. Two operators with the same fixed bath follow the same schedule whatever their rates. */
int gdx_col_set_fixed_bath(gdx_col_handle h, double duration);
/* launch shapes, no reference counterpart: refresh waves with at most `owners` due owners rate each
 * plain owner with a group of four lanes and screen one owner per lane, so that the few owners of a small
 * wave spread over the whole GPU; larger waves take one lane per owner and four owners per lane
 * (default GDX_COL_NARROW_OWNERS; 0 always takes the wide shapes, < 0 is rejected). In a batch or an
 * ensemble the first handle's value applies. Results do not depend on it. */
/* test access, no reference counterpart: on makes the chain pick every partner from the walk's
 * sequential partial sums; off (default) decides each slot from tree-summed partials wherever they are
 * provably on the same side of the target and falls back to the sequential ones otherwise. Results do
 * not depend on it. */
int gdx_col_set_exact_walk(gdx_col_handle h, int on);
#define GDX_COL_NARROW_OWNERS 65536
int gdx_col_set_narrow_waves(gdx_col_handle h, int owners);
int gdx_col_synchronize(gdx_col_handle h);

/* ---- diagnostics (handles created with GDX_COL_DIAGNOSTICS) ---- */
/* open <dir>collision_local_<microseconds>.jsonl, one schema-1 line per operator call
 * (local_workspace, inc/swarm/_col_chain.cuh:70-75, 488-538) */
int gdx_col_open_log(gdx_col_handle h, const char *dir);
/* simulation clock written into the next log lines (clock_sim of evolve_local_collisions) */
int gdx_col_set_clock(gdx_col_handle h, double clock_sim);
/* controller archive since the last gdx_col_reset_controller, schema 2 (save_col_controller,
 * inc/swarm/_col_bound.cuh:242-299; collision_chain_<frame>.json, swarm_runtime.cu:668-675) */
int gdx_col_save_controller(gdx_col_handle h, const char *file_name);
/* event counts and summed log mass change per outcome category over the last operator call
 * (col_event_sum, src/swarm/col_event_sum.cu:12-44): sticking q <= 1e-6, <= 1e-4, <= 1e-2, larger;
 * fragmentation; remnant erosion; debris */
int gdx_col_event_totals(gdx_col_handle h, unsigned long long counts[7], double log_mass[7]);
/* accumulated kernel time (ms), launches and owners per GDX_COL_KERNEL_* since the last reset; no
 * reference counterpart (the reference's COL_PERF_VAL hooks time host phases, _col_chain.cuh:215-241) */
int gdx_col_kernel_times(gdx_col_handle h, double *ms, long long *launches, long long *owners);
int gdx_col_reset_kernel_times(gdx_col_handle h);

/* ---- inspection (host buffers, particle order) ---- */
int gdx_col_active_count(gdx_col_handle h, int *n_active);
/* neighbour codes 3*j + image (-1 empty; image 0 unless the model is an azimuthal wedge: 2 = the
 * partner rotated by +width, 1 = by -width), n x N_K in (d^2, code) order, and the KNN measure (0 if
 * inactive); the col_cache_get output (inc/swarm/_col_cache.cuh:135-155) */
int gdx_col_get_neighbors(gdx_col_handle h, int *codes, double *measure);
/* indexed records of the last epoch (Morton search: n, plus periodic ghosts in a wedge; k-d search: n,
 * or 3n in a wedge), whether the images of one particle are disjoint under the largest search cap (no
 * deduplication needed) and that cap; the record count, unique_ids and max_search_dist of
 * morton_ghost_index::build (inc/swarm/morton/morton_ghost.cuh:96-168, src/swarm/swarm_runtime.cu:
 * 413-415) and N_T (val/swarm/src/const_defs.cuh:240) */
int gdx_col_record_info(gdx_col_handle h, int *records, int *unique_ids, float *q_max);
/* collision sites: float {x, y, z, search cap} x n, active flags, spatial group */
int gdx_col_get_sites(gdx_col_handle h, float *pts4, unsigned char *active, unsigned char *group);
/* k-d tree of the last geometry epoch (search GDX_COL_SEARCH_KDTREE): record index and split dimension
 * per node (one per record of gdx_col_record_info: in a wedge record r is particle r % n at image
 * r / n) and the world box {lo xyz, hi xyz}; test access to kdtree::buildTree
 * (inc/swarm/kdtree/builder_thrust.h:418-526) */
int gdx_col_get_kdtree(gdx_col_handle h, int *node_idx, int *node_dim, float *world6);
/* Morton keys and record order after the stable sort (one per indexed record: n, or the record count
 * of gdx_col_record_info in a wedge; record i >= n is a ghost) */
int gdx_col_get_morton(gdx_col_handle h, unsigned long long *sorted_keys, int *sorted_idx);
/* bath-start rate moments {rate (-1 invalid), first, second, maximum} per particle (0 for inactive) */
int gdx_col_get_cached(gdx_col_handle h, double *cached4);
/* per-particle random-number draw counters (checkpoint/restore) */
int gdx_col_get_rng(gdx_col_handle h, unsigned long long *counters);
int gdx_col_set_rng(gdx_col_handle h, const unsigned long long *counters);
/* dependency graph rows: one 64-bit group mask per group */
int gdx_col_get_graph(gdx_col_handle h, unsigned long long *rows);
/* bath-start rating pass only (publish + col_bath_rate for every owner); for rate parity checks */
int gdx_col_rate_pass(gdx_col_handle h);
/* every owner's K pair rates at the bath start, host n x N_K in particle order and slot order: the terms
 * col_bath_rate sums (_get_col_chain_rate / measure x outcome factor, src/swarm/col_bath_rate.cu:30-58;
 * inc/swarm/_col_rates.cuh:167-223), 0 for empty slots and inactive owners. Runs gdx_col_rate_pass
 * first, so the reservoir is published from the owners' states. */
int gdx_col_pair_rates(gdx_col_handle h, double *rates);

/* ---- host controller and scheduler, callable without a GPU (same code the operator runs) ---- */
/* _build_col_binmap (inc/swarm/_col_bound.cuh:45-98): returns the merged-bin count */
int gdx_col_host_binmap(const int *raw_count, int raw_bins, int *raw_to_merged, int *merged);
/* _choose_col_bath (_col_bound.cuh:101-123); bins: n x {mass, w_rate, w_change, w_second, owners, invalid} */
int gdx_col_host_choose(const double *bins6, int n, double remaining, double limit_scale, double *duration);
/* _finish_col_bath (_col_bound.cuh:126-216); audit: n x 15 (col_audit_accum order), state: {limit_scale,
 * activity, quiet, floor}, result: {max_f, max_e, max_touched, max_events, max_g, max_g_upper, d_bath,
 * activity_overshoot, distribution_overshoot, persistent_overshoot} */
int gdx_col_host_finish(const double *audit15, int n, double *state4, double *result10);
/* local_schedule (_col_sched.cuh:17-183) run with fixed requests and passing audits: due ticks per wave */
int gdx_col_host_schedule(int n, double horizon, const double *requested, const unsigned int *edges, int max_waves,
                          unsigned long long *ticks, int *due_counts, int *levels, int *waves);
/* the same with a request and an audit verdict per group and wave (wave_requested, wave_passed: max_waves x n;
 * wave w > 0 adapts its due groups with row w), levels after every wave in wave_levels (max_waves x n) */
int gdx_col_host_schedule_adaptive(int n, double horizon, const double *requested, const unsigned int *edges,
                                   int max_waves, const double *wave_requested, const unsigned char *wave_passed,
                                   unsigned long long *ticks, int *due_counts, int *wave_levels, int *waves);

/* ---- stand-alone adaptive Morton index: morton_index::build / view (inc/swarm/morton/morton_index.cuh:160-293) ---- */
typedef struct gdx_morton_ctx *gdx_morton_handle;
/* points: device float4 {x, y, z, unused} x n; active: device flags or NULL */
int gdx_morton_build(const void *points4, const unsigned char *active, int n, float ox, float oy, float oz,
                     float width, int dim, int leaf_target, int max_level, gdx_morton_handle *out);
int gdx_morton_destroy(gdx_morton_handle h);
int gdx_morton_node_count(gdx_morton_handle h, int *nodes);
/* host copies: sorted keys/source indices (n); nodes in the reference layout (lower xyz, width,
 * begin, count, child[8], child_count) */
int gdx_morton_get(gdx_morton_handle h, unsigned long long *keys, int *idx, float *lower3, float *width,
                   int *begin, int *count, int *child8, int *child_count);
/* exact top-K for device float4 queries {x, y, z, search cap}: device outputs codes (3*idx, -1) and d^2 */
int gdx_morton_knn(gdx_morton_handle h, const void *queries4, int nq, int k, int *codes, float *d2);

/* ---- learned rate majorant: no reference counterpart --------------------------------------------
 * An 8 -> 64 -> 64 -> 1 perceptron (SiLU, softplus output) in BF16 with FP32 accumulation, trained on
 * the GPU. For owner i it predicts t_i = R_i / S_i, the exact bath-start rate R_i (col_bath_rate.cu:
 * 13-129) over a cheap partner sum S_i; the operator then draws event times by Poisson thinning against
 * Lambda_i = safety * S_i * softplus(net(f_i)) and accepts with the exact FP64 rate. Features (device
 * rows of GDX_COL_MAJORANT_FEATURES floats) and targets come from gdx_col_majorant_sample. */
#define GDX_COL_MAJORANT_FEATURES 8
#define GDX_COL_MAJORANT_PARAMS   4801   /* W1[8][64], b1[64], W2[64][64], b2[64], w3[64], b3 */

typedef struct gdx_col_majorant_ctx *gdx_col_majorant;

typedef struct {
    int epochs, batch;              /* full minibatches per epoch: n / batch (the remainder is dropped) */
    double lr, beta1, beta2, eps;   /* Adam; beta1, beta2, eps <= 0 select 0.9, 0.999, 1e-8 */
    unsigned long long seed;        /* sample order (one shuffle per epoch) */
} gdx_col_train_config;

int gdx_col_majorant_create(unsigned long long seed, gdx_col_majorant *out);
int gdx_col_majorant_destroy(gdx_col_majorant m);
const char *gdx_col_majorant_last_error(gdx_col_majorant m);
/* host copies of the FP32 parameters and the input normalization {mean[8], scale[8]} (either may be NULL);
 * set_params resets the Adam moments */
int gdx_col_majorant_get_params(gdx_col_majorant m, float *params, float *norm16);
int gdx_col_majorant_set_params(gdx_col_majorant m, const float *params, const float *norm16);
/* device rows in, device softplus outputs out */
int gdx_col_majorant_infer(gdx_col_majorant m, const float *features, long long n, float *out);
/* gradient (host, GDX_COL_MAJORANT_PARAMS) and value of the mean loss (ln softplus(z) - log_target)^2 */
int gdx_col_majorant_grad(gdx_col_majorant m, const float *features, const float *log_target, int n, float *grad,
                          double *loss);
/* No counterpart in the reference. gdx_col_majorant_grad for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same kernels
 * and the same results, with the per-wavefront partial gradient rows and their sum kept in one fixed
 * device array per GPU (about 9.9 MB, held until the process ends) instead of the model's buffers. n above
 * GDX_COL_MAJORANT_GRAD_NOALLOC_MAX returns GDX_COL_ERR_ARG. Every call on a GPU uses that GPU's array:
 * calls that overlap on one GPU (from several threads or models) race. */
#define GDX_COL_MAJORANT_GRAD_NOALLOC_MAX 65536
int gdx_col_majorant_grad_noalloc(gdx_col_majorant m, const float *features, const float *log_target, int n,
                                  float *grad, double *loss);
/* Adam training over device samples; epoch_loss (host, epochs entries, may be NULL) gets each epoch's
 * mean minibatch loss before the updates */
int gdx_col_majorant_train(gdx_col_majorant m, const float *features, const float *log_target, long long n,
                           const gdx_col_train_config *cfg, float *epoch_loss);
/* training samples of the current state in particle order (device outputs, n rows; inactive rows are
 * zero): features, the scale S_i and the exact bath-start rate R_i (negative when invalid). Publishes the
 * reservoir from the owners' states as gdx_col_rate_pass does. */
int gdx_col_majorant_sample(gdx_col_handle h, float *features, double *scale, double *rate);
/* operator calls of h use the majorant (NULL: the plain method); safety >= 1 multiplies the bound. A
 * swarm with fewer than min_owners active owners (<= 0: GDX_COL_MAJORANT_MIN_OWNERS) keeps the plain
 * method: its exact 1/16 subsample (1/32 from 2^21 active owners) would leave the controller's audit too
 * few owners. The model must
 * outlive its use by h. */
#define GDX_COL_MAJORANT_MIN_OWNERS 65536
int gdx_col_use_majorant(gdx_col_handle h, gdx_col_majorant m, double safety, int min_owners);
/* the most representatives a swarm with a majorant may hold: its reverse neighbour lists (which owners'
 * rows hold each slot) use 32-bit offsets over n * N_K entries, so (2^32 - 1) / N_K; callable without a
 * GPU. gdx_col_use_majorant rejects larger swarms with GDX_COL_ERR_ARG. */
int gdx_col_majorant_max_owners(long long *out);
/* bounds across refresh waves and calls: with reuse on (the default) an owner whose state and partner
 * records have not changed since its last evaluation keeps its bound instead of running the network
 * again (a change of the attached model's parameters invalidates every kept bound); off evaluates every
 * due owner. Results do not depend on it. */
int gdx_col_set_bound_reuse(gdx_col_handle h, int on);

#ifdef __cplusplus
}
#endif

#endif /* GDX_COL_H */
