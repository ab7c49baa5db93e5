// CPU oracle for the swarm collision operator.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product path: it may be
// loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. The product
// (gamedev_amd/) never links, loads or calls it.
//
// It restates, in plain scalar C++ on the host, the reference algorithm of bijiaqing/GameDev's
// collision operator. Every function cites the reference file:line it follows. Arrays are in
// particle order (the reference's dev_particle order); neighbour codes are 3*idx_old + image.
//
// Parity pins (what fixes each function independently of this file):
//   orc_philox4x32          Random123 known-answer vectors (Salmon et al., SC'11)
//   orc_morton_key          bit-by-bit interleave in tests (doc/guide_swarm.md key definition)
//   orc_sort_perm           sortedness + stability properties, topology cases
//   orc_hierarchy_build     independent recursive linear-partition split (val/swarm/mod/test_knn/topology.cuh:70-110)
//   orc_topk_brute          knn edge cases with hand-derived answers (knn_edge_tests.cu:95-210), numpy brute force
//   orc_topk_tree           equality with orc_topk_brute
//   orc_kd_build            left-balanced layout invariants (subtree sizes, every left-subtree
//                           coordinate <= the split and every right one >= it) and split
//                           dimensions equal to the widest extent of each subtree's domain
//   orc_kd_query            neighbour sets equal to orc_topk_brute, rows a valid max-heap
//   orc_image_dist_min      2 y_min sin(z) |sin(width/2)| evaluated in the test
//   orc_wedge_records       hand-computed seam distances, ghost codes and rotation angles (atan2/hypot)
//   orc_topk_records        chord lengths across the seam, the nearest-image rule on a narrow wedge, and
//                           a numpy brute force over all three images of every particle (no seam test,
//                           no candidate cap) at disjoint and shared-ball search caps
//   orc_kd_images           rotation angles +-width in the test
//   orc_kd_query_wedge      sorted rows equal to the same all-images brute force, rows a valid max-heap
//   orc_ball_measure        closed-form interior balls, Monte-Carlo cap volumes
//   orc_vrel_t              Ormel & Cuzzi (2007) equations 17/18/26-29 evaluated independently in tests
//   orc_vrel_pair, orc_stokes, orc_chain_rate
//                           drift, settling (St capped at 0.5), Stokes number, rotation support and the
//                           vertically integrated overlap sqrt(2 pi (H_gi^2 + H_gj^2)) restated in numpy
//                           from the closed forms of doc/guide_swarm.md 8.4.1/8.4.3 and doc/guide_basis.md
//                           3-4, at the colphys probe points with unequal sizes and St on both sides of 0.5
//   orc_partner_walk        known answers on crafted rates and targets (zero slots, exact cumulative ties,
//                           roundoff fallback to the last positive slot) and numpy searchsorted
//   planted mistakes        -DORC_PLANT=1..7 plants a factor, cap, sign, index or comparison error in the
//                           functions above; the pin tests build each variant and require a failure
//   orc_outcome_moments     numerical expectation of orc_sample_outcome over u, closed-form sticking moments
//   controller/scheduler    crafted inputs with hand-derived outputs; tick tiling sums to the horizon
//   orc_operator            mass conservation, repeat determinism, Smoluchowski-type rate checks
#ifndef GDX_ORACLE_H
#define GDX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// compile-time model constants of the model directory this oracle was built for
typedef struct {
    int n_k, dim, coag_kernel, n_x, n_y, n_z, local_groups, col_bin_s, raw_bins;
    int morton_leaf_target, morton_max_level, x_wedge;
    double y_min, y_max, z_min, z_max, x_min, x_max, h_search, v_frag;
    double col_bath_max, col_bath_eps, col_bath_alpha, init_smin, rho_0;
} orc_model_info;
void orc_get_model_info(orc_model_info *out);

// ---- random numbers -------------------------------------------------------------------------
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_uniform(uint32_t pidx, uint64_t seed, uint64_t counter);

// ---- physics (per pair / per owner) -------------------------------------------------------
double orc_vrel_t(double R, double stokes_i, double stokes_j, double h_g, double sigma_g);
double orc_vrel_pair(double y, double z, double size_i, double size_j);
double orc_ball_measure(double y, double z, double radius);
double orc_outcome_moments(double si, double sj, int high_speed, double *mean, double *second, double *maximum);
double orc_sample_outcome(double si, double sj, int high_speed, double u, int *category, double *log_mass);
// one pair's chain rate numerator (_get_col_chain_rate) and relative speed
double orc_chain_rate(double y_i, double z_i, double y_j, double z_j, double size_i, double size_j,
                      double numr_j, double lambda_0, double *vrel);
// Stokes number of a grain of `size` at (y, z) (_get_stokes, param_phys.cuh:163-210)
double orc_stokes(double y, double z, double size);
// the chain's partner pick (col_chain_run.cu:299-320); last_positive = -3 derives it from the rates
int orc_partner_walk(const double *rate, int k_count, double target, int last_positive);

// ---- geometry epoch -------------------------------------------------------------------------
// returns the first bad particle index + 1 (0 when every state is finite)
int orc_site_init(int n, const double *x, const double *y, const double *z,
                  const double *size, const double *numr,
                  float *pts /*3n*/, float *posx, float *search_dist, uint8_t *active);
void orc_morton_keys(int n, const float *pts, float ox, float oy, float oz, float width,
                     int max_level, int dim, uint64_t *keys);
void orc_sort_perm(int n, const uint64_t *keys, int *perm);
// level-synchronous hierarchy in the reference numbering; returns node count or -1 if cap too small
int orc_hierarchy_build(int n, const uint64_t *sorted_keys, float ox, float oy, float oz, float width,
                        int dim, int leaf_target, int max_level, int cap,
                        float *lower /*3cap*/, float *nwidth, int *begin, int *count,
                        int *child /*8cap*/, int *child_count);
// exact top-K by (d2, code) among active records with d2 <= q^2; codes are 3*idx (image 0)
void orc_topk_brute(int n, const float *pts, const uint8_t *active, int nq, const int *query_idx,
                    const float *query_pts, const float *query_q, int k, int *out_codes, float *out_d2,
                    int threads);
// same contract through the Morton hierarchy (the CPU-baseline search)
// azimuthal wedge: seam ghosts (rec, code hold 3n; returns the record count), brute-force top-K over
// coded records (dedup: 3K candidates, each particle at its nearest image), k-d image nodes (3n) and the
// k-d query over them with the per-query physical-id heap
float orc_image_dist_min(void);
int orc_wedge_records(int n, const float *pts, const float *posx, float q_max, float *rec, int *code);
void orc_topk_records(int nr, const float *rec, const int *code, const uint8_t *active, int nq, const float *query_pts,
                      const float *query_q, int k, int dedup, int *out_codes, float *out_d2, int threads);
void orc_kd_images(int n, const float *pts, float *out);
void orc_kd_query_wedge(int n_phys, const float *pts, const uint8_t *active, const int *node_idx,
                        const int *node_dim, const float *world6, int nq, const int *query_idx, int k,
                        const float *search_dist, int *out_codes, float *out_d2, int threads);
void orc_kd_build(int n, const float *pts, int *out_idx, int *out_dim, float *world6);
void orc_kd_query(int n, const float *pts, const uint8_t *active, const int *node_idx, const int *node_dim,
                  const float *world6, int nq, const int *query_idx, int k, const float *search_dist,
                  int *out_codes, float *out_d2, int threads);
void orc_topk_tree(int n, const float *pts, const uint8_t *active, const int *sorted_idx,
                   int node_count, const float *lower, const float *nwidth, const int *begin,
                   const int *count, const int *child, const int *child_count, int dim, int max_level,
                   int nq, const int *query_idx, int k, const float *search_dist,
                   int *out_codes, float *out_d2, int threads);
void orc_space_bin(int n, const double *x, const double *y, const double *z, int *spatial);

// ---- host controller and scheduler -----------------------------------------------------------
int orc_change_limit(double A, double B, double eps, double horizon, double *duration, int *reason);
int orc_build_binmap(const int *raw_count, int raw_bins, int *raw_to_merged);
// bins: n x {mass, weighted_rate, weighted_change, weighted_second, owner_count, invalid_count}
int orc_choose_bath(const double *bins6, int n, double remaining, double limit_scale,
                    double *duration, int *binding);
// audit bins: n x 15 in col_audit_accum field order; state: {limit_scale, activity, quiet, floor}
int orc_finish_bath(const double *audit15, int n, double *state4, double *result10);
// power-of-two schedule; edges: n x ceil(n/32) words. Writes the sequence of due ticks for a run
// with fixed requests and passing audits (test helper): returns waves or -1 on error.
int orc_schedule_run(int n, double horizon, const double *requested, const uint32_t *edges,
                     int max_waves, uint64_t *ticks, int *due_counts, int *levels_out);
int orc_schedule_adaptive(int n, double horizon, const double *requested, const uint32_t *edges, int max_waves,
                          const double *wave_requested, const uint8_t *wave_passed, uint64_t *ticks, int *due_counts,
                          int *wave_levels);

// ---- the operator -----------------------------------------------------------------------------
typedef struct {
    int waves, launches_equiv, errors;
    long long events, owner_updates, pair_evals, baths;
    double dt_col, clock_dyn;
} orc_stats;
// state arrays in particle order; nbr holds codes (3*j) or -1, n*K. rng is per-particle counter.
// bath state (limit_scale etc.) and moving size bounds persist in ctl (opaque, orc_ctl_size bytes).
int orc_ctl_size(void);
void orc_ctl_init(void *ctl);
int orc_operator(int n, const double *y, const double *z, double *size, double *numr,
                 const uint8_t *active, const int *spatial, const int *nbr, const double *measure,
                 uint64_t *rng, uint64_t seed, double duration, double total_dust_mass,
                 void *ctl, int threads, orc_stats *stats, char *err, int err_len, void *diag);
// COL_DIAGNOSTICS bookkeeping of orc_operator (pass the object as `diag`, or null)
void *orc_diag_new(void);
void orc_diag_free(void *d);
void orc_diag_reset(void *d);
int orc_diag_json(void *d, int which, char *buf, int len);
// bath-start rate moments for every owner (col_bath_rate with size_old = size)
void orc_bath_rates(int n, const double *y, const double *z, const double *size, const double *numr,
                    const uint8_t *active, const int *nbr, const double *measure, double lambda_0,
                    double *cached4, int threads);
// every owner's K pair terms at the bath start (size_old = size), n x N_K in slot order
void orc_pair_rates(int n, const double *y, const double *z, const double *size, const double *numr,
                    const uint8_t *active, const int *nbr, const double *measure, double lambda_0, double *rates,
                    int threads);

#ifdef __cplusplus
}
#endif

#endif  // GDX_ORACLE_H
