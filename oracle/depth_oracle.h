/*
 * depth_oracle.h - CPU oracle for the chiaro-fusion depth hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of
 * crates/chiaro-fusion/src/depth.rs and the helpers it calls (geometry.rs,
 * math.rs, image.rs Plane::sample, align.rs Warp), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker
 * and CPU baseline. No product code may link, load or call it.
 *
 * The reference is Rust and no Rust toolchain exists in this environment, so
 * the reference cannot be built; this restatement replaces it. It follows Rust
 * semantics exactly: IEEE f32/f64 without contraction (built with
 * -ffp-contract=off), saturating float->integer casts, Rust's min_by/max_by
 * tie rules, stable sorts and total_cmp.
 *
 * Transcendentals: the reference calls the platform libm (glibc on
 * x86_64-linux-gnu) through f32::exp and f64::exp. or_expf/or_exp restate the
 * algorithm glibc uses (ARM optimized-routines expf/exp, table size 32 and
 * 128); or_set_libm_mode(1) makes the oracle call the platform expf/exp
 * instead. The tests check the two agree bit for bit (exhaustively for expf).
 */
#ifndef CHIARO_DEPTH_ORACLE_H
#define CHIARO_DEPTH_ORACLE_H

#include "../include/chiaro_depth.h"

#ifdef __cplusplus
extern "C" {
#endif

/* transcendental functions */
float or_expf(float x);
double or_exp(double x);
float or_libm_expf(float x);
double or_libm_exp(double x);
void or_set_libm_mode(int use_libm);
/* count of inputs in [lo_bits, hi_bits) (f32 bit patterns, stepping by
 * `stride`) where or_expf differs from the platform expf. */
/* Mismatches of got[i] against or_expf of the float with bits lo_bits + i,
 * i < n (NaN equals NaN), on `threads` workers. */
uint64_t or_expf_compare(uint32_t lo_bits, uint64_t n, const float *got, uint32_t threads);
uint64_t or_expf_mismatches(uint32_t lo_bits, uint32_t hi_bits, uint32_t stride,
                            uint32_t threads);

/* math.rs */
int or_inverse3(const double *m, double *out);
double or_determinant3(const double *m);

/* geometry.rs */
int or_camera_canonical(const double k[9], const double rotation_wc[9],
                        const double translation_wc[3], uint32_t width, uint32_t height,
                        const chd_distortion *distortion, chd_camera *out);
void or_distort(const chd_distortion *d, const double in[2], double out[2]);
void or_undistort(const chd_distortion *d, const double in[2], double out[2]);
void or_pixel_to_ray(const chd_camera *c, const double pixel[2], double origin[3],
                     double direction[3]);
int or_project(const chd_camera *c, const double point[3], int bounded, double out[2]);
int or_map_from(const chd_camera *self, const chd_camera *reference, const double pixel[2],
                double depth, double out[2]);

/* image.rs Plane::sample, align.rs Warp */
int or_plane_sample(const float *data, uint32_t width, uint32_t height, float x, float y,
                    float *out);
int or_warp_map(const chd_warp *warp, float x, float y, float out[2]);
/* align_module(refine = false) warp: an 8 px infinity-map grid resampled to
 * grid_step (align.rs:317-324, 456-459). points/confidence caller-allocated
 * for (ceil(width/grid_step)+1) x (ceil(height/grid_step)+1) nodes. */
int or_global_warp(const chd_camera *reference, const chd_camera *target, uint32_t width,
                   uint32_t height, uint32_t grid_step, chd_warp *out);

/* depth.rs pieces */
int or_options_valid(const chd_options *o);
void or_default_options(chd_options *o);
void or_inverse_depth_samples(double near_depth, double far_depth, uint32_t count,
                              double *out);
double or_sublabel_depth(const double *labels, uint32_t label_count, const float *costs,
                         uint32_t best_label);
int or_aggregate(const float *scores, uint32_t count, const chd_options *o, float *out);
int or_warped_patch_zncc(const float *reference, uint32_t rw, uint32_t rh,
                         const float *target, uint32_t tw, uint32_t th,
                         const chd_warp *global, double cx, double cy, double dx,
                         double dy, uint32_t radius, float *out);
void or_semi_global_costs(const float *data, const float *guidance, uint32_t columns,
                          uint32_t rows, uint32_t labels, float *output);
/* diagonal_paths (depth.rs:1313-1338): writes path start offsets
 * (count+1 entries) and node indices; returns the number of paths. */
uint32_t or_diagonal_paths(uint32_t columns, uint32_t rows, int mirrored,
                           uint32_t *offsets, uint32_t *nodes);
void or_select_depths(const double *label_depths, uint32_t labels, const float *scores,
                      const uint8_t *tested, const float *regularised, uint32_t nodes,
                      const chd_options *o, chd_node_depth *field, uint8_t *fillable,
                      uint16_t *best_label);
void or_complete_depth_field(chd_node_depth *field, const float *guidance,
                             const uint8_t *fillable, uint32_t columns, uint32_t rows,
                             const chd_options *o);
uint32_t or_direct_depth_candidates(int seeded, double seed_depth, double inverse_step,
                                    const chd_options *o, double *out);
int or_select_direct_depth(const double *depths, const float *scores,
                           const uint8_t *has_score, uint32_t count, int has_baseline,
                           float baseline, int seeded, const chd_options *o,
                           chd_node_depth *out);
int or_best_depth_candidate(const double *depths, const float *scores,
                            const uint8_t *has_score, uint32_t count, const chd_options *o);
void or_reject_isolated_direct_depths(chd_node_depth *field, const float *guidance,
                                      uint32_t columns, uint32_t rows,
                                      double inverse_tolerance);
void or_reject_small_direct_components(chd_node_depth *field, const float *guidance,
                                       uint32_t columns, uint32_t rows,
                                       double inverse_tolerance, uint32_t minimum_nodes);
void or_fit_local_depth_planes(chd_node_depth *field, const float *guidance,
                               uint32_t columns, uint32_t rows, double inverse_tolerance);
int or_warp_decisions_agree(uint32_t kind, const chd_node_warp *centre,
                            const chd_node_warp *neighbour);
void or_enforce_warp_consensus(chd_node_warp *decisions, uint32_t columns, uint32_t rows);
void or_suppress_warp_boundaries(chd_node_warp *decisions, const float *guidance,
                                 uint32_t columns, uint32_t rows);
void or_diagnostic_samples(const chd_depth_map *map, uint16_t *inverse_depth,
                           uint16_t *provenance_rgb);
void or_visualization_samples(const chd_depth_map *map, uint16_t *rgb);

/* Mosaic::luminance_half (image.rs:401-426) with a unit flat field, after
 * reserve_highlight_headroom (image.rs:69-75) when reserve_headroom != 0;
 * always with the platform logf. */
void or_luminance_half_u16(const uint16_t *samples, uint32_t width, uint32_t height,
                           float black_q6, float white_q6, int reserve_headroom, float *plane);

/* Intermediate results of the full path (any pointer may be NULL). */
typedef struct or_debug {
    float *scores;       /* coarse nodes x labels */
    float *costs;        /* coarse nodes x labels */
    float *guidance;     /* coarse nodes */
    uint8_t *tested;     /* coarse nodes */
    float *aggregated;   /* coarse nodes x labels */
    chd_node_depth *coarse;
    uint8_t *fillable;
    uint16_t *best_label;
    chd_node_depth *direct[4]; /* fine field: measured, isolated, components, planes */
    float *fine_guidance;
    uint8_t *fine_tested;
    double seconds[5];   /* cost, sgm+select, direct, filters, targets */
} or_debug;

/* refine_multiview_depth (depth.rs:312-580) with the reference's row-band
 * threading over `threads` workers (0 = hardware concurrency). Setting
 * stop_after_sgm ends after select_depths (for timing the cost+SGM stage). */
int or_refine_multiview_depth(const chd_view *views, uint32_t count,
                              uint32_t reference_index, const chd_options *o,
                              chd_depth_map *map, chd_target_result *targets,
                              or_debug *debug, uint32_t threads, int stop_after_sgm);

/* build_cost_volume restricted to coarse rows [row_begin, row_end) (the
 * bounded CPU-baseline sample); outputs are row-band slices. */
int or_cost_volume_rows(const chd_view *views, uint32_t count, uint32_t reference_index,
                        const chd_options *o, uint32_t row_begin, uint32_t row_end,
                        uint32_t threads, float *scores, float *costs, float *guidance,
                        uint8_t *tested);

/* measure_direct_depths (depth.rs:715-876) from a given coarse field over
 * fine rows [row_begin, row_end); field, guidance and tested are whole
 * fine-grid arrays of which only those rows are written. */
int or_direct_rows(const chd_view *views, uint32_t count, uint32_t reference_index,
                   const chd_options *o, const chd_node_depth *coarse, uint32_t row_begin,
                   uint32_t row_end, uint32_t threads, chd_node_depth *field, float *guidance,
                   uint8_t *tested);

/* align_module's refinement (align_oracle.c; align.rs:270-519, 644-793,
 * image.rs:469-498, 531-640, math.rs:131-138). */
int or_match_patch(const float *reference, uint32_t rw, uint32_t rh, const float *target,
                   uint32_t tw, uint32_t th, uint32_t rx, uint32_t ry, uint32_t size,
                   uint32_t radius, float shift[2], float *score);
int or_apply_homography(const double *h, const double p[2], double out[2]);
int or_fit_homography_least_squares(const double *pairs, const float *scores, uint32_t count,
                                    double h[9]);
int or_fit_homography_ransac(const double *pairs, const float *scores, uint32_t count,
                             float threshold, double h[9], uint32_t *inliers, float *residuals);
uint32_t or_pyramid_levels(uint32_t width, uint32_t height, uint32_t min_size);
void or_default_align_options(chd_align_options *o);
int or_align_module(const chd_view *reference, const chd_view *target,
                    const chd_align_options *o, chd_warp *warp, chd_alignment_report *report);
int or_refine_resolution_warp(const float *reference, uint32_t rw, uint32_t rh,
                              const float *target, uint32_t tw, uint32_t th, const chd_warp *base,
                              uint32_t width, uint32_t height, chd_warp *out,
                              chd_resolution_report *report);
int or_sample_nearest(const chd_depth_map *map, float x, float y, size_t *index);
int or_band_touches(const chd_synth_source *source, size_t start, size_t end,
                    const chd_canvas *canvas);
int or_canvas_size(const chd_canvas *canvas, uint32_t *width, uint32_t *height);
int or_synthesis_geometry(const chd_depth_map *map, const chd_synth_source *sources,
                          uint32_t count, const chd_canvas *canvas, float *points, float *feather,
                          float *local_confidence, float *base_weight, float *focus_weight);

#ifdef __cplusplus
}
#endif

#endif
