/*
 * chiaro_depth.h - C ABI of the MI355X (gfx950) depth hot path.
 *
 * The path is chiaro-fusion's calibrated multi-view depth refinement:
 * plane-sweep ZNCC cost field -> 8-path semi-global matching -> winner-take-all
 * with sub-label parabola -> direct remeasurement on the fine grid ->
 * consistency filters -> per-target cross-camera decisions.
 *
 * Every entry point is plain C: pointers, sizes and POD structs, no C++ or
 * torch types. Errors never cross the boundary as exceptions: functions return
 * CHD_OK (0), CHD_NONE (1, the reference's `None`), or a negative CHD_ERR_*
 * code, and chd_last_error() holds a message for the calling thread.
 *
 * Reference interface (paths relative to the chiaro repository):
 *   refine_multiview_depth  crates/chiaro-fusion/src/depth.rs:312-317
 *   DepthOptions            crates/chiaro-fusion/src/depth.rs:40-91
 *   DepthAlignmentReport    crates/chiaro-fusion/src/depth.rs:93-113
 *   DenseDepthMap / Node    crates/chiaro-fusion/src/depth.rs:115-141
 *   AlignInput              crates/chiaro-fusion/src/align.rs:247-257
 *   Warp                    crates/chiaro-fusion/src/align.rs:35-45
 *   ResolvedCamera          crates/chiaro-fusion/src/geometry.rs:31-62
 */
#ifndef CHIARO_DEPTH_H
#define CHIARO_DEPTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHD_ABI_VERSION 1

#define CHD_OK 0
#define CHD_NONE 1
#define CHD_ERR_ARGUMENT (-1)
#define CHD_ERR_HIP (-2)
#define CHD_ERR_NO_DEVICE (-3)
#define CHD_ERR_LIMIT (-4)
#define CHD_ERR_STATE (-5)

/* Largest supported patch radius of the ZNCC kernels (the direct pass uses
 * max(2 r, 8), so r <= 8 keeps every pass within this bound). */
#define CHD_MAX_PATCH_RADIUS 16
/* Largest number of views (reference included) in one call. Each
 * luminance plane must also stay below 4 GiB as (width + 1) x (height + 1)
 * f32 (CHD_ERR_LIMIT otherwise). */
#define CHD_MAX_VIEWS 32

/* DepthOptions (depth.rs:40-67); defaults from chd_default_options
 * (depth.rs:69-91). */
typedef struct chd_options {
    int32_t enabled;
    uint32_t grid_step;
    double near_depth;
    double far_depth;
    uint32_t planes;
    uint32_t patch_radius;
    uint32_t minimum_support;
    uint32_t best_view_count;
    float minimum_score;
    float minimum_margin;
    float minimum_improvement;
    uint32_t minimum_neighbour_support;
    uint32_t completion_iterations;
} chd_options;

/* Brown distortion k1, k2, p1, p2, k3 (calibration.rs PolynomialDistortion;
 * geometry.rs:303-370). Missing coefficients count as 0. */
typedef struct chd_distortion {
    double center[2];
    double normalization[2];
    double coeffs[5];
    uint32_t coeff_count;
} chd_distortion;

/* Plain-data form of ResolvedCamera (geometry.rs:31-62). Matrices are
 * row-major m[row * 3 + column]. pose_kind 0 = Canonical
 * (rotation_wc/rotation_cw/translation_wc/center), 1 = Mirror
 * (real_cw/reflect/virtual_center). flip_around_x: -1 = None, 0 = Some(false),
 * 1 = Some(true). */
typedef struct chd_camera {
    uint32_t width;
    uint32_t height;
    double k[9];
    double k_inverse[9];
    int32_t has_distortion;
    chd_distortion distortion;
    int32_t flip_around_x;
    int32_t pose_kind;
    double rotation_wc[9];
    double rotation_cw[9];
    double translation_wc[3];
    double center[3];
    double real_cw[9];
    double reflect[9];
    double virtual_center[3];
    double orientation_correction[9];
    double focal_px;
} chd_camera;

/* Warp (align.rs:35-45): columns*rows target points, (x, y) interleaved,
 * NaN where the mapping is undefined, and one confidence per node. */
typedef struct chd_warp {
    uint32_t step;
    uint32_t columns;
    uint32_t rows;
    float *points;
    float *confidence;
} chd_warp;

/* One module: AlignInput (align.rs:247-257) plus the fields of its
 * ModuleAlignment that the path reads (warp, report.accepted). */
typedef struct chd_view {
    const float *luminance; /* plane_height rows of plane_width f32 */
    uint32_t plane_width;
    uint32_t plane_height;
    uint32_t width; /* raster size */
    uint32_t height;
    const chd_camera *camera; /* NULL = no calibration (None) */
    chd_warp global_warp;     /* alignments[i].warp on entry */
    int32_t accepted;         /* alignments[i].report.accepted */
} chd_view;

/* DepthAlignmentReport (depth.rs:93-113). */
typedef struct chd_depth_report {
    uint64_t tested_nodes;
    uint64_t measured_nodes;
    uint64_t regularized_nodes;
    uint64_t fallback_nodes;
    uint64_t refined_nodes;
    uint64_t occluded_nodes;
    uint64_t boundary_nodes;
    float reconstructed_fraction;
    float refined_fraction;
    float occluded_fraction;
    int32_t has_median_depth;
    double median_depth;
    int32_t has_median_score_improvement;
    float median_score_improvement;
} chd_depth_report;

/* NodeWarp codes (depth.rs:256-271). */
#define CHD_WARP_UNDEFINED 0
#define CHD_WARP_GLOBAL 1
#define CHD_WARP_REFINED 2
#define CHD_WARP_OCCLUDED 3
#define CHD_WARP_BOUNDARY 4

/* Per-target result: the refined warp that replaces alignments[i].warp
 * (depth.rs:526-546). Buffers are caller-allocated on the fine grid. */
typedef struct chd_target_result {
    int32_t replaced;   /* out: 1 when the warp and report were produced */
    float *points;      /* fine columns*rows*2 */
    float *confidence;  /* fine columns*rows */
    uint8_t *decisions; /* optional: NodeWarp code per fine node */
    chd_depth_report report;
} chd_target_result;

/* DepthProvenance (depth.rs:115-121). */
#define CHD_PROVENANCE_UNSUPPORTED 0
#define CHD_PROVENANCE_GLOBAL 1
#define CHD_PROVENANCE_MEASURED 2
#define CHD_PROVENANCE_REGULARIZED 3

/* DenseDepthMap (depth.rs:133-141) with caller-allocated node arrays. */
typedef struct chd_depth_map {
    uint32_t columns;
    uint32_t rows;
    uint32_t step;
    double near_depth;
    double far_depth;
    double *depth;        /* NaN = None */
    float *confidence;
    uint8_t *provenance;  /* CHD_PROVENANCE_* */
} chd_depth_map;

/* Option<NodeDepth> (depth.rs:248-254); valid = 0 is None. */
typedef struct chd_node_depth {
    double depth;
    float confidence;
    float improvement;
    uint8_t valid;
    uint8_t regularized;
    uint8_t pad[6];
} chd_node_depth;

/* Grid shapes of one call (depth.rs:326-328, 726-728, 1196-1202). */
typedef struct chd_grid_sizes {
    uint32_t coarse_step;
    uint32_t coarse_columns;
    uint32_t coarse_rows;
    uint32_t labels;
    uint32_t fine_step;
    uint32_t fine_columns;
    uint32_t fine_rows;
} chd_grid_sizes;

/* ---------------------------------------------------------------- host */

const char *chd_last_error(void);
int chd_abi_version(void);

/* DepthOptions::default (depth.rs:69-91). */
void chd_default_options(chd_options *out);

/* valid_options (depth.rs:1879-1891): 1 when valid. */
int chd_options_valid(const chd_options *options);

/* Grid shapes for a reference raster of width x height. */
int chd_grid_sizes_for(const chd_options *options, uint32_t width, uint32_t height,
                       chd_grid_sizes *out);

/* ResolvedCamera::new for a canonical pose and a single intrinsics bundle
 * (geometry.rs:64-133, calibration.rs:186-198): k_inverse = inverse(k),
 * rotation_cw = transpose(rotation_wc), center = -(rotation_cw t).
 * distortion may be NULL. */
int chd_camera_canonical(const double k[9], const double rotation_wc[9],
                         const double translation_wc[3], uint32_t width,
                         uint32_t height, const chd_distortion *distortion,
                         chd_camera *out);

/* ------------------------------------------------------------- device */

typedef struct chd_context chd_context;
typedef struct chd_session chd_session;

/* One context per thread of use; it owns a HIP stream on `device`.
 *
 * Lifetimes: sessions, batches and proposals made on a context may be
 * destroyed in any order, before or after the context. chd_context_destroy
 * with objects still alive waits for the stream and marks the context
 * released (no new object can be made on it); its memory is freed with the
 * last of those objects. Destroying a session that a batch holds leaves
 * the batch valid for chd_batch_destroy only: chd_batch_run then returns
 * CHD_ERR_STATE. Every handle is destroyed at most once. */
int chd_context_create(int device, chd_context **out);
void chd_context_destroy(chd_context *context);
int chd_context_synchronize(chd_context *context);

/* Drop-in for refine_multiview_depth (depth.rs:312-580). Inputs are host
 * buffers; outputs are written to caller-allocated host buffers sized by
 * chd_grid_sizes_for. `targets` has `count` entries (entry i belongs to
 * view i). Returns CHD_NONE where the reference returns None. */
int chd_refine_multiview_depth(chd_context *context, const chd_view *views,
                               uint32_t count, uint32_t reference_index,
                               const chd_options *options, chd_depth_map *map,
                               chd_target_result *targets);

/* Device-resident form of the same call: create uploads every input to HBM
 * once; run enqueues the whole path on the context's stream (asynchronous);
 * read_* copy results back after a synchronize. */
int chd_session_create(chd_context *context, const chd_view *views, uint32_t count,
                       uint32_t reference_index, const chd_options *options,
                       chd_session **out);
void chd_session_destroy(chd_session *session);

#define CHD_STAGE_COST 1u     /* build_cost_volume, depth.rs:1187-1277 */
#define CHD_STAGE_SGM 2u      /* semi_global_costs + select_depths, :1279-1467 */
#define CHD_STAGE_DIRECT 4u   /* measure_direct_depths + filters, :715-1060 */
#define CHD_STAGE_TARGETS 8u  /* per-target decisions, :420-547 */
#define CHD_STAGE_ALL 15u
/* Run the proposal attached by chd_session_attach_proposal first (network on
 * the session's planes, then its band windows); no counterpart in the
 * reference. Not part of CHD_STAGE_ALL. */
#define CHD_STAGE_PROPOSE 16u
/* Keep the aggregated SGM volume (debug/parity only; costs HBM traffic). */
#define CHD_STAGE_KEEP_AGGREGATED 256u
/* Keep every view's own ZNCC slice of the cost volume (the pairwise scores
 * before aggregate, depth.rs:1761-1770): views x nodes x labels f32. */
#define CHD_STAGE_KEEP_VIEW_SCORES 512u

int chd_session_run(chd_session *session, uint32_t stages);
int chd_session_grid(const chd_session *session, chd_grid_sizes *out);
int chd_session_read_results(chd_session *session, chd_depth_map *map,
                             chd_target_result *targets);
/* Cost volume, label-contiguous [node][label] (depth.rs:1270-1276). NaN
 * scores are the reference's f32::NAN for unsupported labels. NULL outputs
 * are skipped (the reads below do the same). */
int chd_session_read_cost_volume(chd_session *session, float *scores, float *costs,
                                 float *guidance, uint8_t *tested);
int chd_session_read_aggregated(chd_session *session, float *aggregated);
/* select_depths output (depth.rs:1385-1467) plus the WTA label index (0 when
 * the node was not tested). */
int chd_session_read_coarse(chd_session *session, chd_node_depth *field,
                            uint8_t *fillable, uint16_t *best_label);
/* Fine-grid field after measure_direct_depths (stage 0), after
 * reject_isolated_direct_depths (1), after reject_small_direct_components (2)
 * and after fit_local_depth_planes (3); guidance and tested of the fine grid. */
int chd_session_read_direct(chd_session *session, uint32_t stage, chd_node_depth *field,
                            float *guidance, uint8_t *tested);

/* Work the last run of each stage performed, counted by the kernels:
 * ZNCC evaluations that ran the full tap loop (after the parallax gate and
 * the bounds test), and nodes sent to the direct pass's wide fallback. */
typedef struct chd_work {
    uint64_t cost_zncc;   /* (node, label, view), build_cost_volume */
    uint64_t direct_zncc; /* (node, candidate, view), measure_direct_depths */
    uint64_t wide_zncc;   /* (node, candidate, view), its r = 8 fallback */
    uint64_t target_zncc; /* per-target refinement, all targets */
    uint64_t wide_nodes;
} chd_work;
int chd_session_work(chd_session *session, chd_work *out);

typedef struct chd_kernel_time {
    char name[48];
    uint32_t launches;
    double total_ms;
} chd_kernel_time;
/* Enable per-kernel HIP-event timing on the session's stream. */
int chd_session_enable_timing(chd_session *session, int enable);
int chd_session_kernel_times(chd_session *session, chd_kernel_time *out,
                             uint32_t capacity, uint32_t *count);
int chd_session_reset_kernel_times(chd_session *session);

/* Per-view ZNCC slices kept by a run with CHD_STAGE_KEEP_VIEW_SCORES:
 * [view slot][node][label], view slots in view-index order over the
 * accepted, calibrated targets (NaN = None). */
int chd_session_read_view_scores(chd_session *session, float *out);

/* Launch shape of the session's kernels. It changes the schedule only:
 * every shape computes the same results bit for bit. Zero fields take the
 * defaults. No counterpart in the reference (CPU threads). */
typedef struct chd_tuning {
    uint32_t cost_nodes_per_block; /* coarse nodes per cost-volume work item */
    uint32_t cost_label_slab;      /* labels per cost-volume work item (band sweeps size
                                      their passes from the LDS instead) */
    uint32_t sgm_paths_per_block;  /* SGM path sweeps (wavefronts) per block: 1, 2 or 4 */
    uint32_t cost_tile_order;      /* 0: work items in launch order; 1: each of the 8 XCDs
                                      (own L2) takes a contiguous run of node tiles; 2 (the
                                      persistent queue only): a static stride instead of
                                      the queue counters */
} chd_tuning;
int chd_session_set_tuning(chd_session *session, const chd_tuning *tuning);
int chd_session_get_tuning(const chd_session *session, chd_tuning *out);
/* Work-groups per CU that the runtime fits (registers, static and dynamic
 * LDS) for the session's ZNCC kernels at the shapes its runs launch:
 * [0] the cost volume (the band sweep while a proposal's windows are set),
 * [1] the direct pass, [2] the per-target stage. The tap loops are tuned for
 * four. No counterpart in the reference (it runs on the CPU); for tests and
 * launch tuning. */
int chd_session_occupancy(chd_session *session, uint32_t blocks_per_cu[3]);

/* Times candidate shapes on the session's own data (it runs the cost and
 * SGM stages several times) and sets the fastest. With cache_path, results
 * are read from and appended to that text file, keyed by the shape of the
 * work (views, taps, labels, grid, plane sizes), so a later call on the same
 * shape only reads it. */
int chd_autotune(chd_session *session, const char *cache_path, chd_tuning *out);

/* Batched execution of several sessions (captures) on one context, the
 * whole sequence for a given `stages` mask captured once into a hipGraph and
 * replayed by every later run. By default the captures are pipelined over
 * two streams of the same priority: capture k's cost volume on the
 * context's stream, and once it is done, capture k's SGM, select, direct
 * pass and decisions on a second stream, beside capture k + 1's cost volume.
 * CHD_BATCH_ONE_SWEEP runs every capture's SGM path sweep in one launch after
 * all cost volumes instead (one stream). With CHD_BATCH_PERSISTENT the cost
 * volumes of all captures run as one persistent launch fed by a (capture,
 * node tile, label slab) work queue, followed by the one sweep. Results are
 * read per session as after chd_session_run, and are the same bit for bit
 * in every mode. Sessions must share the label count and outlive the batch;
 * no counterpart in the reference (one capture at a time). */
typedef struct chd_batch chd_batch;
#define CHD_BATCH_PERSISTENT 1u
#define CHD_BATCH_ONE_SWEEP 2u
int chd_batch_create(chd_context *context, chd_session *const *sessions, uint32_t count,
                     uint32_t flags, chd_batch **out);
int chd_batch_run(chd_batch *batch, uint32_t stages);
void chd_batch_destroy(chd_batch *batch);

/* Learned band proposal (no counterpart in the reference, which sweeps
 * every hypothesis): a 7-layer 3x3 convolutional encoder-decoder with BF16
 * weights and activations and FP32 accumulation on the matrix cores. Its
 * input is width x height x 16 BF16 (channels 0..2: the reference and the
 * first two target planes as means over cells of (grid step / 2)^2 plane
 * pixels, through the input transform below; the rest zero); per pixel it
 * predicts a centre label, a log band half-width and a confidence. Trained
 * with Adam on full-sweep WTA label maps (0 = none). */
typedef struct chd_proposal chd_proposal;
int chd_proposal_create(chd_context *context, uint32_t width, uint32_t height,
                        uint32_t max_batch, uint64_t seed, chd_proposal **out);
void chd_proposal_destroy(chd_proposal *proposal);
/* All parameters as f32 (per layer: weights [9 * cin_p][cout_p], then
 * biases [cout_p]): with get and set NULL, *count receives the count; else
 * *count must equal it and get receives or set replaces them. */
int chd_proposal_parameters(chd_proposal *proposal, float *get, const float *set,
                            size_t *count);
/* Adam state for checkpoints: 2 x the parameter count f32 (the first
 * moments in chd_proposal_parameters order, then the second moments) and
 * the number of steps taken (bias corrections). With get and set NULL,
 * *count and *step receive the sizes and the step; else *count must equal
 * it, get receives the state and *step, or set and *step replace them.
 * Parameters plus this state restore the optimiser exactly: the next Adam
 * step after a restore equals the one the saved network would take. */
int chd_proposal_optimizer_state(chd_proposal *proposal, float *get, const float *set,
                                 size_t *count, uint64_t *step);
/* The input transform: each input value is (cell mean of a plane - offset) *
 * scale in f32, rounded to BF16 (default offset 0, scale 1/8). Inputs given
 * to set_dataset / predict / gradients must have been made with the same
 * transform; set_proposal applies it on the device. */
int chd_proposal_set_input_transform(chd_proposal *proposal, float offset, float scale);
/* Hypothesis labels L of the sessions the network serves (default 129 = 128
 * planes + the global warp): its centre output maps to (L - 1) / 2 +
 * (L - 1) / 8 * o0 labels. */
int chd_proposal_set_label_count(chd_proposal *proposal, uint32_t label_count);
int chd_proposal_get_input_transform(const chd_proposal *proposal, float *offset, float *scale);
/* Training set (host arrays): inputs count x height x width x 16 BF16 bits,
 * labels count x height x width. */
int chd_proposal_set_dataset(chd_proposal *proposal, const uint16_t *inputs,
                             const uint16_t *labels, uint32_t count);
/* Epochs of shuffled mini-batches with Adam (beta 0.9, 0.999; eps 1e-8),
 * the step size decaying from learning_rate to 5% of it on a cosine over
 * the run; epoch_loss (optional, `epochs` entries) receives the mean loss per
 * labelled pixel of each epoch. */
int chd_proposal_train(chd_proposal *proposal, uint32_t epochs, uint32_t batch,
                       float learning_rate, uint64_t shuffle_seed, float *epoch_loss);
/* Epochs first_epoch .. first_epoch + epochs - 1 of a run of total_epochs
 * (the shuffles and the cosine of the whole run): successive calls covering
 * 0 .. total_epochs - 1 train exactly as one chd_proposal_train call, so a
 * caller can report progress between them. epoch_loss gets `epochs`
 * entries. */
int chd_proposal_train_epochs(chd_proposal *proposal, uint32_t first_epoch, uint32_t epochs,
                              uint32_t total_epochs, uint32_t batch, float learning_rate,
                              uint64_t shuffle_seed, float *epoch_loss);
/* Per pixel of `count` inputs: centre label, log half-width, confidence. */
int chd_proposal_predict(chd_proposal *proposal, const uint16_t *inputs, uint32_t count,
                         float *outputs);
/* Parameter gradients of the summed loss (divided by the batch's pixel
 * count) for one batch, in chd_proposal_parameters order; *loss receives the
 * summed loss. For checks of the backward pass. */
int chd_proposal_gradients(chd_proposal *proposal, const uint16_t *inputs,
                           const uint16_t *labels, uint32_t count, float *gradients,
                           float *loss);
/* One Adam update from the given gradients (for checks of the optimiser). */
int chd_proposal_adam_step(chd_proposal *proposal, const float *gradients,
                           float learning_rate);
/* Runs the network on the session's own planes (on the device) and makes
 * its next cost volumes sweep label 0 and centre +- w only, w = ceil(4 b)
 * for the predicted band scale b, clamped to [half_width, max_half_width]
 * (max_half_width <= half_width: always half_width); tiles of tile x tile
 * coarse nodes whose mean confidence is below `threshold` (fallback tiles)
 * keep the full sweep, or with fallback_half_width > 0 sweep centre +-
 * fallback_half_width. The network grid must be the reference plane at half
 * size.
 * full_tiles / tiles (optional) receive the fallback count and the tile
 * count (full_tiles synchronises). clear restores the full sweep. The
 * network's grid has one pixel per coarse node: the reference plane in cells
 * of (grid step / 2) x (grid step / 2) plane pixels (the grid step must be
 * even), and its label count must equal the session's. */
int chd_session_set_proposal(chd_session *session, chd_proposal *proposal, float threshold,
                             uint32_t half_width, uint32_t max_half_width,
                             uint32_t fallback_half_width, uint32_t tile, uint32_t *full_tiles,
                             uint32_t *tiles);
/* chd_session_set_proposal now (same arguments and results), and keeps the
 * proposal and setting with the session: every later chd_session_run or
 * chd_batch_run with CHD_STAGE_PROPOSE runs the network on the session's
 * current planes and sets the windows before the cost volume, inside the
 * batch's graph, so a batch needs no host call per capture for its bands.
 * chd_session_clear_proposal detaches; destroying the proposal detaches it
 * from every session (a PROPOSE run then returns CHD_ERR_STATE). */
int chd_session_attach_proposal(chd_session *session, chd_proposal *proposal, float threshold,
                                uint32_t half_width, uint32_t max_half_width,
                                uint32_t fallback_half_width, uint32_t tile,
                                uint32_t *full_tiles, uint32_t *tiles);
int chd_session_clear_proposal(chd_session *session);
/* The label window of every coarse node set by the last
 * chd_session_set_proposal: labels 0 and lo..hi are swept (lo = 1 and
 * hi = L - 1 on fallback tiles). CHD_ERR_STATE without a proposal. */
int chd_session_read_band(chd_session *session, uint16_t *lo, uint16_t *hi);

/* Stage entry points on host arrays, for parity checks of one stage. */

/* semi_global_costs (depth.rs:1279-1311). */
int chd_semi_global_costs(chd_context *context, const float *costs,
                          const float *guidance, uint32_t columns, uint32_t rows,
                          uint32_t labels, float *aggregated);

/* select_depths (depth.rs:1385-1467). label_depths[0] is ignored (label 0
 * is the global warp); scores/regularised are [node][label]. */
int chd_select_depths(chd_context *context, const double *label_depths,
                      uint32_t labels, const float *scores, const uint8_t *tested,
                      const float *regularised, uint32_t nodes,
                      const chd_options *options, chd_node_depth *field,
                      uint8_t *fillable, uint16_t *best_label);

/* reject_isolated_direct_depths (depth.rs:881-924),
 * reject_small_direct_components (:933-989), fit_local_depth_planes
 * (:995-1060), each in place on a columns x rows field. */
int chd_reject_isolated_direct_depths(chd_context *context, chd_node_depth *field,
                                      const float *guidance, uint32_t columns,
                                      uint32_t rows, double inverse_tolerance);
int chd_reject_small_direct_components(chd_context *context, chd_node_depth *field,
                                       const float *guidance, uint32_t columns,
                                       uint32_t rows, double inverse_tolerance,
                                       uint32_t minimum_nodes);
/* Largest grid of the allocation-free entry point below (2^21 nodes; a
 * 4160x3120 raster's fine grid has 813,021). */
#define CHD_NOALLOC_MAX_NODES 2097152u

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): reject_small_direct_components (depth.rs:933-989) on
 * device buffers, with the union-find parents and sizes in a static device
 * scratch of CHD_NOALLOC_MAX_NODES nodes (16 MiB per GPU, held until the
 * process ends). field, guidance and out are device buffers of
 * columns x rows (out may equal field); the work is enqueued on the
 * context's stream without synchronising, and nothing is allocated.
 * Grids above CHD_NOALLOC_MAX_NODES return CHD_ERR_LIMIT. Every call on one
 * GPU shares its scratch: calls that overlap on one GPU (other streams or
 * contexts) race. */
int chd_reject_small_direct_components_noalloc(chd_context *context,
                                               const chd_node_depth *field,
                                               const float *guidance, uint32_t columns,
                                               uint32_t rows, double inverse_tolerance,
                                               uint32_t minimum_nodes, chd_node_depth *out);
int chd_fit_local_depth_planes(chd_context *context, chd_node_depth *field,
                               const float *guidance, uint32_t columns, uint32_t rows,
                               double inverse_tolerance);

/* NodeWarp decision record used by the consensus and boundary stages. */
typedef struct chd_node_warp {
    uint32_t code; /* CHD_WARP_* */
    float global[2];
    float point[2];
    float confidence;
    uint32_t measured;
} chd_node_warp;

/* enforce_warp_consensus (depth.rs:582-618) and suppress_warp_boundaries
 * (:620-672), in place. */
int chd_enforce_warp_consensus(chd_context *context, chd_node_warp *decisions,
                               uint32_t columns, uint32_t rows);
int chd_suppress_warp_boundaries(chd_context *context, chd_node_warp *decisions,
                                 const float *guidance, uint32_t columns, uint32_t rows);

/* Input preparation: Mosaic::luminance_half (crates/chiaro-fusion/src/
 * image.rs:401-426, called per module at pipeline.rs:1106-1109) of a u16 CFA
 * mosaic in calibration-raster order (width x height, row-major, the
 * pipeline's Q6 samples), with a unit flat field (no vignetting mesh). With
 * reserve_headroom != 0 the samples, black_q6 and white_q6 are first halved
 * as reserve_highlight_headroom does (image.rs:69-75). Writes the
 * (width / 2) x (height / 2) f32 plane that chd_view.luminance takes.
 * chd_luminance_half copies host buffers in and out; the _device form takes
 * device buffers (samples 4-byte aligned) and only enqueues on the context's
 * stream. */
int chd_luminance_half(chd_context *context, const uint16_t *samples, uint32_t width,
                       uint32_t height, float black_q6, float white_q6, int reserve_headroom,
                       float *plane);
int chd_luminance_half_device(chd_context *context, const uint16_t *samples, uint32_t width,
                              uint32_t height, float black_q6, float white_q6,
                              int reserve_headroom, float *plane);

/* ---------------------------------------------------------- alignment */

/* AlignOptions (align.rs:211-244) without the nested DepthOptions. */
typedef struct chd_align_options {
    uint32_t grid_step;     /* output warp grid, reference raster pixels (32) */
    uint32_t patch;         /* correlation window, plane pixels (32) */
    uint32_t coarse_radius; /* search radius at the coarsest level (12) */
    float min_score;        /* NCC peak for a patch to vote (0.5) */
    float inlier_px;        /* RANSAC inlier threshold, reference pixels (3) */
    float min_inlier_ratio; /* finest-level consensus to accept (0.45) */
    int32_t refine;         /* 0: keep the factory model (align.rs:317-324) */
} chd_align_options;

/* LevelReport (align.rs:202-208). */
typedef struct chd_level_report {
    uint32_t scale;
    uint32_t patches;
    uint32_t inliers;
    float median_residual_px; /* NaN when the level's fit failed */
} chd_level_report;

#define CHD_ALIGN_FROM_REFERENCE 0
#define CHD_ALIGN_FROM_CALIBRATION 1

#define CHD_ALIGN_STATUS_REFERENCE 0
#define CHD_ALIGN_STATUS_REJECTED 1
#define CHD_ALIGN_STATUS_FACTORY_MODEL 2
#define CHD_ALIGN_STATUS_REFINED 3
#define CHD_ALIGN_STATUS_WEAK 4
#define CHD_ALIGN_STATUS_NO_SUPPORT 5
#define CHD_MAX_ALIGN_LEVELS 16

/* The numeric fields of AlignmentReport (align.rs:160-200) that
 * align_module fills, plus ModuleAlignment::gain (align.rs:148-157); the
 * status string's case is a CHD_ALIGN_STATUS_* code (align.rs:488-501). */
typedef struct chd_alignment_report {
    uint32_t initialised_from; /* CHD_ALIGN_FROM_* */
    float coverage;
    float correction_median_px[2];
    uint32_t level_count;
    chd_level_report levels[CHD_MAX_ALIGN_LEVELS];
    uint32_t inliers;
    uint32_t patches;
    float residual_median_px;
    float residual_p90_px;
    float inlier_ratio;
    int32_t accepted;
    uint32_t status;
    float gain;
} chd_alignment_report;

/* AlignOptions::default (align.rs:232-244). */
void chd_default_align_options(chd_align_options *out);

/* align_module (align.rs:270-519) for calibrated modules: the factory
 * model (an 8 px infinity-map grid, align.rs:317-324) refined
 * coarse-to-fine by a reference-space homography fitted by RANSAC to
 * match_patch correlations (image.rs:531-640) between the reference
 * pyramid and the target rendered through the current mapping; writes the
 * warp (grid_step spacing over the reference raster; caller-allocated
 * points and confidence of ceil(w / step) + 1 by ceil(h / step) + 1 nodes)
 * and the report. reference == target (the same view) is the reference's
 * own identity alignment. Both views need cameras (the uncalibrated
 * nominal-focal start is not supported: CHD_ERR_ARGUMENT). The luminance
 * planes may be in host or device memory. The pyramids, rendering,
 * correlations, RANSAC consensus counts and photometric samples run on the
 * device; the RANSAC samples, least-squares fits and medians on the host. */
int chd_align_module(chd_context *context, const chd_view *reference, const chd_view *target,
                     const chd_align_options *options, chd_warp *warp,
                     chd_alignment_report *report);

/* match_patch (image.rs:531-640) on the device for a batch of windows of
 * one reference / target plane pair (same sizes, host arrays): window k at
 * (rx[k], ry[k]) of `size` with search radius `radius`; valid[k] = 0 is
 * None, else shift[2k..2k+1] and score[k]. A stage entry for parity
 * checks. */
int chd_match_patches(chd_context *context, const float *reference, const float *target,
                      uint32_t width, uint32_t height, const uint32_t *rx, const uint32_t *ry,
                      uint32_t count, uint32_t size, uint32_t radius, float *shift, float *score,
                      uint8_t *valid);

/* ResolutionAlignmentReport (crates/chiaro-fusion/src/resolution.rs:87-98). */
typedef struct chd_resolution_report {
    float verified_fraction;    /* nodes with a coherent correlation peak */
    float supported_fraction;   /* nodes with confidence > 0 after completion */
    float median_correction_px; /* median |correction| of verified nodes, raster px */
    float mean_confidence;      /* mean confidence of verified nodes */
} chd_resolution_report;

/* refine_resolution_warp (crates/chiaro-fusion/src/resolution.rs:232-432),
 * called per non-reference module at fusion.rs:384 and pipeline.rs:1185:
 * local 12x12 NCC on a 32 px grid of the reference raster (width x height)
 * between the reference plane and the target plane rendered through `base`
 * at the reference's bandwidth, neighbourhood-median checks and two
 * completion passes. Planes are row-major f32 (pitch = width) in host or
 * device memory; `base` is a host warp. `out` gets step 32, columns = ceil(width / 32) + 1,
 * rows = ceil(height / 32) + 1; its points (2 * columns * rows f32) and
 * confidence (columns * rows f32) arrays are the caller's. */
int chd_refine_resolution_warp(chd_context *context, const float *reference,
                               uint32_t reference_width, uint32_t reference_height,
                               const float *target, uint32_t target_width,
                               uint32_t target_height, const chd_warp *base, uint32_t width,
                               uint32_t height, chd_warp *out, chd_resolution_report *report);

/* The synthesis consumer's depth and warp stage. synthesize
 * (crates/chiaro-fusion/src/synth.rs:591-725) walks the output canvas and,
 * per pixel, reads the depth map (DenseDepthMap::sample_nearest,
 * depth.rs:152-159) and every source's warp before it samples the mosaics;
 * these entries compute that part (the demosaic, colour and blending that
 * follow are outside this library). */

/* A synthesis source (SynthSource, synth.rs:346-370): the fields this stage
 * reads. Sources are the `usable` list of synth.rs:601-605, in order. */
typedef struct chd_synth_source {
    chd_warp warp;          /* alignment.warp (host arrays) */
    uint32_t mosaic_width;  /* mosaic raster */
    uint32_t mosaic_height;
    float magnification;
    float confidence;
    int32_t reference;
    int32_t has_focus_distance;
    double focus_distance;  /* read when has_focus_distance */
} chd_synth_source;

/* CropWindow (synth.rs:106-112), the canvas scale and
 * SynthOptions::feather_px. */
typedef struct chd_canvas {
    float crop_x;
    float crop_y;
    float crop_width;
    float crop_height;
    float scale;
    float feather_px;
} chd_canvas;

/* The canvas size synthesize renders (synth.rs:619-620):
 * round(crop * scale), at least 1. CHD_ERR_ARGUMENT on the reference's
 * "invalid canvas geometry" (NaN or non-positive scale, crop below 1). */
int chd_canvas_size(const chd_canvas *canvas, uint32_t *width, uint32_t *height);

/* For every canvas pixel (u, v) and source s, in [s][v][u] order
 * (width x height from chd_canvas_size): the target point q =
 * warp.map(rx, ry) (points, 2 f32), the border feather, the warp's local
 * confidence, base_weight = feather * local * m * m * confidence and the
 * close-side focus weight from the depth map's nearest node
 * (focus_consistency_weight, synth.rs:1469-1491), for the sources that
 * contribute there (synth.rs:640-725: the row band of 64 canvas rows may
 * reach the module, q is defined and inside the mosaic, feather > 0 and
 * local confidence > 0). A source that does not contribute gets NaN points
 * and zeros. `map` may be NULL (no depth map: every focus weight is 1).
 * Host arrays; an output pointer may be NULL to skip it. */
int chd_synthesis_geometry(chd_context *context, const chd_depth_map *map,
                           const chd_synth_source *sources, uint32_t count,
                           const chd_canvas *canvas, float *points, float *feather,
                           float *local_confidence, float *base_weight, float *focus_weight);

/* Synthetic captures rendered on the GPU. No counterpart in the reference
 * (its only fakes are gradient mock LRIs, crates/chiaro/src/mock.rs:62):
 * a seeded scene of fronto-parallel planes over a Voronoi partition of the
 * reference raster, textured by 4 octaves of value noise, seen by pinhole
 * modules (principal point at the raster centre) and recorded as 12-bit
 * Bayer codes with black level, Poisson shot noise and Gaussian read noise;
 * the codes are shifted left by `shift` bits (the pipeline's Q6 step,
 * pipeline.rs:1026-1030) and converted by luminance_half (image.rs:401-426)
 * into the planes a session takes. */
typedef struct chd_synthetic_scene {
    uint64_t seed;
    uint32_t plane_count;     /* 1..64 */
    double inverse_depth_min; /* 1/mm */
    double inverse_depth_max;
    uint32_t texture_cell_px; /* lattice spacing of the coarsest octave */
} chd_synthetic_scene;

typedef struct chd_synthetic_module {
    double focal_px;
    uint32_t width; /* raster size (even) */
    uint32_t height;
    double center[3];      /* optical centre in the reference frame, mm */
    double rotation_wc[9]; /* world-to-camera rotation, row-major */
} chd_synthetic_module;

typedef struct chd_bayer {
    float black_level; /* DN */
    float white_level;
    float exposure;    /* fraction of (white - black) reached by radiance 1 */
    float electrons_per_dn;
    float read_noise_dn;
} chd_bayer;

/* Device buffers of one rig's frames: a u16 mosaic and an f32 plane per
 * module. create allocates them; render draws a scene into them on the
 * context's stream (asynchronous; reuse for many scenes of the same rig);
 * plane gives a module's device plane (plane_width x plane_height, row
 * pitch plane_width) for chd_session_create_device / chd_session_set_planes;
 * read copies a module's mosaic and/or plane to the host. */
typedef struct chd_frames chd_frames;
int chd_frames_create(chd_context *context, const chd_synthetic_module *modules, uint32_t count,
                      uint32_t reference_index, const chd_bayer *bayer, uint32_t shift,
                      float black_q6, float white_q6, chd_frames **out);
int chd_frames_render(chd_frames *frames, const chd_synthetic_scene *scene);
int chd_frames_plane(const chd_frames *frames, uint32_t index, const float **plane,
                     uint32_t *plane_width, uint32_t *plane_height);
int chd_frames_read(chd_frames *frames, uint32_t index, uint16_t *mosaic, float *plane);
void chd_frames_destroy(chd_frames *frames);

/* chd_session_create with every chd_view.luminance a device pointer (row
 * pitch plane_width), e.g. from chd_frames_plane: the planes never leave
 * HBM. */
int chd_session_create_device(chd_context *context, const chd_view *views, uint32_t count,
                              uint32_t reference_index, const chd_options *options,
                              chd_session **out);
/* Replaces the luminance of every view of a session (same sizes; planes[i]
 * on the device when `device`, else on the host) for the next run: a new
 * capture of the same rig. Earlier results are invalidated. */
int chd_session_set_planes(chd_session *session, const float *const *planes, uint32_t count,
                           int device);
/* The learned proposal's input for the session's current planes (what
 * chd_session_set_proposal feeds the network): proposal width x height x 16
 * BF16 bits into `out` (host). */
int chd_session_network_input(chd_session *session, chd_proposal *proposal, uint16_t *out);

/* Device f32 exp / f64 exp used by the kernels, evaluated on the GPU for
 * n inputs (their parity with the platform libm is checked by the tests). */
int chd_device_expf(chd_context *context, const float *x, float *y, size_t n);
int chd_device_exp(chd_context *context, const double *x, double *y, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* CHIARO_DEPTH_H */
