// depth_internal.hpp - device data layout and kernel launchers of the depth
// path, shared by the kernels and the C ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/chiaro_depth.h"

namespace chd {

// One module as the kernels see it. Pointers are device pointers.
struct ViewDev {
    const float *plane;    // (plane_height + 1) rows of plane_pitch f32, the last
                           // column and row replicated (see DevPlane)
    const float *columns;  // the same padded plane column by column (pitch
                           // plane_height + 1), or nullptr (see DevPlane)
    uint32_t plane_width, plane_height, plane_pitch;
    // target views: the 2 x 2 sample neighbourhood of every stored element of
    // the plane the ZNCC tap loops read (the column copy where there is one,
    // else the rows): element (line, along) holds samples (line, along),
    // (line, along + 1), (line + 1, along), (line + 1, along + 1), so that a
    // bilinear sample is one 16-byte load; null for other views
    const float4 *quad;
    const float2 *warp;   // global warp points (x, y), NaN = undefined
    uint32_t warp_step, warp_columns, warp_rows;
    const chd_camera *camera;  // nullptr = no calibration
};

// ZNCC patch context shared by the cost, direct, wide and target kernels.
struct PatchSpec {
    const ViewDev *views;      // all views
    const uint32_t *vlist;     // view indices scored by this launch
    uint32_t view_count;       // entries of vlist
    uint32_t reference;        // reference view index
    uint32_t radius;           // patch radius (taps = (2r+1)^2)
};

// Grid of nodes at raster positions (column * step, row * step).
struct GridSpec {
    uint32_t columns, rows, step;
};

// Aggregated scores per fine node in the direct pass: the global-warp
// baseline and up to 32 candidate depths (depth.rs:1078-1107).
constexpr uint32_t kDirectSlots = 33;

// Entry of the wide-fallback list of the direct pass (depth.rs:803-846).
struct WideEntry {
    uint32_t node;
    uint32_t count;
    uint32_t seeded;
    uint32_t pad;
    double depth[7];
};

// Kernel launchers (the ZNCC stages in depth_zncc.hip, the rest in
// depth_kernels.hip). All enqueue on `stream` and return the hipError_t of
// the launch.
// One cost-volume job: a capture's coarse grid, labels and outputs, and its
// launch shape (cost_job_shape fills T, nb, slab, tiles, slabs). Work items
// are (tile, slab) pairs; every shape computes the same volume.
struct CostJob {
    PatchSpec S;
    GridSpec G;
    const double *labels;
    uint32_t L;              // labels
    uint32_t T;              // taps
    uint32_t nb;             // nodes per tile
    uint32_t slab;           // labels per slab
    uint32_t tiles, slabs;
    uint32_t minimum_support, best_view_count;
    uint32_t topk;           // views aggregated as a running top-K (set by cost_job_shape)
    uint32_t window_max;     // band sweep: labels of the widest window (0: unknown)
    float *scores, *costs, *guidance;
    uint8_t *tested;         // cleared before the launch
    float *view_scores;      // optional [view slot][node][label] per-view ZNCC, NaN = None
    unsigned long long *work;
    uint32_t xcd_order;      // 1: each XCD takes a contiguous run of work items
    // optional per-node label windows (band sweep): label 0 plus lo .. hi;
    // with them a work item is one tile (one slab)
    const uint16_t *window_lo, *window_hi;
};

constexpr uint32_t kXcds = 8;  // MI355X: workgroups go to the 8 XCDs round robin

// nodes_per_block / label_slab of 0 pick the defaults.
void cost_job_shape(CostJob &job, uint32_t nodes_per_block, uint32_t label_slab);
size_t cost_job_lds(const CostJob &job);

hipError_t launch_cost_volume(hipStream_t stream, const CostJob &job);

// Persistent launch over njobs device-resident jobs; first[] (device, njobs +
// 1 entries) is the prefix of tiles * slabs; lds must cover every job;
// counters are kXcds device words the launch clears first. With xcd_order
// each XCD drains its own contiguous run of items before taking others'.
// mode: cost_job_mode of every job when they agree, else 4 (mixed).
int cost_job_mode(const CostJob &job);
hipError_t launch_cost_volume_queue(hipStream_t stream, const CostJob *jobs,
                                    const uint32_t *first, uint32_t njobs, size_t lds,
                                    uint32_t blocks, uint32_t *counters, uint32_t xcd_order,
                                    int mode);

// paths_per_block: wavefronts (one path each) per block, 1, 2 or 4.
// With band windows (window_lo/hi non-null: label 0 and lo .. hi per node,
// +inf costs elsewhere) the sweep moves only the label chunks that meet each
// node's set; path values outside the sets are then left unwritten.
hipError_t launch_sgm(hipStream_t stream, const float *costs, const float *guidance,
                      uint32_t columns, uint32_t rows, uint32_t labels, float *paths,
                      uint32_t paths_per_block, const uint16_t *window_lo = nullptr,
                      const uint16_t *window_hi = nullptr);

// One capture's SGM inputs and path volumes for the batched sweep.
struct SgmJob {
    const float *costs, *guidance;
    uint32_t columns, rows;
    float *paths;
    const uint16_t *win_lo, *win_hi;  // band windows, or null
};
// Paths of one capture's sweep (2R + 2C + 4(C + R - 1)).
uint32_t sgm_waves(uint32_t columns, uint32_t rows);
// first[]: prefix of the jobs' path counts; same_grids: every job has the
// same columns and rows (their paths are then interleaved by length).
hipError_t launch_sgm_batch(hipStream_t stream, const SgmJob *jobs, const uint32_t *first,
                            uint32_t njobs, uint32_t total_paths, uint32_t labels,
                            uint32_t paths_per_block, bool same_grids, bool windows);

hipError_t launch_sgm_select(hipStream_t stream, const float *paths, const float *scores,
                             const uint8_t *tested, const double *labels, uint32_t columns,
                             uint32_t rows, uint32_t label_count, const chd_options &options,
                             float *aggregated, chd_node_depth *field, uint8_t *fillable,
                             uint16_t *best_label, const uint16_t *window_lo = nullptr,
                             const uint16_t *window_hi = nullptr);  // band windows or null

hipError_t launch_select_depths(hipStream_t stream, const float *regularised,
                                const float *scores, const uint8_t *tested,
                                const double *labels, uint32_t nodes, uint32_t label_count,
                                const chd_options &options, chd_node_depth *field,
                                uint8_t *fillable, uint16_t *best_label);

hipError_t launch_complete_depth_field(hipStream_t stream, const chd_node_depth *previous,
                                       chd_node_depth *field, const float *guidance,
                                       const uint8_t *fillable, uint32_t columns,
                                       uint32_t rows, uint32_t minimum_neighbour_support);

hipError_t launch_direct(hipStream_t stream, const PatchSpec &patch, const GridSpec &fine,
                         const GridSpec &coarse, const chd_node_depth *coarse_field,
                         double inverse_step, const chd_options &direct_options,
                         chd_node_depth *field, float *guidance, uint8_t *tested,
                         WideEntry *wide, uint32_t *wide_count, float *agg,
                         unsigned long long *work);

hipError_t launch_direct_wide(hipStream_t stream, const PatchSpec &patch, const GridSpec &fine,
                              const chd_options &wide_options, const WideEntry *wide,
                              const uint32_t *wide_count, uint32_t wide_capacity,
                              chd_node_depth *field, unsigned long long *work);

hipError_t launch_reject_isolated(hipStream_t stream, const chd_node_depth *source,
                                  chd_node_depth *field, const float *guidance,
                                  uint32_t columns, uint32_t rows, double inverse_tolerance);

hipError_t launch_reject_components(hipStream_t stream, const chd_node_depth *source,
                                    chd_node_depth *field, const float *guidance,
                                    uint32_t columns, uint32_t rows, double inverse_tolerance,
                                    uint32_t minimum_nodes, uint32_t *parent, uint32_t *sizes);

hipError_t launch_fit_planes(hipStream_t stream, const chd_node_depth *source,
                             chd_node_depth *field, const float *guidance, uint32_t columns,
                             uint32_t rows, double inverse_tolerance);

// Blocks per CU the runtime fits (registers, static and dynamic LDS) for
// the cost-volume kernel of `cost` (full or band), k_direct_score and
// k_targets at the launch shapes their launches use.
hipError_t zncc_occupancy(const CostJob &cost, const PatchSpec &direct, const PatchSpec &target,
                          uint32_t fine_step, uint32_t best_view_count, int blocks[3]);

// Every view of patch.vlist is a target; decisions is [view slot][node].
hipError_t launch_targets(hipStream_t stream, const PatchSpec &patch, const GridSpec &fine,
                          const chd_node_depth *field, double inverse_step,
                          const chd_options &options, chd_node_warp *decisions,
                          unsigned long long *work);

hipError_t launch_consensus(hipStream_t stream, const chd_node_warp *source,
                            chd_node_warp *decisions, uint32_t columns, uint32_t rows);

hipError_t launch_boundaries(hipStream_t stream, const chd_node_warp *source,
                             chd_node_warp *decisions, const float *guidance, uint32_t columns,
                             uint32_t rows);

// counts[0..3] += refined, occluded, boundary (device counters)
hipError_t launch_target_outputs(hipStream_t stream, const chd_node_warp *decisions,
                                 uint32_t nodes, float *points, float *confidence,
                                 uint8_t *codes, unsigned long long *counts);

// counts[0..2] = tested, measured, regularized; keys for the medians
hipError_t launch_field_stats(hipStream_t stream, const chd_node_depth *field,
                              const uint8_t *tested, uint32_t nodes,
                              unsigned long long *counts, unsigned long long *depth_keys,
                              unsigned int *improvement_keys, uint32_t *key_counts);

hipError_t launch_final_map(hipStream_t stream, const chd_node_depth *field,
                            const uint8_t *tested, uint32_t nodes, double *depth,
                            float *confidence, uint8_t *provenance);

// Replicates the last column and row of a plane stored with pitch width + 1
// and height + 1 rows.
hipError_t launch_pad_plane(hipStream_t stream, float *plane, uint32_t width, uint32_t height);
// src (pitch width, device) into a padded plane dst (pitch width + 1,
// height + 1 rows): the replicated column and row written with it.
hipError_t launch_pitch_copy(hipStream_t stream, const float *src, uint32_t width, uint32_t height,
                             float *dst);

// Sums of a 2x2 cell of u16 samples index the luminance table.
constexpr uint32_t kLuminanceTableSize = 4u * 65535u + 1u;

// luminance_half of a u16 mosaic (width x height) into a (width / 2) x
// (height / 2) f32 plane through the table of kLuminanceTableSize entries.
hipError_t launch_luminance_half(hipStream_t stream, const uint16_t *samples, uint32_t width,
                                 uint32_t height, const float *table, int headroom,
                                 float *plane);

// Copies a padded row-major plane (width + 1 by height + 1) into column-major
// order (pitch height + 1).
hipError_t launch_transpose_plane(hipStream_t stream, const float *rows, float *columns,
                                  uint32_t width, uint32_t height);

// vertical[k] = 1 when, at the reference centre, the displacement between the
// near and far depth in view vlist[k] runs more vertically than horizontally.
// The quad array of a padded plane (`lines` stored lines of `pitch` f32).
hipError_t launch_quad_plane(hipStream_t stream, const float *plane, uint32_t pitch,
                             uint32_t lines, float4 *quad);
hipError_t launch_epipolar_axes(hipStream_t stream, const ViewDev *views, const uint32_t *vlist,
                                uint32_t count, uint32_t reference, double near_depth,
                                double far_depth, uint32_t *vertical);

hipError_t launch_fill_u32(hipStream_t stream, uint32_t *data, uint32_t value, size_t n);
hipError_t launch_fill_u64(hipStream_t stream, unsigned long long *data,
                           unsigned long long value, size_t n);

hipError_t launch_expf(hipStream_t stream, const float *x, float *y, size_t n);
hipError_t launch_exp(hipStream_t stream, const double *x, double *y, size_t n);

hipError_t launch_fill_u8(hipStream_t stream, uint8_t *data, uint8_t value, size_t n);

// ------------------------------------------------ alignment (align_kernels.hip)

// One pyramid level: dst (dw x dh, row pitch dpitch) from src (sw x sh,
// pitch spitch) by Plane::downsample.
hipError_t launch_downsample(hipStream_t stream, const float *src, uint32_t sw, uint32_t sh,
                             uint32_t spitch, float *dst, uint32_t dw, uint32_t dh, uint32_t dpitch);
// target.map_from(reference, grid point, 1e8) on a columns x rows grid of
// `step` (NaN where undefined); cameras in device memory.
hipError_t launch_initial_grid(hipStream_t stream, const chd_camera *reference,
                               const chd_camera *target, uint32_t columns, uint32_t rows,
                               uint32_t step, float2 *out);
// render_through of a width x height level into out (row pitch width).
hipError_t launch_render_through(hipStream_t stream, uint32_t width, uint32_t height,
                                 uint32_t reference_scale, const struct DevPlane &target,
                                 double target_scale, const struct DevWarp &initial,
                                 const double correction[9], float *out);
size_t match_patches_lds(uint32_t size, uint32_t radius);
// match_patch of `count` windows (device arrays rx, ry) between two planes of
// width x height (row pitch width). `coverage` adds a precondition on the
// target: kCoverWindowNan makes a window whose target pixels hold a NaN None
// (align.rs:550-561), kCoverSearchFinite one with a non-finite target pixel
// anywhere in the window grown by `radius` (resolution.rs:467-483).
enum { kCoverNone = 0, kCoverWindowNan = 1, kCoverSearchFinite = 2 };
hipError_t launch_match_patches(hipStream_t stream, const float *reference, const float *target,
                                uint32_t width, uint32_t height, const uint32_t *rx,
                                const uint32_t *ry, uint32_t count, uint32_t size, uint32_t radius,
                                int coverage, float *shift, float *score, uint8_t *valid);
// photometric_gain's samples: (reference texel, target sample) at the
// gx x gy grid of every 16th reference plane pixel, ok = 0 where the warp or
// the sample is undefined (planes padded, level 0).
hipError_t launch_gain_samples(hipStream_t stream, const struct DevPlane &reference,
                              const struct DevPlane &target, const struct DevWarp &warp,
                              uint32_t gx, uint32_t gy, float2 *samples, uint8_t *ok);
// RANSAC consensus: counts[k] = pairs (count x {a.x, a.y, b.x, b.y} f64,
// device) whose residual under hypothesis k (9 f64, row-major, device) is
// at most `threshold`.
hipError_t launch_ransac_count(hipStream_t stream, const double *pairs, uint32_t count,
                              const double *hypotheses, uint32_t hypothesis_count,
                              double threshold, uint32_t *counts);
// sample_at_reference_bandwidth of the target through `base` at every pixel
// of a width x height reference plane (resolution.rs:254-270): out[y][x] is
// the sample at ((q + 0.5) / 2 - 0.5) for q = base.map(2x + 0.5, 2y + 0.5),
// or NaN. taps = 0 samples bilinearly; otherwise `kernel` holds the
// (2 taps + 1)^2 Gaussian weights, row-major from (-taps, -taps).
struct BandwidthKernel {
    uint32_t taps;
    float weight[81];
};
hipError_t launch_render_bandwidth(hipStream_t stream, uint32_t width, uint32_t height,
                                   const struct DevPlane &target, const struct DevWarp &base,
                                   const BandwidthKernel &kernel, float *out);

// The synthesis stage: per canvas pixel and source the warp point, feather,
// local confidence, base weight and focus weight (outputs [source][row][col];
// any may be null). `touch` holds a flag per 64-row band and source.
hipError_t launch_synthesis_geometry(hipStream_t stream, const struct SynthCanvasDev &canvas,
                                     const struct DepthMapDev &map,
                                     const struct SynthSourceDev *sources, uint32_t count,
                                     const uint8_t *touch, float2 *points, float *feather,
                                     float *local_confidence, float *base_weight,
                                     float *focus_weight);

// ------------------------------------------------ synthetic capture generator

// The scene as the render kernels see it: Voronoi sites, plane inverse
// depths (1/mm) and their nearest-first order, and the texture canvas
// placement (reference raster coordinates of its first point).
constexpr uint32_t kMaxScenePlanes = 64;
struct SceneDev {
    uint64_t seed;
    uint32_t n;
    uint32_t texture_cell_px;
    double site_x[kMaxScenePlanes], site_y[kMaxScenePlanes];
    double inverse[kMaxScenePlanes];
    uint32_t order[kMaxScenePlanes];
    double canvas_x0, canvas_y0;
    uint32_t canvas_w, canvas_h;
};

// A pinhole module: principal point at the raster centre, optical centre in
// mm in the reference frame, world-to-camera rotation (row-major).
struct RenderModule {
    double focal_px;
    double center[3];
    double rotation_wc[9];
    uint32_t width, height;
};

// Sensor model (f32 fields: the noise arithmetic promotes them as written).
struct BayerSpec {
    float black_level, white_level, exposure, electrons_per_dn, read_noise_dn;
};

hipError_t launch_scene_canvas(hipStream_t stream, const SceneDev *scene, uint32_t canvas_w,
                               uint32_t canvas_h, float *canvas, uint8_t *cells);
// Bayer codes of one module, shifted left by `shift`, into out (width x height).
hipError_t launch_render_bayer(hipStream_t stream, const SceneDev *scene, const float *canvas,
                               const uint8_t *cells, const RenderModule &module,
                               const RenderModule &reference, uint32_t view, bool is_reference,
                               const BayerSpec &bayer, uint32_t shift, uint16_t *out);

}  // namespace chd
