// depth_api.hip - C ABI of the depth hot path (host side).
//
// chd_refine_multiview_depth mirrors refine_multiview_depth
// (crates/chiaro-fusion/src/depth.rs:312-580): same inputs, same None cases
// (invalid options, bad reference index, missing reference camera), same
// outputs (dense depth map on the fine grid, replaced warps and reports for
// every accepted calibrated target). A session keeps every input and
// intermediate in HBM so the path can run repeatedly without host traffic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <random>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/chiaro_depth.h"
#include "depth_device.hpp"
#include "depth_internal.hpp"
#include "proposal.hpp"

using namespace chd;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &message) {
    g_error = message;
    return code;
}

#define CHD_HIP(expr)                                                                \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess)                                                        \
            return fail(CHD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

size_t div_ceil(size_t a, size_t b) { return (a + b - 1) / b; }

bool options_valid(const chd_options &o) {  // depth.rs:1879-1891
    return o.enabled && o.grid_step > 0 && std::isfinite(o.near_depth) &&
           std::isfinite(o.far_depth) && o.near_depth > 0.0 && o.far_depth > o.near_depth &&
           o.planes >= 3 && o.patch_radius > 0 && o.minimum_support > 0 &&
           o.best_view_count >= o.minimum_support && o.minimum_neighbour_support > 0;
}

void inverse_depth_samples(double near_depth, double far_depth, uint32_t count,
                           double *out) {  // depth.rs:1834-1843
    const double near_inverse = 1.0 / near_depth, far_inverse = 1.0 / far_depth;
    for (uint32_t i = 0; i < count; i++) {
        const double w = (double)i / (double)(count - 1);
        out[i] = 1.0 / (near_inverse * (1.0 - w) + far_inverse * w);
    }
}

// label 0 = global warp (None), labels 1.. = finite depths far -> near
// (depth.rs:1196-1201)
std::vector<double> label_depths(const chd_options &o) {
    std::vector<double> finite(o.planes), labels(o.planes + 1);
    inverse_depth_samples(o.near_depth, o.far_depth, o.planes, finite.data());
    labels[0] = std::nan("");
    for (uint32_t i = 0; i < o.planes; i++) labels[1 + i] = finite[o.planes - 1 - i];
    return labels;
}

void grid_sizes(const chd_options &o, uint32_t width, uint32_t height, chd_grid_sizes &g) {
    g.coarse_step = std::max(o.grid_step, 4u);
    g.coarse_columns = (uint32_t)div_ceil(width, g.coarse_step) + 1;
    g.coarse_rows = (uint32_t)div_ceil(height, g.coarse_step) + 1;
    g.labels = o.planes + 1;
    g.fine_step = std::max(g.coarse_step / 2, 4u);
    g.fine_columns = (uint32_t)div_ceil(width, g.fine_step) + 1;
    g.fine_rows = (uint32_t)div_ceil(height, g.fine_step) + 1;
}

int32_t tkey32_host(float f) {
    int32_t k;
    std::memcpy(&k, &f, 4);
    return k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
}

// MINIMUM_DIRECT_COMPONENT_NODES (depth.rs:37)
constexpr uint32_t kMinimumComponentNodes = 24;

}  // namespace

// Device scratch of the allocation-free entry point: the union-find parents
// and component sizes of CHD_NOALLOC_MAX_NODES fine nodes (16 MiB), one copy
// per GPU for the life of the process. Host code never passes `scratch`
// itself to a kernel: its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device when the first context on
// that device is created.
static __device__ __align__(256) uint32_t scratch[2 * (size_t)CHD_NOALLOC_MAX_NODES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t *g_scratch[kMaxDevices] = {};

// Records `scratch`'s address on the current device (`device`).
int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(CHD_ERR_LIMIT, "device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return CHD_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(CHD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (uint32_t *)ptr;
    return CHD_OK;
}

uint32_t *scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch[device];
}

struct TimedLaunch {
    const char *name;
    hipEvent_t start, stop;
};

}  // namespace

struct chd_context {
    int device;
    hipStream_t stream;
    hipStream_t side = nullptr;  // second stream of pipelined batches (default priority)
    // luminance_half table for the last (black, white) pair
    float *luminance_table = nullptr;
    float luminance_black = 0.0f, luminance_white = 0.0f;
    // Live objects made on this context. chd_context_destroy with any of
    // them alive only marks the context released; the last child's destroy
    // frees it, so objects may be destroyed in any order.
    std::vector<chd_session *> sessions;
    std::vector<chd_batch *> batches;
    std::vector<chd_proposal *> proposals;
    std::vector<chd_frames *> frames;
    bool released = false;
};

namespace {

template <class T>
void unregister(std::vector<T *> &list, T *item) {
    list.erase(std::remove(list.begin(), list.end(), item), list.end());
}

void context_free(chd_context *c) {
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    if (c->luminance_table) hipFree(c->luminance_table);
    if (c->side) {
        hipStreamSynchronize(c->side);
        hipStreamDestroy(c->side);
    }
    hipStreamDestroy(c->stream);
    delete c;
}

// Frees a released context once its last child is gone.
void context_release_if_unused(chd_context *c) {
    if (c->released && c->sessions.empty() && c->batches.empty() && c->proposals.empty() &&
        c->frames.empty())
        context_free(c);
}

// Drops a destroyed session from a batch and marks the batch unusable.
void batch_forget_session(chd_batch *b, chd_session *s);

}  // namespace

struct chd_session {
    chd_context *context = nullptr;
    chd_options options{};
    uint32_t count = 0, reference = 0;
    bool none = false;
    chd_grid_sizes grid{};
    std::vector<void *> allocations;
    std::vector<uint32_t> active;  // accepted, calibrated, non-reference views
    std::vector<std::pair<uint32_t, uint32_t>> plane_sizes;  // per view
    ViewDev *views = nullptr;
    std::vector<ViewDev> host_views;  // what `views` holds
    uint32_t *vlist = nullptr;     // device copy of `active`
    double *labels = nullptr;
    float *scores = nullptr, *costs = nullptr, *guidance = nullptr;
    uint8_t *tested = nullptr;
    float *paths = nullptr, *aggregated = nullptr;
    chd_node_depth *coarse = nullptr, *coarse_tmp = nullptr;
    uint8_t *fillable = nullptr;
    uint16_t *best_label = nullptr;
    chd_node_depth *field[4] = {nullptr, nullptr, nullptr, nullptr};
    float *fine_guidance = nullptr;
    uint8_t *fine_tested = nullptr;
    WideEntry *wide = nullptr;
    uint32_t *wide_count = nullptr;
    float *direct_scores = nullptr;  // [fine node][kDirectSlots]
    uint32_t *parent = nullptr, *sizes = nullptr;
    chd_node_warp *decisions[2] = {nullptr, nullptr};  // [0]: [target slot][node]
    std::vector<float *> target_points, target_confidence;
    std::vector<uint8_t *> target_codes;
    unsigned long long *target_counts = nullptr;  // 3 per view slot
    unsigned long long *stat_counts = nullptr;    // tested, measured, regularized
    unsigned long long *depth_keys = nullptr, *depth_keys_sorted = nullptr;
    unsigned int *improvement_keys = nullptr, *improvement_keys_sorted = nullptr;
    uint32_t *key_counts = nullptr;
    void *sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    double *map_depth = nullptr;
    float *map_confidence = nullptr;
    uint8_t *map_provenance = nullptr;
    unsigned long long *work = nullptr;  // cost, direct, wide, target ZNCCs
    float *view_scores = nullptr;        // [view slot][node][label], on request
    // band sweep proposed by a chd_proposal: label windows per coarse node
    uint16_t *band_lo = nullptr, *band_hi = nullptr;
    bool band = false;
    uint32_t band_window_max = 0;  // labels of the widest band window (2 max_half + 2)
    // the proposal and band setting CHD_STAGE_PROPOSE runs (attach_proposal)
    chd_proposal *attached = nullptr;
    float attached_threshold = 0.0f;
    uint32_t attached_half = 0, attached_max_half = 0, attached_fallback = 0, attached_tile = 0;
    uint32_t version = 0;  // bumped when launch parameters change (batches recapture)
    std::vector<chd_batch *> batches;  // live batches that hold this session
    chd_tuning tuning{};
    uint32_t stages_done = 0;
    bool timing = false;
    std::vector<TimedLaunch> launches;
    std::vector<chd_kernel_time> times;

    ~chd_session() {
        for (auto &t : launches) {
            hipEventDestroy(t.start);
            hipEventDestroy(t.stop);
        }
        for (void *p : allocations) hipFree(p);
    }
};

namespace {

template <class T>
int dalloc(chd_session *s, T **out, size_t count) {
    void *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess)
        return fail(CHD_ERR_HIP, std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                     "): " + hipGetErrorString(e));
    s->allocations.push_back(p);
    *out = (T *)p;
    return CHD_OK;
}

#define CHD_CHECK(expr)        \
    do {                       \
        int _r = (expr);       \
        if (_r != CHD_OK) return _r; \
    } while (0)

// Launch wrapper: optional HIP-event timing on the session's stream.
template <class F>
int timed(chd_session *s, const char *name, F &&launch) {
    hipStream_t stream = s->context->stream;
    TimedLaunch t{name, nullptr, nullptr};
    if (s->timing) {
        CHD_HIP(hipEventCreate(&t.start));
        CHD_HIP(hipEventCreate(&t.stop));
        CHD_HIP(hipEventRecord(t.start, stream));
    }
    hipError_t e = launch();
    if (e != hipSuccess) {
        if (t.start) hipEventDestroy(t.start);
        if (t.stop) hipEventDestroy(t.stop);
        return fail(CHD_ERR_HIP, std::string(name) + ": " + hipGetErrorString(e));
    }
    if (s->timing) {
        CHD_HIP(hipEventRecord(t.stop, stream));
        s->launches.push_back(t);
    }
    return CHD_OK;
}

int collect_times(chd_session *s) {
    if (s->launches.empty()) return CHD_OK;
    CHD_HIP(hipStreamSynchronize(s->context->stream));
    for (auto &t : s->launches) {
        float ms = 0.0f;
        CHD_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        auto it = std::find_if(s->times.begin(), s->times.end(), [&](const chd_kernel_time &k) {
            return std::strncmp(k.name, t.name, sizeof k.name) == 0;
        });
        if (it == s->times.end()) {
            chd_kernel_time k{};
            std::snprintf(k.name, sizeof k.name, "%s", t.name);
            s->times.push_back(k);
            it = s->times.end() - 1;
        }
        it->launches += 1;
        it->total_ms += ms;
        hipEventDestroy(t.start);
        hipEventDestroy(t.stop);
    }
    s->launches.clear();
    return CHD_OK;
}

PatchSpec patch_spec(const chd_session *s, uint32_t radius) {
    PatchSpec p;
    p.views = s->views;
    p.vlist = s->vlist;
    p.view_count = (uint32_t)s->active.size();
    p.reference = s->reference;
    p.radius = radius;
    return p;
}

// The session's cost-volume job in its current launch shape.
CostJob cost_job(const chd_session *s, float *view_scores) {
    const chd_grid_sizes &g = s->grid;
    CostJob j{};
    j.S = patch_spec(s, s->options.patch_radius);
    j.G = GridSpec{g.coarse_columns, g.coarse_rows, g.coarse_step};
    j.labels = s->labels;
    j.L = g.labels;
    j.minimum_support = s->options.minimum_support;
    j.best_view_count = s->options.best_view_count;
    j.scores = s->scores;
    j.costs = s->costs;
    j.guidance = s->guidance;
    j.tested = s->tested;
    j.view_scores = view_scores;
    j.work = s->work;
    j.xcd_order = s->tuning.cost_tile_order;
    j.window_lo = s->band ? s->band_lo : nullptr;
    j.window_hi = s->band ? s->band_hi : nullptr;
    j.window_max = s->band ? s->band_window_max : 0;
    cost_job_shape(j, s->tuning.cost_nodes_per_block, s->tuning.cost_label_slab);
    return j;
}

// View-score slices are allocated on first request.
int view_score_buffer(chd_session *s, bool keep, float **out) {
    *out = nullptr;
    if (!keep) return CHD_OK;
    if (!s->view_scores) {
        const size_t n = std::max<size_t>(s->active.size(), 1) * (size_t)s->grid.coarse_columns *
                         s->grid.coarse_rows * s->grid.labels;
        CHD_CHECK(dalloc(s, &s->view_scores, n));
    }
    *out = s->view_scores;
    return CHD_OK;
}

// The clears a cost-volume launch needs first.
int prepare_cost(chd_session *s) {
    hipStream_t st = s->context->stream;
    CHD_HIP(launch_fill_u64(st, s->work, 0ull, 1));
    CHD_HIP(launch_fill_u8(st, s->tested, 0, (size_t)s->grid.coarse_columns * s->grid.coarse_rows));
    return CHD_OK;  // band items write NaN / +inf outside their windows themselves
}

int propose_attached(chd_session *s);  // with the proposal entries below

int run_cost(chd_session *s, bool keep_view_scores) {
    float *view_scores = nullptr;
    CHD_CHECK(view_score_buffer(s, keep_view_scores, &view_scores));
    const CostJob job = cost_job(s, view_scores);
    CHD_CHECK(prepare_cost(s));
    return timed(s, "cost_volume", [&] { return launch_cost_volume(s->context->stream, job); });
}

int run_sgm_paths(chd_session *s) {
    const chd_grid_sizes &g = s->grid;
    return timed(s, "sgm_paths", [&] {
        return launch_sgm(s->context->stream, s->costs, s->guidance, g.coarse_columns,
                          g.coarse_rows, g.labels, s->paths, s->tuning.sgm_paths_per_block,
                          s->band ? s->band_lo : nullptr, s->band ? s->band_hi : nullptr);
    });
}

// select_depths and the optional completion after the path sweeps.
int run_sgm_rest(chd_session *s, bool keep_aggregated) {
    const chd_grid_sizes &g = s->grid;
    hipStream_t st = s->context->stream;
    CHD_CHECK(timed(s, "sgm_select", [&] {
        return launch_sgm_select(st, s->paths, s->scores, s->tested, s->labels, g.coarse_columns,
                                 g.coarse_rows, g.labels, s->options,
                                 keep_aggregated ? s->aggregated : nullptr, s->coarse, s->fillable,
                                 s->best_label, s->band ? s->band_lo : nullptr,
                                 s->band ? s->band_hi : nullptr);
    }));
    if (s->options.completion_iterations > 0) {
        const size_t n = (size_t)g.coarse_columns * g.coarse_rows;
        for (uint32_t it = 0; it < s->options.completion_iterations; it++) {
            CHD_HIP(hipMemcpyAsync(s->coarse_tmp, s->coarse, n * sizeof(chd_node_depth),
                                   hipMemcpyDeviceToDevice, st));
            CHD_CHECK(timed(s, "complete_depth_field", [&] {
                return launch_complete_depth_field(st, s->coarse_tmp, s->coarse, s->guidance,
                                                   s->fillable, g.coarse_columns, g.coarse_rows,
                                                   s->options.minimum_neighbour_support);
            }));
        }
    }
    return CHD_OK;
}

double inverse_step_of(const chd_options &o) {  // depth.rs:346-347
    return (1.0 / o.near_depth - 1.0 / o.far_depth) / (double)(o.planes - 1);
}

int run_direct(chd_session *s) {
    const chd_grid_sizes &g = s->grid;
    hipStream_t st = s->context->stream;
    const chd_options &o = s->options;
    chd_options direct = o, wide = o;
    direct.patch_radius = std::max(o.patch_radius, 3u);
    wide.patch_radius = std::max(o.patch_radius * 2, 8u);
    const double inverse_step = inverse_step_of(o);
    GridSpec fine{g.fine_columns, g.fine_rows, g.fine_step};
    GridSpec coarse{g.coarse_columns, g.coarse_rows, g.coarse_step};
    const uint32_t fine_nodes = g.fine_columns * g.fine_rows;
    CHD_HIP(launch_fill_u32(st, s->wide_count, 0u, 1));
    CHD_HIP(launch_fill_u64(st, s->work + 1, 0ull, 2));
    CHD_CHECK(timed(s, "direct", [&] {
        return launch_direct(st, patch_spec(s, direct.patch_radius), fine, coarse, s->coarse,
                             inverse_step, direct, s->field[0], s->fine_guidance, s->fine_tested,
                             s->wide, s->wide_count, s->direct_scores, s->work + 1);
    }));
    CHD_CHECK(timed(s, "direct_wide", [&] {
        return launch_direct_wide(st, patch_spec(s, wide.patch_radius), fine, wide, s->wide,
                                  s->wide_count, fine_nodes, s->field[0], s->work + 2);
    }));
    CHD_CHECK(timed(s, "reject_isolated", [&] {
        return launch_reject_isolated(st, s->field[0], s->field[1], s->fine_guidance,
                                      g.fine_columns, g.fine_rows, inverse_step * 2.5);
    }));
    CHD_CHECK(timed(s, "reject_components", [&] {
        return launch_reject_components(st, s->field[1], s->field[2], s->fine_guidance,
                                        g.fine_columns, g.fine_rows, inverse_step * 2.5,
                                        kMinimumComponentNodes, s->parent, s->sizes);
    }));
    CHD_CHECK(timed(s, "fit_planes", [&] {
        return launch_fit_planes(st, s->field[2], s->field[3], s->fine_guidance, g.fine_columns,
                                 g.fine_rows, inverse_step * 4.0);
    }));
    CHD_HIP(launch_fill_u64(st, s->stat_counts, 0ull, 3));
    CHD_HIP(launch_fill_u32(st, s->key_counts, 0u, 2));
    CHD_CHECK(timed(s, "field_stats", [&] {
        return launch_field_stats(st, s->field[3], s->fine_tested, fine_nodes, s->stat_counts,
                                  s->depth_keys, s->improvement_keys, s->key_counts);
    }));
    return timed(s, "final_map", [&] {
        return launch_final_map(st, s->field[3], s->fine_tested, fine_nodes, s->map_depth,
                                s->map_confidence, s->map_provenance);
    });
}

int run_targets(chd_session *s) {
    const chd_grid_sizes &g = s->grid;
    hipStream_t st = s->context->stream;
    const chd_options &o = s->options;
    const double inverse_step = inverse_step_of(o);
    GridSpec fine{g.fine_columns, g.fine_rows, g.fine_step};
    const uint32_t fine_nodes = g.fine_columns * g.fine_rows;
    CHD_HIP(launch_fill_u64(st, s->target_counts, 0ull, 3 * (size_t)s->count));
    CHD_HIP(launch_fill_u64(st, s->work + 3, 0ull, 1));
    if (s->active.empty()) return CHD_OK;
    CHD_CHECK(timed(s, "targets", [&] {
        return launch_targets(st, patch_spec(s, o.patch_radius), fine, s->field[3], inverse_step,
                              o, s->decisions[0], s->work + 3);
    }));
    for (size_t k = 0; k < s->active.size(); k++) {
        const uint32_t t = s->active[k];
        chd_node_warp *decisions = s->decisions[0] + k * fine_nodes;
        CHD_CHECK(timed(s, "consensus", [&] {
            return launch_consensus(st, decisions, s->decisions[1], g.fine_columns, g.fine_rows);
        }));
        CHD_CHECK(timed(s, "boundaries", [&] {
            return launch_boundaries(st, s->decisions[1], decisions, s->fine_guidance,
                                     g.fine_columns, g.fine_rows);
        }));
        CHD_CHECK(timed(s, "target_outputs", [&] {
            return launch_target_outputs(st, decisions, fine_nodes, s->target_points[t],
                                         s->target_confidence[t], s->target_codes[t],
                                         s->target_counts + 3 * (size_t)t);
        }));
    }
    return CHD_OK;
}

}  // namespace

extern "C" {

const char *chd_last_error(void) { return g_error.c_str(); }
int chd_abi_version(void) { return CHD_ABI_VERSION; }

void chd_default_options(chd_options *o) {  // depth.rs:69-91
    o->enabled = 1;
    o->grid_step = 8;
    o->near_depth = 500.0;
    o->far_depth = 10000000.0;
    o->planes = 96;
    o->patch_radius = 4;
    o->minimum_support = 2;
    o->best_view_count = 3;
    o->minimum_score = 0.45f;
    o->minimum_margin = 0.01f;
    o->minimum_improvement = 0.01f;
    o->minimum_neighbour_support = 2;
    o->completion_iterations = 0;
}

int chd_options_valid(const chd_options *o) { return o && options_valid(*o) ? 1 : 0; }

int chd_grid_sizes_for(const chd_options *o, uint32_t width, uint32_t height,
                       chd_grid_sizes *out) {
    if (!o || !out || o->grid_step == 0) return fail(CHD_ERR_ARGUMENT, "grid sizes: bad argument");
    grid_sizes(*o, width, height, *out);
    return CHD_OK;
}

int chd_camera_canonical(const double k[9], const double rotation_wc[9],
                         const double translation_wc[3], uint32_t width, uint32_t height,
                         const chd_distortion *distortion, chd_camera *out) {
    // ResolvedCamera::new, canonical pose (geometry.rs:64-133)
    if (!k || !rotation_wc || !translation_wc || !out)
        return fail(CHD_ERR_ARGUMENT, "camera: null argument");
    std::memset(out, 0, sizeof *out);
    out->width = width;
    out->height = height;
    std::memcpy(out->k, k, sizeof out->k);
    const double *m = k;
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (std::fabs(det) < 1e-300) return fail(CHD_ERR_ARGUMENT, "singular intrinsic matrix");
    const double inv = 1.0 / det;
    auto c = [&](int r0, int c0, int r1, int c1) {
        return m[r0 * 3 + c0] * m[r1 * 3 + c1] - m[r0 * 3 + c1] * m[r1 * 3 + c0];
    };
    const double ki[9] = {c(1, 1, 2, 2) * inv,  -c(0, 1, 2, 2) * inv, c(0, 1, 1, 2) * inv,
                          -c(1, 0, 2, 2) * inv, c(0, 0, 2, 2) * inv,  -c(0, 0, 1, 2) * inv,
                          c(1, 0, 2, 1) * inv,  -c(0, 0, 2, 1) * inv, c(0, 0, 1, 1) * inv};
    std::memcpy(out->k_inverse, ki, sizeof ki);
    if (distortion) {
        out->has_distortion = 1;
        out->distortion = *distortion;
    }
    out->flip_around_x = -1;
    out->pose_kind = 0;
    std::memcpy(out->rotation_wc, rotation_wc, sizeof out->rotation_wc);
    const double *r = rotation_wc;
    const double rcw[9] = {r[0], r[3], r[6], r[1], r[4], r[7], r[2], r[5], r[8]};
    std::memcpy(out->rotation_cw, rcw, sizeof rcw);
    std::memcpy(out->translation_wc, translation_wc, sizeof out->translation_wc);
    const double *t = translation_wc;
    for (int i = 0; i < 3; i++) {
        const double v = rcw[i * 3] * t[0] + rcw[i * 3 + 1] * t[1] + rcw[i * 3 + 2] * t[2];
        out->center[i] = v * -1.0;
    }
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::memcpy(out->orientation_correction, I, sizeof I);
    out->focal_px = k[0];
    return CHD_OK;
}

int chd_context_create(int device, chd_context **out) {
    if (!out) return fail(CHD_ERR_ARGUMENT, "context: null output");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(CHD_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(CHD_ERR_ARGUMENT, "device index out of range");
    CHD_HIP(hipSetDevice(device));
    CHD_CHECK(scratch_setup(device));
    auto *c = new chd_context;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(CHD_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return CHD_OK;
}

void chd_context_destroy(chd_context *c) {
    if (!c || c->released) return;
    c->released = true;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    context_release_if_unused(c);
}

int chd_context_synchronize(chd_context *c) {
    if (!c) return fail(CHD_ERR_ARGUMENT, "null context");
    CHD_HIP(hipStreamSynchronize(c->stream));
    return CHD_OK;
}

}  // extern "C"

namespace {

int session_create(chd_context *context, const chd_view *views, uint32_t count,
                   uint32_t reference_index, const chd_options *options, bool device_planes,
                   chd_session **out) {
    if (!context || !views || !options || !out) return fail(CHD_ERR_ARGUMENT, "session: null argument");
    *out = nullptr;
    if (context->released) return fail(CHD_ERR_STATE, "session: the context was destroyed");
    CHD_HIP(hipSetDevice(context->device));
    auto *s = new chd_session;
    s->context = context;
    s->options = *options;
    s->count = count;
    s->reference = reference_index;
    // the reference's None cases (depth.rs:318-324)
    if (!options_valid(*options) || reference_index >= count ||
        views[reference_index].camera == nullptr) {
        s->none = true;
        context->sessions.push_back(s);
        *out = s;
        return CHD_OK;
    }
    auto bail = [&](int code) {
        delete s;
        return code;
    };
    if (count > CHD_MAX_VIEWS) return bail(fail(CHD_ERR_LIMIT, "more than CHD_MAX_VIEWS views"));
    if (options->patch_radius > CHD_MAX_PATCH_RADIUS / 2)
        return bail(fail(CHD_ERR_LIMIT, "patch_radius above CHD_MAX_PATCH_RADIUS / 2"));
    if (options->planes + 1 > 512) return bail(fail(CHD_ERR_LIMIT, "more than 511 planes"));
    for (uint32_t i = 0; i < count; i++) {
        const chd_view &v = views[i];
        if (!v.luminance || v.plane_width == 0 || v.plane_height == 0)
            return bail(fail(CHD_ERR_ARGUMENT, "view " + std::to_string(i) + ": empty luminance plane"));
        // the kernels address a plane (one padding column and row added)
        // through a 32-bit byte range and 24-bit row products
        if (((size_t)v.plane_width + 1) * ((size_t)v.plane_height + 1) * 4 > 0xffffffffull ||
            v.plane_width >= (1u << 24) - 1 || v.plane_height >= (1u << 24) - 1)
            return bail(fail(CHD_ERR_LIMIT, "view " + std::to_string(i) + ": luminance plane above 4 GiB"));
        if (i != reference_index && v.accepted && v.camera &&
            (!v.global_warp.points || v.global_warp.step == 0 || v.global_warp.columns == 0 ||
             v.global_warp.rows == 0))
            return bail(fail(CHD_ERR_ARGUMENT, "view " + std::to_string(i) + ": empty global warp"));
    }
    grid_sizes(*options, views[reference_index].width, views[reference_index].height, s->grid);
    const chd_grid_sizes &g = s->grid;
    const size_t coarse_nodes = (size_t)g.coarse_columns * g.coarse_rows;
    const size_t fine_nodes = (size_t)g.fine_columns * g.fine_rows;
    const size_t volume = coarse_nodes * g.labels;
    if (volume * 8 > ((size_t)1 << 36))
        return bail(fail(CHD_ERR_LIMIT, "cost volume above the session limit"));
    for (uint32_t i = 0; i < count; i++)
        if (i != reference_index && views[i].accepted && views[i].camera) s->active.push_back(i);

    hipStream_t st = context->stream;
    std::vector<ViewDev> host_views(count);
    std::vector<chd_camera *> cams(count, nullptr);
    int r = CHD_OK;
#define CHD_ALLOC(ptr, n)                       \
    do {                                        \
        r = dalloc(s, &(ptr), (n));             \
        if (r != CHD_OK) return bail(r);        \
    } while (0)
#define CHD_UPLOAD(dst, src, bytes)                                                          \
    do {                                                                                     \
        hipError_t _e = hipMemcpyAsync((dst), (src), (bytes), hipMemcpyHostToDevice, st);    \
        if (_e != hipSuccess) return bail(fail(CHD_ERR_HIP, std::string("upload: ") + hipGetErrorString(_e))); \
    } while (0)
    for (uint32_t i = 0; i < count; i++) {
        const chd_view &v = views[i];
        ViewDev &d = host_views[i];
        std::memset(&d, 0, sizeof d);
        // planes are stored with one replicated extra column and row
        float *plane = nullptr;
        const size_t pitch = (size_t)v.plane_width + 1;
        CHD_ALLOC(plane, pitch * ((size_t)v.plane_height + 1));
        {
            hipError_t e2 = hipMemcpy2DAsync(plane, pitch * sizeof(float), v.luminance,
                                             (size_t)v.plane_width * sizeof(float),
                                             (size_t)v.plane_width * sizeof(float),
                                             v.plane_height,
                                             device_planes ? hipMemcpyDeviceToDevice
                                                           : hipMemcpyHostToDevice,
                                             st);
            if (e2 == hipSuccess) e2 = launch_pad_plane(st, plane, v.plane_width, v.plane_height);
            if (e2 != hipSuccess)
                return bail(fail(CHD_ERR_HIP, std::string("plane upload: ") + hipGetErrorString(e2)));
        }
        s->plane_sizes.emplace_back(v.plane_width, v.plane_height);
        d.plane = plane;
        d.plane_width = v.plane_width;
        d.plane_height = v.plane_height;
        d.plane_pitch = (uint32_t)pitch;
        d.columns = nullptr;
        if (v.global_warp.points && v.global_warp.columns && v.global_warp.rows) {
            float2 *warp = nullptr;
            const size_t wn = (size_t)v.global_warp.columns * v.global_warp.rows;
            CHD_ALLOC(warp, wn);
            CHD_UPLOAD(warp, v.global_warp.points, wn * sizeof(float2));
            d.warp = warp;
            d.warp_step = v.global_warp.step;
            d.warp_columns = v.global_warp.columns;
            d.warp_rows = v.global_warp.rows;
        }
        if (v.camera) {
            chd_camera *cam = nullptr;
            CHD_ALLOC(cam, 1);
            CHD_UPLOAD(cam, v.camera, sizeof(chd_camera));
            d.camera = cam;
        }
    }
    CHD_ALLOC(s->views, count);
    CHD_UPLOAD(s->views, host_views.data(), count * sizeof(ViewDev));
    CHD_ALLOC(s->vlist, std::max<size_t>(s->active.size(), 1));
    if (!s->active.empty()) CHD_UPLOAD(s->vlist, s->active.data(), s->active.size() * sizeof(uint32_t));
    if (!s->active.empty()) {
        // Target planes whose epipolar lines run mostly vertically get a
        // column-major copy, so that the cost volume's samples of
        // neighbouring hypotheses (one per lane) share cache lines. The
        // samples are the same.
        uint32_t *vertical_d = nullptr;
        CHD_ALLOC(vertical_d, s->active.size());
        std::vector<uint32_t> vertical(s->active.size(), 0);
        hipError_t e = launch_epipolar_axes(st, s->views, s->vlist, (uint32_t)s->active.size(),
                                            reference_index, options->near_depth,
                                            options->far_depth, vertical_d);
        if (e == hipSuccess)
            e = hipMemcpyAsync(vertical.data(), vertical_d, vertical.size() * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            return bail(fail(CHD_ERR_HIP, std::string("epipolar axes: ") + hipGetErrorString(e)));
        bool any = false;
        for (size_t k = 0; k < s->active.size(); k++) {
            if (!vertical[k]) continue;
            ViewDev &d = host_views[s->active[k]];
            float *columns = nullptr;
            CHD_ALLOC(columns, ((size_t)d.plane_width + 1) * ((size_t)d.plane_height + 1));
            e = launch_transpose_plane(st, d.plane, columns, d.plane_width, d.plane_height);
            if (e != hipSuccess)
                return bail(fail(CHD_ERR_HIP, std::string("plane transpose: ") + hipGetErrorString(e)));
            d.columns = columns;
            any = true;
        }
        // every target view's quad array, of the plane its tap loops read
        for (size_t k = 0; k < s->active.size(); k++) {
            ViewDev &d = host_views[s->active[k]];
            const uint32_t pitch = d.columns ? d.plane_height + 1 : d.plane_pitch;
            const uint32_t lines = d.columns ? d.plane_width + 1 : d.plane_height + 1;
            float4 *quad = nullptr;
            CHD_ALLOC(quad, (size_t)pitch * lines);
            e = launch_quad_plane(st, d.columns ? d.columns : d.plane, pitch, lines, quad);
            if (e != hipSuccess)
                return bail(fail(CHD_ERR_HIP, std::string("quad plane: ") + hipGetErrorString(e)));
            d.quad = quad;
            any = true;
        }
        if (any) CHD_UPLOAD(s->views, host_views.data(), count * sizeof(ViewDev));
    }
    s->host_views = host_views;
    std::vector<double> labels = label_depths(*options);
    CHD_ALLOC(s->labels, labels.size());
    CHD_UPLOAD(s->labels, labels.data(), labels.size() * sizeof(double));

    CHD_ALLOC(s->scores, volume);
    CHD_ALLOC(s->costs, volume);
    CHD_ALLOC(s->guidance, coarse_nodes);
    CHD_ALLOC(s->tested, coarse_nodes);
    CHD_ALLOC(s->paths, volume * 8);
    CHD_ALLOC(s->aggregated, volume);
    CHD_ALLOC(s->coarse, coarse_nodes);
    CHD_ALLOC(s->coarse_tmp, coarse_nodes);
    CHD_ALLOC(s->fillable, coarse_nodes);
    CHD_ALLOC(s->best_label, coarse_nodes);
    for (int k = 0; k < 4; k++) CHD_ALLOC(s->field[k], fine_nodes);
    CHD_ALLOC(s->fine_guidance, fine_nodes);
    CHD_ALLOC(s->fine_tested, fine_nodes);
    CHD_ALLOC(s->wide, fine_nodes);
    CHD_ALLOC(s->wide_count, 1);
    CHD_ALLOC(s->direct_scores, (size_t)fine_nodes * kDirectSlots);
    CHD_ALLOC(s->parent, fine_nodes);
    CHD_ALLOC(s->sizes, fine_nodes);
    CHD_ALLOC(s->decisions[0], std::max<size_t>(s->active.size(), 1) * fine_nodes);
    CHD_ALLOC(s->decisions[1], fine_nodes);
    s->target_points.assign(count, nullptr);
    s->target_confidence.assign(count, nullptr);
    s->target_codes.assign(count, nullptr);
    for (uint32_t t : s->active) {
        CHD_ALLOC(s->target_points[t], fine_nodes * 2);
        CHD_ALLOC(s->target_confidence[t], fine_nodes);
        CHD_ALLOC(s->target_codes[t], fine_nodes);
    }
    CHD_ALLOC(s->target_counts, 3 * (size_t)count);
    CHD_ALLOC(s->stat_counts, 3);
    CHD_ALLOC(s->depth_keys, fine_nodes);
    CHD_ALLOC(s->depth_keys_sorted, fine_nodes);
    CHD_ALLOC(s->improvement_keys, fine_nodes);
    CHD_ALLOC(s->improvement_keys_sorted, fine_nodes);
    CHD_ALLOC(s->key_counts, 2);
    CHD_ALLOC(s->work, 4);
    CHD_ALLOC(s->map_depth, fine_nodes);
    CHD_ALLOC(s->map_confidence, fine_nodes);
    CHD_ALLOC(s->map_provenance, fine_nodes);
    {
        size_t a = 0, b = 0;
        (void)hipcub::DeviceRadixSort::SortKeys(nullptr, a, s->depth_keys, s->depth_keys_sorted,
                                          (int)fine_nodes, 0, 64, st);
        (void)hipcub::DeviceRadixSort::SortKeys(nullptr, b, s->improvement_keys,
                                          s->improvement_keys_sorted, (int)fine_nodes, 0, 32, st);
        s->sort_temp_bytes = std::max(a, b);
        unsigned char *tmp = nullptr;
        CHD_ALLOC(tmp, s->sort_temp_bytes);
        s->sort_temp = tmp;
    }
#undef CHD_ALLOC
#undef CHD_UPLOAD
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bail(fail(CHD_ERR_HIP, std::string("upload sync: ") + hipGetErrorString(e)));
    context->sessions.push_back(s);
    *out = s;
    return CHD_OK;
}

}  // namespace

extern "C" {

int chd_session_create(chd_context *context, const chd_view *views, uint32_t count,
                       uint32_t reference_index, const chd_options *options,
                       chd_session **out) {
    return session_create(context, views, count, reference_index, options, false, out);
}

int chd_session_create_device(chd_context *context, const chd_view *views, uint32_t count,
                              uint32_t reference_index, const chd_options *options,
                              chd_session **out) {
    return session_create(context, views, count, reference_index, options, true, out);
}

int chd_session_set_planes(chd_session *s, const float *const *planes, uint32_t count,
                           int device) {
    if (!s || !planes) return fail(CHD_ERR_ARGUMENT, "set_planes: bad argument");
    if (s->none) return CHD_NONE;
    if (count != s->count) return fail(CHD_ERR_ARGUMENT, "set_planes: one plane per view");
    for (uint32_t i = 0; i < count; i++)
        if (!planes[i]) return fail(CHD_ERR_ARGUMENT, "set_planes: null plane");
    CHD_HIP(hipSetDevice(s->context->device));
    hipStream_t st = s->context->stream;
    for (uint32_t i = 0; i < count; i++) {
        const ViewDev &d = s->host_views[i];
        float *plane = const_cast<float *>(d.plane);
        CHD_HIP(hipMemcpy2DAsync(plane, (size_t)d.plane_pitch * sizeof(float), planes[i],
                                 (size_t)d.plane_width * sizeof(float),
                                 (size_t)d.plane_width * sizeof(float), d.plane_height,
                                 device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        CHD_HIP(launch_pad_plane(st, plane, d.plane_width, d.plane_height));
        if (d.columns)
            CHD_HIP(launch_transpose_plane(st, d.plane, const_cast<float *>(d.columns),
                                           d.plane_width, d.plane_height));
        if (d.quad)
            CHD_HIP(launch_quad_plane(st, d.columns ? d.columns : d.plane,
                                      d.columns ? d.plane_height + 1 : d.plane_pitch,
                                      d.columns ? d.plane_width + 1 : d.plane_height + 1,
                                      const_cast<float4 *>(d.quad)));
    }
    if (!device) CHD_HIP(hipStreamSynchronize(st));  // the host planes may go away
    s->stages_done = 0;
    return CHD_OK;
}

void chd_session_destroy(chd_session *s) {
    if (!s) return;
    chd_context *c = s->context;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    // a batch that holds this session can no longer run (chd_batch_run
    // returns CHD_ERR_STATE); it stays valid for chd_batch_destroy
    for (chd_batch *b : s->batches) batch_forget_session(b, s);
    unregister(c->sessions, s);
    delete s;
    context_release_if_unused(c);
}

int chd_session_grid(const chd_session *s, chd_grid_sizes *out) {
    if (!s || !out) return fail(CHD_ERR_ARGUMENT, "null argument");
    if (s->none) return CHD_NONE;
    *out = s->grid;
    return CHD_OK;
}

int chd_session_run(chd_session *s, uint32_t stages) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    if (s->none) return CHD_NONE;
    CHD_HIP(hipSetDevice(s->context->device));
    const bool keep = (stages & CHD_STAGE_KEEP_AGGREGATED) != 0;
    const bool keep_views = (stages & CHD_STAGE_KEEP_VIEW_SCORES) != 0;
    if (stages & CHD_STAGE_PROPOSE) {
        if (!s->attached) return fail(CHD_ERR_STATE, "propose: no proposal attached");
        CHD_CHECK(propose_attached(s));
    }
    if (stages & CHD_STAGE_COST) {
        CHD_CHECK(run_cost(s, keep_views));
        s->stages_done = (s->stages_done & ~CHD_STAGE_KEEP_VIEW_SCORES) | CHD_STAGE_COST |
                         (keep_views ? CHD_STAGE_KEEP_VIEW_SCORES : 0);
    }
    if (stages & CHD_STAGE_SGM) {
        if (!(s->stages_done & CHD_STAGE_COST)) return fail(CHD_ERR_STATE, "SGM before cost volume");
        CHD_CHECK(run_sgm_paths(s));
        CHD_CHECK(run_sgm_rest(s, keep));
        s->stages_done = (s->stages_done & ~CHD_STAGE_KEEP_AGGREGATED) | CHD_STAGE_SGM |
                         (keep ? CHD_STAGE_KEEP_AGGREGATED : 0);
    }
    if (stages & CHD_STAGE_DIRECT) {
        if (!(s->stages_done & CHD_STAGE_SGM)) return fail(CHD_ERR_STATE, "direct pass before SGM");
        CHD_CHECK(run_direct(s));
        s->stages_done |= CHD_STAGE_DIRECT;
    }
    if (stages & CHD_STAGE_TARGETS) {
        if (!(s->stages_done & CHD_STAGE_DIRECT)) return fail(CHD_ERR_STATE, "targets before direct pass");
        CHD_CHECK(run_targets(s));
        s->stages_done |= CHD_STAGE_TARGETS;
    }
    return CHD_OK;
}

extern "C++" {
namespace {

int download(chd_session *s, void *dst, const void *src, size_t bytes) {
    if (!dst) return CHD_OK;
    CHD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->context->stream));
    return CHD_OK;
}

// k-th smallest key of `n` device keys (radix sort), or false when n == 0.
template <class K>
int median_key(chd_session *s, K *keys, K *sorted, uint32_t n, int bits, K &out, bool &found) {
    found = false;
    if (n == 0) return CHD_OK;
    hipStream_t st = s->context->stream;
    size_t temp = s->sort_temp_bytes;
    CHD_HIP(hipcub::DeviceRadixSort::SortKeys(s->sort_temp, temp, keys, sorted, (int)n, 0, bits, st));
    CHD_HIP(hipMemcpyAsync(&out, sorted + n / 2, sizeof(K), hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    found = true;
    return CHD_OK;
}

}  // namespace
}  // extern "C++"

int chd_session_read_results(chd_session *s, chd_depth_map *map, chd_target_result *targets) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    if (s->none) return CHD_NONE;
    if (!(s->stages_done & CHD_STAGE_DIRECT)) return fail(CHD_ERR_STATE, "results before the direct pass");
    const chd_grid_sizes &g = s->grid;
    const size_t fine = (size_t)g.fine_columns * g.fine_rows;
    hipStream_t st = s->context->stream;
    if (map) {
        map->columns = g.fine_columns;
        map->rows = g.fine_rows;
        map->step = g.fine_step;
        map->near_depth = s->options.near_depth;
        map->far_depth = s->options.far_depth;
        CHD_CHECK(download(s, map->depth, s->map_depth, fine * sizeof(double)));
        CHD_CHECK(download(s, map->confidence, s->map_confidence, fine * sizeof(float)));
        CHD_CHECK(download(s, map->provenance, s->map_provenance, fine));
    }
    if (targets) {
        unsigned long long stats[3], counts[3 * CHD_MAX_VIEWS];
        uint32_t key_counts[2];
        CHD_HIP(hipMemcpyAsync(stats, s->stat_counts, sizeof stats, hipMemcpyDeviceToHost, st));
        CHD_HIP(hipMemcpyAsync(counts, s->target_counts, 3 * s->count * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, st));
        CHD_HIP(hipMemcpyAsync(key_counts, s->key_counts, sizeof key_counts, hipMemcpyDeviceToHost, st));
        CHD_HIP(hipStreamSynchronize(st));
        unsigned long long dkey = 0;
        unsigned int ikey = 0;
        bool has_depth = false, has_improvement = false;
        CHD_CHECK(median_key(s, s->depth_keys, s->depth_keys_sorted, key_counts[0], 64, dkey, has_depth));
        CHD_CHECK(median_key(s, s->improvement_keys, s->improvement_keys_sorted, key_counts[1], 32,
                             ikey, has_improvement));
        double median_depth = 0.0;
        float median_improvement = 0.0f;
        if (has_depth) {
            int64_t k = (int64_t)(dkey ^ 0x8000000000000000ull);
            k ^= (int64_t)(((uint64_t)(k >> 63)) >> 1);
            std::memcpy(&median_depth, &k, 8);
        }
        if (has_improvement) {
            int32_t k = (int32_t)(ikey ^ 0x80000000u);
            k ^= (int32_t)(((uint32_t)(k >> 31)) >> 1);
            std::memcpy(&median_improvement, &k, 4);
        }
        const bool targets_ran = (s->stages_done & CHD_STAGE_TARGETS) != 0;
        const unsigned long long tested = stats[0], measured = stats[1], regularized = stats[2];
        const unsigned long long fallback =
            tested > measured + regularized ? tested - (measured + regularized) : 0;
        auto fraction = [](unsigned long long c, unsigned long long total) {
            return total == 0 ? 0.0f : (float)c / (float)total;
        };
        for (uint32_t t = 0; t < s->count; t++) targets[t].replaced = 0;
        for (uint32_t t : s->active) {
            chd_target_result &out = targets[t];
            if (!targets_ran) continue;
            CHD_CHECK(download(s, out.points, s->target_points[t], fine * 2 * sizeof(float)));
            CHD_CHECK(download(s, out.confidence, s->target_confidence[t], fine * sizeof(float)));
            CHD_CHECK(download(s, out.decisions, s->target_codes[t], fine));
            out.replaced = 1;
            chd_depth_report &r = out.report;
            std::memset(&r, 0, sizeof r);
            r.tested_nodes = tested;
            r.measured_nodes = measured;
            r.regularized_nodes = regularized;
            r.fallback_nodes = fallback;
            r.refined_nodes = counts[3 * t + 0];
            r.occluded_nodes = counts[3 * t + 1];
            r.boundary_nodes = counts[3 * t + 2];
            r.reconstructed_fraction = fraction(measured + regularized, tested);
            r.refined_fraction = fraction(r.refined_nodes, tested);
            r.occluded_fraction = fraction(r.occluded_nodes, tested);
            r.has_median_depth = has_depth;
            r.median_depth = median_depth;
            r.has_median_score_improvement = has_improvement;
            r.median_score_improvement = median_improvement;
        }
    }
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

int chd_session_read_cost_volume(chd_session *s, float *scores, float *costs, float *guidance,
                                 uint8_t *tested) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    if (s->none) return CHD_NONE;
    if (!(s->stages_done & CHD_STAGE_COST)) return fail(CHD_ERR_STATE, "cost volume not built");
    const size_t nodes = (size_t)s->grid.coarse_columns * s->grid.coarse_rows;
    const size_t volume = nodes * s->grid.labels;
    CHD_CHECK(download(s, scores, s->scores, volume * sizeof(float)));
    CHD_CHECK(download(s, costs, s->costs, volume * sizeof(float)));
    CHD_CHECK(download(s, guidance, s->guidance, nodes * sizeof(float)));
    CHD_CHECK(download(s, tested, s->tested, nodes));
    CHD_HIP(hipStreamSynchronize(s->context->stream));
    return CHD_OK;
}

int chd_session_read_aggregated(chd_session *s, float *aggregated) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    if (s->none) return CHD_NONE;
    if (!(s->stages_done & CHD_STAGE_KEEP_AGGREGATED))
        return fail(CHD_ERR_STATE, "aggregated volume not kept (run with CHD_STAGE_KEEP_AGGREGATED)");
    const size_t volume = (size_t)s->grid.coarse_columns * s->grid.coarse_rows * s->grid.labels;
    CHD_CHECK(download(s, aggregated, s->aggregated, volume * sizeof(float)));
    CHD_HIP(hipStreamSynchronize(s->context->stream));
    return CHD_OK;
}

int chd_session_read_view_scores(chd_session *s, float *out) {
    if (!s || !out) return fail(CHD_ERR_ARGUMENT, "read_view_scores: bad argument");
    if (s->none) return CHD_NONE;
    if (!(s->stages_done & CHD_STAGE_KEEP_VIEW_SCORES))
        return fail(CHD_ERR_STATE, "view scores not kept (run with CHD_STAGE_KEEP_VIEW_SCORES)");
    const size_t n = s->active.size() * (size_t)s->grid.coarse_columns * s->grid.coarse_rows *
                     s->grid.labels;
    CHD_CHECK(download(s, out, s->view_scores, n * sizeof(float)));
    CHD_HIP(hipStreamSynchronize(s->context->stream));
    return CHD_OK;
}

int chd_session_set_tuning(chd_session *s, const chd_tuning *t) {
    if (!s || !t) return fail(CHD_ERR_ARGUMENT, "set_tuning: bad argument");
    if (t->sgm_paths_per_block != 0 && t->sgm_paths_per_block != 1 &&
        t->sgm_paths_per_block != 2 && t->sgm_paths_per_block != 4)
        return fail(CHD_ERR_ARGUMENT, "set_tuning: sgm_paths_per_block must be 1, 2 or 4");
    if (t->cost_nodes_per_block > 64)
        return fail(CHD_ERR_LIMIT, "set_tuning: more than 64 nodes per block");
    if (t->cost_tile_order > 2) return fail(CHD_ERR_ARGUMENT, "set_tuning: cost_tile_order is 0, 1 or 2");
    if (!s->none) {
        chd_tuning old = s->tuning;
        s->tuning = *t;
        if (cost_job_lds(cost_job(s, nullptr)) > 160 * 1024) {
            s->tuning = old;
            return fail(CHD_ERR_LIMIT, "set_tuning: cost work item above 160 KB of LDS");
        }
    }
    s->tuning = *t;
    s->version++;
    return CHD_OK;
}

int chd_session_get_tuning(const chd_session *s, chd_tuning *out) {
    if (!s || !out) return fail(CHD_ERR_ARGUMENT, "get_tuning: bad argument");
    *out = s->tuning;
    return CHD_OK;
}

int chd_session_occupancy(chd_session *s, uint32_t blocks_per_cu[3]) {
    if (!s || !blocks_per_cu) return fail(CHD_ERR_ARGUMENT, "occupancy: bad argument");
    if (s->none) return CHD_NONE;
    CHD_HIP(hipSetDevice(s->context->device));
    const chd_options &o = s->options;
    int blocks[3] = {0, 0, 0};
    CHD_HIP(zncc_occupancy(cost_job(s, nullptr),
                           patch_spec(s, std::max(o.patch_radius, 3u)),
                           patch_spec(s, o.patch_radius), s->grid.fine_step,
                           o.best_view_count, blocks));
    for (int k = 0; k < 3; k++) blocks_per_cu[k] = (uint32_t)blocks[k];
    return CHD_OK;
}

int chd_session_read_coarse(chd_session *s, chd_node_depth *field, uint8_t *fillable,
                            uint16_t *best_label) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    if (s->none) return CHD_NONE;
    if (!(s->stages_done & CHD_STAGE_SGM)) return fail(CHD_ERR_STATE, "SGM not run");
    const size_t nodes = (size_t)s->grid.coarse_columns * s->grid.coarse_rows;
    CHD_CHECK(download(s, field, s->coarse, nodes * sizeof(chd_node_depth)));
    CHD_CHECK(download(s, fillable, s->fillable, nodes));
    CHD_CHECK(download(s, best_label, s->best_label, nodes * sizeof(uint16_t)));
    CHD_HIP(hipStreamSynchronize(s->context->stream));
    return CHD_OK;
}

int chd_session_read_direct(chd_session *s, uint32_t stage, chd_node_depth *field,
                            float *guidance, uint8_t *tested) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    if (s->none) return CHD_NONE;
    if (stage > 3) return fail(CHD_ERR_ARGUMENT, "direct stage must be 0..3");
    if (!(s->stages_done & CHD_STAGE_DIRECT)) return fail(CHD_ERR_STATE, "direct pass not run");
    const size_t nodes = (size_t)s->grid.fine_columns * s->grid.fine_rows;
    CHD_CHECK(download(s, field, s->field[stage], nodes * sizeof(chd_node_depth)));
    CHD_CHECK(download(s, guidance, s->fine_guidance, nodes * sizeof(float)));
    CHD_CHECK(download(s, tested, s->fine_tested, nodes));
    CHD_HIP(hipStreamSynchronize(s->context->stream));
    return CHD_OK;
}

int chd_session_work(chd_session *s, chd_work *out) {
    if (!s || !out) return fail(CHD_ERR_ARGUMENT, "null argument");
    if (s->none) return CHD_NONE;
    unsigned long long w[4];
    uint32_t wide = 0;
    hipStream_t st = s->context->stream;
    CHD_HIP(hipMemcpyAsync(w, s->work, sizeof w, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipMemcpyAsync(&wide, s->wide_count, sizeof wide, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    out->cost_zncc = w[0];
    out->direct_zncc = w[1];
    out->wide_zncc = w[2];
    out->target_zncc = w[3];
    out->wide_nodes = wide;
    return CHD_OK;
}

int chd_session_enable_timing(chd_session *s, int enable) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    s->timing = enable != 0;
    return CHD_OK;
}

int chd_session_kernel_times(chd_session *s, chd_kernel_time *out, uint32_t capacity,
                             uint32_t *count) {
    if (!s || !count) return fail(CHD_ERR_ARGUMENT, "null argument");
    CHD_CHECK(collect_times(s));
    *count = (uint32_t)s->times.size();
    for (uint32_t i = 0; i < capacity && i < s->times.size(); i++) out[i] = s->times[i];
    return CHD_OK;
}

int chd_session_reset_kernel_times(chd_session *s) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    CHD_CHECK(collect_times(s));
    s->times.clear();
    return CHD_OK;
}

int chd_refine_multiview_depth(chd_context *context, const chd_view *views, uint32_t count,
                               uint32_t reference_index, const chd_options *options,
                               chd_depth_map *map, chd_target_result *targets) {
    chd_session *s = nullptr;
    int r = chd_session_create(context, views, count, reference_index, options, &s);
    if (r != CHD_OK) return r;
    if (s->none) {
        chd_session_destroy(s);
        return CHD_NONE;
    }
    r = chd_session_run(s, CHD_STAGE_ALL);
    if (r == CHD_OK) r = chd_session_read_results(s, map, targets);
    chd_session_destroy(s);
    return r;
}

// ------------------------------------------------- stage entry points

extern "C++" {
namespace {

// The luminance_half table (image.rs:401-426 with a unit flat field, after
// reserve_highlight_headroom's halving of black and white when `headroom`):
// entry s is the plane value of a cell whose samples sum to s, computed in
// f32 with the platform libm's logf as the reference computes it. Cached per
// context for the last (black, white) pair.
int luminance_table(chd_context *c, float black, float white, bool headroom) {
    if (headroom) {
        black *= 0.5f;
        white *= 0.5f;
    }
    if (c->luminance_table && std::memcmp(&black, &c->luminance_black, 4) == 0 &&
        std::memcmp(&white, &c->luminance_white, 4) == 0)
        return CHD_OK;
    std::vector<float> table(kLuminanceTableSize);
    const float range = std::fmax(white - black, 1.0f);  // f32::max: NaN gives 1
    for (uint32_t s = 0; s < kLuminanceTableSize; s++) {
        const float linear = std::fmax(((float)s * 0.25f - black) / range, 0.0f) * 1.0f;
        table[s] = logf(1.0f + 1000.0f * linear);
    }
    CHD_HIP(hipStreamSynchronize(c->stream));  // the previous table may be in use
    if (!c->luminance_table) {
        hipError_t e = hipMalloc(&c->luminance_table, kLuminanceTableSize * sizeof(float));
        if (e != hipSuccess) {
            c->luminance_table = nullptr;
            return fail(CHD_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
    }
    CHD_HIP(hipMemcpyAsync(c->luminance_table, table.data(), table.size() * sizeof(float),
                           hipMemcpyHostToDevice, c->stream));
    CHD_HIP(hipStreamSynchronize(c->stream));
    c->luminance_black = black;
    c->luminance_white = white;
    return CHD_OK;
}

struct Scratch {
    std::vector<void *> p;
    ~Scratch() {
        for (void *q : p) hipFree(q);
    }
    template <class T>
    int alloc(T **out, size_t n) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return fail(CHD_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        p.push_back(q);
        *out = (T *)q;
        return CHD_OK;
    }
};

}  // namespace
}  // extern "C++"

int chd_semi_global_costs(chd_context *c, const float *costs, const float *guidance,
                          uint32_t columns, uint32_t rows, uint32_t labels, float *aggregated) {
    if (!c || !costs || !guidance || !aggregated || columns == 0 || rows == 0 || labels == 0)
        return fail(CHD_ERR_ARGUMENT, "semi_global_costs: bad argument");
    if (labels > 512) return fail(CHD_ERR_LIMIT, "semi_global_costs: more than 512 labels");
    CHD_HIP(hipSetDevice(c->device));
    const size_t nodes = (size_t)columns * rows, volume = nodes * labels;
    Scratch sc;
    float *d_costs, *d_guidance, *d_paths, *d_out;
    CHD_CHECK(sc.alloc(&d_costs, volume));
    CHD_CHECK(sc.alloc(&d_guidance, nodes));
    CHD_CHECK(sc.alloc(&d_paths, volume * 8));
    CHD_CHECK(sc.alloc(&d_out, volume));
    hipStream_t st = c->stream;
    CHD_HIP(hipMemcpyAsync(d_costs, costs, volume * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(d_guidance, guidance, nodes * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(launch_sgm(st, d_costs, d_guidance, columns, rows, labels, d_paths, 4));
    // the path sum with every node untested: select_depths writes None
    double *d_labels;
    uint8_t *d_tested, *d_fill;
    chd_node_depth *d_field;
    CHD_CHECK(sc.alloc(&d_labels, labels));
    CHD_CHECK(sc.alloc(&d_tested, nodes));
    CHD_CHECK(sc.alloc(&d_fill, nodes));
    CHD_CHECK(sc.alloc(&d_field, nodes));
    CHD_HIP(hipMemsetAsync(d_tested, 0, nodes, st));
    chd_options o;
    chd_default_options(&o);
    CHD_HIP(launch_sgm_select(st, d_paths, d_costs, d_tested, d_labels, columns, rows, labels, o,
                              d_out, d_field, d_fill, nullptr));
    CHD_HIP(hipMemcpyAsync(aggregated, d_out, volume * 4, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

int chd_select_depths(chd_context *c, const double *label_depths, uint32_t labels,
                      const float *scores, const uint8_t *tested, const float *regularised,
                      uint32_t nodes, const chd_options *options, chd_node_depth *field,
                      uint8_t *fillable, uint16_t *best_label) {
    if (!c || !label_depths || !scores || !tested || !regularised || !options || !field || !fillable)
        return fail(CHD_ERR_ARGUMENT, "select_depths: bad argument");
    if (labels > 512) return fail(CHD_ERR_LIMIT, "select_depths: more than 512 labels");
    CHD_HIP(hipSetDevice(c->device));
    const size_t volume = (size_t)nodes * labels;
    Scratch sc;
    double *d_labels;
    float *d_scores, *d_reg;
    uint8_t *d_tested, *d_fill;
    chd_node_depth *d_field;
    uint16_t *d_best;
    CHD_CHECK(sc.alloc(&d_labels, labels));
    CHD_CHECK(sc.alloc(&d_scores, volume));
    CHD_CHECK(sc.alloc(&d_reg, volume));
    CHD_CHECK(sc.alloc(&d_tested, nodes));
    CHD_CHECK(sc.alloc(&d_fill, nodes));
    CHD_CHECK(sc.alloc(&d_field, nodes));
    CHD_CHECK(sc.alloc(&d_best, nodes));
    hipStream_t st = c->stream;
    CHD_HIP(hipMemcpyAsync(d_labels, label_depths, labels * 8, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(d_scores, scores, volume * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(d_reg, regularised, volume * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(d_tested, tested, nodes, hipMemcpyHostToDevice, st));
    CHD_HIP(launch_select_depths(st, d_reg, d_scores, d_tested, d_labels, nodes, labels, *options,
                                 d_field, d_fill, d_best));
    CHD_HIP(hipMemcpyAsync(field, d_field, nodes * sizeof(chd_node_depth), hipMemcpyDeviceToHost, st));
    CHD_HIP(hipMemcpyAsync(fillable, d_fill, nodes, hipMemcpyDeviceToHost, st));
    if (best_label) CHD_HIP(hipMemcpyAsync(best_label, d_best, nodes * 2, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

extern "C++" {
namespace {

template <class Launch>
int field_stage(chd_context *c, chd_node_depth *field, const float *guidance, uint32_t columns,
                uint32_t rows, Launch &&launch) {
    if (!c || !field || !guidance || columns == 0 || rows == 0)
        return fail(CHD_ERR_ARGUMENT, "field stage: bad argument");
    CHD_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)columns * rows;
    Scratch sc;
    chd_node_depth *src, *dst;
    float *g;
    uint32_t *parent, *sizes;
    CHD_CHECK(sc.alloc(&src, n));
    CHD_CHECK(sc.alloc(&dst, n));
    CHD_CHECK(sc.alloc(&g, n));
    CHD_CHECK(sc.alloc(&parent, n));
    CHD_CHECK(sc.alloc(&sizes, n));
    hipStream_t st = c->stream;
    CHD_HIP(hipMemcpyAsync(src, field, n * sizeof(chd_node_depth), hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(g, guidance, n * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(launch(st, src, dst, g, parent, sizes));
    CHD_HIP(hipMemcpyAsync(field, dst, n * sizeof(chd_node_depth), hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

template <class Launch>
int decision_stage(chd_context *c, chd_node_warp *decisions, const float *guidance,
                   uint32_t columns, uint32_t rows, Launch &&launch) {
    if (!c || !decisions || columns == 0 || rows == 0)
        return fail(CHD_ERR_ARGUMENT, "decision stage: bad argument");
    CHD_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)columns * rows;
    Scratch sc;
    chd_node_warp *src, *dst;
    float *g;
    CHD_CHECK(sc.alloc(&src, n));
    CHD_CHECK(sc.alloc(&dst, n));
    CHD_CHECK(sc.alloc(&g, n));
    hipStream_t st = c->stream;
    CHD_HIP(hipMemcpyAsync(src, decisions, n * sizeof(chd_node_warp), hipMemcpyHostToDevice, st));
    if (guidance) CHD_HIP(hipMemcpyAsync(g, guidance, n * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(launch(st, src, dst, g));
    CHD_HIP(hipMemcpyAsync(decisions, dst, n * sizeof(chd_node_warp), hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

}  // namespace
}  // extern "C++"

int chd_reject_isolated_direct_depths(chd_context *c, chd_node_depth *field,
                                      const float *guidance, uint32_t columns, uint32_t rows,
                                      double tol) {
    return field_stage(c, field, guidance, columns, rows,
                       [&](hipStream_t st, chd_node_depth *src, chd_node_depth *dst, float *g,
                           uint32_t *, uint32_t *) {
                           return launch_reject_isolated(st, src, dst, g, columns, rows, tol);
                       });
}

int chd_reject_small_direct_components(chd_context *c, chd_node_depth *field,
                                       const float *guidance, uint32_t columns, uint32_t rows,
                                       double tol, uint32_t minimum_nodes) {
    return field_stage(c, field, guidance, columns, rows,
                       [&](hipStream_t st, chd_node_depth *src, chd_node_depth *dst, float *g,
                           uint32_t *parent, uint32_t *sizes) {
                           return launch_reject_components(st, src, dst, g, columns, rows, tol,
                                                           minimum_nodes, parent, sizes);
                       });
}

int chd_reject_small_direct_components_noalloc(chd_context *c, const chd_node_depth *field,
                                               const float *guidance, uint32_t columns,
                                               uint32_t rows, double tol, uint32_t minimum_nodes,
                                               chd_node_depth *out) {
    // the size limit comes first: it needs no device
    if ((uint64_t)columns * rows > CHD_NOALLOC_MAX_NODES)
        return fail(CHD_ERR_LIMIT, "reject_small_direct_components_noalloc: more than "
                                   "CHD_NOALLOC_MAX_NODES nodes");
    if (!c || !field || !guidance || !out || columns == 0 || rows == 0)
        return fail(CHD_ERR_ARGUMENT, "reject_small_direct_components_noalloc: bad argument");
    CHD_HIP(hipSetDevice(c->device));
    uint32_t *ptr = scratch_of(c->device);
    if (!ptr) return fail(CHD_ERR_STATE, "reject_small_direct_components_noalloc: no scratch on this device");
    uint32_t *parent = ptr, *sizes = ptr + CHD_NOALLOC_MAX_NODES;
    CHD_HIP(launch_reject_components(c->stream, field, out, guidance, columns, rows, tol,
                                     minimum_nodes, parent, sizes));
    return CHD_OK;
}

int chd_fit_local_depth_planes(chd_context *c, chd_node_depth *field, const float *guidance,
                               uint32_t columns, uint32_t rows, double tol) {
    return field_stage(c, field, guidance, columns, rows,
                       [&](hipStream_t st, chd_node_depth *src, chd_node_depth *dst, float *g,
                           uint32_t *, uint32_t *) {
                           return launch_fit_planes(st, src, dst, g, columns, rows, tol);
                       });
}

int chd_enforce_warp_consensus(chd_context *c, chd_node_warp *decisions, uint32_t columns,
                               uint32_t rows) {
    return decision_stage(c, decisions, nullptr, columns, rows,
                          [&](hipStream_t st, chd_node_warp *src, chd_node_warp *dst, float *) {
                              return launch_consensus(st, src, dst, columns, rows);
                          });
}

int chd_suppress_warp_boundaries(chd_context *c, chd_node_warp *decisions, const float *guidance,
                                 uint32_t columns, uint32_t rows) {
    if (!guidance) return fail(CHD_ERR_ARGUMENT, "suppress_warp_boundaries: null guidance");
    return decision_stage(c, decisions, guidance, columns, rows,
                          [&](hipStream_t st, chd_node_warp *src, chd_node_warp *dst, float *g) {
                              return launch_boundaries(st, src, dst, g, columns, rows);
                          });
}

int chd_luminance_half_device(chd_context *c, const uint16_t *samples, uint32_t width,
                              uint32_t height, float black_q6, float white_q6,
                              int reserve_headroom, float *plane) {
    if (!c || !samples || !plane || width < 2 || height < 2)
        return fail(CHD_ERR_ARGUMENT, "luminance_half: bad argument");
    if (((uintptr_t)samples & 3u) != 0)
        return fail(CHD_ERR_ARGUMENT, "luminance_half: samples must be 4-byte aligned");
    CHD_HIP(hipSetDevice(c->device));
    CHD_CHECK(luminance_table(c, black_q6, white_q6, reserve_headroom != 0));
    CHD_HIP(launch_luminance_half(c->stream, samples, width, height, c->luminance_table,
                                  reserve_headroom != 0, plane));
    return CHD_OK;
}

int chd_luminance_half(chd_context *c, const uint16_t *samples, uint32_t width, uint32_t height,
                       float black_q6, float white_q6, int reserve_headroom, float *plane) {
    if (!c || !samples || !plane || width < 2 || height < 2)
        return fail(CHD_ERR_ARGUMENT, "luminance_half: bad argument");
    CHD_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)width * height, m = (size_t)(width / 2) * (height / 2);
    Scratch sc;
    uint16_t *d_samples;
    float *d_plane;
    CHD_CHECK(sc.alloc(&d_samples, n));
    CHD_CHECK(sc.alloc(&d_plane, m));
    CHD_HIP(hipMemcpyAsync(d_samples, samples, n * 2, hipMemcpyHostToDevice, c->stream));
    CHD_CHECK(chd_luminance_half_device(c, d_samples, width, height, black_q6, white_q6,
                                        reserve_headroom, d_plane));
    CHD_HIP(hipMemcpyAsync(plane, d_plane, m * 4, hipMemcpyDeviceToHost, c->stream));
    CHD_HIP(hipStreamSynchronize(c->stream));
    return CHD_OK;
}

}  // extern "C"

// ------------------------------------------------------------ alignment

extern "C++" {
namespace {

// Warp::map (align.rs:84-111) on host points (x, y interleaved).
bool host_warp_map(const float *points, uint32_t step, uint32_t columns, uint32_t rows, float x,
                   float y, float out[2]) {
    const float fx = x / (float)step, fy = y / (float)step;
    if (fx < 0.0f || fy < 0.0f) return false;
    auto index = [](float f, uint32_t n) -> uint32_t {
        const float fl = std::floor(f);
        uint64_t v = !(fl > 0.0f) ? 0 : fl >= 18446744073709551616.0f ? UINT64_MAX : (uint64_t)fl;
        return (uint32_t)std::min<uint64_t>(v, n - 1);
    };
    const uint32_t c0 = index(fx, columns), r0 = index(fy, rows);
    const uint32_t c1 = std::min(c0 + 1, columns - 1), r1 = std::min(r0 + 1, rows - 1);
    const float tx = fx - (float)c0, ty = fy - (float)r0;
    const float *a = points + 2 * ((size_t)r0 * columns + c0), *b = points + 2 * ((size_t)r0 * columns + c1);
    const float *c = points + 2 * ((size_t)r1 * columns + c0), *d = points + 2 * ((size_t)r1 * columns + c1);
    for (int k = 0; k < 2; k++) {
        const float top = a[k] * (1.0f - tx) + b[k] * tx;
        const float bottom = c[k] * (1.0f - tx) + d[k] * tx;
        out[k] = top * (1.0f - ty) + bottom * ty;
    }
    return !(out[0] != out[0] || out[1] != out[1]);
}

// Plane::sample (image.rs:452-466) on a host plane.
bool host_sample(const float *data, uint32_t width, uint32_t height, float x, float y, float &out) {
    if (!(x >= 0.0f && y >= 0.0f && x <= (float)(width - 1) && y <= (float)(height - 1)))
        return false;
    const uint32_t x0 = (uint32_t)std::floor(x), y0 = (uint32_t)std::floor(y);
    const uint32_t x1 = std::min(x0 + 1, width - 1), y1 = std::min(y0 + 1, height - 1);
    const float tx = x - (float)x0, ty = y - (float)y0;
    auto at = [&](uint32_t i, uint32_t j) { return data[(size_t)j * width + i]; };
    const float top = at(x0, y0) * (1.0f - tx) + at(x1, y0) * tx;
    const float bottom = at(x0, y1) * (1.0f - tx) + at(x1, y1) * tx;
    out = top * (1.0f - ty) + bottom * ty;
    return true;
}

// apply_homography (math.rs:131-138) and Mat3 products (math.rs:48-56)
bool homography_apply(const double *h, double x, double y, double &ox, double &oy) {
    const double v0 = h[0] * x + h[1] * y + h[2] * 1.0;
    const double v1 = h[3] * x + h[4] * y + h[5] * 1.0;
    const double v2 = h[6] * x + h[7] * y + h[8] * 1.0;
    if (std::fabs(v2) < 1e-12) return false;
    ox = v0 / v2;
    oy = v1 / v2;
    return true;
}

void mat3_mul(const double *a, const double *b, double *out) {
    double r[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r[i * 3 + j] = a[i * 3 + 0] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
    std::memcpy(out, r, sizeof r);
}

int64_t total_key64(double d) {
    int64_t k;
    std::memcpy(&k, &d, 8);
    return k ^ (int64_t)(((uint64_t)(k >> 63)) >> 1);
}
int32_t total_key32(float f) {
    int32_t k;
    std::memcpy(&k, &f, 4);
    return k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
}

// solve_linear (align.rs:765-793): Gaussian elimination, partial pivoting
// on the last maximal |a[i][column]|
bool solve8(double a[8][8], double b[8], double x[8]) {
    for (int column = 0; column < 8; column++) {
        int pivot = column;
        for (int i = column + 1; i < 8; i++)
            if (total_key64(std::fabs(a[i][column])) >= total_key64(std::fabs(a[pivot][column])))
                pivot = i;
        if (std::fabs(a[pivot][column]) < 1e-12) return false;
        for (int k = 0; k < 8; k++) std::swap(a[column][k], a[pivot][k]);
        std::swap(b[column], b[pivot]);
        double pivot_row[8];
        std::memcpy(pivot_row, a[column], sizeof pivot_row);
        for (int row = column + 1; row < 8; row++) {
            const double factor = a[row][column] / pivot_row[column];
            for (int k = column; k < 8; k++) a[row][k] -= factor * pivot_row[k];
            b[row] -= factor * b[column];
        }
    }
    for (int row = 7; row >= 0; row--) {
        double sum = b[row];
        for (int k = row + 1; k < 8; k++) sum -= a[row][k] * x[k];
        x[row] = sum / a[row][row];
    }
    return true;
}

struct Pair {
    double a[2], b[2];
    float score;
};

// fit_homography_least_squares (align.rs:713-762): weighted DLT on
// centroid-normalised coordinates, h33 = 1
bool homography_least_squares(const Pair *pairs, size_t count, double h[9]) {
    if (count < 4) return false;
    const double n = (double)count;
    double s[4] = {-0.0, -0.0, -0.0, -0.0};  // Iterator::sum starts from -0.0
    for (int k = 0; k < 4; k++)
        for (size_t i = 0; i < count; i++) s[k] += k < 2 ? pairs[i].a[k] : pairs[i].b[k - 2];
    const double ax = s[0] / n, ay = s[1] / n, bx = s[2] / n, by = s[3] / n;
    double sa = -0.0, sb = -0.0;
    for (size_t i = 0; i < count; i++) {
        const double dx = pairs[i].a[0] - ax, dy = pairs[i].a[1] - ay;
        sa += std::sqrt(dx * dx + dy * dy);
    }
    for (size_t i = 0; i < count; i++) {
        const double dx = pairs[i].b[0] - bx, dy = pairs[i].b[1] - by;
        sb += std::sqrt(dx * dx + dy * dy);
    }
    sa = std::fmax(sa / n, 1e-9);
    sb = std::fmax(sb / n, 1e-9);
    double ata[8][8] = {}, atb[8] = {};
    for (size_t i = 0; i < count; i++) {
        const double w = std::fmax((double)pairs[i].score, 0.05);
        const double x = (pairs[i].a[0] - ax) / sa, y = (pairs[i].a[1] - ay) / sa;
        const double u = (pairs[i].b[0] - bx) / sb, v = (pairs[i].b[1] - by) / sb;
        const double rows[2][8] = {{x, y, 1.0, 0.0, 0.0, 0.0, -u * x, -u * y},
                                   {0.0, 0.0, 0.0, x, y, 1.0, -v * x, -v * y}};
        const double rhs[2] = {u, v};
        for (int r = 0; r < 2; r++)
            for (int p = 0; p < 8; p++) {
                atb[p] += w * rows[r][p] * rhs[r];
                for (int q = 0; q < 8; q++) ata[p][q] += w * rows[r][p] * rows[r][q];
            }
    }
    double x[8];
    if (!solve8(ata, atb, x)) return false;
    const double normalized[9] = {x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], 1.0};
    const double t_a[9] = {1.0 / sa, 0.0, -ax / sa, 0.0, 1.0 / sa, -ay / sa, 0.0, 0.0, 1.0};
    const double t_b_inv[9] = {sb, 0.0, bx, 0.0, sb, by, 0.0, 0.0, 1.0};
    double tmp[9], hh[9];
    mat3_mul(t_b_inv, normalized, tmp);
    mat3_mul(tmp, t_a, hh);
    const double scale = hh[8];
    if (std::fabs(scale) < 1e-12 || !std::isfinite(scale)) return false;
    for (int k = 0; k < 9; k++) h[k] = hh[k] / scale;
    return true;
}

// Device buffers of the RANSAC consensus counts: `capacity` pairs and the
// 300 hypotheses.
constexpr int kRansacIterations = 300;
struct RansacDevice {
    hipStream_t stream = nullptr;
    double *pairs = nullptr;       // capacity x {a.x, a.y, b.x, b.y}
    double *hypotheses = nullptr;  // kRansacIterations x 9
    uint32_t *counts = nullptr;    // kRansacIterations
    size_t capacity = 0;
};

int ransac_device_alloc(Scratch &sc, hipStream_t stream, size_t capacity, RansacDevice &dev) {
    dev.stream = stream;
    dev.capacity = capacity;
    CHD_CHECK(sc.alloc(&dev.pairs, 4 * std::max<size_t>(capacity, 1)));
    CHD_CHECK(sc.alloc(&dev.hypotheses, 9 * (size_t)kRansacIterations));
    CHD_CHECK(sc.alloc(&dev.counts, (size_t)kRansacIterations));
    return CHD_OK;
}

// fit_homography_ransac (align.rs:644-711): 300 LCG-sampled 4-point fits
// (host), their consensus counts over all pairs (device), the first
// hypothesis with the most inliers, its inliers refitted by least squares
// and re-selected under the refit; *found = false is the reference's None.
int homography_ransac(const RansacDevice &dev, const std::vector<Pair> &pairs, float threshold,
                      double h[9], std::vector<float> &residuals, bool &found) {
    found = false;
    const size_t count = pairs.size();
    if (count < 6) return CHD_OK;
    if (count > dev.capacity) return fail(CHD_ERR_LIMIT, "RANSAC: more pairs than the buffers hold");
    uint64_t state = 0x9E3779B97F4A7C15ull ^ (uint64_t)count;
    std::vector<double> hypotheses(9 * (size_t)kRansacIterations, 0.0);
    std::vector<uint8_t> fitted(kRansacIterations, 0);
    for (int iteration = 0; iteration < kRansacIterations; iteration++) {
        uint32_t sample[4];
        for (int slot = 0; slot < 4; slot++)
            for (;;) {
                state = state * 6364136223846793005ull + 1442695040888963407ull;
                const uint32_t candidate = (uint32_t)((state >> 33) % (uint64_t)count);
                bool seen = false;
                for (int k = 0; k < slot; k++) seen = seen || sample[k] == candidate;
                if (!seen) {
                    sample[slot] = candidate;
                    break;
                }
            }
        Pair subset[4];
        for (int k = 0; k < 4; k++) subset[k] = pairs[sample[k]];
        fitted[iteration] = homography_least_squares(subset, 4, &hypotheses[9 * (size_t)iteration]);
    }
    std::vector<double> packed(4 * count);
    for (size_t i = 0; i < count; i++) {
        packed[4 * i] = pairs[i].a[0];
        packed[4 * i + 1] = pairs[i].a[1];
        packed[4 * i + 2] = pairs[i].b[0];
        packed[4 * i + 3] = pairs[i].b[1];
    }
    std::vector<uint32_t> counts(kRansacIterations);
    CHD_HIP(hipMemcpyAsync(dev.pairs, packed.data(), packed.size() * 8, hipMemcpyHostToDevice,
                           dev.stream));
    CHD_HIP(hipMemcpyAsync(dev.hypotheses, hypotheses.data(), hypotheses.size() * 8,
                           hipMemcpyHostToDevice, dev.stream));
    CHD_HIP(launch_ransac_count(dev.stream, dev.pairs, (uint32_t)count, dev.hypotheses,
                                kRansacIterations, (double)threshold, dev.counts));
    CHD_HIP(hipMemcpyAsync(counts.data(), dev.counts, counts.size() * 4, hipMemcpyDeviceToHost,
                           dev.stream));
    CHD_HIP(hipStreamSynchronize(dev.stream));
    // the first hypothesis whose inlier count beats every earlier one
    int best = -1;
    uint32_t best_count = 0;
    for (int k = 0; k < kRansacIterations; k++)
        if (fitted[k] && counts[k] > best_count) {
            best = k;
            best_count = counts[k];
        }
    if (best < 0 || best_count < 6) return CHD_OK;
    const double *hb = &hypotheses[9 * (size_t)best];
    std::vector<Pair> inliers;
    inliers.reserve(best_count);
    for (const Pair &p : pairs) {
        double px, py;
        if (homography_apply(hb, p.a[0], p.a[1], px, py)) {
            const double dx = px - p.b[0], dy = py - p.b[1];
            if (std::sqrt(dx * dx + dy * dy) <= (double)threshold) inliers.push_back(p);
        }
    }
    if (inliers.size() != best_count)
        return fail(CHD_ERR_STATE, "RANSAC: device and host consensus counts differ");
    double hf[9];
    if (!homography_least_squares(inliers.data(), inliers.size(), hf)) return CHD_OK;
    residuals.clear();
    for (const Pair &p : pairs) {
        double px, py;
        if (homography_apply(hf, p.a[0], p.a[1], px, py)) {
            const double dx = px - p.b[0], dy = py - p.b[1];
            const float r = (float)std::sqrt(dx * dx + dy * dy);
            if (r <= threshold) residuals.push_back(r);
        }
    }
    if (residuals.size() < 6) return CHD_OK;
    std::stable_sort(residuals.begin(), residuals.end(),
                     [](float x, float y) { return total_key32(x) < total_key32(y); });
    std::memcpy(h, hf, sizeof hf);
    found = true;
    return CHD_OK;
}

// photometric_gain (align.rs:795-824): median linear-luminance ratio of
// reference / target at every 16th plane pixel. The texel pairs are sampled
// on the device from the level-0 planes; the ratios use the platform expf.
struct DevicePlane;
int photometric_gain(chd_context *c, Scratch &sc, const DevicePlane &reference,
                     const DevicePlane &target, const chd_warp &warp, float &gain);

// A padded device plane (pitch width + 1, height + 1 rows).
struct DevicePlane {
    float *data = nullptr;
    uint32_t width = 0, height = 0;
    DevPlane view() const {
        DevPlane p;
        p.data = data;
        p.width = width;
        p.height = height;
        p.pitch = width + 1;
        p.lines = height + 1;
        p.transposed = 0;
        return p;
    }
};

// A plane of pitch `width` in host or device memory into a padded device
// plane: one contiguous copy (its direction comes from the pointer; a
// pitched copy from pageable host memory goes row by row and is many times
// slower) and a device-side repack that writes the padding too.
int upload_padded(hipStream_t st, Scratch &sc, const float *src, uint32_t width, uint32_t height,
                  float *dst) {
    float *staging;
    CHD_CHECK(sc.alloc(&staging, (size_t)width * height));
    CHD_HIP(hipMemcpyAsync(staging, src, (size_t)width * height * 4, hipMemcpyDefault, st));
    CHD_HIP(launch_pitch_copy(st, staging, width, height, dst));
    return CHD_OK;
}

// A view's luminance as a padded device plane; the source may be host or
// device memory.
int device_plane(chd_context *c, Scratch &sc, const chd_view &v, DevicePlane &out) {
    hipStream_t st = c->stream;
    out.width = v.plane_width;
    out.height = v.plane_height;
    CHD_CHECK(sc.alloc(&out.data, (size_t)(out.width + 1) * (out.height + 1)));
    CHD_CHECK(upload_padded(st, sc, v.luminance, out.width, out.height, out.data));
    return CHD_OK;
}

// Plane::pyramid (image.rs:487-498) on the device from a padded base plane.
int device_pyramid(chd_context *c, Scratch &sc, const DevicePlane &base, uint32_t min_size,
                   std::vector<DevicePlane> &levels) {
    hipStream_t st = c->stream;
    levels.push_back(base);
    const uint32_t floor_size = std::max(min_size, 8u);
    while (std::min(levels.back().width, levels.back().height) / 2 >= floor_size) {
        const DevicePlane &last = levels.back();
        DevicePlane next;
        next.width = std::max(last.width / 2, 1u);
        next.height = std::max(last.height / 2, 1u);
        CHD_CHECK(sc.alloc(&next.data, (size_t)(next.width + 1) * (next.height + 1)));
        CHD_HIP(launch_downsample(st, last.data, last.width, last.height, last.width + 1, next.data,
                                  next.width, next.height, next.width + 1));
        CHD_HIP(launch_pad_plane(st, next.data, next.width, next.height));
        levels.push_back(next);
    }
    return CHD_OK;
}

int photometric_gain(chd_context *c, Scratch &sc, const DevicePlane &reference,
                     const DevicePlane &target, const chd_warp &warp, float &gain) {
    hipStream_t st = c->stream;
    gain = 1.0f;
    uint32_t gx = 0, gy = 0;
    while (16 * (gx + 1) + 16 < reference.width) gx++;
    while (16 * (gy + 1) + 16 < reference.height) gy++;
    const size_t n = (size_t)gx * gy, nodes = (size_t)warp.columns * warp.rows;
    if (n == 0) return CHD_OK;
    float2 *d_samples, *d_points;
    uint8_t *d_ok;
    CHD_CHECK(sc.alloc(&d_samples, n));
    CHD_CHECK(sc.alloc(&d_ok, n));
    CHD_CHECK(sc.alloc(&d_points, nodes));
    CHD_HIP(hipMemcpyAsync(d_points, warp.points, nodes * 8, hipMemcpyHostToDevice, st));
    DevWarp dw;
    dw.points = d_points;
    dw.step = warp.step;
    dw.columns = warp.columns;
    dw.rows = warp.rows;
    dw.inverse_step = (warp.step & (warp.step - 1)) == 0 ? 1.0f / (float)warp.step : 0.0f;
    CHD_HIP(launch_gain_samples(st, reference.view(), target.view(), dw, gx, gy, d_samples, d_ok));
    std::vector<float> samples(2 * n);
    std::vector<uint8_t> ok(n);
    CHD_HIP(hipMemcpyAsync(samples.data(), d_samples, n * 8, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipMemcpyAsync(ok.data(), d_ok, n, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    std::vector<float> ratios;
    for (size_t k = 0; k < n; k++) {  // rows of x, as the reference walks them
        if (!ok[k]) continue;
        const float r = samples[2 * k], t = samples[2 * k + 1];
        const float lr = (expf(r) - 1.0f) / 1000.0f, lt = (expf(t) - 1.0f) / 1000.0f;
        if (lr > 0.01f && lt > 0.01f && lr < 0.95f && lt < 0.95f) ratios.push_back(lr / lt);
    }
    if (ratios.size() < 16) return CHD_OK;
    std::stable_sort(ratios.begin(), ratios.end(),
                     [](float x, float y) { return total_key32(x) < total_key32(y); });
    gain = std::min(std::max(ratios[ratios.size() / 2], 0.25f), 4.0f);
    return CHD_OK;
}

// Warp::from_fn (align.rs:50-80) of initial(correction(p)), or the identity.
void fill_warp(uint32_t width, uint32_t height, uint32_t step, const std::vector<float> *grid,
               uint32_t gc, uint32_t gr, const double *correction, chd_warp &w) {
    w.step = step;
    w.columns = (uint32_t)div_ceil(width, step) + 1;
    w.rows = (uint32_t)div_ceil(height, step) + 1;
    for (uint32_t r = 0; r < w.rows; r++)
        for (uint32_t col = 0; col < w.columns; col++) {
            const size_t i = (size_t)r * w.columns + col;
            const double px = (double)((size_t)col * step), py = (double)((size_t)r * step);
            double q[2] = {px, py};
            bool ok = true;
            if (grid) {
                double cx, cy;
                float g[2];
                ok = homography_apply(correction, px, py, cx, cy) &&
                     host_warp_map(grid->data(), 8, gc, gr, (float)cx, (float)cy, g);
                q[0] = g[0];
                q[1] = g[1];
            }
            w.points[2 * i] = ok ? (float)q[0] : NAN;
            w.points[2 * i + 1] = ok ? (float)q[1] : NAN;
            if (w.confidence) w.confidence[i] = ok ? 1.0f : 0.0f;
        }
}

}  // namespace
}  // extern "C++"

extern "C" {

void chd_default_align_options(chd_align_options *o) {  // align.rs:232-244
    if (!o) return;
    o->grid_step = 32;
    o->patch = 32;
    o->coarse_radius = 12;
    o->min_score = 0.5f;
    o->inlier_px = 3.0f;
    o->min_inlier_ratio = 0.45f;
    o->refine = 1;
}

int chd_match_patches(chd_context *c, const float *reference, const float *target, uint32_t width,
                      uint32_t height, const uint32_t *rx, const uint32_t *ry, uint32_t count,
                      uint32_t size, uint32_t radius, float *shift, float *score, uint8_t *valid) {
    if (!c || !reference || !target || !rx || !ry || !shift || !score || !valid || size == 0)
        return fail(CHD_ERR_ARGUMENT, "match_patches: bad argument");
    if (size > 64 || radius > 32) return fail(CHD_ERR_LIMIT, "match_patches: size above 64 or radius above 32");
    if (count == 0) return CHD_OK;
    CHD_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const size_t n = (size_t)width * height;
    Scratch sc;
    float *d_ref, *d_tgt, *d_shift, *d_score;
    uint32_t *d_rx, *d_ry;
    uint8_t *d_valid;
    CHD_CHECK(sc.alloc(&d_ref, n));
    CHD_CHECK(sc.alloc(&d_tgt, n));
    CHD_CHECK(sc.alloc(&d_rx, count));
    CHD_CHECK(sc.alloc(&d_ry, count));
    CHD_CHECK(sc.alloc(&d_shift, 2 * (size_t)count));
    CHD_CHECK(sc.alloc(&d_score, count));
    CHD_CHECK(sc.alloc(&d_valid, count));
    CHD_HIP(hipMemcpyAsync(d_ref, reference, n * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(d_tgt, target, n * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(d_rx, rx, count * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(d_ry, ry, count * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(launch_match_patches(st, d_ref, d_tgt, width, height, d_rx, d_ry, count, size, radius,
                                 kCoverNone, d_shift, d_score, d_valid));
    CHD_HIP(hipMemcpyAsync(shift, d_shift, 8 * (size_t)count, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipMemcpyAsync(score, d_score, 4 * (size_t)count, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipMemcpyAsync(valid, d_valid, count, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

int chd_align_module(chd_context *c, const chd_view *reference, const chd_view *target,
                     const chd_align_options *o, chd_warp *warp, chd_alignment_report *report) {
    if (!c || !reference || !target || !o || !warp || !report || !warp->points)
        return fail(CHD_ERR_ARGUMENT, "align_module: bad argument");
    if (o->grid_step == 0 || !reference->luminance || !target->luminance)
        return fail(CHD_ERR_ARGUMENT, "align_module: zero grid step or missing plane");
    if (o->patch > 64 || o->coarse_radius > 32)
        return fail(CHD_ERR_LIMIT, "align_module: patch above 64 or coarse radius above 32");
    std::memset(report, 0, sizeof *report);
    const uint32_t width = reference->width, height = reference->height;
    if (reference == target) {  // align.rs:291-309
        fill_warp(width, height, o->grid_step, nullptr, 0, 0, nullptr, *warp);
        report->initialised_from = CHD_ALIGN_FROM_REFERENCE;
        report->coverage = 1.0f;
        report->inlier_ratio = 1.0f;
        report->accepted = 1;
        report->status = CHD_ALIGN_STATUS_REFERENCE;
        report->gain = 1.0f;
        return CHD_OK;
    }
    if (!reference->camera || !target->camera)
        return fail(CHD_ERR_ARGUMENT, "align_module: both modules need cameras");
    CHD_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    report->initialised_from = CHD_ALIGN_FROM_CALIBRATION;
    Scratch sc;
    // the factory model on an 8 px grid (align.rs:317-324)
    const uint32_t gc = (uint32_t)div_ceil(width, 8) + 1, gr = (uint32_t)div_ceil(height, 8) + 1;
    chd_camera *cams;
    float2 *d_grid;
    CHD_CHECK(sc.alloc(&cams, 2));
    CHD_CHECK(sc.alloc(&d_grid, (size_t)gc * gr));
    CHD_HIP(hipMemcpyAsync(cams, reference->camera, sizeof(chd_camera), hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(cams + 1, target->camera, sizeof(chd_camera), hipMemcpyHostToDevice, st));
    CHD_HIP(launch_initial_grid(st, cams, cams + 1, gc, gr, 8, d_grid));
    std::vector<float> grid((size_t)gc * gr * 2);
    CHD_HIP(hipMemcpyAsync(grid.data(), d_grid, grid.size() * 4, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    auto initial = [&](double x, double y, double out[2]) {
        float q[2];
        if (!host_warp_map(grid.data(), 8, gc, gr, (float)x, (float)y, q)) return false;
        out[0] = q[0];
        out[1] = q[1];
        return true;
    };
    double correction[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::vector<float> finest;
    DevicePlane reference0, target0;
    CHD_CHECK(device_plane(c, sc, *reference, reference0));
    CHD_CHECK(device_plane(c, sc, *target, target0));
    if (o->refine) {
        std::vector<DevicePlane> rp, tp;
        CHD_CHECK(device_pyramid(c, sc, reference0, 96, rp));
        CHD_CHECK(device_pyramid(c, sc, target0, 48, tp));
        double magnification;
        {
            const double cx = (double)width / 2.0, cy = (double)height / 2.0, h = 64.0;
            double a[2], b[2], d[2];
            if (initial(cx, cy, a) && initial(cx + h, cy, b) && initial(cx, cy + h, d)) {
                const double dx = std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1])) / h;
                const double dy = std::sqrt((d[0] - a[0]) * (d[0] - a[0]) + (d[1] - a[1]) * (d[1] - a[1])) / h;
                magnification = std::sqrt(dx * dy);
            } else {
                magnification = target->camera->focal_px / reference->camera->focal_px;
            }
        }
        DevWarp dw;
        dw.points = d_grid;
        dw.step = 8;
        dw.columns = gc;
        dw.rows = gr;
        dw.inverse_step = 0.125f;
        float *rendered, *d_shift, *d_score;
        uint32_t *d_rx, *d_ry;
        uint8_t *d_valid;
        const size_t finest_points = (size_t)rp[0].width * rp[0].height;
        CHD_CHECK(sc.alloc(&rendered, finest_points));
        // window capacity: every level has at most the finest level's windows
        const size_t capacity = finest_points / 16 + 64;
        CHD_CHECK(sc.alloc(&d_rx, capacity));
        CHD_CHECK(sc.alloc(&d_ry, capacity));
        CHD_CHECK(sc.alloc(&d_shift, 2 * capacity));
        CHD_CHECK(sc.alloc(&d_score, capacity));
        CHD_CHECK(sc.alloc(&d_valid, capacity));
        RansacDevice ransac;
        CHD_CHECK(ransac_device_alloc(sc, st, capacity, ransac));
        float *unpadded;
        CHD_CHECK(sc.alloc(&unpadded, finest_points));
        for (int level = (int)rp.size() - 1; level >= 0; level--) {
            const uint32_t scale = 1u << level;
            const DevicePlane &ref_plane = rp[(size_t)level];
            const double l2 = std::fmax(std::round(std::log2((double)scale * magnification)), 0.0);
            const uint32_t last = (uint32_t)tp.size() - 1;
            const uint32_t target_level = l2 >= (double)last ? last : (uint32_t)l2;
            const double target_scale = (double)(1u << target_level);
            const uint32_t W = ref_plane.width, H = ref_plane.height;
            CHD_HIP(launch_render_through(st, W, H, scale, tp[target_level].view(), target_scale, dw,
                                          correction, rendered));
            // the reference level without its padding column
            CHD_HIP(hipMemcpy2DAsync(unpadded, (size_t)W * 4, ref_plane.data, (size_t)(W + 1) * 4,
                                     (size_t)W * 4, H, hipMemcpyDeviceToDevice, st));
            const uint32_t radius = level + 1 == (int)rp.size() ? o->coarse_radius : 3;
            const uint32_t patch = std::max(std::min(o->patch, W / 4), 8u);
            const uint32_t stride = patch / 2;
            std::vector<uint32_t> xs, ys;
            for (uint32_t y = radius; y + patch + radius <= H; y += stride)
                for (uint32_t x = radius; x + patch + radius <= W; x += stride) {
                    xs.push_back(x);
                    ys.push_back(y);
                }
            const uint32_t count = (uint32_t)xs.size();
            if (count > capacity) return fail(CHD_ERR_LIMIT, "align_module: too many windows");
            std::vector<float> shifts(2 * (size_t)count), scores(count);
            std::vector<uint8_t> valid(count);
            if (count > 0) {
                CHD_HIP(hipMemcpyAsync(d_rx, xs.data(), count * 4, hipMemcpyHostToDevice, st));
                CHD_HIP(hipMemcpyAsync(d_ry, ys.data(), count * 4, hipMemcpyHostToDevice, st));
                CHD_HIP(launch_match_patches(st, unpadded, rendered, W, H, d_rx, d_ry, count, patch,
                                             radius, kCoverWindowNan, d_shift, d_score, d_valid));
                CHD_HIP(hipMemcpyAsync(shifts.data(), d_shift, 8 * (size_t)count,
                                       hipMemcpyDeviceToHost, st));
                CHD_HIP(hipMemcpyAsync(scores.data(), d_score, 4 * (size_t)count,
                                       hipMemcpyDeviceToHost, st));
                CHD_HIP(hipMemcpyAsync(valid.data(), d_valid, count, hipMemcpyDeviceToHost, st));
            }
            CHD_HIP(hipStreamSynchronize(st));
            std::vector<Pair> pairs;
            const double s = (double)scale;
            for (uint32_t k = 0; k < count; k++) {
                if (!valid[k] || !(scores[k] >= o->min_score)) continue;
                const float cx = (float)xs[k] + (float)patch / 2.0f, cy = (float)ys[k] + (float)patch / 2.0f;
                const float sx = cx + shifts[2 * k], sy = cy + shifts[2 * k + 1];
                Pair p;
                p.a[0] = (double)cx * s * 2.0 + s - 0.5;
                p.a[1] = (double)cy * s * 2.0 + s - 0.5;
                p.b[0] = (double)sx * s * 2.0 + s - 0.5;
                p.b[1] = (double)sy * s * 2.0 + s - 0.5;
                p.score = scores[k];
                pairs.push_back(p);
            }
            float threshold = o->inlier_px * (float)scale * 2.0f;
            if (!(threshold >= o->inlier_px)) threshold = o->inlier_px;
            double update[9];
            std::vector<float> residuals;
            chd_level_report *lr = report->level_count < CHD_MAX_ALIGN_LEVELS
                                       ? &report->levels[report->level_count++]
                                       : nullptr;
            bool fitted = false;
            CHD_CHECK(homography_ransac(ransac, pairs, threshold, update, residuals, fitted));
            if (fitted) {
                mat3_mul(correction, update, correction);
                if (lr) {
                    lr->scale = scale * 2;
                    lr->patches = (uint32_t)pairs.size();
                    lr->inliers = (uint32_t)residuals.size();
                    lr->median_residual_px = residuals[residuals.size() / 2];
                }
                if (level == 0) {
                    finest = residuals;
                    report->patches = (uint32_t)pairs.size();
                    report->inliers = (uint32_t)residuals.size();
                }
            } else if (lr) {
                lr->scale = scale * 2;
                lr->patches = (uint32_t)pairs.size();
                lr->inliers = 0;
                lr->median_residual_px = NAN;
            }
        }
    }
    fill_warp(width, height, o->grid_step, &grid, gc, gr, correction, *warp);
    size_t covered = 0, total = (size_t)warp->columns * warp->rows;
    for (size_t i = 0; i < total; i++) {
        const float x = warp->points[2 * i], y = warp->points[2 * i + 1];
        if (std::isfinite(x) && x >= 0.0f && y >= 0.0f && x <= (float)(target->width - 1) &&
            y <= (float)(target->height - 1))
            covered++;
    }
    report->coverage = (float)covered / (float)total;
    if (!finest.empty()) {
        report->residual_median_px = finest[finest.size() / 2];
        report->residual_p90_px = finest[std::min(finest.size() * 9 / 10, finest.size() - 1)];
    }
    {
        const double cx = (double)(width / 2), cy = (double)(height / 2);
        double mx, my;
        if (!homography_apply(correction, cx, cy, mx, my)) {
            mx = cx;
            my = cy;
        }
        report->correction_median_px[0] = (float)(mx - cx);
        report->correction_median_px[1] = (float)(my - cy);
    }
    report->inlier_ratio =
        report->patches == 0 ? 0.0f : (float)report->inliers / (float)report->patches;
    report->accepted = !o->refine || report->inlier_ratio >= o->min_inlier_ratio;
    if (!report->accepted) report->status = CHD_ALIGN_STATUS_REJECTED;
    else if (!o->refine) report->status = CHD_ALIGN_STATUS_FACTORY_MODEL;
    else if (report->inliers >= 12 && report->residual_median_px < o->inlier_px)
        report->status = CHD_ALIGN_STATUS_REFINED;
    else if (report->inliers > 0) report->status = CHD_ALIGN_STATUS_WEAK;
    else report->status = CHD_ALIGN_STATUS_NO_SUPPORT;
    CHD_CHECK(photometric_gain(c, sc, reference0, target0, *warp, report->gain));
    return CHD_OK;
}

}  // extern "C"

// --------------------------------------------------- refine_resolution_warp

extern "C++" {
namespace {

constexpr uint32_t kLocalWarpStep = 32;     // LOCAL_WARP_STEP
constexpr uint32_t kLocalPatch = 12;        // LOCAL_MATCH_PATCH
constexpr uint32_t kLocalRadius = 2;        // LOCAL_MATCH_RADIUS
constexpr uint32_t kMagnifiedRadius = 5;    // MAGNIFIED_MATCH_RADIUS
constexpr float kLocalMinimumScore = 0.72f; // LOCAL_MINIMUM_SCORE

// Warp::confidence (align.rs:112-131); 0 without a confidence array.
float host_warp_confidence(const chd_warp &w, float x, float y) {
    const float fx = x / (float)w.step, fy = y / (float)w.step;
    if (fx < 0.0f || fy < 0.0f || !w.confidence) return 0.0f;
    auto index = [](float f, uint32_t n) -> uint32_t {
        const float fl = std::floor(f);
        uint64_t v = !(fl > 0.0f) ? 0 : fl >= 18446744073709551616.0f ? UINT64_MAX : (uint64_t)fl;
        return (uint32_t)std::min<uint64_t>(v, n - 1);
    };
    const uint32_t c0 = index(fx, w.columns), r0 = index(fy, w.rows);
    const uint32_t c1 = std::min(c0 + 1, w.columns - 1), r1 = std::min(r0 + 1, w.rows - 1);
    const float tx = fx - (float)c0, ty = fy - (float)r0;
    auto at = [&](uint32_t col, uint32_t row) { return w.confidence[(size_t)row * w.columns + col]; };
    const float top = at(c0, r0) * (1.0f - tx) + at(c1, r0) * tx;
    const float bottom = at(c0, r1) * (1.0f - tx) + at(c1, r1) * tx;
    const float v = top * (1.0f - ty) + bottom * ty;
    return v < 0.0f ? 0.0f : v > 1.0f ? 1.0f : v;  // f32::clamp keeps NaN
}

float clamp_keep_nan(float v, float lo, float hi) { return v < lo ? lo : v > hi ? hi : v; }

// smoothstep (resolution.rs:486-489)
float smoothstep01(float v) {
    v = clamp_keep_nan(v, 0.0f, 1.0f);
    return v * v * (3.0f - 2.0f * v);
}

// The medians of the x and y components of a 3x3 neighbourhood's entries
// (sort_by(total_cmp), element len / 2), sorted lists kept for callers.
struct Neighbourhood {
    float xs[9], ys[9];
    uint32_t n = 0;
    void push(float x, float y) {
        xs[n] = x;
        ys[n] = y;
        n++;
    }
    void sort() {
        auto by_key = [](float a, float b) { return total_key32(a) < total_key32(b); };
        std::sort(xs, xs + n, by_key);
        std::sort(ys, ys + n, by_key);
    }
};

}  // namespace
}  // extern "C++"

extern "C" {

int chd_refine_resolution_warp(chd_context *c, const float *reference, uint32_t rw, uint32_t rh,
                               const float *target, uint32_t tw, uint32_t th, const chd_warp *base,
                               uint32_t width, uint32_t height, chd_warp *out,
                               chd_resolution_report *report) {
    if (!c || !reference || !target || !base || !base->points || !out || !out->points ||
        !out->confidence || !report)
        return fail(CHD_ERR_ARGUMENT, "refine_resolution_warp: bad argument");
    if (base->step == 0 || base->columns == 0 || base->rows == 0 || rw == 0 || rh == 0 ||
        tw == 0 || th == 0)
        return fail(CHD_ERR_ARGUMENT, "refine_resolution_warp: empty plane or base warp");
    CHD_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const chd_warp &bw = *base;
    auto base_map = [&](float x, float y, float q[2]) {
        return host_warp_map(bw.points, bw.step, bw.columns, bw.rows, x, y, q);
    };
    // Warp::magnification at the raster centre (align.rs:135-143), default 1
    float magnification = 1.0f;
    {
        const float cx = (float)width * 0.5f, cy = (float)height * 0.5f, h = (float)bw.step;
        float a[2], b[2], d[2];
        if (base_map(cx, cy, a) && base_map(cx + h, cy, b) && base_map(cx, cy + h, d)) {
            const float dx = std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1])) / h;
            const float dy = std::sqrt((d[0] - a[0]) * (d[0] - a[0]) + (d[1] - a[1]) * (d[1] - a[1])) / h;
            magnification = std::fmax(std::sqrt(dx * dy), 1e-6f);
        }
    }
    magnification = clamp_keep_nan(magnification, 0.5f, 4.0f);
    const uint32_t radius = magnification > 1.35f ? kMagnifiedRadius : kLocalRadius;
    // the render footprint (resolution.rs:441-465): bilinear up to 1.15x,
    // else a Gaussian of sigma max(0.42 m, 0.55) over ceil(2 sigma) taps
    BandwidthKernel footprint{};
    if (magnification > 1.15f) {
        const float sigma = std::fmax(0.42f * magnification, 0.55f);
        const uint32_t taps = (uint32_t)clamp_keep_nan(std::ceil(2.0f * sigma), 1.0f, 4.0f);
        footprint.taps = taps;
        uint32_t k = 0;
        for (int dy = -(int)taps; dy <= (int)taps; dy++)
            for (int dx = -(int)taps; dx <= (int)taps; dx++)
                footprint.weight[k++] = std::exp(-0.5f * (float)(dx * dx + dy * dy) / (sigma * sigma));
    }

    Scratch sc;
    float *d_ref, *d_rendered, *d_shift, *d_score;
    float2 *d_base;
    uint32_t *d_rx, *d_ry;
    uint8_t *d_valid;
    DevicePlane tp;
    tp.width = tw;
    tp.height = th;
    const size_t plane_points = (size_t)rw * rh, base_nodes = (size_t)bw.columns * bw.rows;
    CHD_CHECK(sc.alloc(&d_ref, plane_points));
    CHD_CHECK(sc.alloc(&d_rendered, plane_points));
    CHD_CHECK(sc.alloc(&tp.data, (size_t)(tw + 1) * (th + 1)));
    CHD_CHECK(sc.alloc(&d_base, base_nodes));
    // the planes may be host or device memory: the copies' direction comes
    // from the pointers
    CHD_HIP(hipMemcpyAsync(d_ref, reference, plane_points * 4, hipMemcpyDefault, st));
    CHD_CHECK(upload_padded(st, sc, target, tw, th, tp.data));
    CHD_HIP(hipMemcpyAsync(d_base, bw.points, base_nodes * 8, hipMemcpyHostToDevice, st));
    DevWarp dw;
    dw.points = d_base;
    dw.step = bw.step;
    dw.columns = bw.columns;
    dw.rows = bw.rows;
    dw.inverse_step = (bw.step & (bw.step - 1)) == 0 ? 1.0f / (float)bw.step : 0.0f;
    CHD_HIP(launch_render_bandwidth(st, rw, rh, tp.view(), dw, footprint, d_rendered));

    // one 12 x 12 window per node of the 32 px grid whose search stays
    // inside the plane (resolution.rs:272-301)
    const uint32_t columns = (uint32_t)div_ceil(width, kLocalWarpStep) + 1;
    const uint32_t rows = (uint32_t)div_ceil(height, kLocalWarpStep) + 1;
    const size_t nodes = (size_t)columns * rows;
    std::vector<uint32_t> xs, ys, node_of;
    for (uint32_t row = 0; row < rows; row++)
        for (uint32_t column = 0; column < columns; column++) {
            const float rx = (float)(column * kLocalWarpStep), ry = (float)(row * kLocalWarpStep);
            if (rx >= (float)width || ry >= (float)height) continue;
            const int64_t x = (int64_t)std::round((rx - 0.5f) * 0.5f) - (int64_t)kLocalPatch / 2;
            const int64_t y = (int64_t)std::round((ry - 0.5f) * 0.5f) - (int64_t)kLocalPatch / 2;
            if (x < (int64_t)radius || y < (int64_t)radius ||
                x + kLocalPatch + radius > (int64_t)rw || y + kLocalPatch + radius > (int64_t)rh)
                continue;
            xs.push_back((uint32_t)x);
            ys.push_back((uint32_t)y);
            node_of.push_back(row * columns + column);
        }
    const uint32_t count = (uint32_t)xs.size();
    std::vector<float> shifts(2 * (size_t)count), scores(count);
    std::vector<uint8_t> valid(count);
    CHD_CHECK(sc.alloc(&d_rx, count));
    CHD_CHECK(sc.alloc(&d_ry, count));
    CHD_CHECK(sc.alloc(&d_shift, 2 * (size_t)count));
    CHD_CHECK(sc.alloc(&d_score, count));
    CHD_CHECK(sc.alloc(&d_valid, count));
    if (count > 0) {
        CHD_HIP(hipMemcpyAsync(d_rx, xs.data(), count * 4, hipMemcpyHostToDevice, st));
        CHD_HIP(hipMemcpyAsync(d_ry, ys.data(), count * 4, hipMemcpyHostToDevice, st));
        CHD_HIP(launch_match_patches(st, d_ref, d_rendered, rw, rh, d_rx, d_ry, count, kLocalPatch,
                                     radius, kCoverSearchFinite, d_shift, d_score, d_valid));
        CHD_HIP(hipMemcpyAsync(shifts.data(), d_shift, 8 * (size_t)count, hipMemcpyDeviceToHost, st));
        CHD_HIP(hipMemcpyAsync(scores.data(), d_score, 4 * (size_t)count, hipMemcpyDeviceToHost, st));
        CHD_HIP(hipMemcpyAsync(valid.data(), d_valid, count, hipMemcpyDeviceToHost, st));
    }
    CHD_HIP(hipStreamSynchronize(st));

    // measured shifts in raster px (resolution.rs:295-300)
    std::vector<uint8_t> measured(nodes, 0);
    std::vector<float> mshift(2 * nodes, 0.0f), mscore(nodes, 0.0f);
    for (uint32_t k = 0; k < count; k++)
        if (valid[k] && scores[k] >= kLocalMinimumScore) {
            const uint32_t i = node_of[k];
            measured[i] = 1;
            mshift[2 * i] = shifts[2 * k] * 2.0f;
            mshift[2 * i + 1] = shifts[2 * k + 1] * 2.0f;
            mscore[i] = scores[k];
        }
    auto gather = [&](uint32_t row, uint32_t column, const std::vector<uint8_t> &have,
                      const std::vector<float> &vec, Neighbourhood &nb,
                      const std::vector<float> *conf, float *conf_sum) {
        const uint32_t r0 = row > 0 ? row - 1 : 0, r1 = std::min(row + 1, rows - 1);
        const uint32_t c0 = column > 0 ? column - 1 : 0, c1 = std::min(column + 1, columns - 1);
        for (uint32_t r = r0; r <= r1; r++)
            for (uint32_t cc = c0; cc <= c1; cc++) {
                const size_t j = (size_t)r * columns + cc;
                if (!have[j]) continue;
                nb.push(vec[2 * j], vec[2 * j + 1]);
                if (conf) *conf_sum += (*conf)[j];
            }
    };

    // neighbourhood-median verification (resolution.rs:303-360)
    out->step = kLocalWarpStep;
    out->columns = columns;
    out->rows = rows;
    std::vector<uint8_t> corrected(nodes, 0);
    std::vector<float> correction(2 * nodes, 0.0f), corrections;
    float confidence_sum = 0.0f;
    for (uint32_t row = 0; row < rows; row++)
        for (uint32_t column = 0; column < columns; column++) {
            const size_t i = (size_t)row * columns + column;
            const float rx = (float)(column * kLocalWarpStep), ry = (float)(row * kLocalWarpStep);
            bool refined = false;
            float point[2], local = 0.0f, median[2];
            if (measured[i]) {
                Neighbourhood nb;
                gather(row, column, measured, mshift, nb, nullptr, nullptr);
                if (nb.n >= 3) {
                    nb.sort();
                    median[0] = nb.xs[nb.n / 2];
                    median[1] = nb.ys[nb.n / 2];
                    if (!(std::hypot(mshift[2 * i] - median[0], mshift[2 * i + 1] - median[1]) > 1.5f) &&
                        base_map(rx + median[0], ry + median[1], point)) {
                        local = smoothstep01((mscore[i] - kLocalMinimumScore) / 0.20f) *
                                host_warp_confidence(bw, rx, ry);
                        refined = true;
                    }
                }
            }
            if (refined) {
                out->points[2 * i] = point[0];
                out->points[2 * i + 1] = point[1];
                out->confidence[i] = local;
                corrected[i] = 1;
                correction[2 * i] = median[0];
                correction[2 * i + 1] = median[1];
                corrections.push_back(std::hypot(median[0], median[1]));
                confidence_sum += local;
            } else {
                float q[2];
                if (!base_map(rx, ry, q)) q[0] = q[1] = NAN;
                out->points[2 * i] = q[0];
                out->points[2 * i + 1] = q[1];
                out->confidence[i] = 0.0f;
            }
        }
    // two completion passes over snapshots (resolution.rs:362-412)
    for (int pass = 0; pass < 2; pass++) {
        const std::vector<uint8_t> have = corrected;
        const std::vector<float> snapshot = correction;
        const std::vector<float> conf(out->confidence, out->confidence + nodes);
        for (uint32_t row = 0; row < rows; row++)
            for (uint32_t column = 0; column < columns; column++) {
                const size_t i = (size_t)row * columns + column;
                if (have[i]) continue;
                Neighbourhood nb;
                float csum = 0.0f;
                gather(row, column, have, snapshot, nb, &conf, &csum);
                if (nb.n < 3) continue;
                nb.sort();
                const float cx = nb.xs[nb.n / 2], cy = nb.ys[nb.n / 2];
                // the reference pairs the two independently sorted lists
                bool coherent = true;
                for (uint32_t k = 0; k < nb.n; k++)
                    coherent &= std::hypot(nb.xs[k] - cx, nb.ys[k] - cy) <= 1.5f;
                if (!coherent) continue;
                float p[2];
                if (!base_map((float)(column * kLocalWarpStep) + cx,
                              (float)(row * kLocalWarpStep) + cy, p))
                    continue;
                out->points[2 * i] = p[0];
                out->points[2 * i + 1] = p[1];
                out->confidence[i] = 0.45f * csum / (float)nb.n;
                corrected[i] = 1;
                correction[2 * i] = cx;
                correction[2 * i + 1] = cy;
            }
    }
    std::sort(corrections.begin(), corrections.end(),
              [](float a, float b) { return total_key32(a) < total_key32(b); });
    size_t supported = 0;
    for (size_t i = 0; i < nodes; i++) supported += out->confidence[i] > 0.0f;
    const float denominator = (float)std::max<size_t>(nodes, 1);
    report->verified_fraction = (float)corrections.size() / denominator;
    report->supported_fraction = (float)supported / denominator;
    report->median_correction_px = corrections.empty() ? 0.0f : corrections[corrections.size() / 2];
    report->mean_confidence = confidence_sum / (float)std::max<size_t>(corrections.size(), 1);
    return CHD_OK;
}

}  // extern "C"

// ------------------------------------------------ synthesis depth/warp stage

extern "C++" {
namespace {

// `as usize` of an f32 (NaN and negatives to 0, saturating)
uint64_t usize_of(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)f;
}

// band_touches_module (synth.rs:1688-1721) for canvas rows start..end
bool band_touches(const chd_synth_source &s, uint64_t start, uint64_t end, const chd_canvas &cv) {
    const chd_warp &w = s.warp;
    const float margin = (float)((uint64_t)w.step * 2) * std::fmax(s.magnification, 1.0f);
    const uint64_t rows[3] = {start, (start + end) / 2, end > 0 ? end - 1 : 0};
    for (uint64_t v : rows) {
        const float ry = cv.crop_y + ((float)v + 0.5f) / cv.scale - 0.5f;
        const uint64_t samples = std::max<uint64_t>(usize_of(cv.crop_width / (float)w.step), 8) + 1;
        for (uint64_t i = 0; i <= samples; i++) {
            const float rx = cv.crop_x + ((float)i / (float)samples) * (cv.crop_width - 1.0f);
            float q[2];
            if (host_warp_map(w.points, w.step, w.columns, w.rows, rx, ry, q) && q[0] >= -margin &&
                q[1] >= -margin && q[0] <= (float)s.mosaic_width + margin &&
                q[1] <= (float)s.mosaic_height + margin)
                return true;
        }
    }
    return false;
}

}  // namespace
}  // extern "C++"

extern "C" {

int chd_canvas_size(const chd_canvas *cv, uint32_t *width, uint32_t *height) {
    if (!cv || !width || !height) return fail(CHD_ERR_ARGUMENT, "canvas_size: null argument");
    // synth.rs:611-620
    if (cv->scale != cv->scale || cv->scale <= 0.0f || cv->crop_width < 1.0f ||
        cv->crop_height < 1.0f)
        return fail(CHD_ERR_ARGUMENT, "invalid canvas geometry");
    const uint64_t w = usize_of(std::fmax(std::round(cv->crop_width * cv->scale), 1.0f));
    const uint64_t h = usize_of(std::fmax(std::round(cv->crop_height * cv->scale), 1.0f));
    if (w > (1u << 20) || h > (1u << 20)) return fail(CHD_ERR_LIMIT, "canvas above 2^20 pixels a side");
    *width = (uint32_t)w;
    *height = (uint32_t)h;
    return CHD_OK;
}

int chd_synthesis_geometry(chd_context *c, const chd_depth_map *map,
                           const chd_synth_source *sources, uint32_t count,
                           const chd_canvas *cv, float *points, float *feather,
                           float *local_confidence, float *base_weight, float *focus_weight) {
    if (!c || !sources || !cv) return fail(CHD_ERR_ARGUMENT, "synthesis_geometry: null argument");
    if (count == 0) return fail(CHD_ERR_ARGUMENT, "no modules to synthesise from");
    for (uint32_t s = 0; s < count; s++) {
        const chd_warp &w = sources[s].warp;
        if (!w.points || w.step == 0 || w.columns == 0 || w.rows == 0 ||
            sources[s].mosaic_width == 0 || sources[s].mosaic_height == 0)
            return fail(CHD_ERR_ARGUMENT, "synthesis_geometry: empty warp or mosaic");
    }
    if (map && (!map->depth || !map->confidence))
        return fail(CHD_ERR_ARGUMENT, "synthesis_geometry: depth map without arrays");
    uint32_t width, height;
    CHD_CHECK(chd_canvas_size(cv, &width, &height));
    CHD_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const size_t plane = (size_t)width * height, total = plane * count;
    const uint32_t bands = (height + kSynthBandRows - 1) / kSynthBandRows;
    std::vector<uint8_t> touch((size_t)bands * count);
    for (uint32_t b = 0; b < bands; b++) {
        const uint64_t start = (uint64_t)b * kSynthBandRows;
        const uint64_t end = std::min<uint64_t>(start + kSynthBandRows, height);
        for (uint32_t s = 0; s < count; s++)
            touch[(size_t)b * count + s] = band_touches(sources[s], start, end, *cv) ? 1 : 0;
    }
    Scratch sc;
    std::vector<SynthSourceDev> dev(count);
    for (uint32_t s = 0; s < count; s++) {
        const chd_synth_source &src = sources[s];
        const size_t nodes = (size_t)src.warp.columns * src.warp.rows;
        float2 *pts;
        CHD_CHECK(sc.alloc(&pts, nodes));
        CHD_HIP(hipMemcpyAsync(pts, src.warp.points, nodes * 8, hipMemcpyHostToDevice, st));
        float *conf = nullptr;
        if (src.warp.confidence) {
            CHD_CHECK(sc.alloc(&conf, nodes));
            CHD_HIP(hipMemcpyAsync(conf, src.warp.confidence, nodes * 4, hipMemcpyHostToDevice, st));
        }
        SynthSourceDev &d = dev[s];
        d.warp.points = pts;
        d.warp.step = src.warp.step;
        d.warp.columns = src.warp.columns;
        d.warp.rows = src.warp.rows;
        d.warp.inverse_step =
            (src.warp.step & (src.warp.step - 1)) == 0 ? 1.0f / (float)src.warp.step : 0.0f;
        d.confidence = conf;
        d.mosaic_width = src.mosaic_width;
        d.mosaic_height = src.mosaic_height;
        d.magnification = src.magnification;
        d.source_confidence = src.confidence;
        d.reference = src.reference;
        d.has_focus_distance = src.has_focus_distance;
        d.focus_distance = src.focus_distance;
    }
    SynthSourceDev *d_sources;
    uint8_t *d_touch;
    CHD_CHECK(sc.alloc(&d_sources, count));
    CHD_CHECK(sc.alloc(&d_touch, touch.size()));
    CHD_HIP(hipMemcpyAsync(d_sources, dev.data(), count * sizeof(SynthSourceDev),
                           hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(d_touch, touch.data(), touch.size(), hipMemcpyHostToDevice, st));
    DepthMapDev dm{};
    if (map) {
        const size_t nodes = (size_t)map->columns * map->rows;
        double *depth;
        float *conf;
        CHD_CHECK(sc.alloc(&depth, nodes));
        CHD_CHECK(sc.alloc(&conf, nodes));
        CHD_HIP(hipMemcpyAsync(depth, map->depth, nodes * 8, hipMemcpyHostToDevice, st));
        CHD_HIP(hipMemcpyAsync(conf, map->confidence, nodes * 4, hipMemcpyHostToDevice, st));
        dm.depth = depth;
        dm.confidence = conf;
        dm.columns = map->columns;
        dm.rows = map->rows;
        dm.step = map->step;
    }
    SynthCanvasDev canvas;
    canvas.crop_x = cv->crop_x;
    canvas.crop_y = cv->crop_y;
    canvas.scale = cv->scale;
    canvas.feather_scale = std::fmax(cv->feather_px, 1.0f);
    canvas.width = width;
    canvas.height = height;
    float2 *d_points = nullptr;
    float *d_feather = nullptr, *d_local = nullptr, *d_base = nullptr, *d_focus = nullptr;
    if (points) CHD_CHECK(sc.alloc(&d_points, total));
    if (feather) CHD_CHECK(sc.alloc(&d_feather, total));
    if (local_confidence) CHD_CHECK(sc.alloc(&d_local, total));
    if (base_weight) CHD_CHECK(sc.alloc(&d_base, total));
    if (focus_weight) CHD_CHECK(sc.alloc(&d_focus, total));
    CHD_HIP(launch_synthesis_geometry(st, canvas, dm, d_sources, count, d_touch, d_points,
                                      d_feather, d_local, d_base, d_focus));
    if (points) CHD_HIP(hipMemcpyAsync(points, d_points, total * 8, hipMemcpyDeviceToHost, st));
    if (feather) CHD_HIP(hipMemcpyAsync(feather, d_feather, total * 4, hipMemcpyDeviceToHost, st));
    if (local_confidence)
        CHD_HIP(hipMemcpyAsync(local_confidence, d_local, total * 4, hipMemcpyDeviceToHost, st));
    if (base_weight) CHD_HIP(hipMemcpyAsync(base_weight, d_base, total * 4, hipMemcpyDeviceToHost, st));
    if (focus_weight)
        CHD_HIP(hipMemcpyAsync(focus_weight, d_focus, total * 4, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

}  // extern "C"

// ------------------------------------------------------------ synthetic frames

struct chd_frames {
    chd_context *context = nullptr;
    std::vector<RenderModule> modules;
    uint32_t reference = 0;
    BayerSpec bayer{};
    uint32_t shift = 0;
    float black = 0.0f, white = 0.0f;
    std::vector<uint16_t *> mosaics;
    std::vector<float *> planes;
    std::vector<void *> allocations;
    SceneDev host_scene{};          // source of the last upload (kept alive)
    SceneDev *scene = nullptr;      // device copy
    float *canvas = nullptr;
    uint8_t *cells = nullptr;
    size_t canvas_points = 0;

    ~chd_frames() {
        if (canvas) hipFree(canvas);
        if (cells) hipFree(cells);
        for (void *p : allocations) hipFree(p);
    }
};

namespace {

uint64_t splitmix64_host(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
uint64_t hash3_host(uint64_t seed, uint64_t stream, uint64_t index) {
    return splitmix64_host(splitmix64_host(seed ^ splitmix64_host(stream * 0x632be59bd9b4e019ULL)) ^
                           index);
}
double unit_host(uint64_t h) { return (double)(h >> 11) * 0x1p-53; }

// The scene of `sc` for a rig: Voronoi sites over the reference raster,
// inverse depths, nearest-first order, and a canvas with a margin of the
// largest parallax of the rig (plus 64 px) around the reference.
bool scene_state(const chd_synthetic_scene &sc, const std::vector<RenderModule> &modules,
                 uint32_t reference, SceneDev &s) {
    if (sc.plane_count == 0 || sc.plane_count > kMaxScenePlanes) return false;
    const RenderModule &ref = modules[reference];
    std::memset(&s, 0, sizeof s);
    s.seed = sc.seed;
    s.n = sc.plane_count;
    s.texture_cell_px = sc.texture_cell_px;
    for (uint32_t i = 0; i < s.n; i++) {
        s.site_x[i] = unit_host(hash3_host(sc.seed, 1, i)) * (double)ref.width;
        s.site_y[i] = unit_host(hash3_host(sc.seed, 2, i)) * (double)ref.height;
        s.inverse[i] = sc.inverse_depth_min +
                       (sc.inverse_depth_max - sc.inverse_depth_min) * unit_host(hash3_host(sc.seed, 3, i));
        s.order[i] = i;
    }
    for (uint32_t i = 1; i < s.n; i++) {  // nearest first, ties by index
        uint32_t v = s.order[i], j = i;
        while (j > 0 && s.inverse[s.order[j - 1]] < s.inverse[v]) {
            s.order[j] = s.order[j - 1];
            j--;
        }
        s.order[j] = v;
    }
    double worst = 0.0;
    for (const RenderModule &m : modules) {
        const double b = std::sqrt(m.center[0] * m.center[0] + m.center[1] * m.center[1]);
        const double d = ref.focal_px * b * sc.inverse_depth_max;
        if (d > worst) worst = d;
    }
    const uint32_t margin = (uint32_t)std::ceil(worst) + 64;
    s.canvas_x0 = -(double)margin;
    s.canvas_y0 = -(double)margin;
    s.canvas_w = ref.width + 2 * margin;
    s.canvas_h = ref.height + 2 * margin;
    return true;
}

}  // namespace

extern "C" {

int chd_frames_create(chd_context *c, const chd_synthetic_module *modules, uint32_t count,
                      uint32_t reference_index, const chd_bayer *bayer, uint32_t shift,
                      float black_q6, float white_q6, chd_frames **out) {
    if (!c || !modules || !bayer || !out || count == 0 || reference_index >= count || shift > 4)
        return fail(CHD_ERR_ARGUMENT, "frames_create: bad argument");
    *out = nullptr;
    if (c->released) return fail(CHD_ERR_STATE, "frames_create: the context was destroyed");
    for (uint32_t i = 0; i < count; i++)
        if (modules[i].width < 2 || modules[i].height < 2 || modules[i].width % 2 ||
            modules[i].height % 2 || !(modules[i].focal_px > 0.0))
            return fail(CHD_ERR_ARGUMENT, "frames_create: module sizes must be even and focal lengths positive");
    CHD_HIP(hipSetDevice(c->device));
    auto *f = new chd_frames;
    f->context = c;
    f->reference = reference_index;
    f->bayer = BayerSpec{bayer->black_level, bayer->white_level, bayer->exposure,
                         bayer->electrons_per_dn, bayer->read_noise_dn};
    f->shift = shift;
    f->black = black_q6;
    f->white = white_q6;
    for (uint32_t i = 0; i < count; i++) {
        RenderModule m{};
        m.focal_px = modules[i].focal_px;
        std::memcpy(m.center, modules[i].center, sizeof m.center);
        std::memcpy(m.rotation_wc, modules[i].rotation_wc, sizeof m.rotation_wc);
        m.width = modules[i].width;
        m.height = modules[i].height;
        f->modules.push_back(m);
    }
    auto alloc = [&](size_t bytes, void **p) {
        hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
        if (e != hipSuccess) return fail(CHD_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        f->allocations.push_back(*p);
        return CHD_OK;
    };
    int r = CHD_OK;
    void *p = nullptr;
    if (r == CHD_OK) r = alloc(sizeof(SceneDev), &p);
    f->scene = (SceneDev *)p;
    for (uint32_t i = 0; i < count && r == CHD_OK; i++) {
        const RenderModule &m = f->modules[i];
        r = alloc((size_t)m.width * m.height * sizeof(uint16_t), &p);
        f->mosaics.push_back((uint16_t *)p);
        if (r == CHD_OK) r = alloc((size_t)(m.width / 2) * (m.height / 2) * sizeof(float), &p);
        f->planes.push_back((float *)p);
    }
    if (r != CHD_OK) {
        delete f;
        return r;
    }
    c->frames.push_back(f);
    *out = f;
    return CHD_OK;
}

int chd_frames_render(chd_frames *f, const chd_synthetic_scene *scene) {
    if (!f || !scene) return fail(CHD_ERR_ARGUMENT, "frames_render: bad argument");
    chd_context *c = f->context;
    CHD_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    SceneDev next;
    if (!scene_state(*scene, f->modules, f->reference, next))
        return fail(CHD_ERR_ARGUMENT, "frames_render: plane_count must be 1..64");
    if (scene->texture_cell_px == 0) return fail(CHD_ERR_ARGUMENT, "frames_render: zero texture cell");
    const size_t points = (size_t)next.canvas_w * next.canvas_h;
    // the previous render may still read the canvas and the scene copy
    CHD_HIP(hipStreamSynchronize(st));
    if (points > f->canvas_points) {
        if (f->canvas) hipFree(f->canvas);
        if (f->cells) hipFree(f->cells);
        f->canvas = nullptr;
        f->cells = nullptr;
        f->canvas_points = 0;
        CHD_HIP(hipMalloc(&f->canvas, points * sizeof(float)));
        CHD_HIP(hipMalloc(&f->cells, points));
        f->canvas_points = points;
    }
    f->host_scene = next;
    CHD_HIP(hipMemcpyAsync(f->scene, &f->host_scene, sizeof(SceneDev), hipMemcpyHostToDevice, st));
    CHD_HIP(launch_scene_canvas(st, f->scene, next.canvas_w, next.canvas_h, f->canvas, f->cells));
    CHD_CHECK(luminance_table(c, f->black, f->white, false));
    const RenderModule &ref = f->modules[f->reference];
    for (uint32_t i = 0; i < f->modules.size(); i++) {
        const RenderModule &m = f->modules[i];
        CHD_HIP(launch_render_bayer(st, f->scene, f->canvas, f->cells, m, ref, i, i == f->reference,
                                    f->bayer, f->shift, f->mosaics[i]));
        CHD_HIP(launch_luminance_half(st, f->mosaics[i], m.width, m.height, c->luminance_table,
                                      0, f->planes[i]));
    }
    return CHD_OK;
}

int chd_frames_plane(const chd_frames *f, uint32_t index, const float **plane,
                     uint32_t *plane_width, uint32_t *plane_height) {
    if (!f || !plane || index >= f->modules.size())
        return fail(CHD_ERR_ARGUMENT, "frames_plane: bad argument");
    *plane = f->planes[index];
    if (plane_width) *plane_width = f->modules[index].width / 2;
    if (plane_height) *plane_height = f->modules[index].height / 2;
    return CHD_OK;
}

int chd_frames_read(chd_frames *f, uint32_t index, uint16_t *mosaic, float *plane) {
    if (!f || index >= f->modules.size()) return fail(CHD_ERR_ARGUMENT, "frames_read: bad argument");
    chd_context *c = f->context;
    CHD_HIP(hipSetDevice(c->device));
    const RenderModule &m = f->modules[index];
    if (mosaic)
        CHD_HIP(hipMemcpyAsync(mosaic, f->mosaics[index], (size_t)m.width * m.height * 2,
                               hipMemcpyDeviceToHost, c->stream));
    if (plane)
        CHD_HIP(hipMemcpyAsync(plane, f->planes[index], (size_t)(m.width / 2) * (m.height / 2) * 4,
                               hipMemcpyDeviceToHost, c->stream));
    CHD_HIP(hipStreamSynchronize(c->stream));
    return CHD_OK;
}

void chd_frames_destroy(chd_frames *f) {
    if (!f) return;
    chd_context *c = f->context;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    unregister(c->frames, f);
    delete f;
    context_release_if_unused(c);
}

int chd_device_expf(chd_context *c, const float *x, float *y, size_t n) {
    if (!c || !x || !y) return fail(CHD_ERR_ARGUMENT, "expf: bad argument");
    CHD_HIP(hipSetDevice(c->device));
    Scratch sc;
    float *dx, *dy;
    CHD_CHECK(sc.alloc(&dx, n));
    CHD_CHECK(sc.alloc(&dy, n));
    CHD_HIP(hipMemcpyAsync(dx, x, n * 4, hipMemcpyHostToDevice, c->stream));
    CHD_HIP(launch_expf(c->stream, dx, dy, n));
    CHD_HIP(hipMemcpyAsync(y, dy, n * 4, hipMemcpyDeviceToHost, c->stream));
    CHD_HIP(hipStreamSynchronize(c->stream));
    return CHD_OK;
}

int chd_device_exp(chd_context *c, const double *x, double *y, size_t n) {
    if (!c || !x || !y) return fail(CHD_ERR_ARGUMENT, "exp: bad argument");
    CHD_HIP(hipSetDevice(c->device));
    Scratch sc;
    double *dx, *dy;
    CHD_CHECK(sc.alloc(&dx, n));
    CHD_CHECK(sc.alloc(&dy, n));
    CHD_HIP(hipMemcpyAsync(dx, x, n * 8, hipMemcpyHostToDevice, c->stream));
    CHD_HIP(launch_exp(c->stream, dx, dy, n));
    CHD_HIP(hipMemcpyAsync(y, dy, n * 8, hipMemcpyDeviceToHost, c->stream));
    CHD_HIP(hipStreamSynchronize(c->stream));
    return CHD_OK;
}

// ------------------------------------------------------------ autotuning

extern "C++" {
namespace {

// Key of the work's shape for the tuning cache.
std::string shape_key(const chd_session *s) {
    const chd_grid_sizes &g = s->grid;
    const uint32_t r = s->options.patch_radius;
    std::string key = "chd-tune-2 views=" + std::to_string(s->active.size()) +
                      " taps=" + std::to_string((2 * r + 1) * (2 * r + 1)) +
                      " labels=" + std::to_string(g.labels) + " grid=" +
                      std::to_string(g.coarse_columns) + "x" + std::to_string(g.coarse_rows);
    for (uint32_t v : s->active)
        key += " " + std::to_string(s->plane_sizes[v].first) + "x" +
               std::to_string(s->plane_sizes[v].second);
    return key;
}

bool read_tuning_cache(const char *path, const std::string &key, chd_tuning &out) {
    if (!path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char line[4096];
    bool found = false;
    while (std::fgets(line, sizeof line, f)) {
        char *tab = std::strchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        chd_tuning t{};
        if (key == line && std::sscanf(tab + 1, "%u %u %u %u", &t.cost_nodes_per_block,
                                       &t.cost_label_slab, &t.sgm_paths_per_block,
                                       &t.cost_tile_order) == 4) {
            out = t;  // the last entry for a key wins
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

void append_tuning_cache(const char *path, const std::string &key, const chd_tuning &t) {
    if (!path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s\t%u %u %u %u\n", key.c_str(), t.cost_nodes_per_block,
                 t.cost_label_slab, t.sgm_paths_per_block, t.cost_tile_order);
    std::fclose(f);
}

// Mean milliseconds of `reps` runs of `launch` on the session's stream.
template <class F>
int time_launches(chd_session *s, int reps, float &ms, F &&launch) {
    hipStream_t st = s->context->stream;
    hipEvent_t a, b;
    CHD_HIP(hipEventCreate(&a));
    CHD_HIP(hipEventCreate(&b));
    int r = launch();  // warm-up
    if (r == CHD_OK) r = hipEventRecord(a, st) == hipSuccess ? CHD_OK : CHD_ERR_HIP;
    for (int i = 0; r == CHD_OK && i < reps; i++) r = launch();
    if (r == CHD_OK) r = hipEventRecord(b, st) == hipSuccess ? CHD_OK : CHD_ERR_HIP;
    if (r == CHD_OK) r = hipEventSynchronize(b) == hipSuccess ? CHD_OK : CHD_ERR_HIP;
    if (r == CHD_OK) r = hipEventElapsedTime(&ms, a, b) == hipSuccess ? CHD_OK : CHD_ERR_HIP;
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (r != CHD_OK) return r == CHD_ERR_HIP ? fail(CHD_ERR_HIP, "autotune timing") : r;
    ms /= (float)reps;
    return CHD_OK;
}

}  // namespace
}  // extern "C++"

int chd_autotune(chd_session *s, const char *cache_path, chd_tuning *out) {
    if (!s || !out) return fail(CHD_ERR_ARGUMENT, "autotune: bad argument");
    if (s->none) return CHD_NONE;
    CHD_HIP(hipSetDevice(s->context->device));
    const std::string key = shape_key(s);
    chd_tuning best{};
    if (read_tuning_cache(cache_path, key, best)) {
        CHD_CHECK(chd_session_set_tuning(s, &best));
        *out = best;
        return CHD_OK;
    }
    // every trial shape is set on the session directly; the session ends
    // on its original shape (on failure) or on `best` through
    // chd_session_set_tuning, which bumps the version so that batches
    // holding the session rebuild their job tables
    const chd_tuning original = s->tuning;
    const bool timing = s->timing;
    s->timing = false;
    const uint32_t L = s->grid.labels;
    float best_ms = 0.0f;
    bool have = false;
    int r = CHD_OK;
    const uint32_t slabs[2] = {L, (L + 1) / 2};
    const uint32_t nodes[7] = {2, 4, 6, 8, 10, 12, 16};
    for (uint32_t oi = 0; r == CHD_OK && oi < 2; oi++)
    for (uint32_t si = 0; r == CHD_OK && si < 2; si++) {
        for (uint32_t ni = 0; r == CHD_OK && ni < 7; ni++) {
            chd_tuning t = s->tuning;
            t.cost_nodes_per_block = nodes[ni];
            t.cost_label_slab = slabs[si];
            t.cost_tile_order = oi;
            s->tuning = t;
            if (cost_job_lds(cost_job(s, nullptr)) > 64 * 1024) continue;
            float ms = 0.0f;
            r = time_launches(s, 2, ms, [&] { return run_cost(s, false); });
            if (r == CHD_OK && (!have || ms < best_ms)) {
                best = t;
                best_ms = ms;
                have = true;
            }
        }
    }
    if (r == CHD_OK && !have) r = fail(CHD_ERR_LIMIT, "autotune: no cost shape fits");
    if (r == CHD_OK) {
        s->tuning = best;
        const uint32_t wpb[3] = {1, 2, 4};
        float sgm_best = 0.0f;
        for (int k = 0; r == CHD_OK && k < 3; k++) {
            chd_tuning t = best;
            t.sgm_paths_per_block = wpb[k];
            s->tuning = t;
            float ms = 0.0f;
            r = time_launches(s, 3, ms, [&] { return run_sgm_paths(s); });
            if (r == CHD_OK && (k == 0 || ms < sgm_best)) {
                sgm_best = ms;
                best.sgm_paths_per_block = wpb[k];
            }
        }
    }
    s->timing = timing;
    s->tuning = original;
    if (r != CHD_OK) return r;
    CHD_CHECK(chd_session_set_tuning(s, &best));
    append_tuning_cache(cache_path, key, best);
    *out = best;
    return CHD_OK;
}

// ------------------------------------------------------------ batches

struct chd_batch {
    chd_context *context = nullptr;
    std::vector<chd_session *> sessions;
    std::vector<void *> allocations;
    CostJob *jobs[2] = {nullptr, nullptr};  // without / with per-view score slices
    bool views_ready = false;
    uint32_t *first = nullptr, *counter = nullptr;
    size_t lds = 0;
    uint32_t blocks = 0;
    SgmJob *sgm_jobs = nullptr;
    uint32_t *sgm_first = nullptr;
    uint32_t sgm_total = 0, labels = 0, sgm_paths_per_block = 4, xcd_order = 0;
    bool sgm_same_grids = true;  // every capture on the same coarse grid
    bool sgm_windows = false;    // some capture sweeps band windows
    bool persistent = false;  // one queue launch for every cost volume
    bool one_sweep = false;   // every capture's SGM sweep in one launch, no pipelining
    std::vector<hipEvent_t> cost_done;  // per capture: its cost volume finished
    hipEvent_t joined = nullptr;        // the side stream's work finished
    bool broken = false;      // a session was destroyed under the batch
    int queue_mode = 4;       // the jobs' common kind for the queue kernel (4: mixed)
    std::vector<std::pair<uint32_t, hipGraphExec_t>> graphs;
    std::vector<hipGraph_t> templates;
    std::vector<uint32_t> versions;  // sessions' versions the graphs were captured with

    void drop_graphs() {
        for (auto &g : graphs) hipGraphExecDestroy(g.second);
        for (hipGraph_t g : templates) hipGraphDestroy(g);
        graphs.clear();
        templates.clear();
    }

    ~chd_batch() {
        for (auto &g : graphs) hipGraphExecDestroy(g.second);
        for (hipGraph_t g : templates) hipGraphDestroy(g);
        for (void *p : allocations) hipFree(p);
        for (hipEvent_t e : cost_done) hipEventDestroy(e);
        if (joined) hipEventDestroy(joined);
    }
};

namespace {

void batch_forget_session(chd_batch *b, chd_session *s) {
    for (chd_session *&m : b->sessions)
        if (m == s) m = nullptr;
    b->broken = true;
    b->drop_graphs();  // their nodes point into the session's buffers
}

}  // namespace

extern "C++" {
namespace {

template <class T>
int balloc(chd_batch *b, T **out, size_t count) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return fail(CHD_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    b->allocations.push_back(p);
    *out = (T *)p;
    return CHD_OK;
}

// Points the context's stream at another one for the stage functions that
// launch on it, and back at the end of the scope.
struct StreamSwap {
    chd_context *c;
    hipStream_t saved;
    StreamSwap(chd_context *context, hipStream_t other) : c(context), saved(context->stream) {
        c->stream = other;
    }
    ~StreamSwap() { c->stream = saved; }
};

// Enqueues the stages of every session; no host synchronisation, so that
// the sequence can be captured into a graph.
int batch_enqueue(chd_batch *b, uint32_t stages) {
    hipStream_t st = b->context->stream;
    const bool keep = (stages & CHD_STAGE_KEEP_AGGREGATED) != 0;
    const bool keep_views = (stages & CHD_STAGE_KEEP_VIEW_SCORES) != 0;
    const uint32_t n = (uint32_t)b->sessions.size();
    // every capture's network and band windows first (one proposal's
    // buffers serve them in turn on this stream)
    if (stages & CHD_STAGE_PROPOSE)
        for (chd_session *s : b->sessions) CHD_CHECK(propose_attached(s));
    if (!b->persistent && !b->one_sweep && (stages & CHD_STAGE_COST) && (stages & CHD_STAGE_SGM)) {
        // pipelined: capture k's later stages on the side stream once its
        // cost volume is done, beside capture k + 1's cost volume
        hipStream_t side = b->context->side;
        for (uint32_t k = 0; k < n; k++) {
            chd_session *s = b->sessions[k];
            CHD_CHECK(prepare_cost(s));
            float *vs = nullptr;
            CHD_CHECK(view_score_buffer(s, keep_views, &vs));
            CHD_HIP(launch_cost_volume(st, cost_job(s, vs)));
            CHD_HIP(hipEventRecord(b->cost_done[k], st));
            CHD_HIP(hipStreamWaitEvent(side, b->cost_done[k], 0));
            StreamSwap swap(b->context, side);
            CHD_CHECK(run_sgm_paths(s));
            CHD_CHECK(run_sgm_rest(s, keep));
            if (stages & CHD_STAGE_DIRECT) CHD_CHECK(run_direct(s));
            if (stages & CHD_STAGE_TARGETS) CHD_CHECK(run_targets(s));
        }
        CHD_HIP(hipEventRecord(b->joined, side));
        CHD_HIP(hipStreamWaitEvent(st, b->joined, 0));
        return CHD_OK;
    }
    if (stages & CHD_STAGE_COST) {
        for (chd_session *s : b->sessions) CHD_CHECK(prepare_cost(s));
        if (b->persistent) {
            CHD_HIP(launch_cost_volume_queue(st, b->jobs[keep_views ? 1 : 0], b->first, n,
                                             b->lds, b->blocks, b->counter, b->xcd_order,
                                             b->queue_mode));
        } else {
            for (chd_session *s : b->sessions) {
                float *vs = nullptr;
                CHD_CHECK(view_score_buffer(s, keep_views, &vs));
                CHD_HIP(launch_cost_volume(st, cost_job(s, vs)));
            }
        }
    }
    if (stages & CHD_STAGE_SGM) {
        CHD_HIP(launch_sgm_batch(st, b->sgm_jobs, b->sgm_first, n, b->sgm_total, b->labels,
                                 b->sgm_paths_per_block, b->sgm_same_grids, b->sgm_windows));
        for (chd_session *s : b->sessions) CHD_CHECK(run_sgm_rest(s, keep));
    }
    if (stages & CHD_STAGE_DIRECT)
        for (chd_session *s : b->sessions) CHD_CHECK(run_direct(s));
    if (stages & CHD_STAGE_TARGETS)
        for (chd_session *s : b->sessions) CHD_CHECK(run_targets(s));
    return CHD_OK;
}

}  // namespace
}  // extern "C++"

extern "C++" {
namespace {

// Job tables, item prefixes and launch sizes from the sessions' current
// launch parameters; drops captured graphs. Synchronous.
int batch_prepare(chd_batch *b) {
    chd_context *c = b->context;
    hipStream_t st = c->stream;
    CHD_HIP(hipStreamSynchronize(st));
    b->drop_graphs();
    b->views_ready = false;
    const uint32_t count = (uint32_t)b->sessions.size();
    std::vector<CostJob> plain(count);
    std::vector<uint32_t> first(count + 1, 0);
    b->lds = 0;
    b->versions.clear();
    b->queue_mode = -1;
    for (uint32_t k = 0; k < count; k++) {
        plain[k] = cost_job(b->sessions[k], nullptr);
        first[k + 1] = first[k] + plain[k].tiles * plain[k].slabs;
        const int mode = cost_job_mode(plain[k]);
        b->queue_mode = b->queue_mode < 0 || b->queue_mode == mode ? mode : 4;
        b->lds = std::max(b->lds, cost_job_lds(plain[k]));
        b->versions.push_back(b->sessions[k]->version);
    }
    if (b->lds > 160 * 1024) return fail(CHD_ERR_LIMIT, "batch: cost work item above 160 KB of LDS");
    int cus = 0;
    CHD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const uint32_t per_cu = std::max<uint32_t>(
        1, std::min<uint32_t>(4, (uint32_t)(160 * 1024 / std::max<size_t>(b->lds, 1))));
    b->blocks = (uint32_t)std::max(cus, 1) * per_cu;
    std::vector<SgmJob> sgm(count);
    std::vector<uint32_t> sgm_first(count + 1, 0);
    b->sgm_windows = false;
    for (uint32_t k = 0; k < count; k++) {
        chd_session *s = b->sessions[k];
        sgm[k] = SgmJob{s->costs, s->guidance, s->grid.coarse_columns, s->grid.coarse_rows,
                        s->paths, s->band ? s->band_lo : nullptr, s->band ? s->band_hi : nullptr};
        if (s->band) b->sgm_windows = true;
        sgm_first[k + 1] = sgm_first[k] + sgm_waves(s->grid.coarse_columns, s->grid.coarse_rows);
        if (s->grid.coarse_columns != sgm[0].columns || s->grid.coarse_rows != sgm[0].rows)
            b->sgm_same_grids = false;
    }
    b->sgm_total = sgm_first[count];
    const uint32_t wpb = b->sessions[0]->tuning.sgm_paths_per_block;
    b->sgm_paths_per_block = wpb ? wpb : 4;
    b->xcd_order = b->sessions[0]->tuning.cost_tile_order;
    CHD_HIP(hipMemcpyAsync(b->jobs[0], plain.data(), count * sizeof(CostJob), hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(b->first, first.data(), (count + 1) * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(b->sgm_jobs, sgm.data(), count * sizeof(SgmJob), hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(b->sgm_first, sgm_first.data(), (count + 1) * 4, hipMemcpyHostToDevice, st));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

}  // namespace
}  // extern "C++"

int chd_batch_create(chd_context *c, chd_session *const *sessions, uint32_t count,
                     uint32_t flags, chd_batch **out) {
    if (!c || !sessions || !out || count == 0) return fail(CHD_ERR_ARGUMENT, "batch: bad argument");
    if (flags & ~(CHD_BATCH_PERSISTENT | CHD_BATCH_ONE_SWEEP))
        return fail(CHD_ERR_ARGUMENT, "batch: unknown flags");
    *out = nullptr;
    if (c->released) return fail(CHD_ERR_STATE, "batch: the context was destroyed");
    CHD_HIP(hipSetDevice(c->device));
    auto *b = new chd_batch;
    b->context = c;
    b->persistent = (flags & CHD_BATCH_PERSISTENT) != 0;
    b->one_sweep = (flags & CHD_BATCH_ONE_SWEEP) != 0;
    auto bail = [&](int code) {
        delete b;
        return code;
    };
    for (uint32_t k = 0; k < count; k++) {
        chd_session *s = sessions[k];
        if (!s || s->context != c)
            return bail(fail(CHD_ERR_ARGUMENT, "batch: session missing or on another context"));
        if (s->none) return bail(fail(CHD_ERR_ARGUMENT, "batch: a session is the reference's None"));
        if (k > 0 && s->grid.labels != b->labels)
            return bail(fail(CHD_ERR_ARGUMENT, "batch: sessions differ in label count"));
        b->labels = s->grid.labels;
        b->sessions.push_back(s);
    }
    int r = CHD_OK;
    if (r == CHD_OK) r = balloc(b, &b->jobs[0], count);
    if (r == CHD_OK) r = balloc(b, &b->jobs[1], count);
    if (r == CHD_OK) r = balloc(b, &b->first, count + 1);
    if (r == CHD_OK) r = balloc(b, &b->counter, kXcds);
    if (r == CHD_OK) r = balloc(b, &b->sgm_jobs, count);
    if (r == CHD_OK) r = balloc(b, &b->sgm_first, count + 1);
    if (r == CHD_OK && !c->side) {
        hipError_t e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
        if (e != hipSuccess) {
            c->side = nullptr;
            r = fail(CHD_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
    }
    for (uint32_t k = 0; r == CHD_OK && k <= count; k++) {
        hipEvent_t e = nullptr;
        hipError_t he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (he != hipSuccess) {
            r = fail(CHD_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(he));
        } else if (k < count) {
            b->cost_done.push_back(e);
        } else {
            b->joined = e;
        }
    }
    if (r == CHD_OK) r = batch_prepare(b);
    if (r != CHD_OK) return bail(r);
    for (chd_session *s : b->sessions)
        if (std::find(s->batches.begin(), s->batches.end(), b) == s->batches.end())
            s->batches.push_back(b);
    c->batches.push_back(b);
    *out = b;
    return CHD_OK;
}

int chd_batch_run(chd_batch *b, uint32_t stages) {
    if (!b) return fail(CHD_ERR_ARGUMENT, "null batch");
    if (b->broken) return fail(CHD_ERR_STATE, "batch: one of its sessions was destroyed");
    CHD_HIP(hipSetDevice(b->context->device));
    for (chd_session *s : b->sessions) {
        const uint32_t done = s->stages_done | (stages & (CHD_STAGE_COST | CHD_STAGE_SGM | CHD_STAGE_DIRECT));
        if ((stages & CHD_STAGE_SGM) && !(done & CHD_STAGE_COST)) return fail(CHD_ERR_STATE, "SGM before cost volume");
        if ((stages & CHD_STAGE_DIRECT) && !(done & CHD_STAGE_SGM)) return fail(CHD_ERR_STATE, "direct pass before SGM");
        if ((stages & CHD_STAGE_TARGETS) && !(done & CHD_STAGE_DIRECT)) return fail(CHD_ERR_STATE, "targets before direct pass");
        if ((stages & CHD_STAGE_PROPOSE) && !s->attached) return fail(CHD_ERR_STATE, "propose: no proposal attached");
    }
    hipStream_t st = b->context->stream;
    // a session whose launch parameters changed (tuning, band) needs new
    // job tables and graphs
    std::vector<uint32_t> versions;
    for (chd_session *s : b->sessions) versions.push_back(s->version);
    if (versions != b->versions) CHD_CHECK(batch_prepare(b));
    if ((stages & CHD_STAGE_KEEP_VIEW_SCORES) && !b->views_ready) {
        // the per-view slices are allocated on first request, outside any capture
        std::vector<CostJob> views;
        for (chd_session *s : b->sessions) {
            float *vs = nullptr;
            CHD_CHECK(view_score_buffer(s, true, &vs));
            views.push_back(cost_job(s, vs));
        }
        CHD_HIP(hipMemcpyAsync(b->jobs[1], views.data(), views.size() * sizeof(CostJob),
                               hipMemcpyHostToDevice, st));
        CHD_HIP(hipStreamSynchronize(st));
        b->views_ready = true;
    }
    hipGraphExec_t exec = nullptr;
    for (auto &g : b->graphs)
        if (g.first == stages) exec = g.second;
    if (!exec) {
        std::vector<bool> timing;
        for (chd_session *s : b->sessions) {
            timing.push_back(s->timing);
            s->timing = false;  // no event records inside the graph
        }
        CHD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int r = batch_enqueue(b, stages);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(st, &graph);
        for (size_t k = 0; k < b->sessions.size(); k++) b->sessions[k]->timing = timing[k];
        if (r != CHD_OK) {
            if (graph) hipGraphDestroy(graph);
            return r;
        }
        if (e != hipSuccess) return fail(CHD_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hipGraphDestroy(graph);
            return fail(CHD_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        b->templates.push_back(graph);
        b->graphs.emplace_back(stages, exec);
    }
    CHD_HIP(hipGraphLaunch(exec, st));
    const bool keep = (stages & CHD_STAGE_KEEP_AGGREGATED) != 0;
    const bool keep_views = (stages & CHD_STAGE_KEEP_VIEW_SCORES) != 0;
    for (chd_session *s : b->sessions) {
        if (stages & CHD_STAGE_COST)
            s->stages_done = (s->stages_done & ~CHD_STAGE_KEEP_VIEW_SCORES) | CHD_STAGE_COST |
                             (keep_views ? CHD_STAGE_KEEP_VIEW_SCORES : 0);
        if (stages & CHD_STAGE_SGM)
            s->stages_done = (s->stages_done & ~CHD_STAGE_KEEP_AGGREGATED) | CHD_STAGE_SGM |
                             (keep ? CHD_STAGE_KEEP_AGGREGATED : 0);
        s->stages_done |= stages & (CHD_STAGE_DIRECT | CHD_STAGE_TARGETS);
    }
    return CHD_OK;
}

void chd_batch_destroy(chd_batch *b) {
    if (!b) return;
    chd_context *c = b->context;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (chd_session *s : b->sessions)
        if (s) unregister(s->batches, b);
    unregister(c->batches, b);
    delete b;
    context_release_if_unused(c);
}

// ------------------------------------------------------------ proposal

struct chd_proposal {
    chd_context *context = nullptr;
    uint32_t width = 0, height = 0, max_batch = 0;
    float input_offset = 0.0f, input_scale = 0.125f;  // input = (cell mean - offset) * scale
    uint32_t label_count = 129;  // centre label = (L - 1) / 2 + (L - 1) / 8 * o0
    ConvLayer layers[7];
    struct Param {
        float *w, *b, *mw, *vw, *mb, *vb, *gw, *gb;
        uint16_t *wk, *wt;
    } params[7];
    uint16_t *act[7] = {};  // act[0] the input, act[i] the output of layer i (i <= 6)
    float *out = nullptr;   // output of layer 7 (f32)
    float *gact[7] = {};    // d loss / d act[i] (f32), i = 1 .. 6
    float *gout = nullptr;  // d loss / d out
    uint16_t *gbf = nullptr;  // BF16 gradient of the layer being back-propagated
    uint16_t *labels = nullptr;
    float *totals = nullptr;
    uint16_t *data_inputs = nullptr, *data_labels = nullptr;
    uint32_t data_count = 0;
    uint64_t step = 0;
    uint32_t *full_tiles = nullptr;
    std::vector<void *> allocations;
    std::vector<void *> dataset;  // allocations of the current training set

    ~chd_proposal() {
        for (void *p : dataset) hipFree(p);
        for (void *p : allocations) hipFree(p);
    }
};

extern "C++" {
namespace {

template <class T>
int palloc(chd_proposal *p, T **out, size_t count, std::vector<void *> *list = nullptr) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return fail(CHD_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    (list ? *list : p->allocations).push_back(q);
    *out = (T *)q;
    return CHD_OK;
}

// The network: 3x3 convolutions, stride 2 down twice, nearest x2 up twice.
void proposal_layers(uint32_t w, uint32_t h, ConvLayer *L) {
    const uint32_t w2 = (w + 1) / 2, h2 = (h + 1) / 2, w4 = (w2 + 1) / 2, h4 = (h2 + 1) / 2;
    struct Spec { uint32_t cin, cout, hin, win, hout, wout, stride, up, relu; };
    const Spec spec[7] = {
        {3, 32, h, w, h, w, 1, 0, 1},       {32, 32, h, w, h2, w2, 2, 0, 1},
        {32, 32, h2, w2, h2, w2, 1, 0, 1},  {32, 32, h2, w2, h4, w4, 2, 0, 1},
        {32, 32, h4, w4, h4, w4, 1, 0, 1},  {32, 32, h4, w4, h2, w2, 1, 1, 1},
        {32, 3, h2, w2, h, w, 1, 1, 0},
    };
    for (int i = 0; i < 7; i++) {
        const Spec &s = spec[i];
        L[i] = ConvLayer{s.cin, (s.cin + 15) / 16 * 16, s.cout, (s.cout + 15) / 16 * 16,
                         s.hin, s.win, s.hout, s.wout, s.stride, s.up, s.relu, 0};
        L[i].k = 9 * L[i].cin_p;
    }
}

size_t layer_out(const ConvLayer &L, uint32_t batch) {
    return (size_t)batch * L.hout * L.wout * L.cout_p;
}

int proposal_refresh_weights(chd_proposal *p) {
    for (int i = 0; i < 7; i++)
        CHD_HIP(launch_weights_bf16(p->context->stream, p->layers[i], p->params[i].w,
                                    p->params[i].wk, p->params[i].wt));
    return CHD_OK;
}

// Forward pass over `batch` inputs already in act[0].
int proposal_forward(chd_proposal *p, uint32_t batch) {
    hipStream_t st = p->context->stream;
    for (int i = 0; i < 7; i++) {
        const ConvLayer &L = p->layers[i];
        void *out = i == 6 ? (void *)p->out : (void *)p->act[i + 1];
        CHD_HIP(launch_conv_fwd(st, L, batch, p->act[i], p->params[i].wt, p->params[i].b, out,
                                i == 6));
    }
    return CHD_OK;
}

// Loss, backward pass and gradients (gw, gb) for `batch` inputs/labels.
int proposal_backward(chd_proposal *p, uint32_t batch) {
    hipStream_t st = p->context->stream;
    const ConvLayer &last = p->layers[6];
    const uint32_t M = batch * last.hout * last.wout;
    for (int i = 0; i < 7; i++) {
        const ConvLayer &L = p->layers[i];
        CHD_HIP(launch_fill_u32(st, (uint32_t *)p->params[i].gw, 0u, (size_t)L.k * L.cout_p));
        CHD_HIP(launch_fill_u32(st, (uint32_t *)p->params[i].gb, 0u, L.cout_p));
    }
    CHD_HIP(launch_proposal_loss(st, p->out, p->labels, M, last.cout_p, p->label_count, (float)M,
                                 p->gout, p->totals));
    for (int i = 6; i >= 0; i--) {
        const ConvLayer &L = p->layers[i];
        const size_t n = layer_out(L, batch);
        const float *g = i == 6 ? p->gout : p->gact[i + 1];
        CHD_HIP(launch_grad_bf16(st, g, i == 6 ? nullptr : p->act[i + 1], n, i != 6, p->gbf));
        CHD_HIP(launch_conv_dw(st, L, batch, p->act[i], p->gbf, p->params[i].gw,
                               p->params[i].gb));
        if (i > 0) CHD_HIP(launch_conv_dx(st, L, batch, p->gbf, p->params[i].wk, p->gact[i]));
    }
    return CHD_OK;
}

int proposal_adam(chd_proposal *p, float lr) {
    hipStream_t st = p->context->stream;
    p->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - std::pow(b1, (float)p->step), bc2 = 1.0f - std::pow(b2, (float)p->step);
    for (int i = 0; i < 7; i++) {
        const ConvLayer &L = p->layers[i];
        auto &q = p->params[i];
        CHD_HIP(launch_adam(st, q.w, q.gw, q.mw, q.vw, L.k * L.cout_p, lr, b1, b2, eps, bc1, bc2));
        CHD_HIP(launch_adam(st, q.b, q.gb, q.mb, q.vb, L.cout_p, lr, b1, b2, eps, bc1, bc2));
    }
    return proposal_refresh_weights(p);
}

size_t proposal_parameter_count(const chd_proposal *p) {
    size_t n = 0;
    for (int i = 0; i < 7; i++) n += (size_t)p->layers[i].k * p->layers[i].cout_p + p->layers[i].cout_p;
    return n;
}

}  // namespace
}  // extern "C++"

int chd_proposal_create(chd_context *c, uint32_t width, uint32_t height, uint32_t max_batch,
                        uint64_t seed, chd_proposal **out) {
    if (!c || !out || width < 4 || height < 4 || max_batch == 0)
        return fail(CHD_ERR_ARGUMENT, "proposal: bad argument");
    if ((size_t)width * height * max_batch > ((size_t)1 << 26))
        return fail(CHD_ERR_LIMIT, "proposal: batch x width x height above 2^26 pixels");
    *out = nullptr;
    if (c->released) return fail(CHD_ERR_STATE, "proposal: the context was destroyed");
    CHD_HIP(hipSetDevice(c->device));
    auto *p = new chd_proposal;
    p->context = c;
    p->width = width;
    p->height = height;
    p->max_batch = max_batch;
    proposal_layers(width, height, p->layers);
    auto bail = [&](int code) {
        delete p;
        return code;
    };
    int r = CHD_OK;
    // He-normal weights from a fixed generator and zero biases for the hidden
    // layers; the output layer starts at a tenth of that scale and a band
    // scale of 8 labels (log-scale bias ln 8), so that the first steps'
    // large errors cannot drive the hidden ReLUs dead
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> normal(0.0f, 1.0f);
    for (int i = 0; i < 7 && r == CHD_OK; i++) {
        const ConvLayer &L = p->layers[i];
        auto &q = p->params[i];
        const size_t nw = (size_t)L.k * L.cout_p;
        for (float **f : {&q.w, &q.mw, &q.vw, &q.gw})
            if (r == CHD_OK) r = palloc(p, f, nw);
        for (float **f : {&q.b, &q.mb, &q.vb, &q.gb})
            if (r == CHD_OK) r = palloc(p, f, L.cout_p);
        if (r == CHD_OK) r = palloc(p, &q.wk, nw);
        if (r == CHD_OK) r = palloc(p, &q.wt, nw);
        if (r != CHD_OK) break;
        const bool output = i == 6;
        std::vector<float> w(nw, 0.0f), zeros(nw, 0.0f), b(L.cout_p, 0.0f);
        const float scale = std::sqrt(2.0f / (9.0f * (float)L.cin));
        for (size_t k = 0; k < nw; k++) w[k] = normal(rng) * (output ? 0.1f * scale : scale);
        if (output) b[1] = std::log(8.0f);
        hipStream_t st = c->stream;
        hipError_t e = hipMemcpyAsync(q.w, w.data(), nw * 4, hipMemcpyHostToDevice, st);
        for (float *f : {q.mw, q.vw})
            if (e == hipSuccess) e = hipMemcpyAsync(f, zeros.data(), nw * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(q.b, b.data(), L.cout_p * 4, hipMemcpyHostToDevice, st);
        for (float *f : {q.mb, q.vb})
            if (e == hipSuccess) e = hipMemcpyAsync(f, zeros.data(), L.cout_p * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) r = fail(CHD_ERR_HIP, std::string("proposal init: ") + hipGetErrorString(e));
    }
    const ConvLayer *L = p->layers;
    if (r == CHD_OK) r = palloc(p, &p->act[0], (size_t)max_batch * L[0].hin * L[0].win * L[0].cin_p);
    for (int i = 1; i < 7 && r == CHD_OK; i++) {
        r = palloc(p, &p->act[i], layer_out(L[i - 1], max_batch));
        if (r == CHD_OK) r = palloc(p, &p->gact[i], layer_out(L[i - 1], max_batch));
    }
    size_t largest = 0;
    for (int i = 0; i < 7; i++) largest = std::max(largest, layer_out(L[i], max_batch));
    if (r == CHD_OK) r = palloc(p, &p->out, layer_out(L[6], max_batch));
    if (r == CHD_OK) r = palloc(p, &p->gout, layer_out(L[6], max_batch));
    if (r == CHD_OK) r = palloc(p, &p->gbf, largest);
    if (r == CHD_OK) r = palloc(p, &p->labels, (size_t)max_batch * width * height);
    if (r == CHD_OK) r = palloc(p, &p->totals, 2);
    if (r == CHD_OK) r = palloc(p, &p->full_tiles, 1);
    if (r == CHD_OK) r = proposal_refresh_weights(p);
    if (r != CHD_OK) return bail(r);
    // act[0]'s padding channels must read zero
    hipError_t e = launch_fill_u32(c->stream, (uint32_t *)p->act[0], 0u,
                                   (size_t)max_batch * width * height * 16 / 2);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bail(fail(CHD_ERR_HIP, std::string("proposal init: ") + hipGetErrorString(e)));
    c->proposals.push_back(p);
    *out = p;
    return CHD_OK;
}

void chd_proposal_destroy(chd_proposal *p) {
    if (!p) return;
    chd_context *c = p->context;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (chd_session *s : c->sessions)
        if (s->attached == p) {  // batches holding s recapture without it
            s->attached = nullptr;
            s->version++;
        }
    unregister(c->proposals, p);
    delete p;
    context_release_if_unused(c);
}

int chd_proposal_parameters(chd_proposal *p, float *get, const float *set, size_t *count) {
    if (!p || !count) return fail(CHD_ERR_ARGUMENT, "proposal_parameters: bad argument");
    const size_t n = proposal_parameter_count(p);
    if (!get && !set) {
        *count = n;
        return CHD_OK;
    }
    if (*count != n) return fail(CHD_ERR_ARGUMENT, "proposal_parameters: wrong count");
    hipStream_t st = p->context->stream;
    size_t at = 0;
    for (int i = 0; i < 7; i++) {
        const ConvLayer &L = p->layers[i];
        const size_t nw = (size_t)L.k * L.cout_p;
        if (get) {
            CHD_HIP(hipMemcpyAsync(get + at, p->params[i].w, nw * 4, hipMemcpyDeviceToHost, st));
            CHD_HIP(hipMemcpyAsync(get + at + nw, p->params[i].b, L.cout_p * 4, hipMemcpyDeviceToHost, st));
        } else {
            CHD_HIP(hipMemcpyAsync(p->params[i].w, set + at, nw * 4, hipMemcpyHostToDevice, st));
            CHD_HIP(hipMemcpyAsync(p->params[i].b, set + at + nw, L.cout_p * 4, hipMemcpyHostToDevice, st));
        }
        at += nw + L.cout_p;
    }
    if (set) CHD_CHECK(proposal_refresh_weights(p));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

int chd_proposal_optimizer_state(chd_proposal *p, float *get, const float *set, size_t *count,
                                 uint64_t *step) {
    if (!p || !count || !step) return fail(CHD_ERR_ARGUMENT, "proposal_optimizer_state: bad argument");
    const size_t n = 2 * proposal_parameter_count(p);
    if (!get && !set) {
        *count = n;
        *step = p->step;
        return CHD_OK;
    }
    if (*count != n) return fail(CHD_ERR_ARGUMENT, "proposal_optimizer_state: wrong count");
    CHD_HIP(hipSetDevice(p->context->device));
    hipStream_t st = p->context->stream;
    const size_t half = n / 2;
    size_t at = 0;
    for (int i = 0; i < 7; i++) {
        const ConvLayer &L = p->layers[i];
        const size_t nw = (size_t)L.k * L.cout_p;
        auto &q = p->params[i];
        // first moments in parameter order, then second moments
        float *const dev[4] = {q.mw, q.mb, q.vw, q.vb};
        const size_t off[4] = {at, at + nw, half + at, half + at + nw};
        const size_t len[4] = {nw, L.cout_p, nw, L.cout_p};
        for (int k = 0; k < 4; k++) {
            if (get)
                CHD_HIP(hipMemcpyAsync(get + off[k], dev[k], len[k] * 4, hipMemcpyDeviceToHost, st));
            else
                CHD_HIP(hipMemcpyAsync(dev[k], set + off[k], len[k] * 4, hipMemcpyHostToDevice, st));
        }
        at += nw + L.cout_p;
    }
    CHD_HIP(hipStreamSynchronize(st));
    if (get) *step = p->step;
    else p->step = *step;
    return CHD_OK;
}

int chd_proposal_set_input_transform(chd_proposal *p, float offset, float scale) {
    if (!p || !(scale > 0.0f) || !std::isfinite(scale) || !std::isfinite(offset))
        return fail(CHD_ERR_ARGUMENT, "proposal_set_input_transform: bad argument");
    p->input_offset = offset;
    p->input_scale = scale;
    return CHD_OK;
}

int chd_proposal_set_label_count(chd_proposal *p, uint32_t label_count) {
    if (!p || label_count < 3 || label_count > 512)
        return fail(CHD_ERR_ARGUMENT, "proposal_set_label_count: bad argument");
    p->label_count = label_count;
    return CHD_OK;
}

int chd_proposal_get_input_transform(const chd_proposal *p, float *offset, float *scale) {
    if (!p || !offset || !scale) return fail(CHD_ERR_ARGUMENT, "proposal_get_input_transform: bad argument");
    *offset = p->input_offset;
    *scale = p->input_scale;
    return CHD_OK;
}

int chd_proposal_set_dataset(chd_proposal *p, const uint16_t *inputs, const uint16_t *labels,
                             uint32_t count) {
    if (!p || !inputs || !labels || count == 0)
        return fail(CHD_ERR_ARGUMENT, "proposal_set_dataset: bad argument");
    CHD_HIP(hipStreamSynchronize(p->context->stream));
    for (void *q : p->dataset) hipFree(q);
    p->dataset.clear();
    p->data_count = 0;
    const size_t pixels = (size_t)p->width * p->height;
    CHD_CHECK(palloc(p, &p->data_inputs, pixels * 16 * count, &p->dataset));
    CHD_CHECK(palloc(p, &p->data_labels, pixels * count, &p->dataset));
    CHD_HIP(hipMemcpy(p->data_inputs, inputs, pixels * 16 * count * 2, hipMemcpyHostToDevice));
    CHD_HIP(hipMemcpy(p->data_labels, labels, pixels * count * 2, hipMemcpyHostToDevice));
    p->data_count = count;
    return CHD_OK;
}

int chd_proposal_train(chd_proposal *p, uint32_t epochs, uint32_t batch, float learning_rate,
                       uint64_t shuffle_seed, float *epoch_loss) {
    return chd_proposal_train_epochs(p, 0, epochs, epochs, batch, learning_rate, shuffle_seed,
                                     epoch_loss);
}

int chd_proposal_train_epochs(chd_proposal *p, uint32_t first_epoch, uint32_t epochs,
                              uint32_t total_epochs, uint32_t batch, float learning_rate,
                              uint64_t shuffle_seed, float *epoch_loss) {
    if (!p || batch == 0 || batch > p->max_batch || first_epoch + epochs > total_epochs)
        return fail(CHD_ERR_ARGUMENT, "proposal_train: bad argument");
    if (p->data_count == 0) return fail(CHD_ERR_STATE, "proposal_train: no training set");
    CHD_HIP(hipSetDevice(p->context->device));
    hipStream_t st = p->context->stream;
    const size_t pixels = (size_t)p->width * p->height;
    std::vector<uint32_t> order(p->data_count);
    // cosine decay from learning_rate to 5% of it over the whole run
    const double per_epoch = (double)((p->data_count + batch - 1) / batch);
    const double total = (double)total_epochs * per_epoch;
    double step = (double)first_epoch * per_epoch;
    for (uint32_t e = first_epoch; e < first_epoch + epochs; e++) {
        for (uint32_t k = 0; k < p->data_count; k++) order[k] = k;
        std::mt19937_64 rng(shuffle_seed + e);
        std::shuffle(order.begin(), order.end(), rng);
        CHD_HIP(launch_fill_u32(st, (uint32_t *)p->totals, 0u, 2));
        for (uint32_t first = 0; first < p->data_count; first += batch) {
            const uint32_t n = std::min(batch, p->data_count - first);
            for (uint32_t k = 0; k < n; k++) {
                const uint32_t s = order[first + k];
                CHD_HIP(hipMemcpyAsync(p->act[0] + k * pixels * 16, p->data_inputs + s * pixels * 16,
                                       pixels * 32, hipMemcpyDeviceToDevice, st));
                CHD_HIP(hipMemcpyAsync(p->labels + k * pixels, p->data_labels + s * pixels,
                                       pixels * 2, hipMemcpyDeviceToDevice, st));
            }
            CHD_CHECK(proposal_forward(p, n));
            CHD_CHECK(proposal_backward(p, n));
            const double decay = 0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * step / total));
            step += 1.0;
            CHD_CHECK(proposal_adam(p, (float)(learning_rate * decay)));
        }
        float totals[2] = {0.0f, 0.0f};
        CHD_HIP(hipMemcpyAsync(totals, p->totals, 8, hipMemcpyDeviceToHost, st));
        CHD_HIP(hipStreamSynchronize(st));
        if (epoch_loss) epoch_loss[e - first_epoch] = totals[1] > 0.0f ? totals[0] / totals[1] : 0.0f;
    }
    return CHD_OK;
}

int chd_proposal_predict(chd_proposal *p, const uint16_t *inputs, uint32_t count,
                         float *outputs) {
    if (!p || !inputs || !outputs || count == 0) return fail(CHD_ERR_ARGUMENT, "proposal_predict: bad argument");
    CHD_HIP(hipSetDevice(p->context->device));
    hipStream_t st = p->context->stream;
    const size_t pixels = (size_t)p->width * p->height;
    std::vector<float> raw(pixels * 16 * p->max_batch);
    for (uint32_t first = 0; first < count; first += p->max_batch) {
        const uint32_t n = std::min(p->max_batch, count - first);
        CHD_HIP(hipMemcpyAsync(p->act[0], inputs + (size_t)first * pixels * 16, n * pixels * 32,
                               hipMemcpyHostToDevice, st));
        CHD_CHECK(proposal_forward(p, n));
        CHD_HIP(hipMemcpyAsync(raw.data(), p->out, n * pixels * 16 * 4, hipMemcpyDeviceToHost, st));
        CHD_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < n * pixels; i++) {
            float *o = outputs + ((size_t)first * pixels + i) * 3;
            o[0] = (float)(p->label_count - 1) * 0.5f +
                   (float)(p->label_count - 1) * 0.125f * raw[i * 16];
            o[1] = raw[i * 16 + 1];
            o[2] = 1.0f / (1.0f + std::exp(-raw[i * 16 + 2]));
        }
    }
    return CHD_OK;
}

int chd_proposal_gradients(chd_proposal *p, const uint16_t *inputs, const uint16_t *labels,
                           uint32_t count, float *gradients, float *loss) {
    if (!p || !inputs || !labels || !gradients || count == 0 || count > p->max_batch)
        return fail(CHD_ERR_ARGUMENT, "proposal_gradients: bad argument");
    CHD_HIP(hipSetDevice(p->context->device));
    hipStream_t st = p->context->stream;
    const size_t pixels = (size_t)p->width * p->height;
    CHD_HIP(hipMemcpyAsync(p->act[0], inputs, count * pixels * 32, hipMemcpyHostToDevice, st));
    CHD_HIP(hipMemcpyAsync(p->labels, labels, count * pixels * 2, hipMemcpyHostToDevice, st));
    CHD_HIP(launch_fill_u32(st, (uint32_t *)p->totals, 0u, 2));
    CHD_CHECK(proposal_forward(p, count));
    CHD_CHECK(proposal_backward(p, count));
    size_t at = 0;
    for (int i = 0; i < 7; i++) {
        const ConvLayer &L = p->layers[i];
        const size_t nw = (size_t)L.k * L.cout_p;
        CHD_HIP(hipMemcpyAsync(gradients + at, p->params[i].gw, nw * 4, hipMemcpyDeviceToHost, st));
        CHD_HIP(hipMemcpyAsync(gradients + at + nw, p->params[i].gb, L.cout_p * 4, hipMemcpyDeviceToHost, st));
        at += nw + L.cout_p;
    }
    float totals[2];
    CHD_HIP(hipMemcpyAsync(totals, p->totals, 8, hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    if (loss) *loss = totals[0];
    return CHD_OK;
}

int chd_proposal_adam_step(chd_proposal *p, const float *gradients, float learning_rate) {
    if (!p || !gradients) return fail(CHD_ERR_ARGUMENT, "proposal_adam_step: bad argument");
    hipStream_t st = p->context->stream;
    size_t at = 0;
    for (int i = 0; i < 7; i++) {
        const ConvLayer &L = p->layers[i];
        const size_t nw = (size_t)L.k * L.cout_p;
        CHD_HIP(hipMemcpyAsync(p->params[i].gw, gradients + at, nw * 4, hipMemcpyHostToDevice, st));
        CHD_HIP(hipMemcpyAsync(p->params[i].gb, gradients + at + nw, L.cout_p * 4, hipMemcpyHostToDevice, st));
        at += nw + L.cout_p;
    }
    CHD_CHECK(proposal_adam(p, learning_rate));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

}  // extern "C"

namespace {

// The network input of the session's planes into the proposal's first
// batch slot: the reference and the first two targets, means over cells of
// (grid step / 2)^2 plane pixels through the input transform.
int proposal_input(chd_session *s, chd_proposal *p) {
    if (s->context != p->context) return fail(CHD_ERR_ARGUMENT, "proposal: another context");
    if (s->active.empty()) return fail(CHD_ERR_STATE, "set_proposal: no target view");
    const auto ref = s->plane_sizes[s->reference];
    // one network pixel per coarse node: cells of step / 2 plane pixels
    const uint32_t factor = s->grid.coarse_step / 2;
    if (s->grid.coarse_step % 2 != 0 || ref.first / factor != p->width ||
        ref.second / factor != p->height)
        return fail(CHD_ERR_ARGUMENT, "set_proposal: the network's grid is not the reference plane in cells of half the grid step");
    if (s->grid.labels != p->label_count)
        return fail(CHD_ERR_ARGUMENT, "set_proposal: the network's label count differs from the session's");
    CHD_HIP(hipSetDevice(s->context->device));
    hipStream_t st = s->context->stream;
    const size_t nodes = (size_t)s->grid.coarse_columns * s->grid.coarse_rows;
    if (!s->band_lo) {
        CHD_CHECK(dalloc(s, &s->band_lo, nodes));
        CHD_CHECK(dalloc(s, &s->band_hi, nodes));
    }
    // input: the reference and the first two targets, halved again
    const std::vector<ViewDev> &views = s->host_views;
    const float *planes[3] = {views[s->reference].plane, nullptr, nullptr};
    uint32_t pitches[3] = {views[s->reference].plane_pitch, 0, 0};
    for (size_t k = 0; k < 2 && k < s->active.size(); k++) {
        const ViewDev &v = views[s->active[k]];
        if (v.plane_width != ref.first || v.plane_height != ref.second)
            return fail(CHD_ERR_ARGUMENT, "set_proposal: target planes differ in size from the reference");
        planes[1 + k] = v.plane;
        pitches[1 + k] = v.plane_pitch;
    }
    CHD_HIP(launch_quarter_input(st, planes, pitches, ref.first, ref.second, p->width, p->height,
                                 factor, p->input_offset, p->input_scale, p->act[0]));
    return CHD_OK;
}

// The network on the session's planes and the band windows it gives, on the
// context's stream without host synchronisation (a batch captures it).
int propose_windows(chd_session *s, chd_proposal *p, float threshold, uint32_t half_width,
                    uint32_t max_half_width, uint32_t fallback_half_width, uint32_t tile) {
    CHD_CHECK(proposal_input(s, p));
    hipStream_t st = s->context->stream;
    CHD_CHECK(proposal_forward(p, 1));
    CHD_HIP(launch_fill_u32(st, p->full_tiles, 0u, 1));
    CHD_HIP(launch_band_windows(st, p->out, p->layers[6].cout_p, p->width, p->height,
                                s->grid.coarse_columns, s->grid.coarse_rows, s->grid.labels,
                                threshold, half_width, std::max(half_width, max_half_width),
                                tile, fallback_half_width, s->band_lo, s->band_hi,
                                p->full_tiles));
    return CHD_OK;
}

// CHD_STAGE_PROPOSE: the attached proposal with its setting (the window
// shape is the one attach_proposal set, so no launch parameter changes).
int propose_attached(chd_session *s) {
    return propose_windows(s, s->attached, s->attached_threshold, s->attached_half,
                           s->attached_max_half, s->attached_fallback, s->attached_tile);
}

}  // namespace

extern "C" {

int chd_session_network_input(chd_session *s, chd_proposal *p, uint16_t *out) {
    if (!s || !p || !out) return fail(CHD_ERR_ARGUMENT, "network_input: bad argument");
    if (s->none) return CHD_NONE;
    CHD_CHECK(proposal_input(s, p));
    hipStream_t st = s->context->stream;
    CHD_HIP(hipMemcpyAsync(out, p->act[0], (size_t)p->width * p->height * 16 * sizeof(uint16_t),
                           hipMemcpyDeviceToHost, st));
    CHD_HIP(hipStreamSynchronize(st));
    return CHD_OK;
}

int chd_session_set_proposal(chd_session *s, chd_proposal *p, float threshold,
                             uint32_t half_width, uint32_t max_half_width,
                             uint32_t fallback_half_width, uint32_t tile, uint32_t *full_tiles,
                             uint32_t *tiles) {
    if (!s || !p || tile == 0 || tile > 16 || half_width == 0)
        return fail(CHD_ERR_ARGUMENT, "set_proposal: bad argument");
    if (s->none) return CHD_NONE;
    hipStream_t st = s->context->stream;
    CHD_CHECK(propose_windows(s, p, threshold, half_width, max_half_width, fallback_half_width,
                              tile));
    const uint32_t window_max = 2 * std::max(half_width, max_half_width) + 2;
    if (!s->band || s->band_window_max != window_max) {
        s->band = true;
        s->band_window_max = window_max;  // shapes the band items
        s->version++;
    }
    if (tiles)
        *tiles = ((s->grid.coarse_columns + tile - 1) / tile) * ((s->grid.coarse_rows + tile - 1) / tile);
    if (full_tiles) {
        CHD_HIP(hipMemcpyAsync(full_tiles, p->full_tiles, 4, hipMemcpyDeviceToHost, st));
        CHD_HIP(hipStreamSynchronize(st));
    }
    return CHD_OK;
}

int chd_session_read_band(chd_session *s, uint16_t *lo, uint16_t *hi) {
    if (!s || !lo || !hi) return fail(CHD_ERR_ARGUMENT, "read_band: bad argument");
    if (s->none) return CHD_NONE;
    if (!s->band) return fail(CHD_ERR_STATE, "read_band: the session sweeps every label");
    CHD_HIP(hipSetDevice(s->context->device));
    const size_t nodes = (size_t)s->grid.coarse_columns * s->grid.coarse_rows;
    CHD_CHECK(download(s, lo, s->band_lo, nodes * 2));
    CHD_CHECK(download(s, hi, s->band_hi, nodes * 2));
    CHD_HIP(hipStreamSynchronize(s->context->stream));
    return CHD_OK;
}

int chd_session_attach_proposal(chd_session *s, chd_proposal *p, float threshold,
                                uint32_t half_width, uint32_t max_half_width,
                                uint32_t fallback_half_width, uint32_t tile,
                                uint32_t *full_tiles, uint32_t *tiles) {
    if (s && s->none) return CHD_NONE;
    CHD_CHECK(chd_session_set_proposal(s, p, threshold, half_width, max_half_width,
                                       fallback_half_width, tile, full_tiles, tiles));
    s->attached = p;
    s->attached_threshold = threshold;
    s->attached_half = half_width;
    s->attached_max_half = max_half_width;
    s->attached_fallback = fallback_half_width;
    s->attached_tile = tile;
    s->version++;  // batch graphs recapture with this step
    return CHD_OK;
}

int chd_session_clear_proposal(chd_session *s) {
    if (!s) return fail(CHD_ERR_ARGUMENT, "null session");
    if (s->band || s->attached) {
        s->band = false;
        s->attached = nullptr;
        s->version++;
    }
    return CHD_OK;
}

}  // extern "C"
