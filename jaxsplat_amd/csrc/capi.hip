// extern "C" entry points of libjaxsplat_amd.so (declared in
// include/jaxsplat_amd.h): argument validation, descriptor -> camera
// conversion, workspace carving, and the four reference-compatible XLA
// custom-call targets.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/jaxsplat_amd.h"
#include "common.hpp"

namespace jsa {
hipError_t raster_empty(const Camera& c, const float* background, float* final_Ts,
                        int* final_idx, float* out_img, hipStream_t s, int nch = 3);
hipError_t zero_grads(float* const* ptrs, const size_t* counts, int k, hipStream_t s);
}  // namespace jsa

static_assert(sizeof(jaxsplat_descriptor) == 80, "descriptor must match the reference layout");
static_assert(offsetof(jaxsplat_descriptor, intrins) == 16, "descriptor layout");
static_assert(offsetof(jaxsplat_descriptor, block_dim_2d) == 64, "descriptor layout");

namespace {

thread_local std::string g_err;

}  // namespace

namespace jsa {
// The thread's jaxsplat_last_error() message, for entry points in other files.
void set_last_error(const std::string& msg) { g_err = msg; }

// Stage marks (jaxsplat_set_stage_marks): events of this thread recorded at
// internal stage boundaries.
namespace {
thread_local hipEvent_t g_marks[kNumStageMarks] = {};
}  // namespace

void stage_mark(int k, hipStream_t s) {
    if (k >= 0 && k < kNumStageMarks && g_marks[k]) (void)hipEventRecord(g_marks[k], s);
}
}  // namespace jsa

namespace {

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    return fail(JAXSPLAT_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

#define JSA_CHECK_HIP(expr, where)                  \
    do {                                            \
        hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return hip_fail(e_, where); \
    } while (0)

int validate(const jaxsplat_descriptor* d) {
    if (!d) return fail(JAXSPLAT_ERR_INVALID, "descriptor is NULL");
    if (d->block_width < 1 || d->block_width > 16)
        return fail(JAXSPLAT_ERR_INVALID, "block_width must be in [1, 16], got %u", d->block_width);
    if (d->img_shape.x == 0 || d->img_shape.y == 0)
        return fail(JAXSPLAT_ERR_INVALID, "img_shape must be positive");
    if (!(d->clip_thresh >= 0.f))
        return fail(JAXSPLAT_ERR_INVALID, "clip_thresh must be >= 0");
    if (d->num_points > (1u << 30))
        return fail(JAXSPLAT_ERR_INVALID, "num_points too large");
    return JAXSPLAT_OK;
}

jsa::Camera camera_of(const jaxsplat_descriptor* d) {
    jsa::Camera c;
    c.num_points = (int)d->num_points;
    c.img_w = (int)d->img_shape.x;
    c.img_h = (int)d->img_shape.y;
    c.block_width = (int)d->block_width;
    c.tiles_x = (c.img_w + c.block_width - 1) / c.block_width;
    c.tiles_y = (c.img_h + c.block_width - 1) / c.block_width;
    c.fx = d->intrins[0];
    c.fy = d->intrins[1];
    c.cx = d->intrins[2];
    c.cy = d->intrins[3];
    // forward.cu:70-71 evaluates 0.5 * W / fx in double before rounding to float
    c.tan_fovx = (float)(0.5 * (double)c.img_w / (double)c.fx);
    c.tan_fovy = (float)(0.5 * (double)c.img_h / (double)c.fy);
    c.glob_scale = d->glob_scale;
    c.clip_thresh = d->clip_thresh;
    return c;
}

const jaxsplat_descriptor* unpack(const char* opaque, size_t len) {
    if (!opaque || len != sizeof(jaxsplat_descriptor)) {
        fail(JAXSPLAT_ERR_INVALID, "invalid opaque descriptor size %zu", len);
        return nullptr;
    }
    return reinterpret_cast<const jaxsplat_descriptor*>(opaque);
}

// Grow-only device workspace per (device, stream) for the custom-call targets.
struct Cached {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<std::pair<int, void*>, Cached> g_ws;

void* cached_workspace(void* stream, size_t bytes, int* status) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        *status = fail(JAXSPLAT_ERR_HIP, "hipGetDevice failed");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Cached& c = g_ws[{dev, stream}];
    if (c.bytes < bytes) {
        if (c.ptr) {
            (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
            (void)hipFree(c.ptr);
            c.ptr = nullptr;
            c.bytes = 0;
        }
        const size_t want = bytes + bytes / 4 + (1 << 20);
        hipError_t e = hipMalloc(&c.ptr, want);
        if (e != hipSuccess) {
            *status = hip_fail(e, "workspace hipMalloc");
            c.ptr = nullptr;
            return nullptr;
        }
        c.bytes = want;
    }
    *status = JAXSPLAT_OK;
    return c.ptr;
}

int read_num_intersects(const int32_t* cum, uint32_t n, void* stream, uint32_t* out) {
    *out = 0;
    if (n == 0) return JAXSPLAT_OK;
    int32_t v = 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    JSA_CHECK_HIP(hipMemcpyAsync(&v, cum + (n - 1), sizeof v, hipMemcpyDeviceToHost, s),
                  "read num_intersects");
    JSA_CHECK_HIP(hipStreamSynchronize(s), "read num_intersects");
    if (v < 0) return fail(JAXSPLAT_ERR_INVALID, "negative intersection count");
    *out = (uint32_t)v;
    return JAXSPLAT_OK;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" {

const char* jaxsplat_version(void) { return "jaxsplat_amd 0.1.0 (gfx950)"; }

const char* jaxsplat_last_error(void) { return g_err.c_str(); }

void jaxsplat_set_stage_marks(void* const* events, int count) {
    for (int k = 0; k < jsa::kNumStageMarks; ++k)
        jsa::g_marks[k] = (events && k < count) ? static_cast<hipEvent_t>(events[k]) : nullptr;
}

int jaxsplat_make_descriptor(jaxsplat_descriptor* out, uint32_t num_points, uint32_t img_h,
                             uint32_t img_w, float fx, float fy, float cx, float cy,
                             float glob_scale, float clip_thresh, uint32_t block_width) {
    g_err.clear();
    if (!out) return fail(JAXSPLAT_ERR_INVALID, "out is NULL");
    std::memset(out, 0, sizeof *out);
    out->num_points = num_points;
    out->img_shape = {img_w, img_h, 1};
    out->intrins[0] = fx;
    out->intrins[1] = fy;
    out->intrins[2] = cx;
    out->intrins[3] = cy;
    out->glob_scale = glob_scale;
    out->clip_thresh = clip_thresh;
    out->block_width = block_width;
    out->block_dim_1d = block_width * block_width;
    out->grid_dim_1d = out->block_dim_1d ? (num_points + out->block_dim_1d - 1) / out->block_dim_1d : 0;
    out->block_dim_2d = {block_width, block_width, 1};
    out->grid_dim_2d = {block_width ? (img_w + block_width - 1) / block_width : 0,
                        block_width ? (img_h + block_width - 1) / block_width : 0, 1};
    return validate(out);
}

size_t jaxsplat_project_workspace_size(const jaxsplat_descriptor* d) {
    return d ? align256(jsa::project_fwd_scratch_bytes((int)d->num_points)) : 0;
}

size_t jaxsplat_tile_rects_workspace_size(const jaxsplat_descriptor* d) {
    return jaxsplat_project_workspace_size(d);
}

int jaxsplat_tile_rects(const jaxsplat_descriptor* d, const float* xys, const int32_t* radii,
                        const float* conics, const float* opacities, uint32_t* tile_rects,
                        int32_t* num_tiles_hit, int32_t* cum_tiles_hit, void* workspace,
                        size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!xys || !radii || !conics || !opacities || !tile_rects || !cum_tiles_hit)
        return fail(JAXSPLAT_ERR_INVALID, "tile_rects: NULL required pointer");
    if (!workspace || workspace_bytes < jaxsplat_tile_rects_workspace_size(d))
        return fail(JAXSPLAT_ERR_WORKSPACE, "tile_rects: workspace too small");
    JSA_CHECK_HIP(jsa::tile_rects(camera_of(d), xys, radii, conics, opacities,
                                  reinterpret_cast<uint2*>(tile_rects), num_tiles_hit,
                                  cum_tiles_hit, workspace, static_cast<hipStream_t>(stream)),
                  "tile_rects");
    return JAXSPLAT_OK;
}

int jaxsplat_project_forward(const jaxsplat_descriptor* d, const float* means3d,
                             const float* scales, const float* quats, const float* viewmat,
                             float* cov3d, float* xys, float* depths, int32_t* radii,
                             float* conics, float* compensation, int32_t* num_tiles_hit,
                             int32_t* cum_tiles_hit, void* workspace, size_t workspace_bytes,
                             void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!means3d || !scales || !quats || !viewmat || !xys || !depths || !radii || !conics ||
        !num_tiles_hit || !cum_tiles_hit)
        return fail(JAXSPLAT_ERR_INVALID, "project_forward: NULL required pointer");
    if (!workspace || workspace_bytes < jaxsplat_project_workspace_size(d))
        return fail(JAXSPLAT_ERR_WORKSPACE, "project_forward: workspace too small");
    const jsa::Camera c = camera_of(d);
    JSA_CHECK_HIP(jsa::project_fwd(c, means3d, scales, quats, viewmat, cov3d, xys, depths, radii,
                                   conics, compensation, num_tiles_hit, cum_tiles_hit, nullptr,
                                   nullptr, workspace, static_cast<hipStream_t>(stream)),
                  "project_forward");
    return JAXSPLAT_OK;
}

int jaxsplat_project_forward_noalloc(const jaxsplat_descriptor* d, const float* means3d,
                                     const float* scales, const float* quats,
                                     const float* viewmat, float* cov3d, float* xys,
                                     float* depths, int32_t* radii, float* conics,
                                     float* compensation, int32_t* num_tiles_hit,
                                     int32_t* cum_tiles_hit, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (d->num_points > (uint32_t)jsa::kStaticProjectPoints)
        return fail(JAXSPLAT_ERR_CAPACITY,
                    "project_forward_noalloc: %u Gaussians, at most %d with the library's "
                    "own scratch (use jaxsplat_project_forward with a workspace)",
                    d->num_points, jsa::kStaticProjectPoints);
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!means3d || !scales || !quats || !viewmat || !xys || !depths || !radii || !conics ||
        !num_tiles_hit || !cum_tiles_hit)
        return fail(JAXSPLAT_ERR_INVALID, "project_forward_noalloc: NULL required pointer");
    void* scratch = nullptr;
    JSA_CHECK_HIP(jsa::project_static_scratch(&scratch), "project_forward_noalloc: scratch");
    JSA_CHECK_HIP(jsa::project_fwd(camera_of(d), means3d, scales, quats, viewmat, cov3d, xys,
                                   depths, radii, conics, compensation, num_tiles_hit,
                                   cum_tiles_hit, nullptr, nullptr, scratch,
                                   static_cast<hipStream_t>(stream)),
                  "project_forward_noalloc");
    return JAXSPLAT_OK;
}

int jaxsplat_project_forward_rects(const jaxsplat_descriptor* d, const float* means3d,
                                   const float* scales, const float* quats,
                                   const float* viewmat, const float* opacities, float* cov3d,
                                   float* xys, float* depths, int32_t* radii, float* conics,
                                   float* compensation, int32_t* num_tiles_hit,
                                   int32_t* cum_tiles_hit, uint32_t* tile_rects,
                                   void* workspace, size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!means3d || !scales || !quats || !viewmat || !opacities || !xys || !depths || !radii ||
        !conics || !num_tiles_hit || !cum_tiles_hit || !tile_rects)
        return fail(JAXSPLAT_ERR_INVALID, "project_forward_rects: NULL required pointer");
    if (!workspace || workspace_bytes < jaxsplat_project_workspace_size(d))
        return fail(JAXSPLAT_ERR_WORKSPACE, "project_forward_rects: workspace too small");
    const jsa::Camera c = camera_of(d);
    JSA_CHECK_HIP(jsa::project_fwd(c, means3d, scales, quats, viewmat, cov3d, xys, depths, radii,
                                   conics, compensation, num_tiles_hit, cum_tiles_hit, opacities,
                                   reinterpret_cast<uint2*>(tile_rects), workspace,
                                   static_cast<hipStream_t>(stream)),
                  "project_forward_rects");
    return JAXSPLAT_OK;
}

int jaxsplat_project_forward_splats(const jaxsplat_descriptor* d, const float* means3d,
                                    const float* scales, const float* quats,
                                    const float* viewmat, const float* opacities,
                                    const float* colors, uint32_t channels, float* xys,
                                    float* depths, int32_t* radii, float* conics,
                                    float* compensation, int32_t* num_tiles_hit,
                                    int32_t* cum_tiles_hit, uint32_t* tile_rects, float* splats,
                                    void* workspace, size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (channels < 1 || channels > 3)
        return fail(JAXSPLAT_ERR_INVALID,
                    "project_forward_splats: channels must be in [1, 3], got %u", channels);
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!means3d || !scales || !quats || !viewmat || !opacities || !colors || !depths ||
        !radii || !conics || !cum_tiles_hit || !tile_rects || !splats)
        return fail(JAXSPLAT_ERR_INVALID, "project_forward_splats: NULL required pointer");
    if (!workspace || workspace_bytes < jaxsplat_project_workspace_size(d))
        return fail(JAXSPLAT_ERR_WORKSPACE, "project_forward_splats: workspace too small");
    JSA_CHECK_HIP(jsa::project_fwd(camera_of(d), means3d, scales, quats, viewmat, nullptr, xys,
                                   depths, radii, conics, compensation, num_tiles_hit,
                                   cum_tiles_hit, opacities,
                                   reinterpret_cast<uint2*>(tile_rects), workspace,
                                   static_cast<hipStream_t>(stream), colors, (int)channels,
                                   reinterpret_cast<float4*>(splats)),
                  "project_forward_splats");
    return JAXSPLAT_OK;
}

int jaxsplat_project_backward(const jaxsplat_descriptor* d, const float* means3d,
                              const float* scales, const float* quats, const float* viewmat,
                              const int32_t* radii, const float* conics,
                              const float* compensation, const float* v_xy,
                              const float* v_depth, const float* v_conic,
                              const float* v_compensation, float* v_mean3d, float* v_scale,
                              float* v_quat, float* v_cov2d, float* v_cov3d, int adjoint,
                              void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!means3d || !scales || !quats || !viewmat || !radii || !conics || !v_xy || !v_conic ||
        !v_mean3d || !v_scale || !v_quat)
        return fail(JAXSPLAT_ERR_INVALID, "project_backward: NULL required pointer");
    if (v_compensation && !compensation)
        return fail(JAXSPLAT_ERR_INVALID, "project_backward: v_compensation needs compensation");
    if (adjoint != JAXSPLAT_ADJOINT_REFERENCE && adjoint != JAXSPLAT_ADJOINT_CONSISTENT)
        return fail(JAXSPLAT_ERR_INVALID, "unknown adjoint mode %d", adjoint);
    const jsa::Camera c = camera_of(d);
    JSA_CHECK_HIP(jsa::project_bwd(c, means3d, scales, quats, viewmat, radii, conics,
                                   compensation, v_xy, v_depth, v_conic, v_compensation,
                                   v_mean3d, v_scale, v_quat, v_cov2d, v_cov3d,
                                   adjoint == JAXSPLAT_ADJOINT_CONSISTENT,
                                   static_cast<hipStream_t>(stream)),
                  "project_backward");
    return JAXSPLAT_OK;
}

size_t jaxsplat_bin_workspace_size(const jaxsplat_descriptor* d, uint32_t num_intersects) {
    if (!d || validate(d)) return 0;
    const jsa::BinPlan p = jsa::bin_plan(camera_of(d), (int)num_intersects);
    return align256(p.scratch_bytes);
}

int jaxsplat_bin_and_sort(const jaxsplat_descriptor* d, const float* xys, const float* depths,
                          const int32_t* radii, uint32_t num_intersects,
                          int32_t* gaussian_ids_sorted, int32_t* tile_ids_sorted,
                          int32_t* tile_bins, void* workspace, size_t workspace_bytes,
                          void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (num_intersects >= (1u << 30))
        return fail(JAXSPLAT_ERR_CAPACITY, "num_intersects %u exceeds 2^30", num_intersects);
    if (!tile_bins) return fail(JAXSPLAT_ERR_INVALID, "bin_and_sort: tile_bins is NULL");
    const jsa::Camera c = camera_of(d);
    const jsa::BinPlan p = jsa::bin_plan(c, (int)num_intersects);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_intersects == 0 || d->num_points == 0) {
        JSA_CHECK_HIP(jsa::fill_zero(tile_bins, 8ull * p.num_tiles, s), "bin_and_sort");
        return JAXSPLAT_OK;
    }
    if (!xys || !depths || !radii || !gaussian_ids_sorted)
        return fail(JAXSPLAT_ERR_INVALID, "bin_and_sort: NULL required pointer");
    if (!workspace || workspace_bytes < align256(p.scratch_bytes))
        return fail(JAXSPLAT_ERR_WORKSPACE, "bin_and_sort: workspace too small (%zu < %zu)",
                    workspace_bytes, align256(p.scratch_bytes));
    JSA_CHECK_HIP(jsa::bin_and_sort(p, c, xys, depths, radii, (int)num_intersects,
                                    gaussian_ids_sorted, tile_ids_sorted, tile_bins, workspace, s),
                  "bin_and_sort");
    return JAXSPLAT_OK;
}

size_t jaxsplat_bin_depth_workspace_size(const jaxsplat_descriptor* d) {
    if (!d || validate(d)) return 0;
    return align256(jsa::bin_plan(camera_of(d), 0).depth_bytes);
}

int jaxsplat_bin_depth_sort(const jaxsplat_descriptor* d, const float* xys, const float* depths,
                            const int32_t* radii, void* depth_workspace,
                            size_t depth_workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (d->num_points == 0) return JAXSPLAT_OK;
    const jsa::Camera c = camera_of(d);
    const jsa::BinPlan p = jsa::bin_plan(c, 0);
    if (!xys || !depths || !radii)
        return fail(JAXSPLAT_ERR_INVALID, "bin_depth_sort: NULL required pointer");
    if (!depth_workspace || depth_workspace_bytes < p.depth_bytes)
        return fail(JAXSPLAT_ERR_WORKSPACE, "bin_depth_sort: workspace too small (%zu < %zu)",
                    depth_workspace_bytes, p.depth_bytes);
    JSA_CHECK_HIP(jsa::bin_depth(p, c, xys, depths, radii, nullptr, depth_workspace,
                                 static_cast<hipStream_t>(stream)),
                  "bin_depth_sort");
    return JAXSPLAT_OK;
}

int jaxsplat_bin_depth_sort_rects(const jaxsplat_descriptor* d, const float* depths,
                                  const int32_t* radii, const uint32_t* tile_rects,
                                  void* depth_workspace, size_t depth_workspace_bytes,
                                  void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (d->num_points == 0) return JAXSPLAT_OK;
    const jsa::Camera c = camera_of(d);
    const jsa::BinPlan p = jsa::bin_plan(c, 0);
    if (!depths || !radii || !tile_rects)
        return fail(JAXSPLAT_ERR_INVALID, "bin_depth_sort_rects: NULL required pointer");
    if (!depth_workspace || depth_workspace_bytes < p.depth_bytes)
        return fail(JAXSPLAT_ERR_WORKSPACE,
                    "bin_depth_sort_rects: workspace too small (%zu < %zu)",
                    depth_workspace_bytes, p.depth_bytes);
    JSA_CHECK_HIP(jsa::bin_depth(p, c, nullptr, depths, radii,
                                 reinterpret_cast<const uint2*>(tile_rects), depth_workspace,
                                 static_cast<hipStream_t>(stream)),
                  "bin_depth_sort_rects");
    return JAXSPLAT_OK;
}

size_t jaxsplat_bin_emit_workspace_size(const jaxsplat_descriptor* d, uint32_t num_intersects) {
    if (!d || validate(d)) return 0;
    return align256(jsa::bin_plan(camera_of(d), (int)num_intersects).emit_bytes);
}

int jaxsplat_bin_emit_sort(const jaxsplat_descriptor* d, uint32_t num_intersects,
                           const void* depth_workspace, int32_t* gaussian_ids_sorted,
                           int32_t* tile_ids_sorted, int32_t* tile_bins, void* workspace,
                           size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (num_intersects >= (1u << 30))
        return fail(JAXSPLAT_ERR_CAPACITY, "num_intersects %u exceeds 2^30", num_intersects);
    if (!tile_bins) return fail(JAXSPLAT_ERR_INVALID, "bin_emit_sort: tile_bins is NULL");
    const jsa::Camera c = camera_of(d);
    const jsa::BinPlan p = jsa::bin_plan(c, (int)num_intersects);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_intersects == 0 || d->num_points == 0) {
        JSA_CHECK_HIP(jsa::fill_zero(tile_bins, 8ull * p.num_tiles, s), "bin_emit_sort");
        return JAXSPLAT_OK;
    }
    if (!depth_workspace || !gaussian_ids_sorted)
        return fail(JAXSPLAT_ERR_INVALID, "bin_emit_sort: NULL required pointer");
    if (!workspace || workspace_bytes < p.emit_bytes)
        return fail(JAXSPLAT_ERR_WORKSPACE, "bin_emit_sort: workspace too small (%zu < %zu)",
                    workspace_bytes, p.emit_bytes);
    JSA_CHECK_HIP(jsa::bin_emit(p, c, (int)num_intersects, depth_workspace,
                                gaussian_ids_sorted, tile_ids_sorted, tile_bins, workspace, s),
                  "bin_emit_sort");
    return JAXSPLAT_OK;
}

int jaxsplat_bin_rects_sort(const jaxsplat_descriptor* d, const float* depths,
                            const uint32_t* tile_rects, const int32_t* cum_tiles_hit,
                            uint32_t num_intersects, int32_t* gaussian_ids_sorted,
                            int32_t* tile_ids_sorted, int32_t* tile_bins, void* workspace,
                            size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (num_intersects >= (1u << 30))
        return fail(JAXSPLAT_ERR_CAPACITY, "num_intersects %u exceeds 2^30", num_intersects);
    if (!tile_bins) return fail(JAXSPLAT_ERR_INVALID, "bin_rects_sort: tile_bins is NULL");
    const jsa::Camera c = camera_of(d);
    const jsa::BinPlan p = jsa::bin_plan(c, (int)num_intersects);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_intersects == 0 || d->num_points == 0) {
        JSA_CHECK_HIP(jsa::fill_zero(tile_bins, 8ull * p.num_tiles, s), "bin_rects_sort");
        return JAXSPLAT_OK;
    }
    if (!depths || !tile_rects || !cum_tiles_hit || !gaussian_ids_sorted)
        return fail(JAXSPLAT_ERR_INVALID, "bin_rects_sort: NULL required pointer");
    if (!workspace || workspace_bytes < p.emit_bytes)
        return fail(JAXSPLAT_ERR_WORKSPACE, "bin_rects_sort: workspace too small (%zu < %zu)",
                    workspace_bytes, p.emit_bytes);
    JSA_CHECK_HIP(jsa::bin_rects(p, c, depths, reinterpret_cast<const uint2*>(tile_rects),
                                 cum_tiles_hit, (int)num_intersects, gaussian_ids_sorted,
                                 tile_ids_sorted, tile_bins, workspace, s),
                  "bin_rects_sort");
    return JAXSPLAT_OK;
}

int jaxsplat_bin_check(const jaxsplat_descriptor* d, uint32_t num_intersects,
                       const void* workspace, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (num_intersects == 0 || d->num_points == 0 || !workspace) return JAXSPLAT_OK;
    const jsa::BinPlan p = jsa::bin_plan(camera_of(d), (int)num_intersects);
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t err = 0;
    JSA_CHECK_HIP(hipMemcpyAsync(&err, jsa::bin_error_word(p, workspace), 4,
                                 hipMemcpyDeviceToHost, s),
                  "bin_check");
    JSA_CHECK_HIP(hipStreamSynchronize(s), "bin_check");
    if (err) return fail(JAXSPLAT_ERR_CAPACITY, "binning error (code %u)", err);
    return JAXSPLAT_OK;
}

int jaxsplat_rasterize_forward_binned(const jaxsplat_descriptor* d,
                                      const int32_t* gaussian_ids_sorted,
                                      const int32_t* tile_bins, uint32_t num_intersects,
                                      const float* xys, const float* conics,
                                      const float* colors, const float* opacities,
                                      const float* background, float* final_Ts,
                                      int32_t* final_idx, float* out_img, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (!background || !final_Ts || !final_idx || !out_img)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_forward: NULL required pointer");
    const jsa::Camera c = camera_of(d);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_intersects == 0) {
        JSA_CHECK_HIP(jsa::raster_empty(c, background, final_Ts, final_idx, out_img, s),
                      "rasterize_forward");
        return JAXSPLAT_OK;
    }
    if (!gaussian_ids_sorted || !tile_bins || !xys || !conics || !colors || !opacities)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_forward: NULL required pointer");
    JSA_CHECK_HIP(jsa::raster_fwd(c, gaussian_ids_sorted, tile_bins, xys, conics, colors,
                                  opacities, background, final_Ts, final_idx, out_img, s),
                  "rasterize_forward");
    return JAXSPLAT_OK;
}

int jaxsplat_rasterize_backward_binned(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* xys, const float* conics, const float* colors,
    const float* opacities, const float* background, const float* final_Ts,
    const int32_t* final_idx, const float* v_out_img, const float* v_out_img_alpha,
    float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs, float* v_conic,
    int adjoint, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (adjoint != JAXSPLAT_ADJOINT_REFERENCE && adjoint != JAXSPLAT_ADJOINT_CONSISTENT)
        return fail(JAXSPLAT_ERR_INVALID, "unknown adjoint mode %d", adjoint);
    // (an empty gradient array, N = 0, may come as NULL)
    if (d->num_points > 0 && (!v_colors || !v_opacity || !v_xy || !v_conic))
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward: NULL gradient pointer");
    const jsa::Camera c = camera_of(d);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t n = d->num_points;
    float* ptrs[5] = {v_colors, v_opacity, v_xy, v_conic, v_xy_abs};
    const size_t cnt[5] = {3 * n, n, 2 * n, 3 * n, v_xy_abs ? 2 * n : 0};
    JSA_CHECK_HIP(jsa::zero_grads(ptrs, cnt, 5, s), "rasterize_backward");
    if (num_intersects == 0 || n == 0) return JAXSPLAT_OK;
    if (!gaussian_ids_sorted || !tile_bins || !xys || !conics || !colors || !opacities ||
        !background || !final_Ts || !final_idx || !v_out_img)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward: NULL required pointer");
    JSA_CHECK_HIP(jsa::raster_bwd(c, gaussian_ids_sorted, tile_bins, xys, conics, colors,
                                  opacities, background, final_Ts, final_idx, v_out_img,
                                  v_out_img_alpha, v_colors, v_opacity, v_xy, v_xy_abs, v_conic,
                                  adjoint == JAXSPLAT_ADJOINT_CONSISTENT, s),
                  "rasterize_backward");
    return JAXSPLAT_OK;
}

int jaxsplat_rasterize_forward_nd(const jaxsplat_descriptor* d,
                                  const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
                                  uint32_t num_intersects, const float* xys, const float* conics,
                                  const float* colors, const float* opacities,
                                  const float* background, uint32_t channels, float* final_Ts,
                                  int32_t* final_idx, float* out_img, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (channels < 1 || channels > 4096)
        return fail(JAXSPLAT_ERR_INVALID, "channels must be in [1, 4096], got %u", channels);
    if (!background || !final_Ts || !final_idx || !out_img)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_forward_nd: NULL required pointer");
    const jsa::Camera c = camera_of(d);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_intersects == 0) {
        JSA_CHECK_HIP(jsa::raster_empty(c, background, final_Ts, final_idx, out_img, s,
                                        (int)channels),
                      "rasterize_forward_nd");
        return JAXSPLAT_OK;
    }
    if (!gaussian_ids_sorted || !tile_bins || !xys || !conics || !colors || !opacities)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_forward_nd: NULL required pointer");
    jsa::RasterGrid g = jsa::single_grid(c);
    g.ch_stride = (int)channels;
    // channels in passes of up to three: the blend decisions are the same in every pass
    for (int off = 0; off < (int)channels; off += 3) {
        g.ch_off = off;
        g.ch_n = std::min(3, (int)channels - off);
        JSA_CHECK_HIP(jsa::raster_fwd_grid(g, gaussian_ids_sorted, tile_bins, xys, conics, colors,
                                           opacities, background, final_Ts, final_idx, out_img,
                                           s),
                      "rasterize_forward_nd");
    }
    return JAXSPLAT_OK;
}

int jaxsplat_rasterize_backward_nd(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* xys, const float* conics, const float* colors,
    const float* opacities, const float* background, uint32_t channels, const float* final_Ts,
    const int32_t* final_idx, const float* v_out_img, const float* v_out_img_alpha,
    float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs, float* v_conic,
    int adjoint, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (channels < 1 || channels > 4096)
        return fail(JAXSPLAT_ERR_INVALID, "channels must be in [1, 4096], got %u", channels);
    if (v_xy_abs && channels > 3)
        return fail(JAXSPLAT_ERR_INVALID,
                    "v_xy_abs is only available for up to 3 channels (it is not additive over "
                    "channel passes)");
    if (adjoint != JAXSPLAT_ADJOINT_REFERENCE && adjoint != JAXSPLAT_ADJOINT_CONSISTENT)
        return fail(JAXSPLAT_ERR_INVALID, "unknown adjoint mode %d", adjoint);
    if (d->num_points > 0 && (!v_colors || !v_opacity || !v_xy || !v_conic))
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward_nd: NULL gradient pointer");
    const jsa::Camera c = camera_of(d);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t n = d->num_points;
    float* ptrs[5] = {v_colors, v_opacity, v_xy, v_conic, v_xy_abs};
    const size_t cnt[5] = {(size_t)channels * n, n, 2 * n, 3 * n, v_xy_abs ? 2 * n : 0};
    JSA_CHECK_HIP(jsa::zero_grads(ptrs, cnt, 5, s), "rasterize_backward_nd");
    if (num_intersects == 0 || n == 0) return JAXSPLAT_OK;
    if (!gaussian_ids_sorted || !tile_bins || !xys || !conics || !colors || !opacities ||
        !background || !final_Ts || !final_idx || !v_out_img)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward_nd: NULL required pointer");
    jsa::RasterGrid g = jsa::single_grid(c);
    g.ch_stride = (int)channels;
    // geometry gradients are linear in v_alpha, which sums over channels: the
    // passes add into the same arrays; the alpha cotangent enters once
    for (int off = 0; off < (int)channels; off += 3) {
        g.ch_off = off;
        g.ch_n = std::min(3, (int)channels - off);
        JSA_CHECK_HIP(jsa::raster_bwd_grid(g, gaussian_ids_sorted, tile_bins, xys, conics, colors,
                                           opacities, background, final_Ts, final_idx, v_out_img,
                                           off == 0 ? v_out_img_alpha : nullptr, v_colors,
                                           v_opacity, v_xy, v_xy_abs, v_conic,
                                           adjoint == JAXSPLAT_ADJOINT_CONSISTENT, s),
                      "rasterize_backward_nd");
    }
    return JAXSPLAT_OK;
}

size_t jaxsplat_rasterize_backward_staged_workspace_size(const jaxsplat_descriptor* d,
                                                         uint32_t num_intersects) {
    if (!d || validate(d)) return 0;
    return align256(jsa::raster_bwd_staged_bytes((int)num_intersects));
}

int jaxsplat_rasterize_backward_staged(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* xys, const float* conics, const float* colors,
    const float* opacities, const float* background, uint32_t channels, const int32_t* radii,
    const int32_t* cum_tiles_hit, const float* final_Ts, const int32_t* final_idx,
    const float* v_out_img, const float* v_out_img_alpha, float* v_colors, float* v_opacity,
    float* v_xy, float* v_xy_abs, float* v_conic, int adjoint, void* workspace,
    size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (channels < 1 || channels > 3)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward_staged: channels must be in [1, 3], "
                    "got %u", channels);
    if (num_intersects >= (1u << 30))
        return fail(JAXSPLAT_ERR_CAPACITY, "num_intersects %u exceeds 2^30", num_intersects);
    if (adjoint != JAXSPLAT_ADJOINT_REFERENCE && adjoint != JAXSPLAT_ADJOINT_CONSISTENT)
        return fail(JAXSPLAT_ERR_INVALID, "unknown adjoint mode %d", adjoint);
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!v_colors || !v_opacity || !v_xy || !v_conic)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward_staged: NULL gradient pointer");
    if (!xys || !radii || !cum_tiles_hit ||
        (num_intersects > 0 && (!gaussian_ids_sorted || !tile_bins || !conics || !colors ||
                                !opacities || !background || !final_Ts || !final_idx ||
                                !v_out_img)))
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward_staged: NULL required pointer");
    const size_t need = align256(jsa::raster_bwd_staged_bytes((int)num_intersects));
    if (!workspace || workspace_bytes < need)
        return fail(JAXSPLAT_ERR_WORKSPACE,
                    "rasterize_backward_staged: workspace too small (%zu < %zu)",
                    workspace_bytes, need);
    JSA_CHECK_HIP(jsa::raster_bwd_staged(camera_of(d), (int)channels, gaussian_ids_sorted,
                                         tile_bins, (int)num_intersects, xys, conics, colors,
                                         opacities, background, final_Ts, final_idx, v_out_img,
                                         v_out_img_alpha, radii, nullptr, cum_tiles_hit,
                                         workspace, v_colors, v_opacity, v_xy, v_xy_abs, v_conic,
                                         adjoint == JAXSPLAT_ADJOINT_CONSISTENT,
                                         static_cast<hipStream_t>(stream)),
                  "rasterize_backward_staged");
    return JAXSPLAT_OK;
}

int jaxsplat_render_backward(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* means3d, const float* scales, const float* quats,
    const float* viewmat, const float* splats, const float* conics, const float* background,
    uint32_t channels, const int32_t* radii, const uint32_t* tile_rects,
    const int32_t* cum_tiles_hit, const float* final_Ts, const int32_t* final_idx,
    const float* v_out_img, float* v_colors, float* v_opacity, float* v_mean3d, float* v_scale,
    float* v_quat, int adjoint, void* workspace, size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (channels < 1 || channels > 3)
        return fail(JAXSPLAT_ERR_INVALID, "render_backward: channels must be in [1, 3], got %u",
                    channels);
    if (num_intersects >= (1u << 30))
        return fail(JAXSPLAT_ERR_CAPACITY, "num_intersects %u exceeds 2^30", num_intersects);
    if (adjoint != JAXSPLAT_ADJOINT_REFERENCE && adjoint != JAXSPLAT_ADJOINT_CONSISTENT)
        return fail(JAXSPLAT_ERR_INVALID, "unknown adjoint mode %d", adjoint);
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!v_colors || !v_opacity || !v_mean3d || !v_scale || !v_quat)
        return fail(JAXSPLAT_ERR_INVALID, "render_backward: NULL gradient pointer");
    if (!means3d || !scales || !quats || !viewmat || !splats || !conics || !radii ||
        !tile_rects || !cum_tiles_hit ||
        (num_intersects > 0 && (!gaussian_ids_sorted || !tile_bins || !background ||
                                !final_Ts || !final_idx || !v_out_img)))
        return fail(JAXSPLAT_ERR_INVALID, "render_backward: NULL required pointer");
    const size_t need = align256(jsa::raster_bwd_staged_bytes((int)num_intersects));
    if (!workspace || workspace_bytes < need)
        return fail(JAXSPLAT_ERR_WORKSPACE, "render_backward: workspace too small (%zu < %zu)",
                    workspace_bytes, need);
    const jsa::Camera c = camera_of(d);
    const uint2* rects = reinterpret_cast<const uint2*>(tile_rects);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int consistent = adjoint == JAXSPLAT_ADJOINT_CONSISTENT;
    JSA_CHECK_HIP(jsa::raster_bwd_staged(c, (int)channels, gaussian_ids_sorted, tile_bins,
                                         (int)num_intersects, nullptr, nullptr, nullptr, nullptr,
                                         background, final_Ts, final_idx, v_out_img, nullptr,
                                         radii, rects, cum_tiles_hit, workspace, nullptr, nullptr,
                                         nullptr, nullptr, nullptr, consistent, s, true,
                                         reinterpret_cast<const float4*>(splats)),
                  "render_backward (records)");
    JSA_CHECK_HIP(jsa::render_bwd_tail(c, (int)channels, (int)num_intersects, means3d, scales,
                                       quats, viewmat, radii, rects, cum_tiles_hit, conics,
                                       workspace, v_colors, v_opacity, v_mean3d, v_scale, v_quat,
                                       consistent, s),
                  "render_backward (tail)");
    return JAXSPLAT_OK;
}

int jaxsplat_rasterize_forward_splats(const jaxsplat_descriptor* d,
                                      const int32_t* gaussian_ids_sorted,
                                      const int32_t* tile_bins, uint32_t num_intersects,
                                      const float* splats, const float* background,
                                      uint32_t channels, float* final_Ts, int32_t* final_idx,
                                      float* out_img, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (channels < 1 || channels > 3)
        return fail(JAXSPLAT_ERR_INVALID,
                    "rasterize_forward_splats: channels must be in [1, 3], got %u", channels);
    if (!background || !final_Ts || !final_idx || !out_img)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_forward_splats: NULL required pointer");
    const jsa::Camera c = camera_of(d);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_intersects == 0 || d->num_points == 0) {
        JSA_CHECK_HIP(jsa::raster_empty(c, background, final_Ts, final_idx, out_img, s,
                                        (int)channels),
                      "rasterize_forward_splats");
        return JAXSPLAT_OK;
    }
    if (!gaussian_ids_sorted || !tile_bins || !splats)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_forward_splats: NULL required pointer");
    JSA_CHECK_HIP(jsa::raster_fwd_splats(c, (int)channels, gaussian_ids_sorted, tile_bins,
                                         reinterpret_cast<const float4*>(splats), background,
                                         final_Ts, final_idx, out_img, s),
                  "rasterize_forward_splats");
    return JAXSPLAT_OK;
}

int jaxsplat_rasterize_backward_staged_rects(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* xys, const float* conics, const float* colors,
    const float* opacities, const float* background, uint32_t channels, const int32_t* radii,
    const uint32_t* tile_rects, const int32_t* cum_tiles_hit, const float* final_Ts,
    const int32_t* final_idx, const float* v_out_img, const float* v_out_img_alpha,
    float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs, float* v_conic, int adjoint,
    void* workspace, size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = validate(d)) return st;
    if (channels < 1 || channels > 3)
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward_staged_rects: channels must be in "
                    "[1, 3], got %u", channels);
    if (num_intersects >= (1u << 30))
        return fail(JAXSPLAT_ERR_CAPACITY, "num_intersects %u exceeds 2^30", num_intersects);
    if (adjoint != JAXSPLAT_ADJOINT_REFERENCE && adjoint != JAXSPLAT_ADJOINT_CONSISTENT)
        return fail(JAXSPLAT_ERR_INVALID, "unknown adjoint mode %d", adjoint);
    if (d->num_points == 0) return JAXSPLAT_OK;
    if (!v_colors || !v_opacity || !v_xy || !v_conic)
        return fail(JAXSPLAT_ERR_INVALID,
                    "rasterize_backward_staged_rects: NULL gradient pointer");
    if (!xys || !radii || !tile_rects || !cum_tiles_hit ||
        (num_intersects > 0 && (!gaussian_ids_sorted || !tile_bins || !conics || !colors ||
                                !opacities || !background || !final_Ts || !final_idx ||
                                !v_out_img)))
        return fail(JAXSPLAT_ERR_INVALID, "rasterize_backward_staged_rects: NULL required pointer");
    const size_t need = align256(jsa::raster_bwd_staged_bytes((int)num_intersects));
    if (!workspace || workspace_bytes < need)
        return fail(JAXSPLAT_ERR_WORKSPACE,
                    "rasterize_backward_staged_rects: workspace too small (%zu < %zu)",
                    workspace_bytes, need);
    JSA_CHECK_HIP(jsa::raster_bwd_staged(camera_of(d), (int)channels, gaussian_ids_sorted,
                                         tile_bins, (int)num_intersects, xys, conics, colors,
                                         opacities, background, final_Ts, final_idx, v_out_img,
                                         v_out_img_alpha, radii,
                                         reinterpret_cast<const uint2*>(tile_rects),
                                         cum_tiles_hit, workspace, v_colors, v_opacity, v_xy,
                                         v_xy_abs, v_conic, adjoint == JAXSPLAT_ADJOINT_CONSISTENT,
                                         static_cast<hipStream_t>(stream)),
                  "rasterize_backward_staged_rects");
    return JAXSPLAT_OK;
}

// ---------------------------------------------------------------------------
// Reference-compatible custom-call targets
// ---------------------------------------------------------------------------
// The legacy signature carries no status, so a failing call must not leave
// its results holding whatever the buffers held before (XLA would consume
// them as valid): every result buffer is filled with kPoison, the bits of a
// quiet NaN (an out-of-range value in the integer results), and the message
// is kept in jaxsplat_last_error().  A descriptor of the wrong length cannot
// size the results; such a call only sets the message.

}  // extern "C"

namespace {

constexpr uint32_t kPoison = 0x7FC00000u;

struct Result {
    int index;          // position in `buffers`
    size_t words;       // 32-bit words
};

// Results can be sized from the descriptor unless its sizes are out of range
// (a descriptor that valid buffers could not have come with).
bool sizes_ok(const jaxsplat_descriptor* d) {
    return d->num_points <= (1u << 30) && d->img_shape.x <= (1u << 16) &&
           d->img_shape.y <= (1u << 16);
}

void poison(void* stream, void** buffers, const Result* res, int k,
            const jaxsplat_descriptor* d) {
    if (!sizes_ok(d)) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int i = 0; i < k; ++i)
        if (buffers[res[i].index] && res[i].words)
            (void)jsa::fill_words(buffers[res[i].index], res[i].words, kPoison, s);
}

int legacy_bin(const jaxsplat_descriptor* d, void* stream, const float* xys,
               const float* depths, const int32_t* radii, uint32_t isect, int32_t** gids,
               int32_t** bins) {
    const jsa::Camera c = camera_of(d);
    const size_t ntiles = (size_t)c.tiles_x * c.tiles_y;
    const size_t bin_ws = jaxsplat_bin_workspace_size(d, isect);
    const size_t gid_bytes = align256(4ull * isect);
    const size_t bins_bytes = align256(8ull * ntiles);
    int st = 0;
    uint8_t* ws = static_cast<uint8_t*>(cached_workspace(stream, bin_ws + gid_bytes + bins_bytes, &st));
    if (!ws) return st;
    *gids = reinterpret_cast<int32_t*>(ws + bin_ws);
    *bins = reinterpret_cast<int32_t*>(ws + bin_ws + gid_bytes);
    if ((st = jaxsplat_bin_and_sort(d, xys, depths, radii, isect, *gids, nullptr, *bins, ws,
                                    bin_ws, stream)))
        return st;
    // an overflow (cum_tiles_hit inconsistent with xys / radii) is an error here:
    // the caller has no other way to learn of it
    return jaxsplat_bin_check(d, isect, ws, stream);
}

// operands: means3d, scales, quats, viewmat; results: cov3d, xys, depths,
// radii, conics, compensation, num_tiles_hit, cum_tiles_hit (ops.cu:246-258)
int legacy_project_fwd(const jaxsplat_descriptor* d, void* stream, void** buffers) {
    if (int st = validate(d)) return st;
    int st = 0;
    void* ws = cached_workspace(stream, jaxsplat_project_workspace_size(d), &st);
    if (!ws) return st;
    auto F = [&](int i) { return static_cast<float*>(buffers[i]); };
    auto I = [&](int i) { return static_cast<int32_t*>(buffers[i]); };
    return jaxsplat_project_forward(d, F(0), F(1), F(2), F(3), F(4), F(5), F(6), I(7), F(8),
                                    F(9), I(10), I(11), ws, jaxsplat_project_workspace_size(d),
                                    stream);
}

// operands: colors, opacities, background, xys, depths, radii, conics,
// cum_tiles_hit; results: final_Ts, final_idx, out_img (ops.cu:408-419)
int legacy_rasterize_fwd(const jaxsplat_descriptor* d, void* stream, void** buffers) {
    if (int st = validate(d)) return st;
    auto F = [&](int i) { return static_cast<float*>(buffers[i]); };
    auto I = [&](int i) { return static_cast<int32_t*>(buffers[i]); };
    uint32_t isect = 0;
    if (int st = read_num_intersects(I(7), d->num_points, stream, &isect)) return st;
    int32_t *gids = nullptr, *bins = nullptr;
    if (isect > 0) {
        if (int st = legacy_bin(d, stream, F(3), F(4), I(5), isect, &gids, &bins)) return st;
    }
    return jaxsplat_rasterize_forward_binned(d, gids, bins, isect, F(3), F(6), F(0), F(1), F(2),
                                             F(8), I(9), F(10), stream);
}

// operands: colors, opacities, background, xys, depths, radii, conics,
// cum_tiles_hit, final_Ts, final_idx, v_out_img, v_out_img_alpha; results:
// v_colors, v_opacity, v_xy, v_xy_abs, v_conic (ops.cu:502-519)
int legacy_rasterize_bwd(const jaxsplat_descriptor* d, void* stream, void** buffers) {
    if (int st = validate(d)) return st;
    auto F = [&](int i) { return static_cast<float*>(buffers[i]); };
    auto I = [&](int i) { return static_cast<int32_t*>(buffers[i]); };
    uint32_t isect = 0;
    if (int st = read_num_intersects(I(7), d->num_points, stream, &isect)) return st;
    int32_t *gids = nullptr, *bins = nullptr;
    if (isect > 0) {
        if (int st = legacy_bin(d, stream, F(3), F(4), I(5), isect, &gids, &bins)) return st;
    }
    return jaxsplat_rasterize_backward_binned(d, gids, bins, isect, F(3), F(6), F(0), F(1), F(2),
                                              F(8), I(9), F(10), F(11), F(12), F(13), F(14),
                                              F(15), F(16), JAXSPLAT_ADJOINT_REFERENCE, stream);
}

}  // namespace

extern "C" {

void jaxsplat_project_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
    g_err.clear();
    const jaxsplat_descriptor* d = unpack(opaque, opaque_len);
    if (!d) return;
    if (legacy_project_fwd(d, stream, buffers)) {
        const size_t n = d->num_points;
        const Result res[] = {{4, 6 * n}, {5, 2 * n}, {6, n},     {7, n},
                              {8, 3 * n}, {9, n},     {10, n},    {11, n}};
        poison(stream, buffers, res, 8, d);
    }
}

void jaxsplat_project_bwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
    g_err.clear();
    const jaxsplat_descriptor* d = unpack(opaque, opaque_len);
    if (!d) return;
    // operands: means3d, scales, quats, viewmat, cov3ds, xys, radii, conics,
    // compensation, v_compensation, v_xy, v_depth, v_conic; results: v_mean3d,
    // v_scale, v_quat, v_cov2d, v_cov3d (ops.cu:336-354)
    auto F = [&](int i) { return static_cast<float*>(buffers[i]); };
    if (jaxsplat_project_backward(d, F(0), F(1), F(2), F(3), static_cast<int32_t*>(buffers[6]),
                                  F(7), F(8), F(10), F(11), F(12), F(9), F(13), F(14), F(15),
                                  F(16), F(17), JAXSPLAT_ADJOINT_REFERENCE, stream)) {
        const size_t n = d->num_points;
        const Result res[] = {{13, 3 * n}, {14, 3 * n}, {15, 4 * n}, {16, 3 * n}, {17, 6 * n}};
        poison(stream, buffers, res, 5, d);
    }
}

void jaxsplat_rasterize_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
    g_err.clear();
    const jaxsplat_descriptor* d = unpack(opaque, opaque_len);
    if (!d) return;
    if (legacy_rasterize_fwd(d, stream, buffers)) {
        const size_t px = (size_t)d->img_shape.x * d->img_shape.y;
        const Result res[] = {{8, px}, {9, px}, {10, 3 * px}};
        poison(stream, buffers, res, 3, d);
    }
}

void jaxsplat_rasterize_bwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
    g_err.clear();
    const jaxsplat_descriptor* d = unpack(opaque, opaque_len);
    if (!d) return;
    if (legacy_rasterize_bwd(d, stream, buffers)) {
        const size_t n = d->num_points;
        const Result res[] = {{12, 3 * n}, {13, n}, {14, 2 * n}, {15, 2 * n}, {16, 3 * n}};
        poison(stream, buffers, res, 5, d);
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched rendering
// ---------------------------------------------------------------------------

struct jaxsplat_batch {
    int device = 0;
    int num_scenes = 0, num_views = 0;
    int n_items = 0, n_gauss = 0, max_view_n = 0, max_scene_n = 0;
    int total_tiles = 0, max_view_tiles = 0;
    long long num_pixels = 0;
    float glob_scale = 1.f, clip_thresh = 0.f;
    int block_width = 16;
    int fwd_depth = 0, bwd_depth = 0;     // Gaussians per LDS batch (0: the defaults)
    int schedule = JAXSPLAT_SCHEDULE_GRID;
    jsa::BatchBinPlan bin{};
    std::vector<jsa::ViewRec> views;
    std::vector<jsa::SceneRec> scenes;
    jsa::ViewRec* d_views = nullptr;
    jsa::SceneRec* d_scenes = nullptr;
    // workspace layout
    size_t xys = 0, conics = 0, radii = 0, dkeys = 0, rects = 0, v_xy = 0, v_conic = 0,
           items = 0, bins = 0, bin_scratch = 0, item_first = 0, v_rgbo = 0, staged = 0,
           order = 0, okeys = 0, okeys_out = 0, otmp_k = 0, otmp_v = 0, osort = 0, qcount = 0,
           ws_bytes = 0;
    int isect_cap = 0;
};

namespace {

// The batch's rasterizer grid; with the work-queue schedule, the tile order
// and counters live in the workspace w.
jsa::RasterGrid batch_grid(const jaxsplat_batch* b, uint8_t* w) {
    jsa::RasterGrid g{};
    g.v0 = b->views[0];
    g.views = b->d_views;
    g.num_views = b->num_views;
    g.total_tiles = b->total_tiles;
    g.block_width = b->block_width;
    g.n_items = b->n_items;
    g.fwd_depth = b->fwd_depth;
    g.bwd_depth = b->bwd_depth;
    g.ch_stride = 3;
    g.ch_off = 0;
    g.ch_n = 3;
    if (b->schedule == JAXSPLAT_SCHEDULE_QUEUE) {
        g.order = reinterpret_cast<const int*>(w + b->order);
        g.queue_counters = reinterpret_cast<int*>(w + b->qcount);
    }
    return g;
}

int check_batch_ws(const jaxsplat_batch* b, const void* ws, size_t bytes, const char* where) {
    if (!b) return fail(JAXSPLAT_ERR_INVALID, "%s: batch is NULL", where);
    if (!ws || bytes < b->ws_bytes)
        return fail(JAXSPLAT_ERR_WORKSPACE, "%s: workspace too small (%zu < %zu)", where, bytes,
                    b->ws_bytes);
    return JAXSPLAT_OK;
}

}  // namespace

extern "C" {

int jaxsplat_batch_create(jaxsplat_batch** out, const jaxsplat_scene* scenes, int32_t num_scenes,
                          const jaxsplat_view* views, int32_t num_views, float glob_scale,
                          float clip_thresh, uint32_t block_width, int64_t isect_capacity) {
    g_err.clear();
    if (!out) return fail(JAXSPLAT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!scenes || num_scenes < 1) return fail(JAXSPLAT_ERR_INVALID, "need at least one scene");
    if (!views || num_views < 1 || num_views > 4096)
        return fail(JAXSPLAT_ERR_INVALID, "num_views must be in [1, 4096], got %d", num_views);
    if (block_width < 1 || block_width > 16)
        return fail(JAXSPLAT_ERR_INVALID, "block_width must be in [1, 16], got %u", block_width);
    if (!(clip_thresh >= 0.f)) return fail(JAXSPLAT_ERR_INVALID, "clip_thresh must be >= 0");
    if (isect_capacity < 1 || isect_capacity > 0x7FFFFF00LL)
        return fail(JAXSPLAT_ERR_INVALID, "isect_capacity must be in [1, 2^31 - 256]");
    auto b = new jaxsplat_batch();
    b->num_scenes = num_scenes;
    b->num_views = num_views;
    b->glob_scale = glob_scale;
    b->clip_thresh = clip_thresh;
    b->block_width = (int)block_width;
    long long g = 0;
    for (int s = 0; s < num_scenes; ++s) {
        if (scenes[s].num_gaussians < 0 || scenes[s].first_gaussian != g) {
            delete b;
            return fail(JAXSPLAT_ERR_INVALID,
                        "scene %d: scenes must hold consecutive rows (first_gaussian %d, "
                        "expected %lld)", s, scenes[s].first_gaussian, g);
        }
        g += scenes[s].num_gaussians;
        jsa::SceneRec r{scenes[s].first_gaussian, scenes[s].num_gaussians, 0, 0};
        b->scenes.push_back(r);
        b->max_scene_n = std::max(b->max_scene_n, scenes[s].num_gaussians);
    }
    if (g > (1LL << 30)) {
        delete b;
        return fail(JAXSPLAT_ERR_INVALID, "too many Gaussians (%lld)", g);
    }
    b->n_gauss = (int)g;
    long long items = 0, tiles = 0, pix = 0;
    int prev_scene = 0;
    for (int s = 0; s < num_scenes; ++s) b->scenes[s].v_begin = b->scenes[s].v_end = 0;
    for (int v = 0; v < num_views; ++v) {
        const jaxsplat_view& w = views[v];
        if (w.scene < prev_scene || w.scene >= num_scenes) {
            delete b;
            return fail(JAXSPLAT_ERR_INVALID, "view %d: views must be grouped by ascending scene",
                        v);
        }
        if (w.img_h < 1 || w.img_w < 1 || !(w.fx > 0.f) || !(w.fy > 0.f)) {
            delete b;
            return fail(JAXSPLAT_ERR_INVALID, "view %d: bad image size or focal length", v);
        }
        if (w.scene != prev_scene || v == 0) b->scenes[w.scene].v_begin = v;
        b->scenes[w.scene].v_end = v + 1;
        prev_scene = w.scene;
        jsa::ViewRec r{};
        r.img_w = w.img_w;
        r.img_h = w.img_h;
        r.tiles_x = (w.img_w + (int)block_width - 1) / (int)block_width;
        r.tiles_y = (w.img_h + (int)block_width - 1) / (int)block_width;
        r.fx = w.fx;
        r.fy = w.fy;
        r.cx = w.cx;
        r.cy = w.cy;
        r.tan_fovx = (float)(0.5 * (double)w.img_w / (double)w.fx);
        r.tan_fovy = (float)(0.5 * (double)w.img_h / (double)w.fy);
        r.g0 = scenes[w.scene].first_gaussian;
        r.n = scenes[w.scene].num_gaussians;
        r.item_off = (int)items;
        r.tile_off = (int)tiles;
        r.view = v;
        r.pix_off = pix;
        items += r.n;
        tiles += (long long)r.tiles_x * r.tiles_y;
        pix += (long long)w.img_w * w.img_h;
        b->max_view_n = std::max(b->max_view_n, r.n);
        b->max_view_tiles = std::max(b->max_view_tiles, r.tiles_x * r.tiles_y);
        b->views.push_back(r);
    }
    if (items > 0x7FFFFF00LL || tiles > (1LL << 30)) {
        delete b;
        return fail(JAXSPLAT_ERR_INVALID, "batch too large (%lld items, %lld tiles)", items,
                    tiles);
    }
    b->n_items = (int)items;
    b->total_tiles = (int)tiles;
    b->num_pixels = pix;
    b->bin = jsa::batch_bin_plan(b->n_items, num_views, b->max_view_tiles, (int)isect_capacity);
    if (b->bin.tile_bits + b->bin.view_bits > 31) {
        delete b;
        return fail(JAXSPLAT_ERR_INVALID, "views x tiles exceed 2^31 bins");
    }
    for (int v = 0; v < num_views; ++v) b->views[v].bin_off = v << b->bin.tile_bits;
    // workspace: item rows, backward item gradients, sorted ids, bins, binning
    // scratch, each item's first intersection slot and the staged backward's
    // records (48 bytes per intersection slot) and slot bitmap
    b->isect_cap = (int)isect_capacity;
    size_t o = 0;
    const size_t n = (size_t)(b->n_items > 0 ? b->n_items : 1);
    b->xys = o; o = align256(o + 8 * n);
    b->conics = o; o = align256(o + 12 * n);
    b->radii = o; o = align256(o + 4 * n);
    b->dkeys = o; o = align256(o + 4 * n);
    b->rects = o; o = align256(o + 8 * n);
    b->v_xy = o; o = align256(o + 8 * n);
    b->v_conic = o; o = align256(o + 12 * n);
    b->items = o; o = align256(o + 4 * (size_t)isect_capacity);
    b->bins = o; o = align256(o + 8 * (size_t)b->bin.num_bins);
    b->bin_scratch = o; o = align256(o + b->bin.scratch_bytes);
    b->item_first = o; o = align256(o + 4 * n);
    b->v_rgbo = o; o = align256(o + 16 * n);
    b->staged = o; o = align256(o + jsa::raster_bwd_staged_bytes((int)isect_capacity));
    // work-queue schedule: the tile order, its sort buffers and two counters
    const size_t t4 = 4 * (size_t)(b->total_tiles > 0 ? b->total_tiles : 1);
    b->order = o; o = align256(o + t4);
    b->okeys = o; o = align256(o + t4);
    b->okeys_out = o; o = align256(o + t4);
    b->otmp_k = o; o = align256(o + t4);
    b->otmp_v = o; o = align256(o + t4);
    b->osort = o; o = align256(o + jsa::order_sort_plan(b->total_tiles).scratch_bytes);
    b->qcount = o; o = align256(o + 16);
    b->ws_bytes = o;
    // device tables
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipMalloc(&b->d_views, sizeof(jsa::ViewRec) * num_views);
    if (e == hipSuccess) e = hipMalloc(&b->d_scenes, sizeof(jsa::SceneRec) * num_scenes);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_views, b->views.data(), sizeof(jsa::ViewRec) * num_views,
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_scenes, b->scenes.data(), sizeof(jsa::SceneRec) * num_scenes,
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        jaxsplat_batch_destroy(b);
        return hip_fail(e, "batch_create");
    }
    *out = b;
    return JAXSPLAT_OK;
}

void jaxsplat_batch_destroy(jaxsplat_batch* b) {
    if (!b) return;
    if (b->d_views) (void)hipFree(b->d_views);
    if (b->d_scenes) (void)hipFree(b->d_scenes);
    delete b;
}

size_t jaxsplat_batch_workspace_size(const jaxsplat_batch* b) { return b ? b->ws_bytes : 0; }

int jaxsplat_batch_set_depths(jaxsplat_batch* b, int32_t fwd_depth, int32_t bwd_depth) {
    g_err.clear();
    if (!b) return fail(JAXSPLAT_ERR_INVALID, "batch is NULL");
    auto ok = [](int d) { return d == 0 || d == 64 || d == 128 || d == 256; };
    if (!ok(fwd_depth) || !ok(bwd_depth))
        return fail(JAXSPLAT_ERR_INVALID, "depths must be 0, 64, 128 or 256 (got %d, %d)",
                    fwd_depth, bwd_depth);
    b->fwd_depth = fwd_depth;
    b->bwd_depth = bwd_depth;
    return JAXSPLAT_OK;
}
int jaxsplat_batch_set_schedule(jaxsplat_batch* b, int32_t schedule) {
    g_err.clear();
    if (!b) return fail(JAXSPLAT_ERR_INVALID, "batch is NULL");
    if (schedule != JAXSPLAT_SCHEDULE_GRID && schedule != JAXSPLAT_SCHEDULE_QUEUE)
        return fail(JAXSPLAT_ERR_INVALID, "unknown schedule %d", schedule);
    b->schedule = schedule;
    return JAXSPLAT_OK;
}
int64_t jaxsplat_batch_num_pixels(const jaxsplat_batch* b) { return b ? b->num_pixels : 0; }
int64_t jaxsplat_batch_num_gaussians(const jaxsplat_batch* b) { return b ? b->n_gauss : 0; }
int64_t jaxsplat_batch_pixel_offset(const jaxsplat_batch* b, int32_t view) {
    if (!b || view < 0 || view > b->num_views) return -1;
    return view == b->num_views ? b->num_pixels : b->views[view].pix_off;
}

int jaxsplat_batch_forward(const jaxsplat_batch* b, const float* means3d, const float* scales,
                           const float* quats, const float* colors, const float* opacities,
                           const float* viewmats, const float* backgrounds, float* final_Ts,
                           int32_t* final_idx, float* out_img, void* workspace,
                           size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = check_batch_ws(b, workspace, workspace_bytes, "batch_forward")) return st;
    if (!viewmats || !backgrounds || !final_Ts || !final_idx || !out_img ||
        (b->n_gauss > 0 && (!means3d || !scales || !quats || !colors || !opacities)))
        return fail(JAXSPLAT_ERR_INVALID, "batch_forward: NULL required pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint8_t* w = static_cast<uint8_t*>(workspace);
    float* xys = reinterpret_cast<float*>(w + b->xys);
    float* conics = reinterpret_cast<float*>(w + b->conics);
    int* radii = reinterpret_cast<int*>(w + b->radii);
    uint32_t* dkeys = reinterpret_cast<uint32_t*>(w + b->dkeys);
    uint2* rects = reinterpret_cast<uint2*>(w + b->rects);
    int* items = reinterpret_cast<int*>(w + b->items);
    int* bins = reinterpret_cast<int*>(w + b->bins);
    JSA_CHECK_HIP(jsa::project_batch_fwd(b->d_views, b->num_views, b->max_view_n, b->glob_scale,
                                         b->clip_thresh, b->block_width, means3d, scales, quats,
                                         viewmats, opacities, xys, conics, radii, dkeys, rects,
                                         s),
                  "batch_forward: project");
    JSA_CHECK_HIP(jsa::batch_bin(b->bin, b->d_views, rects, dkeys, items, bins,
                                 reinterpret_cast<uint32_t*>(w + b->item_first),
                                 w + b->bin_scratch, s),
                  "batch_forward: bin");
    const jsa::RasterGrid g = batch_grid(b, w);
    if (g.order) {
        // the work queue's tile order, longest lists first (the backward reuses it)
        auto U = [&](size_t off) { return reinterpret_cast<uint32_t*>(w + off); };
        JSA_CHECK_HIP(jsa::tile_order(g, bins, U(b->okeys), U(b->okeys_out), U(b->otmp_k),
                                      U(b->otmp_v), reinterpret_cast<int*>(w + b->order),
                                      w + b->osort, s),
                      "batch_forward: tile order");
    }
    JSA_CHECK_HIP(jsa::raster_fwd_grid(g, items, bins, xys, conics, colors, opacities,
                                       backgrounds, final_Ts, final_idx, out_img, s),
                  "batch_forward: rasterize");
    return JAXSPLAT_OK;
}

int jaxsplat_batch_backward(const jaxsplat_batch* b, const float* means3d, const float* scales,
                            const float* quats, const float* colors, const float* opacities,
                            const float* viewmats, const float* backgrounds,
                            const float* final_Ts, const int32_t* final_idx,
                            const float* v_out_img, float* v_means3d, float* v_scales,
                            float* v_quats, float* v_colors, float* v_opacities, int adjoint,
                            void* workspace, size_t workspace_bytes, void* stream) {
    g_err.clear();
    if (int st = check_batch_ws(b, workspace, workspace_bytes, "batch_backward")) return st;
    if (adjoint != JAXSPLAT_ADJOINT_REFERENCE && adjoint != JAXSPLAT_ADJOINT_CONSISTENT)
        return fail(JAXSPLAT_ERR_INVALID, "unknown adjoint mode %d", adjoint);
    if (!viewmats || !backgrounds || !final_Ts || !final_idx || !v_out_img ||
        (b->n_gauss > 0 && (!means3d || !scales || !quats || !colors || !opacities ||
                            !v_means3d || !v_scales || !v_quats || !v_colors || !v_opacities)))
        return fail(JAXSPLAT_ERR_INVALID, "batch_backward: NULL required pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint8_t* w = static_cast<uint8_t*>(workspace);
    const float* xys = reinterpret_cast<const float*>(w + b->xys);
    const float* conics = reinterpret_cast<const float*>(w + b->conics);
    const int* radii = reinterpret_cast<const int*>(w + b->radii);
    float* v_xy = reinterpret_cast<float*>(w + b->v_xy);
    float* v_conic = reinterpret_cast<float*>(w + b->v_conic);
    const int* items = reinterpret_cast<const int*>(w + b->items);
    const int* bins = reinterpret_cast<const int*>(w + b->bins);
    const uint32_t* item_first = reinterpret_cast<const uint32_t*>(w + b->item_first);
    float* v_rgbo = reinterpret_cast<float*>(w + b->v_rgbo);
    const int cons = adjoint == JAXSPLAT_ADJOINT_CONSISTENT;
    // staged: per-(item, tile) records summed per item (v_xy, v_conic, v_rgbo),
    // then per Gaussian over the views in the projection backward; every
    // gradient row is written, no float atomics
    JSA_CHECK_HIP(jsa::raster_bwd_batch_staged(batch_grid(b, w), b->max_view_n, items, bins,
                                               b->isect_cap, xys, conics, colors, opacities,
                                               backgrounds, final_Ts, final_idx, v_out_img,
                                               radii, reinterpret_cast<const uint2*>(w + b->rects),
                                               item_first, w + b->staged, v_xy, v_conic, v_rgbo,
                                               cons, s),
                  "batch_backward: rasterize");
    JSA_CHECK_HIP(jsa::project_batch_bwd(b->d_scenes, b->num_scenes, b->max_scene_n, b->d_views,
                                         b->glob_scale, b->clip_thresh, means3d, scales, quats,
                                         viewmats, radii, conics, v_xy, v_conic, v_rgbo,
                                         v_means3d, v_scales, v_quats, v_colors, v_opacities,
                                         cons, s),
                  "batch_backward: project");
    return JAXSPLAT_OK;
}

const uint32_t* jaxsplat_batch_status_word(const jaxsplat_batch* b, const void* workspace) {
    if (!b || !workspace) return nullptr;
    return jsa::batch_bin_error_word(
        b->bin, static_cast<const uint8_t*>(workspace) + b->bin_scratch);
}

int jaxsplat_batch_check(const jaxsplat_batch* b, const void* workspace,
                         int64_t* num_intersects, void* stream) {
    g_err.clear();
    if (int st = check_batch_ws(b, workspace, b ? b->ws_bytes : 0, "batch_check")) return st;
    const uint8_t* w = static_cast<const uint8_t*>(workspace) + b->bin_scratch;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t err = 0;
    unsigned long long total = 0;
    JSA_CHECK_HIP(hipMemcpyAsync(&err, jsa::batch_bin_error_word(b->bin, w), 4,
                                 hipMemcpyDeviceToHost, s),
                  "batch_check");
    JSA_CHECK_HIP(hipMemcpyAsync(&total, jsa::batch_bin_total(b->bin, w), 8,
                                 hipMemcpyDeviceToHost, s),
                  "batch_check");
    JSA_CHECK_HIP(hipStreamSynchronize(s), "batch_check");
    if (num_intersects) *num_intersects = (int64_t)total;
    if (err & 8u)
        return fail(JAXSPLAT_ERR_CAPACITY, "intersections %llu exceed the batch capacity %d",
                    total, b->bin.isect_cap);
    if (err) return fail(JAXSPLAT_ERR_CAPACITY, "binning error (code %u)", err);
    return JAXSPLAT_OK;
}

}  // extern "C"
