// extern "C" entry points of libjaxsplat_amd.so (declared in
// include/jaxsplat_amd.h): argument validation, descriptor -> camera
// conversion, workspace carving, and the four reference-compatible XLA
// custom-call targets.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "../../include/jaxsplat_amd.h"
#include "common.hpp"

namespace jsa {
hipError_t raster_empty(const Camera& c, const float* background, float* final_Ts,
                        int* final_idx, float* out_img, hipStream_t s);
hipError_t zero_grads(float* const* ptrs, const size_t* counts, int k, hipStream_t s);
}  // namespace jsa

static_assert(sizeof(jaxsplat_descriptor) == 80, "descriptor must match the reference layout");
static_assert(offsetof(jaxsplat_descriptor, intrins) == 16, "descriptor layout");
static_assert(offsetof(jaxsplat_descriptor, block_dim_2d) == 64, "descriptor layout");

namespace {

thread_local std::string g_err;

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
                                   conics, compensation, num_tiles_hit, cum_tiles_hit, workspace,
                                   static_cast<hipStream_t>(stream)),
                  "project_forward");
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
    JSA_CHECK_HIP(jsa::bin_and_sort(p, c, xys, depths, radii, (int)num_intersects, nullptr,
                                    gaussian_ids_sorted, tile_ids_sorted, tile_bins, workspace, s),
                  "bin_and_sort");
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
    if (err) return fail(JAXSPLAT_ERR_CAPACITY, "look-back wait limit exceeded (code %u)", err);
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
    if (!v_colors || !v_opacity || !v_xy || !v_conic)
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

// ---------------------------------------------------------------------------
// Reference-compatible custom-call targets
// ---------------------------------------------------------------------------

void jaxsplat_project_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
    g_err.clear();
    const jaxsplat_descriptor* d = unpack(opaque, opaque_len);
    if (!d) return;
    int st = 0;
    void* ws = cached_workspace(stream, jaxsplat_project_workspace_size(d), &st);
    if (!ws) return;
    // operands: means3d, scales, quats, viewmat; results: cov3d, xys, depths,
    // radii, conics, compensation, num_tiles_hit, cum_tiles_hit (ops.cu:246-258)
    auto F = [&](int i) { return static_cast<float*>(buffers[i]); };
    auto I = [&](int i) { return static_cast<int32_t*>(buffers[i]); };
    jaxsplat_project_forward(d, F(0), F(1), F(2), F(3), F(4), F(5), F(6), I(7), F(8), F(9), I(10),
                             I(11), ws, jaxsplat_project_workspace_size(d), stream);
}

void jaxsplat_project_bwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
    g_err.clear();
    const jaxsplat_descriptor* d = unpack(opaque, opaque_len);
    if (!d) return;
    // operands: means3d, scales, quats, viewmat, cov3ds, xys, radii, conics,
    // compensation, v_compensation, v_xy, v_depth, v_conic; results: v_mean3d,
    // v_scale, v_quat, v_cov2d, v_cov3d (ops.cu:336-354)
    auto F = [&](int i) { return static_cast<float*>(buffers[i]); };
    jaxsplat_project_backward(d, F(0), F(1), F(2), F(3), static_cast<int32_t*>(buffers[6]), F(7),
                              F(8), F(10), F(11), F(12), F(9), F(13), F(14), F(15), F(16), F(17),
                              JAXSPLAT_ADJOINT_REFERENCE, stream);
}

static int legacy_bin(const jaxsplat_descriptor* d, void* stream, const float* xys,
                      const float* depths, const int32_t* radii, uint32_t isect,
                      int32_t** gids, int32_t** bins) {
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
    return jaxsplat_bin_and_sort(d, xys, depths, radii, isect, *gids, nullptr, *bins, ws, bin_ws,
                                 stream);
}

void jaxsplat_rasterize_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
    g_err.clear();
    const jaxsplat_descriptor* d = unpack(opaque, opaque_len);
    if (!d || validate(d)) return;
    // operands: colors, opacities, background, xys, depths, radii, conics,
    // cum_tiles_hit; results: final_Ts, final_idx, out_img (ops.cu:408-419)
    auto F = [&](int i) { return static_cast<float*>(buffers[i]); };
    auto I = [&](int i) { return static_cast<int32_t*>(buffers[i]); };
    uint32_t isect = 0;
    if (read_num_intersects(I(7), d->num_points, stream, &isect)) return;
    int32_t *gids = nullptr, *bins = nullptr;
    if (isect > 0 && legacy_bin(d, stream, F(3), F(4), I(5), isect, &gids, &bins)) return;
    jaxsplat_rasterize_forward_binned(d, gids, bins, isect, F(3), F(6), F(0), F(1), F(2), F(8),
                                      I(9), F(10), stream);
}

void jaxsplat_rasterize_bwd(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
    g_err.clear();
    const jaxsplat_descriptor* d = unpack(opaque, opaque_len);
    if (!d || validate(d)) return;
    // operands: colors, opacities, background, xys, depths, radii, conics,
    // cum_tiles_hit, final_Ts, final_idx, v_out_img, v_out_img_alpha; results:
    // v_colors, v_opacity, v_xy, v_xy_abs, v_conic (ops.cu:502-519)
    auto F = [&](int i) { return static_cast<float*>(buffers[i]); };
    auto I = [&](int i) { return static_cast<int32_t*>(buffers[i]); };
    uint32_t isect = 0;
    if (read_num_intersects(I(7), d->num_points, stream, &isect)) return;
    int32_t *gids = nullptr, *bins = nullptr;
    if (isect > 0 && legacy_bin(d, stream, F(3), F(4), I(5), isect, &gids, &bins)) return;
    jaxsplat_rasterize_backward_binned(d, gids, bins, isect, F(3), F(6), F(0), F(1), F(2), F(8),
                                       I(9), F(10), F(11), F(12), F(13), F(14), F(15), F(16),
                                       JAXSPLAT_ADJOINT_REFERENCE, stream);
}

}  // extern "C"
