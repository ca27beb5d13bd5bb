/*
 * jaxsplat_amd -- C ABI of the MI355X (gfx950) differentiable Gaussian-splatting
 * path: projection, tile binning/sort and rasterization, forward and backward.
 *
 * Every pointer argument is a device pointer (HBM) unless stated otherwise;
 * `stream` is a hipStream_t (NULL = default stream).  Arrays are dense,
 * row-major float32/int32 in the layout of the reference's JAX API
 * (yklcs/jaxsplat jaxsplat/_project/abstract.py:7-21,57-77 and
 * jaxsplat/_rasterize/abstract.py:7-24,62-85).  No call allocates device
 * memory except the four reference-compatible custom-call targets, which keep
 * a grow-only workspace per (device, stream) because their signature carries
 * none.  All typed entry points return a jaxsplat_status; on error
 * jaxsplat_last_error() returns a thread-local message.
 */
#ifndef JAXSPLAT_AMD_H
#define JAXSPLAT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#define JAXSPLAT_ALIGN16 alignas(16)
#else
#define JAXSPLAT_ALIGN16 _Alignas(16)
#endif

typedef enum jaxsplat_status {
    JAXSPLAT_OK = 0,
    JAXSPLAT_ERR_INVALID = 1,   /* bad descriptor / argument */
    JAXSPLAT_ERR_HIP = 2,       /* HIP runtime error */
    JAXSPLAT_ERR_CAPACITY = 3,  /* intersection overflow or a binning error */
    JAXSPLAT_ERR_WORKSPACE = 4  /* workspace missing or too small */
} jaxsplat_status;

/* Adjoint convention of the backward entry points. */
typedef enum jaxsplat_adjoint {
    JAXSPLAT_ADJOINT_REFERENCE = 0,  /* bit-for-bit reference semantics */
    JAXSPLAT_ADJOINT_CONSISTENT = 1  /* exact adjoint of the forward */
} jaxsplat_adjoint;

typedef struct jaxsplat_dim3 {
    uint32_t x, y, z;
} jaxsplat_dim3;

/*
 * Opaque call descriptor.  Same 80-byte layout as the reference's
 * ops::Descriptor (lib/ops.h:9-22) packed by make_descriptor
 * (lib/ffi.cu:15-52): img_shape is (W, H, 1), intrins is (fx, fy, cx, cy),
 * grid_dim_2d is the tile grid.
 */
typedef struct jaxsplat_descriptor {
    uint32_t num_points;
    jaxsplat_dim3 img_shape;
    JAXSPLAT_ALIGN16 float intrins[4];
    float glob_scale;
    float clip_thresh;
    uint32_t block_width;
    uint32_t grid_dim_1d;
    uint32_t block_dim_1d;
    jaxsplat_dim3 grid_dim_2d;
    jaxsplat_dim3 block_dim_2d;
} jaxsplat_descriptor;

/* Library version string (host memory). */
const char* jaxsplat_version(void);
/* Thread-local message of the last failing call, "" if none (host memory). */
const char* jaxsplat_last_error(void);

/*
 * Profiling marks (no counterpart in the reference, which has no tracing).
 * events: up to 4 hipEvent_t (NULL entries allowed) that later calls from
 * this thread record, on their stream, at internal stage boundaries:
 * [0] in jaxsplat_bin_emit_sort / jaxsplat_bin_and_sort after the emission,
 * [1] there after the tile sort (before the bin edges), [2] in the staged
 * rasterize backward after the per-intersection records (before the
 * per-Gaussian reduction), [3] in the binning after the bin edges (before
 * the per-tile depth order).  events == NULL (or count 0) clears them.  The
 * caller keeps the events alive while they are set.
 */
void jaxsplat_set_stage_marks(void* const* events, int count);

/*
 * Fill *out; replaces make_descriptor (lib/ffi.cu:15-52).  img_h/img_w are the
 * reference's img_shape (H, W).  Rejects block_width outside [1, 16] (the
 * reference's LDS arrays hold 256 entries, lib/common.h:9) and a negative
 * clip_thresh (reference keys would sign-extend into the tile id,
 * lib/kernels/forward.cu:168-173).
 */
int jaxsplat_make_descriptor(jaxsplat_descriptor* out, uint32_t num_points, uint32_t img_h,
                             uint32_t img_w, float fx, float fy, float cx, float cy,
                             float glob_scale, float clip_thresh, uint32_t block_width);

/* ------------------------------------------------------------------------ */
/* Reference-compatible XLA custom-call targets.                             */
/* Signature of ops::*::xla (lib/ops.h:45-95), registered by the reference   */
/* as "project_fwd", "project_bwd", "rasterize_fwd", "rasterize_bwd"         */
/* (lib/ffi.cu:4-13).  `buffers` holds the operands then the results in the   */
/* reference lowering order (lib/ops.cu:246-258, 336-354, 408-419,           */
/* 502-519); `opaque` is a jaxsplat_descriptor.  The signature has no        */
/* status: on any error the message is kept in jaxsplat_last_error() and     */
/* every result buffer is filled with 0x7FC00000 on `stream` (a quiet NaN in */
/* the float results, an out-of-range value in the integer ones), so that no */
/* stale memory passes as a result; only a descriptor of the wrong length,   */
/* which cannot size the results, leaves them untouched.  The two rasterize  */
/* targets also report a binning overflow (cum_tiles_hit inconsistent with   */
/* xys and radii) this way, after a synchronisation of the stream.           */
/* ------------------------------------------------------------------------ */
void jaxsplat_project_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len);
void jaxsplat_project_bwd(void* stream, void** buffers, const char* opaque, size_t opaque_len);
void jaxsplat_rasterize_fwd(void* stream, void** buffers, const char* opaque, size_t opaque_len);
void jaxsplat_rasterize_bwd(void* stream, void** buffers, const char* opaque, size_t opaque_len);

/* ------------------------------------------------------------------------ */
/* Typed entry points (caller-owned workspace, nothing allocated).           */
/* ------------------------------------------------------------------------ */

/* Workspace bytes for jaxsplat_project_forward. */
size_t jaxsplat_project_workspace_size(const jaxsplat_descriptor* d);

/*
 * Projection forward; replaces ops::project::fwd::xla (lib/ops.cu:12-61) and
 * project_gaussians_fwd (lib/kernels/forward.cu:13-131) + the CUB inclusive
 * scan (lib/ops.cu:607-643).  Shapes: means3d/scales (N,3), quats (N,4),
 * viewmat (4,4); outputs cov3d (N,6) [nullable], xys (N,2), depths (N),
 * radii (N), conics (N,3), compensation (N) [nullable], num_tiles_hit (N),
 * cum_tiles_hit (N, inclusive).  Every output row is written (culled
 * Gaussians get the reference's zero values).
 */
int jaxsplat_project_forward(const jaxsplat_descriptor* d, const float* means3d,
                             const float* scales, const float* quats, const float* viewmat,
                             float* cov3d, float* xys, float* depths, int32_t* radii,
                             float* conics, float* compensation, int32_t* num_tiles_hit,
                             int32_t* cum_tiles_hit, void* workspace, size_t workspace_bytes,
                             void* stream);

/*
 * Projection forward that allocates nothing and takes no workspace.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as jaxsplat_project_forward
 * without the workspace; the scan's scratch is a library-held device array
 * of 512 KiB per GPU, so num_points is at most 2^26 (JAXSPLAT_ERR_CAPACITY
 * above).  All calls on one GPU share that array: calls that overlap on one
 * GPU (other streams, other threads) race and must be serialised by the
 * caller.
 */
int jaxsplat_project_forward_noalloc(const jaxsplat_descriptor* d, const float* means3d,
                                     const float* scales, const float* quats,
                                     const float* viewmat, float* cov3d, float* xys,
                                     float* depths, int32_t* radii, float* conics,
                                     float* compensation, int32_t* num_tiles_hit,
                                     int32_t* cum_tiles_hit, void* stream);

/*
 * Projection backward; replaces ops::project::bwd::xla (lib/ops.cu:63-105) and
 * project_gaussians_bwd (lib/kernels/backward.cu:345-418).  v_depth,
 * v_compensation (with compensation), v_cov2d and v_cov3d are nullable.
 */
int jaxsplat_project_backward(const jaxsplat_descriptor* d, const float* means3d,
                              const float* scales, const float* quats, const float* viewmat,
                              const int32_t* radii, const float* conics,
                              const float* compensation, const float* v_xy,
                              const float* v_depth, const float* v_conic,
                              const float* v_compensation, float* v_mean3d, float* v_scale,
                              float* v_quat, float* v_cov2d, float* v_cov3d, int adjoint,
                              void* stream);

/* Workspace bytes for jaxsplat_bin_and_sort with up to num_intersects pairs. */
size_t jaxsplat_bin_workspace_size(const jaxsplat_descriptor* d, uint32_t num_intersects);

/*
 * Intersection emission, (tile, depth) sort and tile bin edges; replaces
 * ops::sort_and_bin (lib/ops.cu:645-793) with map_gaussian_to_intersects
 * (lib/kernels/forward.cu:135-179), CUB SortPairs and get_tile_bin_edges
 * (forward.cu:184-208).  num_intersects = cum_tiles_hit[N-1].  Outputs:
 * gaussian_ids_sorted (num_intersects), tile_ids_sorted (num_intersects,
 * nullable; the reference key is (tile << 32) | bits(depths[id])) and
 * tile_bins (tiles, 2) with empty tiles (0, 0).
 */
int jaxsplat_bin_and_sort(const jaxsplat_descriptor* d, const float* xys, const float* depths,
                          const int32_t* radii, uint32_t num_intersects,
                          int32_t* gaussian_ids_sorted, int32_t* tile_ids_sorted,
                          int32_t* tile_bins, void* workspace, size_t workspace_bytes,
                          void* stream);

/*
 * The same binning in two stages, so that a caller that reads num_intersects
 * back from cum_tiles_hit can queue the first stage before it waits for the
 * count (the reference reads it between projection and sort_and_bin,
 * lib/ops.cu:645-660).  jaxsplat_bin_depth_sort needs no count: the
 * Gaussians' depth keys and tile rectangles, into a workspace that depends on
 * num_points only.  jaxsplat_bin_emit_sort then emits (in Gaussian id order,
 * as the reference), sorts by tile, bins and orders every tile by depth from
 * that workspace, with the outputs of jaxsplat_bin_and_sort; its workspace
 * (num_intersects dependent) is the one jaxsplat_bin_check reads.
 *
 * In both binning calls num_intersects is a capacity: the length of
 * gaussian_ids_sorted / tile_ids_sorted and what the workspace is sized for.
 * The kernels sort and bin the intersections actually emitted,
 * min(total, num_intersects), counted on the device, so a caller may pass a
 * capacity instead of the exact cum_tiles_hit[N-1] and skip reading the count
 * back (no host synchronisation).  A total above the capacity is an overflow:
 * only the first num_intersects in emission order (the order of
 * cum_tiles_hit) are binned, and bit 8 of
 * the error word is set.  After the call (in stream order) the workspace
 * holds the error word in bytes [0, 4) and the exact total, a uint64, in
 * bytes [8, 16), for callers that check them on the device or later.
 */
size_t jaxsplat_bin_depth_workspace_size(const jaxsplat_descriptor* d);
int jaxsplat_bin_depth_sort(const jaxsplat_descriptor* d, const float* xys, const float* depths,
                            const int32_t* radii, void* depth_workspace,
                            size_t depth_workspace_bytes, void* stream);
size_t jaxsplat_bin_emit_workspace_size(const jaxsplat_descriptor* d, uint32_t num_intersects);
int jaxsplat_bin_emit_sort(const jaxsplat_descriptor* d, uint32_t num_intersects,
                           const void* depth_workspace, int32_t* gaussian_ids_sorted,
                           int32_t* tile_ids_sorted, int32_t* tile_bins, void* workspace,
                           size_t workspace_bytes, void* stream);

/*
 * Opacity-limited tile rectangles.  No counterpart in the reference; for a
 * fused projection + rasterization (the reference's render,
 * jaxsplat/__init__.py:6-70) whose intermediate tile lists the caller never
 * sees.  Per Gaussian: its tile rectangle (the reference's,
 * lib/kernels/helpers.h:9-32) restricted to the tiles holding a pixel centre
 * where o exp(-sigma) >= 1/255 can hold (the bounding box of that ellipse,
 * with margins), packed as two words (x0 | y0 << 16, w | h << 16); the tile
 * counts (num_tiles_hit, nullable) and their inclusive sums (cum_tiles_hit).
 * The tiles dropped are those where the rasterizer's per-pixel test rejects
 * every pixel, so an image rendered through these lists equals the one
 * rendered through the reference's, while the lists are shorter (78% of the
 * intersections at the 1080p, 1M-Gaussian workload).  Binned with
 * jaxsplat_bin_depth_sort_rects + jaxsplat_bin_emit_sort (num_intersects =
 * the last cum_tiles_hit), rasterized by the usual forward and backward, or
 * by jaxsplat_rasterize_backward_staged_rects with the same rectangles.
 */
size_t jaxsplat_tile_rects_workspace_size(const jaxsplat_descriptor* d);
int jaxsplat_tile_rects(const jaxsplat_descriptor* d, const float* xys, const int32_t* radii,
                        const float* conics, const float* opacities, uint32_t* tile_rects,
                        int32_t* num_tiles_hit, int32_t* cum_tiles_hit, void* workspace,
                        size_t workspace_bytes, void* stream);
/* jaxsplat_project_forward whose num_tiles_hit and cum_tiles_hit count the
 * opacity-limited rectangles of the opacities (N), written to tile_rects: the
 * projection and jaxsplat_tile_rects in one pass, for a fused render.  No
 * counterpart in the reference.  Same workspace as jaxsplat_project_forward. */
int jaxsplat_project_forward_rects(const jaxsplat_descriptor* d, const float* means3d,
                                   const float* scales, const float* quats,
                                   const float* viewmat, const float* opacities, float* cov3d,
                                   float* xys, float* depths, int32_t* radii, float* conics,
                                   float* compensation, int32_t* num_tiles_hit,
                                   int32_t* cum_tiles_hit, uint32_t* tile_rects,
                                   void* workspace, size_t workspace_bytes, void* stream);
/* jaxsplat_bin_depth_sort with the tile rectangles given (jaxsplat_tile_rects)
 * instead of computed from xys and radii; same workspace. */
int jaxsplat_bin_depth_sort_rects(const jaxsplat_descriptor* d, const float* depths,
                                  const int32_t* radii, const uint32_t* tile_rects,
                                  void* depth_workspace, size_t depth_workspace_bytes,
                                  void* stream);

/*
 * Binning of given tile rectangles in one call, for a fused render (the
 * reference's render, jaxsplat/__init__.py:6-70, whose tile lists the caller
 * never sees); no counterpart in the reference.  tile_rects and
 * cum_tiles_hit as written by jaxsplat_project_forward_rects /
 * jaxsplat_project_forward_splats (cum_tiles_hit[g] - cum_tiles_hit[g-1] must
 * be the count of tile_rects[g]); depths as written by the projection.  The
 * intersections are emitted from the rectangles in Gaussian id order, sorted
 * by tile and ordered by (depth bits, id) inside every tile: the same outputs
 * as jaxsplat_bin_depth_sort_rects + jaxsplat_bin_emit_sort, with the same
 * workspace (jaxsplat_bin_emit_workspace_size), capacity semantics and error
 * word, without the depth stage.
 */
int jaxsplat_bin_rects_sort(const jaxsplat_descriptor* d, const float* depths,
                            const uint32_t* tile_rects, const int32_t* cum_tiles_hit,
                            uint32_t num_intersects, int32_t* gaussian_ids_sorted,
                            int32_t* tile_ids_sorted, int32_t* tile_bins, void* workspace,
                            size_t workspace_bytes, void* stream);

/*
 * Synchronise `stream` and report an error recorded in the workspace of the
 * last jaxsplat_bin_and_sort or jaxsplat_bin_emit_sort (JAXSPLAT_ERR_CAPACITY):
 * more intersections than num_intersects, or a tile key out of range.
 */
int jaxsplat_bin_check(const jaxsplat_descriptor* d, uint32_t num_intersects,
                       const void* workspace, void* stream);

/*
 * Rasterize forward over binned intersections; replaces rasterize_fwd
 * (lib/kernels/forward.cu:339-462) and the zero-intersection path
 * (lib/ops.cu:119-127).  Outputs final_Ts (H,W), final_idx (H,W),
 * out_img (H,W,3).
 */
int jaxsplat_rasterize_forward_binned(const jaxsplat_descriptor* d,
                                      const int32_t* gaussian_ids_sorted,
                                      const int32_t* tile_bins, uint32_t num_intersects,
                                      const float* xys, const float* conics,
                                      const float* colors, const float* opacities,
                                      const float* background, float* final_Ts,
                                      int32_t* final_idx, float* out_img, void* stream);

/*
 * Rasterize backward over binned intersections; replaces rasterize_bwd
 * (lib/kernels/backward.cu:153-343) and lib/ops.cu:184-186.  Gradient outputs
 * are overwritten (zeroed, then accumulated); v_xy_abs is nullable and a NULL
 * v_out_img_alpha means a zero cotangent for the alpha channel.
 */
int jaxsplat_rasterize_backward_binned(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* xys, const float* conics, const float* colors,
    const float* opacities, const float* background, const float* final_Ts,
    const int32_t* final_idx, const float* v_out_img, const float* v_out_img_alpha,
    float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs, float* v_conic,
    int adjoint, void* stream);

/*
 * N-channel rasterization (colors (N, C), background (C), out_img (H, W, C),
 * v_out_img (H, W, C), v_colors (N, C)); replaces nd_rasterize_forward
 * (lib/kernels/forward.cu:213-337) and nd_rasterize_backward_kernel
 * (lib/kernels/backward.cu:27-151), which the reference builds but never
 * calls.  The blend and its adjoint are those of the 3-channel entry points
 * for every channel, accumulated in fp32 (the reference's dead N-channel code
 * accumulates in fp16 and its backward skips the last contributor); C = 3
 * gives exactly the 3-channel results.  Channels run in passes of up to
 * three.  v_xy_abs must be NULL for C > 3 (it does not add over passes).
 */
int jaxsplat_rasterize_forward_nd(const jaxsplat_descriptor* d,
                                  const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
                                  uint32_t num_intersects, const float* xys, const float* conics,
                                  const float* colors, const float* opacities,
                                  const float* background, uint32_t channels, float* final_Ts,
                                  int32_t* final_idx, float* out_img, void* stream);
int jaxsplat_rasterize_backward_nd(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* xys, const float* conics, const float* colors,
    const float* opacities, const float* background, uint32_t channels, const float* final_Ts,
    const int32_t* final_idx, const float* v_out_img, const float* v_out_img_alpha,
    float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs, float* v_conic,
    int adjoint, void* stream);

/*
 * The backward of jaxsplat_rasterize_backward_nd for up to three channels
 * without float atomics; same adjoint as rasterize_bwd
 * (lib/kernels/backward.cu:153-343).  Each (Gaussian, tile) pair writes its
 * partial sums to a record at its intersection slot in the reference's
 * emission order (cum_tiles_hit[g] - num_tiles_hit[g] + the tile's row-major
 * offset in the Gaussian's tile rectangle, lib/kernels/forward.cu:135-179), and
 * one lane per Gaussian sums its records in slot order.  The result does not
 * depend on scheduling (bitwise reproducible) and every gradient row is
 * written.  radii and cum_tiles_hit are the projection's outputs for the same
 * xys; the workspace holds 48 bytes of record and one bit per intersection
 * (jaxsplat_rasterize_backward_staged_workspace_size).
 */
size_t jaxsplat_rasterize_backward_staged_workspace_size(const jaxsplat_descriptor* d,
                                                         uint32_t num_intersects);
int jaxsplat_rasterize_backward_staged(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* xys, const float* conics, const float* colors,
    const float* opacities, const float* background, uint32_t channels, const int32_t* radii,
    const int32_t* cum_tiles_hit, const float* final_Ts, const int32_t* final_idx,
    const float* v_out_img, const float* v_out_img_alpha, float* v_colors, float* v_opacity,
    float* v_xy, float* v_xy_abs, float* v_conic, int adjoint, void* workspace,
    size_t workspace_bytes, void* stream);
/* The same with the lists binned from tile_rects (jaxsplat_tile_rects; its
 * cum_tiles_hit): the slots follow those rectangles.  No counterpart in the
 * reference.  Same workspace size. */
int jaxsplat_rasterize_backward_staged_rects(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* xys, const float* conics, const float* colors,
    const float* opacities, const float* background, uint32_t channels, const int32_t* radii,
    const uint32_t* tile_rects, const int32_t* cum_tiles_hit, const float* final_Ts,
    const int32_t* final_idx, const float* v_out_img, const float* v_out_img_alpha,
    float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs, float* v_conic, int adjoint,
    void* workspace, size_t workspace_bytes, void* stream);

/*
 * The fused render (no counterpart as one call in the reference, whose render,
 * jaxsplat/__init__.py:6-70, is project then rasterize): the projection with
 * the opacity-limited rectangles also writes one 48-byte splat record per
 * Gaussian (N, 12 floats): (x, y, opacity, log2(255 opacity)), (conic a, b, c,
 * rectangle word 0), (colour, zero-padded to 3 channels, rectangle word 1);
 * the rasterize forward and backward stage their list entries from these
 * records instead of gathering xys, conics, colours, opacities and
 * rectangles separately.  Results are those of jaxsplat_project_forward_rects
 * + jaxsplat_rasterize_forward_binned / the backward below, bit for bit.
 * xys and num_tiles_hit may be NULL (the splat records hold the positions and
 * cum_tiles_hit the counts): the fused render writes neither.
 */
int jaxsplat_project_forward_splats(const jaxsplat_descriptor* d, const float* means3d,
                                    const float* scales, const float* quats,
                                    const float* viewmat, const float* opacities,
                                    const float* colors, uint32_t channels, float* xys,
                                    float* depths, int32_t* radii, float* conics,
                                    float* compensation, int32_t* num_tiles_hit,
                                    int32_t* cum_tiles_hit, uint32_t* tile_rects, float* splats,
                                    void* workspace, size_t workspace_bytes, void* stream);
/* Rasterize forward of the fused render; replaces rasterize_fwd
 * (lib/kernels/forward.cu:339-462) for channels in [1, 3]. */
int jaxsplat_rasterize_forward_splats(const jaxsplat_descriptor* d,
                                      const int32_t* gaussian_ids_sorted,
                                      const int32_t* tile_bins, uint32_t num_intersects,
                                      const float* splats, const float* background,
                                      uint32_t channels, float* final_Ts, int32_t* final_idx,
                                      float* out_img, void* stream);
/*
 * Backward of the fused render: the rasterize adjoint (lib/kernels/
 * backward.cu:153-343) writes the staged records of
 * jaxsplat_rasterize_backward_staged_rects into the workspace, staging from
 * the splat records; one per-Gaussian pass then sums them and runs the
 * projection VJP (backward.cu:345-570) in registers, writing v_colors
 * (N, channels), v_opacity (N), v_mean3d (N,3), v_scale (N,3) and v_quat
 * (N,4), every row.  Results equal jaxsplat_rasterize_backward_staged_rects
 * followed by jaxsplat_project_backward bit for bit.  d: the projection's
 * descriptor (its intrinsics); workspace:
 * jaxsplat_rasterize_backward_staged_workspace_size.
 */
int jaxsplat_render_backward(
    const jaxsplat_descriptor* d, const int32_t* gaussian_ids_sorted, const int32_t* tile_bins,
    uint32_t num_intersects, const float* means3d, const float* scales, const float* quats,
    const float* viewmat, const float* splats, const float* conics, const float* background,
    uint32_t channels, const int32_t* radii, const uint32_t* tile_rects,
    const int32_t* cum_tiles_hit, const float* final_Ts, const int32_t* final_idx,
    const float* v_out_img, float* v_colors, float* v_opacity, float* v_mean3d, float* v_scale,
    float* v_quat, int adjoint, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Batched rendering: many views of one or more scenes in one pass.
 *
 * No counterpart in the reference, which renders one view per call
 * (jaxsplat/__init__.py:6-70, two sorts and several host syncs each).  A
 * batch plan fixes the layout; every forward/backward then runs the whole
 * batch with kernel launches only (no allocation, no host synchronisation),
 * so a caller may capture a full step in one HIP graph.
 *
 * Layout: the parameter arrays (means3d (G,3), scales (G,3), quats (G,4),
 * colors (G,3), opacities (G,1)) hold the scenes end to end
 * (scene s = rows [first_gaussian, first_gaussian + num_gaussians), with
 * first_gaussian of scene s equal to the rows of scenes 0..s-1).  Views are
 * grouped by ascending scene; view v has a row-major 4x4 viewmats[v] and a
 * background backgrounds[v] (3 floats).  The image-sized buffers (out_img
 * (P,3), final_Ts (P), final_idx (P)) hold the views' H x W images end to end;
 * jaxsplat_batch_pixel_offset gives each view's first pixel.
 *
 * Per view the result is the plain path's render of that view (same kernels,
 * same per-tile order), except that a view with no intersections has
 * final_Ts = 1 (the plain path writes 0 there, as the reference does).
 * ------------------------------------------------------------------------ */
typedef struct jaxsplat_scene {
    int32_t first_gaussian;
    int32_t num_gaussians;
} jaxsplat_scene;

typedef struct jaxsplat_view {
    int32_t scene;
    int32_t img_h, img_w;
    float fx, fy, cx, cy;
} jaxsplat_view;

typedef struct jaxsplat_batch jaxsplat_batch;

/* Validates the layout and uploads the view table (allocates a small device
 * table on the current device; this is the only allocation of a batch).
 * isect_capacity bounds the total (view, tile) intersections of one forward. */
int jaxsplat_batch_create(jaxsplat_batch** out, const jaxsplat_scene* scenes, int32_t num_scenes,
                          const jaxsplat_view* views, int32_t num_views, float glob_scale,
                          float clip_thresh, uint32_t block_width, int64_t isect_capacity);
void jaxsplat_batch_destroy(jaxsplat_batch* b);
size_t jaxsplat_batch_workspace_size(const jaxsplat_batch* b);
/* Gaussians staged in LDS per batch by the rasterize forward / backward:
 * 64, 128 or 256, or 0 for the defaults (forward: the workgroup size, bw^2
 * rounded up to whole waves, rounded up again to 64, 128 or 256; backward:
 * 64).  The results do not depend on it; the autotuner of the Python layer
 * picks the fastest. */
int jaxsplat_batch_set_depths(jaxsplat_batch* b, int32_t fwd_depth, int32_t bwd_depth);

/* How the batch rasterizer (forward and backward) assigns its (view, tile)
 * items to workgroups: JAXSPLAT_SCHEDULE_GRID (default) launches one
 * workgroup per tile; JAXSPLAT_SCHEDULE_QUEUE orders the tiles longest list
 * first after each binning (one stable radix pass over the tiles) and runs a
 * persistent grid that takes them from a device counter, which evens out
 * batches of uneven scenes.  The results do not depend on it (every tile runs
 * the same code on the same list; the backward sums in a fixed order). */
typedef enum jaxsplat_schedule {
    JAXSPLAT_SCHEDULE_GRID = 0,
    JAXSPLAT_SCHEDULE_QUEUE = 1
} jaxsplat_schedule;
int jaxsplat_batch_set_schedule(jaxsplat_batch* b, int32_t schedule);
int64_t jaxsplat_batch_num_pixels(const jaxsplat_batch* b);
int64_t jaxsplat_batch_num_gaussians(const jaxsplat_batch* b);
int64_t jaxsplat_batch_pixel_offset(const jaxsplat_batch* b, int32_t view);

/* Forward of every view.  The workspace keeps the residuals the backward
 * reads: pass the same workspace to jaxsplat_batch_backward with no other
 * forward in between. */
int jaxsplat_batch_forward(const jaxsplat_batch* b, const float* means3d, const float* scales,
                           const float* quats, const float* colors, const float* opacities,
                           const float* viewmats, const float* backgrounds, float* final_Ts,
                           int32_t* final_idx, float* out_img, void* workspace,
                           size_t workspace_bytes, void* stream);

/* Backward of every view for the cotangent v_out_img (P,3): writes the
 * gradients of all G Gaussians (summed over the views of their scene).  No
 * float atomics: the per-(view, tile) partial sums are staged in the
 * workspace (48 bytes per intersection of capacity) and summed in a fixed
 * order, so two calls on the same inputs give bit-identical gradients and
 * the outputs need no zeroing. */
int jaxsplat_batch_backward(const jaxsplat_batch* b, const float* means3d, const float* scales,
                            const float* quats, const float* colors, const float* opacities,
                            const float* viewmats, const float* backgrounds,
                            const float* final_Ts, const int32_t* final_idx,
                            const float* v_out_img, float* v_means3d, float* v_scales,
                            float* v_quats, float* v_colors, float* v_opacities, int adjoint,
                            void* workspace, size_t workspace_bytes, void* stream);

/* Synchronous status of the last forward in `workspace`: the exact number of
 * intersections, and JAXSPLAT_ERR_CAPACITY if it exceeded isect_capacity (the
 * forward's outputs are then invalid) or a tile key was out of range. */
int jaxsplat_batch_check(const jaxsplat_batch* b, const void* workspace,
                         int64_t* num_intersects, void* stream);

/* Device address of the status word of the last forward in `workspace`
 * (non-zero after a capacity overflow or a binning error): kernels that must
 * not consume an invalid render, like jaxsplat_fit_adam, read it. */
const uint32_t* jaxsplat_batch_status_word(const jaxsplat_batch* b, const void* workspace);

/* ------------------------------------------------------------------------
 * Image-fitting loop (the training step of examples/single_image.py:26-40,
 * 143-166): parameter activations, the L2 loss gradient and Adam, fused for
 * the device.  With the batched render they make a step of kernel launches
 * only, which a caller may capture in one HIP graph.
 * ------------------------------------------------------------------------ */

/* quats / |quats|, sigmoid(colors), sigmoid(opacities) for n Gaussians
 * (single_image.py:143-166; means and scales are used raw). */
int jaxsplat_fit_activate(int64_t n, const float* quats_raw, const float* colors_raw,
                          const float* opacities_raw, float* quats, float* colors,
                          float* opacities, void* stream);

/* v_img = 2 (img - target) / n and loss_sum = sum (img - target)^2 over n
 * floats (the gradient of mean((img - target)^2), single_image.py:26-29);
 * loss_sum is written (not accumulated) and does not depend on scheduling.
 * The per-workgroup partial sums live in the library (one set per GPU), so
 * calls on one GPU must not overlap: issue them in stream order. */
int jaxsplat_fit_l2(int64_t n, const float* img, const float* target, float* v_img,
                    float* loss_sum, void* stream);

/* One Adam step (optax.adam: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
 * p -= lr m^ / (sqrt(v^) + eps)) on the raw parameters, with g the gradient
 * of the activated parameters pulled back through the activations.  moment1 /
 * moment2 hold 14 floats per Gaussian (means 3, scales 3, quats 4, colors 3,
 * opacity 1); *step (device) counts applied steps for the bias correction.  A
 * non-NULL skip word that reads non-zero makes the step a no-op. */
int jaxsplat_fit_adam(int64_t n, float* means3d, float* scales, float* quats_raw,
                      float* colors_raw, float* opacities_raw, const float* v_means3d,
                      const float* v_scales, const float* v_quats, const float* v_colors,
                      const float* v_opacities, float* moment1, float* moment2, int32_t* step,
                      const uint32_t* skip, float lr, float beta1, float beta2, float eps,
                      void* stream);

#ifdef __cplusplus
}
#endif

#endif /* JAXSPLAT_AMD_H */
