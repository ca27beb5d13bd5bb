// Per-Gaussian projection, forward and backward, for gfx950.
//
// Forward semantics follow the reference kernel project_gaussians_fwd
// (lib/kernels/forward.cu:13-131) with scale_rot_to_cov3d (forward.cu:547-568),
// project_cov3d_ewa (forward.cu:465-544) and the helpers of
// lib/kernels/helpers.h:9-57,101-153,194-211.  Backward follows
// project_gaussians_bwd (lib/kernels/backward.cu:345-418) with
// project_cov3d_ewa_vjp (:421-524), scale_rot_to_cov3d_vjp (:528-570) and the
// VJP helpers (helpers.h:60-98,131-138,155-192).
//
// This translation unit is compiled with -ffp-contract=off: every float
// operation is rounded on its own, in the order written, so the float32
// restatement of the oracle reproduces the forward bit for bit.  Products of
// 3x3 matrices always sum over the inner index in the order 0, 1, 2.
//
// Design:
//  * forward: four Gaussians per lane in coalesced 256-wide slabs, 256-lane
//    workgroups; outputs of every Gaussian are written explicitly (culled ones
//    get zeros), so no output memsets are needed;
//  * the inclusive scan of num_tiles_hit (cum_tiles_hit) is a reduce-then-scan:
//    the forward kernel writes partition-local sums and partition totals, one
//    workgroup scans the totals and a fix-up pass adds the offsets (a
//    single-pass look-back scan was latency bound: its status words cross the
//    XCDs);
//  * cov3d and compensation outputs are optional (nullptr skips the store);
//    the backward (one lane per Gaussian) recomputes cov3d from scales/quats;
//  * batched variants (views of scenes, a device table of ViewRec): the
//    forward runs one grid row per view with the same per-Gaussian code and
//    writes the depth sort key instead of a scan; the backward runs one lane
//    per Gaussian over all views of its scene, summing the per-view
//    contributions in registers.
#include <mutex>

#include "common.hpp"

namespace jsa {
namespace {

struct M3 {
    float m[3][3];  // [row][col]
};

__device__ __forceinline__ M3 mm(const M3& a, const M3& b) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}

__device__ __forceinline__ M3 tr(const M3& a) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}

// helpers.h:140-153; quaternion (w, x, y, z), not normalised.
__device__ __forceinline__ M3 quat_to_rotmat(float w, float x, float y, float z) {
    M3 R;
    R.m[0][0] = 1.f - 2.f * (y * y + z * z);
    R.m[0][1] = 2.f * (x * y - w * z);
    R.m[0][2] = 2.f * (x * z + w * y);
    R.m[1][0] = 2.f * (x * y + w * z);
    R.m[1][1] = 1.f - 2.f * (x * x + z * z);
    R.m[1][2] = 2.f * (y * z - w * x);
    R.m[2][0] = 2.f * (x * z - w * y);
    R.m[2][1] = 2.f * (y * z + w * x);
    R.m[2][2] = 1.f - 2.f * (x * x + y * y);
    return R;
}

__device__ __forceinline__ M3 scale_mat(float sx, float sy, float sz, float g) {
    M3 S;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) S.m[i][j] = 0.f;
    S.m[0][0] = g * sx;
    S.m[1][1] = g * sy;
    S.m[2][2] = g * sz;
    return S;
}

// forward.cu:547-568: upper triangle of (R S)(R S)^T.
__device__ __forceinline__ void cov3d_of(const float* s, const float* q, float g, float cov[6]) {
    const M3 R = quat_to_rotmat(q[0], q[1], q[2], q[3]);
    const M3 S = scale_mat(s[0], s[1], s[2], g);
    const M3 M = mm(R, S);
    const M3 V = mm(M, tr(M));
    cov[0] = V.m[0][0];
    cov[1] = V.m[0][1];
    cov[2] = V.m[0][2];
    cov[3] = V.m[1][1];
    cov[4] = V.m[1][2];
    cov[5] = V.m[2][2];
}

__device__ __forceinline__ void load_view(const float* __restrict__ vm, M3& W, float p[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) W.m[r][c] = vm[4 * r + c];
        p[r] = vm[4 * r + 3];
    }
}

// helpers.h:101-108
__device__ __forceinline__ void transform(const M3& W, const float p[3], float x, float y,
                                          float z, float t[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) t[r] = W.m[r][0] * x + W.m[r][1] * y + W.m[r][2] * z + p[r];
}

__device__ __forceinline__ M3 sym3(const float c[6]) {
    M3 V;
    V.m[0][0] = c[0]; V.m[0][1] = c[1]; V.m[0][2] = c[2];
    V.m[1][0] = c[1]; V.m[1][1] = c[3]; V.m[1][2] = c[4];
    V.m[2][0] = c[2]; V.m[2][1] = c[4]; V.m[2][2] = c[5];
    return V;
}

__device__ __forceinline__ M3 ewa_jacobian(float fx, float fy, float tx, float ty, float rz,
                                           float rz2) {
    M3 J;
    J.m[0][0] = fx * rz; J.m[0][1] = 0.f; J.m[0][2] = -fx * tx * rz2;
    J.m[1][0] = 0.f; J.m[1][1] = fy * rz; J.m[1][2] = -fy * ty * rz2;
    J.m[2][0] = 0.f; J.m[2][1] = 0.f; J.m[2][2] = 0.f;
    return J;
}

constexpr int kProjPer = 4;                   // Gaussians per lane
constexpr int kProjPart = 256 * kProjPer;    // Gaussians per scan partition

// Projected attributes of one Gaussian (culled ones keep the zeros).
struct Proj {
    float cov[6], conic[3];
    float cx, cy, depth, comp;
    int rad, count;
    uint2 rect;         // pack_rect of the tile rectangle, 0 if count == 0
};

// Projection of Gaussian idx (forward.cu:13-131).
__device__ __forceinline__ Proj project_core(const Camera& c, int idx, const M3& W,
                                             const float* pt, const float* __restrict__ means3d,
                                             const float* __restrict__ scales,
                                             const float* __restrict__ quats) {
    const float mx = means3d[3 * idx + 0];
    const float my = means3d[3 * idx + 1];
    const float mz = means3d[3 * idx + 2];
    float t[3];
    transform(W, pt, mx, my, mz, t);                  // clip_near_plane, helpers.h:204-211
    float cov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float conic[3] = {0.f, 0.f, 0.f};
    float cx = 0.f, cy = 0.f, depth = 0.f, comp = 0.f;
    int rad_i = 0, count = 0;
    uint2 rect = make_uint2(0u, 0u);
    if (!(t[2] <= c.clip_thresh)) {
        const float s[3] = {scales[3 * idx + 0], scales[3 * idx + 1], scales[3 * idx + 2]};
        const float4 q4 = reinterpret_cast<const float4*>(quats)[idx];
        const float q[4] = {q4.x, q4.y, q4.z, q4.w};
        cov3d_of(s, q, c.glob_scale, cov);
        // project_cov3d_ewa, forward.cu:465-544
        const float lim_x = 1.3f * c.tan_fovx;
        const float lim_y = 1.3f * c.tan_fovy;
        const float tx = t[2] * fminf(lim_x, fmaxf(-lim_x, t[0] / t[2]));
        const float ty = t[2] * fminf(lim_y, fmaxf(-lim_y, t[1] / t[2]));
        const float rz = 1.f / t[2];
        const float rz2 = rz * rz;
        const M3 J = ewa_jacobian(c.fx, c.fy, tx, ty, rz, rz2);
        const M3 T = mm(J, W);
        const M3 C2 = mm(mm(T, sym3(cov)), tr(T));
        const float c00 = C2.m[0][0], c11 = C2.m[1][1], c01 = C2.m[1][0];
        const float det_orig = c00 * c11 - c01 * c01;
        const float a = c00 + 0.3f, b = c01, cc = c11 + 0.3f;
        const float det_blur = a * cc - b * b;
        const float comp_v = sqrtf(fmaxf(0.f, det_orig / det_blur));
        // compute_cov2d_bounds, helpers.h:34-57
        const float det = a * cc - b * b;
        if (det != 0.f) {
            const float inv_det = 1.f / det;
            conic[0] = cc * inv_det;
            conic[1] = -b * inv_det;
            conic[2] = a * inv_det;
            const float hb = 0.5f * (a + cc);
            const float v1 = hb + sqrtf(fmaxf(0.1f, hb * hb - det));
            const float v2 = hb - sqrtf(fmaxf(0.1f, hb * hb - det));
            const float radius = ceilf(3.f * sqrtf(fmaxf(v1, v2)));
            // project_pix, helpers.h:122-128
            const float rw = 1.f / (t[2] + 1e-6f);
            const float px = t[0] * rw * c.fx + c.cx;
            const float py = t[1] * rw * c.fy + c.cy;
            // get_tile_bbox, helpers.h:9-32
            const float bwf = (float)c.block_width;
            const float tcx = px / bwf, tcy = py / bwf, tr_ = radius / bwf;
            const int x0 = min(max(0, (int)(tcx - tr_)), c.tiles_x);
            const int x1 = min(max(0, (int)(tcx + tr_ + 1.f)), c.tiles_x);
            const int y0 = min(max(0, (int)(tcy - tr_)), c.tiles_y);
            const int y1 = min(max(0, (int)(tcy + tr_ + 1.f)), c.tiles_y);
            const int area = (x1 - x0) * (y1 - y0);
            if (area > 0) {
                count = area;
                rect = pack_rect(x0, x1, y0, y1);
                depth = t[2];
                rad_i = (int)radius;
                cx = px;
                cy = py;
                comp = comp_v;
            }
        }
    }
    Proj r;
#pragma unroll
    for (int k = 0; k < 6; ++k) r.cov[k] = cov[k];
    r.conic[0] = conic[0];
    r.conic[1] = conic[1];
    r.conic[2] = conic[2];
    r.cx = cx;
    r.cy = cy;
    r.depth = depth;
    r.comp = comp;
    r.rad = rad_i;
    r.count = count;
    r.rect = rect;
    return r;
}

// Plain projection of Gaussian idx: writes every output row (xys and
// num_tiles_hit only where given) and returns num_tiles_hit.
__device__ __forceinline__ int project_one(
    const Camera& c, int idx, const M3& W, const float* pt, const float* __restrict__ means3d,
    const float* __restrict__ scales, const float* __restrict__ quats,
    float* __restrict__ cov3d_out, float* __restrict__ xys, float* __restrict__ depths,
    int* __restrict__ radii, float* __restrict__ conics, float* __restrict__ comp_out,
    int* __restrict__ num_tiles_hit, const float* __restrict__ opacities,
    uint2* __restrict__ rects, const float* __restrict__ colors, int nch,
    float4* __restrict__ splats) {
    const Proj r = project_core(c, idx, W, pt, means3d, scales, quats);
    const float* cov = r.cov;
    const float* conic = r.conic;
    if (cov3d_out) {
        float2* o = reinterpret_cast<float2*>(cov3d_out + 6 * (size_t)idx);
        o[0] = make_float2(cov[0], cov[1]);
        o[1] = make_float2(cov[2], cov[3]);
        o[2] = make_float2(cov[4], cov[5]);
    }
    if (xys) reinterpret_cast<float2*>(xys)[idx] = make_float2(r.cx, r.cy);
    depths[idx] = r.depth;
    radii[idx] = r.rad;
    conics[3 * idx + 0] = conic[0];
    conics[3 * idx + 1] = conic[1];
    conics[3 * idx + 2] = conic[2];
    if (comp_out) comp_out[idx] = r.comp;
    int count = r.count;
    if (rects) {        // opacity-limited rectangle and its count (tile_rects_kernel's)
        const uint2 rc = r.rad > 0 ? opacity_tile_rect(c.block_width, c.tiles_x, c.tiles_y,
                                                       make_float2(r.cx, r.cy), r.rad, conic[0],
                                                       conic[1], conic[2], opacities[idx])
                                   : make_uint2(0u, 0u);
        rects[idx] = rc;
        count = (int)rect_count(rc);
        if (splats) {   // the rasterizers' staged inputs in one 48-byte record
            const float o = opacities[idx];
            const float* col = colors + (size_t)nch * idx;
            float4* sp = splats + 3 * (size_t)idx;
            sp[0] = make_float4(r.cx, r.cy, o, alpha_limit2(o));
            sp[1] = make_float4(conic[0], conic[1], conic[2], __uint_as_float(rc.x));
            sp[2] = make_float4(col[0], nch > 1 ? col[1] : 0.f, nch > 2 ? col[2] : 0.f,
                                __uint_as_float(rc.y));
        }
    }
    if (num_tiles_hit) num_tiles_hit[idx] = count;
    return count;
}

// Camera of one view of a batch (the block width and scalars are shared).
__device__ __forceinline__ Camera view_camera(const ViewRec& v, float glob_scale,
                                              float clip_thresh, int block_width) {
    Camera c;
    c.num_points = v.n;
    c.img_w = v.img_w;
    c.img_h = v.img_h;
    c.block_width = block_width;
    c.tiles_x = v.tiles_x;
    c.tiles_y = v.tiles_y;
    c.fx = v.fx;
    c.fy = v.fy;
    c.cx = v.cx;
    c.cy = v.cy;
    c.tan_fovx = v.tan_fovx;
    c.tan_fovy = v.tan_fovy;
    c.glob_scale = glob_scale;
    c.clip_thresh = clip_thresh;
    return c;
}

// Batched forward: blockIdx.y = view, kProjPer Gaussians per lane of the
// view's scene.  Writes the item rows the binning and rasterizer read, and the
// depth sort key; no scan (the emission scans in depth order itself).
__global__ __launch_bounds__(256) void project_batch_fwd_kernel(
    const ViewRec* __restrict__ views, float glob_scale, float clip_thresh, int block_width,
    const float* __restrict__ means3d, const float* __restrict__ scales,
    const float* __restrict__ quats, const float* __restrict__ viewmats,
    const float* __restrict__ opacities, float* __restrict__ xys, float* __restrict__ conics,
    int* __restrict__ radii, uint32_t* __restrict__ dkeys, uint2* __restrict__ rects) {
    const ViewRec vr = views[blockIdx.y];
    const int first = blockIdx.x * kProjPart;
    if (first >= vr.n) return;
    const Camera c = view_camera(vr, glob_scale, clip_thresh, block_width);
    M3 W;
    float pt[3];
    load_view(viewmats + 16 * vr.view, W, pt);
    const float* m3 = means3d + 3 * (size_t)vr.g0;
    const float* sc = scales + 3 * (size_t)vr.g0;
    const float* qt = quats + 4 * (size_t)vr.g0;
#pragma unroll
    for (int k = 0; k < kProjPer; ++k) {
        const int local = first + k * 256 + (int)threadIdx.x;
        if (local >= vr.n) break;
        const Proj r = project_core(c, local, W, pt, m3, sc, qt);
        const size_t it = (size_t)vr.item_off + local;
        reinterpret_cast<float2*>(xys)[it] = make_float2(r.cx, r.cy);
        conics[3 * it + 0] = r.conic[0];
        conics[3 * it + 1] = r.conic[1];
        conics[3 * it + 2] = r.conic[2];
        radii[it] = r.rad;
        // opacity-limited: the tiles where the per-pixel alpha test can pass
        const uint2 rc = (opacities && r.rad > 0)
                             ? opacity_tile_rect(block_width, c.tiles_x, c.tiles_y,
                                                 make_float2(r.cx, r.cy), r.rad, r.conic[0],
                                                 r.conic[1], r.conic[2],
                                                 opacities[(size_t)vr.g0 + local])
                             : r.rect;
        rects[it] = rc;
        // items that emit nothing get the key the depth sort drops
        dkeys[it] = (r.rad > 0 && rect_count(rc) > 0) ? __float_as_uint(r.depth) : 0xFFFFFFFFu;
    }
}

__global__ __launch_bounds__(256) void project_fwd_kernel(
    Camera c, const float* __restrict__ means3d, const float* __restrict__ scales,
    const float* __restrict__ quats, const float* __restrict__ viewmat,
    float* __restrict__ cov3d_out, float* __restrict__ xys, float* __restrict__ depths,
    int* __restrict__ radii, float* __restrict__ conics, float* __restrict__ comp_out,
    int* __restrict__ num_tiles_hit, int* __restrict__ cum_tiles_hit,
    const float* __restrict__ opacities, uint2* __restrict__ rects,
    uint32_t* __restrict__ part_tot, const float* __restrict__ colors, int nch,
    float4* __restrict__ splats) {
    __shared__ uint32_t s_wave_tot[kProjPer][4];
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const int part = blockIdx.x;
    M3 W;
    float pt[3];
    load_view(viewmat, W, pt);

    // element k*256 + tid of the partition, k = 0..kProjPer-1 (coalesced slabs)
    int incl[kProjPer];
    int cnt[kProjPer];
#pragma unroll
    for (int k = 0; k < kProjPer; ++k) {
        const int idx = part * kProjPart + k * 256 + tid;
        cnt[k] = idx < c.num_points
                     ? project_one(c, idx, W, pt, means3d, scales, quats, cov3d_out, xys,
                                   depths, radii, conics, comp_out, num_tiles_hit, opacities,
                                   rects, colors, nch, splats)
                     : 0;
        incl[k] = wave_inclusive_sum(cnt[k]);
        if (lane == 63) s_wave_tot[k][wave] = (uint32_t)incl[k];
    }
    __syncthreads();
    // partition-local inclusive sums; project_scan_add_kernel adds the
    // partition's offset
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < kProjPer; ++k) {
        uint32_t w_off = 0, slab = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t v = s_wave_tot[k][w];
            w_off += (w < wave) ? v : 0u;
            slab += v;
        }
        const int idx = part * kProjPart + k * 256 + tid;
        if (idx < c.num_points) cum_tiles_hit[idx] = (int)(total + w_off + (uint32_t)incl[k]);
        total += slab;
    }
    if (tid == 0) part_tot[part] = total;
}

// Opacity-limited tile rectangles (opacity_tile_rect) of the projected
// Gaussians, their tile counts and the partition-local inclusive sums of the
// counts (finished by project_scan_add_kernel), laid
// out like project_fwd_kernel's.  Gaussians with radius 0 get empty rectangles.
__global__ __launch_bounds__(256) void tile_rects_kernel(
    Camera c, const float2* __restrict__ xys, const int* __restrict__ radii,
    const float* __restrict__ conics, const float* __restrict__ opacities,
    uint2* __restrict__ rects, int* __restrict__ num_tiles_hit, int* __restrict__ cum_tiles_hit,
    uint32_t* __restrict__ part_tot) {
    __shared__ uint32_t s_wave_tot[kProjPer][4];
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const int part = blockIdx.x;
    int incl[kProjPer];
    int cnt[kProjPer];
#pragma unroll
    for (int k = 0; k < kProjPer; ++k) {
        const int idx = part * kProjPart + k * 256 + tid;
        cnt[k] = 0;
        if (idx < c.num_points) {
            const int rad = radii[idx];
            uint2 rc = make_uint2(0u, 0u);
            if (rad > 0)
                rc = opacity_tile_rect(c.block_width, c.tiles_x, c.tiles_y, xys[idx], rad,
                                       conics[3 * idx + 0], conics[3 * idx + 1],
                                       conics[3 * idx + 2], opacities[idx]);
            rects[idx] = rc;
            cnt[k] = (int)rect_count(rc);
            if (num_tiles_hit) num_tiles_hit[idx] = cnt[k];
        }
        incl[k] = wave_inclusive_sum(cnt[k]);
        if (lane == 63) s_wave_tot[k][wave] = (uint32_t)incl[k];
    }
    __syncthreads();
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < kProjPer; ++k) {
        uint32_t w_off = 0, slab = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t v = s_wave_tot[k][w];
            w_off += (w < wave) ? v : 0u;
            slab += v;
        }
        const int idx = part * kProjPart + k * 256 + tid;
        if (idx < c.num_points) cum_tiles_hit[idx] = (int)(total + w_off + (uint32_t)incl[k]);
        total += slab;
    }
    if (tid == 0) part_tot[part] = total;
}

// cum_tiles_hit += its partition's offset, one workgroup per partition
// p >= 1: the workgroup sums the totals of partitions 0..p-1 itself (at most a
// few KB from L2 at 1M Gaussians), so no separate scan kernel runs between
// the projection and this one.
__global__ __launch_bounds__(256) void project_scan_add_kernel(
    int n, const uint32_t* __restrict__ part_tot, int* __restrict__ cum_tiles_hit) {
    __shared__ uint32_t s_w[4];
    const int tid = threadIdx.x;
    const int p = blockIdx.x + 1;
    uint32_t sum = 0;
    for (int q = tid; q < p; q += 256) sum += part_tot[q];
    sum = wave_inclusive_sum(sum);
    if (lane_id() == 63) s_w[tid >> 6] = sum;
    __syncthreads();
    const int off = (int)((s_w[0] + s_w[1]) + (s_w[2] + s_w[3]));
#pragma unroll
    for (int k = 0; k < kProjPer; ++k) {
        const int idx = p * kProjPart + k * 256 + tid;
        if (idx < n) cum_tiles_hit[idx] += off;
    }
}

// Fused dot products of two and three terms, in one fixed order (the
// backward's helpers write every multiply-add explicitly: this file is built
// without contraction, so the bits do not depend on the surrounding code).
__device__ __forceinline__ float fdot2(float a0, float b0, float a1, float b1) {
    return __builtin_fmaf(a0, b0, a1 * b1);
}
__device__ __forceinline__ float fdot3(float a0, float b0, float a1, float b1, float a2,
                                       float b2) {
    return __builtin_fmaf(a0, b0, __builtin_fmaf(a1, b1, a2 * b2));
}

// helpers.h:155-192 with vR[row][col]; the reference indexes v_R[col][row].
__device__ __forceinline__ void quat_vjp(const float q[4], const M3& vR, float out[4]) {
    const float w = q[0], x = q[1], y = q[2], z = q[3];
#define VR(col, row) vR.m[row][col]
    out[0] = 2.f * fdot3(x, VR(1, 2) - VR(2, 1), y, VR(2, 0) - VR(0, 2), z, VR(0, 1) - VR(1, 0));
    out[1] = 2.f * __builtin_fmaf(-2.f * x, VR(1, 1) + VR(2, 2),
                                  fdot3(y, VR(0, 1) + VR(1, 0), z, VR(0, 2) + VR(2, 0), w,
                                        VR(1, 2) - VR(2, 1)));
    out[2] = 2.f * __builtin_fmaf(-2.f * y, VR(0, 0) + VR(2, 2),
                                  fdot3(x, VR(0, 1) + VR(1, 0), z, VR(1, 2) + VR(2, 1), w,
                                        VR(2, 0) - VR(0, 2)));
    out[3] = 2.f * __builtin_fmaf(-2.f * z, VR(0, 0) + VR(1, 1),
                                  fdot3(x, VR(0, 2) + VR(2, 0), y, VR(1, 2) + VR(2, 1), w,
                                        VR(0, 1) - VR(1, 0)));
#undef VR
}

// The backward's helpers below multiply out only the non-zero structure of
// the matrices (the EWA Jacobian's zero row, the 2D cotangent's zero row and
// column, the diagonal scale matrix), use the symmetry of the 2D cotangent G
// and of the covariance V, and fuse products into multiply-adds (fdot2 /
// fdot3): the gradients follow the reference formulas exactly in real
// arithmetic but not its operation order (the forward keeps its order bit
// for bit).

// The rotation R, the scaled diagonal g s and cov3d = (R S)(R S)^T of one
// Gaussian, for its VJPs (forward.cu:547-568; equal to cov3d_of up to rounding).
struct CovParts {
    M3 R;
    float sg[3];
    float cov[6];
};

__device__ __forceinline__ CovParts cov3d_parts(const float s[3], const float q[4], float g) {
    CovParts p;
    p.R = quat_to_rotmat(q[0], q[1], q[2], q[3]);
    M3 M;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        p.sg[j] = g * s[j];
#pragma unroll
        for (int i = 0; i < 3; ++i) M.m[i][j] = p.R.m[i][j] * p.sg[j];
    }
    const auto dot3 = [&](int i, int j) {
        return fdot3(M.m[i][0], M.m[j][0], M.m[i][1], M.m[j][1], M.m[i][2], M.m[j][2]);
    };
    p.cov[0] = dot3(0, 0);
    p.cov[1] = dot3(0, 1);
    p.cov[2] = dot3(0, 2);
    p.cov[3] = dot3(1, 1);
    p.cov[4] = dot3(1, 2);
    p.cov[5] = dot3(2, 2);
    return p;
}

// VJP of one view's projection of one Gaussian with camera-space mean from
// (W, pt, m) and 3D covariance cov: the contribution to v_mean3d (from v_xy,
// v_depth and the EWA Jacobian) and v_cov3d, plus v_cov2d.
// backward.cu:345-524 with helpers.h:60-98,131-138.
__device__ __forceinline__ void view_vjp(const Camera& c, const M3& W, const float* pt,
                                         const float m[3], const float cov[6], float a, float b,
                                         float cc, float2 gxy, float vz, float ga, float gb_raw,
                                         float gc, bool has_comp, float cp, float vcomp,
                                         int consistent, float vmean[3], float vc2[3],
                                         float vc3[6]) {
    float t[3];
    transform(W, pt, m[0], m[1], m[2], t);
    // v_mean from v_xy: project_pix_vjp (helpers.h:131-138) then R^T (:90-98);
    // reciprocals to 1 ulp (v_rcp_f32) instead of correctly rounded divisions
    const float rw = __builtin_amdgcn_rcpf(t[2] + 1e-6f);
    const float vpx = c.fx * gxy.x, vpy = c.fy * gxy.y;
    const float vv0 = vpx * rw, vv1 = vpy * rw;
    const float vv2 = __builtin_fmaf(-fdot2(vpx, t[0], vpy, t[1]) * rw, rw, vz);
#pragma unroll
    for (int k = 0; k < 3; ++k) vmean[k] = fdot3(W.m[0][k], vv0, W.m[1][k], vv1, W.m[2][k], vv2);
    // v_cov2d (helpers.h:60-87)
    const float gb = gb_raw * 0.5f;
    const float xg00 = fdot2(a, ga, b, gb), xg01 = fdot2(a, gb, b, gc);
    const float xg10 = fdot2(b, ga, cc, gb), xg11 = fdot2(b, gb, cc, gc);
    vc2[0] = -fdot2(xg00, a, xg01, b);
    vc2[1] = -(fdot2(xg00, b, xg01, cc) + fdot2(xg10, a, xg11, b));
    vc2[2] = -fdot2(xg10, b, xg11, cc);
    if (has_comp) {
        const float inv_det = a * cc - b * b;
        const float omc = 1.f - cp * cp;
        const float vsq = vcomp * 0.5f / (cp + 1e-6f);
        vc2[0] += vsq * (omc * a - 0.3f * inv_det);
        vc2[1] += 2.f * vsq * (omc * b);
        vc2[2] += vsq * (omc * cc - 0.3f * inv_det);
    }
    // EWA VJP, backward.cu:421-524
    const float rz = __builtin_amdgcn_rcpf(t[2]);
    float tx = t[0], ty = t[1];
    bool inx = true, iny = true;
    float cxv = 0.f, cyv = 0.f;
    if (consistent) {
        const float lim_x = 1.3f * c.tan_fovx, lim_y = 1.3f * c.tan_fovy;
        const float ux = t[0] * rz, uy = t[1] * rz;
        inx = ux > -lim_x && ux < lim_x;
        iny = uy > -lim_y && uy < lim_y;
        cxv = fminf(lim_x, fmaxf(-lim_x, ux));
        cyv = fminf(lim_y, fmaxf(-lim_y, uy));
        tx = t[2] * cxv;
        ty = t[2] * cyv;
    }
    const float rz2 = rz * rz;
    // J = [[fx rz, 0, -fx tx rz2], [0, fy rz, -fy ty rz2], [0, 0, 0]]; T = J W
    // has rows 0 and 1 only
    const float j00 = c.fx * rz, j02 = -c.fx * tx * rz2;
    const float j11 = c.fy * rz, j12 = -c.fy * ty * rz2;
    float T0[3], T1[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T0[j] = fdot2(j00, W.m[0][j], j02, W.m[2][j]);
        T1[j] = fdot2(j11, W.m[1][j], j12, W.m[2][j]);
    }
    // G = [[g00, g01, 0], [g01, g11, 0], [0, 0, 0]]; H = G T (rows 0, 1)
    const float g00 = vc2[0], g01 = 0.5f * vc2[1], g11 = vc2[2];
    float H0[3], H1[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        H0[j] = fdot2(g00, T0[j], g01, T1[j]);
        H1[j] = fdot2(g01, T0[j], g11, T1[j]);
    }
    // v_cov3d from vV = T^T G T = T^T H (symmetric: the off-diagonal pairs
    // are twice one entry)
    const auto vv_ = [&](int i, int j) { return fdot2(T0[i], H0[j], T1[i], H1[j]); };
    vc3[0] = vv_(0, 0);
    vc3[1] = 2.f * vv_(0, 1);
    vc3[2] = 2.f * vv_(0, 2);
    vc3[3] = vv_(1, 1);
    vc3[4] = 2.f * vv_(1, 2);
    vc3[5] = vv_(2, 2);
    // vT = G T V^T + G^T T V = 2 H V (G, V symmetric), rows 0 and 1; only
    // vJ = vT W^T entries (0,0), (0,2), (1,1), (1,2) are used
    const M3 V = sym3(cov);
    float P0[3], P1[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        P0[j] = 2.f * fdot3(H0[0], V.m[0][j], H0[1], V.m[1][j], H0[2], V.m[2][j]);
        P1[j] = 2.f * fdot3(H1[0], V.m[0][j], H1[1], V.m[1][j], H1[2], V.m[2][j]);
    }
    const auto wdot = [&](const float* P, int r) {
        return fdot3(P[0], W.m[r][0], P[1], W.m[r][1], P[2], W.m[r][2]);
    };
    const float vJ00 = wdot(P0, 0), vJ02 = wdot(P0, 2), vJ11 = wdot(P1, 1), vJ12 = wdot(P1, 2);
    const float rz3 = rz2 * rz;
    float vtx = -c.fx * rz2 * vJ02;
    float vty = -c.fy * rz2 * vJ12;
    float vtz = -c.fx * rz2 * vJ00 + 2.f * c.fx * tx * rz3 * vJ02 - c.fy * rz2 * vJ11 +
                2.f * c.fy * ty * rz3 * vJ12;
    if (consistent) {
        vtz = vtz + (inx ? 0.f : cxv) * vtx + (iny ? 0.f : cyv) * vty;
        vtx = inx ? vtx : 0.f;
        vty = iny ? vty : 0.f;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
        vmean[k] = vmean[k] + fdot3(vtx, W.m[0][k], vty, W.m[1][k], vtz, W.m[2][k]);
}

// Scale/rotation VJP of cov3d = (R S)(R S)^T, backward.cu:528-570, with R and
// the scaled diagonal g s of cov3d_parts.
__device__ __forceinline__ void cov3d_vjp(const CovParts& cp, const float q[4], float glob_scale,
                                          const float vc3[6], float vsc[3], float vq[4]) {
    const M3& R = cp.R;
    // vM = 2 vVs M with M = R diag(g s) and vVs the symmetric v_cov3d
    const float h1 = 0.5f * vc3[1], h2 = 0.5f * vc3[2], h4 = 0.5f * vc3[4];
    const float Vs[3][3] = {{vc3[0], h1, h2}, {h1, vc3[3], h4}, {h2, h4, vc3[5]}};
    M3 vM;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            vM.m[i][j] = 2.f * cp.sg[j] *
                         fdot3(Vs[i][0], R.m[0][j], Vs[i][1], R.m[1][j], Vs[i][2], R.m[2][j]);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        vsc[k] = fdot3(R.m[0][k], vM.m[0][k], R.m[1][k], vM.m[1][k], R.m[2][k], vM.m[2][k]) *
                 glob_scale;
    // vR = vM S
    M3 vR;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) vR.m[i][j] = vM.m[i][j] * cp.sg[j];
    quat_vjp(q, vR, vq);
}

__global__ __launch_bounds__(256) void project_bwd_kernel(
    Camera c, const float* __restrict__ means3d, const float* __restrict__ scales,
    const float* __restrict__ quats, const float* __restrict__ viewmat,
    const int* __restrict__ radii, const float* __restrict__ conics,
    const float* __restrict__ compensation, const float* __restrict__ v_xy,
    const float* __restrict__ v_depth, const float* __restrict__ v_conic,
    const float* __restrict__ v_compensation, float* __restrict__ v_mean3d,
    float* __restrict__ v_scale, float* __restrict__ v_quat, float* __restrict__ v_cov2d_out,
    float* __restrict__ v_cov3d_out, int consistent) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= c.num_points) return;
    float vmean[3] = {0.f, 0.f, 0.f}, vsc[3] = {0.f, 0.f, 0.f}, vq[4] = {0.f, 0.f, 0.f, 0.f};
    float vc2[3] = {0.f, 0.f, 0.f}, vc3[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (radii[idx] > 0) {                                      // backward.cu:369-371
        M3 W;
        float pt[3];
        load_view(viewmat, W, pt);
        const float m[3] = {means3d[3 * idx + 0], means3d[3 * idx + 1], means3d[3 * idx + 2]};
        const float s[3] = {scales[3 * idx + 0], scales[3 * idx + 1], scales[3 * idx + 2]};
        const float4 q4 = reinterpret_cast<const float4*>(quats)[idx];
        const float q[4] = {q4.x, q4.y, q4.z, q4.w};
        const CovParts cp = cov3d_parts(s, q, c.glob_scale);
        const bool has_comp = v_compensation != nullptr;
        view_vjp(c, W, pt, m, cp.cov, conics[3 * idx + 0], conics[3 * idx + 1], conics[3 * idx + 2],
                 reinterpret_cast<const float2*>(v_xy)[idx], v_depth ? v_depth[idx] : 0.f,
                 v_conic[3 * idx + 0], v_conic[3 * idx + 1], v_conic[3 * idx + 2], has_comp,
                 has_comp ? compensation[idx] : 0.f, has_comp ? v_compensation[idx] : 0.f,
                 consistent, vmean, vc2, vc3);
        cov3d_vjp(cp, q, c.glob_scale, vc3, vsc, vq);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v_mean3d[3 * idx + k] = vmean[k];
        v_scale[3 * idx + k] = vsc[k];
    }
    reinterpret_cast<float4*>(v_quat)[idx] = make_float4(vq[0], vq[1], vq[2], vq[3]);
    if (v_cov2d_out) {
#pragma unroll
        for (int k = 0; k < 3; ++k) v_cov2d_out[3 * idx + k] = vc2[k];
    }
    if (v_cov3d_out) {
#pragma unroll
        for (int k = 0; k < 6; ++k) v_cov3d_out[6 * idx + k] = vc3[k];
    }
}

// Backward tail of the fused render: per Gaussian, the sum of its staged
// rasterize records in slot order (raster_bwd_reduce_kernel's sum, over the
// slots [cum - count, cum) of its opacity-limited rectangle) gives v_colors and
// v_opacity, written, and v_xy and v_conic, kept in registers for the
// projection VJP of project_bwd_kernel (no depth or compensation cotangent).
// Saves writing the per-Gaussian v_xy / v_conic rows and reading them back
// in a second launch.  No early exit: the wave-wide pass of sum_records needs
// every lane.
__global__ __launch_bounds__(256) void render_bwd_tail_kernel(
    Camera c, int num_isect, const float* __restrict__ means3d, const float* __restrict__ scales,
    const float* __restrict__ quats, const float* __restrict__ viewmat,
    const int* __restrict__ radii, const uint2* __restrict__ rects, const int* __restrict__ cum,
    const float* __restrict__ conics, const float* __restrict__ rec,
    const uint32_t* __restrict__ bits, int nch, float* __restrict__ v_colors,
    float* __restrict__ v_opacity, float* __restrict__ v_mean3d, float* __restrict__ v_scale,
    float* __restrict__ v_quat, int consistent) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int lane = lane_id();
    const bool own = idx < c.num_points;
    // every per-Gaussian input is loaded up front, so that only the bitmap
    // and the records wait on the rectangle and the count
    const int gi = own ? idx : 0;
    const int rad = own ? radii[gi] : 0;
    const uint2 rc = rects[gi];
    const int cend = cum[gi];
    const float m[3] = {means3d[3 * gi + 0], means3d[3 * gi + 1], means3d[3 * gi + 2]};
    const float sc[3] = {scales[3 * gi + 0], scales[3 * gi + 1], scales[3 * gi + 2]};
    const float4 q4 = reinterpret_cast<const float4*>(quats)[gi];
    const float cn[3] = {conics[3 * gi + 0], conics[3 * gi + 1], conics[3 * gi + 2]};
    const bool vis = own && rad > 0;
    float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0;
    int off = 0, end = 0;
    if (vis) {
        int x0, x1, y0, y1;
        unpack_rect(rc, x0, x1, y0, y1);
        end = cend;
        off = end - (x1 - x0) * (y1 - y0);
        if (off < 0 || end > num_isect) off = end = 0;     // inconsistent input: zeros
    }
    sum_records<false>(off, end, rec, bits, lane, a0, a1, a2);
    if (!own) return;
    float* vcol = v_colors + (size_t)nch * idx;
    vcol[0] = a0.x;
    if (nch > 1) vcol[1] = a0.y;
    if (nch > 2) vcol[2] = a0.z;
    v_opacity[idx] = a2.x;
    float vmean[3] = {0.f, 0.f, 0.f}, vsc[3] = {0.f, 0.f, 0.f}, vq[4] = {0.f, 0.f, 0.f, 0.f};
    if (vis) {                                                  // backward.cu:369-371
        M3 W;
        float pt[3];
        load_view(viewmat, W, pt);
        const float q[4] = {q4.x, q4.y, q4.z, q4.w};
        const CovParts cp = cov3d_parts(sc, q, c.glob_scale);
        float vc2[3], vc3[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        view_vjp(c, W, pt, m, cp.cov, cn[0], cn[1], cn[2], make_float2(a1.z, a1.w), 0.f, a0.w,
                 a1.x, a1.y, false, 0.f, 0.f, consistent, vmean, vc2, vc3);
        cov3d_vjp(cp, q, c.glob_scale, vc3, vsc, vq);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v_mean3d[3 * idx + k] = vmean[k];
        v_scale[3 * idx + k] = vsc[k];
    }
    reinterpret_cast<float4*>(v_quat)[idx] = make_float4(vq[0], vq[1], vq[2], vq[3]);
}

// Batched backward: blockIdx.y = scene, one lane per Gaussian; the views of
// the scene are walked in order and their v_mean3d / v_cov3d contributions
// summed in registers (the scale/rotation VJP is linear in v_cov3d, so it
// runs once on the sum), and likewise the items' colour and opacity gradients
// when v_rgbo is given.  No atomics; every Gaussian of the scene is written.
__global__ __launch_bounds__(256) void project_batch_bwd_kernel(
    const SceneRec* __restrict__ scenes, const ViewRec* __restrict__ views, float glob_scale,
    float clip_thresh, const float* __restrict__ means3d, const float* __restrict__ scales,
    const float* __restrict__ quats, const float* __restrict__ viewmats,
    const int* __restrict__ radii, const float* __restrict__ conics,
    const float* __restrict__ v_xy, const float* __restrict__ v_conic,
    const float4* __restrict__ v_rgbo, float* __restrict__ v_mean3d,
    float* __restrict__ v_scale, float* __restrict__ v_quat, float* __restrict__ v_colors,
    float* __restrict__ v_opacity, int consistent) {
    const SceneRec sc = scenes[blockIdx.y];
    const int k = blockIdx.x * 256 + (int)threadIdx.x;
    if (k >= sc.n) return;
    const size_t g = (size_t)sc.g0 + k;
    const float m[3] = {means3d[3 * g + 0], means3d[3 * g + 1], means3d[3 * g + 2]};
    const float s[3] = {scales[3 * g + 0], scales[3 * g + 1], scales[3 * g + 2]};
    const float4 q4 = reinterpret_cast<const float4*>(quats)[g];
    const float q[4] = {q4.x, q4.y, q4.z, q4.w};
    float vmean[3] = {0.f, 0.f, 0.f}, vc3[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float4 rgbo = make_float4(0.f, 0.f, 0.f, 0.f);
    CovParts cp;
    bool any = false;
    for (int v = sc.v_begin; v < sc.v_end; ++v) {
        const ViewRec vr = views[v];
        const size_t it = (size_t)vr.item_off + k;
        if (radii[it] <= 0) continue;
        if (v_rgbo) {
            const float4 c = v_rgbo[it];
            rgbo.x += c.x;
            rgbo.y += c.y;
            rgbo.z += c.z;
            rgbo.w += c.w;
        }
        if (!any) {
            cp = cov3d_parts(s, q, glob_scale);
            any = true;
        }
        const Camera c = view_camera(vr, glob_scale, clip_thresh, 1);
        M3 W;
        float pt[3];
        load_view(viewmats + 16 * vr.view, W, pt);
        float vm[3], v2[3], v3[6];
        view_vjp(c, W, pt, m, cp.cov, conics[3 * it + 0], conics[3 * it + 1], conics[3 * it + 2],
                 reinterpret_cast<const float2*>(v_xy)[it], 0.f, v_conic[3 * it + 0],
                 v_conic[3 * it + 1], v_conic[3 * it + 2], false, 0.f, 0.f, consistent, vm, v2,
                 v3);
#pragma unroll
        for (int j = 0; j < 3; ++j) vmean[j] += vm[j];
#pragma unroll
        for (int j = 0; j < 6; ++j) vc3[j] += v3[j];
    }
    float vsc[3] = {0.f, 0.f, 0.f}, vq[4] = {0.f, 0.f, 0.f, 0.f};
    if (any) cov3d_vjp(cp, q, glob_scale, vc3, vsc, vq);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        v_mean3d[3 * g + j] = vmean[j];
        v_scale[3 * g + j] = vsc[j];
    }
    reinterpret_cast<float4*>(v_quat)[g] = make_float4(vq[0], vq[1], vq[2], vq[3]);
    if (v_rgbo) {
        v_colors[3 * g + 0] = rgbo.x;
        v_colors[3 * g + 1] = rgbo.y;
        v_colors[3 * g + 2] = rgbo.z;
        v_opacity[g] = rgbo.w;
    }
}

// n words set to v, then tail (0..3) zero bytes after them.
__global__ void fill_word_kernel(uint32_t* __restrict__ p, size_t n, int tail, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
    if (blockIdx.x == 0 && (int)threadIdx.x < tail)
        reinterpret_cast<uint8_t*>(p + n)[threadIdx.x] = 0;
}

__global__ void fill_rgb_kernel(float* __restrict__ dst, size_t n_pix, const float* __restrict__ rgb) {
    const float r = rgb[0], g = rgb[1], b = rgb[2];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_pix;
         i += (size_t)gridDim.x * blockDim.x) {
        dst[3 * i + 0] = r;
        dst[3 * i + 1] = g;
        dst[3 * i + 2] = b;
    }
}

}  // namespace

hipError_t fill_zero(void* p, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 4;
    const int tail = (int)(bytes % 4);
    if (bytes == 0) return hipSuccess;
    const size_t nb = (n + 255) / 256;
    const int blocks = (int)(nb == 0 ? 1 : (nb < 2048 ? nb : 2048));
    fill_word_kernel<<<blocks, 256, 0, s>>>(static_cast<uint32_t*>(p), n, tail, 0u);
    return hipGetLastError();
}

hipError_t fill_words(void* p, size_t n_words, uint32_t value, hipStream_t s) {
    if (n_words == 0) return hipSuccess;
    const size_t nb = (n_words + 255) / 256;
    const int blocks = (int)(nb < 2048 ? nb : 2048);
    fill_word_kernel<<<blocks, 256, 0, s>>>(static_cast<uint32_t*>(p), n_words, 0, value);
    return hipGetLastError();
}

hipError_t fill_value_f32x3(float* dst, size_t n_pixels, const float* rgb, hipStream_t s) {
    if (n_pixels == 0) return hipSuccess;
    const int blocks = (int)((n_pixels + 255) / 256 < 2048 ? (n_pixels + 255) / 256 : 2048);
    fill_rgb_kernel<<<blocks, 256, 0, s>>>(dst, n_pixels, rgb);
    return hipGetLastError();
}

// scratch layout: [counter u32 | err u32 | pad to 16 | status u64 x parts]
size_t project_fwd_scratch_bytes(int n) {
    const size_t parts = ((size_t)n + kProjPart - 1) / kProjPart;
    return 16 + 8 * (parts > 0 ? parts : 1);
}

// Projection scratch held by the library, for callers that cannot allocate
// (jaxsplat_project_forward_noalloc): the fused scan's partition totals and
// offsets for up to kStaticProjectPoints Gaussians, 512 KiB + 256 B.  Each
// device holds its own copy until the process ends.  In host code the array's
// name is not a device address: kernels get the address that
// hipGetSymbolAddress returns for the current device.  Nothing in it needs
// to start at zero (the forward kernel writes every partition total before
// the scan reads it); calls that overlap on one device race on it.
constexpr size_t kStaticProjectScratch =
    (16 + 8 * (size_t)(kStaticProjectPoints / kProjPart) + 255) & ~(size_t)255;
static_assert(kStaticProjectPoints % kProjPart == 0, "whole partitions");
static __device__ __align__(256) uint8_t scratch[kStaticProjectScratch];

hipError_t project_static_scratch(void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* ptr[kMaxDevices] = {};
    *out = nullptr;
    if (project_fwd_scratch_bytes(kStaticProjectPoints) > kStaticProjectScratch)
        return hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptr[dev] = p;
    }
    *out = ptr[dev];
    return hipSuccess;
}

hipError_t project_fwd(const Camera& c, const float* means3d, const float* scales,
                       const float* quats, const float* viewmat, float* cov3d, float* xys,
                       float* depths, int* radii, float* conics, float* compensation,
                       int* num_tiles_hit, int* cum_tiles_hit, const float* opacities,
                       uint2* rects, void* scratch, hipStream_t s, const float* colors, int nch,
                       float4* splats) {
    if (c.num_points <= 0) return hipSuccess;
    if (splats && (!opacities || !rects || !colors || nch < 1 || nch > 3))
        return hipErrorInvalidValue;
    const int parts = (c.num_points + kProjPart - 1) / kProjPart;
    uint32_t* part_tot = static_cast<uint32_t*>(scratch);
    project_fwd_kernel<<<parts, 256, 0, s>>>(c, means3d, scales, quats, viewmat, cov3d, xys,
                                             depths, radii, conics, compensation, num_tiles_hit,
                                             cum_tiles_hit, opacities, rects, part_tot, colors,
                                             nch, splats);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || parts == 1) return e;
    project_scan_add_kernel<<<parts - 1, 256, 0, s>>>(c.num_points, part_tot, cum_tiles_hit);
    return hipGetLastError();
}

hipError_t tile_rects(const Camera& c, const float* xys, const int* radii, const float* conics,
                      const float* opacities, uint2* rects, int* num_tiles_hit,
                      int* cum_tiles_hit, void* scratch, hipStream_t s) {
    if (c.num_points <= 0) return hipSuccess;
    const int parts = (c.num_points + kProjPart - 1) / kProjPart;
    uint32_t* part_tot = static_cast<uint32_t*>(scratch);
    tile_rects_kernel<<<parts, 256, 0, s>>>(c, reinterpret_cast<const float2*>(xys), radii,
                                            conics, opacities, rects, num_tiles_hit,
                                            cum_tiles_hit, part_tot);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || parts == 1) return e;
    project_scan_add_kernel<<<parts - 1, 256, 0, s>>>(c.num_points, part_tot, cum_tiles_hit);
    return hipGetLastError();
}

hipError_t project_bwd(const Camera& c, const float* means3d, const float* scales,
                       const float* quats, const float* viewmat, const int* radii,
                       const float* conics, const float* compensation, const float* v_xy,
                       const float* v_depth, const float* v_conic, const float* v_compensation,
                       float* v_mean3d, float* v_scale, float* v_quat, float* v_cov2d,
                       float* v_cov3d, int consistent, hipStream_t s) {
    if (c.num_points <= 0) return hipSuccess;
    const int blocks = (c.num_points + 255) / 256;
    project_bwd_kernel<<<blocks, 256, 0, s>>>(c, means3d, scales, quats, viewmat, radii, conics,
                                              compensation, v_xy, v_depth, v_conic,
                                              v_compensation, v_mean3d, v_scale, v_quat, v_cov2d,
                                              v_cov3d, consistent);
    return hipGetLastError();
}

hipError_t render_bwd_tail(const Camera& c, int nch, int num_isect, const float* means3d,
                           const float* scales, const float* quats, const float* viewmat,
                           const int* radii, const uint2* rects, const int* cum,
                           const float* conics, const void* ws, float* v_colors,
                           float* v_opacity, float* v_mean3d, float* v_scale, float* v_quat,
                           int consistent, hipStream_t s) {
    if (c.num_points <= 0) return hipSuccess;
    if (nch < 1 || nch > 3) return hipErrorInvalidValue;
    // the layout of raster_bwd_staged's workspace: records, then the bitmap
    const uint8_t* w = static_cast<const uint8_t*>(ws);
    const size_t n = (size_t)(num_isect > 0 ? num_isect : 1);
    const float* rec = reinterpret_cast<const float*>(w);
    const uint32_t* bits = reinterpret_cast<const uint32_t*>(w + ((48 * n + 255) & ~size_t(255)));
    const int blocks = (c.num_points + 255) / 256;
    render_bwd_tail_kernel<<<blocks, 256, 0, s>>>(c, num_isect, means3d, scales, quats, viewmat,
                                                  radii, rects, cum, conics, rec, bits, nch,
                                                  v_colors, v_opacity, v_mean3d, v_scale,
                                                  v_quat, consistent);
    return hipGetLastError();
}

hipError_t project_batch_fwd(const ViewRec* views, int num_views, int max_view_n,
                             float glob_scale, float clip_thresh, int block_width,
                             const float* means3d, const float* scales, const float* quats,
                             const float* viewmats, const float* opacities, float* xys,
                             float* conics, int* radii, uint32_t* depth_keys, uint2* item_rects,
                             hipStream_t s) {
    if (num_views <= 0 || max_view_n <= 0) return hipSuccess;
    const dim3 grid((max_view_n + kProjPart - 1) / kProjPart, num_views);
    project_batch_fwd_kernel<<<grid, 256, 0, s>>>(views, glob_scale, clip_thresh, block_width,
                                                  means3d, scales, quats, viewmats, opacities,
                                                  xys, conics, radii, depth_keys, item_rects);
    return hipGetLastError();
}

hipError_t project_batch_bwd(const SceneRec* scenes, int num_scenes, int max_scene_n,
                             const ViewRec* views, float glob_scale, float clip_thresh,
                             const float* means3d, const float* scales, const float* quats,
                             const float* viewmats, const int* radii, const float* conics,
                             const float* v_xy, const float* v_conic, const float* v_rgbo,
                             float* v_mean3d, float* v_scale, float* v_quat, float* v_colors,
                             float* v_opacity, int consistent, hipStream_t s) {
    if (num_scenes <= 0 || max_scene_n <= 0) return hipSuccess;
    const dim3 grid((max_scene_n + 255) / 256, num_scenes);
    project_batch_bwd_kernel<<<grid, 256, 0, s>>>(
        scenes, views, glob_scale, clip_thresh, means3d, scales, quats, viewmats, radii, conics,
        v_xy, v_conic, reinterpret_cast<const float4*>(v_rgbo), v_mean3d, v_scale, v_quat,
        v_colors, v_opacity, consistent);
    return hipGetLastError();
}

}  // namespace jsa
