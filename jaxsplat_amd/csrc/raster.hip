// Tile rasterization, forward and backward, for gfx950.
//
// Forward semantics: reference rasterize_fwd (lib/kernels/forward.cu:339-462):
// per pixel, front to back over the tile's depth-sorted list,
// sigma = 0.5*(a dx^2 + c dy^2) + b dx dy, alpha = min(0.999, o*exp(-sigma)),
// skip if sigma < 0 or alpha < 1/255, stop without adding once
// T*(1-alpha) <= 1e-4, C += c*alpha*T; final_idx = global sorted index of the
// last added Gaussian.
// Backward semantics: reference rasterize_bwd (lib/kernels/backward.cu:153-343):
// back to front over indices <= final_idx, T recovered by division with
// alpha = min(0.99, o*vis) (reference mode) or min(0.999, .) with the clamp's
// zero slope (consistent mode, the exact adjoint of the forward).
//
// Design for CDNA4:
//  * one workgroup per tile, rounded up to whole 64-lane waves (bw*bw pixels;
//    padding lanes take part in loads and reductions but own no pixel);
//  * Gaussian batches staged in LDS as three float4 records
//    (x, y, opacity, id) (a, b, c, -) (r, g, b, -) so each lane reads a batch
//    entry with 128-bit broadcast LDS loads;
//  * per-wave early exit from the batch loop (64-lane ballot) and a per-tile
//    exit once every pixel is done;
//  * backward: the batch walk starts at the tile's largest final_idx (nothing
//    after it can contribute), per-Gaussian per-wave partials are reduced with
//    DPP (row shuffles + row broadcasts), combined across the tile's waves
//    with LDS float atomics and flushed with one set of global float atomics
//    per (Gaussian, tile) instead of one per (Gaussian, wave);
//  * the tile index is remapped so that each of the 8 XCDs (blocks are issued
//    round-robin over XCDs) renders a contiguous band of tiles, keeping
//    neighbouring tiles' Gaussians in the same L2.
#include "common.hpp"

namespace jsa {
namespace {

constexpr float kAlphaMin = 1.f / 255.f;
constexpr float kTEps = 1e-4f;

__device__ __forceinline__ int xcd_tile(int b, int num_tiles) {
    const int per = (num_tiles + 7) / 8;
    return (b & 7) * per + (b >> 3);
}

__device__ __forceinline__ int threads_for(int bw) { return ((bw * bw + 63) / 64) * 64; }

__global__ __launch_bounds__(256) void raster_fwd_kernel(
    Camera c, const int* __restrict__ gids, const int2* __restrict__ bins,
    const float2* __restrict__ xys, const float* __restrict__ conics,
    const float* __restrict__ colors, const float* __restrict__ opac,
    const float* __restrict__ bg, float* __restrict__ final_Ts, int* __restrict__ final_idx,
    float* __restrict__ out_img) {
    __shared__ float4 s_a[kMaxTileThreads];
    __shared__ float4 s_b[kMaxTileThreads];
    __shared__ float4 s_c[kMaxTileThreads];
    __shared__ int s_cnt[4];

    const int num_tiles = c.tiles_x * c.tiles_y;
    const int tile = xcd_tile(blockIdx.x, num_tiles);
    if (tile >= num_tiles) return;
    const int bw = c.block_width;
    const int nthr = blockDim.x;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const int nwaves = nthr >> 6;
    const int ty = tile / c.tiles_x, tx = tile - ty * c.tiles_x;
    const bool pixel = tid < bw * bw;
    const int i = ty * bw + (pixel ? tid / bw : 0);
    const int j = tx * bw + (pixel ? tid % bw : 0);
    const bool inside = pixel && i < c.img_h && j < c.img_w;
    const float px = (float)j + 0.5f;
    const float py = (float)i + 0.5f;

    const int2 range = bins[tile];
    float T = 1.f;
    float cr = 0.f, cg = 0.f, cb = 0.f;
    int cur = 0;
    bool done = !inside;

    for (int bstart = range.x; bstart < range.y; bstart += nthr) {
        const uint64_t dmask = ballot(done);
        if (lane == 0) s_cnt[wave] = __popcll(dmask);
        __syncthreads();
        int ndone = 0;
        for (int w = 0; w < nwaves; ++w) ndone += s_cnt[w];
        if (ndone == nthr) break;
        const int idx = bstart + tid;
        if (idx < range.y) {
            const int g = gids[idx];
            const float2 xy = xys[g];
            s_a[tid] = make_float4(xy.x, xy.y, opac[g], __int_as_float(g));
            s_b[tid] = make_float4(conics[3 * g + 0], conics[3 * g + 1], conics[3 * g + 2], 0.f);
            s_c[tid] = make_float4(colors[3 * g + 0], colors[3 * g + 1], colors[3 * g + 2], 0.f);
        }
        __syncthreads();
        const int nb = min(nthr, range.y - bstart);
        for (int t = 0; t < nb; ++t) {
            if (ballot(!done) == 0) break;
            const float4 a = s_a[t];
            const float4 b = s_b[t];
            const float dx = a.x - px, dy = a.y - py;
            const float sigma = 0.5f * (b.x * dx * dx + b.z * dy * dy) + b.y * dx * dy;
            const float alpha = fminf(0.999f, a.z * __expf(-sigma));
            if (done || sigma < 0.f || alpha < kAlphaMin) continue;
            const float next_T = T * (1.f - alpha);
            if (next_T <= kTEps) {
                done = true;
                continue;
            }
            const float vis = alpha * T;
            const float4 col = s_c[t];
            cr += col.x * vis;
            cg += col.y * vis;
            cb += col.z * vis;
            T = next_T;
            cur = bstart + t;
        }
        // protect s_cnt / LDS batch before the next round overwrites them
        __syncthreads();
    }
    if (inside) {
        const int pix = i * c.img_w + j;
        final_Ts[pix] = T;
        final_idx[pix] = cur;
        out_img[3 * pix + 0] = cr + T * bg[0];
        out_img[3 * pix + 1] = cg + T * bg[1];
        out_img[3 * pix + 2] = cb + T * bg[2];
    }
}

template <bool CONSISTENT, bool ABS>
__global__ __launch_bounds__(256) void raster_bwd_kernel(
    Camera c, const int* __restrict__ gids, const int2* __restrict__ bins,
    const float2* __restrict__ xys, const float* __restrict__ conics,
    const float* __restrict__ colors, const float* __restrict__ opac,
    const float* __restrict__ bg, const float* __restrict__ final_Ts,
    const int* __restrict__ final_idx, const float* __restrict__ v_out_img,
    const float* __restrict__ v_out_alpha, float* __restrict__ v_rgb,
    float* __restrict__ v_opacity, float* __restrict__ v_xy, float* __restrict__ v_xy_abs,
    float* __restrict__ v_conic) {
    constexpr int NV = ABS ? 11 : 9;     // rgb3, conic3, xy2, [xy_abs2], opacity
    constexpr float kClamp = CONSISTENT ? 0.999f : 0.99f;
    __shared__ float4 s_a[kMaxTileThreads];
    __shared__ float4 s_b[kMaxTileThreads];
    __shared__ float4 s_c[kMaxTileThreads];
    __shared__ float s_g[kMaxTileThreads][NV];
    __shared__ int s_flag[kMaxTileThreads];
    __shared__ int s_wmax[4];

    const int num_tiles = c.tiles_x * c.tiles_y;
    const int tile = xcd_tile(blockIdx.x, num_tiles);
    if (tile >= num_tiles) return;
    const int bw = c.block_width;
    const int nthr = blockDim.x;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const int nwaves = nthr >> 6;
    const int ty = tile / c.tiles_x, tx = tile - ty * c.tiles_x;
    const bool pixel = tid < bw * bw;
    const int i = ty * bw + (pixel ? tid / bw : 0);
    const int j = tx * bw + (pixel ? tid % bw : 0);
    const bool inside = pixel && i < c.img_h && j < c.img_w;
    const float px = (float)j + 0.5f;
    const float py = (float)i + 0.5f;
    const int pix = inside ? i * c.img_w + j : 0;

    const float T_final = inside ? final_Ts[pix] : 0.f;
    float T = T_final;
    const int bin_final = inside ? final_idx[pix] : -1;
    const float vo_r = inside ? v_out_img[3 * pix + 0] : 0.f;
    const float vo_g = inside ? v_out_img[3 * pix + 1] : 0.f;
    const float vo_b = inside ? v_out_img[3 * pix + 2] : 0.f;
    const float vo_a = (inside && v_out_alpha) ? v_out_alpha[pix] : 0.f;
    // background term of v_alpha is loop invariant: T_final * ra * (v_out_alpha - bg . v_out)
    const float bgdot = T_final * (vo_a - (bg[0] * vo_r + bg[1] * vo_g + bg[2] * vo_b));
    float buf_r = 0.f, buf_g = 0.f, buf_b = 0.f;

    const int wave_fin = wave_max_i(bin_final);
    if (lane == 0) s_wmax[wave] = wave_fin;
    __syncthreads();
    int block_fin = -1;
    for (int w = 0; w < nwaves; ++w) block_fin = max(block_fin, s_wmax[w]);

    const int2 range = bins[tile];
    const int last = min(range.y - 1, block_fin);
    for (int bend = last; bend >= range.x; bend -= nthr) {
        __syncthreads();
        const int idx = bend - tid;
        if (idx >= range.x) {
            const int g = gids[idx];
            const float2 xy = xys[g];
            s_a[tid] = make_float4(xy.x, xy.y, opac[g], __int_as_float(g));
            s_b[tid] = make_float4(conics[3 * g + 0], conics[3 * g + 1], conics[3 * g + 2], 0.f);
            s_c[tid] = make_float4(colors[3 * g + 0], colors[3 * g + 1], colors[3 * g + 2], 0.f);
        }
        for (int q = 0; q < NV; ++q) s_g[tid][q] = 0.f;
        s_flag[tid] = 0;
        __syncthreads();
        const int nb = min(nthr, bend - range.x + 1);
        const int t0 = max(0, bend - wave_fin);
        for (int t = t0; t < nb; ++t) {
            const int k = bend - t;
            const float4 a = s_a[t];
            const float4 b = s_b[t];
            const float dx = a.x - px, dy = a.y - py;
            const float sigma = 0.5f * (b.x * dx * dx + b.z * dy * dy) + b.y * dx * dy;
            const float vis = __expf(-sigma);
            const float raw = a.z * vis;
            const float alpha = fminf(kClamp, raw);
            const bool valid = (k <= bin_final) && sigma >= 0.f && alpha >= kAlphaMin;
            if (ballot(valid) == 0) continue;
            float g[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) g[q] = 0.f;
            if (valid) {
                const float ra = 1.f / (1.f - alpha);
                T *= ra;
                const float fac = alpha * T;
                const float4 col = s_c[t];
                g[0] = fac * vo_r;
                g[1] = fac * vo_g;
                g[2] = fac * vo_b;
                const float v_alpha = (col.x * T - buf_r * ra) * vo_r +
                                      (col.y * T - buf_g * ra) * vo_g +
                                      (col.z * T - buf_b * ra) * vo_b + ra * bgdot;
                buf_r += col.x * fac;
                buf_g += col.y * fac;
                buf_b += col.z * fac;
                const float slope = (CONSISTENT && raw >= kClamp) ? 0.f : 1.f;
                const float v_sigma = -a.z * vis * v_alpha * slope;
                g[3] = 0.5f * v_sigma * dx * dx;
                g[4] = v_sigma * dx * dy;
                g[5] = 0.5f * v_sigma * dy * dy;
                g[6] = v_sigma * (b.x * dx + b.y * dy);
                g[7] = v_sigma * (b.y * dx + b.z * dy);
                if (ABS) {
                    g[8] = fabsf(g[6]);
                    g[9] = fabsf(g[7]);
                }
                g[NV - 1] = vis * v_alpha * slope;
            }
#pragma unroll
            for (int q = 0; q < NV; ++q) g[q] = wave_sum_to_lane63(g[q]);
            if (lane == 63) {
#pragma unroll
                for (int q = 0; q < NV; ++q) atomicAdd(&s_g[t][q], g[q]);
                s_flag[t] = 1;
            }
        }
        __syncthreads();
        if (tid < nb && s_flag[tid]) {
            const int gid = __float_as_int(s_a[tid].w);
            unsafeAtomicAdd(v_rgb + 3 * gid + 0, s_g[tid][0]);
            unsafeAtomicAdd(v_rgb + 3 * gid + 1, s_g[tid][1]);
            unsafeAtomicAdd(v_rgb + 3 * gid + 2, s_g[tid][2]);
            unsafeAtomicAdd(v_conic + 3 * gid + 0, s_g[tid][3]);
            unsafeAtomicAdd(v_conic + 3 * gid + 1, s_g[tid][4]);
            unsafeAtomicAdd(v_conic + 3 * gid + 2, s_g[tid][5]);
            unsafeAtomicAdd(v_xy + 2 * gid + 0, s_g[tid][6]);
            unsafeAtomicAdd(v_xy + 2 * gid + 1, s_g[tid][7]);
            if (ABS) {
                unsafeAtomicAdd(v_xy_abs + 2 * gid + 0, s_g[tid][8]);
                unsafeAtomicAdd(v_xy_abs + 2 * gid + 1, s_g[tid][9]);
            }
            unsafeAtomicAdd(v_opacity + gid, s_g[tid][NV - 1]);
        }
    }
}

struct ZeroList {
    float* p[5];
    size_t n[5];
};

__global__ void zero_many_kernel(ZeroList z) {
    for (int a = 0; a < 5; ++a) {
        float* p = z.p[a];
        if (!p) continue;
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < z.n[a];
             i += (size_t)gridDim.x * blockDim.x)
            p[i] = 0.f;
    }
}

__global__ void raster_empty_kernel(size_t n_pix, const float* __restrict__ bg,
                                    float* __restrict__ final_Ts, int* __restrict__ final_idx,
                                    float* __restrict__ out_img) {
    const float r = bg[0], g = bg[1], b = bg[2];
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n_pix;
         p += (size_t)gridDim.x * blockDim.x) {
        final_Ts[p] = 0.f;
        final_idx[p] = 0;
        out_img[3 * p + 0] = r;
        out_img[3 * p + 1] = g;
        out_img[3 * p + 2] = b;
    }
}

int tile_grid(const Camera& c) {
    const int nt = c.tiles_x * c.tiles_y;
    return 8 * ((nt + 7) / 8);
}

}  // namespace

hipError_t raster_empty(const Camera& c, const float* background, float* final_Ts,
                        int* final_idx, float* out_img, hipStream_t s) {
    const size_t n = (size_t)c.img_w * c.img_h;
    if (n == 0) return hipSuccess;
    const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    raster_empty_kernel<<<blocks, 256, 0, s>>>(n, background, final_Ts, final_idx, out_img);
    return hipGetLastError();
}

hipError_t zero_grads(float* const* ptrs, const size_t* counts, int k, hipStream_t s) {
    ZeroList z{};
    size_t mx = 0;
    for (int a = 0; a < 5; ++a) {
        z.p[a] = a < k ? ptrs[a] : nullptr;
        z.n[a] = a < k ? counts[a] : 0;
        mx = z.n[a] > mx ? z.n[a] : mx;
    }
    if (mx == 0) return hipSuccess;
    const int blocks = (int)((mx + 255) / 256 < 2048 ? (mx + 255) / 256 : 2048);
    zero_many_kernel<<<blocks, 256, 0, s>>>(z);
    return hipGetLastError();
}

hipError_t raster_fwd(const Camera& c, const int* gids_sorted, const int* tile_bins,
                      const float* xys, const float* conics, const float* colors,
                      const float* opacities, const float* background, float* final_Ts,
                      int* final_idx, float* out_img, hipStream_t s) {
    const int nthr = ((c.block_width * c.block_width + 63) / 64) * 64;
    raster_fwd_kernel<<<tile_grid(c), nthr, 0, s>>>(
        c, gids_sorted, reinterpret_cast<const int2*>(tile_bins),
        reinterpret_cast<const float2*>(xys), conics, colors, opacities, background, final_Ts,
        final_idx, out_img);
    return hipGetLastError();
}

hipError_t raster_bwd(const Camera& c, const int* gids_sorted, const int* tile_bins,
                      const float* xys, const float* conics, const float* colors,
                      const float* opacities, const float* background, const float* final_Ts,
                      const int* final_idx, const float* v_out, const float* v_out_alpha,
                      float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs,
                      float* v_conic, int consistent, hipStream_t s) {
    const int nthr = ((c.block_width * c.block_width + 63) / 64) * 64;
    const int grid = tile_grid(c);
    const int2* b2 = reinterpret_cast<const int2*>(tile_bins);
    const float2* x2 = reinterpret_cast<const float2*>(xys);
#define JSA_LAUNCH(CONS, ABS)                                                                 \
    raster_bwd_kernel<CONS, ABS><<<grid, nthr, 0, s>>>(                                      \
        c, gids_sorted, b2, x2, conics, colors, opacities, background, final_Ts, final_idx,  \
        v_out, v_out_alpha, v_colors, v_opacity, v_xy, v_xy_abs, v_conic)
    if (consistent) {
        if (v_xy_abs) JSA_LAUNCH(true, true); else JSA_LAUNCH(true, false);
    } else {
        if (v_xy_abs) JSA_LAUNCH(false, true); else JSA_LAUNCH(false, false);
    }
#undef JSA_LAUNCH
    return hipGetLastError();
}

}  // namespace jsa
