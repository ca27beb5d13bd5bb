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
//  * per-(Gaussian, 64-pixel group) culling bits built once per batch: the
//    exact minimum of sigma over the group's pixel rectangle against
//    ln(255 o); each wave then walks only its set bits (scalar s_ff1);
//  * per-wave early exit from the batch loop (64-lane ballot) and a per-tile
//    exit once every pixel is done;
//  * backward: the batch walk starts at the tile's largest final_idx (nothing
//    after it can contribute) and each 64-pixel group skips the entries after
//    its own largest final_idx; per-Gaussian per-wave partials are reduced with
//    DPP (row shuffles + row broadcasts), combined across the tile's waves
//    with LDS float atomics and flushed with one set of global float atomics
//    per (Gaussian, tile) instead of one per (Gaussian, wave);
//  * the tile index is remapped so that each of the 8 XCDs (blocks are issued
//    round-robin over XCDs) renders a contiguous band of tiles, keeping
//    neighbouring tiles' Gaussians in the same L2.
#include "common.hpp"

namespace jsa {
namespace {

// Pixels per lane of the backward: 4 for tiles of more than 128 pixels, else
// 2, so that one wave covers any tile (bw <= 16) and each Gaussian is reduced
// once per tile.  At 16 x 16 the 4-slot form cuts the VALU work per
// (Gaussian, tile) by about a third: 0.49 -> 0.46 ms for the hd1m backward
// stage.  The kernel keeps one wave's sums per entry, so it is always
// launched with 64 threads.
inline int bwd_slots(int block_width) { return block_width * block_width > 128 ? 4 : 2; }
inline int bwd_threads(int block_width, int slots) {
    (void)block_width;
    (void)slots;
    return 64;
}
static_assert((kMaxTileThreads + 3) / 4 <= 64 && 128 / 2 <= 64,
              "one wave covers a tile at 4 (more than 128 pixels) or 2 pixels per lane");

constexpr float kTEps = 1e-4f;
// The batches hold the conic scaled to base 2: (a, b, c) * (log2(e)/2,
// log2(e), log2(e)/2), so that sigma2 = sigma * log2(e) =
// dx (a' dx + b' dy) + c' dy^2 and exp(-sigma) = exp2(-sigma2) (v_exp_f32).
constexpr float kLog2e = 1.44269504088896341f;
constexpr float kHalfLog2e = 0.5f * kLog2e;

// sigma * log2(e) of the offset (dx, dy) under a base-2 scaled conic, in one
// fixed order of roundings, fma(dy, c' dy, dx * fma(a', dx, b' dy)): the
// forward and the backward must evaluate it bit for bit alike.
__device__ __forceinline__ float sigma2_terms(float ax, float dx, float bdy, float cdy,
                                             float dy) {
    return __builtin_fmaf(dy, cdy, dx * __builtin_fmaf(ax, dx, bdy));
}
__device__ __forceinline__ float sigma2(float4 b, float dx, float dy) {
    return sigma2_terms(b.x, dx, b.y * dy, b.z * dy, dy);
}

// Per-pixel acceptance of a (Gaussian, pixel) pair, one rule for the forward
// and the backward (the backward recovers T by undoing the forward's blends,
// so both must take the same pairs): o exp(-sigma) >= 1/255 with sigma >= 0 is
// sigma2 in [0, lim2], lim2 = log2(255 o).  Gaussians with lim2 < 0 are culled
// whole (CullInfo::any), so lim2 >= 0 here, and the range test is one unsigned
// compare of the float bits (negative sigma2 and NaN have larger patterns).
// (alpha_limit2 is in common.hpp.)
__device__ __forceinline__ bool pair_accepted(float s2, float lim2) {
    return __float_as_uint(s2) <= __float_as_uint(lim2);
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Mask of bits lo..63 (all bits for lo <= 0, none for lo >= 64).
__device__ __forceinline__ uint64_t bits_from(int lo) {
    return lo <= 0 ? ~0ull : (lo >= 64 ? 0ull : ~((1ull << lo) - 1ull));
}

// The view that owns tile gt of the grid's tile numbering (binary search over
// the ascending tile_off of the view table; uniform per workgroup).
__device__ __forceinline__ ViewRec resolve_view(const RasterGrid& g, int gt) {
    if (g.views == nullptr) return g.v0;
    int lo = 0, hi = g.num_views;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (g.views[mid].tile_off <= gt) lo = mid;
        else hi = mid;
    }
    return g.views[lo];
}

// Up to three channels (rg.ch_off .. + rg.ch_n) of row g of a (rows, ch_stride)
// array; absent channels read as 0.  The default pass is all of an (N, 3) array.
__device__ __forceinline__ float4 load_color(const RasterGrid& rg, const float* __restrict__ a,
                                             size_t g) {
    const float* p = a + (size_t)rg.ch_stride * g + rg.ch_off;
    return make_float4(p[0], rg.ch_n > 1 ? p[1] : 0.f, rg.ch_n > 2 ? p[2] : 0.f, 0.f);
}

// One staged entry's inputs: from the splat records when the grid has them
// (one 48-byte record per Gaussian), else gathered from the separate arrays
// (item-indexed xys / conics, Gaussian-indexed colours / opacities).
struct SplatIn {
    float x, y, o, lim2, ca, cb, cc;
    float4 col;
    uint2 rect;         // splat records only
};

__device__ __forceinline__ SplatIn stage_splat(const RasterGrid& rg, int it, int g,
                                               const float2* __restrict__ xys,
                                               const float* __restrict__ conics,
                                               const float* __restrict__ colors,
                                               const float* __restrict__ opac) {
    SplatIn sp;
    if (rg.splats) {
        const float4 a = rg.splats[3 * (size_t)it + 0];
        const float4 b = rg.splats[3 * (size_t)it + 1];
        const float4 c = rg.splats[3 * (size_t)it + 2];
        sp.x = a.x;
        sp.y = a.y;
        sp.o = a.z;
        sp.lim2 = a.w;
        sp.ca = b.x;
        sp.cb = b.y;
        sp.cc = b.z;
        sp.col = make_float4(c.x, c.y, c.z, 0.f);
        sp.rect = make_uint2(__float_as_uint(b.w), __float_as_uint(c.w));
        return sp;
    }
    const float2 xy = xys[it];
    sp.x = xy.x;
    sp.y = xy.y;
    sp.o = opac[g];
    sp.lim2 = alpha_limit2(sp.o);
    sp.ca = conics[3 * it + 0];
    sp.cb = conics[3 * it + 1];
    sp.cc = conics[3 * it + 2];
    sp.col = load_color(rg, colors, g);
    sp.rect = make_uint2(0u, 0u);
    return sp;
}

// Tile-local (row, column) of pixel number p of a bw x bw tile.  Pixels are
// handled in groups of 64 (one wave, or one slot of a wave).  For 16 x 16 tiles
// group q = p / 64 is the 8 x 8 quadrant at rows 8 (q / 2).., columns 8 (q % 2)..:
// a square group has the smallest extent around a round footprint, so the
// rectangle cull below rejects about 6% more (Gaussian, group) pairs than with
// 4 x 16 row strips.  Other widths number pixels row-major.
__device__ __forceinline__ int2 tile_pixel(int p, int bw) {
    if (bw == 16) {
        const int q = p >> 6, w = p & 63;
        return make_int2(((q >> 1) << 3) + (w >> 3), ((q & 1) << 3) + (w & 7));
    }
    return make_int2(p / bw, p % bw);
}

// Pixel-centre box of the 64 tile pixels p0..p0+63 (p0 a multiple of 64,
// numbered as by tile_pixel), clipped to the image; ok = false if none is inside.
__device__ __forceinline__ float4 group_rect(int p0, int bw, int x_org, int y_org, int W, int H,
                                             bool& ok) {
    const int last = min(p0 + 63, bw * bw - 1);
    ok = p0 <= last;
    int r0, r1, c0, c1;
    if (bw == 16) {
        const int q = p0 >> 6;
        r0 = (q >> 1) << 3;
        c0 = (q & 1) << 3;
        r1 = r0 + 7;
        c1 = c0 + 7;
    } else {
        r0 = p0 / bw;
        r1 = last / bw;
        c0 = (r0 == r1) ? p0 - r0 * bw : 0;
        c1 = (r0 == r1) ? last - r1 * bw : bw - 1;
    }
    const int x0 = x_org + c0, x1 = min(x_org + c1, W - 1);
    const int y0 = y_org + r0, y1 = min(y_org + r1, H - 1);
    ok = ok && x0 <= x1 && y0 <= y1;
    return make_float4((float)x0 + 0.5f, (float)x1 + 0.5f, (float)y0 + 0.5f, (float)y1 + 0.5f);
}

// Pixel-centre box of 8 x 8 quadrant q (rows 8 (q / 2).., columns 8 (q % 2)..)
// of a bw x bw tile, bw <= 16, clipped to the tile and the image; ok = false
// if no pixel is inside.
__device__ __forceinline__ float4 quad_rect(int q, int bw, int x_org, int y_org, int W, int H,
                                            bool& ok) {
    const int r0 = (q >> 1) << 3, c0 = (q & 1) << 3;
    const int x0 = x_org + c0, x1 = min(x_org + min(c0 + 7, bw - 1), W - 1);
    const int y0 = y_org + r0, y1 = min(y_org + min(r0 + 7, bw - 1), H - 1);
    ok = r0 < bw && c0 < bw && x0 <= x1 && y0 <= y1;
    return make_float4((float)x0 + 0.5f, (float)x1 + 0.5f, (float)y0 + 0.5f, (float)y1 + 0.5f);
}

// Culling of a (Gaussian, pixel group) pair.  A pixel can pass the per-pixel
// test only where o * exp(-sigma) >= 1/255, i.e. sigma <= ln(255 o).  The test
// compares that threshold (plus 0.01, far above float32 rounding of sigma)
// with the exact minimum of the quadratic form sigma over the group's
// pixel-centre rectangle, so a culled pair is one the per-pixel test would
// reject anyway.  Conics that are not positive definite are never culled.
struct CullInfo {
    float gx, gy, a, b, c, ra, rc, smax;
    float lim2;     // alpha_limit2(o)
    bool any;       // false: no pixel can pass (lim2 < 0, i.e. 255 o < 1)
    bool exact;     // false: degenerate conic, keep every group
};

__device__ __forceinline__ CullInfo cull_info(float gx, float gy, float a, float b, float c,
                                              float lim2) {
    CullInfo k;
    k.gx = gx;
    k.gy = gy;
    k.a = a;
    k.b = b;
    k.c = c;
    k.lim2 = lim2;
    k.smax = k.lim2 * 0.69314718f + 0.01f;        // ln(255 o) + margin
    k.any = k.lim2 >= 0.f;
    k.exact = a > 0.f && c > 0.f && a * c - b * b > 0.f;
    k.ra = __builtin_amdgcn_rcpf(a);
    k.rc = __builtin_amdgcn_rcpf(c);
    return k;
}

__device__ __forceinline__ float quad(const CullInfo& k, float dx, float dy) {
    return 0.5f * (k.a * dx * dx + k.c * dy * dy) + k.b * dx * dy;
}

__device__ __forceinline__ bool group_hit(const CullInfo& k, float4 r) {
    if (!k.any) return false;
    if (!k.exact) return true;
    if (k.gx >= r.x && k.gx <= r.y && k.gy >= r.z && k.gy <= r.w) return true;
    // centre outside: the minimum lies on the boundary; minimise along each edge
    float best;
    {
        const float dx = r.x - k.gx;
        const float dy = fminf(fmaxf(-k.b * dx * k.rc, r.z - k.gy), r.w - k.gy);
        best = quad(k, dx, dy);
    }
    {
        const float dx = r.y - k.gx;
        const float dy = fminf(fmaxf(-k.b * dx * k.rc, r.z - k.gy), r.w - k.gy);
        best = fminf(best, quad(k, dx, dy));
    }
    {
        const float dy = r.z - k.gy;
        const float dx = fminf(fmaxf(-k.b * dy * k.ra, r.x - k.gx), r.y - k.gx);
        best = fminf(best, quad(k, dx, dy));
    }
    {
        const float dy = r.w - k.gy;
        const float dx = fminf(fmaxf(-k.b * dy * k.ra, r.x - k.gx), r.y - k.gx);
        best = fminf(best, quad(k, dx, dy));
    }
    return best <= k.smax;
}

// ---------------------------------------------------------------------------
// Forward: one pixel per lane, one 64-pixel group per wave.
// ---------------------------------------------------------------------------
// Tile gt (< rg.total_tiles) of the grid's tile numbering, rendered by the
// whole workgroup; the flat-grid kernel and the work-queue kernel below both
// run this body.
template <int DEPTH>
__device__ __forceinline__ void raster_fwd_tile(
    const RasterGrid& rg, int gt, const int* __restrict__ gids, const int2* __restrict__ bins,
    const float2* __restrict__ xys, const float* __restrict__ conics,
    const float* __restrict__ colors, const float* __restrict__ opac,
    const float* __restrict__ bgs, float* __restrict__ final_Ts, int* __restrict__ final_idx,
    float* __restrict__ out_img) {
    constexpr int NW = DEPTH / 64;            // 64-entry words per batch
    __shared__ float4 s_a[DEPTH];             // tile-centred x, y, opacity, id
    __shared__ float4 s_b[DEPTH];             // base-2 conic a', b', c', alpha_limit2(opacity)
    __shared__ float4 s_c[DEPTH];             // colour
    __shared__ uint64_t s_mask[4][NW];        // [group][64-entry word]
    __shared__ float4 s_rect[4];
    __shared__ int s_rok[4];
    __shared__ int s_cnt[4];

    const ViewRec vr = resolve_view(rg, gt);
    const int tile = gt - vr.tile_off;
    const int gshift = vr.g0 - vr.item_off;      // gaussian = item + gshift
    const int bw = rg.block_width;
    const int nthr = blockDim.x;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const int nwaves = nthr >> 6;
    const int ty = tile / vr.tiles_x, tx = tile - ty * vr.tiles_x;
    const int x_org = tx * bw, y_org = ty * bw;
    const bool pixel = tid < bw * bw;
    const int2 rc = tile_pixel(pixel ? tid : 0, bw);
    const int i = y_org + rc.x;
    const int j = x_org + rc.y;
    const bool inside = pixel && i < vr.img_h && j < vr.img_w;
    // pixel centre and staged Gaussian positions relative to the tile's centre
    // pixel corner (x_org + bw / 2, y_org + bw / 2), as in the backward
    const int cx0 = x_org + bw / 2, cy0 = y_org + bw / 2;
    const float px = (float)(j - cx0) + 0.5f;
    const float py = (float)(i - cy0) + 0.5f;
    if (tid < nwaves) {
        bool ok;
        s_rect[tid] = group_rect(tid * 64, bw, x_org, y_org, vr.img_w, vr.img_h, ok);
        s_rok[tid] = ok;
    }

    // wave-uniform, so that the batch bounds and sorted indices stay scalar
    const int2 rb = bins[vr.bin_off + tile];
    const int2 range = make_int2(uniform_i(rb.x), uniform_i(rb.y));
    // A pixel is done once it has terminated (or from the start when it lies
    // outside the image).  The done set is a wave-uniform lane mask: the
    // branches below test ballots of plain compares against it in scalar
    // registers, and a done pixel takes alpha = 0, so it blends nothing, keeps
    // T and never terminates again.  The common pair, on which no blending
    // pixel terminates, then updates every lane unconditionally.
    float T = 1.f;
    float cr = 0.f, cg = 0.f, cb = 0.f;
    int cur = 0;
    uint64_t dm = ballot(!inside);

    for (int bstart = range.x; bstart < range.y; bstart += DEPTH) {
        if (lane == 0) s_cnt[wave] = __popcll(dm);
        __syncthreads();
        int ndone = 0;
        for (int w = 0; w < nwaves; ++w) ndone += s_cnt[w];
        // uniform exit, so that bstart stays a scalar
        if (uniform_i(ndone) == nthr) break;
        // stage entries e = e0 + tid of the batch (several per lane when DEPTH > nthr)
        for (int e0 = 0; e0 < DEPTH; e0 += nthr) {
            const int e = e0 + tid;
            const int idx = bstart + e;
            uint32_t bits = 0;
            const int it = (e < DEPTH && idx < range.y) ? gids[idx] : -1;
            if ((unsigned)(it - vr.item_off) < (unsigned)vr.n) {    // an item of this view
                const SplatIn sp = stage_splat(rg, it, it + gshift, xys, conics, colors, opac);
                const CullInfo ci = cull_info(sp.x, sp.y, sp.ca, sp.cb, sp.cc, sp.lim2);
                s_a[e] = make_float4(sp.x - (float)cx0, sp.y - (float)cy0, sp.o,
                                     __int_as_float(it));
                s_b[e] = make_float4(sp.ca * kHalfLog2e, sp.cb * kLog2e, sp.cc * kHalfLog2e,
                                     sp.lim2);
                s_c[e] = sp.col;
                for (int G = 0; G < nwaves; ++G)
                    if (s_rok[G] && group_hit(ci, s_rect[G])) bits |= 1u << G;
            }
            const int word = (e0 >> 6) + wave;
            for (int G = 0; G < nwaves; ++G) {
                const uint64_t m = ballot((bits >> G) & 1u);
                if (lane == 0 && word < NW) s_mask[G][word] = m;
            }
        }
        __syncthreads();
        // Branch-free per pixel: the decisions are lane masks and every branch
        // is uniform (no exec-mask updates per evaluation).  A pair that no
        // blending pixel accepts ends after the acceptance ballot; otherwise
        // every lane blends with its alpha (0 where not accepted, which keeps
        // T and the colour exactly), and only a pair that terminates some
        // pixel takes the selects that keep it from blending.
        // Software-pipelined over the set bits of each mask word: the next
        // entry's LDS reads are issued before the current pair is evaluated,
        // which takes their latency out of the dependent chain per pair.  Two
        // register sets alternate (unrolled by two), so nothing is copied.
        bool live = dm != ~0ull;
        int cur_off = -1;          // 16 x the batch entry last blended, -1 for none
        for (int q = 0; q < NW && live; ++q) {
            uint64_t m = uniform64(s_mask[wave][q]);
            // branch-free: with no bit left it re-reads entry t (a valid index),
            // so the LDS waits before each evaluation stay exact.  The entry's
            // byte offset 16 t is moved to a vector register once: it is the
            // LDS address of the three reads and the value that marks the
            // pixel's last blended entry (cur_off).
            auto fetch = [&](int& t, int& toff, float4& a, float4& b, float4& c) -> bool {
                const bool h = m != 0;
                t = h ? q * 64 + __ffsll((unsigned long long)m) - 1 : t;
                m &= m - 1;
                __asm__("v_mov_b32 %0, %1" : "=v"(toff) : "s"(t * (int)sizeof(float4)));
                a = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_a) + toff);
                b = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_b) + toff);
                c = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_c) + toff);
                return h;
            };
            // one pair; true when every pixel of the wave is done
            auto eval = [&](int toff, const float4& a, const float4& b, const float4& col) -> bool {
                const float dx = a.x - px, dy = a.y - py;
                const float s2 = sigma2(b, dx, dy);
                const bool ok = pair_accepted(s2, b.w);
                const uint64_t take = ballot(ok) & ~dm;     // accepted by a blending pixel
                if (take == 0) return false;
                const float alpha = fminf(0.999f, a.z * __builtin_amdgcn_exp2f(-s2));
                const float al = lane_in(take) ? alpha : 0.f;
                const float vis = al * T;
                // T (1 - alpha) as T - alpha T, in place; exactly T where al = 0
                T = T - vis;
                // a pixel with al = 0 keeps its T > 1e-4: fin is within take
                const uint64_t fin = ballot(T <= kTEps);
                uint64_t add = take;
                float v = vis;
                if (fin != 0) {
                    // a terminating pixel adds nothing and keeps its transmittance
                    // (T + alpha T restores it to within an ulp)
                    const bool f = lane_in(fin);
                    T = f ? T + vis : T;
                    v = f ? 0.f : vis;
                    add = take & ~fin;
                    dm |= fin;
                }
                cr += col.x * v;
                cg += col.y * v;
                cb += col.z * v;
                cur_off = lane_in(add) ? toff : cur_off;
                return dm == ~0ull;
            };
            int t0 = 0, t1 = 0, o0, o1;
            float4 a0, b0, c0, a1, b1, c1;
            bool h0 = fetch(t0, o0, a0, b0, c0);
            while (h0) {
                const bool h1 = fetch(t1, o1, a1, b1, c1);
                if (eval(o0, a0, b0, c0)) {
                    live = false;
                    break;
                }
                if (!h1) break;
                h0 = fetch(t0, o0, a0, b0, c0);
                if (eval(o1, a1, b1, c1)) {
                    live = false;
                    break;
                }
            }
        }
        // the batch's last blended entry, if any, as a sorted index
        cur = cur_off >= 0 ? bstart + (cur_off >> 4) : cur;
        __syncthreads();
    }
    if (inside) {
        const size_t pix = (size_t)vr.pix_off + (size_t)i * vr.img_w + j;
        const float4 bg = load_color(rg, bgs, vr.view);
        if (rg.ch_off == 0) {            // identical for every channel pass
            final_Ts[pix] = T;
            final_idx[pix] = cur;
        }
        float* o = out_img + (size_t)rg.ch_stride * pix + rg.ch_off;
        o[0] = cr + T * bg.x;
        if (rg.ch_n > 1) o[1] = cg + T * bg.y;
        if (rg.ch_n > 2) o[2] = cb + T * bg.z;
    }
}

// Flat grid: one workgroup per tile, XCD-aware tile order.
template <int DEPTH>
__global__ __launch_bounds__(256) void raster_fwd_kernel(
    RasterGrid rg, const int* __restrict__ gids, const int2* __restrict__ bins,
    const float2* __restrict__ xys, const float* __restrict__ conics,
    const float* __restrict__ colors, const float* __restrict__ opac,
    const float* __restrict__ bgs, float* __restrict__ final_Ts, int* __restrict__ final_idx,
    float* __restrict__ out_img) {
    const int gt = xcd_tile(blockIdx.x, rg.total_tiles);
    if (gt >= rg.total_tiles) return;
    raster_fwd_tile<DEPTH>(rg, gt, gids, bins, xys, conics, colors, opac, bgs, final_Ts,
                           final_idx, out_img);
}

// Next item of a work queue, taken by one workgroup: order[i] for the i-th
// atomic increment of *counter, -1 once all n items are taken.  Every
// workgroup ends after at most one item past n, so the grid always drains.
__device__ __forceinline__ int next_queue_item(int* __restrict__ counter,
                                               const int* __restrict__ order, int n) {
    __shared__ int s_item;
    __syncthreads();                      // the previous item's LDS reads are done
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int i = s_item;
    return i < n ? order[i] : -1;
}

// Work queue: a persistent grid whose workgroups take tiles in the given order
// (longest tile lists first) until none is left.
template <int DEPTH>
__global__ __launch_bounds__(256) void raster_fwd_queue_kernel(
    RasterGrid rg, const int* __restrict__ order, int* __restrict__ counter,
    const int* __restrict__ gids, const int2* __restrict__ bins,
    const float2* __restrict__ xys, const float* __restrict__ conics,
    const float* __restrict__ colors, const float* __restrict__ opac,
    const float* __restrict__ bgs, float* __restrict__ final_Ts, int* __restrict__ final_idx,
    float* __restrict__ out_img) {
    for (;;) {
        const int gt = next_queue_item(counter, order, rg.total_tiles);
        if (gt == -1) return;
        if ((unsigned)gt >= (unsigned)rg.total_tiles) continue;   // never for a valid order
        raster_fwd_tile<DEPTH>(rg, gt, gids, bins, xys, conics, colors, opac, bgs, final_Ts,
                               final_idx, out_img);
    }
}

// ---------------------------------------------------------------------------
// Backward: NSLOT pixels per lane (slots); 64-pixel groups = (slot, wave).
// ---------------------------------------------------------------------------
struct PixState {
    // px, py: pixel centre relative to the tile's centre corner (exact);
    // pxx, pxy, pyy: its products, the weights of the moment sums
    float px, py, pxx, pxy, pyy, vo_r, vo_g, vo_b;
    // U = 1 / T for the transmittance T in front of the Gaussians walked so
    // far (back to front), and w = B / T for their blended colour behind,
    // B = (sum of colour * alpha * T, dotted with v_out) - T_final (v_out_alpha
    // - background . v_out), the only form the adjoint uses.  With these,
    // stepping over a Gaussian of alpha a is U' = U (1 - a) (one FMA, no
    // division), T' = 1 / U', v_alpha = T' (c.v_out - w), w' = w + a (c.v_out - w).
    float U, w;
    int fin;
    bool inside;
};

// One pixel's contribution to the per-Gaussian partial sums, branch-free
// (an invalid pair contributes zeros and leaves the pixel state unchanged).
// Only pixel-dependent sums are accumulated; the per-Gaussian factors are
// applied once per (Gaussian, tile) by finish_grads:
//   g[0..2] = sum fac * v_out                       (v_rgb)
//   g[3..5] = sum vs dx^2, vs dx dy, vs dy^2         (v_conic, x 0.5 / 1 / 0.5)
//   g[6..7] = sum vs dx, vs dy                       (v_xy = conic * (sum vs d))
//   [g[8..9] = sum |v_xy per pixel|]                 (v_xy_abs)
//   g[NV-1] = sum vs                                 (v_opacity = -sum vs / o)
// where without v_xy_abs g[3..7] hold the moments sum vs px^2, vs px py,
// vs py^2, vs px, vs py of the tile-centred pixel positions instead (one
// multiply-add each; finish_grads expands (gx - px)^2 etc. once per entry)
// with vs = dL/dsigma = -o * vis * v_alpha (times the clamp's zero slope in
// consistent mode).  The returned mask of lanes with a valid pair comes from
// ballots taken inside, so every call must run in wave-uniform control flow
// with all 64 lanes active (the call sites branch on scalar mask bits only):
// a lane masked off around the call would be missing from the ballots, and
// its pairs from the per-Gaussian sums.
template <bool CONSISTENT, int NV>
__device__ __forceinline__ uint64_t bwd_pixel(PixState& s, float dx, float dy, float s2,
                                              const float4 a, const float4 b, const float4 col,
                                              int k, float* g) {
#define JSA_ACC(i, v) (g[i] += (v))
    constexpr float kClamp = CONSISTENT ? 0.999f : 0.99f;
    const float raw = a.z * __builtin_amdgcn_exp2f(-s2);
    // a pixel outside the image has fin = -1 (k >= 0), so live implies inside
    const bool live = k <= s.fin;
    const bool acc = pair_accepted(s2, col.w);
    const bool valid = live && acc;
    // an invalid pair acts as alpha = 0: U and w unchanged exactly, no weight
    const float rv = valid ? raw : 0.f;
    const float al = fminf(kClamp, rv);
    const float U = __builtin_fmaf(-s.U, al, s.U);
    s.U = U;
    const float T = __builtin_amdgcn_rcpf(U);    // transmittance in front of this Gaussian
    const float fac = al * T;
    JSA_ACC(0, fac * s.vo_r);
    JSA_ACC(1, fac * s.vo_g);
    JSA_ACC(2, fac * s.vo_b);
    // v_alpha = sum_ch (col T - buf / (1 - alpha)) v_out
    //           + T_final (v_out_alpha - bg . v_out) / (1 - alpha) = T (cv - w)
    const float cv = col.x * s.vo_r + col.y * s.vo_g + col.z * s.vo_b;
    const float dcw = cv - s.w;
    const float v_alpha = T * dcw;
    s.w = __builtin_fmaf(al, dcw, s.w);
    // the clamp's zero slope applies in consistent mode only
    const float vs = (CONSISTENT && raw >= kClamp) ? 0.f : -rv * v_alpha;
    if (NV == 11) {
        const float p = vs * dx, q = vs * dy;
        JSA_ACC(3, p * dx);
        JSA_ACC(4, p * dy);
        JSA_ACC(5, q * dy);
        JSA_ACC(6, p);
        JSA_ACC(7, q);
        // log2(e) |conic (p, q)|, rescaled by finish_grads
        JSA_ACC(8, fabsf(2.f * b.x * p + b.y * q));
        JSA_ACC(9, fabsf(b.y * p + 2.f * b.z * q));
    } else {
        // moments of vs over the pixel positions; finish_grads turns them
        // into the sums over the offsets (dx, dy) = Gaussian - pixel
        JSA_ACC(3, vs * s.pxx);
        JSA_ACC(4, vs * s.pxy);
        JSA_ACC(5, vs * s.pyy);
        JSA_ACC(6, vs * s.px);
        JSA_ACC(7, vs * s.py);
    }
    JSA_ACC(NV - 1, vs);
#undef JSA_ACC
    // the lanes with a valid pair, from ballots of the plain compares
    return ballot(live) & ballot(acc);
}

// Per-Gaussian factors of the partial sums of bwd_pixel (the batch's base-2
// scaled conic and opacity o of the Gaussian): r[] becomes (v_rgb 3, v_conic 3,
// v_xy 2, [v_xy_abs 2], v_opacity).  o > 0 whenever a pair was valid
// (alpha >= 1/255).
template <int NV>
__device__ __forceinline__ void finish_grads(float* r, float4 conic2, float o, float gx,
                                             float gy) {
    constexpr float kInvLog2e = 1.f / kLog2e;
    if (NV == 9) {
        // moments -> offset sums, with (gx, gy) the tile-centred Gaussian
        // position: sum vs (gx - px) = gx S - Sx, sum vs (gx - px)^2 =
        // gx (gx S - 2 Sx) + Sxx, sum vs (gx - px)(gy - py) =
        // gx (gy S - Sy) - gy Sx + Sxy
        const float S = r[NV - 1], Sx = r[6], Sy = r[7];
        const float ex = fmaf(gx, S, -Sx), ey = fmaf(gy, S, -Sy);
        r[3] = fmaf(gx, ex - Sx, r[3]);
        r[4] = fmaf(gx, ey, fmaf(-gy, Sx, r[4]));
        r[5] = fmaf(gy, ey - Sy, r[5]);
        r[6] = ex;
        r[7] = ey;
    }
    const float ca = conic2.x * (2.f * kInvLog2e), cb = conic2.y * kInvLog2e,
                cc = conic2.z * (2.f * kInvLog2e);
    r[3] *= 0.5f;
    r[5] *= 0.5f;
    const float sp = r[6], sq = r[7];
    r[6] = ca * sp + cb * sq;
    r[7] = cb * sp + cc * sq;
    if (NV == 11) {
        r[8] *= kInvLog2e;
        r[9] *= kInvLog2e;
    }
    r[NV - 1] = o > 0.f ? -r[NV - 1] / o : 0.f;
}

// Staged gradient output of a single-view backward (no float atomics).  The
// flush of (Gaussian g, tile) writes its sums as one record (9 floats, or 12
// with the v_xy_abs sums; load_record / store_record) at the pair's
// intersection slot in the reference's emission order, cum_tiles_hit[g] -
// num_tiles_hit[g] + the tile's row-major offset in g's tile rectangle, and
// sets the slot's bit in a bitmap (one 32-bit atomic OR per flush, where the
// direct path adds nine floats atomically); raster_bwd_reduce_kernel (or the
// fused render's tail) then sums each Gaussian's flagged records in slot
// order, so the result does not depend on the order in which tiles finish.
struct StagedOut {
    float* rec;             // one record per slot (load_record / store_record)
    uint32_t* bits;         // one bit per slot, zero before the backward
    const int* cum;         // cum_tiles_hit (inclusive), used when first is null
    const uint32_t* first;  // a batch's first slot per item (exclusive)
    const uint2* rects;     // packed tile rectangles; null: from xys and radii
    const int* radii;
    int n_slots;            // number of intersections; a slot outside is not written
};

// NSLOT pixels per lane (slots), so each (Gaussian, wave) reduction covers
// 64 * NSLOT pixels; 64-pixel groups = (slot, wave).
template <bool CONSISTENT, bool ABS, int DEPTH, bool STAGED, int NSLOT>
__device__ __forceinline__ void raster_bwd_tile(
    const RasterGrid& rg, int gt, const int* __restrict__ gids, const int2* __restrict__ bins,
    const float2* __restrict__ xys, const float* __restrict__ conics,
    const float* __restrict__ colors, const float* __restrict__ opac,
    const float* __restrict__ bgs, const float* __restrict__ final_Ts,
    const int* __restrict__ final_idx, const float* __restrict__ v_out_img,
    const float* __restrict__ v_out_alpha, float* __restrict__ v_rgb,
    float* __restrict__ v_opacity, float* __restrict__ v_xy, float* __restrict__ v_xy_abs,
    float* __restrict__ v_conic, const StagedOut& st) {
    constexpr int NV = ABS ? 11 : 9;     // rgb3, conic3, xy2, [xy_abs2], opacity
    constexpr int NW = DEPTH / 64;       // 64-entry words per batch
    __shared__ float4 s_a[DEPTH];        // tile-centred x, y, opacity, id
    __shared__ float4 s_b[DEPTH];        // base-2 conic a', b', c', staged slot
    __shared__ float4 s_c[DEPTH];        // colour, alpha_limit2(opacity)
    // one wave per workgroup (bwd_threads: NSLOT pixels per lane cover the
    // tile in 64 lanes); its per-entry sums, kept until the batch's flush
    __shared__ float s_g[DEPTH][NV];
    __shared__ int s_flag[DEPTH];
    __shared__ uint64_t s_mask[4][NW];   // [group = slot * nwaves + wave][64-entry word]
    __shared__ float4 s_rect[4];
    __shared__ int s_rok[4];
    __shared__ int s_wmax[2];

    const ViewRec vr = resolve_view(rg, gt);
    const int tile = gt - vr.tile_off;
    const int gshift = vr.g0 - vr.item_off;      // gaussian = item + gshift
    const int bw = rg.block_width;
    constexpr int nthr = 64;             // one wave (bwd_threads)
    const int tid = threadIdx.x;
    const int lane = lane_id();
    constexpr int wave = 0;
    constexpr int nwaves = 1;
    constexpr int ngroups = NSLOT;
    // Four slots (tiles of 12..16 pixels): slot sl of lane l is pixel
    // (l / 8, l % 8) of the 8 x 8 quadrant sl (lanes past a smaller tile own no
    // pixel), so slots 0/2 and 1/3 share a column and 0/1 and 2/3 a row.  Two
    // slots (up to 11 pixels): pixel sl * 64 + l, row-major.
    constexpr bool quad = NSLOT == 4;
    const int ty = tile / vr.tiles_x, tx = tile - ty * vr.tiles_x;
    const int x_org = tx * bw, y_org = ty * bw;
    // positions relative to the tile's centre corner, as in the forward
    const int cx0 = x_org + bw / 2, cy0 = y_org + bw / 2;
    const float4 bg4 = load_color(rg, bgs, vr.view);
    const float bgr = bg4.x, bgg = bg4.y, bgb = bg4.z;

    PixState ps[NSLOT];
    int slot_fin[NSLOT];                 // per slot: largest final index over the wave
#pragma unroll
    for (int sl = 0; sl < NSLOT; ++sl) {
        int r, c;
        bool pixel;
        if constexpr (quad) {
            r = ((sl >> 1) << 3) + (tid >> 3);
            c = ((sl & 1) << 3) + (tid & 7);
            pixel = r < bw && c < bw;
        } else {
            const int p = sl * nthr + tid;
            pixel = p < bw * bw;
            r = pixel ? p / bw : 0;
            c = pixel ? p - r * bw : 0;
        }
        const int i = y_org + r;
        const int j = x_org + c;
        PixState& s = ps[sl];
        s.inside = pixel && i < vr.img_h && j < vr.img_w;
        s.px = (float)(j - cx0) + 0.5f;
        s.py = (float)(i - cy0) + 0.5f;
        s.pxx = s.px * s.px;
        s.pxy = s.px * s.py;
        s.pyy = s.py * s.py;
        const size_t pix = s.inside ? (size_t)vr.pix_off + (size_t)i * vr.img_w + j : 0;
        // the forward leaves T_final > 1e-4; a zero from elsewhere gives a
        // huge finite U (T ~ 0: no weight) rather than inf (inf * 0 = NaN)
        const float T_final = s.inside ? final_Ts[pix] : 1.f;
        s.U = T_final > 0.f ? 1.f / T_final : 3.0e38f;
        s.fin = s.inside ? final_idx[pix] : -1;
        const float4 vo = s.inside ? load_color(rg, v_out_img, pix) : make_float4(0, 0, 0, 0);
        s.vo_r = vo.x;
        s.vo_g = vo.y;
        s.vo_b = vo.z;
        const float vo_a = (s.inside && v_out_alpha) ? v_out_alpha[pix] : 0.f;
        // w = B / T_final with B's background term only
        s.w = (bgr * s.vo_r + bgg * s.vo_g + bgb * s.vo_b) - vo_a;
        slot_fin[sl] = uniform_i(wave_max_i(s.fin));
    }
    if (tid < ngroups) {
        bool ok;
        s_rect[tid] = quad ? quad_rect(tid, bw, x_org, y_org, vr.img_w, vr.img_h, ok)
                           : group_rect(tid * 64, bw, x_org, y_org, vr.img_w, vr.img_h, ok);
        s_rok[tid] = ok;
    }
    for (int e = tid; e < DEPTH; e += nthr) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            s_g[e][v] = 0.f;
        }
        s_flag[e] = 0;
    }
    int wave_fin = slot_fin[0];
#pragma unroll
    for (int sl = 1; sl < NSLOT; ++sl) wave_fin = max(wave_fin, slot_fin[sl]);
    if (lane == 0) s_wmax[wave] = wave_fin;
    __syncthreads();
    int block_fin = -1;
    for (int w = 0; w < nwaves; ++w) block_fin = max(block_fin, s_wmax[w]);

    const int2 range = bins[vr.bin_off + tile];
    // wave-uniform (block_fin comes from LDS), so that the batch bounds and the
    // culling masks below stay in scalar registers
    const int last = uniform_i(min(range.y - 1, block_fin));
    for (int bend = last; bend >= range.x; bend -= DEPTH) {
        // stage entries e = e0 + tid (sorted index bend - e, back to front)
        for (int e0 = 0; e0 < DEPTH; e0 += nthr) {
            const int e = e0 + tid;
            const int idx = bend - e;
            uint32_t bits = 0;
            const int it = (e < DEPTH && idx >= range.x) ? gids[idx] : -1;
            if ((unsigned)(it - vr.item_off) < (unsigned)vr.n) {    // an item of this view
                const SplatIn sp = stage_splat(rg, it, it + gshift, xys, conics, colors, opac);
                s_a[e] = make_float4(sp.x - (float)cx0, sp.y - (float)cy0, sp.o,
                                     __int_as_float(it));
                int slot = -1;
                if (STAGED) {
                    int x0, x1, y0, y1;
                    if (rg.splats)
                        unpack_rect(sp.rect, x0, x1, y0, y1);
                    else if (st.rects)
                        unpack_rect(st.rects[it], x0, x1, y0, y1);
                    else
                        tile_rect(bw, vr.tiles_x, vr.tiles_y, make_float2(sp.x, sp.y),
                                  st.radii[it], x0, x1, y0, y1);
                    const int w = x1 - x0, lx = tx - x0, ly = ty - y0;
                    const int base = st.first ? (int)st.first[it] : st.cum[it] - w * (y1 - y0);
                    const int sl = base + ly * w + lx;
                    // radii / cum_tiles_hit inconsistent with the binning: no slot
                    if (lx >= 0 && lx < w && ly >= 0 && ly < y1 - y0 && sl >= 0 &&
                        sl < st.n_slots)
                        slot = sl;
                }
                s_b[e] = make_float4(sp.ca * kHalfLog2e, sp.cb * kLog2e, sp.cc * kHalfLog2e,
                                     __int_as_float(slot));
                const CullInfo ci = cull_info(sp.x, sp.y, sp.ca, sp.cb, sp.cc, sp.lim2);
                float4 col = sp.col;
                col.w = ci.lim2;                 // colour, acceptance limit
                s_c[e] = col;
                for (int G = 0; G < ngroups; ++G)
                    if (s_rok[G] && group_hit(ci, s_rect[G])) bits |= 1u << G;
            }
            const int word = (e0 >> 6) + wave;
            for (int G = 0; G < ngroups; ++G) {
                const uint64_t m = ballot((bits >> G) & 1u);
                if (lane == 0 && word < NW) s_mask[G][word] = m;
            }
        }
        __syncthreads();
        const int nb = uniform_i(min(DEPTH, bend - range.x + 1));
        // entries t < t0[sl] lie after every final index of the slot's 64 pixels
        int t0[NSLOT];
        int tmin = bend + 1;
#pragma unroll
        for (int sl = 0; sl < NSLOT; ++sl) {
            t0[sl] = max(0, bend - slot_fin[sl]);
            tmin = min(tmin, t0[sl]);
        }
        const int nwords = (nb + 63) >> 6;
        for (int q = tmin >> 6; q < nwords; ++q) {
            uint64_t ms[NSLOT];
            uint64_t m = 0;
#pragma unroll
            for (int sl = 0; sl < NSLOT; ++sl) {
                ms[sl] = uniform64(s_mask[sl * nwaves + wave][q]) & bits_from(t0[sl] - q * 64);
                m |= ms[sl];
            }
            while (m) {
                const int b = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const int t = q * 64 + b;
                const int k = bend - t;
                const float4 a = s_a[t];
                const float4 bb = s_b[t];
                const float4 col = s_c[t];
                float g[NV];
#pragma unroll
                // -0 + x is x exactly, so the first slot's sums need no add
                for (int v = 0; v < NV; ++v) g[v] = -0.f;
                uint64_t any = 0;
                bool shared = false;
                if constexpr (quad) {
                    {
                        // quadrant slots: slots 0/2 and 1/3 share a
                        // pixel column, 0/1 and 2/3 a row, so dx, dy and the
                        // row terms b' dy, c' dy of sigma are computed once per
                        // column / row (the same roundings as sigma2 per pixel)
                        shared = true;
                        const float dxa = a.x - ps[0].px, dxb = a.x - ps[1].px;
#pragma unroll
                        for (int r = 0; r < 2; ++r) {
                            const int s0 = 2 * r;
                            const bool h0 = (ms[s0] >> b) & 1ull, h1 = (ms[s0 + 1] >> b) & 1ull;
                            if (!(h0 || h1)) continue;
                            const float dy = a.y - ps[s0].py;
                            const float bdy = bb.y * dy, cdy = bb.z * dy;
                            if (h0)
                                any |= bwd_pixel<CONSISTENT, NV>(
                                    ps[s0], dxa, dy, sigma2_terms(bb.x, dxa, bdy, cdy, dy), a,
                                    bb, col, k, g);
                            if (h1)
                                any |= bwd_pixel<CONSISTENT, NV>(
                                    ps[s0 + 1], dxb, dy, sigma2_terms(bb.x, dxb, bdy, cdy, dy),
                                    a, bb, col, k, g);
                        }
                    }
                }
                if (!shared) {
#pragma unroll
                    for (int sl = 0; sl < NSLOT; ++sl)
                        if ((ms[sl] >> b) & 1ull) {
                            const float dx = a.x - ps[sl].px, dy = a.y - ps[sl].py;
                            any |= bwd_pixel<CONSISTENT, NV>(ps[sl], dx, dy, sigma2(bb, dx, dy),
                                                             a, bb, col, k, g);
                        }
                }
                if (any == 0) continue;
                // lane j of every row gets the row sum of value j, then the
                // four rows are summed with the row swaps; lanes 0..NV-1 store
                // the wave's sums
                float val = row_sums_split<NV>(g, lane);
                {
                    const auto r = __builtin_amdgcn_permlane16_swap(
                        __float_as_uint(val), __float_as_uint(val), false, false);
                    val = __uint_as_float(r[0]) + __uint_as_float(r[1]);
                }
                {
                    const auto r = __builtin_amdgcn_permlane32_swap(
                        __float_as_uint(val), __float_as_uint(val), false, false);
                    val = __uint_as_float(r[0]) + __uint_as_float(r[1]);
                }
                if (lane < NV) s_g[t][lane] = val;
                if (lane == 0) s_flag[t] = 1;

            }
        }
        __syncthreads();
        for (int e = tid; e < nb; e += nthr) {
            if (!s_flag[e]) continue;
            float r[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                r[v] = s_g[e][v];
                s_g[e][v] = 0.f;
            }
            s_flag[e] = 0;
            finish_grads<NV>(r, s_b[e], s_a[e].z, s_a[e].x, s_a[e].y);
            if (STAGED) {
                const int slot = __float_as_int(s_b[e].w);
                if (slot < 0) continue;
                store_record<ABS>(st.rec, (size_t)slot, make_float4(r[0], r[1], r[2], r[3]),
                                  make_float4(r[4], r[5], r[6], r[7]),
                                  make_float4(r[NV - 1], ABS ? r[8] : 0.f, ABS ? r[9] : 0.f,
                                              0.f));
                atomicOr(&st.bits[slot >> 5], 1u << (slot & 31));
                continue;
            }
            const int it = __float_as_int(s_a[e].w);
            const int gid = it + gshift;
            float* vc = v_rgb + (size_t)rg.ch_stride * gid + rg.ch_off;
            unsafeAtomicAdd(vc + 0, r[0]);
            if (rg.ch_n > 1) unsafeAtomicAdd(vc + 1, r[1]);
            if (rg.ch_n > 2) unsafeAtomicAdd(vc + 2, r[2]);
            unsafeAtomicAdd(v_conic + 3 * it + 0, r[3]);
            unsafeAtomicAdd(v_conic + 3 * it + 1, r[4]);
            unsafeAtomicAdd(v_conic + 3 * it + 2, r[5]);
            unsafeAtomicAdd(v_xy + 2 * it + 0, r[6]);
            unsafeAtomicAdd(v_xy + 2 * it + 1, r[7]);
            if (ABS) {
                unsafeAtomicAdd(v_xy_abs + 2 * it + 0, r[8]);
                unsafeAtomicAdd(v_xy_abs + 2 * it + 1, r[9]);
            }
            unsafeAtomicAdd(v_opacity + gid, r[NV - 1]);
        }
        __syncthreads();
    }
}

template <bool CONSISTENT, bool ABS, int DEPTH, bool STAGED, int NSLOT>
__global__ __launch_bounds__(64) void raster_bwd_kernel(
    RasterGrid rg, const int* __restrict__ gids, const int2* __restrict__ bins,
    const float2* __restrict__ xys, const float* __restrict__ conics,
    const float* __restrict__ colors, const float* __restrict__ opac,
    const float* __restrict__ bgs, const float* __restrict__ final_Ts,
    const int* __restrict__ final_idx, const float* __restrict__ v_out_img,
    const float* __restrict__ v_out_alpha, float* __restrict__ v_rgb,
    float* __restrict__ v_opacity, float* __restrict__ v_xy, float* __restrict__ v_xy_abs,
    float* __restrict__ v_conic, StagedOut st) {
    const int gt = xcd_tile(blockIdx.x, rg.total_tiles);
    if (gt >= rg.total_tiles) return;
    raster_bwd_tile<CONSISTENT, ABS, DEPTH, STAGED, NSLOT>(
        rg, gt, gids, bins, xys, conics, colors, opac, bgs, final_Ts, final_idx, v_out_img,
        v_out_alpha, v_rgb, v_opacity, v_xy, v_xy_abs, v_conic, st);
}

// Work-queue form of the staged batch backward (see raster_fwd_queue_kernel).
template <bool CONSISTENT, int DEPTH, int NSLOT>
__global__ __launch_bounds__(64) void raster_bwd_queue_kernel(
    RasterGrid rg, const int* __restrict__ order, int* __restrict__ counter,
    const int* __restrict__ gids, const int2* __restrict__ bins,
    const float2* __restrict__ xys, const float* __restrict__ conics,
    const float* __restrict__ colors, const float* __restrict__ opac,
    const float* __restrict__ bgs, const float* __restrict__ final_Ts,
    const int* __restrict__ final_idx, const float* __restrict__ v_out_img, StagedOut st) {
    for (;;) {
        const int gt = next_queue_item(counter, order, rg.total_tiles);
        if (gt == -1) return;
        if ((unsigned)gt >= (unsigned)rg.total_tiles) continue;   // never for a valid order
        raster_bwd_tile<CONSISTENT, false, DEPTH, true, NSLOT>(
            rg, gt, gids, bins, xys, conics, colors, opac, bgs, final_Ts, final_idx, v_out_img,
            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, st);
    }
}

// Per Gaussian: the sum of its flagged staged records, in slot order, as the
// final gradients (every output row is written; a Gaussian with no flagged
// record gets zeros).  num_isect bounds the slots (a radius or cum_tiles_hit
// inconsistent with it yields zeros rather than an out-of-range read).
template <bool ABS>
__global__ __launch_bounds__(256) void raster_bwd_reduce_kernel(
    Camera c, int num_isect, const float2* __restrict__ xys, const int* __restrict__ radii,
    const uint2* __restrict__ rects, const int* __restrict__ cum, const float* __restrict__ rec,
    const uint32_t* __restrict__ bits, int nch, float* __restrict__ v_rgb,
    float* __restrict__ v_opacity, float* __restrict__ v_xy, float* __restrict__ v_xy_abs,
    float* __restrict__ v_conic) {
    // no early exit: the wave-wide pass of sum_records needs every lane
    const int g = blockIdx.x * 256 + threadIdx.x;
    const int lane = lane_id();
    const bool own = g < c.num_points;
    float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0;
    int off = 0, end = 0;
    if (own && radii[g] > 0) {
        int x0, x1, y0, y1;
        if (rects)
            unpack_rect(rects[g], x0, x1, y0, y1);
        else
            tile_rect(c.block_width, c.tiles_x, c.tiles_y, xys[g], radii[g], x0, x1, y0, y1);
        end = cum[g];
        off = end - (x1 - x0) * (y1 - y0);
        if (off < 0 || end > num_isect) off = end = 0;     // inconsistent input: zeros
    }
    sum_records<ABS>(off, end, rec, bits, lane, a0, a1, a2);
    if (!own) return;
    float* vc = v_rgb + (size_t)nch * g;
    vc[0] = a0.x;
    if (nch > 1) vc[1] = a0.y;
    if (nch > 2) vc[2] = a0.z;
    v_conic[3 * (size_t)g + 0] = a0.w;
    v_conic[3 * (size_t)g + 1] = a1.x;
    v_conic[3 * (size_t)g + 2] = a1.y;
    v_xy[2 * (size_t)g + 0] = a1.z;
    v_xy[2 * (size_t)g + 1] = a1.w;
    v_opacity[g] = a2.x;
    if (ABS) {
        v_xy_abs[2 * (size_t)g + 0] = a2.y;
        v_xy_abs[2 * (size_t)g + 1] = a2.z;
    }
}

// Per item of a batch (one grid row per view): the sum of the item's flagged
// records, slots [item_first, item_first + its rectangle's tile count), as the
// item's v_xy and v_conic and its colour and opacity gradients
// v_rgbo = (rgb, opacity); every item row of every view is written (zeros for
// items without records).  Slots beyond n_slots yield zeros.
__global__ __launch_bounds__(256) void raster_bwd_batch_reduce_kernel(
    const ViewRec* __restrict__ views, int n_slots, const int* __restrict__ radii,
    const uint2* __restrict__ item_rects, const uint32_t* __restrict__ item_first,
    const float* __restrict__ rec, const uint32_t* __restrict__ bits, float* __restrict__ v_xy,
    float* __restrict__ v_conic, float4* __restrict__ v_rgbo) {
    const ViewRec vr = views[blockIdx.y];
    if ((int)blockIdx.x * 256 >= vr.n) return;        // uniform over the block
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int lane = lane_id();
    const bool own = k < vr.n;
    const size_t it = (size_t)vr.item_off + (own ? k : 0);
    float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0;
    int off = 0, end = 0;
    if (own && radii[it] > 0) {
        const int cnt = (int)rect_count(item_rects[it]);
        if (cnt > 0) {
            const long long f = item_first[it];
            if (f + cnt <= (long long)n_slots) {
                off = (int)f;
                end = off + cnt;
            }
        }
    }
    sum_records<false>(off, end, rec, bits, lane, a0, a1, a2);
    if (!own) return;
    v_conic[3 * it + 0] = a0.w;
    v_conic[3 * it + 1] = a1.x;
    v_conic[3 * it + 2] = a1.y;
    v_xy[2 * it + 0] = a1.z;
    v_xy[2 * it + 1] = a1.w;
    v_rgbo[it] = make_float4(a0.x, a0.y, a0.z, a2.x);
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

__global__ void raster_empty_kernel(size_t n_pix, int nch, const float* __restrict__ bg,
                                    float* __restrict__ final_Ts, int* __restrict__ final_idx,
                                    float* __restrict__ out_img) {
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n_pix;
         p += (size_t)gridDim.x * blockDim.x) {
        final_Ts[p] = 0.f;
        final_idx[p] = 0;
        for (int k = 0; k < nch; ++k) out_img[(size_t)nch * p + k] = bg[k];
    }
}

int tile_grid(int total_tiles) { return 8 * ((total_tiles + 7) / 8); }

// Workgroups of a persistent work-queue grid: the CUs times the workgroups of
// `kernel` one CU holds at once (at most one per tile).
template <typename K>
hipError_t queue_grid(K kernel, int threads, int total_tiles, int* grid) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0);
    if (e != hipSuccess) return e;
    const long long n = (long long)(cus > 0 ? cus : 1) * (per_cu > 0 ? per_cu : 1);
    *grid = (int)(n < total_tiles ? n : total_tiles);
    return hipSuccess;
}

// Stable-pass key of tile gt: the leading zeros of its list length (31 for
// an empty list), so that a stable ascending pass puts the longest lists
// first; 5 bits.
__global__ __launch_bounds__(256) void tile_order_keys_kernel(RasterGrid g,
                                                              const int2* __restrict__ bins,
                                                              uint32_t* __restrict__ keys) {
    const int gt = blockIdx.x * 256 + threadIdx.x;
    if (gt >= g.total_tiles) return;
    const ViewRec vr = resolve_view(g, gt);
    const int2 b = bins[vr.bin_off + gt - vr.tile_off];
    const int len = b.y > b.x ? b.y - b.x : 0;
    keys[gt] = len > 0 ? (uint32_t)__clz(len) : 31u;
}

}  // namespace

SortPlan order_sort_plan(int total_tiles) { return sort_plan(total_tiles, 0, 5); }

hipError_t tile_order(const RasterGrid& g, const int* tile_bins, uint32_t* keys,
                      uint32_t* keys_out, uint32_t* tmp_k, uint32_t* tmp_v, int* order,
                      void* sort_scratch, hipStream_t s) {
    if (g.total_tiles <= 0) return hipSuccess;
    tile_order_keys_kernel<<<(g.total_tiles + 255) / 256, 256, 0, s>>>(
        g, reinterpret_cast<const int2*>(tile_bins), keys);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sort_pairs(order_sort_plan(g.total_tiles), keys, nullptr, keys_out,
                      reinterpret_cast<uint32_t*>(order), tmp_k, tmp_v, g.total_tiles, nullptr,
                      sort_scratch, nullptr, s);
}

ViewRec single_view(const Camera& c) {
    ViewRec v{};
    v.img_w = c.img_w;
    v.img_h = c.img_h;
    v.tiles_x = c.tiles_x;
    v.tiles_y = c.tiles_y;
    v.fx = c.fx;
    v.fy = c.fy;
    v.cx = c.cx;
    v.cy = c.cy;
    v.tan_fovx = c.tan_fovx;
    v.tan_fovy = c.tan_fovy;
    v.n = c.num_points;
    return v;
}

RasterGrid single_grid(const Camera& c) {
    RasterGrid g{};
    g.v0 = single_view(c);
    g.views = nullptr;
    g.num_views = 1;
    g.total_tiles = c.tiles_x * c.tiles_y;
    g.block_width = c.block_width;
    g.n_items = c.num_points;
    g.fwd_depth = 0;
    g.bwd_depth = 0;
    g.ch_stride = 3;
    g.ch_off = 0;
    g.ch_n = 3;
    return g;
}

hipError_t raster_empty(const Camera& c, const float* background, float* final_Ts,
                        int* final_idx, float* out_img, hipStream_t s, int nch) {
    const size_t n = (size_t)c.img_w * c.img_h;
    if (n == 0) return hipSuccess;
    const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    raster_empty_kernel<<<blocks, 256, 0, s>>>(n, nch, background, final_Ts, final_idx, out_img);
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

hipError_t raster_fwd_grid(const RasterGrid& g, const int* items_sorted, const int* tile_bins,
                           const float* xys, const float* conics, const float* colors,
                           const float* opacities, const float* backgrounds, float* final_Ts,
                           int* final_idx, float* out_img, hipStream_t s) {
    if (g.total_tiles <= 0) return hipSuccess;
    const int nthr = ((g.block_width * g.block_width + 63) / 64) * 64;
    // default: the smallest batch depth that holds one entry per thread (a
    // 12- or 13-pixel tile has 192 threads: 256, two staging rounds)
    const int depth = g.fwd_depth > 0 ? g.fwd_depth : (nthr <= 64 ? 64 : nthr <= 128 ? 128 : 256);
    const int grid = tile_grid(g.total_tiles);
    const int2* b2 = reinterpret_cast<const int2*>(tile_bins);
    const float2* x2 = reinterpret_cast<const float2*>(xys);
    if (g.order) {
        // work queue: a persistent grid takes the tiles in g.order
        hipError_t e = fill_zero(g.queue_counters, 4, s);
        if (e != hipSuccess) return e;
        int qgrid = 0;
#define JSA_QUEUE(D)                                                                         \
    do {                                                                                     \
        if ((e = queue_grid(raster_fwd_queue_kernel<D>, nthr, g.total_tiles, &qgrid)) !=     \
            hipSuccess)                                                                      \
            return e;                                                                        \
        raster_fwd_queue_kernel<D><<<qgrid, nthr, 0, s>>>(                                   \
            g, g.order, g.queue_counters, items_sorted, b2, x2, conics, colors, opacities,   \
            backgrounds, final_Ts, final_idx, out_img);                                      \
    } while (0)
        switch (depth) {
            case 64: JSA_QUEUE(64); break;
            case 128: JSA_QUEUE(128); break;
            case 256: JSA_QUEUE(256); break;
            default: return hipErrorInvalidValue;
        }
#undef JSA_QUEUE
        return hipGetLastError();
    }
#define JSA_LAUNCH(D)                                                                        \
    raster_fwd_kernel<D><<<grid, nthr, 0, s>>>(g, items_sorted, b2, x2, conics, colors,     \
                                               opacities, backgrounds, final_Ts, final_idx,  \
                                               out_img)
    switch (depth) {
        case 64: JSA_LAUNCH(64); break;
        case 128: JSA_LAUNCH(128); break;
        case 256: JSA_LAUNCH(256); break;
        default: return hipErrorInvalidValue;
    }
#undef JSA_LAUNCH
    return hipGetLastError();
}

hipError_t raster_fwd_splats(const Camera& c, int nch, const int* gids_sorted,
                             const int* tile_bins, const float4* splats,
                             const float* background, float* final_Ts, int* final_idx,
                             float* out_img, hipStream_t s) {
    if (nch < 1 || nch > 3 || !splats) return hipErrorInvalidValue;
    RasterGrid g = single_grid(c);
    g.ch_stride = nch;
    g.ch_n = nch;
    g.splats = splats;
    return raster_fwd_grid(g, gids_sorted, tile_bins, nullptr, nullptr, nullptr, nullptr,
                           background, final_Ts, final_idx, out_img, s);
}

hipError_t raster_fwd(const Camera& c, const int* gids_sorted, const int* tile_bins,
                      const float* xys, const float* conics, const float* colors,
                      const float* opacities, const float* background, float* final_Ts,
                      int* final_idx, float* out_img, hipStream_t s) {
    return raster_fwd_grid(single_grid(c), gids_sorted, tile_bins, xys, conics, colors,
                           opacities, background, final_Ts, final_idx, out_img, s);
}

hipError_t raster_bwd_grid(const RasterGrid& g, const int* items_sorted, const int* tile_bins,
                           const float* xys, const float* conics, const float* colors,
                           const float* opacities, const float* background,
                           const float* final_Ts, const int* final_idx, const float* v_out,
                           const float* v_out_alpha, float* v_colors, float* v_opacity,
                           float* v_xy, float* v_xy_abs, float* v_conic, int consistent,
                           hipStream_t s) {
    if (g.total_tiles <= 0) return hipSuccess;
    const int slots = bwd_slots(g.block_width);
    const int nthr = bwd_threads(g.block_width, slots);
    const int grid = tile_grid(g.total_tiles);
    // 64 Gaussians per batch by default: the LDS partial sums (depth x 9..11 x 4
    // floats) otherwise limit the workgroups per CU (measured fastest on the
    // 1080p, 800x800 and mixed-size workloads)
    const int depth = g.bwd_depth > 0 ? g.bwd_depth : 64;
    if (depth != 64 && depth != 128 && depth != 256) return hipErrorInvalidValue;
    const int* gids_sorted = items_sorted;
    const int2* b2 = reinterpret_cast<const int2*>(tile_bins);
    const float2* x2 = reinterpret_cast<const float2*>(xys);
#define JSA_LAUNCH_S(CONS, ABS, D, S)                                                        \
    raster_bwd_kernel<CONS, ABS, D, false, S><<<grid, nthr, 0, s>>>(                        \
        g, gids_sorted, b2, x2, conics, colors, opacities, background, final_Ts, final_idx,  \
        v_out, v_out_alpha, v_colors, v_opacity, v_xy, v_xy_abs, v_conic, StagedOut{})
#define JSA_LAUNCH_D(CONS, ABS, D)                                                           \
    do {                                                                                     \
        if (slots == 4) JSA_LAUNCH_S(CONS, ABS, D, 4);                                       \
        else JSA_LAUNCH_S(CONS, ABS, D, 2);                                                  \
    } while (0)
#define JSA_LAUNCH(CONS, ABS)                                                                \
    do {                                                                                     \
        if (depth == 64) JSA_LAUNCH_D(CONS, ABS, 64);                                        \
        else if (depth == 128) JSA_LAUNCH_D(CONS, ABS, 128);                                 \
        else JSA_LAUNCH_D(CONS, ABS, 256);                                                   \
    } while (0)
    if (consistent) {
        if (v_xy_abs) JSA_LAUNCH(true, true); else JSA_LAUNCH(true, false);
    } else {
        if (v_xy_abs) JSA_LAUNCH(false, true); else JSA_LAUNCH(false, false);
    }
#undef JSA_LAUNCH
#undef JSA_LAUNCH_D
#undef JSA_LAUNCH_S
    return hipGetLastError();
}

size_t raster_bwd_staged_bytes(int num_isect) {
    const size_t n = (size_t)(num_isect > 0 ? num_isect : 1);
    return ((48 * n + 255) & ~size_t(255)) + ((4 * ((n + 31) / 32) + 255) & ~size_t(255));
}

hipError_t raster_bwd_staged(const Camera& c, int nch, const int* gids_sorted,
                             const int* tile_bins, int num_isect, const float* xys,
                             const float* conics, const float* colors, const float* opacities,
                             const float* background, const float* final_Ts,
                             const int* final_idx, const float* v_out, const float* v_out_alpha,
                             const int* radii, const uint2* rects, const int* cum_tiles_hit,
                             void* ws,
                             float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs,
                             float* v_conic, int consistent, hipStream_t s, bool records_only,
                             const float4* splats) {
    if (c.num_points <= 0) return hipSuccess;
    if (nch < 1 || nch > 3) return hipErrorInvalidValue;
    uint8_t* w = static_cast<uint8_t*>(ws);
    const size_t n = (size_t)(num_isect > 0 ? num_isect : 1);
    StagedOut st;
    st.rec = reinterpret_cast<float*>(w);
    st.bits = reinterpret_cast<uint32_t*>(w + ((48 * n + 255) & ~size_t(255)));
    st.cum = cum_tiles_hit;
    st.first = nullptr;
    st.rects = rects;
    st.radii = radii;
    st.n_slots = num_isect;
    hipError_t e = hipSuccess;
    if (num_isect > 0) {
        const size_t words = ((size_t)num_isect + 31) / 32;
        if ((e = fill_zero(st.bits, 4 * words, s)) != hipSuccess) return e;
        RasterGrid g = single_grid(c);
        g.ch_stride = nch;
        g.ch_n = nch;
        g.splats = splats;
        const int slots = bwd_slots(g.block_width);
        const int nthr = bwd_threads(g.block_width, slots);
        const int grid = tile_grid(g.total_tiles);
        const int2* b2 = reinterpret_cast<const int2*>(tile_bins);
        const float2* x2 = reinterpret_cast<const float2*>(xys);
#define JSA_LAUNCH_S(CONS, ABS, S)                                                           \
    raster_bwd_kernel<CONS, ABS, 64, true, S><<<grid, nthr, 0, s>>>(                        \
        g, gids_sorted, b2, x2, conics, colors, opacities, background, final_Ts, final_idx,  \
        v_out, v_out_alpha, nullptr, nullptr, nullptr, nullptr, nullptr, st)
#define JSA_LAUNCH(CONS, ABS)                                                                \
    do {                                                                                     \
        if (slots == 4) JSA_LAUNCH_S(CONS, ABS, 4);                                          \
        else JSA_LAUNCH_S(CONS, ABS, 2);                                                     \
    } while (0)
        if (consistent) {
            if (v_xy_abs) JSA_LAUNCH(true, true); else JSA_LAUNCH(true, false);
        } else {
            if (v_xy_abs) JSA_LAUNCH(false, true); else JSA_LAUNCH(false, false);
        }
#undef JSA_LAUNCH
#undef JSA_LAUNCH_S
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    stage_mark(kMarkRecords, s);
    if (records_only) return hipSuccess;
    const int blocks = (c.num_points + 255) / 256;
    const float2* x2 = reinterpret_cast<const float2*>(xys);
    if (v_xy_abs)
        raster_bwd_reduce_kernel<true><<<blocks, 256, 0, s>>>(
            c, num_isect, x2, radii, rects, cum_tiles_hit, st.rec, st.bits, nch, v_colors,
            v_opacity, v_xy, v_xy_abs, v_conic);
    else
        raster_bwd_reduce_kernel<false><<<blocks, 256, 0, s>>>(
            c, num_isect, x2, radii, rects, cum_tiles_hit, st.rec, st.bits, nch, v_colors,
            v_opacity, v_xy, nullptr, v_conic);
    return hipGetLastError();
}

hipError_t raster_bwd_batch_staged(const RasterGrid& g, int max_view_n, const int* items_sorted,
                                   const int* tile_bins, int cap, const float* xys,
                                   const float* conics, const float* colors,
                                   const float* opacities, const float* backgrounds,
                                   const float* final_Ts, const int* final_idx,
                                   const float* v_out, const int* radii, const uint2* item_rects,
                                   const uint32_t* item_first, void* ws, float* v_xy,
                                   float* v_conic, float* v_rgbo, int consistent,
                                   hipStream_t s) {
    if (g.total_tiles <= 0 || g.num_views <= 0 || max_view_n <= 0) return hipSuccess;
    if (g.ch_n != 3 || g.ch_stride != 3 || g.ch_off != 0) return hipErrorInvalidValue;
    const int depth = g.bwd_depth > 0 ? g.bwd_depth : 64;
    if (depth != 64 && depth != 128 && depth != 256) return hipErrorInvalidValue;
    uint8_t* w = static_cast<uint8_t*>(ws);
    const size_t n = (size_t)(cap > 0 ? cap : 1);
    StagedOut st;
    st.rec = reinterpret_cast<float*>(w);
    st.bits = reinterpret_cast<uint32_t*>(w + ((48 * n + 255) & ~size_t(255)));
    st.cum = nullptr;
    st.first = item_first;
    st.rects = item_rects;
    st.radii = radii;
    st.n_slots = cap;
    hipError_t e = fill_zero(st.bits, 4 * ((n + 31) / 32), s);
    if (e != hipSuccess) return e;
    const int slots = bwd_slots(g.block_width);
    const int nthr = bwd_threads(g.block_width, slots);
    const int grid = tile_grid(g.total_tiles);
    const int2* b2 = reinterpret_cast<const int2*>(tile_bins);
    const float2* x2 = reinterpret_cast<const float2*>(xys);
    // work queue (g.order): a persistent grid takes the tiles in that order
    if (g.order && (e = fill_zero(g.queue_counters + 1, 4, s)) != hipSuccess) return e;
#define JSA_LAUNCH_S(CONS, D, S)                                                             \
    do {                                                                                     \
        if (g.order) {                                                                       \
            int qgrid = 0;                                                                   \
            if ((e = queue_grid(raster_bwd_queue_kernel<CONS, D, S>, nthr, g.total_tiles,    \
                                &qgrid)) != hipSuccess)                                      \
                return e;                                                                    \
            raster_bwd_queue_kernel<CONS, D, S><<<qgrid, nthr, 0, s>>>(                      \
                g, g.order, g.queue_counters + 1, items_sorted, b2, x2, conics, colors,      \
                opacities, backgrounds, final_Ts, final_idx, v_out, st);                     \
        } else {                                                                             \
            raster_bwd_kernel<CONS, false, D, true, S><<<grid, nthr, 0, s>>>(                \
                g, items_sorted, b2, x2, conics, colors, opacities, backgrounds, final_Ts,   \
                final_idx, v_out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, st); \
        }                                                                                    \
    } while (0)
#define JSA_LAUNCH_D(CONS, D)                                                                \
    do {                                                                                     \
        if (slots == 4) JSA_LAUNCH_S(CONS, D, 4);                                            \
        else JSA_LAUNCH_S(CONS, D, 2);                                                       \
    } while (0)
#define JSA_LAUNCH(CONS)                                                                     \
    do {                                                                                     \
        if (depth == 64) JSA_LAUNCH_D(CONS, 64);                                             \
        else if (depth == 128) JSA_LAUNCH_D(CONS, 128);                                      \
        else JSA_LAUNCH_D(CONS, 256);                                                        \
    } while (0)
    if (consistent) JSA_LAUNCH(true); else JSA_LAUNCH(false);
#undef JSA_LAUNCH
#undef JSA_LAUNCH_D
#undef JSA_LAUNCH_S
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const dim3 rgrid((max_view_n + 255) / 256, g.num_views);
    raster_bwd_batch_reduce_kernel<<<rgrid, 256, 0, s>>>(
        g.views, cap, radii, item_rects, item_first, st.rec, st.bits, v_xy, v_conic,
        reinterpret_cast<float4*>(v_rgbo));
    return hipGetLastError();
}

hipError_t raster_bwd(const Camera& c, const int* gids_sorted, const int* tile_bins,
                      const float* xys, const float* conics, const float* colors,
                      const float* opacities, const float* background, const float* final_Ts,
                      const int* final_idx, const float* v_out, const float* v_out_alpha,
                      float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs,
                      float* v_conic, int consistent, hipStream_t s) {
    return raster_bwd_grid(single_grid(c), gids_sorted, tile_bins, xys, conics, colors,
                           opacities, background, final_Ts, final_idx, v_out, v_out_alpha,
                           v_colors, v_opacity, v_xy, v_xy_abs, v_conic, consistent, s);
}

}  // namespace jsa
