// Shared device utilities for the gfx950 (CDNA4) Gaussian-splatting kernels:
// wave64 reductions and scans on DPP, ballot helpers, the packed tile
// rectangle, and the internal host-side launch interface shared by the .hip
// translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace jsa {

constexpr int kWave = 64;
constexpr int kMaxBlockWidth = 16;                 // tile side, bw*bw <= 256 threads
constexpr int kMaxTileThreads = kMaxBlockWidth * kMaxBlockWidth;

// ---------------------------------------------------------------------------
// Error codes shared with the C ABI
// ---------------------------------------------------------------------------
enum Status : int {
    kOk = 0,
    kErrInvalid = 1,
    kErrHip = 2,
    kErrCapacity = 3,
    kErrWorkspace = 4,
};

// ---------------------------------------------------------------------------
// Lane helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// The lane's bit of a wave-uniform mask as a lane predicate (the mask itself
// serves as the select mask: no per-lane shift or compare).
__device__ __forceinline__ bool lane_in(uint64_t m) {
    return __builtin_amdgcn_inverse_ballot_w64(m);
}

// DPP controls (GFX9 encoding).
constexpr int kDppQuadPerm1032 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppQuadPerm2301 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppRowRor4 = 0x124;
constexpr int kDppRowRor8 = 0x128;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
        0, __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xF, false));
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xF, false);
}

// Sum over all 64 lanes; the total ends up in lane 63.  Requires a full EXEC
// mask (call from wave-uniform control flow).
__device__ __forceinline__ float wave_sum_to_lane63(float v) {
    v += dpp_f<kDppQuadPerm1032>(v);
    v += dpp_f<kDppQuadPerm2301>(v);
    v += dpp_f<kDppRowRor4>(v);
    v += dpp_f<kDppRowRor8>(v);
    v += dpp_f<kDppRowBcast15, 0xA>(v);
    v += dpp_f<kDppRowBcast31, 0xC>(v);
    return v;
}

// Sum over each 16-lane row; every lane of a row ends with the row's total.
__device__ __forceinline__ float row_sum16(float v) {
    v += dpp_f<kDppQuadPerm1032>(v);
    v += dpp_f<kDppQuadPerm2301>(v);
    v += dpp_f<kDppRowRor4>(v);
    v += dpp_f<kDppRowRor8>(v);
    return v;
}

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Orders LDS traffic between the lanes of one wave: without it the compiler
// may forward a lane's own earlier store to a later load of the same address
// and miss a store made by another lane in between.  Emits no instruction
// (a wave's LDS operations already execute in program order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive max-scan over the 64 lanes for non-negative values (0 is the
// identity): Hillis-Steele row shifts, then the two row broadcasts.
__device__ __forceinline__ int wave_max_scan_nonneg(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true));   // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true));   // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true));   // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true));   // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppRowBcast15, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppRowBcast31, 0xC, 0xF, false));
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(
        __builtin_bit_cast(int, wave_sum_to_lane63(v)), 63));
}

__device__ __forceinline__ int wave_max_i(int v) {
    v = max(v, __shfl_xor(v, 1));
    v = max(v, __shfl_xor(v, 2));
    v = max(v, __shfl_xor(v, 4));
    v = max(v, __shfl_xor(v, 8));
    v = max(v, __shfl_xor(v, 16));
    v = max(v, __shfl_xor(v, 32));
    return v;
}

// Inclusive prefix sum over the wave (Kogge-Stone on shuffles).
template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T n = __shfl_up(v, o);
        if (lane >= o) v += n;
    }
    return v;
}

// Tile of workgroup b in a grid of 8 * ceil(num_tiles / 8) workgroups:
// workgroups are dealt to the 8 XCDs round robin, so XCD b & 7 gets the
// contiguous band [x * per, (x + 1) * per) of tiles (indexes past num_tiles
// are the caller's to skip).
__device__ __forceinline__ int xcd_tile(int b, int num_tiles) {
    const int per = (num_tiles + 7) / 8;
    return (b & 7) * per + (b >> 3);
}

// Mask of the lanes of this wave whose `bits`-bit digit equals this lane's
// (restricted to lanes in `valid`): one ballot per digit bit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, int bits, uint64_t valid) {
    uint64_t peers = valid;
    for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    return peers;
}

// Exclusive scan of one value per thread over a 256-thread block (s_tmp: 4
// words of LDS; two barriers).
__device__ __forceinline__ uint32_t block_exclusive_scan256(uint32_t v, uint32_t* s_tmp) {
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const uint32_t incl = wave_inclusive_sum(v);
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    uint32_t off = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) off += (w < wave) ? s_tmp[w] : 0u;
    __syncthreads();
    return off + incl - v;
}

// Tile rectangle [x0, x1) x [y0, y1) of a Gaussian from its centre and integer
// radius: forward.cu:152-161 / helpers.h:9-32 (the binning and the staged
// backward must agree bit for bit, so both use this one function).
__device__ __forceinline__ void tile_rect(int block_width, int tiles_x, int tiles_y, float2 xy,
                                          int rad, int& x0, int& x1, int& y0, int& y1) {
    const float bwf = (float)block_width;
    const float tcx = xy.x / bwf, tcy = xy.y / bwf, tr = (float)rad / bwf;
    x0 = min(max(0, (int)(tcx - tr)), tiles_x);
    x1 = min(max(0, (int)(tcx + tr + 1.f)), tiles_x);
    y0 = min(max(0, (int)(tcy - tr)), tiles_y);
    y1 = min(max(0, (int)(tcy + tr + 1.f)), tiles_y);
}

// Tile rectangle [x0, x1) x [y0, y1) packed as (x0 | y0 << 16, w | h << 16),
// w = x1 - x0, h = y1 - y0; it covers w * h tiles (tile counts below 2^16 per
// axis).
__device__ __forceinline__ uint2 pack_rect(int x0, int x1, int y0, int y1) {
    return make_uint2((uint32_t)x0 | ((uint32_t)y0 << 16),
                      (uint32_t)(x1 - x0) | ((uint32_t)(y1 - y0) << 16));
}

__device__ __forceinline__ uint32_t rect_count(uint2 rc) {
    return (rc.y & 0xFFFFu) * (rc.y >> 16);
}

__device__ __forceinline__ void unpack_rect(uint2 rc, int& x0, int& x1, int& y0, int& y1) {
    x0 = (int)(rc.x & 0xFFFFu);
    y0 = (int)(rc.x >> 16);
    x1 = x0 + (int)(rc.y & 0xFFFFu);
    y1 = y0 + (int)(rc.y >> 16);
}

// log2(255 o): a pair passes the per-pixel test o exp(-sigma) >= 1/255 with
// sigma >= 0 iff sigma * log2(e) lies in [0, alpha_limit2(o)].  One function
// for every kernel that evaluates or precomputes it (the projection's splat
// records, the rectangles below and both rasterizers must agree bit for bit).
__device__ __forceinline__ float alpha_limit2(float o) {
    return __builtin_amdgcn_logf(255.f * o);
}

// Tile rectangle of a Gaussian (centre xy, radius rad > 0, conic (a, b, c),
// opacity o) restricted to the tiles holding a pixel centre it can blend
// into: those where o exp(-sigma) >= 1/255, i.e. the ellipse
// sigma = (a dx^2 + c dy^2) / 2 + b dx dy <= ln(255 o), whose extents are
// sqrt(2 ln(255 o) cov_xx) and sqrt(2 ln(255 o) cov_yy) (cov = conic^-1).
// The limit and the extents get generous margins (1% + 0.05 on the limit,
// 0.1% + 0.05 px on the extents: float32 rounding of sigma at a pixel stays
// far below, even with cancellation in strongly elongated conics).  A
// Gaussian with log2(255 o) < 0 -- culled whole by the rasterizer -- gets an
// empty rectangle; a conic that is not positive definite keeps the full one.
// Packed as pack_rect.
__device__ __forceinline__ uint2 opacity_tile_rect(int bw, int tiles_x, int tiles_y, float2 xy,
                                                   int rad, float a, float b, float c, float o) {
    int x0, x1, y0, y1;
    tile_rect(bw, tiles_x, tiles_y, xy, rad, x0, x1, y0, y1);
    const float lim2 = alpha_limit2(o);
    if (!(lim2 >= 0.f)) return make_uint2((uint32_t)x0 | ((uint32_t)y0 << 16), 0u);
    const float det = a * c - b * b;
    if (a > 0.f && c > 0.f && det > 0.f) {
        const float s = lim2 * 0.69314718f * 1.01f + 0.05f;     // ln(255 o) with margin
        const float hx = sqrtf(2.f * s * c / det) * 1.001f + 0.05f;
        const float hy = sqrtf(2.f * s * a / det) * 1.001f + 0.05f;
        const float bwf = (float)bw, inv = 1.f / bwf;
        // tile t holds the pixel centres [t bw + 0.5, t bw + bw - 0.5]
        const float fx0 = fminf(fmaxf(ceilf((xy.x - hx - bwf + 0.5f) * inv), -1.f), 65535.f);
        const float fx1 = fminf(fmaxf(floorf((xy.x + hx - 0.5f) * inv) + 1.f, -1.f), 65535.f);
        const float fy0 = fminf(fmaxf(ceilf((xy.y - hy - bwf + 0.5f) * inv), -1.f), 65535.f);
        const float fy1 = fminf(fmaxf(floorf((xy.y + hy - 0.5f) * inv) + 1.f, -1.f), 65535.f);
        x0 = max(x0, (int)fx0);
        x1 = min(x1, (int)fx1);
        y0 = max(y0, (int)fy0);
        y1 = min(y1, (int)fy1);
        if (x1 < x0) x1 = x0;
        if (y1 < y0) y1 = y0;
    }
    return pack_rect(x0, x1, y0, y1);
}

// ---------------------------------------------------------------------------
// Staged rasterize-backward records (raster.hip writes them; the reduction
// kernels in raster.hip and project.hip sum them)
// ---------------------------------------------------------------------------
// A record holds one (Gaussian, tile) pair's sums at the pair's intersection
// slot k: 9 floats at rec + 9 k, (rgb 3, conic 3, xy 2, opacity), or with the
// v_xy_abs sums (ABS) 12 floats at rec + 12 k, (rgb 3, conic a), (conic b,
// c, xy 2), (opacity, xy_abs 2, -).  bits holds one bit per slot, set when
// the slot's record was written.  The workspace reserves 48 bytes per slot.
template <bool ABS>
__device__ __forceinline__ const float* record_at(const float* __restrict__ rec, size_t k) {
    return rec + (ABS ? 12 : 9) * k;
}

template <bool ABS>
__device__ __forceinline__ void load_record_from(const float* __restrict__ q, float4& a0,
                                                 float4& a1, float4& a2) {
    if (ABS) {
        const float4* q4 = reinterpret_cast<const float4*>(q);
        a0 = q4[0];
        a1 = q4[1];
        a2 = q4[2];
    } else {
        a0 = make_float4(q[0], q[1], q[2], q[3]);
        a1 = make_float4(q[4], q[5], q[6], q[7]);
        a2 = make_float4(q[8], 0.f, 0.f, 0.f);
    }
}

template <bool ABS>
__device__ __forceinline__ void load_record(const float* __restrict__ rec, size_t k, float4& a0,
                                            float4& a1, float4& a2) {
    load_record_from<ABS>(record_at<ABS>(rec, k), a0, a1, a2);
}

// An all-zero record (zero-initialised, never written): the source of the
// absent records' loads in sum_records, whose +0 terms leave a sum unchanged.
static __device__ __align__(16) float kZeroRecord[12];

template <bool ABS>
__device__ __forceinline__ void store_record(float* __restrict__ rec, size_t k, float4 a0,
                                             float4 a1, float4 a2) {
    if (ABS) {
        float4* q = reinterpret_cast<float4*>(rec) + 3 * k;
        q[0] = a0;
        q[1] = a1;
        q[2] = a2;
    } else {
        float* q = rec + 9 * k;
        q[0] = a0.x;
        q[1] = a0.y;
        q[2] = a0.z;
        q[3] = a0.w;
        q[4] = a1.x;
        q[5] = a1.y;
        q[6] = a1.z;
        q[7] = a1.w;
        q[8] = a2.x;
    }
}
// Row (16-lane) sums of v[0..NV) by recursive halving.  At each DPP level
// (partner lanes l^1, l^2, l+-4, l+-8 within the row: each flips one lane bit
// and keeps the lower ones) the values are taken in pairs: a lane whose level
// bit is set keeps the second value of a pair and passes the first, the other
// lanes the reverse, and each lane adds what its partner passes.  A pair costs
// two selects and one DPP add (two DPP adds as whole row sums, plus the
// selects that pick lane j's value afterwards); an unpaired value is summed
// whole.  Lane j of every row ends with the row sum of value j, j < NV <= 16.
template <int L>
__device__ __forceinline__ float dpp_level(float x) {
    if constexpr (L == 0) return dpp_f<kDppQuadPerm1032>(x);
    else if constexpr (L == 1) return dpp_f<kDppQuadPerm2301>(x);
    else if constexpr (L == 2) return dpp_f<kDppRowRor4>(x);
    else return dpp_f<kDppRowRor8>(x);
}

template <int L, int M>
__device__ __forceinline__ void halve_level(float* v, int lane) {
    const bool hi = (lane >> L) & 1;
#pragma unroll
    for (int i = 0; i < M / 2; ++i) {
        const float a = v[2 * i], b = v[2 * i + 1];
        const float keep = hi ? b : a, pass = hi ? a : b;
        v[i] = keep + dpp_level<L>(pass);
    }
    if (M & 1) v[M / 2] = v[M - 1] + dpp_level<L>(v[M - 1]);
}

template <int NV>
__device__ __forceinline__ float row_sums_split(float* v, int lane) {
    constexpr int M1 = (NV + 1) / 2, M2 = (M1 + 1) / 2, M3 = (M2 + 1) / 2;
    static_assert(NV <= 16 && (M3 + 1) / 2 == 1, "at most 16 values per row");
    halve_level<0, NV>(v, lane);
    halve_level<1, M1>(v, lane);
    halve_level<2, M2>(v, lane);
    halve_level<3, M3>(v, lane);
    return v[0];
}

// Wave sums of v[0..NV): row_sums_split, then the four rows summed with the
// gfx950 row swaps; lane j < NV returns the wave's sum of value j.
template <int NV>
__device__ __forceinline__ float wave_sums_split(float* v, int lane) {
    float val = row_sums_split<NV>(v, lane);
    {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(val),
                                                        __float_as_uint(val), false, false);
        val = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(val),
                                                        __float_as_uint(val), false, false);
        val = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    return val;
}

// Sum of the flagged staged records of slots [off, end) (empty: off == end),
// in slot order, into (a0, a1, a2) of the calling lane.  Ranges of at most 16
// slots are summed by their own lane: the (at most two) bitmap words covering
// the range as one mask (bits outside the range masked off), the set bits
// taken four records at a time with their loads in flight together (a missing
// one reads the zero record).  Longer ranges are summed one at a time
// by the whole wave -- lane l takes slots off + l, off + l + 64, ... -- and
// combined in a fixed tree (wave_sums_split), so that one large Gaussian does
// not hold its wave for tens of rounds; the results stay independent of
// scheduling.
// Called by every lane of the wave.
template <bool ABS>
__device__ __forceinline__ void sum_records(int off, int end, const float* __restrict__ rec,
                                            const uint32_t* __restrict__ bits, int lane,
                                            float4& a0, float4& a1, float4& a2) {
    // 16: 56 us for the backward tail of 1M Gaussians at 1080p (24 the same,
    // 32: 60 us; 12: 62 us and 8: 92 us measured before the two-word mask):
    // longer per-lane runs hold their wave, and each wave-wide range costs one
    // transposed wave reduction
    constexpr int kBig = 16;
    const bool big = end - off > kBig;
    const int cnt = end - off;
    static_assert(kBig <= 33, "a per-lane range spans at most two bitmap words");
    if (cnt > 0 && !big) {
        // the (at most two) bitmap words of the range as one 64-bit mask, so
        // that four set bits of either word are taken per step
        const int wfirst = off >> 5, wlast = (end - 1) >> 5;
        const uint32_t lo_mask = ~0u << (off & 31);
        const int hb = (end - 1) & 31;
        const uint32_t hi_mask = hb == 31 ? ~0u : ((2u << hb) - 1u);
        const uint32_t wa = bits[wfirst];
        const uint32_t wb = wlast > wfirst ? bits[wlast] : 0u;
        uint64_t v = wlast > wfirst ? ((uint64_t)(wb & hi_mask) << 32) | (wa & lo_mask)
                                    : (uint64_t)(wa & lo_mask & hi_mask);
        const int kb = wfirst << 5;
        while (v) {
            int k[4];
            bool ok[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ok[i] = v != 0u;
                k[i] = ok[i] ? kb + __ffsll((unsigned long long)v) - 1 : 0;
                v &= v - 1u;
            }
            float4 r[4][3];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                load_record_from<ABS>(ok[i] ? record_at<ABS>(rec, (size_t)k[i]) : kZeroRecord,
                                      r[i][0], r[i][1], r[i][2]);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float4 q0 = r[i][0], q1 = r[i][1], q2 = r[i][2];
                a0.x += q0.x; a0.y += q0.y; a0.z += q0.z; a0.w += q0.w;
                a1.x += q1.x; a1.y += q1.y; a1.z += q1.z; a1.w += q1.w;
                a2.x += q2.x;
                if (ABS) {
                    a2.y += q2.y;
                    a2.z += q2.z;
                }
            }
        }
    }
    uint64_t bigm = ballot(big);
    while (bigm) {
        const int src = __ffsll((unsigned long long)bigm) - 1;
        bigm &= bigm - 1;
        const int boff = __builtin_amdgcn_readlane(off, src);
        const int bend = __builtin_amdgcn_readlane(end, src);
        float4 b0 = make_float4(0.f, 0.f, 0.f, 0.f), b1 = b0, b2 = b0;
        for (int k = boff + lane; k < bend; k += 64) {
            if (!((bits[k >> 5] >> (k & 31)) & 1u)) continue;
            float4 r0, r1, r2;
            load_record<ABS>(rec, (size_t)k, r0, r1, r2);
            b0.x += r0.x; b0.y += r0.y; b0.z += r0.z; b0.w += r0.w;
            b1.x += r1.x; b1.y += r1.y; b1.z += r1.z; b1.w += r1.w;
            b2.x += r2.x; b2.y += r2.y; b2.z += r2.z;
        }
        float v[11] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w, b2.x, b2.y, b2.z};
        // lane j < NV ends with the wave sum of value j (xy_abs sums only with ABS)
        constexpr int NV = ABS ? 11 : 9;
        const float val = wave_sums_split<NV>(v, lane);
        float t[11];
#pragma unroll
        for (int i = 0; i < NV; ++i)
            t[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(val), i));
        if (lane == src) {
            a0 = make_float4(t[0], t[1], t[2], t[3]);
            a1 = make_float4(t[4], t[5], t[6], t[7]);
            a2 = make_float4(t[8], ABS ? t[9] : 0.f, ABS ? t[10] : 0.f, 0.f);
        }
    }
}


// ---------------------------------------------------------------------------
// Host-side internal interface (implemented in the .hip files)
// ---------------------------------------------------------------------------
struct Camera {
    int num_points;
    int img_w, img_h;
    int block_width;
    int tiles_x, tiles_y;
    float fx, fy, cx, cy;
    float tan_fovx, tan_fovy;
    float glob_scale, clip_thresh;
};

// One view of a render.  A plain render is a single record with zero offsets;
// a batch is a device table of records, ordered by scene, whose items, tiles,
// tile bins and pixels are laid end to end.
struct alignas(16) ViewRec {
    int img_w, img_h, tiles_x, tiles_y;
    float fx, fy, cx, cy;
    float tan_fovx, tan_fovy;
    int g0, n;            // the scene's Gaussians: rows [g0, g0 + n) of the parameter arrays
    int item_off;         // first projected item; item = item_off + (gaussian - g0)
    int tile_off;         // first tile in the batch's tile numbering
    int bin_off;          // first tile bin (view << tile_bits)
    int view;             // row of viewmats and backgrounds
    long long pix_off;    // first pixel in the image-sized buffers
    long long reserved;
};
static_assert(sizeof(ViewRec) == 80, "ViewRec layout");

// Scenes of a batch: Gaussians [g0, g0 + n) seen by views [v_begin, v_end).
struct SceneRec {
    int g0, n, v_begin, v_end;
};

// Launch description of the rasterizer: one view (views == nullptr, v0 used)
// or a device table of num_views records covering total_tiles tiles.
struct RasterGrid {
    ViewRec v0;
    const ViewRec* views;
    int num_views;
    int total_tiles;
    int block_width;
    int n_items;          // sorted entries name items in [0, n_items)
    int fwd_depth;        // Gaussians staged per batch: 64, 128 or 256 (0: workgroup size)
    int bwd_depth;        // likewise for the backward (0: 64)
    // channel pass: channels [ch_off, ch_off + ch_n) (ch_n <= 3) of colors,
    // backgrounds, images and v_colors stored with ch_stride channels per row
    int ch_stride, ch_off, ch_n;
    // work queue (optional): a permutation of the tiles in the order a
    // persistent grid takes them, and two device counters (forward,
    // backward), zeroed by the launch; null: one workgroup per tile
    const int* order;
    int* queue_counters;
    // splat records (optional, single view): per Gaussian three float4
    // (x, y, opacity, alpha_limit2(opacity)), (conic a, b, c, tile rectangle
    // word 0), (colour 3, rectangle word 1), staged instead of gathering
    // xys, opacities, conics, colours and rectangles separately
    const float4* splats;
};

ViewRec single_view(const Camera& c);
RasterGrid single_grid(const Camera& c);

// Profiling marks (jaxsplat_set_stage_marks): stage_mark(k, s) records the
// calling thread's event k, if one is set, on stream s.
enum StageMark : int {
    kMarkEmitted = 0,       // bin_emit: emission done, tile sort next
    kMarkTileSorted = 1,    // bin_emit: tile sort done, bin edges next
    kMarkRecords = 2,       // staged rasterize backward: records done, reduction next
    kMarkBinned = 3,        // bin_emit: bin edges done, per-tile depth order next
    kNumStageMarks = 4,
};
void stage_mark(int k, hipStream_t s);

// Zeroes exactly `bytes` bytes at p (4-byte aligned) with a kernel (no memset
// node when captured in a graph).
hipError_t fill_zero(void* p, size_t bytes, hipStream_t s);
// Sets n_words 32-bit words at p to value, with a kernel.
hipError_t fill_words(void* p, size_t n_words, uint32_t value, hipStream_t s);
hipError_t fill_value_f32x3(float* dst, size_t n_pixels, const float* rgb, hipStream_t s);

// ---- projection ----
size_t project_fwd_scratch_bytes(int n);
// Largest num_points of the projection forward with the library-held scratch.
constexpr int kStaticProjectPoints = 1 << 26;
// *out = the current device's address of the library-held projection scratch
// (project_fwd_scratch_bytes(kStaticProjectPoints) bytes, looked up once per
// device).  Every caller on a device shares it.
hipError_t project_static_scratch(void** out);
// With opacities and rects (both optional), the tile counts and their sums are
// those of the opacity-limited rectangles (opacity_tile_rect), written to rects.
// With opacities, colors (nch <= 3 channels per row) and splats as well, the
// splat records of RasterGrid::splats are written too.
hipError_t project_fwd(const Camera& c, const float* means3d, const float* scales,
                       const float* quats, const float* viewmat, float* cov3d,
                       float* xys, float* depths, int* radii, float* conics,
                       float* compensation, int* num_tiles_hit, int* cum_tiles_hit,
                       const float* opacities, uint2* rects, void* scratch, hipStream_t s,
                       const float* colors = nullptr, int nch = 0, float4* splats = nullptr);
// Opacity-limited tile rectangles (opacity_tile_rect) of projected Gaussians,
// their counts (num_tiles_hit optional) and inclusive sums; scratch
// project_fwd_scratch_bytes(c.num_points).
hipError_t tile_rects(const Camera& c, const float* xys, const int* radii, const float* conics,
                      const float* opacities, uint2* rects, int* num_tiles_hit,
                      int* cum_tiles_hit, void* scratch, hipStream_t s);
hipError_t project_bwd(const Camera& c, const float* means3d, const float* scales,
                       const float* quats, const float* viewmat, const int* radii,
                       const float* conics, const float* compensation,
                       const float* v_xy, const float* v_depth, const float* v_conic,
                       const float* v_compensation, float* v_mean3d, float* v_scale,
                       float* v_quat, float* v_cov2d, float* v_cov3d, int consistent,
                       hipStream_t s);

// Backward tail of the fused render (projection, binning into the
// opacity-limited rectangles, rasterization): per Gaussian, the sum of its
// staged rasterize records (ws of raster_bwd_staged with records_only, over
// num_isect slots of the rectangles rects / inclusive counts cum) gives
// v_colors (nch channels), v_opacity, and, in registers, v_xy and v_conic,
// whose projection VJP gives v_mean3d, v_scale and v_quat; every row written.
hipError_t render_bwd_tail(const Camera& c, int nch, int num_isect, const float* means3d,
                           const float* scales, const float* quats, const float* viewmat,
                           const int* radii, const uint2* rects, const int* cum,
                           const float* conics, const void* ws, float* v_colors,
                           float* v_opacity, float* v_mean3d, float* v_scale, float* v_quat,
                           int consistent, hipStream_t s);

// ---- LSD radix sort (uint32 keys, uint32 values) ----
struct SortPlan {
    int n_cap;          // capacity (upper bound on n) used for launch shapes
    int begin_bit, end_bit;
    int passes;
    int parts;          // partitions per pass at capacity
    int tile;           // keys per partition
    int first_shift, first_bits;   // digit of the first pass
    size_t scratch_bytes;
};
SortPlan sort_plan(int n_cap, int begin_bit, int end_bit);
// The first pass's per-partition digit counts, digit-major (counts[d * parts +
// part], 256 digits), inside the plan's scratch: a producer of the keys may
// fill them itself and pass counted = true to sort_pairs, which then skips
// that pass's counting kernel.
uint32_t* sort_first_counts(const SortPlan& p, void* scratch);
// Sorts (keys_in, vals_in) -> (keys_out, vals_out) stably on [begin, end)
// (with unstable_first, equal keys may leave in any order).
// vals_in == nullptr means values = index.  n is read from n_dev when non-null.
// tmp_keys/tmp_vals must hold n_cap entries (ping-pong buffers).  The scratch
// area must be zeroed by the caller (sort_zero_bytes of it) before the call.
size_t sort_zero_bytes(const SortPlan& p);
hipError_t sort_pairs(const SortPlan& p, const uint32_t* keys_in, const uint32_t* vals_in,
                      uint32_t* keys_out, uint32_t* vals_out, uint32_t* tmp_keys,
                      uint32_t* tmp_vals, int n_host, const int* n_dev, void* scratch,
                      uint32_t* err, hipStream_t s, bool counted = false,
                      bool unstable_first = false);

// ---- binning ----
struct BinPlan {
    int n;              // Gaussians
    int isect_cap;      // intersection capacity
    int num_tiles;
    int tile_bits;
    SortPlan tile_sort;
    size_t depth_bytes;     // depth stage workspace (depends on n only)
    size_t emit_bytes;      // emission stage workspace
    size_t scratch_bytes;   // both, for bin_and_sort: emission region first
};
BinPlan bin_plan(const Camera& c, int isect_cap);
// Depth stage: per-Gaussian depth keys (depth bits; 0xFFFFFFFF for Gaussians
// that emit nothing) and packed tile rectangles.  rects (optional, packed as
// pack_rect) replaces the rectangles computed from xys and radii.
hipError_t bin_depth(const BinPlan& p, const Camera& c, const float* xys, const float* depths,
                     const int* radii, const uint2* rects, void* depth_ws, hipStream_t s);
// Emission stage over the depth stage's workspace: emission in id order, the
// stable tile sort, edges, and the per-tile depth order.
// p.isect_cap bounds the slots written; the sort and the edges cover the slots
// actually written, min(total, p.isect_cap), read on the device, so that a
// num_isect_host larger than the total never sorts unwritten slots.
hipError_t bin_emit(const BinPlan& p, const Camera& c, int num_isect_host,
                    const void* depth_ws, int* gids_sorted, int* tiles_sorted, int* tile_bins,
                    void* scratch, hipStream_t s);
// Binning of given packed tile rectangles with their inclusive counts cum
// (consistent: cum[g] - cum[g-1] = the rectangle's tile count), the depth
// keys read from depths (float bits); the emission workspace (p.emit_bytes).
hipError_t bin_rects(const BinPlan& p, const Camera& c, const float* depths,
                     const uint2* rects, const int* cum, int num_isect_host, int* gids_sorted,
                     int* tiles_sorted, int* tile_bins, void* scratch, hipStream_t s);
hipError_t bin_and_sort(const BinPlan& p, const Camera& c, const float* xys,
                        const float* depths, const int* radii, int num_isect_host,
                        int* gids_sorted, int* tiles_sorted, int* tile_bins, void* scratch,
                        hipStream_t s);
// status word of the last bin_and_sort using `scratch`: bit 4 bad tile key,
// bit 8 more intersections than the capacity
const uint32_t* bin_error_word(const BinPlan& p, const void* scratch);

// ---- batched projection (views of scenes, see ViewRec) ----
// Forward: items of every view; writes xys, conics, radii, the depth sort
// key (depth bits, 0xFFFFFFFF for items that emit nothing) and the packed tile
// rectangle (pack_rect; 0 for items that emit nothing): with opacities
// (optional) the opacity-limited one (opacity_tile_rect), as the single-view
// rasterize path bins.
hipError_t project_batch_fwd(const ViewRec* views, int num_views, int max_view_n,
                             float glob_scale, float clip_thresh, int block_width,
                             const float* means3d, const float* scales, const float* quats,
                             const float* viewmats, const float* opacities, float* xys,
                             float* conics, int* radii, uint32_t* depth_keys, uint2* item_rects,
                             hipStream_t s);
// Backward: one lane per Gaussian, summing the contributions of its scene's
// views; writes v_mean3d, v_scale, v_quat for every Gaussian of every scene,
// and with v_rgbo (per item: rgb, opacity; optional) also v_colors and
// v_opacity as the sums of the Gaussian's items over the views.
hipError_t project_batch_bwd(const SceneRec* scenes, int num_scenes, int max_scene_n,
                             const ViewRec* views, float glob_scale, float clip_thresh,
                             const float* means3d, const float* scales, const float* quats,
                             const float* viewmats, const int* radii, const float* conics,
                             const float* v_xy, const float* v_conic, const float* v_rgbo,
                             float* v_mean3d, float* v_scale, float* v_quat, float* v_colors,
                             float* v_opacity, int consistent, hipStream_t s);

// ---- batched binning ----
struct BatchBinPlan {
    int n_items;
    int num_views;
    int isect_cap;
    int tile_bits, view_bits;
    int num_bins;       // num_views << tile_bits
    SortPlan tile_sort;
    size_t scratch_bytes;
};
BatchBinPlan batch_bin_plan(int n_items, int num_views, int max_view_tiles, int isect_cap);
// Sorted item ids of every (view, tile) and bins[view << tile_bits | tile].
// The scratch keeps the intersection count clamped to the capacity (the
// length the sorts and bins use), the exact total (64 bits) and an error word:
// bit 4 bad tile key, bit 8 capacity overflow.  item_first (optional, one per
// item) receives the first intersection slot of every item with a non-empty
// tile rectangle; its slots are the rectangle's tiles in row-major order.
hipError_t batch_bin(const BatchBinPlan& p, const ViewRec* views, const uint2* item_rects,
                     const uint32_t* depth_keys, int* items_sorted, int* tile_bins,
                     uint32_t* item_first, void* scratch, hipStream_t s);
const uint32_t* batch_bin_error_word(const BatchBinPlan& p, const void* scratch);
const int* batch_bin_count(const BatchBinPlan& p, const void* scratch);
const unsigned long long* batch_bin_total(const BatchBinPlan& p, const void* scratch);

// ---- rasterization ----
hipError_t raster_fwd(const Camera& c, const int* gids_sorted, const int* tile_bins,
                      const float* xys, const float* conics, const float* colors,
                      const float* opacities, const float* background, float* final_Ts,
                      int* final_idx, float* out_img, hipStream_t s);
hipError_t raster_bwd(const Camera& c, const int* gids_sorted, const int* tile_bins,
                      const float* xys, const float* conics, const float* colors,
                      const float* opacities, const float* background,
                      const float* final_Ts, const int* final_idx, const float* v_out,
                      const float* v_out_alpha, float* v_colors, float* v_opacity,
                      float* v_xy, float* v_xy_abs, float* v_conic, int consistent,
                      hipStream_t s);
// The same kernels over a RasterGrid (one view or a batch).  Item-indexed
// arrays: xys, conics, v_xy, v_xy_abs, v_conic; Gaussian-indexed: colors,
// opacities, v_colors, v_opacity; backgrounds are rows of 3 per view.
hipError_t raster_fwd_grid(const RasterGrid& g, const int* items_sorted, const int* tile_bins,
                           const float* xys, const float* conics, const float* colors,
                           const float* opacities, const float* backgrounds, float* final_Ts,
                           int* final_idx, float* out_img, hipStream_t s);
// Single-view backward with staged gradients instead of float atomics (see
// StagedOut in raster.hip): nch <= 3 channels, workspace raster_bwd_staged_bytes;
// radii and cum_tiles_hit are the projection's.  Writes every gradient row.
size_t raster_bwd_staged_bytes(int num_isect);
// rects (optional, packed as pack_rect) replaces the rectangles computed from
// xys and radii (and must be the ones the binning used).  records_only: stop
// after the records and their bitmap (in ws); render_bwd_tail sums them.
hipError_t raster_bwd_staged(const Camera& c, int nch, const int* gids_sorted,
                             const int* tile_bins, int num_isect, const float* xys,
                             const float* conics, const float* colors, const float* opacities,
                             const float* background, const float* final_Ts,
                             const int* final_idx, const float* v_out, const float* v_out_alpha,
                             const int* radii, const uint2* rects, const int* cum_tiles_hit,
                             void* ws,
                             float* v_colors, float* v_opacity, float* v_xy, float* v_xy_abs,
                             float* v_conic, int consistent, hipStream_t s,
                             bool records_only = false, const float4* splats = nullptr);
// Single-view forward staging from splat records (RasterGrid::splats) of
// nch <= 3 channels: no xys / conics / colours / opacities arrays.
hipError_t raster_fwd_splats(const Camera& c, int nch, const int* gids_sorted,
                             const int* tile_bins, const float4* splats,
                             const float* background, float* final_Ts, int* final_idx,
                             float* out_img, hipStream_t s);
// Batch backward with staged gradients (3 channels): records at the slots
// item_first[item] + the tile's row-major offset in the item's rectangle
// (item_rects, the ones the binning used; workspace raster_bwd_staged_bytes(cap)),
// summed per item into v_xy, v_conic and v_rgbo (4 floats per item: rgb,
// opacity); every item row is written.  g.views is the device view table;
// max_view_n the largest view's item count.
hipError_t raster_bwd_batch_staged(const RasterGrid& g, int max_view_n, const int* items_sorted,
                                   const int* tile_bins, int cap, const float* xys,
                                   const float* conics, const float* colors,
                                   const float* opacities, const float* backgrounds,
                                   const float* final_Ts, const int* final_idx,
                                   const float* v_out, const int* radii, const uint2* item_rects,
                                   const uint32_t* item_first, void* ws, float* v_xy,
                                   float* v_conic, float* v_rgbo, int consistent,
                                   hipStream_t s);
hipError_t raster_bwd_grid(const RasterGrid& g, const int* items_sorted, const int* tile_bins,
                           const float* xys, const float* conics, const float* colors,
                           const float* opacities, const float* backgrounds,
                           const float* final_Ts, const int* final_idx, const float* v_out,
                           const float* v_out_alpha, float* v_colors, float* v_opacity,
                           float* v_xy, float* v_xy_abs, float* v_conic, int consistent,
                           hipStream_t s);

// Work-queue order of a batch's tiles, longest tile lists first: one stable
// radix pass over the tiles keyed by the bit length of their list length
// (the tile numbering order within a key).  keys/keys_out/tmp_* hold
// total_tiles entries; sort_scratch order_sort_plan(total_tiles).scratch_bytes.
SortPlan order_sort_plan(int total_tiles);
hipError_t tile_order(const RasterGrid& g, const int* tile_bins, uint32_t* keys,
                      uint32_t* keys_out, uint32_t* tmp_k, uint32_t* tmp_v, int* order,
                      void* sort_scratch, hipStream_t s);

}  // namespace jsa
