// Tile binning for gfx950: intersection emission, the (tile, depth) sort and
// tile bin edges.
//
// Replaces the reference's ops::sort_and_bin (lib/ops.cu:645-793):
// map_gaussian_to_intersects (lib/kernels/forward.cu:135-179) emits
// key = (tile << 32) | depth_bits per (Gaussian, tile), CUB sorts the 64-bit
// keys stably on [0, 32 + msb) and get_tile_bin_edges (forward.cu:184-208)
// marks [start, end) per tile.  Because emission is in ascending Gaussian id,
// the reference's final order is the lexicographic order
// (tile, depth bits, gaussian id).
//
// This implementation produces exactly that order with far less sort traffic:
//  1. depth_keys: one 32-bit key per Gaussian (depth bits, or 0xFFFFFFFF for
//     Gaussians with radius 0, which emit nothing);
//  2. a stable LSD radix sort of the N Gaussians by depth key (values = index),
//     giving the (depth bits, gaussian id) order;
//  3. emit: a single-pass kernel walks the Gaussians in that order, scans
//     their tile counts with a decoupled look-back and writes
//     (tile id, gaussian id) pairs; each wave spreads its slots evenly over
//     its lanes (owner lookup by an LDS mark and a DPP max-scan), so a
//     Gaussian covering hundreds of tiles does not serialise one lane;
//  4. a stable LSD radix sort of the pairs on the tile id bits only
//     (2 passes at 1080p/16x16 instead of 6 passes over 64-bit keys);
//  5. tile_bins_kernel: [start, end) edges from the sorted tile ids, empty
//     tiles (0, 0).
// Stability of step 4 on input already ordered by (depth bits, id) yields
// (tile, depth bits, id): identical to the reference order.
#include "common.hpp"

namespace jsa {
namespace {

__global__ void depth_keys_kernel(int n, const float* __restrict__ depths,
                                  const int* __restrict__ radii, uint32_t* __restrict__ keys) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    keys[g] = radii[g] > 0 ? __float_as_uint(depths[g]) : 0xFFFFFFFFu;
}

// forward.cu:152-161 / helpers.h:9-32, recomputed from xys and the integer radius.
__device__ __forceinline__ void tile_rect(int block_width, int tiles_x, int tiles_y, float2 xy,
                                          int rad, int& x0, int& x1, int& y0, int& y1) {
    const float bwf = (float)block_width;
    const float tcx = xy.x / bwf, tcy = xy.y / bwf, tr = (float)rad / bwf;
    x0 = min(max(0, (int)(tcx - tr)), tiles_x);
    x1 = min(max(0, (int)(tcx + tr + 1.f)), tiles_x);
    y0 = min(max(0, (int)(tcy - tr)), tiles_y);
    y1 = min(max(0, (int)(tcy + tr + 1.f)), tiles_y);
}

// View of each depth-sorted item (largest v with item_off[v] <= item), the
// key of the stable pass that groups a batch's items by view.
constexpr int kMaxBatchViews = 4096;
__global__ __launch_bounds__(256) void view_keys_kernel(const uint32_t* __restrict__ perm, int n,
                                                        const ViewRec* __restrict__ views,
                                                        int num_views,
                                                        uint32_t* __restrict__ vkeys) {
    __shared__ int s_off[kMaxBatchViews];
    for (int v = threadIdx.x; v < num_views; v += blockDim.x) s_off[v] = views[v].item_off;
    __syncthreads();
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int it = (int)perm[r];
    int lo = 0, hi = num_views;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_off[mid] <= it) lo = mid;
        else hi = mid;
    }
    vkeys[r] = (uint32_t)lo;
}

constexpr int kEmitPer = 4;                  // Gaussians per lane
constexpr int kEmitPart = 256 * kEmitPer;    // Gaussians per scan partition

// Emission: the workgroup owns 1024 consecutive Gaussians of the depth order
// (4 groups of 64 per wave).  A single-pass decoupled look-back scan gives the
// workgroup's first output slot; each wave then expands its groups 64 slots
// at a time: the lane whose tile run starts inside the chunk marks that
// position in a per-wave LDS row, an inclusive max-scan (DPP) gives every
// slot its owning lane, and the owner's (id, rectangle, offset) is read from
// LDS.  Stores of one chunk are 64 consecutive slots.
// BATCH: the sorted entries are items of a batch's views (vkeys gives each
// entry's view); the key written is the view's bin_off + tile, and the last
// partition publishes the total (clamped count for the sorts, exact total).
template <bool BATCH>
__global__ __launch_bounds__(256) void emit_kernel(
    Camera c, const ViewRec* __restrict__ views, const uint32_t* __restrict__ vkeys,
    const uint32_t* __restrict__ perm, const float2* __restrict__ xys,
    const int* __restrict__ radii, int isect_cap, uint32_t* __restrict__ tile_out,
    uint32_t* __restrict__ gid_out, uint64_t* __restrict__ status, uint32_t* __restrict__ counter,
    uint32_t* __restrict__ err, int* __restrict__ count_out,
    unsigned long long* __restrict__ total_out) {
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave_tot[kEmitPer][4];
    __shared__ uint64_t s_excl;
    __shared__ int s_mark[4][64];
    __shared__ int4 s_own[4][64];        // (gid, x0 | y0 << 16, width, group offset)
    __shared__ int2 s_key[4][64];        // (key base, tiles_x) of the owner's view
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    if (tid == 0) s_part = atomicAdd(counter, 1u);
    __syncthreads();
    const int part = (int)s_part;

    // element e = k*256 + tid of the partition (depth rank part*kEmitPart + e)
    int cnt[kEmitPer], incl[kEmitPer], rx0[kEmitPer], ry0[kEmitPer], rw[kEmitPer];
    int kbase[kEmitPer], ktx[kEmitPer];
    uint32_t gg[kEmitPer];
#pragma unroll
    for (int k = 0; k < kEmitPer; ++k) {
        const int r = part * kEmitPart + k * 256 + tid;
        int count = 0, x0 = 0, y0 = 0, w = 1, base = 0, tiles_x = c.tiles_x;
        uint32_t g = 0;
        if (r < c.num_points) {
            g = perm[r];
            const int rad = radii[g];
            if (rad > 0) {
                int tiles_y = c.tiles_y;
                if (BATCH) {
                    const ViewRec& vr = views[vkeys[r]];
                    tiles_x = vr.tiles_x;
                    tiles_y = vr.tiles_y;
                    base = vr.bin_off;
                }
                int x1, y1;
                tile_rect(c.block_width, tiles_x, tiles_y, xys[g], rad, x0, x1, y0, y1);
                w = max(x1 - x0, 1);
                count = (x1 - x0) * (y1 - y0);
            }
        }
        cnt[k] = count;
        rx0[k] = x0;
        ry0[k] = y0;
        rw[k] = w;
        gg[k] = g;
        kbase[k] = base;
        ktx[k] = tiles_x;
        incl[k] = wave_inclusive_sum(count);
        if (lane == 63) s_wave_tot[k][wave] = (uint32_t)incl[k];
    }
    __syncthreads();
    uint32_t total = 0, grp_off[kEmitPer];
#pragma unroll
    for (int k = 0; k < kEmitPer; ++k) {
        uint32_t w_off = 0, slab = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t v = s_wave_tot[k][w];
            w_off += (w < wave) ? v : 0u;
            slab += v;
        }
        grp_off[k] = total + w_off;
        total += slab;
    }
    if (tid == 0) {
        store_agent(status + part, (part == 0 ? kFlagPre : kFlagAgg) | (uint64_t)total);
        if (part == 0) s_excl = 0;
    }
    if (wave == 0 && part > 0) {
        const uint64_t excl = lookback64(status, part, err);
        if (lane == 0) {
            store_agent(status + part, kFlagPre | (excl + total));
            s_excl = excl;
        }
    }
    __syncthreads();
    const uint64_t gbase = s_excl;
    if (BATCH && tid == 0 && part == (int)gridDim.x - 1) {
        const uint64_t all = gbase + total;
        *total_out = (unsigned long long)all;
        *count_out = all > (uint64_t)isect_cap ? isect_cap : (int)all;
        if (all > (uint64_t)isect_cap) atomicOr(err, 8u);
    }

#pragma unroll
    for (int k = 0; k < kEmitPer; ++k) {
        const int gtotal = uniform_i(__shfl(incl[k], 63));
        const int ex = incl[k] - cnt[k];
        s_own[wave][lane] = make_int4((int)gg[k], rx0[k] | (ry0[k] << 16), rw[k], ex);
        if (BATCH) s_key[wave][lane] = make_int2(kbase[k], ktx[k]);
        wave_lds_sync();
        const uint64_t obase = gbase + uniform_i((int)grp_off[k]);
        int carry = 0;
        for (int chunk = 0; chunk < gtotal; chunk += 64) {
            s_mark[wave][lane] = 0;
            wave_lds_sync();
            if (cnt[k] > 0 && ex >= chunk && ex < chunk + 64) s_mark[wave][ex - chunk] = lane + 1;
            wave_lds_sync();
            int v = wave_max_scan_nonneg(s_mark[wave][lane]);
            const int owner = v > 0 ? v - 1 : carry;
            carry = __builtin_amdgcn_readlane(owner, 63);
            const int slot = chunk + lane;
            if (slot < gtotal) {
                const int4 od = s_own[wave][owner];
                const int2 kb = BATCH ? s_key[wave][owner] : make_int2(0, c.tiles_x);
                const int kk = slot - od.w;
                const int q = kk / od.z;
                const int ty = (od.y >> 16) + q;
                const int tx = (od.y & 0xFFFF) + (kk - q * od.z);
                const uint64_t pos = obase + (uint64_t)slot;
                if (pos < (uint64_t)isect_cap) {
                    tile_out[pos] = (uint32_t)(kb.x + ty * kb.y + tx);
                    gid_out[pos] = (uint32_t)od.x;
                }
            }
            wave_lds_sync();
        }
    }
}

__global__ void tile_bins_kernel(const uint32_t* __restrict__ tiles, int n_host,
                                 const int* __restrict__ n_dev, int num_tiles,
                                 int2* __restrict__ bins, uint32_t* __restrict__ err) {
    const int n = n_dev ? *n_dev : n_host;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = tiles[i];
    const uint32_t p_ = i > 0 ? tiles[i - 1] : 0u;
    if (t >= (uint32_t)num_tiles || p_ >= (uint32_t)num_tiles) {   // never for valid input
        atomicOr(err, 4u);
        return;
    }
    if (i == 0) bins[t].x = 0;
    if (i == n - 1) bins[t].y = n;
    if (i > 0) {
        const uint32_t p = tiles[i - 1];
        if (p != t) {
            bins[p].y = i;
            bins[t].x = i;
        }
    }
}

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct BinLayout {
    // zeroed region
    size_t err, emit_counter, emit_status, zero_end;
    // plain region
    size_t depth_sort, tile_sort, dkeys, dtmp_k, dtmp_v, dsorted_k, perm, tiles, gids, ttmp_k,
        ttmp_v, tiles_sorted, end;
};

BinLayout layout(const BinPlan& p) {
    BinLayout L;
    size_t o = 0;
    L.err = o; o += 16;
    L.emit_counter = o; o += 16;
    L.emit_status = o; o = align256(o + 8ull * ((p.n + kEmitPart - 1) / kEmitPart + 1));
    L.zero_end = o;
    L.depth_sort = o; o = align256(o + p.depth_sort.scratch_bytes);
    L.tile_sort = o; o = align256(o + p.tile_sort.scratch_bytes);
    const size_t n4 = align256(4ull * (p.n > 0 ? p.n : 1));
    const size_t i4 = align256(4ull * (p.isect_cap > 0 ? p.isect_cap : 1));
    L.dkeys = o; o += n4;
    L.dtmp_k = o; o += n4;
    L.dtmp_v = o; o += n4;
    L.dsorted_k = o; o += n4;
    L.perm = o; o += n4;
    L.tiles = o; o += i4;
    L.gids = o; o += i4;
    L.ttmp_k = o; o += i4;
    L.ttmp_v = o; o += i4;
    L.tiles_sorted = o; o += i4;
    L.end = o;
    return L;
}

int bitlen(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

}  // namespace

BinPlan bin_plan(const Camera& c, int isect_cap) {
    BinPlan p;
    p.n = c.num_points;
    p.isect_cap = isect_cap;
    p.num_tiles = c.tiles_x * c.tiles_y;
    p.tile_bits = bitlen((uint32_t)(p.num_tiles > 1 ? p.num_tiles - 1 : 1));
    p.depth_sort = sort_plan(p.n, 0, 32);
    p.tile_sort = sort_plan(isect_cap, 0, p.tile_bits);
    p.scratch_bytes = layout(p).end;
    return p;
}

const uint32_t* bin_error_word(const BinPlan& p, const void* scratch) {
    return reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(scratch) + layout(p).err);
}

hipError_t bin_and_sort(const BinPlan& p, const Camera& c, const float* xys, const float* depths,
                        const int* radii, int num_isect_host, const int* num_isect_dev,
                        int* gids_sorted, int* tiles_sorted, int* tile_bins, void* scratch,
                        hipStream_t s) {
    const BinLayout L = layout(p);
    uint8_t* b = static_cast<uint8_t*>(scratch);
    hipError_t e = fill_zero(b, L.zero_end, s);
    if (e != hipSuccess) return e;
    e = fill_zero(tile_bins, 8ull * p.num_tiles, s);
    if (e != hipSuccess) return e;
    if (p.n <= 0) return hipSuccess;
    uint32_t* err = reinterpret_cast<uint32_t*>(b + L.err);
    auto U = [&](size_t off) { return reinterpret_cast<uint32_t*>(b + off); };

    depth_keys_kernel<<<(p.n + 255) / 256, 256, 0, s>>>(p.n, depths, radii, U(L.dkeys));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = sort_pairs(p.depth_sort, U(L.dkeys), nullptr, U(L.dsorted_k), U(L.perm), U(L.dtmp_k),
                   U(L.dtmp_v), p.n, nullptr, b + L.depth_sort, err, s);
    if (e != hipSuccess) return e;
    emit_kernel<false><<<(p.n + kEmitPart - 1) / kEmitPart, 256, 0, s>>>(
        c, nullptr, nullptr, U(L.perm), reinterpret_cast<const float2*>(xys), radii, p.isect_cap,
        U(L.tiles), U(L.gids), reinterpret_cast<uint64_t*>(b + L.emit_status), U(L.emit_counter),
        err, nullptr, nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t* tiles_out = tiles_sorted ? reinterpret_cast<uint32_t*>(tiles_sorted) : U(L.tiles_sorted);
    e = sort_pairs(p.tile_sort, U(L.tiles), U(L.gids), tiles_out,
                   reinterpret_cast<uint32_t*>(gids_sorted), U(L.ttmp_k), U(L.ttmp_v),
                   num_isect_host, num_isect_dev, b + L.tile_sort, err, s);
    if (e != hipSuccess) return e;
    const int nb = (p.isect_cap + 255) / 256;
    tile_bins_kernel<<<nb > 0 ? nb : 1, 256, 0, s>>>(tiles_out, num_isect_host, num_isect_dev,
                                                   p.num_tiles,
                                                   reinterpret_cast<int2*>(tile_bins), err);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched binning (items of many views; see ViewRec)
// ---------------------------------------------------------------------------
//  1. stable sort of the items by depth key (4 passes): (depth bits, item);
//  2. view_keys + one stable pass on the view bits: (view, depth bits, item),
//     the emission order;
//  3. emit_kernel<true>: keys view << tile_bits | tile, values = item ids,
//     with the count published by the last partition;
//  4. stable sort on the tile bits only (the view bits ride along): order
//     (tile, view, depth bits, item), so every (view, tile) run is contiguous
//     and in the reference's per-tile order;
//  5. tile_bins_kernel on the full keys.
namespace {

struct BatchLayout {
    size_t err, count, total, emit_counter, emit_status, zero_end;
    size_t depth_sort, view_sort, tile_sort, dsorted_k, perm, dtmp_k, dtmp_v, vkeys, vsorted,
        perm2, keys, vals, ttmp_k, ttmp_v, keys_sorted, end;
};

BatchLayout batch_layout(const BatchBinPlan& p) {
    BatchLayout L;
    size_t o = 0;
    L.err = o; o += 16;
    L.count = o; o += 16;
    L.total = o; o += 16;
    L.emit_counter = o; o += 16;
    L.emit_status = o; o = align256(o + 8ull * ((p.n_items + kEmitPart - 1) / kEmitPart + 1));
    L.zero_end = o;
    L.depth_sort = o; o = align256(o + p.depth_sort.scratch_bytes);
    L.view_sort = o; o = align256(o + p.view_sort.scratch_bytes);
    L.tile_sort = o; o = align256(o + p.tile_sort.scratch_bytes);
    const size_t n4 = align256(4ull * (p.n_items > 0 ? p.n_items : 1));
    const size_t i4 = align256(4ull * (p.isect_cap > 0 ? p.isect_cap : 1));
    L.dsorted_k = o; o += n4;
    L.perm = o; o += n4;
    L.dtmp_k = o; o += n4;
    L.dtmp_v = o; o += n4;
    L.vkeys = o; o += n4;
    L.vsorted = o; o += n4;
    L.perm2 = o; o += n4;
    L.keys = o; o += i4;
    L.vals = o; o += i4;
    L.ttmp_k = o; o += i4;
    L.ttmp_v = o; o += i4;
    L.keys_sorted = o; o += i4;
    L.end = o;
    return L;
}

}  // namespace

BatchBinPlan batch_bin_plan(int n_items, int num_views, int max_view_tiles, int isect_cap) {
    BatchBinPlan p;
    p.n_items = n_items;
    p.num_views = num_views;
    p.isect_cap = isect_cap;
    p.tile_bits = bitlen((uint32_t)(max_view_tiles > 1 ? max_view_tiles - 1 : 1));
    p.view_bits = bitlen((uint32_t)(num_views > 1 ? num_views - 1 : 0));
    p.num_bins = num_views << p.tile_bits;
    p.depth_sort = sort_plan(n_items, 0, 32);
    p.view_sort = sort_plan(n_items, 0, p.view_bits);
    p.tile_sort = sort_plan(isect_cap, 0, p.tile_bits);
    p.scratch_bytes = batch_layout(p).end;
    return p;
}

const uint32_t* batch_bin_error_word(const BatchBinPlan& p, const void* scratch) {
    return reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(scratch) +
                                             batch_layout(p).err);
}

const int* batch_bin_count(const BatchBinPlan& p, const void* scratch) {
    return reinterpret_cast<const int*>(static_cast<const uint8_t*>(scratch) +
                                        batch_layout(p).count);
}

const unsigned long long* batch_bin_total(const BatchBinPlan& p, const void* scratch) {
    return reinterpret_cast<const unsigned long long*>(static_cast<const uint8_t*>(scratch) +
                                                       batch_layout(p).total);
}

hipError_t batch_bin(const BatchBinPlan& p, const ViewRec* views, int block_width,
                     const float* xys, const int* radii, const uint32_t* depth_keys,
                     int* items_sorted, int* tile_bins, void* scratch, hipStream_t s) {
    const BatchLayout L = batch_layout(p);
    uint8_t* b = static_cast<uint8_t*>(scratch);
    hipError_t e = fill_zero(b, L.zero_end, s);
    if (e != hipSuccess) return e;
    e = fill_zero(tile_bins, 8ull * p.num_bins, s);
    if (e != hipSuccess) return e;
    if (p.n_items <= 0) return hipSuccess;
    if (p.num_views > kMaxBatchViews) return hipErrorInvalidValue;
    uint32_t* err = reinterpret_cast<uint32_t*>(b + L.err);
    int* count = reinterpret_cast<int*>(b + L.count);
    auto U = [&](size_t off) { return reinterpret_cast<uint32_t*>(b + off); };

    e = sort_pairs(p.depth_sort, depth_keys, nullptr, U(L.dsorted_k), U(L.perm), U(L.dtmp_k),
                   U(L.dtmp_v), p.n_items, nullptr, b + L.depth_sort, err, s);
    if (e != hipSuccess) return e;
    view_keys_kernel<<<(p.n_items + 255) / 256, 256, 0, s>>>(U(L.perm), p.n_items, views,
                                                              p.num_views, U(L.vkeys));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t* order = U(L.perm);
    const uint32_t* order_views = U(L.vkeys);
    if (p.view_bits > 0) {
        e = sort_pairs(p.view_sort, U(L.vkeys), U(L.perm), U(L.vsorted), U(L.perm2),
                       U(L.dtmp_k), U(L.dtmp_v), p.n_items, nullptr, b + L.view_sort, err, s);
        if (e != hipSuccess) return e;
        order = U(L.perm2);
        order_views = U(L.vsorted);
    }
    Camera c{};
    c.num_points = p.n_items;
    c.block_width = block_width;
    emit_kernel<true><<<(p.n_items + kEmitPart - 1) / kEmitPart, 256, 0, s>>>(
        c, views, order_views, order, reinterpret_cast<const float2*>(xys), radii, p.isect_cap,
        U(L.keys), U(L.vals), reinterpret_cast<uint64_t*>(b + L.emit_status), U(L.emit_counter),
        err, count, reinterpret_cast<unsigned long long*>(b + L.total));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = sort_pairs(p.tile_sort, U(L.keys), U(L.vals), U(L.keys_sorted),
                   reinterpret_cast<uint32_t*>(items_sorted), U(L.ttmp_k), U(L.ttmp_v), 0, count,
                   b + L.tile_sort, err, s);
    if (e != hipSuccess) return e;
    const int nb = (p.isect_cap + 255) / 256;
    tile_bins_kernel<<<nb > 0 ? nb : 1, 256, 0, s>>>(U(L.keys_sorted), 0, count, p.num_bins,
                                                   reinterpret_cast<int2*>(tile_bins), err);
    return hipGetLastError();
}

}  // namespace jsa
