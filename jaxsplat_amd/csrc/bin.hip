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
// Single view, with far less sort traffic than a 64-bit key sort:
//  1. depth_keys_kernel: per Gaussian its depth bits and packed tile
//     rectangle;
//  2. emission in Gaussian id order: a reduce-then-scan of the Gaussians'
//     tile counts gives each Gaussian its first output slot (the reference's
//     cum_tiles_hit order), then every wave writes whole 64-slot chunks of
//     (tile id, gaussian id) pairs (owner lookup by an LDS mark and a DPP
//     max-scan), so neither a Gaussian covering hundreds of tiles nor a run
//     of large Gaussians serialises a lane or a wave; the expansion also
//     counts the tile sort's first digit;
//  3. an LSD radix sort of the pairs on the tile id bits only (2 passes at
//     1080p/16x16; the first need not be stable, the order inside a tile
//     being set in step 5);
//  4. tile_bins_kernel: [start, end) edges from the sorted tile ids, empty
//     tiles (0, 0);
//  5. tile_depth_sort_kernel: every tile's run ordered by (depth bits, id)
//     with LDS buckets linear in the run's key range (global scratch for
//     runs longer than the LDS holds): identical to the reference order.
// The depth order is thus established per tile (about 800 entries each at
// 1080p with 1M Gaussians, sorted where they lie in LDS) instead of by a
// global sort of the Gaussians and a stable tile sort after it.
#include "common.hpp"

namespace jsa {
namespace {

// Per Gaussian (id order): the depth sort key and the packed tile rectangle.
__global__ void depth_keys_kernel(Camera c, const float* __restrict__ depths,
                                  const int* __restrict__ radii, const float2* __restrict__ xys,
                                  const uint2* __restrict__ rects_in, uint32_t* __restrict__ keys,
                                  uint2* __restrict__ rects) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= c.num_points) return;
    const int rad = radii[g];
    uint2 rc = make_uint2(0u, 0u);
    if (rad > 0) {
        if (rects_in) {
            rc = rects_in[g];
        } else {
            int x0, x1, y0, y1;
            tile_rect(c.block_width, c.tiles_x, c.tiles_y, xys[g], rad, x0, x1, y0, y1);
            rc = pack_rect(x0, x1, y0, y1);
        }
    }
    // Gaussians that emit nothing get the key the depth sort drops
    keys[g] = rect_count(rc) > 0 ? __float_as_uint(depths[g]) : 0xFFFFFFFFu;
    rects[g] = rc;
}

// View of each item of the emission order (largest v with item_off[v] <=
// item); perm == nullptr: the items 0..n-1 in order, n = n_host.
constexpr int kMaxBatchViews = 4096;
__global__ __launch_bounds__(256) void view_keys_kernel(const uint32_t* __restrict__ perm,
                                                        int n_host,
                                                        const int* __restrict__ n_dev,
                                                        const ViewRec* __restrict__ views,
                                                        int num_views,
                                                        uint32_t* __restrict__ vkeys) {
    __shared__ int s_off[kMaxBatchViews];
    for (int v = threadIdx.x; v < num_views; v += blockDim.x) s_off[v] = views[v].item_off;
    __syncthreads();
    const int n = n_dev ? *n_dev : n_host;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int it = perm ? (int)perm[r] : r;
    int lo = 0, hi = num_views;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_off[mid] <= it) lo = mid;
        else hi = mid;
    }
    vkeys[r] = (uint32_t)lo;
}

// ---------------------------------------------------------------------------
// Emission: reduce-then-scan over the depth order, then slot-balanced expansion
// ---------------------------------------------------------------------------
// Tile counts per Gaussian range from 1 to the whole image, and an order such
// as depth order puts the large, near Gaussians together, so splitting the
// work by Gaussians leaves a few waves with several times the average number
// of slots (at 1080p with 1M Gaussians in depth order the busiest wave
// expanded 231 chunks of 64 slots against a mean of 35).  Here every wave of
// the expansion owns whole 64-slot chunks of the output instead:
//  1. emit_count_kernel: each depth-ordered entry's tile rectangle, packed
//     (x0 | y0 << 16, w | h << 16), and per-partition totals of slots and of
//     entries that emit (w * h > 0);
//  2. emit_scan_kernel (one workgroup): partition offsets of both, the total,
//     the number of slots written, min(total, capacity), and of emitting
//     entries;
//  3. emit_compact_kernel: the emitting entries, compacted in order, with
//     their first slot (clamped to the slots written), rectangle, id and (for
//     a batch) key base and tiles_x; and for every 64-slot chunk the compacted
//     entry that owns its first slot.  Compaction matters for a batch, whose
//     order (view, depth) leaves a run of culled items at the end of every
//     view;
//  4. emit_expand_kernel: every wave writes an equal, contiguous range of
//     chunks, sliding a window of 64 compacted entries from the owner of its
//     first slot; an LDS mark per entry start and a DPP max-scan give every
//     slot of a chunk its entry, and the 64 slots are stored contiguously.
// BATCH: the entries are items of a batch's views (vkeys gives each entry's
// view) and the key written is the view's bin_off + tile.
constexpr int kEmitPer = 4;                  // entries per lane in passes 1 and 3
constexpr int kEmitPart = 256 * kEmitPer;    // entries per scan partition
constexpr int kExpandWaves = 8192;           // waves of the expansion (a full chip)

// Pass 1 over the entries perm[0..n): gathers each entry's packed rectangle
// (one 8-byte read at a random id) into emission order and sums slots and
// emitting entries per partition.  The lane's loads are issued stage by stage
// so that they overlap.  perm == nullptr is the id order: the rectangles are
// read where they lie and not copied.
__global__ __launch_bounds__(256) void emit_count_kernel(
    int n_host, const int* __restrict__ n_dev, const uint32_t* __restrict__ perm,
    const uint2* __restrict__ rect_by_id, uint2* __restrict__ rects,
    uint32_t* __restrict__ part_tot, uint32_t* __restrict__ part_nz) {
    __shared__ uint32_t s_w[4], s_z[4];
    const int n = n_dev ? *n_dev : n_host;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int r0 = blockIdx.x * kEmitPart + tid;
    uint32_t g[kEmitPer];
    uint2 rc[kEmitPer];
#pragma unroll
    for (int k = 0; k < kEmitPer; ++k)
        g[k] = r0 + k * 256 < n ? (perm ? perm[r0 + k * 256] : (uint32_t)(r0 + k * 256))
                                : 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < kEmitPer; ++k)
        rc[k] = g[k] != 0xFFFFFFFFu ? rect_by_id[g[k]] : make_uint2(0u, 0u);
    uint32_t sum = 0, nz = 0;
#pragma unroll
    for (int k = 0; k < kEmitPer; ++k) {
        if (perm && r0 + k * 256 < n) rects[r0 + k * 256] = rc[k];
        const uint32_t cnt = rect_count(rc[k]);
        sum += cnt;
        nz += cnt > 0 ? 1u : 0u;
    }
    const uint32_t ws = wave_inclusive_sum(sum);
    const uint32_t wz = wave_inclusive_sum(nz);
    if (lane == 63) {
        s_w[tid >> 6] = ws;
        s_z[tid >> 6] = wz;
    }
    __syncthreads();
    if (tid == 0) {
        part_tot[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        part_nz[blockIdx.x] = s_z[0] + s_z[1] + s_z[2] + s_z[3];
    }
}

// One workgroup of 1024 threads; thread t scans a contiguous run of partitions.
// info[0] = slots written (min(total, cap)), info[1] = emitting entries.
__global__ __launch_bounds__(1024) void emit_scan_kernel(
    const uint32_t* __restrict__ part_tot, const uint32_t* __restrict__ part_nz, int nparts,
    uint64_t* __restrict__ part_off, uint32_t* __restrict__ part_nzoff, int cap,
    int* __restrict__ info, unsigned long long* __restrict__ total_out,
    uint32_t* __restrict__ err) {
    __shared__ uint64_t s_w[16];
    __shared__ uint32_t s_z[16];
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const int per = (nparts + 1023) / 1024;
    const int b0 = min(tid * per, nparts), b1 = min(b0 + per, nparts);
    uint64_t sum = 0;
    uint32_t nz = 0;
    for (int i = b0; i < b1; ++i) {
        sum += part_tot[i];
        nz += part_nz[i];
    }
    const uint64_t incl = wave_inclusive_sum(sum);
    const uint32_t zincl = wave_inclusive_sum(nz);
    if (lane == 63) {
        s_w[wave] = incl;
        s_z[wave] = zincl;
    }
    __syncthreads();
    uint64_t run = incl - sum;
    uint32_t zrun = zincl - nz;
    for (int w = 0; w < wave; ++w) {
        run += s_w[w];
        zrun += s_z[w];
    }
    for (int i = b0; i < b1; ++i) {
        part_off[i] = run;
        part_nzoff[i] = zrun;
        run += part_tot[i];
        zrun += part_nz[i];
    }
    if (tid == 1023) {                   // run and zrun are now the grand totals
        if (total_out) *total_out = (unsigned long long)run;
        info[0] = run > (uint64_t)cap ? cap : (int)run;
        info[1] = (int)zrun;
        if (run > (uint64_t)cap) atomicOr(err, 8u);
    }
}

// Compacted emitting entries (see the overview above).
struct EmitEntries {
    uint32_t* off;      // first slot, clamped to the slots written
    uint2* rect;        // packed rectangle
    uint32_t* gid;      // gaussian / item id
    int2* key;          // BATCH: (bin_off, tiles_x) of the entry's view
    uint32_t* chunk_owner;
    uint32_t* first_by_id;  // optional: first slot by id, for emitting entries
    // with cum set: the entries are all Gaussians in id order (entry j is
    // Gaussian j, emitting or not), their rectangles rect_by_id and first
    // slots cum[j] - count (cum: inclusive sums of the rectangles' counts)
    const int* cum;
    const uint2* rect_by_id;
};

// The tile sort's first-pass digit counts, produced by the expansion (see
// sort_first_counts): with counts set, workgroup b of the expansion writes
// exactly the slots of sort partition b (tile slots) and stores the counts of
// digit (key >> shift) & (2^bits - 1) of the keys it wrote at
// counts[d * parts + b] for all 256 digits, as the sort's counting kernel
// would.  The grid is then parts workgroups.
struct FirstCounts {
    uint32_t* counts;
    int tile, parts, shift, bits;
};

FirstCounts first_counts(const SortPlan& sp, uint8_t* sort_scratch) {
    FirstCounts fc;
    fc.counts = sp.passes > 0 ? sort_first_counts(sp, sort_scratch) : nullptr;
    fc.tile = sp.tile;
    fc.parts = sp.parts;
    fc.shift = sp.first_shift;
    fc.bits = sp.first_bits;
    return fc;
}

template <bool BATCH>
__global__ __launch_bounds__(256) void emit_compact_kernel(
    int n_host, const int* __restrict__ n_dev, const ViewRec* __restrict__ views,
    const uint32_t* __restrict__ vkeys, const uint32_t* __restrict__ perm,
    const uint2* __restrict__ rects, const uint64_t* __restrict__ part_off,
    const uint32_t* __restrict__ part_nzoff, const int* __restrict__ info, EmitEntries E) {
    __shared__ uint32_t s_w[kEmitPer][4], s_z[kEmitPer][4];
    const int n = n_dev ? *n_dev : n_host;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint2 rc[kEmitPer];
    uint32_t cnt[kEmitPer], incl[kEmitPer];
    uint64_t nzb[kEmitPer];
#pragma unroll
    for (int k = 0; k < kEmitPer; ++k) {
        const int r = blockIdx.x * kEmitPart + k * 256 + tid;
        rc[k] = r < n ? rects[r] : make_uint2(0u, 0u);
        cnt[k] = rect_count(rc[k]);
        incl[k] = wave_inclusive_sum(cnt[k]);
        nzb[k] = ballot(cnt[k] > 0);
        if (lane == 63) {
            s_w[k][wave] = incl[k];
            s_z[k][wave] = (uint32_t)__popcll(nzb[k]);
        }
    }
    __syncthreads();
    const uint64_t limit = (uint64_t)info[0];
    uint64_t base = part_off[blockIdx.x];
    uint32_t zbase = part_nzoff[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kEmitPer; ++k) {
        uint64_t w_off = 0, slab = 0;
        uint32_t z_off = 0, zslab = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t v = s_w[k][w], z = s_z[k][w];
            w_off += (w < wave) ? v : 0u;
            z_off += (w < wave) ? z : 0u;
            slab += v;
            zslab += z;
        }
        const int r = blockIdx.x * kEmitPart + k * 256 + tid;
        const uint64_t off = base + w_off + (incl[k] - cnt[k]);
        const uint32_t j = zbase + z_off + (uint32_t)__popcll(nzb[k] & below);
        base += slab;
        zbase += zslab;
        if (cnt[k] == 0) continue;          // also every r >= n
        const uint32_t id = perm ? perm[r] : (uint32_t)r;
        E.off[j] = (uint32_t)(off < limit ? off : limit);
        E.rect[j] = rc[k];
        E.gid[j] = id;
        if (E.first_by_id) E.first_by_id[id] = (uint32_t)(off < limit ? off : limit);
        if (BATCH) {
            const ViewRec& vr = views[vkeys[r]];
            E.key[j] = make_int2(vr.bin_off, vr.tiles_x);
        }
        if (off < limit) {
            // chunks whose first slot lies in [off, min(off + cnt, limit))
            const uint64_t end = off + cnt[k] < limit ? off + cnt[k] : limit;
            for (uint64_t ch = (off + 63) >> 6; (ch << 6) < end; ++ch)
                E.chunk_owner[ch] = j;
        }
    }
}

// Wave gw expands the chunks [c0, c0 + per) of the output, an equal share of
// the slots.  It slides a window of 64 consecutive compacted entries (one per
// lane, loaded once and staged in LDS) from the owner of its first slot on;
// inside a window each 64-slot chunk gets its owners from an LDS mark per
// entry start and a DPP max-scan, slots before the chunk's first mark
// belonging to the owner of the previous chunk's last slot.  A mark is
// (step << 7) | (lane + 1) with a step number that grows with every chunk the
// wave processes, so older marks never win the max-scan and the mark row is
// cleared only once.  With compacted entries every entry emits at least one
// slot, so a window covers at least 64 slots; with the entries of E.cum (all
// Gaussians) a window may cover fewer, even none, and the wave slides on.
template <bool BATCH>
__global__ __launch_bounds__(256) void emit_expand_kernel(
    int tiles_x_plain, const EmitEntries E, const int* __restrict__ info,
    uint32_t* __restrict__ tile_out, uint32_t* __restrict__ gid_out, FirstCounts fc) {
    __shared__ int s_mark[4][64];
    __shared__ uint4 s_ent[4][64];      // (gid, x0 | y0 << 16, w, first slot)
    __shared__ int2 s_key[4][64];       // BATCH: (bin_off, tiles_x)
    __shared__ uint32_t s_hist[4][256]; // fc: per-wave digit counts of the keys written
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const uint32_t limit = (uint32_t)info[0];
    const uint32_t nnz = (uint32_t)info[1];
    const uint32_t nchunks = (limit + 63u) >> 6;
    uint32_t per, c0;
    if (fc.counts) {                    // workgroup = sort partition of fc.tile slots
        per = (uint32_t)fc.tile >> 8;
        c0 = blockIdx.x * ((uint32_t)fc.tile >> 6) + (uint32_t)wave * per;
#pragma unroll
        for (int w = 0; w < 4; ++w) s_hist[w][threadIdx.x] = 0u;
        __syncthreads();
    } else {
        const uint32_t nw = gridDim.x * 4u;
        per = (nchunks + nw - 1u) / nw;
        c0 = (blockIdx.x * 4u + (uint32_t)wave) * per;
    }
    const uint32_t dmask = (1u << fc.bits) - 1u;
    if (c0 < nchunks) {
        const uint32_t end = min((c0 + per) << 6, limit);   // limit < 2^31
        uint32_t pos = c0 << 6;                               // first slot not yet written
        uint32_t cur = (uint32_t)uniform_i((int)E.chunk_owner[c0]);   // owns slot pos
        int step = 0;       // < 2^24: at most per + windows steps, per <= 2^25 / 64
        s_mark[wave][lane] = 0;
        while (pos < end) {
            const uint32_t j = cur + (uint32_t)lane;
            uint32_t off = limit, gid = 0u;
            uint2 rc = make_uint2(0u, 0u);
            if (j < nnz) {
                if (E.cum) {
                    rc = E.rect_by_id[j];
                    gid = j;
                    const uint32_t first = (uint32_t)E.cum[j] - rect_count(rc);
                    off = first < limit ? first : limit;
                } else {
                    off = E.off[j];
                    rc = E.rect[j];
                    gid = E.gid[j];
                }
                if (BATCH) s_key[wave][lane] = E.key[j];
            }
            const uint32_t cnt = rect_count(rc);
            s_ent[wave][lane] = make_uint4(gid, rc.x, rc.y & 0xFFFFu, off);
            // the window's slots end where its last entry's run ends
            const uint32_t run_end = cnt > limit - off ? limit : off + cnt;
            const uint32_t seg = min((uint32_t)__builtin_amdgcn_readlane((int)run_end, 63), end);
            int carry = 0;      // lane 0 owns pos in the first window; later windows start at pos
            for (uint32_t s0 = pos & ~63u; s0 < seg; s0 += 64u) {
                ++step;
                const int tag = step << 7;
                if (cnt > 0 && off >= s0 && off < s0 + 64u && off < seg)
                    s_mark[wave][off - s0] = tag | (lane + 1);
                wave_lds_sync();
                const int v = wave_max_scan_nonneg(s_mark[wave][lane]);
                const int owner = v >= tag ? (v & 127) - 1 : carry;
                carry = __builtin_amdgcn_readlane(owner, 63);
                const uint4 od = s_ent[wave][owner];
                const uint32_t slot = s0 + (uint32_t)lane;
                if (slot >= pos && slot < seg) {
                    // q = kk / w from a float estimate (|error| < 1 for q < 2^16) and one correction
                    const uint32_t kk = slot - od.w;
                    uint32_t q = (uint32_t)((float)kk * __builtin_amdgcn_rcpf((float)od.z));
                    int rem = (int)(kk - q * od.z);
                    if (rem < 0) {
                        --q;
                        rem += (int)od.z;
                    } else if (rem >= (int)od.z) {
                        ++q;
                        rem -= (int)od.z;
                    }
                    const uint32_t ty = (od.y >> 16) + q;
                    const uint32_t tx = (od.y & 0xFFFFu) + (uint32_t)rem;
                    const int2 kb = BATCH ? s_key[wave][owner] : make_int2(0, tiles_x_plain);
                    const uint32_t key = (uint32_t)kb.x + ty * (uint32_t)kb.y + tx;
                    tile_out[slot] = key;
                    gid_out[slot] = od.x;
                    if (fc.counts) atomicAdd(&s_hist[wave][(key >> fc.shift) & dmask], 1u);
                }
            }
            pos = seg > pos ? seg : pos;
            cur += 64u;
            wave_lds_sync();    // the next window overwrites s_ent / s_key
        }
    }
    if (fc.counts) {
        __syncthreads();
        const int d = threadIdx.x;
        fc.counts[(size_t)d * fc.parts + blockIdx.x] =
            (s_hist[0][d] + s_hist[1][d]) + (s_hist[2][d] + s_hist[3][d]);
    }
}

// kTileBinsPer consecutive sorted ids per lane (one 16-byte load when
// aligned); an edge between ids i - 1 and i closes bin tiles[i - 1] and opens
// bin tiles[i].
constexpr int kTileBinsPer = 4;
__global__ __launch_bounds__(256) void tile_bins_kernel(const uint32_t* __restrict__ tiles,
                                                        int n_host, const int* __restrict__ n_dev,
                                                        int num_tiles, int2* __restrict__ bins,
                                                        uint32_t* __restrict__ err) {
    const int n = n_dev ? *n_dev : n_host;
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * kTileBinsPer;
    if (i0 >= n) return;
    uint32_t k[kTileBinsPer];
    if (i0 + kTileBinsPer <= n && (reinterpret_cast<uintptr_t>(tiles) & 15u) == 0) {
        const uint4 v = *reinterpret_cast<const uint4*>(tiles + i0);
        k[0] = v.x;
        k[1] = v.y;
        k[2] = v.z;
        k[3] = v.w;
    } else {
#pragma unroll
        for (int j = 0; j < kTileBinsPer; ++j) k[j] = i0 + j < n ? tiles[i0 + j] : 0u;
    }
    uint32_t prev = i0 > 0 ? tiles[i0 - 1] : 0u;
#pragma unroll
    for (int j = 0; j < kTileBinsPer; ++j) {
        const int i = i0 + j;
        if (i >= n) break;
        const uint32_t t = k[j];
        if (t >= (uint32_t)num_tiles || prev >= (uint32_t)num_tiles) {   // never for valid input
            atomicOr(err, 4u);
            return;
        }
        if (i == 0) bins[t].x = 0;
        if (i == n - 1) bins[t].y = n;
        if (i > 0 && prev != t) {
            bins[prev].y = i;
            bins[t].x = i;
        }
        prev = t;
    }
}

// ---------------------------------------------------------------------------
// Per-tile depth order
// ---------------------------------------------------------------------------
// A tile's run [bins[t].x, bins[t].y) of the tile-sorted ids holds the
// Gaussians whose rectangle covers the tile, in any order.  Ordering it by
// (depth bits, id) gives the reference order inside the tile.
//
// Bucket order: every entry gets the composite v = (key - kmin) << ib |
// (id - imin), ib the bit length of the id range, whose order is the
// (key, id) order; the run is split into 2^11 buckets linear in v (the shift
// makes the largest v fit 11 bits, so at least half of the buckets are in
// range).  An LDS histogram with returning atomics gives every entry its
// bucket and a slot in it, a scan the bucket starts, and the entries are
// placed by bucket.  Each entry's final position is then its bucket's start
// plus the number of entries of its bucket that precede it in (key, id)
// order, counted by reading the bucket.  Depths inside a tile are spread over
// a range of floats (at 1080p with 1M Gaussians about 800 entries per tile,
// at most 12 in a bucket), and where many keys are equal the ids spread the
// buckets instead; the order is exact whatever the distribution, only the
// counting reads grow with the bucket sizes.
//
// One 256-thread workgroup per tile: runs of up to kSegCap entries are held
// in registers and placed in LDS; longer ones go through global scratch at
// the run's own offsets (L2-resident), with the bucket counts in LDS.
constexpr int kSegThreads = 256;
constexpr int kSegWaves = kSegThreads / kWave;
constexpr int kSegItems = 6;                           // entries per lane
constexpr int kSegCap = kSegThreads * kSegItems;       // 1536
constexpr int kSegBucketBits = 11;
constexpr int kSegBuckets = 1 << kSegBucketBits;

// Unsigned min / max over the 64 lanes (full EXEC): DPP inside the rows,
// then the four rows by readlane.
template <bool MAX>
__device__ __forceinline__ uint32_t wave_extreme_u32(uint32_t v) {
    auto op = [](uint32_t a, uint32_t b) { return MAX ? max(a, b) : min(a, b); };
    int a = (int)v;
    a = (int)op((uint32_t)a, (uint32_t)dpp_i<kDppQuadPerm1032>(a));
    a = (int)op((uint32_t)a, (uint32_t)dpp_i<kDppQuadPerm2301>(a));
    a = (int)op((uint32_t)a, (uint32_t)dpp_i<kDppRowRor4>(a));
    a = (int)op((uint32_t)a, (uint32_t)dpp_i<kDppRowRor8>(a));
    return op(op((uint32_t)__builtin_amdgcn_readlane(a, 0), (uint32_t)__builtin_amdgcn_readlane(a, 16)),
              op((uint32_t)__builtin_amdgcn_readlane(a, 32), (uint32_t)__builtin_amdgcn_readlane(a, 48)));
}

// 20 KB, so that eight workgroups fit a CU.  red and tmp are used before the
// entries are placed, so they share kv's space.
struct SegShared {
    union {
        uint2 kv[kSegCap];                          // placed (key, id)
        struct {
            uint32_t red[4 * kSegWaves];
            uint32_t tmp[kSegWaves];
        };
    };
    uint32_t cnt[kSegBuckets];                      // bucket counts / starts; then the ids
};
static_assert(kSegBuckets >= kSegCap, "the ids are staged in cnt");
static_assert(sizeof(SegShared) <= 20480, "eight workgroups per CU");
static_assert(kSegBuckets % kSegThreads == 0, "bucket scan: whole runs per thread");

// The run's composite bucketing: v = (key - kmin) << ib | (id - imin),
// bucket = v >> shift.
struct Buckets {
    uint32_t kmin, imin;
    int ib, shift;
    __device__ __forceinline__ uint32_t of(uint32_t key, uint32_t id) const {
        const uint64_t v = ((uint64_t)(key - kmin) << ib) | (uint64_t)(id - imin);
        return (uint32_t)(v >> shift);
    }
};

// Block min / max of the keys and ids (ids of a run of two or more entries
// differ, so the composite range is never empty).  Ends with a barrier.
__device__ __forceinline__ Buckets block_buckets(uint32_t kmin, uint32_t kmax, uint32_t imin,
                                                 uint32_t imax, SegShared& sh) {
    const int wave = threadIdx.x >> 6;
    kmin = wave_extreme_u32<false>(kmin);
    kmax = wave_extreme_u32<true>(kmax);
    imin = wave_extreme_u32<false>(imin);
    imax = wave_extreme_u32<true>(imax);
    if (lane_id() == 0) {
        sh.red[wave] = kmin;
        sh.red[kSegWaves + wave] = kmax;
        sh.red[2 * kSegWaves + wave] = imin;
        sh.red[3 * kSegWaves + wave] = imax;
    }
    __syncthreads();
    kmin = sh.red[0];
    kmax = sh.red[kSegWaves];
    imin = sh.red[2 * kSegWaves];
    imax = sh.red[3 * kSegWaves];
#pragma unroll
    for (int w = 1; w < kSegWaves; ++w) {
        kmin = min(kmin, sh.red[w]);
        kmax = max(kmax, sh.red[kSegWaves + w]);
        imin = min(imin, sh.red[2 * kSegWaves + w]);
        imax = max(imax, sh.red[3 * kSegWaves + w]);
    }
    __syncthreads();
    Buckets bk;
    bk.kmin = kmin;
    bk.imin = imin;
    bk.ib = imax > imin ? 32 - __builtin_clz(imax - imin) : 0;
    const uint64_t vmax = ((uint64_t)(kmax - kmin) << bk.ib) | (uint64_t)(imax - imin);
    const int vbits = 64 - __builtin_clzll(vmax);
    bk.shift = vbits > kSegBucketBits ? vbits - kSegBucketBits : 0;
    return bk;
}

// LDS word of bucket b's count.  The scan below has thread t read buckets
// 8t..8t+7, which would put the lanes t, t+8, ... on one bank; rotating the
// low three bits by the 64-word row (a permutation inside each 8-word group)
// gives the 64 lanes of a read 64 different banks.
__device__ __forceinline__ uint32_t bucket_word(uint32_t b) {
    return (b & ~7u) | ((b + (b >> 6)) & 7u);
}

// Counts in sh.cnt[bucket_word(0..kSegBuckets)) -> exclusive starts (the last
// bucket ends at the run's length).  Each thread scans a run of consecutive
// buckets.  Ends with a barrier.
__device__ __forceinline__ void bucket_scan(SegShared& sh) {
    constexpr int per = kSegBuckets / kSegThreads;
    const int tid = threadIdx.x;
    const int b0 = tid * per;
    uint32_t c[per];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        c[k] = sh.cnt[bucket_word(b0 + k)];
        sum += c[k];
    }
    uint32_t run = block_exclusive_scan256(sum, sh.tmp);
#pragma unroll
    for (int k = 0; k < per; ++k) {
        sh.cnt[bucket_word(b0 + k)] = run;
        run += c[k];
    }
    __syncthreads();
}

// End of bucket bk: the next bucket's start, or n for the last one.
__device__ __forceinline__ uint32_t bucket_end(const SegShared& sh, uint32_t bk, int n) {
    return bk + 1 < (uint32_t)kSegBuckets ? sh.cnt[bucket_word(bk + 1)] : (uint32_t)n;
}

// (key, id) order of two entries.
__device__ __forceinline__ bool entry_before(uint2 a, uint32_t key, uint32_t id) {
    return a.x < key || (a.x == key && a.y < id);
}

// Rank of (key, id) among the placed entries kv[s, s + c).
template <typename KV>
__device__ __forceinline__ uint32_t bucket_rank(KV kv, uint32_t s, uint32_t c, uint32_t key,
                                                uint32_t id) {
    uint32_t rank = 0;
    for (uint32_t j = s; j < s + c; ++j) rank += entry_before(kv[j], key, id) ? 1u : 0u;
    return rank;
}

// Runs of more than kSegCap entries: the same order through global scratch
// at the run's offsets: ka the keys, va the slots in the buckets, kv the
// placed (key, id).  Every lane of the block calls it (barriers inside).
__device__ void long_run_order(int base, int n, const uint32_t* __restrict__ dkeys,
                               uint32_t* __restrict__ ids, uint32_t* __restrict__ ka,
                               uint32_t* __restrict__ va, uint2* __restrict__ kv,
                               SegShared& sh) {
    const int tid = threadIdx.x;
    uint32_t kmin = ~0u, kmax = 0u, imin = ~0u, imax = 0u;
    for (int e = tid; e < n; e += kSegThreads) {
        const uint32_t id = ids[base + e];
        const uint32_t k = dkeys[id];
        ka[base + e] = k;
        kmin = min(kmin, k);
        kmax = max(kmax, k);
        imin = min(imin, id);
        imax = max(imax, id);
    }
    const Buckets bk = block_buckets(kmin, kmax, imin, imax, sh);
    for (int i = tid; i < kSegBuckets; i += kSegThreads) sh.cnt[i] = 0u;
    __syncthreads();
    for (int e = tid; e < n; e += kSegThreads)
        va[base + e] = atomicAdd(&sh.cnt[bucket_word(bk.of(ka[base + e], ids[base + e]))], 1u);
    __syncthreads();
    bucket_scan(sh);
    for (int e = tid; e < n; e += kSegThreads) {
        const uint32_t k = ka[base + e];
        const uint32_t id = ids[base + e];
        kv[base + sh.cnt[bucket_word(bk.of(k, id))] + va[base + e]] = make_uint2(k, id);
    }
    __syncthreads();                                    // ids are no longer read
    for (int j = tid; j < n; j += kSegThreads) {
        const uint2 me = kv[base + j];
        const uint32_t b = bk.of(me.x, me.y);
        const uint32_t st = sh.cnt[bucket_word(b)];
        const uint32_t rank = bucket_rank(kv + base, st, bucket_end(sh, b, n) - st, me.x, me.y);
        ids[base + st + rank] = me.y;
    }
}

// Orders run t (all lanes of the block; barriers inside).
__device__ __forceinline__ void order_tile_run(int t, const int2* __restrict__ bins,
                                               const uint32_t* __restrict__ dkeys,
                                               uint32_t* __restrict__ ids,
                                               uint32_t* __restrict__ ka,
                                               uint32_t* __restrict__ va,
                                               uint2* __restrict__ kv_tmp, SegShared& sh) {
    const int2 b = bins[t];
    const int n = b.y - b.x;
    if (n <= 1) return;                                 // uniform
    if (n > kSegCap) {
        long_run_order(b.x, n, dkeys, ids, ka, va, kv_tmp, sh);
        return;
    }
    const int tid = threadIdx.x;
    uint32_t key[kSegItems], val[kSegItems];
    bool valid[kSegItems];
#pragma unroll
    for (int r = 0; r < kSegItems; ++r) {
        const int e = r * kSegThreads + tid;
        valid[r] = e < n;
        val[r] = valid[r] ? ids[b.x + e] : 0u;
    }
    // the bucket counts are cleared while the loads are in flight
    for (int i = tid; i < kSegBuckets; i += kSegThreads) sh.cnt[i] = 0u;
    uint32_t kmin = ~0u, kmax = 0u, imin = ~0u, imax = 0u;
#pragma unroll
    for (int r = 0; r < kSegItems; ++r) {
        key[r] = valid[r] ? dkeys[val[r]] : 0u;
        if (valid[r]) {
            kmin = min(kmin, key[r]);
            kmax = max(kmax, key[r]);
            imin = min(imin, val[r]);
            imax = max(imax, val[r]);
        }
    }
    const Buckets bkt = block_buckets(kmin, kmax, imin, imax, sh);
    uint32_t bk[kSegItems], slot[kSegItems];
#pragma unroll
    for (int r = 0; r < kSegItems; ++r) {
        bk[r] = valid[r] ? bkt.of(key[r], val[r]) : 0u;
        slot[r] = valid[r] ? atomicAdd(&sh.cnt[bucket_word(bk[r])], 1u) : 0u;
    }
    __syncthreads();
    bucket_scan(sh);
    uint32_t st[kSegItems], cnt[kSegItems];
#pragma unroll
    for (int r = 0; r < kSegItems; ++r) {
        st[r] = 0u;
        cnt[r] = 0u;
        if (valid[r]) {
            st[r] = sh.cnt[bucket_word(bk[r])];
            cnt[r] = bucket_end(sh, bk[r], n) - st[r];
            sh.kv[st[r] + slot[r]] = make_uint2(key[r], val[r]);
        }
    }
    __syncthreads();
    uint32_t pos[kSegItems];
#pragma unroll
    for (int r = 0; r < kSegItems; ++r)
        pos[r] = valid[r] ? st[r] + bucket_rank(sh.kv, st[r], cnt[r], key[r], val[r]) : 0u;
    __syncthreads();                                    // sh.cnt is no longer read
#pragma unroll
    for (int r = 0; r < kSegItems; ++r)
        if (valid[r]) sh.cnt[pos[r]] = val[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegItems; ++r) {
        const int e = r * kSegThreads + tid;
        if (valid[r]) ids[b.x + e] = sh.cnt[e];
    }
}

// One workgroup per tile (see above), mapped as the raster kernels map the
// tiles (xcd_tile), so that each XCD orders a band of tiles whose ids its
// raster workgroups then read through the same L2.
__global__ __launch_bounds__(kSegThreads) void tile_depth_sort_kernel(
    const int2* __restrict__ bins, int num_bins, const uint32_t* __restrict__ dkeys,
    uint32_t* __restrict__ ids, uint32_t* __restrict__ ka, uint32_t* __restrict__ va,
    uint2* __restrict__ kv_tmp) {
    __shared__ SegShared sh;
    const int t = xcd_tile(blockIdx.x, num_bins);
    if (t < num_bins) order_tile_run(t, bins, dkeys, ids, ka, va, kv_tmp, sh);
}

// A batch's bins (view << tile_bits | tile: many empty, many short): each
// workgroup takes the runs blockIdx.x, + gridDim.x, ..., so an empty one costs
// a load rather than a workgroup.  At least 6 waves per SIMD (78 VGPRs; the
// loop took 87 unbounded).
__global__ __launch_bounds__(kSegThreads, 6) void tile_depth_sort_loop_kernel(
    const int2* __restrict__ bins, int num_bins, const uint32_t* __restrict__ dkeys,
    uint32_t* __restrict__ ids, uint32_t* __restrict__ ka, uint32_t* __restrict__ va,
    uint2* __restrict__ kv_tmp) {
    __shared__ SegShared sh;
    for (int t = blockIdx.x; t < num_bins; t += gridDim.x) {
        order_tile_run(t, bins, dkeys, ids, ka, va, kv_tmp, sh);
        __syncthreads();                                // LDS is reused by the next run
    }
}

// The per-tile depth order of every run of bins[0..num_bins).  ka, va: free
// buffers of the intersection capacity; kv_tmp: a free one of twice that.
// xcd: the bins are one view's tiles in the raster kernels' numbering (the
// batch's bins, view << tile_bits | tile, are not: a band per XCD would also
// be uneven there).
hipError_t tile_depth_order(const int2* bins, int num_bins, bool xcd, const uint32_t* dkeys,
                            uint32_t* ids, uint32_t* ka, uint32_t* va, uint2* kv_tmp,
                            hipStream_t s) {
    if (num_bins <= 0) return hipSuccess;
    // xcd_tile's grid, or two runs per workgroup (A/B at 6 waves per SIMD:
    // multi32 +7%, orbit64 -1% against one run per workgroup; 4 or 8 runs
    // cost orbit64's long runs more)
    const int grid = xcd ? 8 * ((num_bins + 7) / 8) : (num_bins + 1) / 2;
    if (xcd)
        tile_depth_sort_kernel<<<grid, kSegThreads, 0, s>>>(bins, num_bins, dkeys, ids, ka, va,
                                                           kv_tmp);
    else
        tile_depth_sort_loop_kernel<<<grid, kSegThreads, 0, s>>>(bins, num_bins, dkeys, ids, ka,
                                                                va, kv_tmp);
    return hipGetLastError();
}

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Emission buffers (passes 1-4 above) for n entries and a capacity of cap slots.
struct EmitLayout {
    size_t part_tot, part_nz, part_off, part_nzoff, info, rects, off, rect, gid, key, chunk_owner;
};

size_t emit_layout(size_t o, int n, int cap, bool batch, EmitLayout& E) {
    const size_t parts = (size_t)((n + kEmitPart - 1) / kEmitPart) + 1;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const size_t chunks = (size_t)(cap > 0 ? cap : 1) / 64 + 1;
    E.part_tot = o; o = align256(o + 4 * parts);
    E.part_nz = o; o = align256(o + 4 * parts);
    E.part_off = o; o = align256(o + 8 * parts);
    E.part_nzoff = o; o = align256(o + 4 * parts);
    E.info = o; o = align256(o + 16);
    E.rects = o; o = align256(o + 8 * nn);
    E.off = o; o = align256(o + 4 * nn);
    E.rect = o; o = align256(o + 8 * nn);
    E.gid = o; o = align256(o + 4 * nn);
    E.key = o; o = align256(o + (batch ? 8 * nn : 0));
    E.chunk_owner = o; o = align256(o + 4 * chunks);
    return o;
}

// Emission straight from the inclusive counts (E.cum): every 64-slot chunk's
// owner, the Gaussian whose slots [cum - count, cum) hold the chunk's first
// slot, plus the totals of emit_scan_kernel (info[0] = slots written =
// min(total, cap), info[1] = n, *total_out) and the binning's error word
// (8 on overflow, else 0).  Threads t < num_tiles also clear bins[t] (the
// grid covers max(n, num_tiles) threads).
__global__ __launch_bounds__(256) void cum_owner_kernel(int n, const int* __restrict__ cum,
                                                        const uint2* __restrict__ rects,
                                                        int cap, int* __restrict__ info,
                                                        unsigned long long* __restrict__ total_out,
                                                        uint32_t* __restrict__ err,
                                                        uint32_t* __restrict__ chunk_owner,
                                                        int num_tiles, int2* __restrict__ bins) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < num_tiles) bins[g] = make_int2(0, 0);
    if (g >= n) return;
    const uint32_t total = (uint32_t)cum[n - 1];
    const uint32_t limit = total < (uint32_t)cap ? total : (uint32_t)cap;
    if (g == n - 1) {
        *total_out = total;
        *err = total > (uint32_t)cap ? 8u : 0u;
        info[0] = (int)limit;
        info[1] = n;
    }
    const uint32_t cnt = rect_count(rects[g]);
    if (cnt == 0) return;
    const uint32_t off = (uint32_t)cum[g] - cnt;
    if (off >= limit) return;
    const uint32_t end = cnt > limit - off ? limit : off + cnt;
    for (uint32_t ch = (off + 63u) >> 6; (ch << 6) < end; ++ch) chunk_owner[ch] = (uint32_t)g;
}

// Passes 1-4 over the entries perm[0..n) in emission order (perm == nullptr:
// the ids 0..n-1), n = *n_dev when given, else n_host; launch shapes follow
// n_host, an upper bound.  Writes the slots written (min(total, cap)) to
// info[0] and the number of emitting entries to info[1] (info: the layout's
// slot if null), the exact total to *total if total is not null, and bit 8 of
// err when the total exceeds cap.  first_by_id (optional) receives the first
// slot of every emitting id (clamped to the slots written).
template <bool BATCH>
hipError_t emit(int n, const int* n_dev, int tiles_x_plain, const ViewRec* views,
                const uint32_t* vkeys,
                const uint32_t* perm, const uint2* rect_by_id, int cap, uint8_t* b,
                const EmitLayout& L, int* info, unsigned long long* total, uint32_t* err,
                uint32_t* tile_out, uint32_t* gid_out, uint32_t* first_by_id,
                const FirstCounts& fc, hipStream_t s) {
    const int parts = (n + kEmitPart - 1) / kEmitPart;
    auto P = [&](size_t off) { return b + off; };
    uint32_t* part_tot = reinterpret_cast<uint32_t*>(P(L.part_tot));
    uint32_t* part_nz = reinterpret_cast<uint32_t*>(P(L.part_nz));
    uint64_t* part_off = reinterpret_cast<uint64_t*>(P(L.part_off));
    uint32_t* part_nzoff = reinterpret_cast<uint32_t*>(P(L.part_nzoff));
    if (!info) info = reinterpret_cast<int*>(P(L.info));
    uint2* rects = reinterpret_cast<uint2*>(P(L.rects));
    EmitEntries E;
    E.off = reinterpret_cast<uint32_t*>(P(L.off));
    E.rect = reinterpret_cast<uint2*>(P(L.rect));
    E.gid = reinterpret_cast<uint32_t*>(P(L.gid));
    E.key = BATCH ? reinterpret_cast<int2*>(P(L.key)) : nullptr;
    E.chunk_owner = reinterpret_cast<uint32_t*>(P(L.chunk_owner));
    E.first_by_id = first_by_id;
    E.cum = nullptr;
    E.rect_by_id = nullptr;
    emit_count_kernel<<<parts, 256, 0, s>>>(n, n_dev, perm, rect_by_id, rects, part_tot,
                                            part_nz);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    emit_scan_kernel<<<1, 1024, 0, s>>>(part_tot, part_nz, parts, part_off, part_nzoff, cap,
                                        info, total, err);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    emit_compact_kernel<BATCH><<<parts, 256, 0, s>>>(n, n_dev, views, vkeys, perm,
                                                     perm ? rects : rect_by_id, part_off,
                                                     part_nzoff, info, E);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const long long chunks = ((long long)cap + 63) / 64;
    const int waves = (int)(chunks < kExpandWaves ? chunks : kExpandWaves);
    const int blocks = fc.counts ? fc.parts : (waves + 3) / 4;
    emit_expand_kernel<BATCH><<<blocks > 0 ? blocks : 1, 256, 0, s>>>(tiles_x_plain, E, info,
                                                                      tile_out, gid_out, fc);
    return hipGetLastError();
}

// Depth stage: per-Gaussian depth keys and rectangles.  It needs no
// intersection count, so a caller can run it while it reads the count back;
// its workspace depends on the Gaussian count only.
struct DepthLayout {
    size_t dkeys, rect_by_id, end;
};

DepthLayout depth_layout(const BinPlan& p) {
    DepthLayout L;
    size_t o = 0;
    const size_t n4 = align256(4ull * (p.n > 0 ? p.n : 1));
    L.dkeys = o; o += n4;
    L.rect_by_id = o; o += 2 * n4;
    L.end = o;
    return L;
}

// Emission stage (emission, tile sort, bin edges, per-tile depth order); the
// error word comes first.  The zeroed header is part of the C ABI
// (jaxsplat_bin_emit_sort): the error word at byte 0 and the exact
// intersection total (uint64) at byte 8.
struct BinLayout {
    size_t err, total, zero_end;
    size_t tile_sort, tiles, gids, ttmp_k, ttmp_v, tiles_sorted, end;
    EmitLayout emit;
};

BinLayout layout(const BinPlan& p) {
    BinLayout L;
    size_t o = 0;
    L.err = o;
    L.total = o + 8;
    o += 16;
    L.zero_end = o;
    o = align256(o);
    L.tile_sort = o; o = align256(o + p.tile_sort.scratch_bytes);
    o = emit_layout(o, p.n, p.isect_cap, false, L.emit);
    const size_t i4 = align256(4ull * (p.isect_cap > 0 ? p.isect_cap : 1));
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
    p.tile_sort = sort_plan(isect_cap, 0, p.tile_bits);
    p.depth_bytes = align256(depth_layout(p).end);
    p.emit_bytes = align256(layout(p).end);
    p.scratch_bytes = p.emit_bytes + p.depth_bytes;
    return p;
}

const uint32_t* bin_error_word(const BinPlan& p, const void* scratch) {
    return reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(scratch) + layout(p).err);
}

hipError_t bin_depth(const BinPlan& p, const Camera& c, const float* xys, const float* depths,
                     const int* radii, const uint2* rects, void* depth_ws, hipStream_t s) {
    if (p.n <= 0) return hipSuccess;
    const DepthLayout D = depth_layout(p);
    uint8_t* b = static_cast<uint8_t*>(depth_ws);
    depth_keys_kernel<<<(p.n + 255) / 256, 256, 0, s>>>(
        c, depths, radii, reinterpret_cast<const float2*>(xys), rects,
        reinterpret_cast<uint32_t*>(b + D.dkeys), reinterpret_cast<uint2*>(b + D.rect_by_id));
    return hipGetLastError();
}

hipError_t bin_emit(const BinPlan& p, const Camera& c, int num_isect_host,
                    const void* depth_ws, int* gids_sorted,
                    int* tiles_sorted, int* tile_bins, void* scratch, hipStream_t s) {
    const BinLayout L = layout(p);
    uint8_t* b = static_cast<uint8_t*>(scratch);
    hipError_t e = fill_zero(b, L.zero_end, s);
    if (e != hipSuccess) return e;
    e = fill_zero(tile_bins, 8ull * p.num_tiles, s);
    if (e != hipSuccess) return e;
    if (p.n <= 0) return hipSuccess;
    uint32_t* err = reinterpret_cast<uint32_t*>(b + L.err);
    auto U = [&](size_t off) { return reinterpret_cast<uint32_t*>(b + off); };
    const DepthLayout D = depth_layout(p);
    const uint8_t* d = static_cast<const uint8_t*>(depth_ws);
    const uint32_t* dkeys = reinterpret_cast<const uint32_t*>(d + D.dkeys);
    const uint2* rect_by_id = reinterpret_cast<const uint2*>(d + D.rect_by_id);
    // the sort and the edges cover the slots the emission wrote, min(total,
    // capacity), read on the device: a count given too large (an inconsistent
    // cum_tiles_hit) then never sorts unwritten slots, and one too small is
    // an overflow recorded in the error word
    int* written = reinterpret_cast<int*>(b + L.emit.info);
    // emission in id order; the expansion also counts the tile sort's first
    // digit per sort partition
    const FirstCounts fc = first_counts(p.tile_sort, b + L.tile_sort);
    e = emit<false>(p.n, nullptr, c.tiles_x, nullptr, nullptr, nullptr, rect_by_id,
                    p.isect_cap, b, L.emit, written,
                    reinterpret_cast<unsigned long long*>(b + L.total), err, U(L.tiles),
                    U(L.gids), nullptr, fc, s);
    if (e != hipSuccess) return e;
    stage_mark(kMarkEmitted, s);
    uint32_t* tiles_out = tiles_sorted ? reinterpret_cast<uint32_t*>(tiles_sorted) : U(L.tiles_sorted);
    uint32_t* ids = reinterpret_cast<uint32_t*>(gids_sorted);
    // the order inside a tile comes from the per-tile sort: the first pass
    // need not be stable
    e = sort_pairs(p.tile_sort, U(L.tiles), U(L.gids), tiles_out, ids, U(L.ttmp_k),
                   U(L.ttmp_v), num_isect_host, written, b + L.tile_sort, err, s,
                   fc.counts != nullptr, true);
    if (e != hipSuccess) return e;
    stage_mark(kMarkTileSorted, s);
    const int nb = (int)(((long long)p.isect_cap + 256 * kTileBinsPer - 1) / (256 * kTileBinsPer));
    tile_bins_kernel<<<nb > 0 ? nb : 1, 256, 0, s>>>(tiles_out, num_isect_host, written,
                                                   p.num_tiles,
                                                   reinterpret_cast<int2*>(tile_bins), err);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    stage_mark(kMarkBinned, s);
    // per-tile depth order; the emission's pair buffers and the sort's
    // ping-pong buffers (adjacent) are free again and serve the longer runs
    return tile_depth_order(reinterpret_cast<const int2*>(tile_bins), p.num_tiles, true, dkeys,
                            ids, U(L.tiles), U(L.gids), reinterpret_cast<uint2*>(b + L.ttmp_k),
                            s);
}

// Binning of given tile rectangles with their inclusive counts (the render
// path, whose projection wrote both): the emission reads the rectangles and
// first slots straight from them and the per-tile depth order the depth bits
// straight from `depths`, so neither the depth stage nor the emission's
// count, scan and compaction run.  Same workspace and outputs as bin_emit.
hipError_t bin_rects(const BinPlan& p, const Camera& c, const float* depths,
                     const uint2* rects, const int* cum, int num_isect_host, int* gids_sorted,
                     int* tiles_sorted, int* tile_bins, void* scratch, hipStream_t s) {
    const BinLayout L = layout(p);
    uint8_t* b = static_cast<uint8_t*>(scratch);
    hipError_t e;
    if (p.n <= 0) {
        if ((e = fill_zero(b, L.zero_end, s)) != hipSuccess) return e;
        return fill_zero(tile_bins, 8ull * p.num_tiles, s);
    }
    uint32_t* err = reinterpret_cast<uint32_t*>(b + L.err);
    auto U = [&](size_t off) { return reinterpret_cast<uint32_t*>(b + off); };
    int* written = reinterpret_cast<int*>(b + L.emit.info);
    EmitEntries E{};
    E.chunk_owner = U(L.emit.chunk_owner);
    E.cum = cum;
    E.rect_by_id = rects;
    // also writes the error word and the total, and clears the bins
    const int threads = p.n > p.num_tiles ? p.n : p.num_tiles;
    cum_owner_kernel<<<(threads + 255) / 256, 256, 0, s>>>(
        p.n, cum, rects, p.isect_cap, written,
        reinterpret_cast<unsigned long long*>(b + L.total), err, E.chunk_owner, p.num_tiles,
        reinterpret_cast<int2*>(tile_bins));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const FirstCounts fc = first_counts(p.tile_sort, b + L.tile_sort);
    const long long chunks = ((long long)p.isect_cap + 63) / 64;
    const int waves = (int)(chunks < kExpandWaves ? chunks : kExpandWaves);
    const int blocks = fc.counts ? fc.parts : (waves + 3) / 4;
    emit_expand_kernel<false><<<blocks > 0 ? blocks : 1, 256, 0, s>>>(
        c.tiles_x, E, written, U(L.tiles), U(L.gids), fc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    stage_mark(kMarkEmitted, s);
    uint32_t* tiles_out = tiles_sorted ? reinterpret_cast<uint32_t*>(tiles_sorted) : U(L.tiles_sorted);
    uint32_t* ids = reinterpret_cast<uint32_t*>(gids_sorted);
    // the order inside a tile comes from the per-tile sort: the first pass
    // need not be stable
    e = sort_pairs(p.tile_sort, U(L.tiles), U(L.gids), tiles_out, ids, U(L.ttmp_k),
                   U(L.ttmp_v), num_isect_host, written, b + L.tile_sort, err, s,
                   fc.counts != nullptr, true);
    if (e != hipSuccess) return e;
    stage_mark(kMarkTileSorted, s);
    const int nb = (int)(((long long)p.isect_cap + 256 * kTileBinsPer - 1) / (256 * kTileBinsPer));
    tile_bins_kernel<<<nb > 0 ? nb : 1, 256, 0, s>>>(tiles_out, num_isect_host, written,
                                                   p.num_tiles,
                                                   reinterpret_cast<int2*>(tile_bins), err);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    stage_mark(kMarkBinned, s);
    return tile_depth_order(reinterpret_cast<const int2*>(tile_bins), p.num_tiles, true,
                            reinterpret_cast<const uint32_t*>(depths), ids, U(L.tiles),
                            U(L.gids), reinterpret_cast<uint2*>(b + L.ttmp_k), s);
}

// One workspace: the emission stage's region (error word first), then the
// depth stage's.
hipError_t bin_and_sort(const BinPlan& p, const Camera& c, const float* xys, const float* depths,
                        const int* radii, int num_isect_host,
                        int* gids_sorted, int* tiles_sorted, int* tile_bins, void* scratch,
                        hipStream_t s) {
    uint8_t* depth_ws = static_cast<uint8_t*>(scratch) + p.emit_bytes;
    hipError_t e = bin_depth(p, c, xys, depths, radii, nullptr, depth_ws, s);
    if (e != hipSuccess) return e;
    return bin_emit(p, c, num_isect_host, depth_ws, gids_sorted, tiles_sorted,
                    tile_bins, scratch, s);
}

// ---------------------------------------------------------------------------
// Batched binning (items of many views; see ViewRec)
// ---------------------------------------------------------------------------
// Items are numbered view by view (ViewRec::item_off), so the item order is
// the (view, item) order:
//  1. view_keys: every item's view;
//  2. emission (BATCH) in item order: keys view << tile_bits | tile, values =
//     item ids, with the total and the clamped count published on the device;
//  3. stable sort on the tile bits only (the view bits ride along): order
//     (tile, view, item), so every (view, tile) run is contiguous;
//  4. tile_bins_kernel on the full keys;
//  5. the per-tile depth order of every (view, tile) run (as for one view).
namespace {

struct BatchLayout {
    size_t err, count, total, zero_end;
    size_t tile_sort, vkeys, keys, vals, ttmp_k, ttmp_v, keys_sorted, end;
    EmitLayout emit;
};

BatchLayout batch_layout(const BatchBinPlan& p) {
    BatchLayout L;
    size_t o = 0;
    L.err = o; o += 16;
    L.count = o; o += 16;
    L.total = o; o += 16;
    L.zero_end = o;
    o = align256(o);
    L.tile_sort = o; o = align256(o + p.tile_sort.scratch_bytes);
    o = emit_layout(o, p.n_items, p.isect_cap, true, L.emit);
    const size_t n4 = align256(4ull * (p.n_items > 0 ? p.n_items : 1));
    const size_t i4 = align256(4ull * (p.isect_cap > 0 ? p.isect_cap : 1));
    L.vkeys = o; o += n4;
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

hipError_t batch_bin(const BatchBinPlan& p, const ViewRec* views, const uint2* item_rects,
                     const uint32_t* depth_keys, int* items_sorted, int* tile_bins,
                     uint32_t* item_first, void* scratch, hipStream_t s) {
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

    view_keys_kernel<<<(p.n_items + 255) / 256, 256, 0, s>>>(nullptr, p.n_items, nullptr, views,
                                                              p.num_views, U(L.vkeys));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const FirstCounts fc = first_counts(p.tile_sort, b + L.tile_sort);
    e = emit<true>(p.n_items, nullptr, 0, views, U(L.vkeys), nullptr, item_rects, p.isect_cap,
                   b, L.emit, count, reinterpret_cast<unsigned long long*>(b + L.total), err,
                   U(L.keys), U(L.vals), item_first, fc, s);
    if (e != hipSuccess) return e;
    uint32_t* ids = reinterpret_cast<uint32_t*>(items_sorted);
    e = sort_pairs(p.tile_sort, U(L.keys), U(L.vals), U(L.keys_sorted), ids, U(L.ttmp_k),
                   U(L.ttmp_v), 0, count, b + L.tile_sort, err, s, fc.counts != nullptr);
    if (e != hipSuccess) return e;
    const int nb = (int)(((long long)p.isect_cap + 256 * kTileBinsPer - 1) / (256 * kTileBinsPer));
    tile_bins_kernel<<<nb > 0 ? nb : 1, 256, 0, s>>>(U(L.keys_sorted), 0, count, p.num_bins,
                                                   reinterpret_cast<int2*>(tile_bins), err);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return tile_depth_order(reinterpret_cast<const int2*>(tile_bins), p.num_bins, false,
                            depth_keys, ids, U(L.keys), U(L.vals),
                            reinterpret_cast<uint2*>(b + L.ttmp_k), s);
}

}  // namespace jsa
