// Reduce-then-scan LSD radix sort of (uint32 key, uint32 value) pairs for gfx950.
//
// Replaces the reference's CUB DeviceRadixSort::SortPairs call
// (lib/ops.cu:722-760): a stable least-significant-digit sort over a bit range.
//
// Structure: per 8-bit digit pass, reduce-then-scan in three kernels.
//  * radix_upsweep: each workgroup counts the digits of its partition
//    (1024 or 4096 keys) into per-wave LDS histograms and writes the 256
//    counts digit-major, counts[d * parts + part];
//  * radix_colscan: one workgroup per digit turns its column into exclusive
//    prefixes over the partitions and records the digit total;
//  * radix_downsweep: each workgroup ranks its keys stably -- each wave ranks
//    64 keys per round with 64-lane ballots (one ballot per digit bit gives
//    the mask of lanes holding the same digit; the highest lane of a peer
//    group advances the wave's LDS digit counter) -- and scatters them through
//    LDS so that the global stores of each digit run are contiguous, at
//    digit_start[d] + counts[d * parts + part].
// A decoupled look-back variant (one kernel per pass) was measured slower on
// MI355X: the device-coherent status reads cross the XCDs, so the prefix
// propagates only a few partitions per round trip.
#include "common.hpp"

namespace jsa {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kDigitBits = 8;
constexpr int kBins = 1 << kDigitBits;
constexpr int kMaxPasses = 4;

__device__ __forceinline__ int read_n(int n_host, const int* n_dev) {
    return n_dev ? *n_dev : n_host;
}

// Per-partition digit counts of one pass, stored digit-major:
// counts[d * parts + part].  Per-wave LDS histograms keep same-digit runs
// (frequent in emission order) from piling onto one LDS word.
template <int ITEMS>
__global__ __launch_bounds__(kSortThreads) void radix_upsweep(
    const uint32_t* __restrict__ keys, int n_host, const int* __restrict__ n_dev, int shift,
    int bits, int parts, uint32_t* __restrict__ counts) {
    constexpr int TILE = kSortThreads * ITEMS;
    __shared__ uint32_t s_hist[kSortWaves][kBins];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) s_hist[w][tid] = 0u;
    __syncthreads();
    const int n = read_n(n_host, n_dev);
    const int part = blockIdx.x;
    const int base = part * TILE;
    const uint32_t mask = (1u << bits) - 1u;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int p = base + r * kSortThreads + tid;
        if (p < n) atomicAdd(&s_hist[wave][(keys[p] >> shift) & mask], 1u);
    }
    __syncthreads();
    uint32_t v = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) v += s_hist[w][tid];
    counts[(size_t)tid * parts + part] = v;
}

// In-place exclusive scan of one digit's column counts[d * parts + 0..parts)
// (one workgroup per digit); totals[d] = column sum.  Up to kColPer * 256
// partitions each thread scans a run of consecutive entries and one block
// scan joins the runs (two barriers); longer columns go in 256-entry rounds.
constexpr int kColPer = 16;
__global__ __launch_bounds__(kSortThreads) void radix_colscan(uint32_t* __restrict__ counts,
                                                             int parts,
                                                             uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_tmp[kSortWaves];
    __shared__ uint32_t s_carry;
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    uint32_t* col = counts + (size_t)blockIdx.x * parts;
    const int per = (parts + kSortThreads - 1) / kSortThreads;
    if (per <= kColPer) {
        const int b0 = min(tid * per, parts), b1 = min(b0 + per, parts);
        uint32_t v[kColPer];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kColPer; ++k) {
            v[k] = (k < per && b0 + k < b1) ? col[b0 + k] : 0u;
            sum += v[k];
        }
        const uint32_t incl = wave_inclusive_sum(sum);
        if (lane == 63) s_tmp[wave] = incl;
        __syncthreads();
        uint32_t run = incl - sum, tot = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            run += (w < wave) ? s_tmp[w] : 0u;
            tot += s_tmp[w];
        }
#pragma unroll
        for (int k = 0; k < kColPer; ++k) {
            if (k < per && b0 + k < b1) col[b0 + k] = run;
            run += v[k];
        }
        if (tid == 0) totals[blockIdx.x] = tot;
        return;
    }
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int b = 0; b < parts; b += kSortThreads) {
        const int i = b + tid;
        const uint32_t v = i < parts ? col[i] : 0u;
        const uint32_t incl = wave_inclusive_sum(v);
        if (lane == 63) s_tmp[wave] = incl;
        __syncthreads();
        uint32_t off = s_carry, tot = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            off += (w < wave) ? s_tmp[w] : 0u;
            tot += s_tmp[w];
        }
        if (i < parts) col[i] = off + incl - v;
        __syncthreads();
        if (tid == 0) s_carry += tot;
        __syncthreads();
    }
    if (tid == 0) totals[blockIdx.x] = s_carry;
}

// Ranking + scatter of one partition with known global offsets: stable, or
// (STABLE false) in any order inside a digit, ranked by one returning LDS
// atomic per key instead of a ballot per digit bit (measured 31.5 against
// 41 us for 6.6M keys and 7-bit digits).
template <int ITEMS, bool STABLE>
__global__ __launch_bounds__(kSortThreads) void radix_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int n_host,
    const int* __restrict__ n_dev, int shift, int bits, int parts,
    const uint32_t* __restrict__ counts /* scanned, digit-major */,
    const uint32_t* __restrict__ totals /* per digit */) {
    constexpr int TILE = kSortThreads * ITEMS;
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_whist[kSortWaves][kBins];
    __shared__ uint32_t s_lstart[kBins];
    __shared__ uint32_t s_gbase[kBins];
    __shared__ uint32_t s_tmp[kSortWaves];
    __shared__ uint32_t s_valid;

    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const uint32_t digit_mask = (1u << bits) - 1u;
    const int part = blockIdx.x;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) s_whist[w][tid] = 0u;
    __syncthreads();
    const int n = read_n(n_host, n_dev);
    const int base = part * TILE;
    if (base >= n) return;  // uniform over the block

    // ---- coalesced loads of this wave's ITEMS x 64 consecutive keys ----
    uint32_t key[ITEMS], val[ITEMS];
    uint16_t rank[ITEMS];
    bool keep[ITEMS];
    const uint64_t lt = __lanemask_lt();
    const int wbase = base + wave * (kWave * ITEMS);
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const int p = wbase + r * kWave + lane;
        const bool in = p < n;
        key[r] = in ? keys_in[p] : 0u;
        val[r] = in ? (vals_in ? vals_in[p] : (uint32_t)p) : 0u;
        keep[r] = in;
    }
    // ---- wave-local ranking: stable with 64-lane ballots, else atomics ----
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const bool valid = keep[r];
        const uint32_t d = (key[r] >> shift) & digit_mask;
        if (STABLE) {
            const uint64_t peers = match_digit(d, bits, ballot(valid));
            uint32_t prior = 0;
            if (valid) prior = s_whist[wave][d];
            const bool leader = valid && ((peers & ~(lt | (1ull << lane))) == 0);
            if (leader) s_whist[wave][d] = prior + (uint32_t)__popcll(peers);
            rank[r] = (uint16_t)(prior + (uint32_t)__popcll(peers & lt));
            wave_lds_sync();
        } else {
            rank[r] = valid ? (uint16_t)atomicAdd(&s_whist[wave][d], 1u) : (uint16_t)0;
        }
    }
    __syncthreads();

    // ---- per-digit block totals and wave offsets ----
    const int d = tid;
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t c = s_whist[w][d];
        s_whist[w][d] = total;
        total += c;
    }
    const uint32_t dstart = block_exclusive_scan256(totals[d], s_tmp);
    const uint32_t lstart = block_exclusive_scan256(total, s_tmp);
    s_lstart[d] = lstart;
    s_gbase[d] = dstart + counts[(size_t)d * parts + part];
    if (d == kBins - 1) s_valid = lstart + total;      // keys kept in this partition
    __syncthreads();

    // ---- scatter into LDS in digit order ----
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        if (keep[r]) {
            const uint32_t dd = (key[r] >> shift) & digit_mask;
            const uint32_t lp = s_lstart[dd] + s_whist[wave][dd] + rank[r];
            s_keys[lp] = key[r];
            s_vals[lp] = val[r];
        }
    }
    __syncthreads();

    // ---- runs of equal digits go to contiguous global positions ----
    const int cnt = (int)s_valid;
    for (int i = tid; i < cnt; i += kSortThreads) {
        const uint32_t k = s_keys[i];
        const uint32_t dd = (k >> shift) & digit_mask;
        const uint32_t g = s_gbase[dd] + (uint32_t)i - s_lstart[dd];
        keys_out[g] = k;
        vals_out[g] = s_vals[i];
    }
}

}  // namespace

// Partition size: 16 keys per lane for large arrays (the tile sort of millions
// of pairs), 8 from 512K keys (measured faster than 4 or 16 on a depth sort
// of 1M keys), 4 for small ones (more workgroups in flight).
static int items_for(int n_cap) { return n_cap >= (1 << 22) ? 16 : (n_cap >= (1 << 19) ? 8 : 4); }

SortPlan sort_plan(int n_cap, int begin_bit, int end_bit) {
    SortPlan p;
    p.n_cap = n_cap > 0 ? n_cap : 1;
    p.begin_bit = begin_bit;
    p.end_bit = end_bit > begin_bit ? end_bit : begin_bit;
    p.passes = (p.end_bit - p.begin_bit + kDigitBits - 1) / kDigitBits;
    const int tile = kSortThreads * items_for(p.n_cap);
    p.tile = tile;
    p.parts = (p.n_cap + tile - 1) / tile;
    // the key bits split evenly over the passes (see sort_pairs)
    const int per = p.passes > 0 ? (p.end_bit - p.begin_bit + p.passes - 1) / p.passes : 0;
    p.first_shift = p.begin_bit;
    p.first_bits = min(per, p.end_bit - p.begin_bit);
    // [totals 256][counts 256 * parts]
    p.scratch_bytes = 4ull * kBins + 4ull * kBins * (size_t)p.parts;
    p.scratch_bytes = (p.scratch_bytes + 255) & ~size_t(255);
    return p;
}

size_t sort_zero_bytes(const SortPlan& p) { return 0; }

uint32_t* sort_first_counts(const SortPlan& p, void* scratch) {
    return static_cast<uint32_t*>(scratch) + kBins;
}

hipError_t sort_pairs(const SortPlan& p, const uint32_t* keys_in, const uint32_t* vals_in,
                      uint32_t* keys_out, uint32_t* vals_out, uint32_t* tmp_keys,
                      uint32_t* tmp_vals, int n_host, const int* n_dev, void* scratch,
                      uint32_t* err, hipStream_t s, bool counted, bool unstable_first) {
    (void)err;
    if (p.passes <= 0 || p.passes > kMaxPasses) return hipErrorInvalidValue;
    uint32_t* totals = static_cast<uint32_t*>(scratch);
    uint32_t* counts = totals + kBins;
    const int items = items_for(p.n_cap);
    // the key bits split evenly over the passes (13 bits: 7 + 6), so that no
    // pass scatters into more digits than needed: longer runs per digit and
    // partition, fewer partly written lines
    const int per = (p.end_bit - p.begin_bit + p.passes - 1) / p.passes;
    hipError_t e;
    // ping-pong so that the last pass lands in (keys_out, vals_out)
    const uint32_t* ki = keys_in;
    const uint32_t* vi = vals_in;
    for (int q = 0; q < p.passes; ++q) {
        const bool to_out = ((p.passes - 1 - q) % 2) == 0;
        uint32_t* ko = to_out ? keys_out : tmp_keys;
        uint32_t* vo = to_out ? vals_out : tmp_vals;
        const int shift = p.begin_bit + q * per;
        const int bits = min(per, p.end_bit - shift);
        // the first pass's counts may come from the keys' producer
        const bool count_here = !(counted && q == 0);
        // an LSD sort needs every pass after the first to be stable; the
        // first one only where the caller needs equal keys in input order
        const bool stable = !(unstable_first && q == 0);
#define JSA_PASS(IT)                                                                         \
    do {                                                                                     \
        if (count_here) {                                                                    \
            radix_upsweep<IT><<<p.parts, kSortThreads, 0, s>>>(ki, n_host, n_dev, shift,     \
                                                                bits, p.parts, counts);      \
            if ((e = hipGetLastError()) != hipSuccess) return e;                             \
        }                                                                                    \
        radix_colscan<<<kBins, kSortThreads, 0, s>>>(counts, p.parts, totals);               \
        if ((e = hipGetLastError()) != hipSuccess) return e;                                 \
        if (stable)                                                                          \
            radix_downsweep<IT, true><<<p.parts, kSortThreads, 0, s>>>(                      \
                ki, vi, ko, vo, n_host, n_dev, shift, bits, p.parts, counts, totals);        \
        else                                                                                 \
            radix_downsweep<IT, false><<<p.parts, kSortThreads, 0, s>>>(                     \
                ki, vi, ko, vo, n_host, n_dev, shift, bits, p.parts, counts, totals);        \
        if ((e = hipGetLastError()) != hipSuccess) return e;                                 \
    } while (0)
        if (items == 16) JSA_PASS(16);
        else if (items == 8) JSA_PASS(8);
        else JSA_PASS(4);
#undef JSA_PASS
        ki = ko;
        vi = vo;
    }
    return hipSuccess;
}

}  // namespace jsa
