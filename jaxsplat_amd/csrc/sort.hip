// Onesweep LSD radix sort of (uint32 key, uint32 value) pairs for gfx950.
//
// Replaces the reference's CUB DeviceRadixSort::SortPairs call
// (lib/ops.cu:722-760): a stable least-significant-digit sort over a bit range.
//
// Structure (one histogram kernel + one kernel per 8-bit digit pass):
//  * radix_histogram: one read of the keys builds the digit histograms of all
//    passes at once (per-wave LDS histograms, then global atomics);
//  * onesweep_pass: a workgroup takes a 4096-key partition from an atomic
//    counter (partitions start in order, so the look-back below cannot wait
//    on a partition that has not started).  Each wave ranks 64 keys per round
//    with 64-lane ballots (one ballot per digit bit gives the mask of lanes
//    holding the same digit; the highest lane of a peer group advances the
//    wave's LDS digit counter), which keeps the ranking stable.  The block
//    publishes its 256 digit counts, performs a per-digit decoupled look-back
//    (one lane per digit) and scatters the keys through LDS so that the global
//    stores of each digit run are contiguous.
//  * Every wait in the look-back is bounded; on timeout an error word is set
//    and the kernel still terminates.
#include "common.hpp"

namespace jsa {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kItems = 16;
constexpr int kTile = kSortThreads * kItems;   // 4096 keys per partition
constexpr int kDigitBits = 8;
constexpr int kBins = 1 << kDigitBits;
constexpr int kMaxPasses = 4;

// 32-bit look-back status: bits 31..30 flag (1 aggregate, 2 prefix), 29..0 count.
constexpr uint32_t kSFlagAgg = 1u << 30;
constexpr uint32_t kSFlagPre = 2u << 30;
constexpr uint32_t kSValMask = (1u << 30) - 1;

__device__ __forceinline__ int read_n(int n_host, const int* n_dev) {
    return n_dev ? *n_dev : n_host;
}

__global__ __launch_bounds__(kSortThreads) void radix_histogram(
    const uint32_t* __restrict__ keys, int n_host, const int* __restrict__ n_dev,
    int begin_bit, int passes, uint32_t* __restrict__ hist /* [passes][256] */) {
    __shared__ uint32_t s_hist[kSortWaves][kMaxPasses][kBins];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    for (int i = tid; i < kSortWaves * kMaxPasses * kBins; i += kSortThreads)
        (&s_hist[0][0][0])[i] = 0u;
    __syncthreads();
    const int n = read_n(n_host, n_dev);
    for (int i = blockIdx.x * kSortThreads + tid; i < n; i += gridDim.x * kSortThreads) {
        const uint32_t k = keys[i];
        for (int p = 0; p < passes; ++p) {
            const uint32_t d = (k >> (begin_bit + p * kDigitBits)) & (kBins - 1);
            atomicAdd(&s_hist[wave][p][d], 1u);
        }
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
        uint32_t v = 0;
        for (int w = 0; w < kSortWaves; ++w) v += s_hist[w][p][tid];
        if (v) atomicAdd(&hist[p * kBins + tid], v);
    }
}

// Exclusive scan of one value per thread over the 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan256(uint32_t v, uint32_t* s_tmp) {
    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    const uint32_t incl = wave_inclusive_sum(v);
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    uint32_t off = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) off += (w < wave) ? s_tmp[w] : 0u;
    __syncthreads();
    return off + incl - v;
}

__global__ __launch_bounds__(kSortThreads) void onesweep_pass(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int n_host,
    const int* __restrict__ n_dev, int shift, uint32_t digit_mask,
    const uint32_t* __restrict__ hist /* 256 bins of this pass */,
    uint32_t* __restrict__ status /* [parts][256] */, uint32_t* __restrict__ counter,
    uint32_t* __restrict__ err) {
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t s_whist[kSortWaves][kBins];
    __shared__ uint32_t s_lstart[kBins];
    __shared__ uint32_t s_gbase[kBins];
    __shared__ uint32_t s_tmp[kSortWaves];
    __shared__ uint32_t s_part;

    const int tid = threadIdx.x;
    const int lane = lane_id();
    const int wave = tid >> 6;
    if (tid == 0) s_part = atomicAdd(counter, 1u);
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) s_whist[w][tid] = 0u;
    __syncthreads();
    const int n = read_n(n_host, n_dev);
    const int part = (int)s_part;
    const int base = part * kTile;
    if (base >= n) return;  // uniform over the block

    // ---- stable wave-local ranking with 64-lane ballots ----
    uint32_t key[kItems], val[kItems];
    uint16_t rank[kItems];
    const uint64_t lt = __lanemask_lt();
    const int wbase = base + wave * (kWave * kItems);
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const int p = wbase + r * kWave + lane;
        const bool valid = p < n;
        const uint32_t k = valid ? keys_in[p] : 0u;
        const uint32_t v = valid ? (vals_in ? vals_in[p] : (uint32_t)p) : 0u;
        const uint32_t d = (k >> shift) & digit_mask;
        uint64_t peers = ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        uint32_t prior = 0;
        if (valid) prior = s_whist[wave][d];
        const bool leader = valid && ((peers & ~(lt | (1ull << lane))) == 0);
        if (leader) s_whist[wave][d] = prior + (uint32_t)__popcll(peers);
        rank[r] = (uint16_t)(prior + (uint32_t)__popcll(peers & lt));
        key[r] = k;
        val[r] = v;
    }
    __syncthreads();

    // ---- per-digit block totals, wave offsets, publication ----
    const int d = tid;
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t c = s_whist[w][d];
        s_whist[w][d] = total;
        total += c;
    }
    uint32_t* my_status = status + (size_t)part * kBins + d;
    store_agent(my_status, (part == 0 ? kSFlagPre : kSFlagAgg) | total);

    // global start of this digit for the pass (exclusive scan of the histogram)
    const uint32_t gstart = block_exclusive_scan256(hist[d], s_tmp);
    const uint32_t lstart = block_exclusive_scan256(total, s_tmp);

    // ---- decoupled look-back, one lane per digit ----
    uint32_t excl = 0;
    if (part > 0) {
        int j = part - 1;
        uint32_t spins = 0;
        while (j >= 0) {
            const uint32_t s = load_agent(status + (size_t)j * kBins + d);
            const uint32_t flag = s & ~kSValMask;
            if (flag == 0u) {
                if (++spins > kSpinLimit) {
                    atomicOr(err, 2u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += s & kSValMask;
            if (flag == kSFlagPre) break;
            --j;
        }
        store_agent(my_status, kSFlagPre | (excl + total));
    }
    s_lstart[d] = lstart;
    s_gbase[d] = gstart + excl;
    __syncthreads();

    // ---- scatter into LDS in digit order ----
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const int p = wbase + r * kWave + lane;
        if (p < n) {
            const uint32_t dd = (key[r] >> shift) & digit_mask;
            const uint32_t lp = s_lstart[dd] + s_whist[wave][dd] + rank[r];
            s_keys[lp] = key[r];
            s_vals[lp] = val[r];
        }
    }
    __syncthreads();

    // ---- coalesced-run global stores ----
    const int cnt = min(kTile, n - base);
    for (int i = tid; i < cnt; i += kSortThreads) {
        const uint32_t k = s_keys[i];
        const uint32_t dd = (k >> shift) & digit_mask;
        const uint32_t g = s_gbase[dd] + (uint32_t)i - s_lstart[dd];
        keys_out[g] = k;
        vals_out[g] = s_vals[i];
    }
}

}  // namespace

SortPlan sort_plan(int n_cap, int begin_bit, int end_bit) {
    SortPlan p;
    p.n_cap = n_cap > 0 ? n_cap : 1;
    p.begin_bit = begin_bit;
    p.end_bit = end_bit > begin_bit ? end_bit : begin_bit;
    p.passes = (p.end_bit - p.begin_bit + kDigitBits - 1) / kDigitBits;
    p.parts = (p.n_cap + kTile - 1) / kTile;
    // [err u32, pad][hist passes*256][counters passes][status passes*parts*256]
    p.scratch_bytes = 16 + 4ull * kBins * kMaxPasses + 4ull * kMaxPasses +
                      4ull * (size_t)kMaxPasses * p.parts * kBins;
    p.scratch_bytes = (p.scratch_bytes + 255) & ~size_t(255);
    return p;
}

size_t sort_zero_bytes(const SortPlan& p) { return p.scratch_bytes; }

hipError_t sort_pairs(const SortPlan& p, const uint32_t* keys_in, const uint32_t* vals_in,
                      uint32_t* keys_out, uint32_t* vals_out, uint32_t* tmp_keys,
                      uint32_t* tmp_vals, int n_host, const int* n_dev, void* scratch,
                      uint32_t* err, hipStream_t s) {
    if (p.passes <= 0 || p.passes > kMaxPasses) return hipErrorInvalidValue;
    uint8_t* b = static_cast<uint8_t*>(scratch);
    uint32_t* hist = reinterpret_cast<uint32_t*>(b + 16);
    uint32_t* counters = hist + kBins * kMaxPasses;
    uint32_t* status = counters + kMaxPasses;
    const int hist_blocks = p.parts < 1024 ? (p.parts > 0 ? p.parts : 1) : 1024;
    radix_histogram<<<hist_blocks, kSortThreads, 0, s>>>(keys_in, n_host, n_dev, p.begin_bit,
                                                        p.passes, hist);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // ping-pong so that the last pass lands in (keys_out, vals_out)
    const uint32_t* ki = keys_in;
    const uint32_t* vi = vals_in;
    for (int q = 0; q < p.passes; ++q) {
        const bool to_out = ((p.passes - 1 - q) % 2) == 0;
        uint32_t* ko = to_out ? keys_out : tmp_keys;
        uint32_t* vo = to_out ? vals_out : tmp_vals;
        const int shift = p.begin_bit + q * kDigitBits;
        const int bits = min(kDigitBits, p.end_bit - shift);
        const uint32_t mask = (1u << bits) - 1u;
        onesweep_pass<<<p.parts, kSortThreads, 0, s>>>(
            ki, vi, ko, vo, n_host, n_dev, shift, mask, hist + q * kBins,
            status + (size_t)q * p.parts * kBins, counters + q, err);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        ki = ko;
        vi = vo;
    }
    return hipSuccess;
}

}  // namespace jsa
