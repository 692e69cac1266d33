// Device-wide exclusive scan and LSD radix sort for the front end.
//
// Element counts may live on the device (the producer kernel's append
// counter): kernels are launched for a host-known capacity and every tile
// checks the device count, so a pipeline never waits on the host between
// stages.  The radix sort is stable; a tile is 256 threads x 8 keys staged
// through LDS, ranked per 8-bit digit with wavefront ballots.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcgpu {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16; // elements per thread in the scan tiles
constexpr int kScanTile = kScanBlock * kScanItems;

constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

__device__ __forceinline__ uint64_t read_count(const uint64_t* n_dev, uint64_t n_host)
{
    return n_dev ? *n_dev : n_host;
}

// block-wide exclusive scan of one value per thread (256 threads); returns the total in *total
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wsum[kScanBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
        uint32_t s = wsum[w];
        if (w < wid) off += s;
        tot += s;
    }
    __syncthreads();
    if (total) *total = tot;
    return off + x - v;
}

// pass 1: per-tile sums
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const uint32_t* __restrict__ in, uint32_t* __restrict__ tile_sums,
                                                            uint64_t n)
{
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int i = 0; i < kScanItems; ++i) {
        uint64_t idx = base + (uint64_t)i * kScanBlock + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// pass 2: exclusive scan of the tile sums by one block (any count), total into *total_out
__global__ void __launch_bounds__(kScanBlock) k_scan_tiles(uint32_t* __restrict__ tile_sums, uint32_t ntiles,
                                                           uint32_t* __restrict__ total_out)
{
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += kScanBlock) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < ntiles ? tile_sums[i] : 0;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(v, &tot);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// pass 3: scan inside each tile, offset by the tile's prefix.  The tile goes
// through LDS so that global loads and stores are coalesced; each thread scans
// kScanItems consecutive elements there (stride kScanItems + 1: no bank
// conflicts).
__global__ void __launch_bounds__(kScanBlock) k_scan_apply(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                           const uint32_t* __restrict__ tile_sums, uint64_t n)
{
    __shared__ uint32_t t[kScanBlock * (kScanItems + 1)];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    for (int i = 0; i < kScanItems; ++i) {
        const uint32_t k = (uint32_t)i * kScanBlock + threadIdx.x;
        t[(k / kScanItems) * (kScanItems + 1) + k % kScanItems] = base + k < n ? in[base + k] : 0u;
    }
    __syncthreads();
    uint32_t* mine = t + threadIdx.x * (kScanItems + 1);
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = mine[i];
        s += v[i];
    }
    uint32_t run = block_exclusive_scan(s, nullptr) + tile_sums[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i) {
        mine[i] = run;
        run += v[i];
    }
    __syncthreads();
    for (int i = 0; i < kScanItems; ++i) {
        const uint32_t k = (uint32_t)i * kScanBlock + threadIdx.x;
        if (base + k < n) out[base + k] = t[(k / kScanItems) * (kScanItems + 1) + k % kScanItems];
    }
}

// One-block exclusive scan for radix histograms.  The histogram holds
// kRadix x (tiles in use) entries, where the tile count follows the element
// count (device or host); 1024 threads walk it in coalesced chunks of 4096
// with a carried prefix.  In place.
constexpr int kScan1Block = 1024;
constexpr int kScan1Items = 4;

__global__ void __launch_bounds__(kScan1Block) k_scan_hist(uint32_t* __restrict__ data, const uint64_t* n_dev,
                                                           uint64_t n_host, uint32_t tile)
{
    __shared__ uint32_t wsum[kScan1Block / 64];
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t entries = (uint64_t)kRadix * ((n + tile - 1) / tile);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint64_t base = 0; base < entries; base += (uint64_t)kScan1Block * kScan1Items) {
        const uint64_t b = base + (uint64_t)threadIdx.x * kScan1Items;
        uint32_t v[kScan1Items];
        uint32_t tot = 0;
        for (int q = 0; q < kScan1Items; ++q) {
            v[q] = (b + q < entries) ? data[b + q] : 0u;
            tot += v[q];
        }
        uint32_t x = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        uint32_t off = 0, all = 0;
        for (int w = 0; w < kScan1Block / 64; ++w) {
            const uint32_t t = wsum[w];
            if (w < wid) off += t;
            all += t;
        }
        uint32_t run = carry + off + x - tot;
        for (int q = 0; q < kScan1Items; ++q) {
            if (b + q < entries) data[b + q] = run;
            run += v[q];
        }
        carry += all;
        __syncthreads();
    }
}

// Host launcher: exclusive scan of n (host-known) u32 into out (may alias in);
// tmp needs ceil(n / kScanTile) entries; total (device) optional.
inline void scan_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, uint32_t* total,
                           hipStream_t s)
{
    uint32_t ntiles = (uint32_t)((n + kScanTile - 1) / kScanTile);
    if (ntiles == 0) ntiles = 1;
    k_scan_reduce<<<ntiles, kScanBlock, 0, s>>>(in, tmp, n);
    k_scan_tiles<<<1, kScanBlock, 0, s>>>(tmp, ntiles, total);
    k_scan_apply<<<ntiles, kScanBlock, 0, s>>>(in, out, tmp, n);
}

constexpr uint64_t scan_tmp_elems(uint64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

// ---------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift)
{
    return (uint32_t)(key >> shift) & (kRadix - 1);
}

// histogram of one digit per tile, stored digit-major: hist[d * ntiles + t].
// With ghist, each tile's counts are also added into the totals of its group
// of G consecutive tiles: ghist[d * ngc + t / G].
template <typename K>
__global__ void __launch_bounds__(kSortBlock) k_radix_hist(const K* __restrict__ keys, const uint64_t* n_dev,
                                                           uint64_t n_host, int shift, uint32_t* __restrict__ hist,
                                                           uint32_t ntiles, uint32_t* __restrict__ ghist, uint32_t G,
                                                           uint32_t ngc)
{
    __shared__ uint32_t h[kRadix];
    const uint64_t n = read_count(n_dev, n_host);
    const uint32_t nt = (uint32_t)((n + kSortTile - 1) / kSortTile); // tiles in use (<= ntiles)
    (void)ntiles;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        h[threadIdx.x] = 0;
        __syncthreads();
        const uint64_t base = (uint64_t)t * kSortTile;
        if (base < n) {
            for (int j = 0; j < kSortItems; ++j) {
                uint64_t idx = base + (uint64_t)j * kSortBlock + threadIdx.x;
                if (idx < n) atomicAdd(&h[digit_of(keys[idx], shift)], 1u);
            }
        }
        __syncthreads();
        const uint32_t c = h[threadIdx.x];
        hist[(uint64_t)threadIdx.x * nt + t] = c;
        if (ghist && c) atomicAdd(&ghist[(uint64_t)threadIdx.x * ngc + t / G], c);
        __syncthreads();
    }
}

// stable scatter by digit; offs = exclusive scan of hist (digit-major), or
// with group totals the raw tile histograms.  A tile is ranked digit by digit
// as before, but its keys are first placed in tile order by digit in LDS and
// then written out in that order, so the lanes of a wave write runs of
// consecutive addresses (one run per digit of the tile) instead of one
// scattered element each.
template <typename K, bool HasVal>
__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                                                              K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                              const uint64_t* n_dev, uint64_t n_host, int shift,
                                                              const uint32_t* __restrict__ offs, uint32_t ntiles,
                                                              const uint32_t* __restrict__ ghist, uint32_t G, uint32_t ngc)
{
    __shared__ uint32_t wcount[kSortBlock / 64][kRadix];
    __shared__ uint32_t running[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t lstart[kRadix]; // first tile-local slot of each digit
    __shared__ K lkey[kSortTile];
    __shared__ uint32_t lval[HasVal ? kSortTile : 1];
    const uint64_t n = read_count(n_dev, n_host);
    const uint32_t nt = (uint32_t)((n + kSortTile - 1) / kSortTile);
    (void)ntiles;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // with group totals, offs holds the raw tile histograms: this thread's
    // digit starts after every smaller digit (dstart), then after its own
    // digit in the groups and in the tiles of its group before the tile
    uint32_t dstart = 0;
    const uint32_t ng = ghist ? (nt + G - 1) / G : 0;
    if (ghist) {
        const uint32_t* gd = ghist + (uint64_t)threadIdx.x * ngc;
        uint32_t tot = 0;
        for (uint32_t q = 0; q < ng; ++q) tot += gd[q];
        dstart = block_exclusive_scan(tot, nullptr);
    }
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t base = (uint64_t)t * kSortTile;
        if (base >= n) break; // tiles are visited in increasing order
        const uint32_t cnt = (uint32_t)((n - base) < (uint64_t)kSortTile ? (n - base) : kSortTile);
        running[threadIdx.x] = 0;
        lstart[threadIdx.x] = 0;
        if (ghist) {
            const uint32_t g = t / G;
            const uint32_t* gd = ghist + (uint64_t)threadIdx.x * ngc;
            const uint32_t* hd = offs + (uint64_t)threadIdx.x * nt;
            uint32_t off = dstart;
            for (uint32_t q = 0; q < g; ++q) off += gd[q];
            for (uint32_t q = g * G; q < t; ++q) off += hd[q];
            gbase[threadIdx.x] = off;
        } else {
            gbase[threadIdx.x] = offs[(uint64_t)threadIdx.x * nt + t];
        }
        for (int w = 0; w < kSortBlock / 64; ++w) wcount[w][threadIdx.x] = 0;
        K key[kSortItems];
        uint32_t val[kSortItems];
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const uint32_t li = (uint32_t)j * kSortBlock + threadIdx.x;
            key[j] = li < cnt ? keys_in[base + li] : (K)0;
            val[j] = (HasVal && li < cnt) ? vals_in[base + li] : 0u;
        }
        __syncthreads();
        // the tile's digit counts, then their exclusive scan
#pragma unroll
        for (int j = 0; j < kSortItems; ++j)
            if ((uint32_t)j * kSortBlock + threadIdx.x < cnt) atomicAdd(&lstart[digit_of(key[j], shift)], 1u);
        __syncthreads();
        {
            const uint32_t c = lstart[threadIdx.x];
            __syncthreads();
            lstart[threadIdx.x] = block_exclusive_scan(c, nullptr);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) {
            const bool valid = (uint32_t)j * kSortBlock + threadIdx.x < cnt;
            const uint32_t d = digit_of(key[j], shift);
            // lanes of this wave holding the same digit
            uint64_t peers = __ballot(valid);
            for (int b = 0; b < kRadixBits; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            const uint32_t rank = __popcll(peers & lt_mask);
            if (valid && rank == 0) wcount[wid][d] = __popcll(peers);
            __syncthreads();
            if (valid) {
                uint32_t pos = running[d] + rank;
                for (int w = 0; w < wid; ++w) pos += wcount[w][d];
                lkey[lstart[d] + pos] = key[j];
                if (HasVal) lval[lstart[d] + pos] = val[j];
            }
            __syncthreads();
            {
                uint32_t add = 0;
                for (int w = 0; w < kSortBlock / 64; ++w) {
                    add += wcount[w][threadIdx.x];
                    wcount[w][threadIdx.x] = 0;
                }
                running[threadIdx.x] += add;
            }
            __syncthreads();
        }
        // the tile in digit order: consecutive slots of one digit go to consecutive addresses
        for (uint32_t i = threadIdx.x; i < cnt; i += kSortBlock) {
            const K k = lkey[i];
            const uint32_t d = digit_of(k, shift);
            const uint64_t dst = (uint64_t)gbase[d] + (i - lstart[d]);
            keys_out[dst] = k;
            if (HasVal) vals_out[dst] = lval[i];
        }
        __syncthreads();
    }
}

struct RadixTemp {
    uint32_t* hist;      // radix_hist_elems(cap): tile histograms, then group totals
    uint32_t* scan_tmp;  // scan_tmp_elems(kRadix * ntiles)
};

// Tile counts up to this use group totals in place of a scan between the
// histogram and scatter passes: two launches per digit instead of five.
// Every scatter block reads about 2 sqrt(tiles) counts per digit, so the
// group path stops where that outgrows the scan.
constexpr uint32_t kGroupedMaxTiles = 1024;
constexpr int kMaxPasses = 8;

constexpr uint32_t radix_tiles(uint64_t cap) { return (uint32_t)((cap + kSortTile - 1) / kSortTile) + 1; }
constexpr uint32_t radix_group(uint32_t ntiles)
{
    uint32_t g = 1;
    while (g * g < ntiles) ++g;
    return g;
}
constexpr uint32_t radix_groups(uint32_t ntiles) { return (ntiles + radix_group(ntiles) - 1) / radix_group(ntiles); }
constexpr uint64_t radix_hist_elems(uint64_t cap)
{
    const uint32_t nt = radix_tiles(cap);
    return (uint64_t)kRadix * nt + (uint64_t)kRadix * kMaxPasses * radix_groups(nt);
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0;
}

// Sorts keys[0..n) (n on device or host) by bits [0, bits) with an optional
// u32 payload.  Ping-pongs between (keys, vals) and (keys_alt, vals_alt);
// returns true when the sorted result ends in the *_alt buffers.
// The group totals a sort of (cap, bits) accumulates into, which must be zero
// when it starts: their start and count (0 when the sort takes the scan path).
// A producer kernel on the same stream may clear them and spare the sort its
// clearing launch (radix_sort with ghist_zeroed).
inline uint32_t* radix_ghist(const RadixTemp& tmp, uint64_t cap, int bits, uint64_t* count)
{
    const uint32_t ntiles = radix_tiles(cap);
    const int passes = (bits + kRadixBits - 1) / kRadixBits;
    const bool grouped = ntiles <= kGroupedMaxTiles && passes <= kMaxPasses;
    *count = grouped ? (uint64_t)kRadix * radix_groups(ntiles) * passes : 0;
    return tmp.hist + (uint64_t)kRadix * ntiles;
}

template <typename K>
inline bool radix_sort(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, const uint64_t* n_dev,
                       uint64_t n_host, uint64_t cap, int bits, const RadixTemp& tmp, hipStream_t s,
                       bool ghist_zeroed = false)
{
    const uint32_t ntiles = radix_tiles(cap);
    const uint32_t grid = ntiles < 2048 ? ntiles : 2048;
    const int passes = (bits + kRadixBits - 1) / kRadixBits;
    const bool grouped = ntiles <= kGroupedMaxTiles && passes <= kMaxPasses;
    const uint32_t G = radix_group(ntiles), ngc = radix_groups(ntiles);
    uint32_t* ghist = tmp.hist + (uint64_t)kRadix * ntiles;
    if (grouped && !ghist_zeroed) {
        const uint64_t nz = (uint64_t)kRadix * ngc * passes;
        k_zero_u32<<<(unsigned)((nz + 255) / 256 < 1024 ? (nz + 255) / 256 : 1024), 256, 0, s>>>(ghist, nz);
    }
    bool alt = false;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kRadixBits;
        K* kin = alt ? keys_alt : keys;
        K* kout = alt ? keys : keys_alt;
        uint32_t* vin = alt ? vals_alt : vals;
        uint32_t* vout = alt ? vals : vals_alt;
        uint32_t* gh = grouped ? ghist + (uint64_t)kRadix * ngc * p : nullptr;
        k_radix_hist<K><<<grid, kSortBlock, 0, s>>>(kin, n_dev, n_host, shift, tmp.hist, ntiles, gh, G, ngc);
        if (!grouped) {
            const uint64_t nh_host = (uint64_t)kRadix * ((n_host + kSortTile - 1) / kSortTile);
            if (!n_dev && nh_host > 16384)
                scan_exclusive(tmp.hist, tmp.hist, nh_host, tmp.scan_tmp, nullptr, s);
            else
                k_scan_hist<<<1, kScan1Block, 0, s>>>(tmp.hist, n_dev, n_host, kSortTile);
        }
        if (vals)
            k_radix_scatter<K, true><<<grid, kSortBlock, 0, s>>>(kin, vin, kout, vout, n_dev, n_host, shift, tmp.hist, ntiles,
                                                                 gh, G, ngc);
        else
            k_radix_scatter<K, false><<<grid, kSortBlock, 0, s>>>(kin, nullptr, kout, nullptr, n_dev, n_host, shift, tmp.hist,
                                                                  ntiles, gh, G, ngc);
        alt = !alt;
    }
    return alt;
}

} // namespace mcgpu
