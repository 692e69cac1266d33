// Front-end kernels of the MI355X MCUT intersection front end.
//
// Every kernel that consumes a variable-length list reads the length from the
// device counter its producer wrote and walks it with a grid-stride loop, so a
// dispatch is enqueued end to end without host synchronisation.
#pragma once
#include "kernels.hpp"
#include "sort.hpp"

namespace mcgpu {

// ===========================================================================
// 1. shift parameters (calculate_vertex_parameters, preproc.cpp:2124-2290)
// ===========================================================================

// One 64-lane wavefront per (mesh, coordinate): six independent chains.  Each
// sum must be the reference's left-to-right sum (any other association
// changes the centre's bits and so every shifted coordinate), so the chain is
// inherently serial: one dependent FP64 add per vertex.  The kernel makes each
// step a single instruction:
//  - a 64-value block of one coordinate is held in four registers X[0..3] so
//    that lane c of every 16-lane row holds values 16 q + c in X[q];
//  - step k of row q is v_fmac_f64 with DPP row_newbcast:k, which broadcasts
//    lane k of the row as the first operand: acc = fma(X[q] of lane k, 1.0,
//    acc).  The product x * 1.0 is exact, so the fused result is the single
//    rounding of x + acc: bit for bit the add the reference performs;
//  - blocks arrive through a ring of kSumRing blocks of registers, loaded
//    about 4 us ahead of use with a uniform block base and fixed lane offsets
//    (no per-lane address arithmetic in the chain);
//  - the last n mod 64 values are padded with +0.0 loaded from memory: the
//    sum starts at +0.0 and round-to-nearest never yields -0.0 from it, so
//    adding +0.0 changes no bit.
// A DPP operand needs two wait states after a VALU write of its register;
// every DPP operand here is written only by a load, never by a VALU
// instruction, so the chain needs none.  The accumulator is read as the
// fused add's third operand, not through DPP, and needs none either.  The
// sixteen steps of a row are therefore one asm statement: the compiler
// treats each asm statement as a possible DPP reader of the registers the
// previous one wrote and would put a wait state between single steps, which
// costs a third of the chain's speed.  The wave raises its issue priority,
// and the host runs it on a stream whose CU mask keeps other kernels off its
// CUs.  Measured: 2.44 ns per step (one step per asm statement: 3.8 ns; LDS
// broadcast reads: 5.4 ns), bit for bit the host's left-to-right sum.
#define MCGPU_SUM_DPP(k) "v_fmac_f64_dpp %0, %1, %2 row_newbcast:" #k " row_mask:0xf bank_mask:0xf\n\t"
#define MCGPU_SUM_ROW(acc, x, one)                                                                                     \
    asm volatile(MCGPU_SUM_DPP(0) MCGPU_SUM_DPP(1) MCGPU_SUM_DPP(2) MCGPU_SUM_DPP(3) MCGPU_SUM_DPP(4)                 \
                 MCGPU_SUM_DPP(5) MCGPU_SUM_DPP(6) MCGPU_SUM_DPP(7) MCGPU_SUM_DPP(8) MCGPU_SUM_DPP(9)                  \
                 MCGPU_SUM_DPP(10) MCGPU_SUM_DPP(11) MCGPU_SUM_DPP(12) MCGPU_SUM_DPP(13) MCGPU_SUM_DPP(14)             \
                 MCGPU_SUM_DPP(15)                                                                                     \
                 : "+v"(acc)                                                                                           \
                 : "v"(x), "v"(one))

constexpr int kSumRing = 16;

// left-to-right sum of v[0], v[3], ..., v[3 (n - 1)] by one wavefront (the
// whole wave returns it)
__device__ __forceinline__ double serial_sum(const double* __restrict__ v, uint32_t n, const double* __restrict__ zero)
{
    const uint32_t c = threadIdx.x & 15;
    const uint32_t nfull = n / 64;
    const uint32_t last = nfull ? nfull - 1 : 0;
    double one = 1.0, acc = 0.0;
    if (nfull > 0) {
        double X[kSumRing][4];
#pragma unroll
        for (int r = 0; r < kSumRing; ++r) {
            const double* vb = v + 192ull * ((uint32_t)r < last ? (uint32_t)r : last); // uniform block base
#pragma unroll
            for (int q = 0; q < 4; ++q) X[r][q] = vb[3 * (16 * q + c)];
        }
        for (uint32_t b0 = 0; b0 < nfull; b0 += kSumRing) {
#pragma unroll
            for (int r = 0; r < kSumRing; ++r) {
                if (b0 + r < nfull) {
                    MCGPU_SUM_ROW(acc, X[r][0], one);
                    MCGPU_SUM_ROW(acc, X[r][1], one);
                    MCGPU_SUM_ROW(acc, X[r][2], one);
                    MCGPU_SUM_ROW(acc, X[r][3], one);
                }
                // refill this ring slot kSumRing blocks ahead (clamped to the last block)
                const uint32_t nb = b0 + kSumRing + r;
                const double* vb = v + 192ull * (nb < last ? nb : last);
#pragma unroll
                for (int q = 0; q < 4; ++q) X[r][q] = vb[3 * (16 * q + c)];
            }
        }
    }
    if (n % 64) {
        double T[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t i = nfull * 64 + 16 * q + c;
            T[q] = *(i < n ? v + 3ull * i : zero);
        }
        MCGPU_SUM_ROW(acc, T[0], one);
        MCGPU_SUM_ROW(acc, T[1], one);
        MCGPU_SUM_ROW(acc, T[2], one);
        MCGPU_SUM_ROW(acc, T[3], one);
    }
    return acc;
}

__global__ void __launch_bounds__(64) k_mesh_sums(const double* __restrict__ v0, uint32_t n0,
                                                  const double* __restrict__ v1, uint32_t n1,
                                                  const double* __restrict__ zero, Params* P,
                                                  const uint32_t* __restrict__ only_failed)
{
    __builtin_amdgcn_s_setprio(3);
    const int m = blockIdx.x / 3, j = blockIdx.x % 3;
    if (m == 0 && !v0) return; // a resident source: its sums are set already
    if (only_failed && __builtin_amdgcn_readfirstlane(only_failed[blockIdx.x]) == 0u) return;
    const double acc = serial_sum((m ? v1 : v0) + j, m ? n1 : n0, zero);
    if (threadIdx.x == 0) P->sum[m][j] = acc;
}

// ---------------------------------------------------------------------------
// The same sums from a chain about five times shorter.  A step s = RN(s' + x)
// whose result stays in the binade of s' (|s| and |s'| in [2^e, 2^(e+1)), same
// sign) rounds to the multiples of u = 2^(e-52), of which s' already is one:
// s = s' + RN_u(x), whatever s' is, unless RN_u(x) is a tie.  A run of such
// steps adds sum RN_u(x) to its start value exactly, one rounding-free add.
// Binade changes and ties stay single steps of the chain.
//
// k_sum_blocks and k_sum_items predict the binade of every partial sum from
// an approximate prefix (any summation order) and write each 32-value thread
// range as items: raw values, and for every same-binade run its exact sum
// (integer multiples of u) with the interval [lo, hi] of start values for
// which every partial sum of the run stays in [2^e + u, 2^(e+1)].  k_sum_chain adds the items left to right with the serial sum's DPP
// chain and checkpoints every 64 items; k_sum_verify replays every 16 items
// from the checkpoint before them and checks each run's actual start value
// against its interval.  A start inside it proves the run exact (the prediction only
// chooses the runs); a failure sends that coordinate to the serial sum
// (k_mesh_sums with the failure flags).  On the icosphere-7 pair the longest
// chain holds 35k items for 164k vertices.
constexpr int kSumThreads = 128, kSumPer = 32;
constexpr uint32_t kSumTile = kSumThreads * kSumPer; // values per block and per item region
constexpr int kSumLds = kSumPer + 1;                 // LDS stride of a thread's values (no bank conflicts)

struct SumBufs {
    const double* v[2]; // vertex arrays (xyz), null when that mesh's sums are preset
    uint32_t n[2];      // vertex counts
    uint32_t nblk;      // blocks per chain (of the larger mesh)
    double* bsum;       // [6][nblk]
    double* items;      // [6][nblk * kSumTile]
    double* lo;
    double* hi;
    uint32_t* nitem;    // [6][nblk]: items of each block region, padded to 64
    uint32_t* list;     // [6][lcap]: 64-item blocks in chain order
    uint32_t lcap;
    double* ckpt;       // [6][lcap]: chain value after each 64 items
    uint32_t* fail;     // [6]
    uint32_t* nb64;     // [6]: 64-item blocks of each chain
};

__device__ __forceinline__ bool sum_chain_active(const SumBufs& B, int ch) { return B.v[ch / 3] != nullptr && B.n[ch / 3] > 0; }

__global__ void __launch_bounds__(kSumThreads) k_sum_blocks(SumBufs B)
{
    __shared__ double red[kSumThreads / 64];
    const int ch = blockIdx.y, m = ch / 3, j = ch % 3;
    if (blockIdx.x == 0 && threadIdx.x == 0) B.fail[ch] = 0u;
    if (!sum_chain_active(B, ch)) return;
    const uint32_t n = B.n[m], b = blockIdx.x;
    if ((uint64_t)b * kSumTile >= n) return;
    const double* v = B.v[m] + j;
    const uint64_t i0 = (uint64_t)b * kSumTile;
    double t = 0.0;
#pragma unroll 8
    for (uint32_t k = threadIdx.x; k < kSumTile; k += kSumThreads)
        if (i0 + k < n) t += v[3 * (i0 + k)];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int w = 0; w < kSumThreads / 64; ++w) a += red[w];
        B.bsum[(uint64_t)ch * B.nblk + b] = a;
    }
}

// (sign, exponent) of a predicted partial sum; kSumTiny for zero, |a| < 2^-900
// or |a| >= 2^1001 (keeps 2^(e+3) and 2^(52-e) finite)
constexpr int kSumTiny = 1 << 20;
__device__ __forceinline__ int sum_binade(double a)
{
    const uint64_t bits = (uint64_t)__double_as_longlong(a);
    const int ex = (int)((bits >> 52) & 0x7ff);
    if (ex < 1023 - 900 || ex > 1023 + 1000) return kSumTiny;
    return (ex << 1) | (int)(bits >> 63);
}

struct SumRun {
    long long q, qmin, qmax; // running sum of RN_u(x) / u and its extremes (0 included)
    int key;                 // binade; kSumTiny when no run is open
    bool bad;
};

// one thread's items for its values x[0..cnt) (LDS), starting from the
// approximate prefix a: emission through emit(item, lo, hi)
template <class Emit>
__device__ __forceinline__ void sum_items_walk(const double* x, uint32_t cnt, double a, Emit emit)
{
    SumRun R{0, 0, 0, kSumTiny, false};
    int prev = sum_binade(a);
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    auto close = [&]() {
        if (R.key == kSumTiny) return;
        const int e = (R.key >> 1) - 1023;
        const double u = __longlong_as_double((long long)(e - 52 + 1023) << 52);
        // in units of u: every partial sum v + Q within [2^e + u, 2^(e+1)]
        // (mirrored for a negative binade).  The exact step value is within u/2
        // of v + Q, and there the nearest double is v + Q: inside the binade the
        // doubles are the multiples of u, and just above 2^(e+1) they are 2u
        // apart.  Exact in integers, rounded inward to doubles.
        const long long L = (R.key & 1) ? -(1ll << 53) : (1ll << 52) + 1;
        const long long H = (R.key & 1) ? -((1ll << 52) + 1) : (1ll << 53);
        const long long lo_u = L - R.qmin, hi_u = H - R.qmax;
        double lod = (double)lo_u, hid = (double)hi_u;
        if ((long long)lod < lo_u) lod = nextafter(lod, 1.0e300);
        if ((long long)hid > hi_u) hid = nextafter(hid, -1.0e300);
        double lo = lod * u, hi = hid * u;
        if (R.bad) { // empty: the verification fails
            lo = inf;
            hi = -inf;
        }
        emit((double)R.q * u, lo, hi);
        R.key = kSumTiny;
    };
    for (uint32_t i = 0; i < cnt; ++i) {
        const double xi = x[i];
        a = a + xi;
        const int key = sum_binade(a);
        bool stateless = key == prev && key != kSumTiny;
        double c = 0.0;
        if (stateless) {
            const int e = (key >> 1) - 1023;
            const double inv_u = __longlong_as_double((long long)(52 - e + 1023) << 52);
            const double ax = xi < 0 ? -xi : xi;
            if (!(ax < __longlong_as_double((long long)(e + 3 + 1023) << 52))) {
                stateless = false;
            } else {
                const double q = xi * inv_u;
                c = rint(q);
                if (q - floor(q) == 0.5) stateless = false; // a tie: the parity of s' decides
            }
        }
        if (stateless) {
            if (R.key != key) {
                close();
                R = SumRun{0, 0, 0, key, false};
            }
            R.q += (long long)c;
            R.qmin = R.q < R.qmin ? R.q : R.qmin;
            R.qmax = R.q > R.qmax ? R.q : R.qmax;
            if (R.q > (1ll << 53) || R.q < -(1ll << 53)) R.bad = true;
        } else {
            close();
            emit(xi, -inf, inf);
        }
        prev = key;
    }
    close();
}

__global__ void __launch_bounds__(kSumThreads) k_sum_items(SumBufs B)
{
    __shared__ double xs[kSumThreads * kSumLds];
    __shared__ double wpre[kSumThreads / 64];
    __shared__ uint32_t wcnt[kSumThreads / 64];
    __shared__ double a0s;
    const int ch = blockIdx.y, m = ch / 3, j = ch % 3;
    if (!sum_chain_active(B, ch)) return;
    const uint32_t n = B.n[m], b = blockIdx.x;
    if ((uint64_t)b * kSumTile >= n) {
        if (threadIdx.x == 0) B.nitem[(uint64_t)ch * B.nblk + b] = 0u;
        return;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const double* v = B.v[m] + j;
    const uint64_t i0 = (uint64_t)b * kSumTile;
    const uint32_t nb = (n - i0) < kSumTile ? (uint32_t)(n - i0) : kSumTile;
    // the block's values into LDS (coalesced), value k of thread t at t * kSumLds + k
#pragma unroll 8
    for (uint32_t k = threadIdx.x; k < kSumTile; k += kSumThreads)
        xs[(k / kSumPer) * kSumLds + k % kSumPer] = k < nb ? v[3 * (i0 + k)] : 0.0;
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (uint32_t q = 0; q < b; ++q) a += B.bsum[(uint64_t)ch * B.nblk + q];
        a0s = a;
    }
    __syncthreads();
    const double* x = xs + threadIdx.x * kSumLds;
    const uint32_t t0 = threadIdx.x * kSumPer;
    const uint32_t cnt = t0 < nb ? (nb - t0 < (uint32_t)kSumPer ? nb - t0 : (uint32_t)kSumPer) : 0u;
    double t = 0.0;
    for (uint32_t k = 0; k < cnt; ++k) t += x[k];
    // approximate prefix at this thread's first value
    double y = t;
    for (int o = 1; o < 64; o <<= 1) {
        const double z = __shfl_up(y, o, 64);
        if (lane >= o) y += z;
    }
    if (lane == 63) wpre[wid] = y;
    __syncthreads();
    double a = a0s + (y - t);
    for (int w = 0; w < wid; ++w) a += wpre[w];
    uint32_t nit = 0;
    if (cnt) sum_items_walk(x, cnt, a, [&](double, double, double) { ++nit; });
    // this thread's offset in the block region
    uint32_t r = nit;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t z = __shfl_up(r, o, 64);
        if (lane >= o) r += z;
    }
    if (lane == 63) wcnt[wid] = r;
    __syncthreads();
    uint32_t off = r - nit, tot = 0;
    for (int w = 0; w < kSumThreads / 64; ++w) {
        if (w < wid) off += wcnt[w];
        tot += wcnt[w];
    }
    const uint64_t base = ((uint64_t)ch * B.nblk + b) * kSumTile;
    double* it = B.items + base;
    double* lo = B.lo + base;
    double* hi = B.hi + base;
    if (cnt) {
        uint32_t k = off;
        sum_items_walk(x, cnt, a, [&](double val, double l, double h) {
            it[k] = val;
            lo[k] = l;
            hi[k] = h;
            ++k;
        });
    }
    // +0.0 to the next multiple of 64 (adding +0.0 changes no partial sum)
    const uint32_t padded = (tot + 63u) & ~63u;
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    for (uint32_t k = tot + threadIdx.x; k < padded; k += kSumThreads) {
        it[k] = 0.0;
        lo[k] = -inf;
        hi[k] = inf;
    }
    if (threadIdx.x == 0) B.nitem[(uint64_t)ch * B.nblk + b] = padded;
}

// one wave per chain (on the reserved CUs): the DPP chain over the item
// regions in order, a checkpoint after every 16 items.  A scalar cursor
// (region, 64-item block) walks the regions ahead of the chain, so every
// refill address is a uniform base with no dependent load; the block list is
// written for the verification.
__global__ void __launch_bounds__(64) k_sum_chain(SumBufs B, Params* P)
{
    extern __shared__ uint32_t cnt64[]; // 64-item blocks of each region
    __builtin_amdgcn_s_setprio(3);
    const int ch = blockIdx.x, m = ch / 3, j = ch % 3;
    if (!sum_chain_active(B, ch)) return;
    const uint32_t lane = threadIdx.x, c = lane & 15;
    const uint32_t* nit = B.nitem + (uint64_t)ch * B.nblk;
    uint32_t* list = B.list + (uint64_t)ch * B.lcap;
    uint32_t nb = 0;
    for (uint32_t b = 0; b < B.nblk; ++b) {
        const uint32_t k = nit[b] / 64u;
        if (lane == 0) cnt64[b] = k;
        for (uint32_t q = lane; q < k; q += 64) list[nb + q] = b * (kSumTile / 64) + q;
        nb += k;
    }
    __syncthreads();
    nb = __builtin_amdgcn_readfirstlane(nb);
    if (lane == 0) B.nb64[ch] = nb;
    const double* items = B.items + (uint64_t)ch * B.nblk * kSumTile;
    double* ckpt = B.ckpt + (uint64_t)ch * B.lcap;
    // load cursor: region rb (cur 64-item blocks), block jb within it; it
    // reads a region's count only when it enters the region and stays on the
    // last block at the end
    uint32_t rb = 0, jb = 0, last_blk = 0, cur = __builtin_amdgcn_readfirstlane(cnt64[0]);
    auto next_blk = [&]() -> uint32_t {
        while (jb >= cur && rb < B.nblk) {
            ++rb;
            jb = 0;
            cur = rb < B.nblk ? __builtin_amdgcn_readfirstlane(cnt64[rb]) : 0u;
        }
        if (rb < B.nblk) last_blk = rb * (kSumTile / 64) + jb++;
        return last_blk;
    };
    double one = 1.0, acc = 0.0;
    if (nb > 0) {
        double X[kSumRing][4];
#pragma unroll
        for (int r = 0; r < kSumRing; ++r) {
            const double* vb = items + 64ull * next_blk();
#pragma unroll
            for (int q = 0; q < 4; ++q) X[r][q] = vb[16 * q + c];
        }
        for (uint32_t b0 = 0; b0 < nb; b0 += kSumRing) {
#pragma unroll
            for (int r = 0; r < kSumRing; ++r) {
                if (b0 + r < nb) {
                    MCGPU_SUM_ROW(acc, X[r][0], one);
                    MCGPU_SUM_ROW(acc, X[r][1], one);
                    MCGPU_SUM_ROW(acc, X[r][2], one);
                    MCGPU_SUM_ROW(acc, X[r][3], one);
                    if (lane == 0) ckpt[b0 + r] = acc;
                }
                const double* vb = items + 64ull * next_blk();
#pragma unroll
                for (int q = 0; q < 4; ++q) X[r][q] = vb[16 * q + c];
            }
        }
    }
    if (lane == 0) P->sum[m][j] = acc;
}

// one thread per 16 items: replay from the checkpoint before their 64-item
// block (at most 48 unchecked adds), then check each run's start value in
// the 16 against its interval, and the block's last row against the next
// checkpoint
__global__ void __launch_bounds__(256) k_sum_verify(SumBufs B)
{
    const int ch = blockIdx.y;
    if (!sum_chain_active(B, ch)) return;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 4 * B.nb64[ch]) return;
    const uint32_t k = r / 4, row = r % 4;
    const uint64_t cb = (uint64_t)ch * B.lcap;
    const uint64_t base = (uint64_t)ch * B.nblk * kSumTile + 64ull * B.list[cb + k];
    double v = k ? B.ckpt[cb + k - 1] : 0.0;
    for (uint32_t i = 0; i < 16 * row; ++i) v = v + B.items[base + i];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint64_t at = base + 16 * row + i;
        const double l = B.lo[at], h = B.hi[at], x = B.items[at];
        ok &= (l <= v) & (v <= h);
        v = v + x;
    }
    if (row == 3) ok &= __double_as_longlong(v) == __double_as_longlong(B.ckpt[cb + k]);
    if (!ok) B.fail[ch] = 1u;
}

// Batches: the 6 chains of every pair (pair p, mesh m, coordinate j is chain
// 6 p + 3 m + j), handed out longest first (order[]) by a work queue to a
// persistent grid of waves, so that long chains start early and short ones
// fill in behind them.
__global__ void __launch_bounds__(64) k_batch_sums(const double* __restrict__ vs, const double* __restrict__ vc,
                                                   const BatchPair* __restrict__ bp, const uint32_t* __restrict__ order,
                                                   uint32_t nchains, Counters* C, Params* P)
{
    __builtin_amdgcn_s_setprio(3);
    for (;;) {
        // Everything that steers this loop is a scalar (readfirstlane) and the
        // body has no lane-dependent branch: with a chain index read back from
        // lane 0 by a shuffle, or an `if (lane 0)` around the atomic or the
        // store, the compiler treats the exit as divergent and may structure
        // the loop so that lanes 1-63 keep iterating without lane 0, which
        // never ends.  Every lane issues the atomic (lane 0 adds 1, the others
        // 0) and lane 0's result is the wave's chain.
        const uint32_t w = __builtin_amdgcn_readfirstlane(atomicAdd(&C->sum_next, threadIdx.x == 0 ? 1u : 0u));
        if (w >= nchains) break;
        const uint32_t chain = __builtin_amdgcn_readfirstlane(order[w]);
        const uint32_t p = chain / 6, m = (chain % 6) / 3, j = chain % 3;
        const uint32_t v0 = __builtin_amdgcn_readfirstlane(m ? bp[p].vc0 : bp[p].vs0);
        const uint32_t n = __builtin_amdgcn_readfirstlane(m ? bp[p].nvc : bp[p].nvs);
        const double acc = serial_sum((m ? vc : vs) + 3ull * v0 + j, n, &C->zero);
        P[p].sum[m][j] = acc; // the same value in every lane
    }
}

// com, t, diag and eps of a pair from its coordinate sums (P->sum) and the raw
// boxes of its two meshes (rbox[m] = order keys lo xyz, hi xyz)
__device__ inline void shift_params_of(Params* P, const unsigned long long (*rbox)[6], uint32_t n0, uint32_t n1,
                                       double gp_constant, int eps_absolute)
{
    // raw bounding boxes from the validation pass (min/max do not depend on order)
    for (int m = 0; m < 2; ++m)
        for (int j = 0; j < 3; ++j) {
            P->bmin[m][j] = ord_value(rbox[m][j]);
            P->bmax[m][j] = ord_value(rbox[m][3 + j]);
        }
    double c0[3], c1[3], tpq[3];
    for (int j = 0; j < 3; ++j) {
        c0[j] = P->sum[0][j] / (double)n0;
        c1[j] = P->sum[1][j] / (double)n1;
    }
    for (int j = 0; j < 3; ++j) {
        P->com[j] = (c0[j] + c1[j]) / 2.0;
        const double gmin = (P->bmin[1][j] < P->bmin[0][j]) ? P->bmin[1][j] : P->bmin[0][j];
        tpq[j] = P->com[j] - gmin;
    }
    const double tl = sqrt(dot3(tpq, tpq));
    for (int j = 0; j < 3; ++j) P->t[j] = tpq[j] + tpq[j] / tl;
    double cd[3];
    for (int j = 0; j < 3; ++j) cd[j] = (P->bmax[1][j] + P->t[j]) - (P->bmin[1][j] + P->t[j]);
    P->diag = sqrt(dot3(cd, cd)); // length(cutmesh_bboxmax - cutmesh_bboxmin), preproc.cpp:2518
    P->gp_constant = gp_constant;
    P->eps_absolute = eps_absolute;
    P->eps = (eps_absolute ? 1.0 : P->diag) * gp_constant; // preproc.cpp:2667-2676
}

__global__ void k_shift_params(Params* P, const Counters* __restrict__ C, uint32_t n0, uint32_t n1, double gp_constant,
                               int eps_absolute, int has_pert, double p0, double p1, double p2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    shift_params_of(P, C->rbox, n0, n1, gp_constant, eps_absolute);
    P->has_pert = has_pert;
    P->pert[0] = p0;
    P->pert[1] = p1;
    P->pert[2] = p2;
}

// batches: one thread per pair
__global__ void k_batch_shift_params(Params* P, const unsigned long long (*rbox)[2][6], const BatchPair* __restrict__ bp,
                                     uint32_t npairs, double gp_constant, int eps_absolute)
{
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += gridDim.x * blockDim.x) {
        shift_params_of(P + p, rbox[p], bp[p].nvs, bp[p].nvc, gp_constant, eps_absolute);
        P[p].has_pert = 0;
        P[p].pert[0] = P[p].pert[1] = P[p].pert[2] = 0.0;
    }
}

__global__ void k_set_perturbation(Params* P, int has_pert, double p0, double p1, double p2)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        P->has_pert = has_pert;
        P->pert[0] = p0;
        P->pert[1] = p1;
        P->pert[2] = p2;
    }
}

// Per-axis tolerance that makes the raw-space broad phase find every pair
// whose shifted boxes overlap.  With u = 2^-53 and, per axis, E the extent of
// both meshes' raw box, W its largest magnitude and n the larger vertex count:
//   |x - com| <= R = E + (n + 2) 2^-52 W   (com is a mean of sequentially
//                                           summed coordinates),
//   |t| <= (R + 1)(1 + 4u),
//   s(x) = x - com + t + d(x) with |d(x)| <= u (3R + 1)(1 + 4u),
// and the shifted test s(a_lo) <= fl(s(b_hi) + eps) implies
//   a_lo <= b_hi + eps (1 + u) + u (8R + 4)(1 + 4u),
// which the raw test a_lo <= fl(b_hi + tol) accepts once tol covers that
// bound plus the rounding of b_hi + tol (u (W + tol)).  eps is bounded from
// the cut mesh's raw extent.  tol keeps a factor of 32 over every rounding
// term.  A bound that is not finite marks the input unsupported (bit 1).
__device__ inline bool tolerance_of(const unsigned long long (*rbox)[6], uint32_t nv0, uint32_t nv1, double gp_constant,
                                    int absolute, double* tol, bool* loose = nullptr)
{
    const double u32 = 3.552713678800501e-15; // 2^-48 = 32 u
    const double n = (double)(nv0 > nv1 ? nv0 : nv1);
    double cext2 = 0.0, wmax = 0.0, rmax = 0.0, emax = 0.0, R[3], W[3];
    bool finite = true;
    for (int j = 0; j < 3; ++j) {
        const double l0 = ord_value(rbox[0][j]), l1 = ord_value(rbox[1][j]);
        const double h0 = ord_value(rbox[0][3 + j]), h1 = ord_value(rbox[1][3 + j]);
        const double lo = l0 < l1 ? l0 : l1, hi = h0 > h1 ? h0 : h1;
        W[j] = fmax(fabs(lo), fabs(hi));
        R[j] = (hi - lo) + (n + 2.0) * 2.220446049250313e-16 * W[j];
        emax = fmax(emax, hi - lo);
        const double ce = h1 - l1;
        cext2 += ce * ce;
        wmax = fmax(wmax, W[j]);
        rmax = fmax(rmax, R[j]);
        finite = finite && isfinite(lo) && isfinite(hi);
    }
    // diag = |(bmax + t) - (bmin + t)| of the cut mesh, each term within
    // u (2W + 2|t| + E) of the raw extent
    const double diag_hi = (sqrt(cext2) + u32 * 2.0 * (wmax + 2.0 * rmax + 2.0)) * (1.0 + u32);
    const double eps_hi = (absolute ? 1.0 : diag_hi) * gp_constant * (1.0 + u32);
    bool wide = false;
    for (int j = 0; j < 3; ++j) {
        const double round = u32 * (8.0 * R[j] + W[j] + 4.0);
        tol[j] = eps_hi * (1.0 + u32) + round;
        finite = finite && isfinite(tol[j]);
        // the rounding of the shift, not eps, widens the boxes by more than a
        // sixteenth of the pair's largest extent: coordinates far from the
        // origin relative to the meshes (near 2^52 the shift itself rounds to
        // whole units)
        wide = wide || round * 16.0 > emax;
    }
    if (loose) *loose = wide;
    return finite;
}

// allow_loose: the dispatch can rerun with its broad phase in shifted
// coordinates (a single pair whose source is not resident)
__global__ void k_tolerance(Params* P, Counters* C, uint32_t nv0, uint32_t nv1, double gp_constant, int absolute,
                            int allow_loose)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    bool loose = false;
    if (!tolerance_of(C->rbox, nv0, nv1, gp_constant, absolute, P->tol, &loose)) C->unsupported |= 2u;
    else if (loose && allow_loose) C->loose = 1u;
}

// batches: one thread per pair; any pair without a finite bound rejects the batch
__global__ void k_batch_tolerance(Params* P, const unsigned long long (*rbox)[2][6], const BatchPair* __restrict__ bp,
                                  uint32_t npairs, double gp_constant, int absolute, Counters* C)
{
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += gridDim.x * blockDim.x)
        if (!tolerance_of(rbox[p], bp[p].nvs, bp[p].nvc, gp_constant, absolute, P[p].tol)) atomicOr(&C->unsupported, 2u);
}

__device__ __forceinline__ bool dispatch_aborted(const Counters* C)
{
    // bad face lists, or no finite broad-phase bound: nothing may index vertices
    return C->bad_input != 0 || (C->unsupported & 2u) != 0;
}

// ps coordinates: ((x - com) + t) + 0 for source vertices, + perturbation for
// cut vertices (preproc.cpp:244-276)
__global__ void k_shift_vertices(const double* __restrict__ vs, uint32_t ns, const double* __restrict__ vc, uint32_t nc,
                                 const Params* __restrict__ P, double* __restrict__ ps_v, int cut_only)
{
    const uint64_t total = (uint64_t)(ns + nc) * 3, split = (uint64_t)ns * 3;
    const double com[3] = {P->com[0], P->com[1], P->com[2]};
    const double t[3] = {P->t[0], P->t[1], P->t[2]};
    const int hp = P->has_pert;
    const double pe[3] = {hp ? P->pert[0] : 0.0, hp ? P->pert[1] : 0.0, hp ? P->pert[2] : 0.0};
    const uint64_t start = cut_only ? split : 0;
    for (uint64_t i = start + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const bool cut = i >= split;
        const int j = (int)(i % 3);
        const double x = cut ? vc[i - split] : vs[i];
        const double y = (x - com[j]) + t[j];
        ps_v[i] = y + (cut ? pe[j] : 0.0);
    }
}

// serving: the resident source's sums, raw box and edge count into this
// dispatch's parameters and counters (after the counter reset)
struct ResidentGpuView {
    double sum[3];
    unsigned long long rbox[6];
    unsigned long long ne;
};

__global__ void k_resident_source(Counters* C, Params* P, const ResidentGpuView* r)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int j = 0; j < 3; ++j) P->sum[0][j] = r->sum[j];
    for (int j = 0; j < 6; ++j) C->rbox[0][j] = r->rbox[j];
    C->n_edges_src = r->ne;
}

// batches: every pair's vertices with its own shift (no perturbation)
__global__ void k_batch_shift_vertices(const double* __restrict__ vs, uint32_t ns, const double* __restrict__ vc,
                                       uint32_t nc, const uint32_t* __restrict__ vpair, const Params* __restrict__ P,
                                       double* __restrict__ ps_v)
{
    const uint64_t total = (uint64_t)(ns + nc) * 3, split = (uint64_t)ns * 3;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const bool cut = i >= split;
        const int j = (int)(i % 3);
        const Params* q = P + vpair[i / 3];
        const double x = cut ? vc[i - split] : vs[i];
        const double y = (x - q->com[j]) + q->t[j];
        ps_v[i] = y + 0.0;
    }
}

// ===========================================================================
// 2. fills, offsets, validation
// ===========================================================================

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_triangle_offsets(uint32_t* __restrict__ off, uint32_t nf)
{
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f <= nf; f += (uint64_t)gridDim.x * blockDim.x)
        off[f] = (uint32_t)(3 * f);
}

__global__ void k_offsets_tail(uint32_t* __restrict__ off, uint32_t nf, const uint32_t* __restrict__ total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) off[nf] = *total;
}

__global__ void k_validate(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ off,
                           const uint32_t* __restrict__ sizes, uint32_t nf, uint32_t nv, Counters* C)
{
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = off[f], n = sizes ? sizes[f] : 3u;
        bool bad = n < 3;
        for (uint32_t i = 0; i < n && !bad; ++i) {
            const uint32_t a = idx[b + i];
            bad = a >= nv;
            for (uint32_t k = 0; k < i && !bad; ++k) bad = idx[b + k] == a; // duplicate vertex in a face
        }
        if (bad) atomicOr(&C->bad_input, 1u);
    }
}

// ===========================================================================
// 3. polygon-soup edge numbering (hmesh_t::add_face, hmesh.cpp:523-650)
//
// An edge is created at its first use in face-insertion order and its h0
// follows that use.  Source faces are inserted as given; cut faces are inserted
// from the cut mesh's vertex lists, which start at the second vertex
// (kernel.cpp:1717-1727), so a cut corner's insertion rank is its index minus
// one within the face.  hfirst[slot] = min(rank << 32 | source vertex).
// ===========================================================================

__device__ __forceinline__ uint64_t edge_key(uint32_t a, uint32_t b)
{
    return a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
}

__device__ __forceinline__ uint32_t corner_rank(uint32_t b, uint32_t i, uint32_t n, int is_cut)
{
    return b + (is_cut ? (i == 0 ? n - 1 : i - 1) : i);
}

// corner c -> (face, corner-in-face); triangle meshes need no table
struct CornerMap {
    const uint32_t* off;
    const uint32_t* cface; // polygon meshes: face of each corner; null for triangles
    __device__ __forceinline__ void operator()(uint64_t c, uint32_t& f, uint32_t& b, uint32_t& n, uint32_t& i) const
    {
        if (!cface) {
            f = (uint32_t)(c / 3);
            b = 3 * f;
            n = 3;
        } else {
            f = cface[c];
            b = off[f];
            n = off[f + 1] - b;
        }
        i = (uint32_t)c - b;
    }
};

__global__ void k_corner_faces(const uint32_t* __restrict__ off, uint32_t nf, uint32_t* __restrict__ cface)
{
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x)
        for (uint32_t c = off[f]; c < off[f + 1]; ++c) cface[c] = (uint32_t)f;
}

__global__ void k_edge_insert(const uint32_t* __restrict__ idx, CornerMap cm, uint64_t nk, int is_cut,
                              unsigned long long* __restrict__ hkeys, unsigned long long* __restrict__ hfirst, uint64_t mask)
{
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nk; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f, b, n, i;
        cm(c, f, b, n, i);
        const uint32_t a = idx[c], d = idx[b + (i + 1 == n ? 0 : i + 1)];
        const uint64_t slot = hash_insert(hkeys, mask, edge_key(a, d));
        atomicMin(&hfirst[slot], ((unsigned long long)corner_rank(b, i, n, is_cut) << 32) | a);
    }
}

__global__ void k_edge_flag(const uint32_t* __restrict__ idx, CornerMap cm, uint64_t nk, int is_cut,
                            const unsigned long long* __restrict__ hkeys, const unsigned long long* __restrict__ hfirst,
                            uint64_t mask, uint32_t* __restrict__ flag, Counters* C)
{
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nk; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f, b, n, i;
        cm(c, f, b, n, i);
        const uint32_t a = idx[c], d = idx[b + (i + 1 == n ? 0 : i + 1)];
        const uint32_t rank = corner_rank(b, i, n, is_cut);
        const int64_t slot = hash_find(hkeys, mask, edge_key(a, d));
        if (slot < 0) {
            atomicOr(&C->nonmanifold, 2u);
            continue;
        }
        const unsigned long long first = hfirst[slot];
        const uint32_t r0 = (uint32_t)(first >> 32);
        flag[rank] = (r0 == rank) ? 1u : 0u;
        if (r0 != rank && (uint32_t)first == a) atomicOr(&C->nonmanifold, 1u); // used twice in one direction
    }
}

__global__ void k_store_edge_count(const uint32_t* __restrict__ total, Counters* C, int is_cut)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (is_cut) C->n_edges_cut = *total;
        else C->n_edges_src = *total;
    }
}

// corner -> ps edge; edge endpoints (h0 direction) and incident faces.  A
// third corner on one edge (non-manifold) finds the h1 face already taken.
__global__ void k_edge_assign(const uint32_t* __restrict__ idx, CornerMap cm, uint64_t nk, int is_cut,
                              const unsigned long long* __restrict__ hkeys, const unsigned long long* __restrict__ hfirst,
                              uint64_t mask, const uint32_t* __restrict__ eid_by_rank, Counters* C, uint32_t vofs,
                              uint32_t fofs, uint32_t* __restrict__ cedge, uint32_t* __restrict__ ev0,
                              uint32_t* __restrict__ ev1, int32_t* __restrict__ ef0, int32_t* __restrict__ ef1)
{
    const uint32_t ebase = is_cut ? (uint32_t)C->n_edges_src : 0u;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nk; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f, b, n, i;
        cm(c, f, b, n, i);
        const uint32_t a = idx[c], d = idx[b + (i + 1 == n ? 0 : i + 1)];
        const uint32_t rank = corner_rank(b, i, n, is_cut);
        const int64_t slot = hash_find(hkeys, mask, edge_key(a, d));
        if (slot < 0) continue; // flagged by k_edge_flag
        const uint32_t r0 = (uint32_t)(hfirst[slot] >> 32);
        const uint32_t e = ebase + eid_by_rank[r0];
        cedge[c] = e;
        if (r0 == rank) {
            ev0[e] = vofs + a;
            ev1[e] = vofs + d;
            ef0[e] = (int32_t)(fofs + f);
        } else if (atomicExch(&ef1[e], (int32_t)(fofs + f)) != -1) {
            atomicOr(&C->nonmanifold, 4u);
        }
    }
}

// Packed edge table, for meshes whose vertex ids fit 21 bits and whose corner
// ranks fit 22: one 64-bit word per edge holds (min vertex, max vertex) in
// its top 42 bits and the smallest insertion rank in its low 22, so one
// atomicCAS (plus an atomicMin when the edge is already there) replaces the
// key and first-use tables: 53 us instead of 103 us for the 983,040 corners of
// an icosphere-7.  Both meshes go through one launch per pass: corners
// [0, K0) are the source's, [K0, K0 + K1) the cut mesh's (is_cut = 1).  The
// insert leaves each corner's slot in its cedge entry; the flag pass reads
// it instead of probing and replaces it with the edge's first rank, which the
// assign pass then reads in corner order (the table word is a gather at a
// hashed slot that brings a whole line; the first rank's eid entry lies near
// the corner's own rank) before it overwrites the entry with the edge id.
// Edge ids come from one exclusive scan over both meshes'
// first-use flags, so a cut edge's id already counts the source's edges
// (with a resident source only the cut mesh is scanned, after its ne edges).
constexpr uint32_t kPackedVertexBits = 21, kPackedRankBits = 22;
constexpr unsigned long long kPackedRankMask = (1ull << kPackedRankBits) - 1;

struct TopoMesh {
    const uint32_t* idx;
    CornerMap cm;
    unsigned long long* keys;
    uint64_t mask;
    uint32_t* cedge;
    uint64_t nk;
    uint32_t vofs, fofs; // soup offsets of the mesh's vertices and faces
};

__device__ __forceinline__ uint64_t packed_edge(uint32_t a, uint32_t b)
{
    return a < b ? ((uint64_t)a << kPackedVertexBits | b) : ((uint64_t)b << kPackedVertexBits | a);
}

// the corner whose insertion rank is r: a face's ranks are a permutation of
// its corner positions, so r lies in its own face (the inverse of corner_rank)
__device__ __forceinline__ uint64_t rank_corner(const CornerMap& cm, uint32_t r, int is_cut)
{
    uint32_t f, b, n, i;
    cm(r, f, b, n, i);
    if (!is_cut) return r;
    const uint32_t j = r - b;
    return b + (j == n - 1 ? 0 : j + 1);
}

__global__ void k_fill2_u64(unsigned long long* p0, uint64_t n0, unsigned long long* p1, uint64_t n1, unsigned long long v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n0 + n1; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < n0) p0[i] = v;
        else p1[i - n0] = v;
    }
}

__global__ void k_edge_insert_packed(TopoMesh M0, TopoMesh M1, uint64_t start)
{
    const uint64_t K = M0.nk + M1.nk;
    for (uint64_t q = start + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < K; q += (uint64_t)gridDim.x * blockDim.x) {
        const int m = q >= M0.nk;
        const TopoMesh& M = m ? M1 : M0;
        const uint64_t c = m ? q - M0.nk : q;
        uint32_t f, b, n, i;
        M.cm(c, f, b, n, i);
        const uint32_t a = M.idx[c], d = M.idx[b + (i + 1 == n ? 0 : i + 1)];
        const uint64_t ek = packed_edge(a, d);
        const uint32_t rank = corner_rank(b, i, n, m);
        const unsigned long long w = (ek << kPackedRankBits) | rank;
        uint64_t h = hash64(ek) & M.mask;
        for (uint64_t probe = 0; probe <= M.mask; ++probe) {
            const unsigned long long prev = atomicCAS(&M.keys[h], kEmpty64, w);
            if (prev == kEmpty64) break;
            if ((prev >> kPackedRankBits) == ek) {
                if ((prev & kPackedRankMask) > rank) atomicMin(&M.keys[h], w);
                break;
            }
            h = (h + 1) & M.mask;
        }
        M.cedge[c] = (uint32_t)h;
    }
}

// flag[moff + rank] = 1 at an edge's first use (moff = K0 for the cut mesh)
__global__ void k_edge_flag_packed(TopoMesh M0, TopoMesh M1, uint64_t start, uint32_t* __restrict__ flag, Counters* C)
{
    const uint64_t K = M0.nk + M1.nk;
    for (uint64_t q = start + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < K; q += (uint64_t)gridDim.x * blockDim.x) {
        const int m = q >= M0.nk;
        const TopoMesh& M = m ? M1 : M0;
        const uint64_t c = m ? q - M0.nk : q;
        uint32_t f, b, n, i;
        M.cm(c, f, b, n, i);
        const uint32_t a = M.idx[c];
        const uint32_t rank = corner_rank(b, i, n, m);
        const uint32_t r0 = (uint32_t)(M.keys[M.cedge[c]] & kPackedRankMask);
        M.cedge[c] = r0; // the assign pass reads the edge's first rank here, not through the table
        flag[(m ? M0.nk : 0) + rank] = (r0 == rank) ? 1u : 0u;
        if (r0 != rank && M.idx[rank_corner(M.cm, r0, m)] == a) atomicOr(&C->nonmanifold, 1u); // used twice in one direction
    }
}

// edge counts from the scan: both meshes scanned together, or (resident
// source) the cut mesh alone
__global__ void k_store_edge_counts_packed(const uint32_t* __restrict__ eid, const uint32_t* __restrict__ flag, uint64_t K0,
                                           const uint32_t* __restrict__ total, int cut_only, Counters* C)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (cut_only) {
            C->n_edges_cut = *total;
        } else {
            const uint32_t ns = K0 ? eid[K0 - 1] + flag[K0 - 1] : 0u;
            C->n_edges_src = ns;
            C->n_edges_cut = *total - ns;
        }
    }
}

__global__ void k_edge_assign_packed(TopoMesh M0, TopoMesh M1, uint64_t start, const uint32_t* __restrict__ eid,
                                     int cut_only, Counters* C, uint32_t* __restrict__ ev0, uint32_t* __restrict__ ev1,
                                     int32_t* __restrict__ ef0, int32_t* __restrict__ ef1)
{
    const uint64_t K = M0.nk + M1.nk;
    const uint32_t eb1 = cut_only ? (uint32_t)C->n_edges_src : 0u;
    for (uint64_t q = start + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < K; q += (uint64_t)gridDim.x * blockDim.x) {
        const int m = q >= M0.nk;
        const TopoMesh& M = m ? M1 : M0;
        const uint64_t c = m ? q - M0.nk : q;
        uint32_t f, b, n, i;
        M.cm(c, f, b, n, i);
        const uint32_t a = M.idx[c], d = M.idx[b + (i + 1 == n ? 0 : i + 1)];
        const uint32_t rank = corner_rank(b, i, n, m);
        const uint32_t r0 = M.cedge[c]; // the edge's first rank (flag pass)
        const uint32_t e = (m ? eb1 : 0u) + eid[(m ? M0.nk : 0) + r0];
        M.cedge[c] = e;
        if (r0 == rank) {
            ev0[e] = M.vofs + a;
            ev1[e] = M.vofs + d;
            ef0[e] = (int32_t)(M.fofs + f);
        } else if (atomicExch(&ef1[e], (int32_t)(M.fofs + f)) != -1) {
            atomicOr(&C->nonmanifold, 4u);
        }
    }
}

// Batches: the packed edge table with one region per (pair, mesh).  Every
// corner's edge lies within its pair, so the words of pair p's edges live in a
// region of 2 K_p slots at 2 c_p of its part's table (c_p: the pair's first
// corner in the part, K_p its corner count) and hold pair-local vertex ids and
// insertion ranks (21 and 22 bits; the host takes this path only when every
// pair fits).  A pair's corners are consecutive, so the waves of a launch
// touch the regions of a few pairs at a time and stay in cache, where one
// table hashed over the whole batch misses on every probe.  Flags, scans and
// edge ids are the part-wide ones of the two-table path: the numbering does
// not change.
struct BatchTopo {
    const uint32_t* idx;        // face vertex ids of the part (ids of the concatenated part)
    CornerMap cm;
    unsigned long long* keys;   // 2 nk slots
    uint32_t* cedge;
    uint64_t nk;
    const uint32_t* fpair;      // soup face -> pair
    const BatchPair* bp;
    const uint32_t* off;        // the part's face -> first corner (nf + 1 entries)
    uint32_t fofs;              // soup id of the part's first face (0 or Fs)
    int m;                      // 0 source part, 1 cut part
};

// the region of face f's pair: first vertex id, first corner and corner count
__device__ __forceinline__ void batch_region(const BatchTopo& T, uint32_t f, uint32_t& vb, uint64_t& cb, uint64_t& kp)
{
    const BatchPair b = T.bp[T.fpair[T.fofs + f]];
    const uint32_t f0 = T.m ? b.fc0 : b.fs0, nfp = T.m ? b.nfc : b.nfs;
    vb = T.m ? b.vc0 : b.vs0;
    cb = T.off[f0];
    kp = T.off[f0 + nfp] - cb;
}

__global__ void k_edge_insert_batch(BatchTopo T)
{
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < T.nk; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f, b, n, i;
        T.cm(c, f, b, n, i);
        uint32_t vb;
        uint64_t cb, kp;
        batch_region(T, f, vb, cb, kp);
        const uint32_t a = T.idx[c] - vb, d = T.idx[b + (i + 1 == n ? 0 : i + 1)] - vb;
        const uint64_t ek = packed_edge(a, d);
        const uint32_t rank = corner_rank(b, i, n, T.m) - (uint32_t)cb;
        const unsigned long long w = (ek << kPackedRankBits) | rank;
        const uint64_t base = 2 * cb, size = 2 * kp;
        uint64_t h = base + __umul64hi(hash64(ek), size);
        for (uint64_t probe = 0; probe < size; ++probe) {
            const unsigned long long prev = atomicCAS(&T.keys[h], kEmpty64, w);
            if (prev == kEmpty64) break;
            if ((prev >> kPackedRankBits) == ek) {
                if ((prev & kPackedRankMask) > rank) atomicMin(&T.keys[h], w);
                break;
            }
            h = (h + 1 == base + size) ? base : h + 1;
        }
        T.cedge[c] = (uint32_t)h;
    }
}

// flag[rank] = 1 at an edge's first use (part-local corner ranks)
__global__ void k_edge_flag_batch(BatchTopo T, uint32_t* __restrict__ flag, Counters* C)
{
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < T.nk; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f, b, n, i;
        T.cm(c, f, b, n, i);
        uint32_t vb;
        uint64_t cb, kp;
        batch_region(T, f, vb, cb, kp);
        const uint32_t a = T.idx[c];
        const uint32_t rank = corner_rank(b, i, n, T.m);
        const uint32_t r0 = (uint32_t)cb + (uint32_t)(T.keys[T.cedge[c]] & kPackedRankMask);
        T.cedge[c] = r0; // the assign pass reads the edge's first rank here, not through the table
        flag[rank] = (r0 == rank) ? 1u : 0u;
        if (r0 != rank && T.idx[rank_corner(T.cm, r0, T.m)] == a) atomicOr(&C->nonmanifold, 1u); // used twice in one direction
    }
}

__global__ void k_edge_assign_batch(BatchTopo T, const uint32_t* __restrict__ eid_by_rank, Counters* C, uint32_t vofs,
                                    uint32_t* __restrict__ ev0, uint32_t* __restrict__ ev1, int32_t* __restrict__ ef0,
                                    int32_t* __restrict__ ef1)
{
    const uint32_t ebase = T.m ? (uint32_t)C->n_edges_src : 0u;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < T.nk; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t f, b, n, i;
        T.cm(c, f, b, n, i);
        uint32_t vb;
        uint64_t cb, kp;
        batch_region(T, f, vb, cb, kp);
        const uint32_t a = T.idx[c], d = T.idx[b + (i + 1 == n ? 0 : i + 1)];
        const uint32_t rank = corner_rank(b, i, n, T.m);
        const uint32_t r0 = T.cedge[c]; // the edge's first rank (flag pass)
        const uint32_t e = ebase + eid_by_rank[r0];
        T.cedge[c] = e;
        if (r0 == rank) {
            ev0[e] = vofs + a;
            ev1[e] = vofs + d;
            ef0[e] = (int32_t)(T.fofs + f);
        } else if (atomicExch(&ef1[e], (int32_t)(T.fofs + f)) != -1) {
            atomicOr(&C->nonmanifold, 4u);
        }
    }
}

// Single-component check of each input mesh (check_input_mesh,
// preproc.cpp:505-578 with find_connected_components, kernel.cpp:235-364):
// the reference walks the vertex graph breadth first from every unvisited
// vertex, so every vertex, a vertex no face uses included, belongs to exactly
// one component, and a mesh of more than one is refused (MC_INVALID_VALUE).
// Here hooking and pointer jumping in rounds (Shiloach-Vishkin style), each
// kernel reading what the previous one wrote: every edge whose ends have
// different roots points the larger root at the smaller (a plain store;
// when several edges hook one root, one of them wins and the others cross
// again next round), then every vertex jumps to its root.  A round without a
// hook ends the check, and the roots left are counted per mesh.  Every round
// with a crossing edge removes a root, so rounds continue (the host enqueues
// more) until none crosses.
__global__ void k_cc_init(uint32_t* __restrict__ parent, uint64_t n)
{
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        parent[v] = (uint32_t)v;
}

// cut_only: the source's edges are skipped (a resident source, checked once)
__global__ void k_cc_hook(const uint32_t* __restrict__ ev0, const uint32_t* __restrict__ ev1, Counters* C,
                          int cut_only, uint32_t* parent)
{
    if (C->cc_done) return;
    const uint64_t ne = C->n_edges_src + C->n_edges_cut;
    bool hooked = false;
    for (uint64_t e = (cut_only ? C->n_edges_src : 0) + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < ne;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t pu = parent[ev0[e]], pv = parent[ev1[e]];
        if (pu != pv) {
            parent[pu > pv ? pu : pv] = pu > pv ? pv : pu;
            hooked = true;
        }
    }
    if (__any(hooked) && __lane_id() == 0) C->cc_changed = 1u;
}

__global__ void k_cc_jump(uint32_t* parent, uint64_t start, uint64_t n, const Counters* C)
{
    if (C->cc_done) return;
    for (uint64_t v = start + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x = parent[v], y = parent[x];
        while (x != y) {
            x = y;
            y = parent[x];
        }
        parent[v] = x;
    }
}

__global__ void k_cc_round_end(Counters* C)
{
    if (C->cc_done) return;
    if (!C->cc_changed) C->cc_done = 1u;
    C->cc_changed = 0u;
}

// roots of [start, nv0) and [nv0, n), once the rounds are done
__global__ void k_cc_count(const uint32_t* __restrict__ parent, uint64_t start, uint64_t nv0, uint64_t n, Counters* C)
{
    if (!C->cc_done) return;
    unsigned c0 = 0, c1 = 0;
    for (uint64_t v = start + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        if (parent[v] == (uint32_t)v) (v < nv0 ? c0 : c1)++;
    for (int o = 32; o > 0; o >>= 1) {
        c0 += __shfl_xor(c0, o, 64);
        c1 += __shfl_xor(c1, o, 64);
    }
    if (__lane_id() == 0) {
        if (c0) atomicAdd(&C->ncc[0], c0);
        if (c1) atomicAdd(&C->ncc[1], c1);
    }
}

// ===========================================================================
// 4. face AABBs (bvh.cpp:242-313) and the LBVH over them
// ===========================================================================

__global__ void k_face_boxes(Soup S, const Params* __restrict__ P, const Counters* __restrict__ C, double* __restrict__ box,
                             int cut_only)
{
    const uint32_t F = S.nf[0] + S.nf[1];
    const double eps0 = P->eps;
    const bool aborted = dispatch_aborted(C);
    for (uint64_t g = (cut_only ? S.nf[0] : 0) + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < F;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const FaceView fv = face_view(S, (uint32_t)g);
        double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
        double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
        for (uint32_t i = 0; i < (aborted ? 0u : fv.n); ++i) {
            double p[3];
            fv((int)i, p);
            for (int j = 0; j < 3; ++j) {
                hi[j] = (hi[j] < p[j]) ? p[j] : hi[j];
                lo[j] = (p[j] < lo[j]) ? p[j] : lo[j];
            }
        }
        const double eps = S.fpair ? P[S.fpair[g]].eps : eps0; // batches: the pair's eps
        if (g >= S.nf[0] && eps > 0.0)
            for (int j = 0; j < 3; ++j) {
                hi[j] = hi[j] + eps;
                lo[j] = lo[j] - eps;
            }
        double* o = box + 6 * g;
        o[0] = lo[0], o[1] = lo[1], o[2] = lo[2], o[3] = hi[0], o[4] = hi[1], o[5] = hi[2];
    }
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// raw (unshifted) bounding box of a mesh: ord keys lo[3], hi[3]; one atomic
// per block and coordinate after an LDS reduction
__global__ void __launch_bounds__(256) k_raw_bbox(const double* __restrict__ v, uint32_t n, unsigned long long* __restrict__ box)
{
    __shared__ double red[4][6];
    double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        for (int j = 0; j < 3; ++j) {
            const double x = v[3 * i + j];
            lo[j] = x < lo[j] ? x : lo[j];
            hi[j] = x > hi[j] ? x : hi[j];
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int j = 0; j < 3; ++j) {
            const double a = __shfl_xor(lo[j], o, 64), b = __shfl_xor(hi[j], o, 64);
            lo[j] = a < lo[j] ? a : lo[j];
            hi[j] = b > hi[j] ? b : hi[j];
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int j = 0; j < 3; ++j) {
            red[w][j] = lo[j];
            red[w][3 + j] = hi[j];
        }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int j = threadIdx.x;
        double a = red[0][j], b = red[0][3 + j];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            a = red[q][j] < a ? red[q][j] : a;
            b = red[q][3 + j] > b ? red[q][3 + j] : b;
        }
        atomicMin(&box[j], ord_key(a));
        atomicMax(&box[3 + j], ord_key(b));
    }
}

// Validation of two triangle meshes in three launches instead of nine (every
// launch before the fork delays the broad phase): the counters reset with the
// raw-box minima preset in one block, both raw boxes in one reduction (each
// thread keeps one box per mesh, so a block may straddle the meshes), and both
// meshes' faces checked with their offset tables written in one pass.
__global__ void __launch_bounds__(256) k_reset_counters(Counters* C, int first_mesh)
{
    uint32_t* w = reinterpret_cast<uint32_t*>(C);
    for (uint32_t i = threadIdx.x; i < (uint32_t)(sizeof(Counters) / 4); i += blockDim.x) w[i] = 0u;
    __syncthreads();
    if (threadIdx.x < 6) {
        const int m = (int)threadIdx.x / 3, j = (int)threadIdx.x % 3;
        if (m >= first_mesh) C->rbox[m][j] = ~0ull;
    }
}

__global__ void __launch_bounds__(256) k_raw_bbox2(const double* __restrict__ v0, uint32_t n0, const double* __restrict__ v1,
                                                   uint32_t n1, Counters* C)
{
    __shared__ double red[4][2][6];
    double lo[2][3], hi[2][3];
    for (int m = 0; m < 2; ++m)
        for (int j = 0; j < 3; ++j) {
            lo[m][j] = 1.7976931348623157e308;
            hi[m][j] = -1.7976931348623157e308;
        }
    const uint64_t n = (uint64_t)n0 + n1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int m = i >= n0;
        const double* p = m ? v1 + 3 * (i - n0) : v0 + 3 * i;
        for (int j = 0; j < 3; ++j) {
            const double x = p[j];
            lo[m][j] = x < lo[m][j] ? x : lo[m][j];
            hi[m][j] = x > hi[m][j] ? x : hi[m][j];
        }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int m = 0; m < 2; ++m)
            for (int j = 0; j < 3; ++j) {
                const double a = __shfl_xor(lo[m][j], o, 64), b = __shfl_xor(hi[m][j], o, 64);
                lo[m][j] = a < lo[m][j] ? a : lo[m][j];
                hi[m][j] = b > hi[m][j] ? b : hi[m][j];
            }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int m = 0; m < 2; ++m)
            for (int j = 0; j < 3; ++j) {
                red[w][m][j] = lo[m][j];
                red[w][m][3 + j] = hi[m][j];
            }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int m = (int)threadIdx.x / 3, j = (int)threadIdx.x % 3;
        if (m ? n1 : n0) {
            double a = red[0][m][j], b = red[0][m][3 + j];
            for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
                a = red[q][m][j] < a ? red[q][m][j] : a;
                b = red[q][m][3 + j] > b ? red[q][m][3 + j] : b;
            }
            atomicMin(&C->rbox[m][j], ord_key(a));
            atomicMax(&C->rbox[m][3 + j], ord_key(b));
        }
    }
}

// faces of both triangle meshes (nf0 = 0 skips the source): off[f] = 3 f,
// every vertex index below its mesh's count, no vertex twice in a face
__global__ void k_validate_tri2(const uint32_t* __restrict__ idx0, uint32_t* __restrict__ off0, uint32_t nf0, uint32_t nv0,
                                const uint32_t* __restrict__ idx1, uint32_t* __restrict__ off1, uint32_t nf1, uint32_t nv1,
                                Counters* C)
{
    const uint64_t n = (uint64_t)nf0 + nf1;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < n; g += (uint64_t)gridDim.x * blockDim.x) {
        const int m = g >= nf0;
        const uint32_t f = (uint32_t)(m ? g - nf0 : g);
        uint32_t* off = m ? off1 : off0;
        const uint32_t nf = m ? nf1 : nf0;
        off[f] = 3 * f;
        if (f + 1 == nf) off[nf] = 3 * nf;
        const uint32_t* idx = m ? idx1 : idx0;
        const uint32_t nv = m ? nv1 : nv0;
        const uint32_t a = idx[3 * (uint64_t)f], b = idx[3 * (uint64_t)f + 1], c = idx[3 * (uint64_t)f + 2];
        if (a >= nv || b >= nv || c >= nv || a == b || b == c || a == c) atomicOr(&C->bad_input, 1u);
    }
}

// 21-bit Morton codes (128 cells per axis) of raw face-box centres with the
// mesh as bit 21, so that one sort orders the leaves of both trees, and the
// raw face boxes the broad phase runs on.
//
// The broad phase runs in unshifted coordinates so that it overlaps the serial
// coordinate sums that fix the shift.  The shift s(x) = ((x - com) + t) + 0 is
// a composition of correctly rounded additions, hence non-decreasing in x, so
// a shifted face box is s() of the raw box, and raw boxes whose cut side is
// widened by a tolerance tol covering eps plus every rounding error of the
// shift find a superset of the pairs whose shifted boxes overlap.  The exact
// shifted-box test (k_filter_pairs) then leaves exactly the reference's set.
// one mesh's inputs to the leaf pass
struct MortonMesh {
    const double* v;
    const uint32_t* idx;
    const uint32_t* off;
    const unsigned long long* rbox; // raw box as order keys
    uint32_t nf;
    uint32_t mesh_bit;
    int widen;                      // cut mesh: boxes widened by tol
};

// leaves of both meshes in one launch: faces [start, nf0 + nf1) of the soup
// (start = nf0 when the source's leaves are already made)
__global__ void k_morton_raw(MortonMesh M0, MortonMesh M1, uint32_t start, uint32_t* __restrict__ code,
                             uint32_t* __restrict__ face, float* __restrict__ rawbox, const Params* __restrict__ P,
                             const Counters* __restrict__ C, uint32_t* __restrict__ zero, uint64_t nzero)
{
    // the leaf sort's group totals (radix_ghist), cleared here rather than by a launch of their own
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nzero; i += (uint64_t)gridDim.x * blockDim.x)
        zero[i] = 0u;
    const bool aborted = dispatch_aborted(C) || C->loose; // loose: the rerun in shifted coordinates finds the pairs
    const uint32_t F = M0.nf + M1.nf;
    for (uint64_t g = start + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < F; g += (uint64_t)gridDim.x * blockDim.x) {
        const bool cutm = g >= M0.nf;
        const MortonMesh& M = cutm ? M1 : M0;
        const uint32_t f = (uint32_t)(cutm ? g - M0.nf : g);
        face[g] = f;
        float* o = rawbox + 6 * g;
        if (aborted) { // empty box: the broad phase finds nothing
            code[g] = M.mesh_bit;
            o[0] = o[1] = o[2] = 3.402823466e38f;
            o[3] = o[4] = o[5] = -3.402823466e38f;
            continue;
        }
        const uint32_t b = M.off[f], n = M.off[f + 1] - b;
        double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
        double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
        for (uint32_t i = 0; i < n; ++i) {
            const double* p = M.v + 3 * (uint64_t)M.idx[b + i];
            for (int j = 0; j < 3; ++j) {
                mn[j] = p[j] < mn[j] ? p[j] : mn[j];
                mx[j] = p[j] > mx[j] ? p[j] : mx[j];
            }
        }
        uint32_t c[3];
        for (int j = 0; j < 3; ++j) {
            const double lo = ord_value(M.rbox[j]), hi = ord_value(M.rbox[3 + j]);
            const double ext = hi > lo ? hi - lo : 1.0;
            float x = (float)((((mn[j] + mx[j]) * 0.5) - lo) / ext);
            x = fminf(fmaxf(x * 128.0f, 0.0f), 127.0f);
            c[j] = expand_bits10((uint32_t)x);
        }
        code[g] = (c[0] * 4 + c[1] * 2 + c[2]) | M.mesh_bit;
        // single precision, rounded outward, relative to the source mesh's raw
        // lower corner (a resident source keeps its tree in that frame): far
        // from the origin float32 would otherwise merge whole meshes into one
        // box.  Each of the three double operations errs by at most 2^-53 of
        // its result; pushing the bound out by 2^-48 of |x - origin| + tol
        // covers them, so the float box still contains the exact box and the
        // superset property carries over.
        for (int j = 0; j < 3; ++j) {
            const double t = M.widen ? P->tol[j] : 0.0;
            const double org = ord_value(M0.rbox[j]);
            const double dlo = mn[j] - org, dhi = mx[j] - org;
            o[j] = __double2float_rd((dlo - t) - (fabs(dlo) + t) * 3.552713678800501e-15);
            o[3 + j] = __double2float_ru((dhi + t) + (fabs(dhi) + t) * 3.552713678800501e-15);
        }
    }
}

// Leaves in shifted coordinates, for a pair whose raw-space broad phase is
// loose (Counters::loose): the exact shifted face boxes of k_face_boxes (cut
// faces already widened by eps) rounded outward to single precision, so each
// float box contains its double box and the trees find a superset of the
// pairs the exact filter keeps.  The Morton frame is the shifted image of the
// raw union box (s() is monotone, so it bounds every shifted box up to
// rounding); codes only order the leaves, and the clamp absorbs the rest.
__global__ void k_morton_shifted(const double* __restrict__ box, uint32_t nf0, uint32_t nf1, const Params* __restrict__ P,
                                 const Counters* __restrict__ C, uint32_t* __restrict__ code, uint32_t* __restrict__ face,
                                 float* __restrict__ rawbox, uint32_t* __restrict__ zero, uint64_t nzero)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nzero; i += (uint64_t)gridDim.x * blockDim.x)
        zero[i] = 0u;
    const bool aborted = dispatch_aborted(C);
    double lo[3], ext[3];
    for (int j = 0; j < 3; ++j) {
        const double a = P->bmin[0][j] < P->bmin[1][j] ? P->bmin[0][j] : P->bmin[1][j];
        const double b = P->bmax[0][j] > P->bmax[1][j] ? P->bmax[0][j] : P->bmax[1][j];
        lo[j] = (a - P->com[j]) + P->t[j];
        const double hi = (b - P->com[j]) + P->t[j];
        ext[j] = hi > lo[j] ? hi - lo[j] : 1.0;
    }
    const uint32_t F = nf0 + nf1;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < F; g += (uint64_t)gridDim.x * blockDim.x) {
        const bool cutm = g >= nf0;
        face[g] = (uint32_t)(cutm ? g - nf0 : g);
        float* o = rawbox + 6 * g;
        const uint32_t mesh_bit = cutm ? (1u << 21) : 0u;
        if (aborted) {
            code[g] = mesh_bit;
            o[0] = o[1] = o[2] = 3.402823466e38f;
            o[3] = o[4] = o[5] = -3.402823466e38f;
            continue;
        }
        const double* b = box + 6 * g;
        uint32_t c[3];
        for (int j = 0; j < 3; ++j) {
            float x = (float)((((b[j] + b[3 + j]) * 0.5) - lo[j]) / ext[j]);
            x = fminf(fmaxf(x * 128.0f, 0.0f), 127.0f);
            c[j] = expand_bits10((uint32_t)x);
            o[j] = __double2float_rd(b[j]);
            o[3 + j] = __double2float_ru(b[3 + j]);
        }
        code[g] = (c[0] * 4 + c[1] * 2 + c[2]) | mesh_bit;
    }
}

// Batches: every pair's raw boxes are moved by the pair's offset P[p].boff
// into a cell of their own, so that one pair of trees over all pairs never
// pairs faces of different pairs (the pair filter checks it regardless).  The
// offset is added in double, then the bound is pushed out by 2^-48 of its
// magnitude, which covers the rounding of that addition, before the outward
// rounding to single precision: the float box still contains the exact moved
// box.  Leaf keys: mesh as bit 31, pair above mb bits per axis of Morton
// code of the centre within the pair's own mesh box.
__device__ __forceinline__ float float_below(double x)
{
    return __double2float_rd(x - fabs(x) * 3.552713678800501e-15);
}

__device__ __forceinline__ float float_above(double x)
{
    return __double2float_ru(x + fabs(x) * 3.552713678800501e-15);
}

__global__ void k_batch_morton_raw(const double* __restrict__ v, const uint32_t* __restrict__ idx,
                                   const uint32_t* __restrict__ off, uint32_t nf, int m, uint32_t fofs,
                                   const uint32_t* __restrict__ fpair, const unsigned long long (*rbox)[2][6],
                                   const Params* __restrict__ P, const Counters* __restrict__ C, int mb,
                                   uint32_t* __restrict__ code, uint32_t* __restrict__ face, float* __restrict__ rawbox)
{
    const bool aborted = dispatch_aborted(C);
    const float cells = (float)(1u << mb);
    const uint32_t mesh_bit = m ? (1u << 31) : 0u;
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        face[f] = (uint32_t)f;
        float* o = rawbox + 6 * f;
        if (aborted) {
            code[f] = mesh_bit;
            o[0] = o[1] = o[2] = 3.402823466e38f;
            o[3] = o[4] = o[5] = -3.402823466e38f;
            continue;
        }
        const uint32_t p = fpair[fofs + f];
        const Params* q = P + p;
        const uint32_t b = off[f], n = off[f + 1] - b;
        double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
        double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
        for (uint32_t i = 0; i < n; ++i) {
            const double* x = v + 3 * (uint64_t)idx[b + i];
            for (int j = 0; j < 3; ++j) {
                mn[j] = x[j] < mn[j] ? x[j] : mn[j];
                mx[j] = x[j] > mx[j] ? x[j] : mx[j];
            }
        }
        uint32_t c[3];
        for (int j = 0; j < 3; ++j) {
            const double lo = ord_value(rbox[p][m][j]), hi = ord_value(rbox[p][m][3 + j]);
            const double ext = hi > lo ? hi - lo : 1.0;
            float x = (float)((((mn[j] + mx[j]) * 0.5) - lo) / ext);
            x = fminf(fmaxf(x * cells, 0.0f), cells - 1.0f);
            c[j] = expand_bits10((uint32_t)x);
        }
        code[f] = mesh_bit | (p << (3 * mb)) | (c[0] * 4 + c[1] * 2 + c[2]);
        for (int j = 0; j < 3; ++j) {
            const double t = m ? q->tol[j] : 0.0;
            o[j] = float_below((mn[j] - t) + q->boff[j]);
            o[3 + j] = float_above((mx[j] + t) + q->boff[j]);
        }
    }
}

// Broad phase over caller-given face boxes (min xyz, max xyz, doubles): the
// union box of one mesh's face boxes, as order keys ...
__global__ void __launch_bounds__(256) k_box_bounds(const double* __restrict__ box, uint32_t nf,
                                                    unsigned long long* __restrict__ rbox)
{
    __shared__ double red[4][6];
    double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x)
        for (int j = 0; j < 3; ++j) {
            const double a = box[6 * f + j], b = box[6 * f + 3 + j];
            lo[j] = a < lo[j] ? a : lo[j];
            hi[j] = b > hi[j] ? b : hi[j];
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int j = 0; j < 3; ++j) {
            const double a = __shfl_xor(lo[j], o, 64), b = __shfl_xor(hi[j], o, 64);
            lo[j] = a < lo[j] ? a : lo[j];
            hi[j] = b > hi[j] ? b : hi[j];
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int j = 0; j < 3; ++j) {
            red[w][j] = lo[j];
            red[w][3 + j] = hi[j];
        }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int j = threadIdx.x;
        double a = red[0][j], b = red[0][3 + j];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            a = red[q][j] < a ? red[q][j] : a;
            b = red[q][3 + j] > b ? red[q][3 + j] : b;
        }
        atomicMin(&rbox[j], ord_key(a));
        atomicMax(&rbox[3 + j], ord_key(b));
    }
}

// ... and the leaf keys and single-precision leaf boxes (rounded outward) of
// those face boxes, as k_morton_raw makes them from vertices
__global__ void k_morton_boxes(const double* __restrict__ box, uint32_t nf, const unsigned long long* __restrict__ rbox,
                               uint32_t* __restrict__ code, uint32_t* __restrict__ face, float* __restrict__ fbox,
                               uint32_t mesh_bit)
{
    double lo[3], ext[3];
    for (int j = 0; j < 3; ++j) {
        lo[j] = ord_value(rbox[j]);
        const double h = ord_value(rbox[3 + j]);
        ext[j] = h > lo[j] ? h - lo[j] : 1.0;
    }
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        const double* b = box + 6 * f;
        uint32_t c[3];
        for (int j = 0; j < 3; ++j) {
            float x = (float)((((b[j] + b[3 + j]) * 0.5) - lo[j]) / ext[j]);
            x = fminf(fmaxf(x * 128.0f, 0.0f), 127.0f); // NaN centres clamp to a cell
            c[j] = expand_bits10((uint32_t)x);
        }
        code[f] = (c[0] * 4 + c[1] * 2 + c[2]) | mesh_bit;
        face[f] = (uint32_t)f;
        float* o = fbox + 6 * f;
        for (int j = 0; j < 3; ++j) {
            o[j] = __double2float_rd(b[j]);
            o[3 + j] = __double2float_ru(b[3 + j]);
        }
    }
}

// ---------------------------------------------------------------------------
// Implicit complete binary tree over the Morton-ordered leaves (the layout
// family of the reference's OIBVH, bvh.cpp:81-193): leaf level D holds 2^D
// slots, node (l, k) lives at (2^l - 1) + k and has children (l+1, 2k) and
// (l+1, 2k+1); slots past the last leaf hold empty boxes, which overlap
// nothing.  Boxes are single precision (the raw boxes are already rounded
// outward).  The refit needs no atomics and no parent links: a 512-thread
// block reduces 1024 nodes of one level through LDS into the 10 levels above
// them, and both trees are refitted by the same launches.
// ---------------------------------------------------------------------------
constexpr int kRefitWidth = 1024;
constexpr int kRefitLevels = 10;

__device__ __forceinline__ void empty_box(float* b)
{
    b[0] = b[1] = b[2] = 3.402823466e38f;
    b[3] = b[4] = b[5] = -3.402823466e38f;
}

__device__ __forceinline__ void empty_box(double* b)
{
    b[0] = b[1] = b[2] = 1.7976931348623157e308;
    b[3] = b[4] = b[5] = -1.7976931348623157e308;
}

__device__ __forceinline__ bool boxes_overlap(const float* a, const float* b)
{
    return (a[0] <= b[3] && a[3] >= b[0]) && (a[1] <= b[4] && a[4] >= b[1]) && (a[2] <= b[5] && a[5] >= b[2]);
}

struct RefitTree {
    const float* facebox;   // leaf face boxes of this mesh (mesh-local face order)
    const uint32_t* sface;  // leaf -> mesh-local face
    float* nodes;
    int n, D, lin;          // leaves, leaf level, level this launch starts from (lin <= 0: nothing to do)
    int blocks;             // blocks of this launch for this tree
};

// one block's share of a refit round: 1024 nodes of level lin (blk-th
// group) reduced through LDS into the up to ten levels above them
__device__ inline void refit_block(const RefitTree& T, int lin, int blk)
{
    __shared__ float sb[kRefitWidth * 6];
    const int D = T.D;
    const int width = 1 << lin;
    const int per = width < kRefitWidth ? width : kRefitWidth;
    const int steps = lin < kRefitLevels ? lin : kRefitLevels;
    const int base = blk * per;
    for (int t = threadIdx.x; t < per; t += blockDim.x) {
        const int k = base + t;
        float b[6];
        if (lin == D) {
            if (k < T.n) {
                const float* fb = T.facebox + 6 * (uint64_t)T.sface[k];
                for (int q = 0; q < 6; ++q) b[q] = fb[q];
            } else {
                empty_box(b);
            }
            float* o = T.nodes + 6 * ((uint64_t)(1u << D) - 1 + k);
            for (int q = 0; q < 6; ++q) o[q] = b[q];
        } else {
            const float* src = T.nodes + 6 * ((uint64_t)(1u << lin) - 1 + k);
            for (int q = 0; q < 6; ++q) b[q] = src[q];
        }
        for (int q = 0; q < 6; ++q) sb[6 * t + q] = b[q];
    }
    __syncthreads();
    int cur = per;
    for (int st = 1; st <= steps; ++st) {
        const int half = cur >> 1;
        float u[6];
        const int t = threadIdx.x;
        if (t < half) {
            for (int q = 0; q < 3; ++q) {
                const float a0 = sb[6 * (2 * t) + q], a1 = sb[6 * (2 * t + 1) + q];
                const float b0 = sb[6 * (2 * t) + 3 + q], b1 = sb[6 * (2 * t + 1) + 3 + q];
                u[q] = a1 < a0 ? a1 : a0;
                u[3 + q] = b0 < b1 ? b1 : b0;
            }
        }
        __syncthreads();
        if (t < half) {
            for (int q = 0; q < 6; ++q) sb[6 * t + q] = u[q];
            const int l = lin - st;
            float* o = T.nodes + 6 * ((uint64_t)(1u << l) - 1 + (base >> st) + t);
            for (int q = 0; q < 6; ++q) o[q] = u[q];
        }
        __syncthreads();
        cur = half;
    }
}

// One round for both trees: blocks [0, t0.blocks) refit tree 0, the rest
// tree 1.  (Finishing the top levels in the round's last block would need
// every block to fence its writes at device scope, which on this GPU writes
// back the L2 of its XCD: measured 4x slower, and it slows every kernel
// running beside it.)
__global__ void __launch_bounds__(512) k_refit_levels(RefitTree t0, RefitTree t1)
{
    const bool second = (int)blockIdx.x >= t0.blocks;
    const RefitTree& T = second ? t1 : t0;
    const int blk = second ? (int)blockIdx.x - t0.blocks : (int)blockIdx.x;
    refit_block(T, T.lin, blk);
}

// ===========================================================================
// 5. broad phase traversal over the raw-space trees.  Every (src, cut) face
// pair whose raw boxes overlap is emitted: a superset of the pairs
// bvh.cpp:661-781 finds, which k_filter_pairs reduces to exactly that set.
// Walks are stackless: the node after a subtree follows from index arithmetic
// on the implicit tree.
// ===========================================================================

struct Tree {
    const uint32_t* sface;  // leaf -> mesh-local face
    const float* nodes;     // implicit tree boxes (single precision, rounded outward); leaf k holds
                            // the raw box of face sface[k], so leaf boxes are read in Morton order
    int n, D;
};

// Wavefront-cooperative BVH x BVH traversal.  A 64-lane wave owns
// 2^LL Morton-consecutive source leaves (a source block; kWaveLeaves by
// default, the autotuner may choose 8 or 32); their union is the source-tree
// node at depth D_s - LL.  The wave walks the cut tree breadth
// first with that node's box, all lanes testing children in parallel and
// compacting the survivors into an LDS frontier with ballots; at the leaf
// level every lane owning a source face tests it against the surviving cut
// leaves, staged through LDS.  A walk starts at cut level kStartLevel with all
// of its nodes, tested in one memory round trip, instead of descending the
// nearly empty levels above one round trip each, and then descends
// STRIDE levels per step (kTravStride by default).  The defaults, 32 leaves
// a block and four levels a step, are what the autotuner chose on the
// icosphere-7 pair (0.151 ms against 0.202 ms for 16 leaves and three levels;
// one level a step and 32 leaves took 0.23 ms).
// A frontier wider than kSplitFrontier means the node's leaves are spread
// out (the Morton curve jumped): the wave then splits the source node and
// walks its two children in turn, depth first and without a stack (the next
// node follows from index arithmetic).  A single source leaf keeps going up
// to the LDS capacity kFrontierCap, beyond which it takes a per-lane walk.
// Each lane loads kTravILP children before testing them, so one memory round
// trip covers 64 * kTravILP children.  Found pairs are staged in LDS and
// flushed with one global atomic per kTravStage - 64 pairs or so.  The walk
// is latency-bound, so its LDS (8 KB a wave) and registers (at most 128, four
// waves per SIMD) are sized for occupancy: against 14 KB and 129 registers
// the traversal went from 0.171 to 0.157 ms.
constexpr unsigned long long kMaxListEntries = 0xffffffffull; // list positions are 32-bit
constexpr int kFrontierCap = 512;
constexpr int kSplitFrontier = 256;
constexpr int kWaveLeavesLog = 5;
constexpr int kWaveLeaves = 1 << kWaveLeavesLog;
constexpr int kStartLevel = 8; // 256 nodes: kTravILP per lane
// levels descended per breadth-first step: a node's box contains its
// descendants', so testing the 16 descendants four levels down of a frontier
// node is exact (only less pruned) and takes a quarter of the dependent round
// trips
constexpr int kTravStride = 4;
constexpr int kTravILP = 4;
constexpr int kTravStage = 256;
// default work order: one wave per block, each XCD's blocks in runs of 16
// Morton-consecutive ones (TravQueue below); on the icosphere-7 pair it
// fetches 27% fewer bytes than plain grid order, in the same time
constexpr int kTravQueues = 1;
constexpr int kTravRun = 16;

// writes the wave's staged pairs to the global list (all lanes call it)
__device__ __forceinline__ void flush_pairs(const unsigned long long* stage, int sn, unsigned long long* pairs,
                                            uint64_t cap, Counters* C)
{
    __syncthreads();
    unsigned long long base = 0;
    if (__lane_id() == 0) base = atomicAdd(&C->n_raw, (unsigned long long)sn);
    base = __shfl(base, 0, 64);
    for (int q = __lane_id(); q < sn; q += 64) {
        const unsigned long long at = base + q;
        if (at < cap) pairs[at] = stage[q];
        else atomicOr(&C->ov_pairs, 1u);
    }
    __syncthreads();
}

__device__ __forceinline__ void per_lane_walk(const float* q, uint32_t s, const Tree& cut, unsigned long long* pairs,
                                              uint64_t cap, Counters* C)
{
    const int D = cut.D;
    const long long budget = 2ll << D;
    int l = 0;
    uint32_t k = 0;
    // a walk that starts with the list already full only counts its pairs
    // (one atomic at its end instead of one per pair); the host sizes the
    // next attempt, or refuses, from the count
    const bool counting = __hip_atomic_load(&C->n_raw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cap;
    unsigned long long counted = 0;
    for (long long it = 0; it < budget; ++it) {
        const float* nb = cut.nodes + 6 * ((uint64_t)(1u << l) - 1 + k);
        float b[6];
        for (int j = 0; j < 6; ++j) b[j] = nb[j];
        if (boxes_overlap(q, b)) {
            if (l == D) {
                if (counting) {
                    ++counted;
                } else {
                    const uint32_t c = cut.sface[k];
                    const uint64_t key = ((uint64_t)s << 32) | c;
                    const unsigned long long at = wave_append(&C->n_raw);
                    if (at < cap) pairs[at] = key;
                    else atomicOr(&C->ov_pairs, 1u);
                }
            } else {
                ++l;
                k <<= 1;
                continue;
            }
        }
        while (l > 0 && (k & 1u)) {
            k >>= 1;
            --l;
        }
        if (l == 0) break;
        ++k;
    }
    if (counted) {
        atomicAdd(&C->n_raw, counted);
        atomicOr(&C->ov_pairs, 1u);
    }
}

// How the source blocks (work items of LL-level source nodes) reach the waves.
// queues == 0: one wave per block, grid-stride (the hardware dispatcher hands
// out the waves).  queues == 1: the same, with the grid's blocks renumbered so
// that the blocks the dispatcher sends to one XCD (workgroup b goes to XCD
// b % kXcds) walk runs of `chunk` Morton-consecutive blocks, the part of the
// cut tree a run reaches then staying in that XCD's L2 (the grid is a
// multiple of kXcds * chunk).  queues == Q > 1: a persistent
// grid; the blocks, in Morton
// order, are cut into Q contiguous ranges, each with a head counter
// (Counters::trav_q), and a wave takes `chunk` blocks at a time from the range
// of queue blockIdx % Q, then from the others in turn.  Workgroups are handed
// to the XCDs round robin, so with Q = 8 a range of Morton-adjacent blocks,
// and the part of the cut tree they reach, stays on one XCD's L2 as long as
// its waves find work there.
struct TravQueue {
    int queues;
    int chunk;
};

constexpr uint32_t kXcds = 8; // XCDs of an MI355X; workgroups are dealt to them round robin

// the grid-local block that workgroup b walks first when the blocks of each
// XCD come in runs of `run` consecutive ones: the i-th workgroup XCD x
// receives walks block (i / run) * kXcds * run + x * run + i % run (a
// bijection of [0, g) for a grid g that is a multiple of kXcds * run)
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t run)
{
    const uint32_t x = b % kXcds, i = b / kXcds;
    return (i / run) * kXcds * run + x * run + i % run;
}

template <int LL, int STRIDE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_traverse_wave(Tree src, Tree cut, unsigned long long* __restrict__ pairs,
                                                      uint64_t cap, Counters* C, TravQueue tq)
{
    __shared__ uint32_t front[2][kFrontierCap];
    __shared__ float lbox[64][6];
    __shared__ uint32_t lface[64];
    __shared__ unsigned long long stage[kTravStage];
    int sn = 0; // staged pairs (wave-uniform)
    const int lane = threadIdx.x;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int Ds = src.D;
    const int lw = Ds >= LL ? Ds - LL : 0; // level of a wave's source node
    const uint32_t nwaves = Ds >= LL ? (uint32_t)((src.n + (1 << LL) - 1) >> LL) : 1u;
    const int s0 = cut.D < kStartLevel ? cut.D : kStartLevel;
    const uint32_t leaf_base = (uint32_t)(1u << cut.D) - 1;
    unsigned long long splits = 0, lane_walks = 0;
    const int nq = tq.queues > 1 ? tq.queues : 0; // persistent work queues
    const int home = nq ? (int)(blockIdx.x % (unsigned)nq) : 0;
    int qr = 0; // queues tried so far (persistent)
    // current block; end of the taken chunk
    uint32_t w = nq ? 0u : tq.queues == 1 ? xcd_block(blockIdx.x, (uint32_t)tq.chunk) : blockIdx.x, wend = 0;
    for (;;) {
        if (nq) {
            if (w >= wend) { // take the next chunk: home queue first, then the others
                bool got = false;
                while (qr < nq) {
                    const int qi = (home + qr) % nq;
                    const uint32_t lo = (uint32_t)((uint64_t)nwaves * qi / nq);
                    const uint32_t hi = (uint32_t)((uint64_t)nwaves * (qi + 1) / nq);
                    uint32_t g = 0;
                    if (lane == 0) g = atomicAdd(&C->trav_q[qi], (uint32_t)tq.chunk);
                    g = __builtin_amdgcn_readfirstlane(g);
                    if (g < hi - lo) {
                        w = lo + g;
                        wend = (hi - w) < (uint32_t)tq.chunk ? hi : w + (uint32_t)tq.chunk;
                        got = true;
                        break;
                    }
                    ++qr;
                }
                if (!got) break;
            }
        } else if (w >= nwaves) {
            break;
        }
        // past 2^32 raw pairs the host refuses the dispatch whatever the
        // exact count: the remaining blocks are not walked
        {
            unsigned long long seen = 0;
            if (lane == 0) seen = __hip_atomic_load(&C->n_raw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__shfl(seen, 0, 64) > kMaxListEntries) break;
        }
        int l = lw;
        uint32_t k = (uint32_t)w;
        for (;;) {
            // source node (l, k): leaves [k << span, (k + 1) << span), one per
            // lane of each of the 64 >> span lane groups (span <= LL <= 5);
            // the first group's lane is the leaf's own (valid), the copies let
            // the leaf phase test 64 >> span cut leaves at a time
            const int span = Ds - l;
            const uint32_t mask = (1u << span) - 1;
            const uint32_t i = (k << span) + ((uint32_t)lane & mask);
            const bool rep = i < (uint32_t)src.n;
            const bool valid = lane < (1 << span) && rep;
            float q[6];
            uint32_t s = 0;
            if (rep) {
                s = src.sface[i];
                const float* lb = src.nodes + 6 * ((uint64_t)(1u << Ds) - 1 + i);
                for (int j = 0; j < 6; ++j) q[j] = lb[j];
            } else {
                empty_box(q);
            }
            float wb[6];
            {
                const float* nb = src.nodes + 6 * ((uint64_t)(1u << l) - 1 + k);
                for (int j = 0; j < 6; ++j) wb[j] = nb[j];
            }
            // breadth-first walk of the cut tree with the node box, from the
            // 2^s0 nodes of level s0
            const int limit = span > 0 ? kSplitFrontier : kFrontierCap;
            int cur = 0, nf = 0, level = s0;
            bool overflow = false;
            {
                const int n0 = 1 << s0;
                const uint32_t base0 = (uint32_t)n0 - 1;
                for (int c0 = 0; c0 < n0; c0 += 64 * kTravILP) {
                    float b[kTravILP][6];
#pragma unroll
                    for (int u = 0; u < kTravILP; ++u) {
                        const int c = c0 + 64 * u + lane;
                        if (c < n0) {
                            const float* nb = cut.nodes + 6 * ((uint64_t)base0 + c);
                            for (int j = 0; j < 6; ++j) b[u][j] = nb[j];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kTravILP; ++u) {
                        const int c = c0 + 64 * u + lane;
                        const bool keep = c < n0 && boxes_overlap(wb, b[u]);
                        const uint64_t m = __ballot(keep);
                        const int pos = nf + __popcll(m & lt);
                        if (keep && pos < kFrontierCap) front[0][pos] = (uint32_t)c;
                        nf += __popcll(m);
                    }
                }
                if (nf > limit) overflow = true;
            }
            __syncthreads();
            while (nf > 0 && level < cut.D && !overflow) {
                int nn = 0;
                const int step = cut.D - level < STRIDE ? cut.D - level : STRIDE;
                const int nchild = nf << step;
                const uint32_t base_next = (uint32_t)(1u << (level + step)) - 1;
                for (int c0 = 0; c0 < nchild; c0 += 64 * kTravILP) {
                    float b[kTravILP][6];
                    uint32_t child[kTravILP];
#pragma unroll
                    for (int u = 0; u < kTravILP; ++u) {
                        const int c = c0 + 64 * u + lane;
                        child[u] = 0;
                        if (c < nchild) {
                            child[u] = (front[cur][c >> step] << step) + (uint32_t)(c & ((1 << step) - 1));
                            const float* nb = cut.nodes + 6 * ((uint64_t)base_next + child[u]);
                            for (int j = 0; j < 6; ++j) b[u][j] = nb[j];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kTravILP; ++u) {
                        const int c = c0 + 64 * u + lane;
                        const bool keep = c < nchild && boxes_overlap(wb, b[u]);
                        const uint64_t m = __ballot(keep);
                        const int pos = nn + __popcll(m & lt);
                        if (keep && pos < kFrontierCap) front[cur ^ 1][pos] = child[u];
                        nn += __popcll(m);
                    }
                }
                __syncthreads();
                if (nn > limit) overflow = true;
                nf = nn;
                cur ^= 1;
                level += step;
            }
            if (overflow && span > 0) {
                ++splits; // walk the two halves of this source node instead
                ++l;
                k <<= 1;
                continue;
            }
            if (overflow) {
                ++lane_walks;
                if (valid) per_lane_walk(q, s, cut, pairs, cap, C);
            } else if (nf > 0) {
                for (int t0 = 0; t0 < nf; t0 += 64) {
                    const int t = t0 + lane;
                    if (t < nf) {
                        const uint32_t kk = front[cur][t];
                        const float* nb = cut.nodes + 6 * ((uint64_t)leaf_base + kk);
                        for (int j = 0; j < 6; ++j) lbox[lane][j] = nb[j];
                        lface[lane] = cut.sface[kk];
                    }
                    __syncthreads();
                    const int cnt = (nf - t0) < 64 ? (nf - t0) : 64;
                    const int groups = 64 >> span, grp = lane >> span;
                    for (int u0 = 0; u0 < cnt; u0 += groups) {
                        if (sn > kTravStage - 64) {
                            flush_pairs(stage, sn, pairs, cap, C);
                            sn = 0;
                        }
                        const int u = u0 + grp; // this lane group's cut leaf
                        bool hit = false;
                        if (u < cnt) {
                            float b[6];
                            for (int j = 0; j < 6; ++j) b[j] = lbox[u][j];
                            hit = rep && boxes_overlap(q, b);
                        }
                        const uint64_t m = __ballot(hit);
                        if (hit) stage[sn + __popcll(m & lt)] = ((uint64_t)s << 32) | lface[u];
                        sn += __popcll(m);
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
            // next source node: climb past right children, then the sibling
            while (l > lw && (k & 1u)) {
                k >>= 1;
                --l;
            }
            if (l == lw) break;
            ++k;
        }
        w += nq ? 1u : gridDim.x;
    }
    if (sn > 0) flush_pairs(stage, sn, pairs, cap, C);
    if (lane == 0 && splits) atomicAdd(&C->n_fallback, splits); // rare: statistics only
    if (lane == 0 && lane_walks) atomicAdd(&C->n_lane_walks, lane_walks);
}

// ===========================================================================
// 6. candidate faces and their planes (kernel.cpp:2184-2356)
// ===========================================================================

// The exact broad-phase test on the shifted boxes (bvh.cpp:661-781 with the
// boxes of bvh.cpp:242-313) applied to the raw-space superset.  A surviving
// pair is appended to the candidate-pair list and the pair hash used for test
// dedup, and each of its faces is flagged once and appended to the candidate
// face list.
// The pair hash is sized from the raw pair count: twice the pairs it can
// receive, rounded up to a power of two, within the allocated table.
__device__ __forceinline__ uint64_t pair_hash_slots(const Counters* C, uint64_t cap, uint64_t table)
{
    const uint64_t n = C->n_raw < cap ? C->n_raw : cap;
    uint64_t m = 1024;
    while (m < 2 * n) m <<= 1;
    return m < table ? m : table;
}

// the whole table, for a dispatch whose stream A is idle before the join: the
// filter then sizes the hash itself (pair_hash_slots) and skips the init
__global__ void k_pair_hash_clear(unsigned long long* __restrict__ phash, uint64_t table)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < table; i += (uint64_t)gridDim.x * blockDim.x)
        phash[i] = kEmpty64;
}

__global__ void k_pair_hash_init(unsigned long long* __restrict__ phash, uint64_t cap, uint64_t table, Counters* C)
{
    const uint64_t m = pair_hash_slots(C, cap, table);
    if (blockIdx.x == 0 && threadIdx.x == 0) C->pmask = m - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        phash[i] = kEmpty64;
}

// The staged-append kernels run 1024-thread blocks with one item per thread:
// every item is a short chain of dependent gathers, so more threads in flight
// shorten the kernel, while the tile stays large enough that the one global
// atomic per tile flush does not serialise the device.
constexpr int kStageBlock = 1024;
constexpr int kFilterItems = 1;
constexpr int kFilterTile = kStageBlock * kFilterItems;

__global__ void __launch_bounds__(kStageBlock) k_filter_pairs(const double* __restrict__ box,
                                                         const unsigned long long* __restrict__ raw, uint64_t cap,
                                                         uint32_t Fs, const uint32_t* __restrict__ fpair,
                                                         unsigned long long* __restrict__ pairs,
                                                         unsigned long long* __restrict__ phash,
                                                         uint32_t* __restrict__ cand, uint32_t* __restrict__ cand_list,
                                                         uint64_t table, Counters* C)
{
    __shared__ unsigned long long st_pair[kFilterTile];
    __shared__ uint32_t st_face[2 * kFilterTile];
    __shared__ uint32_t st_np, st_nf;
    __shared__ unsigned long long st_base;
    const uint64_t n = C->n_raw < cap ? C->n_raw : cap;
    // table > 0: the table was cleared whole before the join, and the hash's
    // size follows the raw pair count here
    const uint64_t pmask = table ? pair_hash_slots(C, cap, table) - 1 : C->pmask;
    if (table && blockIdx.x == 0 && threadIdx.x == 0) C->pmask = pmask;
    for (uint64_t t0 = (uint64_t)blockIdx.x * kFilterTile; t0 < n; t0 += (uint64_t)gridDim.x * kFilterTile) {
        if (threadIdx.x == 0) st_np = st_nf = 0;
        __syncthreads();
        for (int rr = 0; rr < kFilterItems; ++rr) {
            const uint64_t i = t0 + (uint64_t)rr * kStageBlock + threadIdx.x;
            if (i >= n) continue;
            const uint64_t k = raw[i];
            const uint32_t f[2] = {(uint32_t)(k >> 32), Fs + (uint32_t)k};
            if (fpair && fpair[f[0]] != fpair[f[1]]) continue; // batches: faces of different pairs
            double a[6], b[6];
            for (int j = 0; j < 6; ++j) {
                a[j] = box[6 * (uint64_t)f[0] + j];
                b[j] = box[6 * (uint64_t)f[1] + j];
            }
            if (!boxes_overlap(a, b)) continue;
            st_pair[wave_reserve(&st_np)] = k; // one per item: the stage cannot overflow
            hash_insert(phash, pmask, k);
            for (int side = 0; side < 2; ++side)
                if (cand[f[side]] == 0u && atomicExch(&cand[f[side]], 1u) == 0u)
                    st_face[wave_reserve(&st_nf)] = f[side]; // at most two per item
        }
        __syncthreads();
        const uint32_t np = st_np, nfc = st_nf;
        const unsigned long long pb = block_reserve(&C->n_pairs, np, &st_base);
        for (uint32_t q = threadIdx.x; q < np; q += kStageBlock) pairs[pb + q] = st_pair[q]; // pb + np <= n <= cap
        __syncthreads();
        const unsigned long long fb = block_reserve(&C->n_cand, nfc, &st_base);
        for (uint32_t q = threadIdx.x; q < nfc; q += kStageBlock) cand_list[fb + q] = st_face[q]; // one entry per face
        __syncthreads();
    }
}

// triplet-scratch units (16 bytes) of a polygon of more than five vertices:
// its projected vertices, its triplets and the two position lists of the
// partitions (k_big_polygons)
__host__ __device__ inline unsigned long long big_polygon_units(unsigned long long n)
{
    const unsigned long long c3 = n * (n - 1) * (n - 2) / 6;
    return n + 2 * c3; // vertices, triplets, and 2 c3 eight-byte positions
}

// orient2d of projected triplets; the first triplet of greatest |orient2d|
// (std::sort of <= 16 triplets is an insertion sort, which keeps ties in order)
template <typename VertexAt>
__device__ inline bool best_triple_small(int n, VertexAt vat, const double* normal, int k, uint16_t* tri)
{
    double M[6], X[2 * 5];
    projection_matrix(normal, k, M);
    for (int i = 0; i < n; ++i) {
        double p[3];
        vat(i, p);
        project2(M, p, X + 2 * i);
    }
    double best = 0.0;
    bool found = false;
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b)
            for (int c = b + 1; c < n; ++c) {
                const double r = orient2d_adaptive(X + 2 * a, X + 2 * b, X + 2 * c);
                if (r == 0.0) continue;
                if (!found || fabs(r) > best) {
                    best = fabs(r);
                    tri[0] = (uint16_t)a, tri[1] = (uint16_t)b, tri[2] = (uint16_t)c;
                    found = true;
                }
            }
    return found;
}

__global__ void k_planes(Soup S, const uint32_t* __restrict__ cand_list, Plane* __restrict__ planes,
                         uint32_t* __restrict__ big_list, unsigned long long* __restrict__ big_off, uint64_t cap_big,
                         Counters* C)
{
    const uint64_t nc = C->n_cand;
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nc; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t g = cand_list[w];
        const FaceView fv = face_view(S, g);
        Plane pl;
        const int k = plane_coefficients((int)fv.n, fv, pl.n, &pl.d);
        pl.k = k;
        pl.tri[0] = 0, pl.tri[1] = 1, pl.tri[2] = 2;
        pl.big = 0;
        if (k < 0) {
            // kernel.cpp:2320-2331 names the mesh with "face id > source face count"
            if (g > S.nf[0]) atomicOr(&C->bad_cut, 1u);
            else atomicOr(&C->bad_src, 1u);
            pair_flag(S, g, g > S.nf[0] ? PF_BAD_CUT : PF_BAD_SRC);
        } else if (fv.n > 3 && fv.n <= 5) {
            if (!best_triple_small((int)fv.n, fv, pl.n, k, pl.tri)) pl.tri[0] = kCollinear;
        } else if (fv.n > 5 && fv.n <= kMaxPoly) {
            // the projected vertices, every triplet and the partition lists, in the triplet scratch
            pl.big = 1;
            const unsigned long long n = fv.n;
            const unsigned long long at = atomicAdd(&C->n_big, 1ull);
            const unsigned long long off = atomicAdd(&C->n_trip, big_polygon_units(n));
            if (at < cap_big) {
                big_list[at] = g;
                big_off[at] = off;
            } else {
                atomicOr(&C->ov_big, 1u);
            }
        } else if (fv.n > kMaxPoly) {
            atomicOr(&C->unsupported, 1u);
            pair_flag(S, g, PF_UNSUPPORTED);
        }
        planes[g] = pl;
    }
}

// Candidate polygons of more than five vertices, one wave each.  The wave
// lists the polygon's C(n, 3) triplets with non-zero orient2d in the
// reference's (i < j < k) lexicographic order, 64 third vertices a step with
// ballot-ranked appends, in a region of the triplet scratch that k_planes
// reserved (big_polygon_units: the n projected vertices as 16-byte records,
// the triplets, then the two position lists of the partitions), and takes the
// element that the reference's std::sort puts first (trip_first in
// tripsort.hpp: the left-most chain of the introsort's partitions), each
// partition computed a wave at a time in the form of pairs that
// tests/support/tripsort_check.cpp holds to std::__unguarded_partition.  A
// region past the scratch's capacity is left undone and flagged; the host
// grows the scratch and runs the dispatch again.
// The unguarded partition of T[1, l) around the pivot T[0] (std::__unguarded_partition
// after the median-of-three move), by the whole wave.  Let L be the positions in
// [1, l) whose element is not greater than the pivot, ascending, and R those
// in [0, l) whose element is not less than it, descending (the pivot itself
// last).  The sequential partition swaps L[k] with R[k] for k = 0, 1, ... while
// L[k] < R[k]; at the first k = K that fails it returns L[0] if K = 0 and
// otherwise the smaller of L[K] (l if L has no K-th entry) and R[K - 1]
// (a later left scan stops at the element swapped into R[K - 1] at the latest).
// Positions of different pairs before K are distinct, so the swaps run in
// parallel.  Lp and Rp hold the lists (at most l entries each).  Every lane
// returns the cut.
__device__ inline long wave_trip_partition(Trip* T, long l, unsigned long long* Lp, unsigned long long* Rp)
{
    const int lane = (int)(threadIdx.x & 63);
    const unsigned long long below = (1ull << lane) - 1ull;
    const Trip P = T[0];
    long nL = 0, nR = 0;
    for (long p0 = 1; p0 < l; p0 += 64) {
        const long p = p0 + lane;
        const bool s = p < l && !trip_less(T[p], P);
        const unsigned long long m = __ballot(s);
        if (s) Lp[nL + __popcll(m & below)] = (unsigned long long)p;
        nL += __popcll(m);
    }
    for (long p0 = l - 1; p0 >= 0; p0 -= 64) {
        const long p = p0 - lane;
        const bool s = p >= 0 && !trip_less(P, T[p]);
        const unsigned long long m = __ballot(s);
        if (s) Rp[nR + __popcll(m & below)] = (unsigned long long)p;
        nR += __popcll(m);
    }
    __syncthreads(); // the lists, written lane by lane, are read across the wave
    const long np = nL < nR ? nL : nR;
    long K = np;
    for (long k0 = 0; k0 < np; k0 += 64) {
        const long k = k0 + lane;
        const bool fail = k < np && !(Lp[k] < Rp[k]);
        const unsigned long long m = __ballot(fail);
        if (m) {
            K = k0 + __ffsll((long long)m) - 1;
            break;
        }
    }
    for (long k = lane; k < K; k += 64) trip_swap(T + Lp[k], T + Rp[k]);
    long cut;
    if (K == 0) {
        cut = (long)Lp[0];
    } else {
        const long lK = K < nL ? (long)Lp[K] : l;
        const long rK = (long)Rp[K - 1];
        cut = lK < rK ? lK : rK;
    }
    __syncthreads(); // the swaps land before the next partition reads the range
    return cut;
}

// trip_first over T[0, m) by the whole wave: lane 0 moves the median to the
// front and, when the depth budget runs out, heap sorts the range; the
// partitions are the wave's
__device__ inline Trip wave_trip_first(Trip* T, long m, unsigned long long* Lp, unsigned long long* Rp)
{
    const int lane = (int)(threadIdx.x & 63);
    long l = m;
    int depth = trip_depth_budget(m);
    while (l > 16) {
        if (depth == 0) {
            if (lane == 0) trip_heap_sort(T, T + l);
            __syncthreads();
            return T[0];
        }
        --depth;
        if (lane == 0) trip_median_to_first(T, T + l);
        __syncthreads();
        l = wave_trip_partition(T, l, Lp, Rp);
    }
    Trip best = T[0];
    for (long p = 1; p < l; ++p)
        if (trip_less(T[p], best)) best = T[p];
    return best;
}

__global__ void __launch_bounds__(64) k_big_polygons(Soup S, const uint32_t* __restrict__ big_list,
                                                     const unsigned long long* __restrict__ big_off,
                                                     const Counters* __restrict__ C, uint64_t cap_big, uint64_t cap_trip,
                                                     Plane* __restrict__ planes, Trip* __restrict__ scratch, Counters* Cw)
{
    const int lane = (int)threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint64_t nb = C->n_big < cap_big ? C->n_big : cap_big;
    for (uint64_t w = blockIdx.x; w < nb; w += gridDim.x) {
        const uint32_t g = big_list[w];
        const FaceView fv = face_view(S, g);
        const uint64_t n = fv.n;
        const uint64_t off = big_off[w];
        if (off + big_polygon_units(n) > cap_trip) {
            if (lane == 0) atomicOr(&Cw->ov_trip, 1u);
            continue;
        }
        Plane pl = planes[g];
        double M[6];
        projection_matrix(pl.n, pl.k, M);
        double* X = reinterpret_cast<double*>(scratch + off); // projected vertex i: X[2i], X[2i + 1]
        for (uint32_t i = lane; i < fv.n; i += 64) {
            double p[3];
            fv((int)i, p);
            project2(M, p, X + 2 * i);
        }
        __syncthreads();
        const uint64_t c3 = n * (n - 1) * (n - 2) / 6;
        Trip* T = scratch + off + n;
        unsigned long long* Lp = reinterpret_cast<unsigned long long*>(scratch + off + n + c3);
        unsigned long long* Rp = Lp + c3;
        long m = 0;
        for (int a = 0; a < (int)n; ++a)
            for (int b = a + 1; b < (int)n; ++b)
                for (int c0 = b + 1; c0 < (int)n; c0 += 64) {
                    const int c = c0 + lane;
                    double r = 0.0;
                    if (c < (int)n) r = orient2d_adaptive(X + 2 * a, X + 2 * b, X + 2 * c);
                    const bool keep = r != 0.0;
                    const unsigned long long mk = __ballot(keep);
                    if (keep) {
                        Trip t;
                        t.a = fabs(r);
                        t.ijk = (uint64_t)a | ((uint64_t)b << 16) | ((uint64_t)c << 32);
                        T[m + __popcll(mk & below)] = t;
                    }
                    m += __popcll(mk);
                }
        __syncthreads();
        if (m == 0) {
            pl.tri[0] = kCollinear;
        } else {
            const Trip best = wave_trip_first(T, m, Lp, Rp);
            pl.tri[0] = (uint16_t)(best.ijk & 0xffff);
            pl.tri[1] = (uint16_t)((best.ijk >> 16) & 0xffff);
            pl.tri[2] = (uint16_t)((best.ijk >> 32) & 0xffff);
        }
        if (lane == 0) planes[g] = pl;
        __syncthreads();
    }
}

// ===========================================================================
// 7. edge-face tests: expansion, dedup and cull (kernel.cpp:1779-2177)
//
// The reference tests edge e against every face paired with a face of e.  A
// test (e, g) is emitted here by the lowest-numbered face of e that is paired
// with g, so each test appears once without a global sort; the cull keeps it
// only when e's box (endpoints, never enlarged) overlaps g's box.
// ===========================================================================

__device__ __forceinline__ bool pair_exists(const unsigned long long* phash, uint64_t pmask, uint32_t s, uint32_t c)
{
    return hash_find(phash, pmask, ((uint64_t)s << 32) | c) >= 0;
}

__device__ __forceinline__ bool edge_face_cull(const Soup& S, uint32_t e, uint32_t g)
{
    const double* p0 = S.ps_v + 3 * (uint64_t)S.ev1[e]; // vertex(e, 0) = target(h0)
    const double* p1 = S.ps_v + 3 * (uint64_t)S.ev0[e];
    double eb[6];
    for (int j = 0; j < 3; ++j) {
        double lo = 1.7976931348623157e308, hi = -1.7976931348623157e308;
        hi = (hi < p0[j]) ? p0[j] : hi;
        lo = (p0[j] < lo) ? p0[j] : lo;
        hi = (hi < p1[j]) ? p1[j] : hi;
        lo = (p1[j] < lo) ? p1[j] : lo;
        eb[j] = lo;
        eb[3 + j] = hi;
    }
    double fb[6];
    for (int j = 0; j < 6; ++j) fb[j] = S.box[6 * (uint64_t)g + j];
    return boxes_overlap(eb, fb);
}

// six work items per pair: (side, corner slot); a polygon's corners beyond
// the third are taken by the same slot in steps of three.  A block takes a
// tile of kTestTile items and stages the tests it emits in LDS.
constexpr int kTestItems = 1;
constexpr int kTestTile = kStageBlock * kTestItems;
constexpr int kTestStage = 2048;

__global__ void __launch_bounds__(kStageBlock) k_tests(Soup S, const unsigned long long* __restrict__ pairs, uint64_t cap_pairs,
                                                  const unsigned long long* __restrict__ phash,
                                                  uint32_t* __restrict__ te, uint32_t* __restrict__ tg, uint64_t cap_tests,
                                                  Counters* C)
{
    __shared__ uint32_t st_e[kTestStage], st_g[kTestStage];
    __shared__ uint32_t st_n;
    __shared__ unsigned long long st_base;
    const uint64_t n = C->n_pairs < cap_pairs ? C->n_pairs : cap_pairs;
    const uint64_t pmask = C->pmask;
    const uint32_t Fs = S.nf[0];
    const uint64_t items = 6 * n;
    for (uint64_t t0 = (uint64_t)blockIdx.x * kTestTile; t0 < items; t0 += (uint64_t)gridDim.x * kTestTile) {
        if (threadIdx.x == 0) st_n = 0;
        __syncthreads();
        for (int rr = 0; rr < kTestItems; ++rr) {
            const uint64_t w = t0 + (uint64_t)rr * kStageBlock + threadIdx.x;
            if (w >= items) continue;
            const uint64_t i = w / 6;
            const uint32_t slot = (uint32_t)(w - 6 * i);
            const int side = slot >= 3;
            const uint32_t j0 = side ? slot - 3 : slot;
            const uint64_t key = pairs[i];
            const uint32_t s = (uint32_t)(key >> 32), c = (uint32_t)key;
            const uint32_t f = side ? c : s;         // face whose edges are tested
            const uint32_t self = side ? Fs + c : s; // its ps id
            const uint32_t g = side ? s : Fs + c;    // the other face (ps id)
            const uint32_t b = S.off[side][f], nn = S.off[side][f + 1] - b;
            for (uint32_t k = j0; k < nn; k += 3) {
                const uint32_t e = S.cedge[side][b + k];
                const int32_t f0 = S.ef0[e], f1 = S.ef1[e];
                const int32_t fo = (f0 == (int32_t)self) ? f1 : f0;
                // the lowest-numbered face of e paired with g emits the test
                if (fo >= 0 && (uint32_t)fo < self) {
                    const bool paired = side ? pair_exists(phash, pmask, s, (uint32_t)fo - Fs)
                                             : pair_exists(phash, pmask, (uint32_t)fo, c);
                    if (paired) continue;
                }
                if (!edge_face_cull(S, e, g)) continue;
                const uint32_t at = wave_reserve(&st_n);
                if (at < (uint32_t)kTestStage) {
                    st_e[at] = e;
                    st_g[at] = g;
                } else { // more than the stage holds (polygons): append directly
                    const unsigned long long ga = wave_append(&C->n_tests);
                    if (ga < cap_tests) {
                        te[ga] = e;
                        tg[ga] = g;
                    } else {
                        atomicOr(&C->ov_tests, 1u);
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t m = st_n < (uint32_t)kTestStage ? st_n : (uint32_t)kTestStage;
        const unsigned long long base = block_reserve(&C->n_tests, m, &st_base);
        for (uint32_t q = threadIdx.x; q < m; q += kStageBlock) {
            const unsigned long long at = base + q;
            if (at < cap_tests) {
                te[at] = st_e[q];
                tg[at] = st_g[q];
            } else {
                atomicOr(&C->ov_tests, 1u);
            }
        }
        __syncthreads();
    }
}

// ===========================================================================
// 8. narrow phase (kernel.cpp:2456-2658)
// ===========================================================================

struct TestOut {
    int8_t* qs;
    int8_t* rs;
    char* type;
    char* pip;
};

struct RegOut {
    unsigned long long* key; // sort key (reg_key)
    double* p;
    uint32_t* val;           // sort payload: the entry's own index
    uint32_t* e;             // edge and face of each entry (ps ids)
    uint32_t* g;
    uint64_t cap;
    int gbits;
    const BatchPair* bp;     // batches: pair ranges (the key orders pairs first)
};

// Registry order: by (edge, face).  In a batch the soup holds all source
// meshes before all cut meshes, so the key uses each id's rank in the
// pair-major order [pair 0 source, pair 0 cut, pair 1 source, ...]: one sort
// then leaves every pair's entries together and in the order of that pair's
// own dispatch (both maps are increasing within a pair).
__device__ __forceinline__ unsigned long long reg_key(const RegOut& R, const Soup& S, uint32_t e, uint32_t g, uint32_t Es)
{
    if (!R.bp) return ((unsigned long long)e << R.gbits) | g;
    const BatchPair b = R.bp[S.fpair[g]];
    const uint32_t rg = g < S.nf[0] ? g + b.fc0 : (g - S.nf[0]) + b.fs0 + b.nfs;
    const uint32_t re = e < Es ? e + b.ec0 : (e - Es) + b.es0 + b.nes;
    return ((unsigned long long)re << R.gbits) | rg;
}

__device__ __forceinline__ void store_registry(const RegOut& R, unsigned long long at, unsigned long long key,
                                               uint32_t e, uint32_t g, const double* p, Counters* C)
{
    if (at < R.cap) {
        R.key[at] = key;
        R.p[3 * at] = p[0];
        R.p[3 * at + 1] = p[1];
        R.p[3 * at + 2] = p[2];
        R.val[at] = (uint32_t)at;
        R.e[at] = e;
        R.g[at] = g;
    } else {
        atomicOr(&C->ov_reg, 1u);
    }
}

// the decision after the two orient3d signs are known; true when (e, g)
// enters the registry with intersection point p
__device__ inline bool resolve_test(const Soup& S, const Plane* planes, uint32_t e, uint32_t g, int sq, int sr,
                                    uint64_t ti, const TestOut& TO, Counters* C, double* p)
{
    const char ty = segment_type(sq, sr);
    if (TO.qs) {
        TO.qs[ti] = (int8_t)sq;
        TO.rs[ti] = (int8_t)sr;
        TO.type[ti] = ty;
        TO.pip[ti] = '-';
    }
    if (ty == '0') return false;
    const Plane pl = planes[g];
    const FaceView fv = face_view(S, g);
    const double* qv = S.ps_v + 3 * (uint64_t)S.ev0[e]; // source(h0)
    const double* rv = S.ps_v + 3 * (uint64_t)S.ev1[e]; // target(h0)
    double q[3], r[3];
    load3(qv, q);
    load3(rv, r);
    if (ty != '1') {
        // an endpoint touches the plane: general position is violated when it
        // lies inside or on the boundary of the face (kernel.cpp:2519-2556)
        bool violated = false;
        char res = '-';
        if (ty == 'q' || ty == 'p') {
            res = point_in_polygon(q, (int)fv.n, fv, pl.n, pl.k);
            violated = res == 'i' || res == 'v' || res == 'e';
        } else {
            res = point_in_polygon(r, (int)fv.n, fv, pl.n, pl.k);
            violated = res == 'i' || res == 'v' || res == 'e';
        }
        if (ty == 'p' && !violated) {
            res = point_in_polygon(r, (int)fv.n, fv, pl.n, pl.k);
            violated = res == 'i' || res == 'v' || res == 'e';
        }
        if (TO.qs) TO.pip[ti] = res;
        if (violated) {
            atomicOr(&C->gp, 1u);
            pair_flag(S, g, PF_GP);
        }
        return false;
    }
    p[0] = p[1] = p[2] = 0.0;
    segment_point(pl.n, pl.d, q, r, p);
    const char res = point_in_polygon(p, (int)fv.n, fv, pl.n, pl.k);
    if (TO.qs) TO.pip[ti] = res;
    if (res == 'v' || res == 'e') {
        atomicOr(&C->gp, 1u);
        pair_flag(S, g, PF_GP);
        return false;
    }
    return res == 'i';
}

// A block takes a tile of kNarrowTile tests; registry entries and escalated
// tests (at most one of either per test) are staged in LDS.
constexpr int kNarrowItems = 1;
constexpr int kNarrowTile = kStageBlock * kNarrowItems;

__global__ void __launch_bounds__(kStageBlock) k_narrow(Soup S, const Plane* __restrict__ planes, const uint32_t* __restrict__ te,
                                                   const uint32_t* __restrict__ tg, uint64_t cap_tests,
                                                   uint32_t* __restrict__ esc, uint64_t cap_esc, TestOut TO, RegOut R,
                                                   Counters* C)
{
    __shared__ uint32_t st_e[kNarrowTile], st_g[kNarrowTile];
    __shared__ double st_p[3 * kNarrowTile];
    __shared__ uint32_t st_esc[kNarrowTile];
    __shared__ uint32_t st_nr, st_ne, st_n32, st_n64;
    __shared__ unsigned long long st_base;
    const uint64_t n = C->n_tests < cap_tests ? C->n_tests : cap_tests;
    const uint32_t Es = (uint32_t)C->n_edges_src;
    if (threadIdx.x == 0) st_n32 = st_n64 = 0;
    for (uint64_t t0 = (uint64_t)blockIdx.x * kNarrowTile; t0 < n; t0 += (uint64_t)gridDim.x * kNarrowTile) {
        if (threadIdx.x == 0) st_nr = st_ne = 0;
        __syncthreads();
        for (int rr = 0; rr < kNarrowItems; ++rr) {
            const uint64_t i = t0 + (uint64_t)rr * kStageBlock + threadIdx.x;
            if (i >= n) continue;
            const uint32_t e = te[i], g = tg[i];
            const Plane pl = planes[g];
            if (pl.tri[0] == kCollinear || pl.k < 0) {
                // all polygon vertices collinear: the type is '0' (math.cpp:404-410)
                if (TO.qs) {
                    TO.qs[i] = 0;
                    TO.rs[i] = 0;
                    TO.type[i] = '0';
                    TO.pip[i] = '-';
                }
                continue;
            }
            const FaceView fv = face_view(S, g);
            double a[3], b[3], c[3], q[3], r[3];
            fv(pl.tri[0], a);
            fv(pl.tri[1], b);
            fv(pl.tri[2], c);
            load3(S.ps_v + 3 * (uint64_t)S.ev0[e], q);
            load3(S.ps_v + 3 * (uint64_t)S.ev1[e], r);
            int stq, str_;
            const int sq = orient3d_staged(a, b, c, q, stq);
            const int sr = orient3d_staged(a, b, c, r, str_);
            if (sq == 0 || sr == 0) {
                st_esc[wave_reserve(&st_ne)] = (uint32_t)i;
                continue;
            }
            // the test's stage: the later of its two signs' stages
            if ((stq | str_) == 0) wave_reserve(&st_n32);
            else wave_reserve(&st_n64);
            double p[3];
            if (resolve_test(S, planes, e, g, sq, sr, i, TO, C, p)) {
                const uint32_t at = wave_reserve(&st_nr);
                st_e[at] = e;
                st_g[at] = g;
                st_p[3 * at] = p[0];
                st_p[3 * at + 1] = p[1];
                st_p[3 * at + 2] = p[2];
            }
        }
        __syncthreads();
        const uint32_t nr = st_nr, ne = st_ne;
        const unsigned long long rb = block_reserve(&C->n_reg, nr, &st_base);
        for (uint32_t q = threadIdx.x; q < nr; q += kStageBlock)
            store_registry(R, rb + q, reg_key(R, S, st_e[q], st_g[q], Es), st_e[q], st_g[q], st_p + 3 * q, C);
        __syncthreads();
        const unsigned long long eb = block_reserve(&C->n_esc, ne, &st_base);
        for (uint32_t q = threadIdx.x; q < ne; q += kStageBlock) {
            if (eb + q < cap_esc) esc[eb + q] = st_esc[q];
            else atomicOr(&C->ov_esc, 1u);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && st_n32) atomicAdd(&C->n_f32, (unsigned long long)st_n32);
    if (threadIdx.x == 0 && st_n64) atomicAdd(&C->n_f64, (unsigned long long)st_n64);
}

// tests whose filter was inconclusive: exact signs, then the same decision
__global__ void __launch_bounds__(64) k_narrow_exact(Soup S, const Plane* __restrict__ planes, const uint32_t* __restrict__ te,
                                                     const uint32_t* __restrict__ tg, const uint32_t* __restrict__ esc,
                                                     uint64_t cap_esc, TestOut TO, RegOut R, Counters* C)
{
    const uint64_t n = C->n_esc < cap_esc ? C->n_esc : cap_esc;
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < n; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = esc[w];
        const uint32_t e = te[i], g = tg[i];
        const Plane pl = planes[g];
        const FaceView fv = face_view(S, g);
        double a[3], b[3], c[3], q[3], r[3];
        fv(pl.tri[0], a);
        fv(pl.tri[1], b);
        fv(pl.tri[2], c);
        load3(S.ps_v + 3 * (uint64_t)S.ev0[e], q);
        load3(S.ps_v + 3 * (uint64_t)S.ev1[e], r);
        int stq, str_;
        int sq = orient3d_staged(a, b, c, q, stq);
        int sr = orient3d_staged(a, b, c, r, str_);
        if (sq == 0) sq = orient3d_exact_sign(a, b, c, q);
        if (sr == 0) sr = orient3d_exact_sign(a, b, c, r);
        double p[3];
        if (resolve_test(S, planes, e, g, sq, sr, i, TO, C, p))
            store_registry(R, wave_append(&C->n_reg), reg_key(R, S, e, g, (uint32_t)C->n_edges_src), e, g, p, C);
    }
}

// The predicate stage alone: both orient3d signs of every edge-face test
// through the staged filter (FP32, FP64, exact), against the face's plane
// triplet, or its first three vertices where the plane has none (a
// degenerate or collinear face, which the narrow phase does not test).
// counts: [decided in FP32, in FP64, exactly, negative, zero, positive] over
// the 2 T predicates; qs / rs (optional) receive the signs per test.
__global__ void __launch_bounds__(256) k_predicates(Soup S, const Plane* __restrict__ planes, const uint32_t* __restrict__ te,
                                                    const uint32_t* __restrict__ tg, uint64_t cap_tests,
                                                    const Counters* __restrict__ C, int8_t* __restrict__ qs,
                                                    int8_t* __restrict__ rs, unsigned long long* __restrict__ counts)
{
    __shared__ unsigned int cnt[6];
    if (threadIdx.x < 6) cnt[threadIdx.x] = 0u;
    __syncthreads();
    unsigned int mine[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    const uint64_t n = C->n_tests < cap_tests ? C->n_tests : cap_tests;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = te[i], g = tg[i];
        const Plane pl = planes[g];
        const bool tri_ok = pl.k >= 0 && pl.tri[0] != kCollinear;
        const FaceView fv = face_view(S, g);
        double a[3], b[3], c[3], x[2][3];
        fv(tri_ok ? pl.tri[0] : 0, a);
        fv(tri_ok ? pl.tri[1] : 1, b);
        fv(tri_ok ? pl.tri[2] : 2, c);
        load3(S.ps_v + 3 * (uint64_t)S.ev0[e], x[0]);
        load3(S.ps_v + 3 * (uint64_t)S.ev1[e], x[1]);
        int sg[2];
        for (int k = 0; k < 2; ++k) {
            int stage;
            int sgn = orient3d_staged(a, b, c, x[k], stage);
            if (stage == 2) sgn = orient3d_exact_sign(a, b, c, x[k]);
            ++mine[stage];
            ++mine[4 + sgn];
            sg[k] = sgn;
        }
        if (qs) {
            qs[i] = (int8_t)sg[0];
            rs[i] = (int8_t)sg[1];
        }
    }
    for (int k = 0; k < 6; ++k)
        if (mine[k]) atomicAdd(&cnt[k], mine[k]);
    __syncthreads();
    if (threadIdx.x < 6 && cnt[threadIdx.x]) atomicAdd(&counts[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

// ===========================================================================
// 9. registry: sorted order, then the status rules that read it
// ===========================================================================

// cutpath_edge_creation_info keys (kernel.cpp:2620-2636): every (source face,
// cut face) pair joined by an intersection point needs at least two points
// (kernel.cpp:3439-3452); a cut-mesh border edge hit marks a partial cut
// (kernel.cpp:2646-2650)
// status hash: 4 slots per registry entry, the power of two above 4 n
__device__ __forceinline__ uint64_t status_slots(uint64_t n)
{
    uint64_t p = 64;
    while (p < 4 * n) p <<= 1;
    return p;
}

// counts one registry entry (e, g) on the face pairs it joins
__device__ __forceinline__ void status_insert(const Soup& S, uint32_t e, uint32_t g, unsigned long long* hkeys,
                                              uint32_t* hcnt, uint64_t mask, Counters* C)
{
    const uint32_t Fs = S.nf[0], Vs = S.nv[0];
    const int32_t f0 = S.ef0[e], f1 = S.ef1[e];
    const bool on_cut = f0 >= (int32_t)Fs;
    uint64_t k0, k1 = kEmpty64;
    if (on_cut) {
        k0 = ((uint64_t)g << 32) | (uint32_t)f0;
        if (f1 >= 0) k1 = ((uint64_t)g << 32) | (uint32_t)f1;
    } else {
        k0 = ((uint64_t)(uint32_t)f0 << 32) | g;
        if (f1 >= 0) k1 = ((uint64_t)(uint32_t)f1 << 32) | g;
    }
    atomicAdd(&hcnt[hash_insert(hkeys, mask, k0)], 1u);
    if (k1 != kEmpty64) atomicAdd(&hcnt[hash_insert(hkeys, mask, k1)], 1u);
    if (S.ev0[e] >= Vs && (f0 < 0 || f1 < 0)) {
        if (!C->partial) atomicOr(&C->partial, 1u);
        pair_flag(S, g, PF_PARTIAL);
    }
}

// Registry order for up to kRegMergeTiles tiles of kRegTile entries, in two
// launches instead of a multi-pass radix sort.  k_reg_tile_sort sorts each
// tile in LDS and clears the status hash; k_reg_merge ranks every entry by its
// place in its own tile plus binary searches in the other tiles (registry keys
// are distinct), writes (edge, face, point) to that rank and counts the
// status face pairs.
//
// The tile sort is an LSD radix sort over the key's kbits bits, 8 bits a
// pass.  Element i of a tile sits in thread i % 1024 as its row i / 1024, so a
// row of one wave holds 64 consecutive elements.  A pass ranks each row of a
// wave by ballots, turns the per-(row, wave, digit) counts into starting
// offsets (rows, then waves, after all smaller digits) and scatters the tile
// through LDS.  That is about 50 instructions per element and pass; a bitonic
// network spent 78 compare-exchange stages on the 4096 elements of one CU and
// took 61 us even for a registry of a few hundred entries.
constexpr int kRegTile = 4096;
constexpr int kRegSortThreads = 1024;
constexpr int kRegPer = kRegTile / kRegSortThreads; // 4
constexpr int kRegWaves = kRegSortThreads / 64;     // 16
constexpr uint64_t kRegMergeTiles = 64;

__global__ void __launch_bounds__(kRegSortThreads) k_reg_tile_sort(const unsigned long long* __restrict__ key,
                                                                   const uint32_t* __restrict__ val,
                                                                   const unsigned long long* __restrict__ n_dev,
                                                                   uint64_t cap, unsigned long long* __restrict__ skey,
                                                                   uint32_t* __restrict__ sval,
                                                                   unsigned long long* __restrict__ hkeys,
                                                                   uint32_t* __restrict__ hcnt, uint64_t max_slots, int kbits)
{
    __shared__ unsigned long long sk[kRegTile];
    __shared__ uint32_t sv[kRegTile];
    __shared__ uint16_t pre[kRegPer][kRegWaves][kRadix]; // wave-row digit counts, then their starting offsets
    __shared__ uint32_t rowtot[kRegPer][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t wsum[4];
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    uint64_t slots = status_slots(n);
    if (slots > max_slots) slots = max_slots;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slots; i += (uint64_t)gridDim.x * blockDim.x) {
        hkeys[i] = kEmpty64;
        hcnt[i] = 0u;
    }
    const uint64_t t0 = (uint64_t)blockIdx.x * kRegTile;
    if (t0 >= n) return; // whole block
    const uint32_t cnt = (n - t0) < (uint64_t)kRegTile ? (uint32_t)(n - t0) : (uint32_t)kRegTile;
    const uint32_t t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    unsigned long long K[kRegPer];
    uint32_t V[kRegPer];
#pragma unroll
    for (int r = 0; r < kRegPer; ++r) {
        const uint32_t i = r * kRegSortThreads + t;
        K[r] = i < cnt ? key[t0 + i] : ~0ull; // padding: all digits 0xff, after the real keys (stable)
        V[r] = i < cnt ? val[t0 + i] : 0u;
    }
    // passes over the bits the tile's keys use (at most kbits): the padding's
    // all-ones low bits still sort it last
    __shared__ unsigned long long wor[kRegWaves];
    {
        unsigned long long o = 0;
#pragma unroll
        for (int r = 0; r < kRegPer; ++r)
            if (r * kRegSortThreads + t < cnt) o |= K[r];
        for (int q = 32; q > 0; q >>= 1) o |= __shfl_xor(o, q, 64);
        if (lane == 0) wor[wid] = o;
    }
    __syncthreads();
    unsigned long long all = 0;
#pragma unroll
    for (int w = 0; w < kRegWaves; ++w) all |= wor[w];
    const int used = all ? 64 - __clzll((long long)all) : 1;
    const int nbits = used < kbits ? used : kbits;
    // the prefix step: thread t sums digit dg over the waves of row pr
    const uint32_t dg = t & (kRadix - 1), pr = t >> kRadixBits;
    for (int shift = 0; shift < nbits; shift += kRadixBits) {
        for (uint32_t q = t; q < (uint32_t)(kRegPer * kRegWaves * kRadix); q += kRegSortThreads) (&pre[0][0][0])[q] = 0;
        __syncthreads();
        uint32_t d[kRegPer], rk[kRegPer];
#pragma unroll
        for (int r = 0; r < kRegPer; ++r) {
            d[r] = digit_of(K[r], shift);
            uint64_t peers = ~0ull;
#pragma unroll
            for (int b = 0; b < kRadixBits; ++b) {
                const bool bit = (d[r] >> b) & 1u;
                const uint64_t m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            rk[r] = (uint32_t)__popcll(peers & lt);
            if (rk[r] == 0) pre[r][wid][d[r]] = (uint16_t)__popcll(peers);
        }
        __syncthreads();
        uint32_t c[kRegWaves], rs = 0;
#pragma unroll
        for (int w = 0; w < kRegWaves; ++w) {
            c[w] = pre[pr][w][dg];
            rs += c[w];
        }
        rowtot[pr][dg] = rs;
        __syncthreads();
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < kRegPer; ++q) {
            const uint32_t x = rowtot[q][dg];
            if (q < (int)pr) before += x;
            tot += x;
        }
        uint32_t x = tot; // exclusive scan of the digit totals by the four waves of row 0
        if (pr == 0) {
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if (lane >= (uint32_t)o) x += y;
            }
            if (lane == 63) wsum[wid] = x;
        }
        __syncthreads();
        if (pr == 0) {
            uint32_t off = 0;
            for (uint32_t w = 0; w < wid; ++w) off += wsum[w];
            dstart[dg] = off + x - tot;
        }
        __syncthreads();
        uint32_t run = dstart[dg] + before;
#pragma unroll
        for (int w = 0; w < kRegWaves; ++w) {
            pre[pr][w][dg] = (uint16_t)run;
            run += c[w];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRegPer; ++r) {
            const uint32_t pos = pre[r][wid][d[r]] + rk[r];
            sk[pos] = K[r];
            sv[pos] = V[r];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRegPer; ++r) {
            K[r] = sk[r * kRegSortThreads + t];
            V[r] = sv[r * kRegSortThreads + t];
        }
    }
#pragma unroll
    for (int r = 0; r < kRegPer; ++r) {
        const uint32_t i = r * kRegSortThreads + t;
        if (i < cnt) {
            skey[t0 + i] = K[r];
            sval[t0 + i] = V[r];
        }
    }
}

// first index of tile u (sorted, cnt_u entries at skey + lo_u) whose key is not
// below k, for up to 8 tiles at once so that their loads overlap
constexpr int kRegMergeILP = 8;

__global__ void k_reg_merge(Soup S, const unsigned long long* __restrict__ skey, const uint32_t* __restrict__ sval,
                            const unsigned long long* __restrict__ n_dev, uint64_t cap, const uint32_t* __restrict__ e_in,
                            const uint32_t* __restrict__ g_in, const double* __restrict__ p_in,
                            uint32_t* __restrict__ re, uint32_t* __restrict__ rg,
                            double* __restrict__ rp, unsigned long long* __restrict__ hkeys, uint32_t* __restrict__ hcnt,
                            uint64_t max_slots, Counters* C)
{
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    const uint64_t ntiles = (n + kRegTile - 1) / kRegTile;
    uint64_t slots = status_slots(n);
    if (slots > max_slots) slots = max_slots;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t t = i / kRegTile;
        const unsigned long long k = skey[i];
        uint64_t rank = i - t * kRegTile;
        for (uint64_t u0 = 0; u0 < ntiles; u0 += kRegMergeILP) {
            uint32_t lo[kRegMergeILP], hi[kRegMergeILP];
#pragma unroll
            for (int q = 0; q < kRegMergeILP; ++q) {
                const uint64_t u = u0 + q;
                const uint64_t base = u * kRegTile;
                lo[q] = 0;
                hi[q] = (u < ntiles && u != t) ? (uint32_t)((n - base) < (uint64_t)kRegTile ? (n - base) : kRegTile) : 0u;
            }
            for (int step = 0; step < 13; ++step) { // 13 halvings cover 4096 entries
#pragma unroll
                for (int q = 0; q < kRegMergeILP; ++q) {
                    if (lo[q] < hi[q]) {
                        const uint32_t mid = (lo[q] + hi[q]) >> 1;
                        if (skey[(u0 + q) * kRegTile + mid] < k) lo[q] = mid + 1;
                        else hi[q] = mid;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < kRegMergeILP; ++q) rank += lo[q];
        }
        const uint32_t j = sval[i];
        const uint32_t e = e_in[j], g = g_in[j];
        re[rank] = e;
        rg[rank] = g;
        const double* p = p_in + 3 * (uint64_t)j;
        rp[3 * rank] = p[0];
        rp[3 * rank + 1] = p[1];
        rp[3 * rank + 2] = p[2];
        status_insert(S, e, g, hkeys, hcnt, slots - 1, C);
    }
}

__global__ void k_clamp_count(const unsigned long long* n_dev, uint64_t cap, unsigned long long* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = *n_dev < cap ? *n_dev : cap;
}

// sort permutation -> (edge, face, point) arrays in registry order
__global__ void k_registry_gather(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ e_in,
                                  const uint32_t* __restrict__ g_in, const double* __restrict__ p_in,
                                  const unsigned long long* __restrict__ n_dev, uint64_t cap, uint32_t* __restrict__ re,
                                  uint32_t* __restrict__ rg, double* __restrict__ rp)
{
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t j = perm[i];
        re[i] = e_in[j];
        rg[i] = g_in[j];
        rp[3 * i] = p_in[3 * (uint64_t)j];
        rp[3 * i + 1] = p_in[3 * (uint64_t)j + 1];
        rp[3 * i + 2] = p_in[3 * (uint64_t)j + 2];
    }
}


__global__ void k_status_clear(const unsigned long long* __restrict__ n_dev, unsigned long long* __restrict__ hkeys,
                               uint32_t* __restrict__ hcnt, uint64_t max_slots)
{
    uint64_t slots = status_slots(*n_dev);
    if (slots > max_slots) slots = max_slots;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slots; i += (uint64_t)gridDim.x * blockDim.x) {
        hkeys[i] = kEmpty64;
        hcnt[i] = 0u;
    }
}

__global__ void k_status_pairs(Soup S, const uint32_t* __restrict__ re, const uint32_t* __restrict__ rg,
                               const unsigned long long* __restrict__ n_dev, uint64_t max_slots,
                               unsigned long long* __restrict__ hkeys, uint32_t* __restrict__ hcnt, Counters* C)
{
    const uint64_t n = *n_dev;
    uint64_t slots = status_slots(n);
    if (slots > max_slots) slots = max_slots;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        status_insert(S, re[i], rg[i], hkeys, hcnt, slots - 1, C);
}

__global__ void k_status_check(Soup S, const unsigned long long* __restrict__ n_dev, const unsigned long long* __restrict__ hkeys,
                               const uint32_t* __restrict__ hcnt, uint64_t max_slots, Counters* C)
{
    uint64_t slots = status_slots(*n_dev);
    if (slots > max_slots) slots = max_slots;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slots; i += (uint64_t)gridDim.x * blockDim.x)
        if (hkeys[i] != kEmpty64 && hcnt[i] < 2) {
            atomicOr(&C->pair_fail, 1u);
            pair_flag(S, (uint32_t)(hkeys[i] >> 32), PF_PAIR_FAIL); // the key's source face
        }
}

// ===========================================================================
// 10. batches: pair tables, input, raw boxes, layout, per-pair results
// ===========================================================================

// out[base + i] = the segment of element i, starts[] being the first element
// of each of the nseg segments (increasing)
__global__ void k_segment_ids(const uint32_t* __restrict__ starts, uint32_t nseg, uint32_t n, uint32_t* __restrict__ out)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nseg; // last segment whose start is <= i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (starts[mid] <= (uint32_t)i) lo = mid;
            else hi = mid;
        }
        out[i] = lo;
    }
}

// Face lists of a batch mesh part: checks every pair's faces against its own
// vertex count (as k_validate does for one mesh) and writes the indices as
// ids of the concatenated part.
__global__ void k_batch_globalize(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ sizes, uint32_t nf, int m, uint32_t fofs,
                                  const uint32_t* __restrict__ fpair, const BatchPair* __restrict__ bp,
                                  uint32_t* __restrict__ out, Counters* C)
{
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        const BatchPair b = bp[fpair[fofs + f]];
        const uint32_t nv = m ? b.nvc : b.nvs, v0 = m ? b.vc0 : b.vs0;
        const uint32_t o = off[f], n = sizes ? sizes[f] : 3u;
        bool bad = n < 3;
        for (uint32_t i = 0; i < n && !bad; ++i) {
            const uint32_t a = idx[o + i];
            bad = a >= nv;
            for (uint32_t k = 0; k < i && !bad; ++k) bad = idx[o + k] == a;
            out[o + i] = v0 + a;
        }
        if (bad) atomicOr(&C->bad_input, 1u);
    }
}

// raw box of every (pair, mesh) segment: one block each, order keys out
__global__ void __launch_bounds__(256) k_batch_bbox(const double* __restrict__ vs, const double* __restrict__ vc,
                                                    const BatchPair* __restrict__ bp, unsigned long long (*rbox)[2][6])
{
    __shared__ double red[4][6];
    const uint32_t p = blockIdx.x >> 1, m = blockIdx.x & 1;
    const BatchPair b = bp[p];
    const double* v = m ? vc + 3ull * b.vc0 : vs + 3ull * b.vs0;
    const uint32_t n = m ? b.nvc : b.nvs;
    double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        for (int j = 0; j < 3; ++j) {
            const double x = v[3ull * i + j];
            lo[j] = x < lo[j] ? x : lo[j];
            hi[j] = x > hi[j] ? x : hi[j];
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int j = 0; j < 3; ++j) {
            const double a = __shfl_xor(lo[j], o, 64), c = __shfl_xor(hi[j], o, 64);
            lo[j] = a < lo[j] ? a : lo[j];
            hi[j] = c > hi[j] ? c : hi[j];
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int j = 0; j < 3; ++j) {
            red[w][j] = lo[j];
            red[w][3 + j] = hi[j];
        }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int j = threadIdx.x;
        double a = red[0][j], c = red[0][3 + j];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            a = red[q][j] < a ? red[q][j] : a;
            c = red[q][3 + j] > c ? red[q][3 + j] : c;
        }
        rbox[p][m][j] = ord_key(a);
        rbox[p][m][3 + j] = ord_key(c);
    }
}

// One block: the cell edge is the largest widened extent of any pair (both
// meshes, cut side widened by tol) times 1.0625, and pair p's cell is the
// one at the Morton-order position p, so that the pair-major leaf order
// walks neighbouring cells.  boff moves the pair's widened low corner to its
// cell's low corner.
__device__ __forceinline__ uint32_t compact_bits3(uint32_t x)
{
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0xff0000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}

__global__ void __launch_bounds__(1024) k_batch_layout(Params* P, const unsigned long long (*rbox)[2][6], uint32_t npairs)
{
    __shared__ double red[16];
    double ext = 0.0;
    for (uint32_t p = threadIdx.x; p < npairs; p += blockDim.x)
        for (int j = 0; j < 3; ++j) {
            const double lo = fmin(ord_value(rbox[p][0][j]), ord_value(rbox[p][1][j])) - P[p].tol[j];
            const double hi = fmax(ord_value(rbox[p][0][3 + j]), ord_value(rbox[p][1][3 + j])) + P[p].tol[j];
            ext = fmax(ext, hi - lo);
        }
    for (int o = 32; o > 0; o >>= 1) ext = fmax(ext, __shfl_xor(ext, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ext;
    __syncthreads();
    double cell = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) cell = fmax(cell, red[q]);
    cell = cell > 0.0 ? cell * 1.0625 : 1.0;
    for (uint32_t p = threadIdx.x; p < npairs; p += blockDim.x) {
        const uint32_t cx = compact_bits3(p >> 2), cy = compact_bits3(p >> 1), cz = compact_bits3(p);
        const double cc[3] = {(double)cx * cell, (double)cy * cell, (double)cz * cell};
        for (int j = 0; j < 3; ++j) {
            const double lo = fmin(ord_value(rbox[p][0][j]), ord_value(rbox[p][1][j])) - P[p].tol[j];
            P[p].boff[j] = cc[j] - lo;
        }
    }
}

// polygon-soup edge ranges of every pair, once the edges are numbered: the
// edges first used by a pair's corners are numbered in one run, starting at
// the count of edges first used before its first corner
__global__ void k_batch_edge_bases(BatchPair* bp, uint32_t npairs, const uint32_t* __restrict__ off_s,
                                   const uint32_t* __restrict__ off_c, uint32_t nfs, uint32_t nfc,
                                   const uint32_t* __restrict__ eid_s, const uint32_t* __restrict__ eid_c,
                                   const Counters* __restrict__ C)
{
    const uint32_t Es = (uint32_t)C->n_edges_src, Ec = (uint32_t)C->n_edges_cut;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += gridDim.x * blockDim.x) {
        const BatchPair b = bp[p];
        const uint32_t s0 = b.fs0 < nfs ? eid_s[off_s[b.fs0]] : Es;
        const uint32_t s1 = b.fs0 + b.nfs < nfs ? eid_s[off_s[b.fs0 + b.nfs]] : Es;
        const uint32_t c0 = b.fc0 < nfc ? eid_c[off_c[b.fc0]] : Ec;
        const uint32_t c1 = b.fc0 + b.nfc < nfc ? eid_c[off_c[b.fc0 + b.nfc]] : Ec;
        bp[p].es0 = s0;
        bp[p].nes = s1 - s0;
        bp[p].ec0 = c0;
        bp[p].nec = c1 - c0;
    }
}

// Pair-major registry -> pair-local ids (the ids of that pair's own dispatch)
// and each pair's range [rstart, rend) (both filled with 0 beforehand)
__global__ void k_batch_localize(const uint32_t* __restrict__ re, const uint32_t* __restrict__ rg,
                                 const unsigned long long* __restrict__ n_dev, uint64_t cap, uint32_t Fs,
                                 const uint32_t* __restrict__ fpair, const BatchPair* __restrict__ bp,
                                 const Counters* __restrict__ C, uint32_t* __restrict__ le, uint32_t* __restrict__ lg,
                                 uint32_t* __restrict__ rstart, uint32_t* __restrict__ rend)
{
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    const uint32_t Es = (uint32_t)C->n_edges_src;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = re[i], g = rg[i];
        const uint32_t p = fpair[g];
        const BatchPair b = bp[p];
        lg[i] = g < Fs ? g - b.fs0 : (g - Fs - b.fc0) + b.nfs;
        le[i] = e < Es ? e - b.es0 : (e - Es - b.ec0) + b.nes;
        if (i == 0 || fpair[rg[i - 1]] != p) rstart[p] = (uint32_t)i;
        if (i + 1 == n || fpair[rg[i + 1]] != p) rend[p] = (uint32_t)(i + 1);
    }
}

// ===========================================================================
// 7. cut-path edges (kernel.cpp:3427-3615): the edges between intersection
// points that cut each (source face, cut face) pair.
//
// The registry's points are grouped by (source face, cut face) pair as the
// narrow phase files them (kernel.cpp:2616-2631): a point on a cut-mesh edge
// joins the groups of the source face it crosses with each face of its edge,
// a point on a source-mesh edge the groups of each face of its edge with the
// cut face; groups are visited in (source face, cut face) order
// (std::map<pair<fd_t>, ...>) and hold their points in registry order.  A
// group of two points gives the edge (first, second); a longer group is sorted
// along its line (linear_projection_sort, kernel.cpp:1496-1531: projections on
// the unit vector from the second point to the first, libstdc++'s std::sort
// ascending) and each consecutive pair becomes an edge when its midpoint lies
// inside exactly two of the faces both points' registry entries name
// (kernel.cpp:3527-3583).  An edge that an earlier group already made is not
// made again.  A group of one point stops the dispatch (kernel.cpp:3445-3459).
// ===========================================================================

constexpr unsigned long long kCutNone = ~0ull;

// device counts of the cut-path stage (the scans write their totals here)
struct CutCounts {
    uint32_t groups, items, candidates, edges;
    uint32_t short_group, unsupported;
};

// a point's two group keys (source face << fbits | cut face); kCutNone for none
__global__ void k_cut_items(uint64_t n, const uint32_t* __restrict__ order, const uint32_t* __restrict__ re,
                            const uint32_t* __restrict__ rg, const int32_t* __restrict__ ef0,
                            const int32_t* __restrict__ ef1, uint32_t Fs, int fbits,
                            unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals)
{
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = order ? order[k] : (uint32_t)k;
        const uint32_t e = re[i];
        const unsigned long long g = rg[i];
        const int32_t f0 = ef0[e], f1 = ef1[e];
        const int32_t ef = f0 >= 0 ? f0 : f1;
        unsigned long long a, b = kCutNone;
        if ((uint32_t)ef >= Fs) { // a cut-mesh edge through source face g
            a = (g << fbits) | (uint32_t)ef;
            const int32_t other = ef == f0 ? f1 : -1;
            if (other >= 0) b = (g << fbits) | (uint32_t)other;
        } else { // a source-mesh edge through cut face g
            a = ((unsigned long long)ef << fbits) | g;
            const int32_t other = ef == f0 ? f1 : f0;
            if (other >= 0) b = ((unsigned long long)other << fbits) | g;
        }
        keys[2 * k] = a;
        keys[2 * k + 1] = b;
        vals[2 * k] = vals[2 * k + 1] = (uint32_t)k;
    }
}

// first item of each group (flag) and, once, the number of grouped items
__global__ void k_cut_flags(const unsigned long long* __restrict__ keys, uint64_t n2, uint32_t* __restrict__ flag,
                            CutCounts* cc)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n2; j += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long key = keys[j];
        flag[j] = (key != kCutNone && (j == 0 || keys[j - 1] != key)) ? 1u : 0u;
        if (key != kCutNone && (j + 1 == n2 || keys[j + 1] == kCutNone)) cc->items = (uint32_t)(j + 1);
    }
}

__global__ void k_cut_starts(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ gid, uint64_t n2,
                             uint32_t* __restrict__ gstart)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n2; j += (uint64_t)gridDim.x * blockDim.x)
        if (flag[j]) gstart[gid[j]] = (uint32_t)j;
}

// candidates per group (consecutive point pairs); groups of one point flagged
__global__ void k_cut_sizes(const uint32_t* __restrict__ gstart, CutCounts* cc, uint32_t* __restrict__ ncand)
{
    const uint32_t G = cc->groups;
    const uint64_t items = cc->items;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < G; g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t len = (g + 1 < G ? gstart[g + 1] : items) - gstart[g];
        ncand[g] = len >= 2 ? (uint32_t)(len - 1) : 0u;
        if (len < 2) atomicOr(&cc->short_group, 1u);
    }
}

struct CutCands {
    uint32_t *src, *tgt;     // points (registry ranks in the caller's order)
    uint32_t *fa, *fb;       // faces the edge is filed under: the group's pair, or the two containing faces
    uint32_t* group;         // group index (map order)
    uint8_t* kind;           // 0: group of two points; 1: consecutive points of a longer group (midpoint test)
    uint8_t* accepted;
};

// One thread per group: its candidates at coff[g].  Longer groups are sorted
// along their line in the triplet scratch (the group's item range).
__global__ void k_cut_groups(const uint32_t* __restrict__ gstart, const CutCounts* __restrict__ cc,
                             const unsigned long long* __restrict__ keys,
                             const uint32_t* __restrict__ vals, const uint32_t* __restrict__ order,
                             const double* __restrict__ rpts, const uint32_t* __restrict__ coff, int fbits,
                             Trip* __restrict__ scratch, CutCands out)
{
    const uint32_t G = cc->groups;
    const uint64_t items = cc->items;
    const unsigned long long fmask = (1ull << fbits) - 1;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < G; g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s0 = gstart[g];
        const uint64_t len = (g + 1 < G ? gstart[g + 1] : items) - s0;
        if (len < 2) continue;
        const unsigned long long key = keys[s0];
        const uint32_t sm = (uint32_t)(key >> fbits), cm = (uint32_t)(key & fmask);
        const uint32_t c0 = coff[g];
        if (len == 2) {
            out.src[c0] = vals[s0];
            out.tgt[c0] = vals[s0 + 1];
            out.fa[c0] = sm;
            out.fb[c0] = cm;
            out.group[c0] = (uint32_t)g;
            out.kind[c0] = 0;
            continue;
        }
        // linear_projection_sort: origin = first point, unit vector from the
        // second point to the first, projections of (origin - point) on it
        double o[3], d[3], u[3];
        {
            const uint32_t i0 = order ? order[vals[s0]] : vals[s0];
            const uint32_t i1 = order ? order[vals[s0 + 1]] : vals[s0 + 1];
            for (int j = 0; j < 3; ++j) {
                o[j] = rpts[3 * (uint64_t)i0 + j];
                d[j] = rpts[3 * (uint64_t)i1 + j];
            }
            double v[3];
            for (int j = 0; j < 3; ++j) v[j] = o[j] - d[j];
            double sq = 0.0;
            for (int j = 0; j < 3; ++j) sq = sq + v[j] * v[j];
            const double len_v = sqrt(sq);
            for (int j = 0; j < 3; ++j) u[j] = v[j] / len_v;
        }
        Trip* T = scratch + s0;
        for (uint64_t q = 0; q < len; ++q) {
            const uint32_t k = vals[s0 + q];
            const uint32_t i = order ? order[k] : k;
            double w[3];
            for (int j = 0; j < 3; ++j) w[j] = o[j] - rpts[3 * (uint64_t)i + j];
            double pr = 0.0;
            for (int j = 0; j < 3; ++j) pr = pr + w[j] * u[j];
            // the replayed sort orders by descending key: negated projections
            // ascend, and negation keeps every comparison's outcome
            T[q].a = -pr;
            T[q].ijk = (uint64_t)k;
        }
        trip_sort(T, T + len);
        for (uint64_t q = 1; q < len; ++q) {
            const uint32_t c = c0 + (uint32_t)(q - 1);
            out.src[c] = (uint32_t)T[q - 1].ijk;
            out.tgt[c] = (uint32_t)T[q].ijk;
            out.fa[c] = sm;
            out.fb[c] = cm;
            out.group[c] = (uint32_t)g;
            out.kind[c] = 1;
        }
    }
}

// the faces a registry entry names: its face, then its edge's two faces
// (ps_get_ivtx_registry_entry_faces, kernel.cpp:1407-1415; -1: none)
__device__ __forceinline__ void entry_faces(uint32_t i, const uint32_t* re, const uint32_t* rg, const int32_t* ef0,
                                            const int32_t* ef1, int32_t* f)
{
    const uint32_t e = re[i];
    f[0] = (int32_t)rg[i];
    f[1] = ef0[e];
    f[2] = ef1[e];
}

// Midpoint tests of the candidates of longer groups (kernel.cpp:3527-3583):
// the faces both points name (in the first point's order, none skipped but
// "no face"), each tested with its plane and vertices; accepted when the
// midpoint lies inside exactly two, which the edge is filed under.
__global__ void k_cut_midpoints(const CutCounts* __restrict__ cc, CutCands cand, const uint32_t* __restrict__ order,
                                const uint32_t* __restrict__ re, const uint32_t* __restrict__ rg,
                                const int32_t* __restrict__ ef0, const int32_t* __restrict__ ef1,
                                const double* __restrict__ rpts, const Plane* __restrict__ planes,
                                const uint32_t* __restrict__ cand_face, Soup S, CutCounts* ccw)
{
    const uint64_t n = cc->candidates;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x) {
        if (cand.kind[c] == 0) {
            cand.accepted[c] = 1;
            continue;
        }
        const uint32_t a = order ? order[cand.src[c]] : cand.src[c];
        const uint32_t b = order ? order[cand.tgt[c]] : cand.tgt[c];
        int32_t fa[3], fb[3];
        entry_faces(a, re, rg, ef0, ef1, fa);
        entry_faces(b, re, rg, ef0, ef1, fb);
        double mid[3];
        for (int j = 0; j < 3; ++j) mid[j] = (rpts[3 * (uint64_t)b + j] + rpts[3 * (uint64_t)a + j]) * 0.5;
        int inside = 0;
        uint32_t in_face[3] = {0, 0, 0};
        for (int x = 0; x < 3; ++x) {
            const int32_t f = fa[x];
            if (f < 0 || !(f == fb[0] || f == fb[1] || f == fb[2])) continue;
            if (!cand_face[f]) { // a face without a plane: the reference would find none either
                atomicOr(&ccw->unsupported, 1u);
                continue;
            }
            const Plane pl = planes[f];
            const FaceView fv = face_view(S, (uint32_t)f);
            if (point_in_polygon(mid, (int)fv.n, fv, pl.n, pl.k) == 'i') {
                if (inside < 3) in_face[inside] = (uint32_t)f;
                ++inside;
            }
        }
        cand.accepted[c] = inside == 2 ? 1 : 0;
        if (inside == 2) {
            cand.fa[c] = in_face[0];
            cand.fb[c] = in_face[1];
        }
    }
}

// an edge's first accepted candidate (candidates are numbered in group order)
__global__ void k_cut_first(const CutCounts* __restrict__ cc, CutCands cand, unsigned long long* __restrict__ table,
                            uint64_t mask)
{
    const uint64_t n = cc->candidates;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x) {
        if (!cand.accepted[c]) continue;
        const uint32_t a = cand.src[c], b = cand.tgt[c];
        const unsigned long long key = a < b ? ((unsigned long long)a << 32 | b) : ((unsigned long long)b << 32 | a);
        // slots hold (edge key, first candidate) in two words: the key by CAS, the candidate by atomicMin
        uint64_t h = hash64(key) & mask;
        for (uint64_t probe = 0; probe <= mask; ++probe) {
            const unsigned long long prev = atomicCAS(&table[2 * h], kCutNone, key);
            if (prev == kCutNone || prev == key) {
                atomicMin(&table[2 * h + 1], (unsigned long long)c);
                break;
            }
            h = (h + 1) & mask;
        }
    }
}

__global__ void k_cut_keep(const CutCounts* __restrict__ cc, CutCands cand, const unsigned long long* __restrict__ table,
                           uint64_t mask, uint32_t* __restrict__ keep)
{
    const uint64_t n = cc->candidates;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t k = 0;
        if (cand.accepted[c]) {
            const uint32_t a = cand.src[c], b = cand.tgt[c];
            const unsigned long long key = a < b ? ((unsigned long long)a << 32 | b) : ((unsigned long long)b << 32 | a);
            uint64_t h = hash64(key) & mask;
            while (table[2 * h] != key) h = (h + 1) & mask; // inserted by k_cut_first
            k = table[2 * h + 1] == c ? 1u : 0u;
        }
        keep[c] = k;
    }
}

struct CutEdges {
    uint32_t *src, *tgt, *fa, *fb, *group;
    uint8_t* kind;
};

__global__ void k_cut_emit(const CutCounts* __restrict__ cc, CutCands cand, const uint32_t* __restrict__ keep,
                           const uint32_t* __restrict__ pos, CutEdges out)
{
    const uint64_t n = cc->candidates;
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x) {
        if (!keep[c]) continue;
        const uint32_t q = pos[c];
        out.src[q] = cand.src[c];
        out.tgt[q] = cand.tgt[c];
        out.fa[q] = cand.fa[c];
        out.fb[q] = cand.fb[c];
        out.group[q] = cand.group[c];
        out.kind[q] = cand.kind[c];
    }
}

} // namespace mcgpu
