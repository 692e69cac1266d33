// Front-end kernels: shift parameters, polygon-soup topology, LBVH broad
// phase, edge-face test expansion, face planes and the narrow phase.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geom.hpp"
#include "tripsort.hpp"

namespace mcgpu {

constexpr uint64_t kEmpty64 = ~0ull;
constexpr int kBlock = 256;
constexpr int kMaxPoly = 65535; // largest candidate polygon (its triplet is kept as three 16-bit vertex indices)
constexpr uint16_t kCollinear = 0xffff; // Plane::tri[0] of a polygon whose vertices are all collinear

struct Params {
    double com[3];      // shift x' = ((x - com) + t) + pert
    double t[3];
    double eps;         // cut-face AABB enlargement
    double diag;
    double pert[3];
    int has_pert;
    int eps_absolute;
    double gp_constant;
    double sum[2][3];   // per-mesh sequential coordinate sums
    double bmin[2][3];  // per-mesh raw bounding boxes
    double bmax[2][3];
    double tol[3];      // raw-space broad-phase tolerance per axis (k_tolerance)
    double boff[3];     // batches: raw-space offset that places this pair's boxes in a cell of its own
};

struct Counters {
    unsigned long long n_pairs, n_tests, n_esc, n_reg, n_big, n_cand;
    unsigned long long n_edges_src, n_edges_cut;
    unsigned long long n_raw;                              // raw-space (superset) candidate pairs
    unsigned long long n_fallback, n_lane_walks; // traversal statistics: node splits, per-lane walks
    unsigned long long n_f32, n_f64;            // tests whose two signs the FP32 / FP64 filter settled
    unsigned long long pmask; // pair hash slots - 1 (set once the raw pairs are known)
    unsigned long long n_trip; // triplet-scratch entries reserved by the polygons of more than five vertices
    double zero;              // +0.0 (the per-dispatch reset clears it): padding the serial sums load
    unsigned int ov_pairs, ov_tests, ov_esc, ov_reg, ov_big;
    // unsupported: bit 0 a candidate polygon of more than kMaxPoly vertices, bit 1
    // coordinates whose broad-phase rounding bound is not finite
    unsigned int gp, bad_src, bad_cut, unsupported, nonmanifold, bad_input, pair_fail, partial;
    unsigned int sum_next; // batches: work-queue head of the serial sums
    unsigned int ov_trip;  // the triplet scratch overflowed
    unsigned int trav_q[8]; // persistent traversal: head of each work queue (blocks taken)
    unsigned int ncc[2];    // connected components of each mesh's vertex graph (all pairs of a batch)
    unsigned int cc_changed, cc_done; // component check: a hook this round; no hook in the last round
    // the raw-space broad phase would be loose: the rounding part of its
    // tolerance exceeds 1/16 of the pair's largest extent (k_tolerance);
    // its leaves are left empty and the host reruns the dispatch with the
    // broad phase in shifted coordinates
    unsigned int loose;
    // raw bounding box of each mesh as order keys (lo xyz, hi xyz); written
    // while the input is validated and kept by the per-dispatch counter reset
    unsigned long long rbox[2][6];
};

// the per-dispatch scalars the host reads back, in one allocation (one copy)
struct DevState {
    Counters C;
    Params P;
};

// One mesh pair of a batch.  A batch is run as one polygon soup whose source
// part is every source mesh in pair order and whose cut part is every cut mesh
// in pair order; these are the pair's ranges in the two parts.
struct BatchPair {
    uint32_t fs0, nfs, fc0, nfc; // faces: first source face, count; first cut face (cut-part index), count
    uint32_t vs0, nvs, vc0, nvc; // vertices, likewise
    uint32_t es0, nes, ec0, nec; // polygon-soup edges, likewise (written once the edges are numbered)
};

// per-pair status bits of a batch (the dispatch-wide flags of Counters, per pair)
enum PairFlag : uint32_t {
    PF_GP = 1u,
    PF_BAD_SRC = 2u,
    PF_BAD_CUT = 4u,
    PF_PAIR_FAIL = 8u,
    PF_PARTIAL = 16u,
    PF_UNSUPPORTED = 32u,
};

// polygon soup view: source faces [0,Fs), cut faces [Fs, Fs+Fc); vertices likewise
struct Soup {
    const uint32_t* idx[2];  // face vertex indices (mesh-local)
    const uint32_t* off[2];  // face offsets (nf+1)
    uint32_t nv[2], nf[2];
    const double* ps_v;      // shifted coordinates, xyz per ps vertex
    const double* box;       // per ps face: min xyz, max xyz (cut faces enlarged)
    const uint32_t* cedge[2];// per corner (mesh corner order): ps edge id
    const uint32_t* ev0;     // per ps edge: h0 = v0 -> v1
    const uint32_t* ev1;
    const int32_t* ef0;      // face of h0, h1 (-1 none)
    const int32_t* ef1;
    const uint32_t* fpair;   // batches: pair of each ps face (null for one pair)
    uint32_t* pflags;        // batches: PairFlag bits per pair
};

// records a status bit on the pair of ps face g (batches only)
__device__ __forceinline__ void pair_flag(const Soup& S, uint32_t g, uint32_t bit)
{
    if (!S.pflags) return;
    uint32_t* f = S.pflags + S.fpair[g];
    if (!(*f & bit)) atomicOr(f, bit);
}

__device__ __forceinline__ void load3(const double* p, double* o)
{
    o[0] = p[0];
    o[1] = p[1];
    o[2] = p[2];
}

// ps vertex list of ps face g (targets of its halfedges): source faces are
// stored [v1..v_{n-1}, v0], cut faces [u2..u0, u1] (hmesh.cpp:705-733,
// kernel.cpp:1717-1727)
struct FaceView {
    const uint32_t* idx;
    uint32_t base, n, rot, vofs;
    const double* ps_v;
    __device__ __forceinline__ uint32_t vid(int i) const
    {
        uint32_t j = (uint32_t)i + rot;
        if (j >= n) j -= n;
        if (j >= n) j -= n;
        return vofs + idx[base + j];
    }
    __device__ __forceinline__ void operator()(int i, double* o) const { load3(ps_v + 3 * (uint64_t)vid(i), o); }
};

__device__ __forceinline__ FaceView face_view(const Soup& S, uint32_t g)
{
    FaceView v;
    const int m = g >= S.nf[0];
    const uint32_t f = m ? g - S.nf[0] : g;
    v.idx = S.idx[m];
    v.base = S.off[m][f];
    v.n = S.off[m][f + 1] - v.base;
    v.rot = m ? 2 : 1;
    v.vofs = m ? S.nv[0] : 0;
    v.ps_v = S.ps_v;
    return v;
}

struct Plane {
    double n[3];
    double d;
    int32_t k;      // largest normal component; -1 = degenerate
    uint16_t tri[3]; // vertex triple for orient3d; tri[0] = kCollinear: all collinear
    uint8_t big;    // triple decided by the large-polygon pass
};

// order-preserving map of doubles to unsigned integers (for atomic min/max)
__device__ __forceinline__ unsigned long long ord_key(double x)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double ord_value(unsigned long long k)
{
    const unsigned long long u = (k & 0x8000000000000000ull) ? (k & ~0x8000000000000000ull) : ~k;
    return __longlong_as_double((long long)u);
}

__device__ __forceinline__ bool boxes_overlap(const double* a, const double* b)
{
    // intersect_bounding_boxes, math.h:931-941 (inclusive)
    return (a[0] <= b[3] && a[3] >= b[0]) && (a[1] <= b[4] && a[4] >= b[1]) && (a[2] <= b[5] && a[5] >= b[2]);
}

__device__ __forceinline__ uint64_t hash64(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// insert-or-find; returns slot
// Tables are sized to at least twice the keys they receive, so probing ends at
// an empty slot; the probe count is still bounded by the table size.
__device__ __forceinline__ uint64_t hash_insert(unsigned long long* keys, uint64_t mask, uint64_t key)
{
    uint64_t h = hash64(key) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        unsigned long long prev = atomicCAS(&keys[h], (unsigned long long)kEmpty64, (unsigned long long)key);
        if (prev == kEmpty64 || prev == key) return h;
        h = (h + 1) & mask;
    }
    return h; // unreachable while the table is at most half full
}

__device__ __forceinline__ int64_t hash_find(const unsigned long long* keys, uint64_t mask, uint64_t key)
{
    uint64_t h = hash64(key) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        unsigned long long k = keys[h];
        if (k == key) return (int64_t)h;
        if (k == kEmpty64) return -1;
        h = (h + 1) & mask;
    }
    return -1;
}

// wavefront-aggregated append: every active lane gets a slot; returns its index
__device__ __forceinline__ unsigned long long wave_append(unsigned long long* counter)
{
    const uint64_t active = __ballot(1);
    const int lane = __lane_id();
    const int leader = __ffsll((long long)active) - 1;
    const uint32_t rank = __popcll(active & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(active));
    base = __shfl(base, leader, 64);
    return base + rank;
}

// Appends to a global list cost one same-address atomic each, and those
// serialise device-wide (about 12 ns apiece on MI355X).  Kernels that emit
// many entries therefore stage them in LDS: a wave reserves staging slots with
// one LDS atomic (wave_reserve), and the block flushes the staged entries with
// one global atomic for the whole tile.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* lds_counter)
{
    const uint64_t active = __ballot(1);
    const int lane = __lane_id();
    const int leader = __ffsll((long long)active) - 1;
    const uint32_t rank = __popcll(active & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(lds_counter, (uint32_t)__popcll(active));
    base = __shfl(base, leader, 64);
    return base + rank;
}

// Block-wide: reserves `m` entries of a global list (one atomic, by thread 0)
// and returns the first index to every thread.  All threads of the block call
// it; `slot` is an LDS word.
__device__ __forceinline__ unsigned long long block_reserve(unsigned long long* counter, uint32_t m,
                                                            unsigned long long* slot)
{
    if (threadIdx.x == 0) *slot = m ? atomicAdd(counter, (unsigned long long)m) : 0ull;
    __syncthreads();
    const unsigned long long base = *slot;
    return base;
}

} // namespace mcgpu
