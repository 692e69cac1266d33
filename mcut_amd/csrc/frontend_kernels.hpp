// Front-end kernels of the MI355X MCUT intersection front end.
//
// Every kernel that consumes a variable-length list reads the length from the
// device counter its producer wrote and walks it with a grid-stride loop, so a
// dispatch is enqueued end to end without host synchronisation.
#pragma once
#include "kernels.hpp"

namespace mcgpu {

// ===========================================================================
// 1. shift parameters (calculate_vertex_parameters, preproc.cpp:2124-2290)
// ===========================================================================

// One 64-lane block per mesh.  The coordinate sums must be the reference's
// left-to-right sums (any other association changes the centre's bits and so
// every shifted coordinate), so the lanes walk the vertex array in order with
// uniform loads; the bounding box is an order-independent lane reduction.
__global__ void __launch_bounds__(64) k_mesh_sums(const double* __restrict__ v0, uint32_t n0,
                                                  const double* __restrict__ v1, uint32_t n1, Params* P)
{
    const int m = blockIdx.x;
    const double* __restrict__ v = m ? v1 : v0;
    const uint32_t n = m ? n1 : n0;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    uint32_t i = 0;
    for (; i + 4 <= n; i += 4) {
        const double* p = v + 3 * (uint64_t)i;
        const double a0 = p[0], a1 = p[1], a2 = p[2], b0 = p[3], b1 = p[4], b2 = p[5];
        const double c0 = p[6], c1 = p[7], c2 = p[8], d0 = p[9], d1 = p[10], d2 = p[11];
        sx += a0;
        sy += a1;
        sz += a2;
        sx += b0;
        sy += b1;
        sz += b2;
        sx += c0;
        sy += c1;
        sz += c2;
        sx += d0;
        sy += d1;
        sz += d2;
    }
    for (; i < n; ++i) {
        sx += v[3 * (uint64_t)i];
        sy += v[3 * (uint64_t)i + 1];
        sz += v[3 * (uint64_t)i + 2];
    }
    double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    for (uint32_t k = threadIdx.x; k < n; k += 64)
        for (int j = 0; j < 3; ++j) {
            const double x = v[3 * (uint64_t)k + j];
            hi[j] = (hi[j] < x) ? x : hi[j];
            lo[j] = (x < lo[j]) ? x : lo[j];
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int j = 0; j < 3; ++j) {
            const double a = __shfl_xor(lo[j], o, 64), b = __shfl_xor(hi[j], o, 64);
            lo[j] = (a < lo[j]) ? a : lo[j];
            hi[j] = (hi[j] < b) ? b : hi[j];
        }
    if (threadIdx.x == 0) {
        P->sum[m][0] = sx;
        P->sum[m][1] = sy;
        P->sum[m][2] = sz;
        for (int j = 0; j < 3; ++j) {
            P->bmin[m][j] = lo[j];
            P->bmax[m][j] = hi[j];
        }
    }
}

__global__ void k_shift_params(Params* P, uint32_t n0, uint32_t n1, double gp_constant, int eps_absolute, int has_pert,
                               double p0, double p1, double p2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
    P->has_pert = has_pert;
    P->pert[0] = p0;
    P->pert[1] = p1;
    P->pert[2] = p2;
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

__global__ void k_edge_insert(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ off, uint32_t nf,
                              int is_cut, unsigned long long* __restrict__ hkeys, unsigned long long* __restrict__ hfirst,
                              uint32_t* __restrict__ hcount, uint64_t mask)
{
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = off[f], n = off[f + 1] - b;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t a = idx[b + i], c = idx[b + (i + 1 == n ? 0 : i + 1)];
            const uint64_t slot = hash_insert(hkeys, mask, edge_key(a, c));
            atomicMin(&hfirst[slot], ((unsigned long long)corner_rank(b, i, n, is_cut) << 32) | a);
            atomicAdd(&hcount[slot], 1u);
        }
    }
}

__global__ void k_edge_flag(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ off, uint32_t nf, int is_cut,
                            const unsigned long long* __restrict__ hkeys, const unsigned long long* __restrict__ hfirst,
                            const uint32_t* __restrict__ hcount, uint64_t mask, uint32_t* __restrict__ flag, Counters* C)
{
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = off[f], n = off[f + 1] - b;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t a = idx[b + i], c = idx[b + (i + 1 == n ? 0 : i + 1)];
            const uint32_t rank = corner_rank(b, i, n, is_cut);
            const int64_t slot = hash_find(hkeys, mask, edge_key(a, c));
            if (slot < 0) {
                atomicOr(&C->nonmanifold, 2u);
                continue;
            }
            const unsigned long long first = hfirst[slot];
            const uint32_t r0 = (uint32_t)(first >> 32);
            flag[rank] = (r0 == rank) ? 1u : 0u;
            if (hcount[slot] > 2) atomicOr(&C->nonmanifold, 1u);
            if (r0 != rank && (uint32_t)first == a) atomicOr(&C->nonmanifold, 1u); // used twice in one direction
        }
    }
}

__global__ void k_store_edge_count(const uint32_t* __restrict__ total, Counters* C, int is_cut)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (is_cut) C->n_edges_cut = *total;
        else C->n_edges_src = *total;
    }
}

// corner -> ps edge; edge endpoints (h0 direction) and incident faces
__global__ void k_edge_assign(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ off, uint32_t nf, int is_cut,
                              const unsigned long long* __restrict__ hkeys, const unsigned long long* __restrict__ hfirst,
                              uint64_t mask, const uint32_t* __restrict__ eid_by_rank, const Counters* __restrict__ C,
                              uint32_t vofs, uint32_t fofs, uint32_t* __restrict__ cedge, uint32_t* __restrict__ ev0,
                              uint32_t* __restrict__ ev1, int32_t* __restrict__ ef0, int32_t* __restrict__ ef1)
{
    const uint32_t ebase = is_cut ? (uint32_t)C->n_edges_src : 0u;
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = off[f], n = off[f + 1] - b;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t a = idx[b + i], c = idx[b + (i + 1 == n ? 0 : i + 1)];
            const uint32_t rank = corner_rank(b, i, n, is_cut);
            const int64_t slot = hash_find(hkeys, mask, edge_key(a, c));
            if (slot < 0) continue; // cannot happen: every corner's key was inserted
            const uint32_t r0 = (uint32_t)(hfirst[slot] >> 32);
            const uint32_t e = ebase + eid_by_rank[r0];
            cedge[b + i] = e;
            if (r0 == rank) {
                ev0[e] = vofs + a;
                ev1[e] = vofs + c;
                ef0[e] = (int32_t)(fofs + f);
            } else {
                ef1[e] = (int32_t)(fofs + f);
            }
        }
    }
}

// ===========================================================================
// 4. face AABBs (bvh.cpp:242-313) and the LBVH over them
// ===========================================================================

__global__ void k_face_boxes(Soup S, const Params* __restrict__ P, double* __restrict__ box, int cut_only)
{
    const uint32_t F = S.nf[0] + S.nf[1];
    const double eps = P->eps;
    for (uint64_t g = (cut_only ? S.nf[0] : 0) + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < F;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const FaceView fv = face_view(S, (uint32_t)g);
        double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
        double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
        for (uint32_t i = 0; i < fv.n; ++i) {
            double p[3];
            fv((int)i, p);
            for (int j = 0; j < 3; ++j) {
                hi[j] = (hi[j] < p[j]) ? p[j] : hi[j];
                lo[j] = (p[j] < lo[j]) ? p[j] : lo[j];
            }
        }
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

// 30-bit Morton codes of face-box centres within the mesh's shifted box.  The
// candidate set does not depend on the tree (every overlapping leaf pair is
// found), so the codes only shape the tree.
__global__ void k_morton(const double* __restrict__ box, const Params* __restrict__ P, int mesh, uint32_t nf,
                         uint32_t* __restrict__ code, uint32_t* __restrict__ face)
{
    double lo[3], ext[3];
    for (int j = 0; j < 3; ++j) {
        lo[j] = (P->bmin[mesh][j] - P->com[j]) + P->t[j] - P->eps;
        const double hi = (P->bmax[mesh][j] - P->com[j]) + P->t[j] + P->eps;
        ext[j] = hi > lo[j] ? hi - lo[j] : 1.0;
    }
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
        const double* b = box + 6 * f;
        uint32_t c[3];
        for (int j = 0; j < 3; ++j) {
            float x = (float)((((b[j] + b[3 + j]) * 0.5) - lo[j]) / ext[j]);
            x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
            c[j] = expand_bits10((uint32_t)x);
        }
        code[f] = c[0] * 4 + c[1] * 2 + c[2];
        face[f] = (uint32_t)f;
    }
}

constexpr uint32_t kLeaf = 0x80000000u;

__device__ __forceinline__ int lbvh_delta(const uint32_t* __restrict__ code, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint32_t a = code[i], b = code[j];
    if (a == b) return 32 + __clz((uint32_t)(i ^ j));
    return __clz(a ^ b);
}

// binary radix tree over sorted codes (Karras 2012); duplicates ordered by index
__global__ void k_lbvh_nodes(const uint32_t* __restrict__ code, int n, uint32_t* __restrict__ child,
                             int32_t* __restrict__ nparent, int32_t* __restrict__ lparent)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        const int d = (lbvh_delta(code, n, i, i + 1) - lbvh_delta(code, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = lbvh_delta(code, n, i, i - d);
        int lmax = 2;
        while (lbvh_delta(code, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (lbvh_delta(code, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = lbvh_delta(code, n, i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (lbvh_delta(code, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int gamma = i + s * d + (d < 0 ? d : 0);
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const uint32_t left = (lo == gamma) ? (kLeaf | (uint32_t)gamma) : (uint32_t)gamma;
        const uint32_t right = (hi == gamma + 1) ? (kLeaf | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
        child[2 * i] = left;
        child[2 * i + 1] = right;
        if (left & kLeaf) lparent[gamma] = i;
        else nparent[gamma] = i;
        if (right & kLeaf) lparent[gamma + 1] = i;
        else nparent[gamma + 1] = i;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) nparent[0] = -1;
}

__device__ __forceinline__ void load_box_coherent(const double* p, double* o)
{
    const volatile double* q = p;
    for (int j = 0; j < 6; ++j) o[j] = q[j];
}

// bottom-up refit: the second child to arrive at a node writes the union
__global__ void k_lbvh_refit(int n, const uint32_t* __restrict__ sface, const double* __restrict__ facebox,
                             const uint32_t* __restrict__ child, const int32_t* __restrict__ nparent,
                             const int32_t* __restrict__ lparent, uint32_t* __restrict__ visits, double* nodebox)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int p = lparent[i];
        for (int level = 0; p >= 0 && level < 128; ++level) {
            __threadfence();
            if (atomicAdd(&visits[p], 1u) == 0) break;
            __threadfence();
            double b[6], c[6];
            for (int side = 0; side < 2; ++side) {
                const uint32_t ch = child[2 * p + side];
                if (ch & kLeaf) {
                    const double* fb = facebox + 6 * (uint64_t)sface[ch & ~kLeaf];
                    for (int j = 0; j < 6; ++j) c[j] = fb[j];
                } else {
                    load_box_coherent(nodebox + 6 * (uint64_t)ch, c);
                }
                if (side == 0) {
                    for (int j = 0; j < 6; ++j) b[j] = c[j];
                } else {
                    for (int j = 0; j < 3; ++j) {
                        b[j] = (c[j] < b[j]) ? c[j] : b[j];
                        b[3 + j] = (b[3 + j] < c[3 + j]) ? c[3 + j] : b[3 + j];
                    }
                }
            }
            for (int j = 0; j < 6; ++j) nodebox[6 * (uint64_t)p + j] = b[j];
            p = nparent[p];
        }
    }
}

// ===========================================================================
// 5. broad phase traversal: every source face box against the cut tree.
// Emits each (src, cut) face pair whose boxes overlap (bvh.cpp:661-781 finds
// exactly these) and records it in the pair hash used for test dedup.
// ===========================================================================

struct Tree {
    const uint32_t* child;
    const uint32_t* sface; // leaf -> mesh-local face
    const double* nodebox;
    const double* facebox; // ps face boxes of this mesh
    int n;
};

__global__ void __launch_bounds__(kBlock) k_traverse(Tree src, Tree cut, unsigned long long* __restrict__ pairs,
                                                     uint64_t cap, unsigned long long* __restrict__ phash, uint64_t pmask,
                                                     Counters* C)
{
    uint32_t stack[64];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < src.n; i += gridDim.x * blockDim.x) {
        const uint32_t s = src.sface[i];
        double q[6];
        for (int j = 0; j < 6; ++j) q[j] = src.facebox[6 * (uint64_t)s + j];
        int sp = 0;
        if (cut.n == 1) {
            stack[sp++] = kLeaf | 0u;
        } else {
            stack[sp++] = 0u;
        }
        // each node is entered at most once per query: the visit count is bounded by the tree size
        for (int visits = 0; sp > 0 && visits < 2 * cut.n + 2; ++visits) {
            const uint32_t node = stack[--sp];
            uint32_t kids[2];
            int nk;
            if (node & kLeaf) {
                kids[0] = node;
                nk = 1;
            } else {
                if (sp == 0 && node == 0u) {
                    // root: test it before descending
                    double r[6];
                    for (int j = 0; j < 6; ++j) r[j] = cut.nodebox[j];
                    if (!boxes_overlap(q, r)) break;
                }
                kids[0] = cut.child[2 * node];
                kids[1] = cut.child[2 * node + 1];
                nk = 2;
            }
            for (int k = 0; k < nk; ++k) {
                const uint32_t ch = kids[k];
                double b[6];
                if (ch & kLeaf) {
                    const uint32_t c = cut.sface[ch & ~kLeaf];
                    for (int j = 0; j < 6; ++j) b[j] = cut.facebox[6 * (uint64_t)c + j];
                    if (boxes_overlap(q, b)) {
                        const uint64_t key = ((uint64_t)s << 32) | c;
                        const unsigned long long at = wave_append(&C->n_pairs);
                        if (at < cap) {
                            pairs[at] = key;
                            hash_insert(phash, pmask, key);
                        } else {
                            atomicOr(&C->ov_pairs, 1u);
                        }
                    }
                } else {
                    for (int j = 0; j < 6; ++j) b[j] = cut.nodebox[6 * (uint64_t)ch + j];
                    if (boxes_overlap(q, b)) {
                        if (sp < 64) stack[sp++] = ch;
                        else atomicOr(&C->ov_pairs, 2u);
                    }
                }
            }
        }
    }
}

// ===========================================================================
// 6. candidate faces and their planes (kernel.cpp:2184-2356)
// ===========================================================================

__global__ void k_mark_candidates(const unsigned long long* __restrict__ pairs, const Counters* __restrict__ C,
                                  uint64_t cap, uint32_t Fs, uint8_t* __restrict__ cand)
{
    const uint64_t n = C->n_pairs < cap ? C->n_pairs : cap;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = pairs[i];
        cand[k >> 32] = 1;
        cand[Fs + (uint32_t)k] = 1;
    }
}

// orient2d of projected triplets; the first triplet of greatest |orient2d|
// (std::sort of <= 16 triplets is an insertion sort, which keeps ties in order)
template <typename VertexAt>
__device__ inline bool best_triple_small(int n, VertexAt vat, const double* normal, int k, uint8_t* tri)
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
                    tri[0] = (uint8_t)a, tri[1] = (uint8_t)b, tri[2] = (uint8_t)c;
                    found = true;
                }
            }
    return found;
}

__global__ void k_planes(Soup S, const uint8_t* __restrict__ cand, Plane* __restrict__ planes,
                         uint32_t* __restrict__ big_list, uint64_t cap_big, Counters* C)
{
    const uint32_t F = S.nf[0] + S.nf[1];
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < F; g += (uint64_t)gridDim.x * blockDim.x) {
        if (!cand[g]) continue;
        const FaceView fv = face_view(S, (uint32_t)g);
        Plane pl;
        const int k = plane_coefficients((int)fv.n, fv, pl.n, &pl.d);
        pl.k = k;
        pl.tri[0] = 0, pl.tri[1] = 1, pl.tri[2] = 2;
        pl.big = 0;
        atomicAdd(&C->n_cand, 1ull);
        if (k < 0) {
            // kernel.cpp:2320-2331 names the mesh with "face id > source face count"
            if ((uint32_t)g > S.nf[0]) atomicOr(&C->bad_cut, 1u);
            else atomicOr(&C->bad_src, 1u);
        } else if (fv.n > 3 && fv.n <= 5) {
            if (!best_triple_small((int)fv.n, fv, pl.n, k, pl.tri)) pl.tri[0] = 0xff;
        } else if (fv.n > 5 && fv.n <= kMaxPoly) {
            pl.big = 1;
            const unsigned long long at = atomicAdd(&C->n_big, 1ull);
            if (at < cap_big) big_list[at] = (uint32_t)g;
            else atomicOr(&C->ov_big, 1u);
        } else if (fv.n > kMaxPoly) {
            atomicOr(&C->unsupported, 1u);
        }
        planes[g] = pl;
    }
}

// Replay of GCC libstdc++'s std::sort on the triplet list (introsort with a
// 16-element insertion-sort threshold, median-of-three pivot moved to the
// front, unguarded partition, final insertion sort), so the first element
// -- the triplet the reference uses -- matches for polygons of 6..16 vertices.
struct Trip {
    double a; // |orient2d|
    int32_t ijk;
};

__device__ __forceinline__ bool trip_less(const Trip& x, const Trip& y) { return x.a > y.a; }

__device__ inline void trip_linear_insert(Trip* last)
{
    Trip val = *last;
    Trip* next = last - 1;
    while (trip_less(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

__device__ inline void trip_insertion_sort(Trip* first, Trip* last)
{
    if (first == last) return;
    for (Trip* i = first + 1; i != last; ++i) {
        if (trip_less(*i, *first)) {
            Trip val = *i;
            for (Trip* p = i; p != first; --p) *p = *(p - 1);
            *first = val;
        } else {
            trip_linear_insert(i);
        }
    }
}

__device__ __forceinline__ void trip_swap(Trip* x, Trip* y)
{
    Trip t = *x;
    *x = *y;
    *y = t;
}

__device__ inline int trip_sort(Trip* first, Trip* last)
{
    const long n = last - first;
    if (n <= 0) return 0;
    int lg = 0;
    while ((1L << (lg + 1)) <= n) ++lg;
    // iterative form of the introsort loop: the recursion on the right part
    // happens before the loop continues on the left part
    Trip* seg_first[64];
    Trip* seg_last[64];
    int seg_depth[64];
    int top = 0;
    seg_first[top] = first;
    seg_last[top] = last;
    seg_depth[top] = 2 * lg;
    ++top;
    while (top > 0) {
        --top;
        Trip* f = seg_first[top];
        Trip* l = seg_last[top];
        int depth = seg_depth[top];
        while (l - f > 16) {
            if (depth == 0) return -1;
            --depth;
            Trip* mid = f + (l - f) / 2;
            Trip *a = f + 1, *b = mid, *c = l - 1;
            if (trip_less(*a, *b)) {
                if (trip_less(*b, *c)) trip_swap(f, b);
                else if (trip_less(*a, *c)) trip_swap(f, c);
                else trip_swap(f, a);
            } else if (trip_less(*a, *c)) {
                trip_swap(f, a);
            } else if (trip_less(*b, *c)) {
                trip_swap(f, c);
            } else {
                trip_swap(f, b);
            }
            Trip* lo = f + 1;
            Trip* hi = l;
            for (;;) {
                while (trip_less(*lo, *f)) ++lo;
                --hi;
                while (trip_less(*f, *hi)) --hi;
                if (!(lo < hi)) break;
                trip_swap(lo, hi);
                ++lo;
            }
            // recurse on [lo, l) first, then continue with [f, lo)
            if (top >= 62) return -1;
            // push the left part to be processed after the right part
            seg_first[top] = f;
            seg_last[top] = lo;
            seg_depth[top] = depth;
            ++top;
            f = lo;
        }
    }
    if (n > 16) {
        trip_insertion_sort(first, first + 16);
        for (Trip* i = first + 16; i != last; ++i) trip_linear_insert(i);
    } else {
        trip_insertion_sort(first, last);
    }
    return 0;
}

__global__ void k_big_polygons(Soup S, const uint32_t* __restrict__ big_list, const Counters* __restrict__ C,
                               uint64_t cap_big, Plane* __restrict__ planes, Trip* __restrict__ scratch, Counters* Cw)
{
    const uint64_t nb = C->n_big < cap_big ? C->n_big : cap_big;
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    Trip* T = scratch + tid * 560;
    for (uint64_t w = tid; w < nb; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t g = big_list[w];
        const FaceView fv = face_view(S, g);
        Plane pl = planes[g];
        double M[6], X[2 * kMaxPoly];
        projection_matrix(pl.n, pl.k, M);
        for (uint32_t i = 0; i < fv.n; ++i) {
            double p[3];
            fv((int)i, p);
            project2(M, p, X + 2 * i);
        }
        int m = 0;
        for (int a = 0; a < (int)fv.n; ++a)
            for (int b = a + 1; b < (int)fv.n; ++b)
                for (int c = b + 1; c < (int)fv.n; ++c) {
                    const double r = orient2d_adaptive(X + 2 * a, X + 2 * b, X + 2 * c);
                    if (r == 0.0) continue;
                    T[m].a = fabs(r);
                    T[m].ijk = a | (b << 8) | (c << 16);
                    ++m;
                }
        if (m == 0) {
            pl.tri[0] = 0xff;
        } else if (trip_sort(T, T + m) < 0) {
            atomicOr(&Cw->unsupported, 2u);
        } else {
            pl.tri[0] = (uint8_t)(T[0].ijk & 0xff);
            pl.tri[1] = (uint8_t)((T[0].ijk >> 8) & 0xff);
            pl.tri[2] = (uint8_t)((T[0].ijk >> 16) & 0xff);
        }
        planes[g] = pl;
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

__global__ void __launch_bounds__(kBlock) k_tests(Soup S, const unsigned long long* __restrict__ pairs, uint64_t cap_pairs,
                                                  const unsigned long long* __restrict__ phash, uint64_t pmask,
                                                  uint32_t* __restrict__ te, uint32_t* __restrict__ tg, uint64_t cap_tests,
                                                  Counters* C)
{
    const uint64_t n = C->n_pairs < cap_pairs ? C->n_pairs : cap_pairs;
    const uint32_t Fs = S.nf[0];
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = pairs[i];
        const uint32_t s = (uint32_t)(key >> 32), c = (uint32_t)key;
        // edges of the source face against the cut face
        {
            const uint32_t b = S.off[0][s], nn = S.off[0][s + 1] - b;
            for (uint32_t k = 0; k < nn; ++k) {
                const uint32_t e = S.cedge[0][b + k];
                const int32_t f0 = S.ef0[e], f1 = S.ef1[e];
                const int32_t fo = (f0 == (int32_t)s) ? f1 : f0;
                if (fo >= 0 && (uint32_t)fo < s && pair_exists(phash, pmask, (uint32_t)fo, c)) continue;
                if (!edge_face_cull(S, e, Fs + c)) continue;
                const unsigned long long at = wave_append(&C->n_tests);
                if (at < cap_tests) {
                    te[at] = e;
                    tg[at] = Fs + c;
                } else {
                    atomicOr(&C->ov_tests, 1u);
                }
            }
        }
        // edges of the cut face against the source face
        {
            const uint32_t b = S.off[1][c], nn = S.off[1][c + 1] - b;
            const int32_t self = (int32_t)(Fs + c);
            for (uint32_t k = 0; k < nn; ++k) {
                const uint32_t e = S.cedge[1][b + k];
                const int32_t f0 = S.ef0[e], f1 = S.ef1[e];
                const int32_t fo = (f0 == self) ? f1 : f0;
                if (fo >= 0 && fo < self && pair_exists(phash, pmask, s, (uint32_t)fo - Fs)) continue;
                if (!edge_face_cull(S, e, s)) continue;
                const unsigned long long at = wave_append(&C->n_tests);
                if (at < cap_tests) {
                    te[at] = e;
                    tg[at] = s;
                } else {
                    atomicOr(&C->ov_tests, 1u);
                }
            }
        }
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
    unsigned long long* key;
    double* p;
    uint64_t cap;
    int gbits;
};

// the decision after the two orient3d signs are known
__device__ inline void resolve_test(const Soup& S, const Plane* planes, uint32_t e, uint32_t g, int sq, int sr,
                                    uint64_t ti, const TestOut& TO, const RegOut& R, Counters* C)
{
    const char ty = segment_type(sq, sr);
    if (TO.qs) {
        TO.qs[ti] = (int8_t)sq;
        TO.rs[ti] = (int8_t)sr;
        TO.type[ti] = ty;
        TO.pip[ti] = '-';
    }
    if (ty == '0') return;
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
        if (violated) atomicOr(&C->gp, 1u);
        return;
    }
    double p[3] = {0.0, 0.0, 0.0};
    segment_point(pl.n, pl.d, q, r, p);
    const char res = point_in_polygon(p, (int)fv.n, fv, pl.n, pl.k);
    if (TO.qs) TO.pip[ti] = res;
    if (res == 'v' || res == 'e') {
        atomicOr(&C->gp, 1u);
        return;
    }
    if (res != 'i') return;
    const unsigned long long at = wave_append(&C->n_reg);
    if (at < R.cap) {
        R.key[at] = ((unsigned long long)e << R.gbits) | g;
        R.p[3 * at] = p[0];
        R.p[3 * at + 1] = p[1];
        R.p[3 * at + 2] = p[2];
    } else {
        atomicOr(&C->ov_reg, 1u);
    }
}

__global__ void __launch_bounds__(kBlock) k_narrow(Soup S, const Plane* __restrict__ planes, const uint32_t* __restrict__ te,
                                                   const uint32_t* __restrict__ tg, uint64_t cap_tests,
                                                   uint32_t* __restrict__ esc, uint64_t cap_esc, TestOut TO, RegOut R,
                                                   Counters* C)
{
    const uint64_t n = C->n_tests < cap_tests ? C->n_tests : cap_tests;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = te[i], g = tg[i];
        const Plane pl = planes[g];
        int sq = 0, sr = 0;
        if (pl.tri[0] == 0xff || pl.k < 0) {
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
        sq = orient3d_filter(a, b, c, q);
        sr = orient3d_filter(a, b, c, r);
        if (sq == 0 || sr == 0) {
            const unsigned long long at = wave_append(&C->n_esc);
            if (at < cap_esc) esc[at] = (uint32_t)i;
            else atomicOr(&C->ov_esc, 1u);
            continue;
        }
        resolve_test(S, planes, e, g, sq, sr, i, TO, R, C);
    }
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
        int sq = orient3d_filter(a, b, c, q);
        int sr = orient3d_filter(a, b, c, r);
        if (sq == 0) sq = orient3d_exact_sign(a, b, c, q);
        if (sr == 0) sr = orient3d_exact_sign(a, b, c, r);
        resolve_test(S, planes, e, g, sq, sr, i, TO, R, C);
    }
}

// ===========================================================================
// 9. registry: sorted order, then the status rules that read it
// ===========================================================================

__global__ void k_iota(uint32_t* v, const unsigned long long* n_dev, uint64_t cap)
{
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}

__global__ void k_clamp_count(const unsigned long long* n_dev, uint64_t cap, unsigned long long* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = *n_dev < cap ? *n_dev : cap;
}

// sorted keys + permutation -> (edge, face, point) arrays; the status
// side structures of kernel.cpp:2603-2654 are keyed the same way
__global__ void k_registry_gather(const unsigned long long* __restrict__ skey, const uint32_t* __restrict__ perm,
                                  const double* __restrict__ p_in, const unsigned long long* __restrict__ n_dev,
                                  uint64_t cap, int gbits, uint32_t* __restrict__ re, uint32_t* __restrict__ rg,
                                  double* __restrict__ rp)
{
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    const uint64_t gmask = (1ull << gbits) - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = skey[i];
        re[i] = (uint32_t)(k >> gbits);
        rg[i] = (uint32_t)(k & gmask);
        const uint32_t j = perm[i];
        rp[3 * i] = p_in[3 * (uint64_t)j];
        rp[3 * i + 1] = p_in[3 * (uint64_t)j + 1];
        rp[3 * i + 2] = p_in[3 * (uint64_t)j + 2];
    }
}

// cutpath_edge_creation_info keys (kernel.cpp:2620-2636): every (source face,
// cut face) pair joined by an intersection point needs at least two points
// (kernel.cpp:3439-3452); a cut-mesh border edge hit marks a partial cut
// (kernel.cpp:2646-2650)
__global__ void k_status_pairs(Soup S, const uint32_t* __restrict__ re, const uint32_t* __restrict__ rg,
                               const unsigned long long* __restrict__ n_dev, uint64_t cap,
                               unsigned long long* __restrict__ hkeys, uint32_t* __restrict__ hcnt, uint64_t mask,
                               Counters* C)
{
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    const uint32_t Fs = S.nf[0], Vs = S.nv[0];
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = re[i], g = rg[i];
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
        if (S.ev0[e] >= Vs && (f0 < 0 || f1 < 0)) atomicOr(&C->partial, 1u);
    }
}

__global__ void k_status_check(const unsigned long long* __restrict__ hkeys, const uint32_t* __restrict__ hcnt,
                               uint64_t slots, Counters* C)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slots; i += (uint64_t)gridDim.x * blockDim.x)
        if (hkeys[i] != kEmpty64 && hcnt[i] < 2) atomicOr(&C->pair_fail, 1u);
}

// packed test keys for the optional sorted test listing
__global__ void k_test_keys(const uint32_t* __restrict__ te, const uint32_t* __restrict__ tg,
                            const unsigned long long* __restrict__ n_dev, uint64_t cap, int gbits,
                            unsigned long long* __restrict__ key)
{
    const uint64_t n = *n_dev < cap ? *n_dev : cap;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        key[i] = ((unsigned long long)te[i] << gbits) | tg[i];
}

} // namespace mcgpu
