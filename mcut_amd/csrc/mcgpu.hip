// Host side of libmcgpu: the C ABI of include/mcgpu.h, device buffers, and
// the enqueue order of the front-end pipeline.
#include "mcgpu.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "frontend_kernels.hpp"
#include "sort.hpp"

using namespace mcgpu;

namespace {

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};

// Stages of one dispatch.  The serial coordinate sums run alone on a stream
// whose CU mask reserves kSumCUs compute units for them; stream A carries the
// chain that depends on the shift; stream B builds both trees and runs the
// broad phase in raw coordinates alongside the sums; stream C numbers the
// polygon-soup edges.  B and C are masked off the reserved CUs: a serial FP64
// chain that shares its CU with memory- or atomic-heavy waves runs 4-8 times
// slower.  B and C also get disjoint CU sets: sharing all of them, the tree
// build's small launches queue behind the edge numbering's large hash grids
// and the sort stretches from 0.12 to 0.25 ms.  B takes kBroadShare of the
// CUs left after the sums (174 of 248 on the 256-CU part): it is the longest
// of the three paths, and the edge numbering finishes well inside it.  A joins B before the exact pair filter and
// C before the edge-face tests.  A batch runs on a second pair of streams
// kBatchBroadShare : rest (mcgpu_set_batch_split before the first batch):
// its edge numbering runs over every pair's corners and is its longest path
// (a 4096-pair batch: 80.2 ms at 0.46, 81.7 ms at 0.40, 85.8 ms at 0.52,
// 86.1 ms at 0.58).  The pair is made on a context's first batch, so a
// context that never batches holds four hardware queues.
constexpr int kSumCUs = 8;
constexpr double kBroadShare = 0.70;
constexpr double kBatchBroadShare = 0.46;
constexpr uint32_t kBatchSumWaves = 128; // persistent waves of the batched serial sums (on the reserved CUs)
enum Stage {
    ST_SUMS,      // A: serial coordinate sums and raw boxes (k_mesh_sums alone)
    ST_PARAMS,    // A: shift parameters
    ST_SHIFT,     // A: shifted vertices, shifted face boxes
    ST_TOPOLOGY,  // C: polygon-soup edge numbering
    ST_BVH_SORT,  // B: raw face boxes, Morton codes, leaf sort
    ST_REFIT,     // B: implicit-tree boxes
    ST_TRAVERSE,  // B: raw-space broad phase (superset of the pairs)
    ST_FILTER,    // A: exact shifted-box test, pair hash, candidate faces
    ST_PLANES,    // A: candidate-face planes
    ST_TESTS,     // A: edge-face tests
    ST_NARROW,    // A: narrow phase
    ST_REGISTRY,  // A: registry sort, status
    ST_TRAV_KERNEL, // B: the traversal kernel alone, from its own dispatch timestamps
    ST_COUNT
};

const char* kStageNames[ST_COUNT] = {"coordinate_sums", "shift_params", "shift_boxes", "topology",
                                     "bvh_sort",        "bvh_refit",    "traversal",   "pair_filter",
                                     "face_planes",     "tests",        "narrow",      "registry",
                                     "traversal_kernel"};

// event slots: stage boundaries on the three streams
enum Ev { EV_A0, EV_S0, EV_A_SUMS, EV_A_PARAMS, EV_A_SHIFT, EV_A_BOXES, EV_A_JOIN, EV_A_FILTER, EV_A_PLANES, EV_A_JOIN2, EV_A_TESTS, EV_A_NARROW,
          EV_A_END, EV_FORK, EV_B0, EV_B_SORT, EV_B_REFIT, EV_B_TRAV, EV_C0, EV_C_FILL, EV_C_TOPO, EV_SUM_ITEMS,
          EV_SUM_CHAIN, EV_TRAV_K0, EV_TRAV_K1, EV_B_CC, EV_COUNT };

// Launch shape of the traversal kernel, chosen per workload class by the
// autotuner (mcgpu_autotune) or set by the caller (mcgpu_set_tuning).
struct TravConfig {
    int ll = kWaveLeavesLog;  // source leaves per block: 2^ll
    int stride = kTravStride; // cut-tree levels per breadth-first step
    int queues = kTravQueues; // 0: one wave per block; 1: the same in per-XCD runs of `chunk`; 8: persistent
                              // waves over XCD work queues
    int chunk = kTravRun;     // blocks a persistent wave takes per queue pop; with queues 1 the run length
};
constexpr int kTravWavesPerCU = 16; // persistent grid: four waves per SIMD (the kernel's occupancy)

// the traversal configurations the autotuner times
constexpr int kTuneLL[] = {3, 4, 5};
constexpr int kTuneStride[] = {2, 3, 4};
constexpr int kTuneQueue[][2] = {{0, 1}, {1, 16}, {8, 1}, {8, 4}}; // (queues, chunk)

// one step of a dispatch's replay program (mcgpu_set_graphs)
struct GraphOp {
    enum Kind { LAUNCH, RECORD, WAIT } kind;
    hipStream_t st;
    int ev;              // RECORD / WAIT: the context's event slot
    hipGraphExec_t exec; // LAUNCH
};

// one autotuned workload class: single pairs or batches, by the bit lengths of
// the two face counts
struct TuneEntry {
    bool batch;
    int bs, bc;
    TravConfig trav;
    mcgpu_tuning report;
};

uint64_t next_pow2(uint64_t x)
{
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

int bits_for(uint64_t n)
{
    int b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

// Serving: the request-independent parts of a resident source mesh, computed
// once by the server's master context and read by every request slot.
using ResidentGpu = ResidentGpuView; // sums, raw box as order keys, source edge count

struct Resident {
    uint32_t nv = 0, nf = 0, ne = 0;
    uint64_t ncorner = 0;
    const double* dv = nullptr;
    const uint32_t* didx = nullptr;
    const uint32_t* dsz = nullptr;
    const uint32_t* off = nullptr;   // face offsets
    const uint32_t* cedge = nullptr; // corner -> edge
    const uint32_t *ev0 = nullptr, *ev1 = nullptr;
    const int32_t *ef0 = nullptr, *ef1 = nullptr; // source edges [0, ne)
    const float* rawbox = nullptr;   // float leaf boxes of the source faces
    const uint32_t* sface = nullptr; // Morton order of the source faces
    const float* nodes = nullptr;    // source tree
    int depth = 0;
    ResidentGpu* dres = nullptr;
};

} // namespace

struct mcgpu_context_t {
    int device = 0;
    hipStream_t s = nullptr, s2 = nullptr, s3 = nullptr, ssum = nullptr;
    hipStream_t s2b = nullptr, s3b = nullptr; // B and C of batched dispatches, made by the first one
    hipEvent_t ev[EV_COUNT];
    hipError_t last_error = hipSuccess;

    Buf in_v[2], in_idx[2], in_sz[2];
    Buf off[2], state, ps_v, box; // state: DevState (counters and shift parameters)
    // pinned host copy of the state, enqueued right behind a dispatch's last
    // kernel so that finish() waits once instead of copying after the wait
    DevState* hstate = nullptr;
    bool state_copied = false;
    Buf hkeys[2], hfirst[2], eflag[2], eid[2], cedge[2], cface[2], etotal;
    Buf ev0, ev1, ef0, ef1;
    Buf mcode, mface, mcode2, mface2; // Morton keys and leaf order of both meshes
    Buf nodes[2], rawbox;
    Buf raw_pairs, pairs, phash, cand, cand_list, planes, big_list, big_off, big_scratch;
    Buf te, tg, tqs, trs, ttype, tpip, esc;
    Buf rkey, rval, rkey2, rval2, rp, re, rg, rpts, rcount, re_in, rg_in;
    Buf shkeys, shcnt;
    Buf pqs, prs, pcounts; // predicate stage: per-test signs, stage counts
    Buf scan_tmp, scan_tmp_c, hist;
    Buf sbsum, sitems, slo, shi, snitem, slist, sckpt, sflags; // coordinate sums from run items
    uint32_t sum_nblk = 0;

    uint64_t cap_pairs = 0, cap_tests = 0, cap_esc = 0, cap_reg = 0, cap_big = 0, cap_trip = 0;

    // state of the last dispatch
    bool have = false;
    bool broad_only = false; // the last call was mcgpu_broad_phase: only its pairs are valid
    uint32_t nv[2] = {0, 0}, nf[2] = {0, 0};
    uint64_t ncorner[2] = {0, 0};
    const double* dv[2] = {nullptr, nullptr};
    const uint32_t* didx[2] = {nullptr, nullptr};
    const uint32_t* dsz[2] = {nullptr, nullptr};
    int gbits = 1, kbits = 2;
    uint32_t flags = 0, enforcement_count = 0;
    double gp_constant = 1e-4;
    int keep_tests = 0;
    bool mface_alt = false;
    int depth[2] = {0, 0};
    bool reg_alt = false;
    Counters C;
    Params P;
    float stage_ms[ST_COUNT];
    bool stages_ready = false;
    int stage_mode = 0; // 0: full dispatch (three streams), 1: perturbed narrow-phase rerun
    bool stage_timing = false; // record the per-stage events (mcgpu_set_stage_timing)
    bool kernel_timing = false; // record the traversal kernel's dispatch timestamps only (mcgpu_set_stage_timing 2)
    bool last_timed = false;   // the last dispatch recorded them
    bool last_kernel_only = false; // ... only the traversal kernel's timestamps
    float total_ms = 0.f;
    int sort_bits = 22;  // leaf key bits

    // batches (mcgpu_dispatch_batch): the soup is every source mesh, then
    // every cut mesh, in pair order
    bool batch = false;
    uint32_t npairs = 0;
    int batch_mb = 7;                    // Morton bits per axis within a pair
    bool batch_packed = false;           // every pair's vertex ids and corner ranks fit the packed edge word
    bool shifted_bp = false;             // this dispatch's broad phase runs in shifted coordinates (Counters::loose)
    bool allow_loose = false;            // a single pair's own broad phase: a loose raw tolerance reruns it shifted
    bool cc_check = false;               // count each mesh's connected components (check_input_mesh)
    bool pending = false;                // an enqueued dispatch (mcgpu_dispatch_frontend_async) not yet waited for
    hipEvent_t ev_user = nullptr;        // recorded on the caller's stream before an enqueued dispatch
    Buf cc_parent;                       // union-find parents of the soup's vertices
    int batch_share_pct = (int)(kBatchBroadShare * 100.0 + 0.5); // split of the batch streams (made on the first batch)
    std::vector<BatchPair> hpairs;       // pair ranges (host copy, edge ranges read back)
    std::vector<uint32_t> hstarts, horder;
    std::vector<uint32_t> hflags, hrstart, hrend;
    std::vector<int32_t> hstatus;
    std::vector<uint32_t> hpartial;
    Buf bpairs, bstarts, fpair, vpair, bparams, brbox, pflags, border, bidx[2], rle, rlg, rstart, rend;
    const uint32_t* bin_idx[2] = {nullptr, nullptr}; // the caller's pair-local face lists

    // the general-position retry's perturbation draws, as the reference makes
    // them on its context's API thread (preproc.cpp:2621-2631): a
    // default_random_engine seeded with 1 seeds an mt19937, whose draws go
    // through uniform_real_distribution(-1, 1); the sequence continues across
    // the context's dispatches
    std::default_random_engine gp_seed_engine{1};
    std::mt19937 gp_twister{gp_seed_engine()};
    std::uniform_real_distribution<double> gp_uniform{-1.0, 1.0};

    // serving: a request slot of an mcgpu_server; one stream (s2, s3, s2b, s3b
    // and ssum are s), the source mesh is the server's resident one
    const Resident* res = nullptr;
    bool own_streams = true;

    // traversal launch shape of the current dispatch; the autotuned classes;
    // a caller's override
    int ncu = 0;                       // compute units of the device
    TravConfig trav;
    bool autotune = false;
    std::vector<TuneEntry> tuned;
    bool forced = false;
    TravConfig forced_trav;
    mcgpu_tuning used{};               // what the last dispatch ran with
    Buf tune_c;                        // counters of the tuner's traversal runs

    // graph replay (mcgpu_set_graphs): the recorded program and the
    // dispatch shape it was recorded for
    bool graphs = false;
    bool recording = false;
    hipError_t rec_error = hipSuccess;
    std::vector<GraphOp> prog;
    std::vector<uint64_t> prog_key;
    uint64_t replays = 0, recordings = 0;

    // cut-path edges of the last dispatch (mcgpu_cutpath_edges)
    Buf cp_order, cp_keys, cp_vals, cp_keys2, cp_vals2, cp_flag, cp_gid, cp_gstart, cp_ncand, cp_coff, cp_trip;
    Buf cp_src, cp_tgt, cp_fa, cp_fb, cp_group, cp_kind, cp_acc, cp_table, cp_keep, cp_pos;
    Buf cp_osrc, cp_otgt, cp_ofa, cp_ofb, cp_ogroup, cp_okind, cp_cnt;
    bool cp_have = false;
    CutCounts cp_counts{};
};

#define CK(x)                                                                                                          \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) {                                                                                        \
            ctx->last_error = e_;                                                                                      \
            return e_ == hipErrorOutOfMemory ? MCGPU_ERROR_OUT_OF_MEMORY : MCGPU_ERROR_DEVICE;                         \
        }                                                                                                              \
    } while (0)

namespace {

template <typename T>
hipError_t ensure(Buf& b, uint64_t count, T** out)
{
    size_t bytes = (size_t)std::max<uint64_t>(count, 1) * sizeof(T);
    if (b.cap < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        size_t want = bytes + bytes / 4;
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) return e;
        b.cap = want;
    }
    *out = (T*)b.p;
    return hipSuccess;
}

template <typename T>
T* P_(Buf& b)
{
    return (T*)b.p;
}

inline unsigned grid_for(uint64_t n, unsigned block = kBlock, unsigned maxg = 4096)
{
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > maxg) g = maxg;
    return (unsigned)g;
}

__global__ void k_reset_narrow(Counters* C)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        C->n_tests = C->n_esc = C->n_reg = C->n_big = C->n_trip = 0; // candidate faces are kept
        C->ov_trip = 0;
        C->n_f32 = C->n_f64 = 0;
        C->ov_tests = C->ov_esc = C->ov_reg = C->ov_big = 0;
        C->gp = C->bad_src = C->bad_cut = C->unsupported = C->pair_fail = C->partial = 0;
    }
}

void fill_u32(void* p, uint32_t v, uint64_t n, hipStream_t s)
{
    k_fill_u32<<<grid_for(n), kBlock, 0, s>>>((uint32_t*)p, v, n);
}

void fill_u64(void* p, uint64_t v, uint64_t n, hipStream_t s)
{
    k_fill_u64<<<grid_for(n), kBlock, 0, s>>>((unsigned long long*)p, (unsigned long long)v, n);
}

mcgpu_result check_mesh(const mcgpu_mesh* m)
{
    if (!m || !m->vertices || !m->face_indices || m->num_vertices < 3 || m->num_faces < 1)
        return MCGPU_ERROR_INVALID_VALUE;
    return MCGPU_SUCCESS;
}

int32_t decide_status(bool bad_cut, bool bad_src, bool gp, uint64_t n_reg, bool partial, bool pair_fail, uint32_t flags,
                      uint32_t enforcement_count)
{
    const bool enforce =
        (flags & (MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION | MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION_ABSOLUTE)) != 0;
    // kernel.cpp:2320-2331 (zero-area candidate face), 2532-2604 (narrow-phase
    // general-position violation), 3211-3229 (empty registry), 3269-3278
    // (partial cut with through-cuts required), 3439-3452 (fewer than two
    // intersection points on a face pair)
    if (bad_cut) return MCGPU_STATUS_INVALID_CUT_MESH;
    if (bad_src) return MCGPU_STATUS_INVALID_SRC_MESH;
    if (gp) return MCGPU_STATUS_GENERAL_POSITION_VIOLATION;
    if (n_reg == 0)
        return (enforce && enforcement_count > 0) ? MCGPU_STATUS_GENERAL_POSITION_VIOLATION : MCGPU_STATUS_SUCCESS;
    if (partial && (flags & MCGPU_DISPATCH_REQUIRE_THROUGH_CUTS)) return MCGPU_STATUS_SUCCESS;
    if (pair_fail)
        return enforce ? MCGPU_STATUS_GENERAL_POSITION_VIOLATION : MCGPU_STATUS_INVALID_MESH_INTERSECTION;
    return MCGPU_STATUS_SUCCESS;
}

int32_t decide_status(const Counters& C, uint32_t flags, uint32_t enforcement_count)
{
    return decide_status(C.bad_cut, C.bad_src, C.gp, C.n_reg, C.partial, C.pair_fail, flags, enforcement_count);
}

Soup make_soup(mcgpu_context ctx)
{
    Soup S;
    for (int m = 0; m < 2; ++m) {
        S.idx[m] = ctx->didx[m];
        S.off[m] = P_<uint32_t>(ctx->off[m]);
        S.nv[m] = ctx->nv[m];
        S.nf[m] = ctx->nf[m];
        S.cedge[m] = P_<uint32_t>(ctx->cedge[m]);
    }
    if (ctx->res) { // the resident source's face offsets and corner edges
        S.off[0] = ctx->res->off;
        S.cedge[0] = ctx->res->cedge;
    }
    S.ps_v = P_<double>(ctx->ps_v);
    S.box = P_<double>(ctx->box);
    S.ev0 = P_<uint32_t>(ctx->ev0);
    S.ev1 = P_<uint32_t>(ctx->ev1);
    S.ef0 = P_<int32_t>(ctx->ef0);
    S.ef1 = P_<int32_t>(ctx->ef1);
    S.fpair = ctx->batch ? P_<uint32_t>(ctx->fpair) : nullptr;
    S.pflags = ctx->batch ? P_<uint32_t>(ctx->pflags) : nullptr;
    return S;
}

// buffers whose size depends only on the meshes
// a stage-boundary event: recorded only when stage timing is on.  Recorded
// events cost host and device time (about 50 us a dispatch for thirteen of
// them on the icosphere-7 pair); the events that order the streams and
// EV_A0 / EV_A_END (device_ms) are always recorded.
void stage_event(mcgpu_context ctx, int e, hipStream_t st)
{
    if (ctx->stage_timing) (void)hipEventRecord(ctx->ev[e], st);
}

mcgpu_result alloc_mesh_buffers(mcgpu_context ctx)
{
    const uint64_t F = (uint64_t)ctx->nf[0] + ctx->nf[1];
    const uint64_t V = (uint64_t)ctx->nv[0] + ctx->nv[1];
    const uint64_t K = ctx->ncorner[0] + ctx->ncorner[1];
    void* p;
    DevState* st;
    CK(ensure(ctx->state, 1, &st));
    double* dp;
    CK(ensure(ctx->ps_v, 3 * V, &dp));
    CK(ensure(ctx->box, 6 * F, &dp));
    float* fp;
    CK(ensure(ctx->rawbox, 6 * F, &fp));
    uint32_t* u;
    int32_t* i32;
    unsigned long long* ull;
    CK(ensure(ctx->ev0, K, &u));
    CK(ensure(ctx->ev1, K, &u));
    CK(ensure(ctx->ef0, K, &i32));
    CK(ensure(ctx->ef1, K, &i32));
    CK(ensure(ctx->etotal, 2, &u));
    uint64_t scan_need = scan_tmp_elems(K + 8);
    // the single pair's packed table and the two-table path hash over a power
    // of two; a packed batch takes 2 slots per corner (one region per pair)
    // and has no first-use table
    const bool packed_batch = ctx->batch && ctx->batch_packed;
    for (int m = 0; m < 2; ++m) {
        const uint64_t nk = ctx->ncorner[m], nf = ctx->nf[m];
        const uint64_t hs = packed_batch ? 2 * nk + 8 : next_pow2(2 * nk + 2);
        CK(ensure(ctx->off[m], nf + 1, &u));
        CK(ensure(ctx->hkeys[m], hs, &ull));
        if (!packed_batch) CK(ensure(ctx->hfirst[m], hs, &ull));
        CK(ensure(ctx->cface[m], nk, &u));
        CK(ensure(ctx->eflag[m], nk, &u));
        CK(ensure(ctx->eid[m], nk, &u));
        CK(ensure(ctx->cedge[m], nk, &u));
        int D = 0;
        while ((1ull << D) < nf) ++D;
        ctx->depth[m] = D;
        CK(ensure(ctx->nodes[m], 6 * ((2ull << D) - 1), &fp));
        scan_need = std::max<uint64_t>(scan_need, scan_tmp_elems(nf + 8));
    }
    CK(ensure(ctx->mcode, F, &u));
    CK(ensure(ctx->mface, F, &u));
    CK(ensure(ctx->mcode2, F, &u));
    CK(ensure(ctx->mface2, F, &u));
    scan_need = std::max<uint64_t>(scan_need, scan_tmp_elems(F + 8));
    CK(ensure(ctx->cand, F + 8, &u));
    CK(ensure(ctx->cand_list, F + 8, &u));
    Plane* pl;
    CK(ensure(ctx->planes, F, &pl));
    (void)p;
    CK(ensure(ctx->scan_tmp, scan_need + 1024, &u));
    if (!ctx->batch) { // coordinate sums from run items (k_sum_blocks .. k_sum_verify), for the chains computed here
        const uint64_t nmax = std::max<uint64_t>(ctx->res ? 0 : ctx->nv[0], ctx->nv[1]);
        const uint32_t nblk = (uint32_t)std::max<uint64_t>(1, (nmax + kSumTile - 1) / kSumTile);
        const uint64_t region = 6ull * nblk * kSumTile, lcap = (uint64_t)nblk * (kSumTile / 64);
        ctx->sum_nblk = nblk;
        CK(ensure(ctx->sbsum, 6ull * nblk, &dp));
        CK(ensure(ctx->sitems, region, &dp));
        CK(ensure(ctx->slo, region, &dp));
        CK(ensure(ctx->shi, region, &dp));
        CK(ensure(ctx->snitem, 6ull * nblk, &u));
        CK(ensure(ctx->slist, 6 * lcap, &u));
        CK(ensure(ctx->sckpt, 6 * lcap, &dp));
        CK(ensure(ctx->sflags, 16, &u));
    }
    // the packed edge numbering scans both meshes' first-use flags as one array
    const uint64_t K2 = ctx->ncorner[0] + ctx->ncorner[1];
    CK(ensure(ctx->eflag[0], K2 + 8, &u));
    CK(ensure(ctx->eid[0], K2 + 8, &u));
    if (ctx->cc_check) CK(ensure(ctx->cc_parent, (uint64_t)ctx->nv[0] + ctx->nv[1] + 8, &u));
    CK(ensure(ctx->scan_tmp_c, scan_tmp_elems(K2 + 8) + 1024, &u));
    return MCGPU_SUCCESS;
}

mcgpu_result alloc_list_buffers(mcgpu_context ctx)
{
    uint32_t* u;
    unsigned long long* ull;
    double* dp;
    CK(ensure(ctx->raw_pairs, ctx->cap_pairs, &ull));
    CK(ensure(ctx->pairs, ctx->cap_pairs, &ull));
    CK(ensure(ctx->phash, next_pow2(2 * ctx->cap_pairs), &ull));
    CK(ensure(ctx->te, ctx->cap_tests, &u));
    CK(ensure(ctx->tg, ctx->cap_tests, &u));
    if (ctx->keep_tests) {
        int8_t* i8;
        char* c8;
        CK(ensure(ctx->tqs, ctx->cap_tests, &i8));
        CK(ensure(ctx->trs, ctx->cap_tests, &i8));
        CK(ensure(ctx->ttype, ctx->cap_tests, &c8));
        CK(ensure(ctx->tpip, ctx->cap_tests, &c8));
    }
    CK(ensure(ctx->esc, ctx->cap_esc, &u));
    CK(ensure(ctx->rkey, ctx->cap_reg, &ull));
    CK(ensure(ctx->rkey2, ctx->cap_reg, &ull));
    CK(ensure(ctx->rval, ctx->cap_reg, &u));
    CK(ensure(ctx->rval2, ctx->cap_reg, &u));
    CK(ensure(ctx->rp, 3 * ctx->cap_reg, &dp));
    CK(ensure(ctx->re, ctx->cap_reg, &u));
    CK(ensure(ctx->rg, ctx->cap_reg, &u));
    CK(ensure(ctx->rpts, 3 * ctx->cap_reg, &dp));
    CK(ensure(ctx->re_in, ctx->cap_reg, &u));
    CK(ensure(ctx->rg_in, ctx->cap_reg, &u));
    if (ctx->batch) {
        CK(ensure(ctx->rle, ctx->cap_reg, &u));
        CK(ensure(ctx->rlg, ctx->cap_reg, &u));
    }
    CK(ensure(ctx->rcount, 1, &ull));
    CK(ensure(ctx->shkeys, next_pow2(4 * ctx->cap_reg), &ull));
    CK(ensure(ctx->shcnt, next_pow2(4 * ctx->cap_reg), &u));
    CK(ensure(ctx->big_list, ctx->cap_big, &u));
    CK(ensure(ctx->big_off, ctx->cap_big, &ull));
    Trip* tr;
    CK(ensure(ctx->big_scratch, ctx->cap_trip, &tr));
    const uint64_t sort_cap = std::max<uint64_t>(ctx->cap_reg, (uint64_t)ctx->nf[0] + ctx->nf[1]);
    CK(ensure(ctx->hist, radix_hist_elems(sort_cap), &u));
    // the scan scratch also serves the mesh-sized scans (sized in alloc_mesh_buffers)
    const uint64_t K = ctx->ncorner[0] + ctx->ncorner[1];
    const uint64_t need = std::max<uint64_t>(scan_tmp_elems(radix_hist_elems(sort_cap)), scan_tmp_elems(K + 8));
    CK(ensure(ctx->scan_tmp, need + 1024, &u));
    return MCGPU_SUCCESS;
}

SumBufs sum_bufs(mcgpu_context ctx)
{
    SumBufs B;
    B.v[0] = ctx->res ? nullptr : ctx->dv[0];
    B.v[1] = ctx->dv[1];
    B.n[0] = ctx->nv[0];
    B.n[1] = ctx->nv[1];
    B.nblk = ctx->sum_nblk;
    B.bsum = P_<double>(ctx->sbsum);
    B.items = P_<double>(ctx->sitems);
    B.lo = P_<double>(ctx->slo);
    B.hi = P_<double>(ctx->shi);
    B.nitem = P_<uint32_t>(ctx->snitem);
    B.list = P_<uint32_t>(ctx->slist);
    B.lcap = ctx->sum_nblk * (kSumTile / 64);
    B.ckpt = P_<double>(ctx->sckpt);
    B.fail = P_<uint32_t>(ctx->sflags);
    B.nb64 = P_<uint32_t>(ctx->sflags) + 8;
    return B;
}

// one packed word per edge (frontend_kernels.hpp, TopoMesh) when every vertex
// id fits 21 bits and every corner rank 22 bits; batches keep the two tables
bool topology_packed(mcgpu_context ctx)
{
    if (ctx->batch) return false;
    for (int m = ctx->res ? 1 : 0; m < 2; ++m)
        if (ctx->nv[m] >= (1u << kPackedVertexBits) || ctx->ncorner[m] >= (1ull << kPackedRankBits)) return false;
    return true;
}

// the part of the pipeline that runs again for a perturbed cut mesh (whose
// candidate pairs and faces are those of the unperturbed dispatch):
// candidate-face planes, tests, narrow phase, registry, status
// ---------------------------------------------------------------------------
// Graph replay of a dispatch (mcgpu_set_graphs).  The enqueue functions order
// their four streams only through ev_record / ev_wait.  When a context
// records, every stream is under stream capture and each of those calls cuts
// its stream's capture: the kernels captured since the stream's previous cut
// become one instantiated graph.  The dispatch is then a program of graph
// launches, event records and event waits in the order the eager enqueue
// issues them, each graph launched on the stream it was captured from, so its
// kernels keep that stream's CU mask; replaying the program is the dispatch.
// ---------------------------------------------------------------------------

void graph_cut(mcgpu_context ctx, hipStream_t st)
{
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (e == hipSuccess && g) {
        size_t n = 0;
        e = hipGraphGetNodes(g, nullptr, &n);
        if (e == hipSuccess && n) {
            hipGraphExec_t ex = nullptr;
            e = hipGraphInstantiateWithFlags(&ex, g, 0);
            if (e == hipSuccess) ctx->prog.push_back(GraphOp{GraphOp::LAUNCH, st, -1, ex});
        }
        (void)hipGraphDestroy(g);
    }
    if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess && ctx->rec_error == hipSuccess) ctx->rec_error = e;
}

void ev_record(mcgpu_context ctx, int e, hipStream_t st)
{
    if (!ctx->recording) {
        (void)hipEventRecord(ctx->ev[e], st);
        return;
    }
    graph_cut(ctx, st);
    ctx->prog.push_back(GraphOp{GraphOp::RECORD, st, e, nullptr});
}

void ev_wait(mcgpu_context ctx, hipStream_t st, int e)
{
    if (!ctx->recording) {
        (void)hipStreamWaitEvent(st, ctx->ev[e], 0);
        return;
    }
    graph_cut(ctx, st);
    ctx->prog.push_back(GraphOp{GraphOp::WAIT, st, e, nullptr});
}

void enqueue_narrow(mcgpu_context ctx, const Soup& S, bool join_topology)
{
    hipStream_t s = ctx->s;
    const uint64_t F = (uint64_t)ctx->nf[0] + ctx->nf[1];
    Counters* C = &P_<DevState>(ctx->state)->C;
    k_planes<<<grid_for(F), kBlock, 0, s>>>(S, P_<uint32_t>(ctx->cand_list), P_<Plane>(ctx->planes),
                                             P_<uint32_t>(ctx->big_list), P_<unsigned long long>(ctx->big_off),
                                             ctx->cap_big, C);
    if (ctx->dsz[0] || ctx->dsz[1]) // faces of more than 5 vertices: none in two triangle meshes
        k_big_polygons<<<(unsigned)std::min<uint64_t>(std::max<uint64_t>(ctx->cap_big, 1), 1024), 64, 0, s>>>(
            S, P_<uint32_t>(ctx->big_list), P_<unsigned long long>(ctx->big_off), C, ctx->cap_big, ctx->cap_trip,
            P_<Plane>(ctx->planes), P_<Trip>(ctx->big_scratch), C);
    stage_event(ctx, EV_A_PLANES, s);
    if (join_topology) ev_wait(ctx, s, EV_C_TOPO);
    stage_event(ctx, EV_A_JOIN2, s);
    k_tests<<<grid_for(6 * ctx->cap_pairs, kTestTile, 2048), kStageBlock, 0, s>>>(S, P_<unsigned long long>(ctx->pairs), ctx->cap_pairs,
                                                        P_<unsigned long long>(ctx->phash), P_<uint32_t>(ctx->te),
                                                        P_<uint32_t>(ctx->tg), ctx->cap_tests, C);
    stage_event(ctx, EV_A_TESTS, s);
    TestOut TO{nullptr, nullptr, nullptr, nullptr};
    if (ctx->keep_tests)
        TO = TestOut{P_<int8_t>(ctx->tqs), P_<int8_t>(ctx->trs), P_<char>(ctx->ttype), P_<char>(ctx->tpip)};
    RegOut R{P_<unsigned long long>(ctx->rkey), P_<double>(ctx->rp), P_<uint32_t>(ctx->rval), P_<uint32_t>(ctx->re_in),
             P_<uint32_t>(ctx->rg_in), ctx->cap_reg, ctx->gbits,
             ctx->batch ? P_<BatchPair>(ctx->bpairs) : nullptr};
    k_narrow<<<grid_for(ctx->cap_tests, kNarrowTile, 2048), kStageBlock, 0, s>>>(S, P_<Plane>(ctx->planes), P_<uint32_t>(ctx->te),
                                                         P_<uint32_t>(ctx->tg), ctx->cap_tests, P_<uint32_t>(ctx->esc),
                                                         ctx->cap_esc, TO, R, C);
    k_narrow_exact<<<grid_for(ctx->cap_esc, 64, 2048), 64, 0, s>>>(S, P_<Plane>(ctx->planes), P_<uint32_t>(ctx->te),
                                                                   P_<uint32_t>(ctx->tg), P_<uint32_t>(ctx->esc),
                                                                   ctx->cap_esc, TO, R, C);
    stage_event(ctx, EV_A_NARROW, s);
    unsigned long long* rcount = P_<unsigned long long>(ctx->rcount);
    const uint64_t hs = next_pow2(4 * ctx->cap_reg);
    unsigned long long* hk = P_<unsigned long long>(ctx->shkeys);
    uint32_t* hc = P_<uint32_t>(ctx->shcnt);
    if (ctx->cap_reg <= kRegMergeTiles * kRegTile) {
        // tile sort + merge by rank: two launches
        const unsigned tiles = (unsigned)((ctx->cap_reg + kRegTile - 1) / kRegTile);
        k_reg_tile_sort<<<tiles, kRegSortThreads, 0, s>>>(P_<unsigned long long>(ctx->rkey), P_<uint32_t>(ctx->rval), &C->n_reg,
                                                         ctx->cap_reg, P_<unsigned long long>(ctx->rkey2),
                                                         P_<uint32_t>(ctx->rval2), hk, hc, hs, ctx->kbits);
        k_reg_merge<<<grid_for(ctx->cap_reg), kBlock, 0, s>>>(S, P_<unsigned long long>(ctx->rkey2), P_<uint32_t>(ctx->rval2),
                                                              &C->n_reg, ctx->cap_reg, R.e, R.g, P_<double>(ctx->rp),
                                                              P_<uint32_t>(ctx->re), P_<uint32_t>(ctx->rg),
                                                              P_<double>(ctx->rpts), hk, hc, hs, C);
        k_status_check<<<grid_for(hs), kBlock, 0, s>>>(S, &C->n_reg, hk, hc, hs, C);
    } else {
        // large registries: LSD radix sort of the keys with the entry index as payload
        k_clamp_count<<<1, 1, 0, s>>>(&C->n_reg, ctx->cap_reg, rcount);
        RadixTemp rt{P_<uint32_t>(ctx->hist), P_<uint32_t>(ctx->scan_tmp)};
        const bool alt = radix_sort<unsigned long long>(P_<unsigned long long>(ctx->rkey), P_<uint32_t>(ctx->rval),
                                                        P_<unsigned long long>(ctx->rkey2), P_<uint32_t>(ctx->rval2),
                                                        (const uint64_t*)rcount, 0, ctx->cap_reg, ctx->kbits, rt, s);
        const uint32_t* sv = P_<uint32_t>(alt ? ctx->rval2 : ctx->rval);
        k_registry_gather<<<grid_for(ctx->cap_reg), kBlock, 0, s>>>(sv, R.e, R.g, P_<double>(ctx->rp), rcount, ctx->cap_reg,
                                                                     P_<uint32_t>(ctx->re), P_<uint32_t>(ctx->rg),
                                                                     P_<double>(ctx->rpts));
        k_status_clear<<<grid_for(hs), kBlock, 0, s>>>(rcount, hk, hc, hs);
        k_status_pairs<<<grid_for(ctx->cap_reg), kBlock, 0, s>>>(S, P_<uint32_t>(ctx->re), P_<uint32_t>(ctx->rg), rcount, hs,
                                                                 hk, hc, C);
        k_status_check<<<grid_for(hs), kBlock, 0, s>>>(S, rcount, hk, hc, hs, C);
    }
    if (ctx->batch) {
        // pair-local ids and each pair's range of the pair-major registry
        fill_u32(ctx->rstart.p, 0u, ctx->npairs, s);
        fill_u32(ctx->rend.p, 0u, ctx->npairs, s);
        k_batch_localize<<<grid_for(ctx->cap_reg), kBlock, 0, s>>>(
            P_<uint32_t>(ctx->re), P_<uint32_t>(ctx->rg), &C->n_reg, ctx->cap_reg, ctx->nf[0], P_<uint32_t>(ctx->fpair),
            P_<BatchPair>(ctx->bpairs), C, P_<uint32_t>(ctx->rle), P_<uint32_t>(ctx->rlg), P_<uint32_t>(ctx->rstart),
            P_<uint32_t>(ctx->rend));
    }
    if (ctx->cc_check && ctx->stage_mode == 0) ev_wait(ctx, s, EV_B_CC); // the component counts are in
    ev_record(ctx, EV_A_END, s);
}

// Per-dispatch reset, raw bounding boxes, face offsets, input validation and
// the broad-phase tolerance, all on stream A ahead of the fork.  Nothing waits
// for the host: the kernels that first read coordinates through the face
// lists write empty boxes when the input was rejected, so no later kernel
// indexes a vertex through a bad face list, and the host reports the
// rejection after the dispatch.
void enqueue_validate(mcgpu_context ctx)
{
    ctx->state_copied = false; // a state copy belongs to the dispatch it follows
    ctx->last_timed = ctx->stage_timing || ctx->kernel_timing;
    ctx->last_kernel_only = !ctx->stage_timing && ctx->kernel_timing;
    hipStream_t s = ctx->s;
    Counters* C = &P_<DevState>(ctx->state)->C;
    uint32_t* etotal = P_<uint32_t>(ctx->etotal);
    ev_record(ctx, EV_A0, s);
    if (!ctx->dsz[0] && !ctx->dsz[1]) { // two triangle meshes: three launches (k_reset_counters)
        const int m0 = ctx->res ? 1 : 0;
        k_reset_counters<<<1, kBlock, 0, s>>>(C, m0);
        if (ctx->res) k_resident_source<<<1, 1, 0, s>>>(C, &P_<DevState>(ctx->state)->P, ctx->res->dres);
        const uint32_t n0 = m0 ? 0u : ctx->nv[0], f0 = m0 ? 0u : ctx->nf[0];
        k_raw_bbox2<<<grid_for((uint64_t)n0 + ctx->nv[1], kBlock, 64), kBlock, 0, s>>>(m0 ? nullptr : ctx->dv[0], n0,
                                                                                         ctx->dv[1], ctx->nv[1], C);
        k_validate_tri2<<<grid_for((uint64_t)f0 + ctx->nf[1]), kBlock, 0, s>>>(
            ctx->didx[0], P_<uint32_t>(ctx->off[0]), f0, ctx->nv[0], ctx->didx[1], P_<uint32_t>(ctx->off[1]), ctx->nf[1],
            ctx->nv[1], C);
    } else {
    fill_u32(C, 0u, sizeof(Counters) / 4, s);
    if (ctx->res) k_resident_source<<<1, 1, 0, s>>>(C, &P_<DevState>(ctx->state)->P, ctx->res->dres);
    for (int m = 0; m < 2; ++m) {
        if (ctx->res && m == 0) continue; // resident source: validated, boxed and summed once
        fill_u64(C->rbox[m], ~0ull, 3, s);
        k_raw_bbox<<<grid_for(ctx->nv[m], kBlock, 64), kBlock, 0, s>>>(ctx->dv[m], ctx->nv[m], C->rbox[m]);
        uint32_t* off = P_<uint32_t>(ctx->off[m]);
        if (ctx->dsz[m]) {
            scan_exclusive(ctx->dsz[m], off, ctx->nf[m], P_<uint32_t>(ctx->scan_tmp), etotal + m, s);
            k_offsets_tail<<<1, 1, 0, s>>>(off, ctx->nf[m], etotal + m);
        } else {
            k_triangle_offsets<<<grid_for((uint64_t)ctx->nf[m] + 1), kBlock, 0, s>>>(off, ctx->nf[m]);
        }
        k_validate<<<grid_for(ctx->nf[m]), kBlock, 0, s>>>(ctx->didx[m], off, ctx->dsz[m], ctx->nf[m], ctx->nv[m], C);
    }
    }
    const bool absolute = (ctx->flags & MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION_ABSOLUTE) != 0;
    k_tolerance<<<1, 1, 0, s>>>(&P_<DevState>(ctx->state)->P, C, ctx->nv[0], ctx->nv[1], ctx->gp_constant,
                                absolute ? 1 : 0, (ctx->allow_loose && !ctx->res) ? 1 : 0);
}

// Batches: the same reset and validation, per pair: segment ids of every
// vertex and face, face lists checked against their own pair's vertex count
// and rewritten as ids of the concatenated part, raw boxes of every (pair,
// mesh), each pair's broad-phase tolerance, and the cell layout.
void enqueue_validate_batch(mcgpu_context ctx)
{
    ctx->state_copied = false;
    ctx->last_timed = ctx->stage_timing || ctx->kernel_timing;
    ctx->last_kernel_only = !ctx->stage_timing && ctx->kernel_timing;
    hipStream_t s = ctx->s;
    Counters* C = &P_<DevState>(ctx->state)->C;
    uint32_t* etotal = P_<uint32_t>(ctx->etotal);
    const uint32_t np = ctx->npairs;
    const uint32_t* st = P_<uint32_t>(ctx->bstarts); // vertex starts (source, cut), face starts (source, cut)
    uint32_t* fp = P_<uint32_t>(ctx->fpair);
    uint32_t* vp = P_<uint32_t>(ctx->vpair);
    const BatchPair* bp = P_<BatchPair>(ctx->bpairs);
    auto* brbox = (unsigned long long(*)[2][6])ctx->brbox.p;
    Params* Pb = P_<Params>(ctx->bparams);
    ev_record(ctx, EV_A0, s);
    fill_u32(C, 0u, sizeof(Counters) / 4, s);
    fill_u32(ctx->pflags.p, 0u, np, s);
    k_segment_ids<<<grid_for(ctx->nv[0]), kBlock, 0, s>>>(st, np, ctx->nv[0], vp);
    k_segment_ids<<<grid_for(ctx->nv[1]), kBlock, 0, s>>>(st + np, np, ctx->nv[1], vp + ctx->nv[0]);
    k_segment_ids<<<grid_for(ctx->nf[0]), kBlock, 0, s>>>(st + 2 * np, np, ctx->nf[0], fp);
    k_segment_ids<<<grid_for(ctx->nf[1]), kBlock, 0, s>>>(st + 3 * np, np, ctx->nf[1], fp + ctx->nf[0]);
    for (int m = 0; m < 2; ++m) {
        uint32_t* off = P_<uint32_t>(ctx->off[m]);
        if (ctx->dsz[m]) {
            scan_exclusive(ctx->dsz[m], off, ctx->nf[m], P_<uint32_t>(ctx->scan_tmp), etotal + m, s);
            k_offsets_tail<<<1, 1, 0, s>>>(off, ctx->nf[m], etotal + m);
        } else {
            k_triangle_offsets<<<grid_for((uint64_t)ctx->nf[m] + 1), kBlock, 0, s>>>(off, ctx->nf[m]);
        }
        k_batch_globalize<<<grid_for(ctx->nf[m]), kBlock, 0, s>>>(ctx->bin_idx[m], off, ctx->dsz[m], ctx->nf[m], m,
                                                                  m ? ctx->nf[0] : 0u, fp, bp, P_<uint32_t>(ctx->bidx[m]), C);
    }
    k_batch_bbox<<<2 * np, 256, 0, s>>>(ctx->dv[0], ctx->dv[1], bp, brbox);
    const bool absolute = (ctx->flags & MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION_ABSOLUTE) != 0;
    k_batch_tolerance<<<grid_for(np), kBlock, 0, s>>>(Pb, brbox, bp, np, ctx->gp_constant, absolute ? 1 : 0, C);
    k_batch_layout<<<1, 1024, 0, s>>>(Pb, brbox, np);
}

// per-pair results of a batch, after finish(): status bits and registry
// ranges read back, statuses decided as for one pair
mcgpu_result finish_batch(mcgpu_context ctx, mcgpu_frontend_info* info)
{
    const uint32_t np = ctx->npairs;
    ctx->hflags.resize(np);
    ctx->hrstart.resize(np);
    ctx->hrend.resize(np);
    ctx->hstatus.resize(np);
    ctx->hpartial.resize(np);
    CK(hipMemcpy(ctx->hflags.data(), ctx->pflags.p, 4ull * np, hipMemcpyDeviceToHost));
    CK(hipMemcpy(ctx->hrstart.data(), ctx->rstart.p, 4ull * np, hipMemcpyDeviceToHost));
    CK(hipMemcpy(ctx->hrend.data(), ctx->rend.p, 4ull * np, hipMemcpyDeviceToHost));
    CK(hipMemcpy(ctx->hpairs.data(), ctx->bpairs.p, sizeof(BatchPair) * np, hipMemcpyDeviceToHost));
    int32_t first = MCGPU_STATUS_SUCCESS;
    for (uint32_t p = 0; p < np; ++p) {
        const uint32_t f = ctx->hflags[p];
        const int32_t st = decide_status((f & PF_BAD_CUT) != 0, (f & PF_BAD_SRC) != 0, (f & PF_GP) != 0,
                                         ctx->hrend[p] - ctx->hrstart[p], (f & PF_PARTIAL) != 0,
                                         (f & PF_PAIR_FAIL) != 0, ctx->flags, ctx->enforcement_count);
        ctx->hstatus[p] = st;
        ctx->hpartial[p] = (st == MCGPU_STATUS_SUCCESS && ctx->hrend[p] > ctx->hrstart[p] && (f & PF_PARTIAL)) ? 1u : 0u;
        if (first == MCGPU_STATUS_SUCCESS) first = st;
    }
    if (info) {
        info->status = first;
        for (int j = 0; j < 3; ++j) info->com[j] = info->translation[j] = 0.0;
        info->eps = 0.0;
    }
    return MCGPU_SUCCESS;
}

// The device arrays of the trees and the raw-space traversal: leaf keys and
// their sorted copies, single-precision leaf boxes, both trees, the sort's
// scratch and the raw pair list.  A resident source (serving) brings its own
// sorted leaves and refitted tree (src_*).
struct BroadView {
    uint32_t *mcode, *mface, *mcode2, *mface2;
    float* rawbox;
    float* nodes[2];
    RadixTemp rt;
    unsigned long long* raw_pairs;
    uint64_t cap_pairs;
    uint32_t nf[2];
    int depth[2];
    int sort_bits;
    const float* src_rawbox;  // resident source, else null
    const uint32_t* src_sface;
    const float* src_nodes;
    int src_depth;
    TravConfig trav; // traversal launch shape
    int cus;         // compute units of the stream the traversal runs on
};

template <int LL, int ST>
void launch_trav_t(dim3 grid, hipStream_t st, hipEvent_t a, hipEvent_t b, const Tree& src, const Tree& cut,
                   unsigned long long* pairs, uint64_t cap, Counters* C, TravQueue q)
{
    if (a) // the launch itself records the kernel's start and end (its dispatch timestamps)
        hipExtLaunchKernelGGL(k_traverse_wave<LL, ST>, grid, dim3(64), 0, st, a, b, 0, src, cut, pairs, cap, C, q);
    else
        k_traverse_wave<LL, ST><<<grid, 64, 0, st>>>(src, cut, pairs, cap, C, q);
}

// The traversal of src against cut with launch shape t on stream st, whose
// mask holds `cus` compute units.  One wave per source block, or a persistent
// grid of kTravWavesPerCU waves per CU over the work queues (whose heads the
// per-dispatch counter reset cleared).
void launch_traversal(const TravConfig& t, int cus, const Tree& src, const Tree& cut, unsigned long long* pairs,
                      uint64_t cap, Counters* C, hipStream_t st, hipEvent_t a = nullptr, hipEvent_t b = nullptr)
{
    const uint64_t nblocks = src.D >= t.ll ? (((uint64_t)src.n + (1u << t.ll) - 1) >> t.ll) : 1;
    const TravQueue q{t.queues, t.chunk < 1 ? 1 : t.chunk};
    uint64_t g = q.queues > 1 ? std::min<uint64_t>((nblocks + q.chunk - 1) / q.chunk,
                                               (uint64_t)std::max(cus, 1) * kTravWavesPerCU)
                          : std::min<uint64_t>(nblocks, 65536);
    if (q.queues == 1) { // whole runs of q.chunk blocks on every XCD; the surplus workgroups find no block
        const uint64_t unit = (uint64_t)kXcds * q.chunk;
        g = (g + unit - 1) / unit * unit;
    }
    const dim3 grid((unsigned)std::max<uint64_t>(g, 1));
    switch (t.ll * 16 + t.stride) {
    case 3 * 16 + 2: launch_trav_t<3, 2>(grid, st, a, b, src, cut, pairs, cap, C, q); break;
    case 3 * 16 + 3: launch_trav_t<3, 3>(grid, st, a, b, src, cut, pairs, cap, C, q); break;
    case 3 * 16 + 4: launch_trav_t<3, 4>(grid, st, a, b, src, cut, pairs, cap, C, q); break;
    case 4 * 16 + 2: launch_trav_t<4, 2>(grid, st, a, b, src, cut, pairs, cap, C, q); break;
    case 4 * 16 + 3: launch_trav_t<4, 3>(grid, st, a, b, src, cut, pairs, cap, C, q); break;
    case 4 * 16 + 4: launch_trav_t<4, 4>(grid, st, a, b, src, cut, pairs, cap, C, q); break;
    case 5 * 16 + 2: launch_trav_t<5, 2>(grid, st, a, b, src, cut, pairs, cap, C, q); break;
    case 5 * 16 + 3: launch_trav_t<5, 3>(grid, st, a, b, src, cut, pairs, cap, C, q); break;
    default: launch_trav_t<5, 4>(grid, st, a, b, src, cut, pairs, cap, C, q); break; // the default shape
    static_assert(kWaveLeavesLog == 5 && kTravStride == 4, "the switch's default is the default shape");
    }
}

bool trav_config_valid(const TravConfig& t)
{
    return t.ll >= 3 && t.ll <= 5 && t.stride >= 2 && t.stride <= 4 && (t.queues == 0 || t.queues == 1 || t.queues == 8) &&
           t.chunk >= 1 && t.chunk <= 64;
}

// The two trees of a broad view once its leaves are sorted (into mface2 when
// `alt`): refit state R and traversal view T.
void trees_of(const BroadView& B, bool alt, Tree* T, RefitTree* R)
{
    const uint32_t Fs = B.nf[0];
    const uint32_t* sorted = alt ? B.mface2 : B.mface;
    for (int m = 0; m < 2; ++m) {
        R[m].facebox = B.rawbox + 6 * (uint64_t)(m ? Fs : 0);
        R[m].sface = sorted + (m ? Fs : 0);
        R[m].nodes = B.nodes[m];
        R[m].n = (int)B.nf[m];
        R[m].D = B.depth[m];
        R[m].lin = B.depth[m];
        if (m == 0 && B.src_nodes) { // a resident source: already refitted
            R[0].facebox = B.src_rawbox;
            R[0].sface = B.src_sface;
            R[0].nodes = const_cast<float*>(B.src_nodes);
            R[0].D = B.src_depth;
            R[0].lin = -1;
        }
        T[m].sface = R[m].sface;
        T[m].nodes = R[m].nodes;
        T[m].n = R[m].n;
        T[m].D = R[m].D;
    }
}

// Both trees and the raw-space traversal, on stream st, once the leaf keys
// (mcode, mface) and single-precision leaf boxes (rawbox) are written: one
// stable sort orders the leaves of both trees (the mesh is the top key bit),
// the implicit trees are refitted, and the traversal appends raw pairs.
// Returns whether the sorted leaves ended in mface2.  b_sort / b_refit: the
// stage events (null: none recorded).
bool enqueue_trees_traverse(const BroadView& B, hipStream_t st, Counters* C, bool ghist_zeroed, hipEvent_t b_sort,
                            hipEvent_t b_refit, hipEvent_t k_start = nullptr, hipEvent_t k_stop = nullptr)
{
    const uint32_t Fs = B.nf[0];
    const uint64_t F = (uint64_t)Fs + B.nf[1];
    // a resident source keeps its own sorted leaves and tree: only the cut
    // leaves are sorted and only the cut tree is refitted
    const bool res = B.src_nodes != nullptr;
    const uint64_t o0 = res ? Fs : 0;
    const bool alt = radix_sort<uint32_t>(B.mcode + o0, B.mface + o0, B.mcode2 + o0, B.mface2 + o0, nullptr, F - o0,
                                          F - o0, B.sort_bits, B.rt, st, ghist_zeroed);
    if (b_sort) (void)hipEventRecord(b_sort, st);
    Tree T[2];
    RefitTree R[2];
    trees_of(B, alt, T, R);
    // each round lifts both trees by up to kRefitLevels levels; a tree of
    // 2^19 leaves needs two rounds
    for (;;) {
        int blocks = 0;
        for (int m = 0; m < 2; ++m) {
            R[m].blocks = 0;
            if (R[m].lin > 0 || (R[m].lin == 0 && R[m].D == 0)) {
                const int width = 1 << R[m].lin;
                R[m].blocks = width / (width < kRefitWidth ? width : kRefitWidth);
            }
            blocks += R[m].blocks;
        }
        if (blocks == 0) break;
        k_refit_levels<<<blocks, 512, 0, st>>>(R[0], R[1]);
        for (int m = 0; m < 2; ++m) {
            if (R[m].blocks == 0) continue;
            if (R[m].lin == 0) R[m].lin = -1; // a single-leaf tree: its leaf is its root
            else R[m].lin -= (R[m].lin < kRefitLevels ? R[m].lin : kRefitLevels);
            if (R[m].lin == 0) R[m].lin = -1;
        }
    }
    if (b_refit) (void)hipEventRecord(b_refit, st);
    launch_traversal(B.trav, B.cus, T[0], T[1], B.raw_pairs, B.cap_pairs, C, st, k_start, k_stop);
    return alt;
}

// compute units of a broad-phase stream that takes pct percent of the CUs
// outside the sums' reservation (the rounding of create_split_streams)
int cus_of_share(int ncu, int pct)
{
    if (ncu <= 4 * kSumCUs) return ncu;
    return (int)(pct / 100.0 * (ncu - kSumCUs) + 0.5);
}

int cus_of(mcgpu_context ctx, bool batch)
{
    if (!ctx->own_streams) return ctx->ncu;
    return cus_of_share(ctx->ncu, batch ? ctx->batch_share_pct : (int)(kBroadShare * 100.0 + 0.5));
}

// the broad view of a context's last (or current) dispatch
BroadView broad_view(mcgpu_context ctx)
{
    BroadView B{};
    B.mcode = P_<uint32_t>(ctx->mcode);
    B.mface = P_<uint32_t>(ctx->mface);
    B.mcode2 = P_<uint32_t>(ctx->mcode2);
    B.mface2 = P_<uint32_t>(ctx->mface2);
    B.rawbox = P_<float>(ctx->rawbox);
    for (int m = 0; m < 2; ++m) {
        B.nodes[m] = P_<float>(ctx->nodes[m]);
        B.nf[m] = ctx->nf[m];
        B.depth[m] = ctx->depth[m];
    }
    B.rt = RadixTemp{P_<uint32_t>(ctx->hist), P_<uint32_t>(ctx->scan_tmp)};
    B.raw_pairs = P_<unsigned long long>(ctx->raw_pairs);
    B.cap_pairs = ctx->cap_pairs;
    B.sort_bits = ctx->sort_bits;
    if (ctx->res) {
        B.src_rawbox = ctx->res->rawbox;
        B.src_sface = ctx->res->sface;
        B.src_nodes = ctx->res->nodes;
        B.src_depth = ctx->res->depth;
    }
    B.trav = ctx->trav;
    B.cus = cus_of(ctx, ctx->batch);
    return B;
}

void enqueue_bvh_traverse(mcgpu_context ctx, hipStream_t st, Counters* C, bool ghist_zeroed)
{
    const BroadView B = broad_view(ctx);
    const bool t = ctx->stage_timing;
    const bool kt = t || ctx->kernel_timing;
    ctx->mface_alt = enqueue_trees_traverse(B, st, C, ghist_zeroed, t ? ctx->ev[EV_B_SORT] : nullptr,
                                            t ? ctx->ev[EV_B_REFIT] : nullptr, kt ? ctx->ev[EV_TRAV_K0] : nullptr,
                                            kt ? ctx->ev[EV_TRAV_K1] : nullptr);
    ev_record(ctx, EV_B_TRAV, st);
}

// Polygon-soup edge numbering on stream s3, and the candidate-face flags the
// pair filter fills: records EV_C_FILL once the flags are clear and
// EV_C_TOPO at the end.
void enqueue_topology(mcgpu_context ctx, hipStream_t s3)
{
    const uint64_t F = (uint64_t)ctx->nf[0] + ctx->nf[1];
    Counters* C = &P_<DevState>(ctx->state)->C;
    stage_event(ctx, EV_C0, s3);
    uint32_t* etotal = P_<uint32_t>(ctx->etotal);
    const uint64_t K = ctx->ncorner[0] + ctx->ncorner[1];
    fill_u32(ctx->cand.p, 0u, F, s3);
    ev_record(ctx, EV_C_FILL, s3);
    const uint32_t res_ne = ctx->res ? ctx->res->ne : 0u;
    if (ctx->res) { // the resident source's edges occupy [0, ne) of the edge table
        (void)hipMemcpyAsync(ctx->ev0.p, ctx->res->ev0, 4ull * res_ne, hipMemcpyDeviceToDevice, s3);
        (void)hipMemcpyAsync(ctx->ev1.p, ctx->res->ev1, 4ull * res_ne, hipMemcpyDeviceToDevice, s3);
        (void)hipMemcpyAsync(ctx->ef0.p, ctx->res->ef0, 4ull * res_ne, hipMemcpyDeviceToDevice, s3);
        (void)hipMemcpyAsync(ctx->ef1.p, ctx->res->ef1, 4ull * res_ne, hipMemcpyDeviceToDevice, s3);
    }
    fill_u32(P_<uint32_t>(ctx->ef1) + res_ne, 0xffffffffu, K - res_ne, s3);
    CornerMap cmap[2];
    for (int m = 0; m < 2; ++m) {
        cmap[m].off = P_<uint32_t>(ctx->off[m]);
        cmap[m].cface = ctx->dsz[m] ? P_<uint32_t>(ctx->cface[m]) : nullptr;
        if (ctx->dsz[m] && !(ctx->res && m == 0))
            k_corner_faces<<<grid_for(ctx->nf[m]), kBlock, 0, s3>>>(cmap[m].off, ctx->nf[m], P_<uint32_t>(ctx->cface[m]));
    }
    // the scans of this stream use their own scratch: stream B's sort uses scan_tmp
    uint32_t* ctmp = P_<uint32_t>(ctx->scan_tmp_c);
    if (topology_packed(ctx)) {
        TopoMesh TM[2];
        for (int m = 0; m < 2; ++m)
            TM[m] = TopoMesh{ctx->didx[m], cmap[m], P_<unsigned long long>(ctx->hkeys[m]),
                             next_pow2(2 * ctx->ncorner[m] + 2) - 1, P_<uint32_t>(ctx->cedge[m]), ctx->ncorner[m],
                             m ? ctx->nv[0] : 0u, m ? ctx->nf[0] : 0u};
        const int cut_only = ctx->res ? 1 : 0;
        const uint64_t K0 = ctx->ncorner[0], start = cut_only ? K0 : 0;
        uint32_t* flag = P_<uint32_t>(ctx->eflag[0]);
        uint32_t* eid = P_<uint32_t>(ctx->eid[0]);
        const uint64_t t0 = cut_only ? 0 : TM[0].mask + 1, t1 = TM[1].mask + 1;
        k_fill2_u64<<<grid_for(t0 + t1), kBlock, 0, s3>>>(TM[0].keys, t0, TM[1].keys, t1, kEmpty64);
        k_edge_insert_packed<<<grid_for(K - start), kBlock, 0, s3>>>(TM[0], TM[1], start);
        k_edge_flag_packed<<<grid_for(K - start), kBlock, 0, s3>>>(TM[0], TM[1], start, flag, C);
        scan_exclusive(flag + start, eid + start, K - start, ctmp, etotal, s3);
        k_store_edge_counts_packed<<<1, 1, 0, s3>>>(eid, flag, K0, etotal, cut_only, C);
        k_edge_assign_packed<<<grid_for(K - start), kBlock, 0, s3>>>(TM[0], TM[1], start, eid, cut_only, C,
                                                                     P_<uint32_t>(ctx->ev0), P_<uint32_t>(ctx->ev1),
                                                                     P_<int32_t>(ctx->ef0), P_<int32_t>(ctx->ef1));
    } else if (ctx->batch && ctx->batch_packed) {
        BatchTopo T[2];
        for (int m = 0; m < 2; ++m) {
            T[m] = BatchTopo{ctx->didx[m], cmap[m], P_<unsigned long long>(ctx->hkeys[m]), P_<uint32_t>(ctx->cedge[m]),
                             ctx->ncorner[m], P_<uint32_t>(ctx->fpair), P_<BatchPair>(ctx->bpairs),
                             P_<uint32_t>(ctx->off[m]), m ? ctx->nf[0] : 0u, m};
            fill_u64(T[m].keys, kEmpty64, 2 * ctx->ncorner[m], s3);
            k_edge_insert_batch<<<grid_for(ctx->ncorner[m]), kBlock, 0, s3>>>(T[m]);
            k_edge_flag_batch<<<grid_for(ctx->ncorner[m]), kBlock, 0, s3>>>(T[m], P_<uint32_t>(ctx->eflag[m]), C);
            scan_exclusive(P_<uint32_t>(ctx->eflag[m]), P_<uint32_t>(ctx->eid[m]), ctx->ncorner[m], ctmp, etotal + m, s3);
            k_store_edge_count<<<1, 1, 0, s3>>>(etotal + m, C, m);
        }
        for (int m = 0; m < 2; ++m)
            k_edge_assign_batch<<<grid_for(ctx->ncorner[m]), kBlock, 0, s3>>>(
                T[m], P_<uint32_t>(ctx->eid[m]), C, m ? ctx->nv[0] : 0u, P_<uint32_t>(ctx->ev0), P_<uint32_t>(ctx->ev1),
                P_<int32_t>(ctx->ef0), P_<int32_t>(ctx->ef1));
    } else {
    for (int m = ctx->res ? 1 : 0; m < 2; ++m) {
        const uint64_t nk = ctx->ncorner[m];
        const uint64_t hs = next_pow2(2 * nk + 2);
        fill_u64(ctx->hkeys[m].p, kEmpty64, hs, s3);
        fill_u64(ctx->hfirst[m].p, kEmpty64, hs, s3);
        k_edge_insert<<<grid_for(nk), kBlock, 0, s3>>>(ctx->didx[m], cmap[m], nk, m, P_<unsigned long long>(ctx->hkeys[m]),
                                                       P_<unsigned long long>(ctx->hfirst[m]), hs - 1);
        k_edge_flag<<<grid_for(nk), kBlock, 0, s3>>>(ctx->didx[m], cmap[m], nk, m, P_<unsigned long long>(ctx->hkeys[m]),
                                                     P_<unsigned long long>(ctx->hfirst[m]), hs - 1,
                                                     P_<uint32_t>(ctx->eflag[m]), C);
        scan_exclusive(P_<uint32_t>(ctx->eflag[m]), P_<uint32_t>(ctx->eid[m]), nk, ctmp, etotal + m, s3);
        k_store_edge_count<<<1, 1, 0, s3>>>(etotal + m, C, m);
    }
    for (int m = ctx->res ? 1 : 0; m < 2; ++m) {
        const uint64_t nk = ctx->ncorner[m];
        const uint64_t hs = next_pow2(2 * nk + 2);
        k_edge_assign<<<grid_for(nk), kBlock, 0, s3>>>(
            ctx->didx[m], cmap[m], nk, m, P_<unsigned long long>(ctx->hkeys[m]), P_<unsigned long long>(ctx->hfirst[m]),
            hs - 1, P_<uint32_t>(ctx->eid[m]), C, m ? ctx->nv[0] : 0, m ? ctx->nf[0] : 0, P_<uint32_t>(ctx->cedge[m]),
            P_<uint32_t>(ctx->ev0), P_<uint32_t>(ctx->ev1), P_<int32_t>(ctx->ef0), P_<int32_t>(ctx->ef1));
    }
    }
    if (ctx->batch)
        k_batch_edge_bases<<<grid_for(ctx->npairs), kBlock, 0, s3>>>(
            P_<BatchPair>(ctx->bpairs), ctx->npairs, P_<uint32_t>(ctx->off[0]), P_<uint32_t>(ctx->off[1]), ctx->nf[0],
            ctx->nf[1], P_<uint32_t>(ctx->eid[0]), P_<uint32_t>(ctx->eid[1]), C);
    ev_record(ctx, EV_C_TOPO, s3);

}

// Rounds of the component check (hook, jump, round end) and the root count,
// on stream st; a resident source was checked once, so only the cut mesh.
constexpr int kCcRounds = 6;
void enqueue_cc_rounds(mcgpu_context ctx, hipStream_t st)
{
    Counters* C = &P_<DevState>(ctx->state)->C;
    const uint64_t V = (uint64_t)ctx->nv[0] + ctx->nv[1], v0 = ctx->res ? ctx->nv[0] : 0;
    const uint64_t K = ctx->ncorner[0] + ctx->ncorner[1];
    uint32_t* parent = P_<uint32_t>(ctx->cc_parent);
    for (int r = 0; r < kCcRounds; ++r) {
        k_cc_hook<<<grid_for(K), kBlock, 0, st>>>(P_<uint32_t>(ctx->ev0), P_<uint32_t>(ctx->ev1), C, ctx->res ? 1 : 0,
                                                  parent);
        k_cc_jump<<<grid_for(V - v0), kBlock, 0, st>>>(parent, v0, V, C);
        k_cc_round_end<<<1, 1, 0, st>>>(C);
    }
    k_cc_count<<<grid_for(V - v0), kBlock, 0, st>>>(parent, v0, ctx->nv[0], V, C);
}

void enqueue_frontend(mcgpu_context ctx)
{
    hipStream_t s = ctx->s, s2 = ctx->batch ? ctx->s2b : ctx->s2, s3 = ctx->batch ? ctx->s3b : ctx->s3;
    const uint32_t Fs = ctx->nf[0];
    const uint64_t F = (uint64_t)Fs + ctx->nf[1], V = (uint64_t)ctx->nv[0] + ctx->nv[1];
    Counters* C = &P_<DevState>(ctx->state)->C;
    Params* Pd = &P_<DevState>(ctx->state)->P;

    // ---- stream A, enqueued first: the serial sums start at once ----
    ev_record(ctx, EV_FORK, s);
    // a single dispatch on its own four streams clears the whole pair hash
    // early; batches (huge tables) and serving slots (one stream) size it
    // after the traversal
    const bool early_clear = ctx->own_streams && !ctx->batch;
    const bool absolute = (ctx->flags & MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION_ABSOLUTE) != 0;
    Params* Pb = P_<Params>(ctx->bparams);
    const auto* brbox = (const unsigned long long(*)[2][6])ctx->brbox.p;
    if (ctx->batch) {
        ev_wait(ctx, ctx->ssum, EV_FORK);
        stage_event(ctx, EV_S0, ctx->ssum);
        const uint32_t nch = 6 * ctx->npairs;
        k_batch_sums<<<std::min<uint32_t>(nch, kBatchSumWaves), 64, 0, ctx->ssum>>>(
            ctx->dv[0], ctx->dv[1], P_<BatchPair>(ctx->bpairs), P_<uint32_t>(ctx->border), nch, C, Pb);
    } else {
        // run items on stream A (every CU), the chain and its check on the
        // reserved CUs, the serial sum only for a coordinate that failed
        if (early_clear) { // the pair hash, before the sum kernels (its fill would slow the sum chain beside it)
            const uint64_t table = next_pow2(2 * ctx->cap_pairs);
            k_pair_hash_clear<<<grid_for(table, kBlock, 2048), kBlock, 0, s>>>(P_<unsigned long long>(ctx->phash), table);
        }
        stage_event(ctx, EV_S0, s);
        const SumBufs SB = sum_bufs(ctx);
        const dim3 grid(SB.nblk, 6);
        k_sum_blocks<<<grid, kSumThreads, 0, s>>>(SB);
        k_sum_items<<<grid, kSumThreads, 0, s>>>(SB);
        ev_record(ctx, EV_SUM_ITEMS, s);
        ev_wait(ctx, ctx->ssum, EV_SUM_ITEMS);

        k_sum_chain<<<6, 64, 4 * SB.nblk, ctx->ssum>>>(SB, Pd);
        ev_record(ctx, EV_SUM_CHAIN, ctx->ssum);
        ev_wait(ctx, s, EV_SUM_CHAIN);
        k_sum_verify<<<dim3(4 * SB.nblk * (kSumTile / 64) / 256, 6), 256, 0, s>>>(SB);
        k_mesh_sums<<<6, 64, 0, s>>>(ctx->res ? nullptr : ctx->dv[0], ctx->nv[0], ctx->dv[1], ctx->nv[1], &C->zero, Pd,
                                     SB.fail);
        ev_record(ctx, EV_A_SUMS, s);
    }
    if (ctx->batch) {
        ev_record(ctx, EV_A_SUMS, ctx->ssum);
        ev_wait(ctx, s, EV_A_SUMS);
    }
    if (ctx->batch) {
        k_batch_shift_params<<<grid_for(ctx->npairs), kBlock, 0, s>>>(Pb, brbox, P_<BatchPair>(ctx->bpairs), ctx->npairs,
                                                                      ctx->gp_constant, absolute ? 1 : 0);
        stage_event(ctx, EV_A_PARAMS, s);
        k_batch_shift_vertices<<<grid_for(3 * V), kBlock, 0, s>>>(ctx->dv[0], ctx->nv[0], ctx->dv[1], ctx->nv[1],
                                                                  P_<uint32_t>(ctx->vpair), Pb, P_<double>(ctx->ps_v));
    } else {
        k_shift_params<<<1, 1, 0, s>>>(Pd, C, ctx->nv[0], ctx->nv[1], ctx->gp_constant, absolute ? 1 : 0, 0, 0.0, 0.0, 0.0);
        stage_event(ctx, EV_A_PARAMS, s);
        k_shift_vertices<<<grid_for(3 * V), kBlock, 0, s>>>(ctx->dv[0], ctx->nv[0], ctx->dv[1], ctx->nv[1], Pd,
                                                            P_<double>(ctx->ps_v), 0);
    }
    const Soup S = make_soup(ctx);
    k_face_boxes<<<grid_for(F), kBlock, 0, s>>>(S, ctx->batch ? Pb : Pd, C, P_<double>(ctx->box), 0);
    stage_event(ctx, EV_A_SHIFT, s);
    if (ctx->shifted_bp) ev_record(ctx, EV_A_BOXES, s);

    // ---- stream B: raw-space trees and broad phase (in shifted coordinates,
    // after the face boxes, for a pair whose raw-space tolerance is loose) ----
    ev_wait(ctx, s2, ctx->shifted_bp ? EV_A_BOXES : EV_FORK);
    stage_event(ctx, EV_B0, s2);
    float* rawbox = P_<float>(ctx->rawbox);
    if (ctx->shifted_bp) {
        uint64_t nzero = 0;
        uint32_t* zero = radix_ghist(RadixTemp{P_<uint32_t>(ctx->hist), P_<uint32_t>(ctx->scan_tmp)}, F, ctx->sort_bits,
                                     &nzero);
        k_morton_shifted<<<grid_for(F), kBlock, 0, s2>>>(P_<double>(ctx->box), ctx->nf[0], ctx->nf[1], Pd, C,
                                                         P_<uint32_t>(ctx->mcode), P_<uint32_t>(ctx->mface), rawbox,
                                                         zero, nzero);
    } else if (ctx->batch) {
        for (int m = 0; m < 2; ++m) {
            const uint32_t nf = ctx->nf[m];
            const uint64_t o = m ? Fs : 0;
            k_batch_morton_raw<<<grid_for(nf), kBlock, 0, s2>>>(ctx->dv[m], ctx->didx[m], P_<uint32_t>(ctx->off[m]), nf, m,
                                                                (uint32_t)o, P_<uint32_t>(ctx->fpair), brbox, Pb, C,
                                                                ctx->batch_mb, P_<uint32_t>(ctx->mcode) + o,
                                                                P_<uint32_t>(ctx->mface) + o, rawbox + 6 * o);
        }
    } else { // both meshes' leaves in one launch (the cut mesh's only, for a resident source)
        MortonMesh M[2];
        for (int m = 0; m < 2; ++m)
            M[m] = MortonMesh{ctx->dv[m], ctx->didx[m], P_<uint32_t>(ctx->off[m]), C->rbox[m], ctx->nf[m],
                              m ? (1u << 21) : 0u, m};
        const uint32_t start = ctx->res ? Fs : 0u;
        uint64_t nzero = 0;
        uint32_t* zero = radix_ghist(RadixTemp{P_<uint32_t>(ctx->hist), P_<uint32_t>(ctx->scan_tmp)}, F - start,
                                     ctx->sort_bits, &nzero);
        k_morton_raw<<<grid_for(F - start), kBlock, 0, s2>>>(M[0], M[1], start, P_<uint32_t>(ctx->mcode),
                                                             P_<uint32_t>(ctx->mface), rawbox, Pd, C, zero, nzero);
    }
    enqueue_bvh_traverse(ctx, s2, C, !ctx->batch); // k_morton_raw cleared the sort's group totals

    // ---- stream C: polygon-soup edge numbering, and the tables the filter fills ----
    ev_wait(ctx, s3, EV_FORK);
    enqueue_topology(ctx, s3);

    // ---- stream B, once its traversal and stream C's edge table are done:
    // one connected component per input mesh (per pair in a batch), in the
    // time B would otherwise idle while A filters and tests ----
    if (ctx->cc_check) {
        ev_wait(ctx, s2, EV_C_TOPO);
        uint32_t* parent = P_<uint32_t>(ctx->cc_parent);
        k_cc_init<<<grid_for((uint64_t)ctx->nv[0] + ctx->nv[1]), kBlock, 0, s2>>>(parent, (uint64_t)ctx->nv[0] + ctx->nv[1]);
        enqueue_cc_rounds(ctx, s2);
        ev_record(ctx, EV_B_CC, s2);
    }

    // ---- stream A: join the broad phase, exact filter, narrow phase ----
    ev_wait(ctx, s, EV_B_TRAV);
    ev_wait(ctx, s, EV_C_FILL);
    stage_event(ctx, EV_A_JOIN, s);
    const uint64_t table = next_pow2(2 * ctx->cap_pairs);
    if (!early_clear)
        k_pair_hash_init<<<grid_for(table, kBlock, 1024), kBlock, 0, s>>>(P_<unsigned long long>(ctx->phash),
                                                                          ctx->cap_pairs, table, C);
    k_filter_pairs<<<grid_for(ctx->cap_pairs, kFilterTile, 2048), kStageBlock, 0, s>>>(P_<double>(ctx->box), P_<unsigned long long>(ctx->raw_pairs),
                                                               ctx->cap_pairs, Fs,
                                                               ctx->batch ? P_<uint32_t>(ctx->fpair) : nullptr,
                                                               P_<unsigned long long>(ctx->pairs),
                                                               P_<unsigned long long>(ctx->phash), P_<uint32_t>(ctx->cand),
                                                               P_<uint32_t>(ctx->cand_list), early_clear ? table : 0,
                                                               C);
    stage_event(ctx, EV_A_FILTER, s);
    enqueue_narrow(ctx, S, true);
}

// The state's copy to pinned host memory, on stream A behind the dispatch
// (stream A joins the other three before its last kernel).
mcgpu_result enqueue_state_copy(mcgpu_context ctx)
{
    ctx->state_copied = false;
    if (!ctx->hstate) CK(hipHostMalloc((void**)&ctx->hstate, sizeof(DevState), hipHostMallocDefault));
    CK(hipMemcpyAsync(ctx->hstate, ctx->state.p, sizeof(DevState), hipMemcpyDeviceToHost, ctx->s));
    ctx->state_copied = true;
    return MCGPU_SUCCESS;
}

mcgpu_result finish(mcgpu_context ctx, mcgpu_frontend_info* info, bool* overflow)
{
    ctx->cp_have = false; // cut-path edges belong to the dispatch they were made from
    // stream A joins the other three before its last kernel, so it finishes last
    CK(hipStreamSynchronize(ctx->s));
    CK(hipGetLastError());
    DevState hs;
    if (ctx->state_copied) { // enqueued with the dispatch: complete with the stream
        std::memcpy(&hs, ctx->hstate, sizeof(DevState));
        ctx->state_copied = false;
    } else {
        CK(hipMemcpyAsync(&hs, ctx->state.p, sizeof(DevState), hipMemcpyDeviceToHost, ctx->s));
        CK(hipStreamSynchronize(ctx->s));
    }
    ctx->C = hs.C;
    ctx->P = hs.P;
    const Counters& C = ctx->C;
    *overflow = false;
    if (getenv("MCGPU_DEBUG"))
        fprintf(stderr,
                "[mcgpu] raw=%llu pairs=%llu tests=%llu esc=%llu reg=%llu big=%llu cand=%llu edges=%llu+%llu ov=%u/%u/%u/%u/%u "
                "gp=%u bad=%u/%u unsup=%u nonman=%u badin=%u pairfail=%u partial=%u caps=%llu/%llu/%llu/%llu "
                "trav splits=%llu lane_walks=%llu\n",
                C.n_raw, C.n_pairs, C.n_tests, C.n_esc, C.n_reg, C.n_big, C.n_cand, C.n_edges_src, C.n_edges_cut, C.ov_pairs,
                C.ov_tests, C.ov_esc, C.ov_reg, C.ov_big, C.gp, C.bad_src, C.bad_cut, C.unsupported, C.nonmanifold,
                C.bad_input, C.pair_fail, C.partial, (unsigned long long)ctx->cap_pairs,
                (unsigned long long)ctx->cap_tests, (unsigned long long)ctx->cap_esc, (unsigned long long)ctx->cap_reg,
                C.n_fallback, C.n_lane_walks);
    if (C.bad_input || C.nonmanifold) return MCGPU_ERROR_INVALID_VALUE;
    if (ctx->cc_check) { // check_input_mesh: one component per mesh (per pair of a batch)
        // rounds until no edge crosses two components (each round removes a root)
        for (uint64_t more = 0; !ctx->C.cc_done; ++more) {
            if (more > ctx->nv[0] + (uint64_t)ctx->nv[1]) return MCGPU_ERROR_DEVICE;
            enqueue_cc_rounds(ctx, ctx->s);
            CK(hipStreamSynchronize(ctx->s));
            CK(hipMemcpyAsync(&ctx->C, &P_<DevState>(ctx->state)->C, sizeof(Counters), hipMemcpyDeviceToHost, ctx->s));
            CK(hipStreamSynchronize(ctx->s));
        }
        const uint32_t want = ctx->batch ? ctx->npairs : 1u;
        if ((!ctx->res && C.ncc[0] != want) || C.ncc[1] != want) return MCGPU_ERROR_INVALID_VALUE;
    }
    if (C.unsupported & 2u) return MCGPU_ERROR_UNSUPPORTED;
    if (C.loose && !ctx->shifted_bp) { // run again with the broad phase in shifted coordinates
        ctx->shifted_bp = true;
        *overflow = true;
        return MCGPU_SUCCESS;
    }
    if (C.ov_pairs || C.ov_tests || C.ov_esc || C.ov_reg || C.ov_big || C.ov_trip) {
        *overflow = true;
        // a refusal leaves the context's capacities as they were
        const uint64_t was[6] = {ctx->cap_pairs, ctx->cap_tests, ctx->cap_esc, ctx->cap_reg, ctx->cap_big,
                                 ctx->cap_trip};
        auto refuse = [&]() {
            ctx->cap_pairs = was[0], ctx->cap_tests = was[1], ctx->cap_esc = was[2];
            ctx->cap_reg = was[3], ctx->cap_big = was[4], ctx->cap_trip = was[5];
            return MCGPU_ERROR_OUT_OF_MEMORY;
        };
        if (C.ov_pairs) ctx->cap_pairs = std::max<uint64_t>(2 * ctx->cap_pairs, C.n_raw + C.n_raw / 4);
        if (C.ov_tests) ctx->cap_tests = std::max<uint64_t>(2 * ctx->cap_tests, C.n_tests + C.n_tests / 4);
        if (C.ov_esc) ctx->cap_esc = std::max<uint64_t>(2 * ctx->cap_esc, C.n_esc + C.n_esc / 4);
        if (C.ov_reg) ctx->cap_reg = std::max<uint64_t>(2 * ctx->cap_reg, C.n_reg + C.n_reg / 4);
        if (C.ov_big) ctx->cap_big = std::max<uint64_t>(2 * ctx->cap_big, C.n_big + C.n_big / 4);
        if (C.ov_trip) ctx->cap_trip = std::max<uint64_t>(2 * ctx->cap_trip, C.n_trip + C.n_trip / 4);
        // the grown lists must fit the device: a raw pair takes up to 48 bytes
        // (raw list, pair list, pair hash of twice the capacity rounded up to a
        // power of two), a test 10, an escalated test 4, a registry entry 100.
        // Raw boxes rounded to float32 can overlap wholesale when coordinates
        // sit far from the origin relative to the mesh (2^52 + a unit sphere:
        // every face pair), and a quadratic pair count is refused here rather
        // than allocated.
        // The kernels keep list positions (a test, an escalated test, a
        // registry entry) in 32 bits: no list may reach 2^32 entries.
        if (ctx->cap_pairs > kMaxListEntries || ctx->cap_tests > kMaxListEntries || ctx->cap_esc > kMaxListEntries ||
            ctx->cap_reg > kMaxListEntries || ctx->cap_big > kMaxListEntries)
            return refuse();
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return MCGPU_ERROR_DEVICE;
        const long double need = 48.0L * ctx->cap_pairs + 10.0L * ctx->cap_tests + 4.0L * ctx->cap_esc +
                                 100.0L * ctx->cap_reg + 16.0L * ctx->cap_trip;
        if (need > 0.9L * (long double)(free_b + ctx->raw_pairs.cap + ctx->pairs.cap + ctx->phash.cap))
            return refuse();
        return MCGPU_SUCCESS;
    }
    if (C.unsupported) return MCGPU_ERROR_UNSUPPORTED;
    auto el = [&](int a, int b) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) ms = 0.f;
        return ms;
    };
    ctx->stages_ready = false; // stage times are read from the events on request
    float tot = el(EV_A0, EV_A_END);
    ctx->total_ms = tot;
    ctx->have = true;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = decide_status(C, ctx->flags, ctx->enforcement_count);
        info->gp_violation = C.gp;
        info->num_candidate_pairs = C.n_pairs;
        info->num_candidate_faces = C.n_cand;
        info->num_edges = C.n_edges_src + C.n_edges_cut;
        info->num_tests = C.n_tests;
        info->num_escalated = C.n_esc;
        info->num_registry = C.n_reg;
        info->num_raw_pairs = C.n_raw;
        info->num_decided_f32 = C.n_f32;
        info->num_decided_f64 = C.n_f64;
        for (int j = 0; j < 3; ++j) {
            info->com[j] = ctx->P.com[j];
            info->translation[j] = ctx->P.t[j];
        }
        info->eps = ctx->P.eps;
        info->device_ms = tot;
        info->partial_cut = (info->status == MCGPU_STATUS_SUCCESS && C.n_reg > 0 && C.partial) ? 1u : 0u;
    }
    return MCGPU_SUCCESS;
}

// per-stage times of the last dispatch from its events (the events stay
// valid until the next dispatch records them again)
void compute_stage_times(mcgpu_context ctx)
{
    auto el = [&](int a, int b) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) ms = 0.f;
        return ms;
    };
    if (ctx->last_kernel_only) { // only the traversal kernel's timestamps were recorded
        for (int i = 0; i < ST_COUNT; ++i) ctx->stage_ms[i] = 0.f;
        ctx->stage_ms[ST_TRAV_KERNEL] = ctx->stage_mode == 0 ? el(EV_TRAV_K0, EV_TRAV_K1) : 0.f;
        ctx->stages_ready = true;
        return;
    }
    // stages of streams B and C run only in a full dispatch; A's stages after a
    // join start at the event recorded once the join completed
    const bool full = ctx->stage_mode == 0;
    ctx->stage_ms[ST_SUMS] = full ? el(EV_S0, EV_A_SUMS) : 0.f;
    ctx->stage_ms[ST_PARAMS] = full ? el(EV_A_SUMS, EV_A_PARAMS) : el(EV_A0, EV_A_PARAMS);
    ctx->stage_ms[ST_SHIFT] = el(EV_A_PARAMS, EV_A_SHIFT);
    ctx->stage_ms[ST_TOPOLOGY] = full ? el(EV_C0, EV_C_TOPO) : 0.f;
    ctx->stage_ms[ST_BVH_SORT] = full ? el(EV_B0, EV_B_SORT) : 0.f;
    ctx->stage_ms[ST_REFIT] = full ? el(EV_B_SORT, EV_B_REFIT) : 0.f;
    ctx->stage_ms[ST_TRAVERSE] = full ? el(EV_B_REFIT, EV_B_TRAV) : 0.f;
    ctx->stage_ms[ST_FILTER] = full ? el(EV_A_JOIN, EV_A_FILTER) : 0.f;
    ctx->stage_ms[ST_PLANES] = el(EV_A_FILTER, EV_A_PLANES);
    ctx->stage_ms[ST_TESTS] = el(EV_A_JOIN2, EV_A_TESTS);
    ctx->stage_ms[ST_NARROW] = el(EV_A_TESTS, EV_A_NARROW);
    ctx->stage_ms[ST_REGISTRY] = el(EV_A_NARROW, EV_A_END);
    ctx->stage_ms[ST_TRAV_KERNEL] = full ? el(EV_TRAV_K0, EV_TRAV_K1) : 0.f;
    ctx->stages_ready = true;
}

// every device buffer a context owns
std::vector<Buf*> context_bufs(mcgpu_context ctx)
{
    return {&ctx->in_v[0], &ctx->in_v[1], &ctx->in_idx[0], &ctx->in_idx[1], &ctx->in_sz[0], &ctx->in_sz[1],
                  &ctx->off[0], &ctx->off[1], &ctx->state, &ctx->ps_v, &ctx->box, &ctx->hkeys[0],
                  &ctx->hkeys[1], &ctx->hfirst[0], &ctx->hfirst[1], &ctx->cface[0], &ctx->cface[1], &ctx->eflag[0],
                  &ctx->eflag[1], &ctx->eid[0], &ctx->eid[1], &ctx->cedge[0], &ctx->cedge[1], &ctx->etotal, &ctx->ev0,
                  &ctx->ev1, &ctx->ef0, &ctx->ef1, &ctx->mcode, &ctx->mface, &ctx->mcode2, &ctx->mface2, &ctx->nodes[0], &ctx->nodes[1],
                  &ctx->rawbox, &ctx->raw_pairs, &ctx->pairs, &ctx->phash, &ctx->cand, &ctx->cand_list,
                  &ctx->planes, &ctx->big_list, &ctx->big_off, &ctx->big_scratch, &ctx->te, &ctx->tg, &ctx->tqs, &ctx->trs,
                  &ctx->ttype, &ctx->tpip, &ctx->esc, &ctx->rkey, &ctx->rval, &ctx->rkey2, &ctx->rval2, &ctx->rp,
                  &ctx->re, &ctx->rg, &ctx->rpts, &ctx->rcount, &ctx->shkeys, &ctx->shcnt, &ctx->scan_tmp, &ctx->scan_tmp_c, &ctx->hist,
                  &ctx->re_in, &ctx->rg_in, &ctx->bpairs, &ctx->bstarts, &ctx->fpair, &ctx->vpair, &ctx->bparams,
                  &ctx->brbox, &ctx->pflags, &ctx->border, &ctx->bidx[0], &ctx->bidx[1], &ctx->rle, &ctx->rlg,
                  &ctx->rstart, &ctx->rend, &ctx->pqs, &ctx->prs, &ctx->pcounts, &ctx->sbsum,
                  &ctx->sitems, &ctx->slo, &ctx->shi, &ctx->snitem, &ctx->slist, &ctx->sckpt, &ctx->sflags,
                  &ctx->tune_c, &ctx->cc_parent, &ctx->cp_order, &ctx->cp_keys, &ctx->cp_vals, &ctx->cp_keys2, &ctx->cp_vals2,
                  &ctx->cp_flag, &ctx->cp_gid, &ctx->cp_gstart, &ctx->cp_ncand, &ctx->cp_coff, &ctx->cp_trip,
                  &ctx->cp_src, &ctx->cp_tgt, &ctx->cp_fa, &ctx->cp_fb, &ctx->cp_group, &ctx->cp_kind, &ctx->cp_acc,
                  &ctx->cp_table, &ctx->cp_keep, &ctx->cp_pos, &ctx->cp_osrc, &ctx->cp_otgt, &ctx->cp_ofa,
                  &ctx->cp_ofb, &ctx->cp_ogroup, &ctx->cp_okind, &ctx->cp_cnt};
}

// The shape of the dispatch about to be enqueued: everything its kernels'
// launch parameters depend on (sizes, capacities, flags, the traversal shape,
// the streams and the address of every buffer).  A recorded program replays
// only for the same shape; the data behind the addresses may differ.
std::vector<uint64_t> graph_key(mcgpu_context ctx)
{
    std::vector<uint64_t> k;
    auto put = [&k](uint64_t x) { k.push_back(x); };
    auto ptr = [&k](const void* p) { k.push_back((uint64_t)(uintptr_t)p); };
    put(ctx->batch), put(ctx->npairs), put(ctx->batch_mb), put(ctx->batch_packed), put(ctx->cc_check), put(ctx->sort_bits),
        put(ctx->shifted_bp),
        put(ctx->gbits), put(ctx->kbits);
    for (int m = 0; m < 2; ++m) {
        put(ctx->nv[m]), put(ctx->nf[m]), put(ctx->ncorner[m]), put(ctx->depth[m]);
        ptr(ctx->dv[m]), ptr(ctx->didx[m]), ptr(ctx->dsz[m]), ptr(ctx->bin_idx[m]);
    }
    put(ctx->cap_pairs), put(ctx->cap_tests), put(ctx->cap_esc), put(ctx->cap_reg), put(ctx->cap_big),
        put(ctx->cap_trip), put(ctx->sum_nblk);
    uint64_t gp;
    std::memcpy(&gp, &ctx->gp_constant, 8);
    put(ctx->flags), put(gp), put(ctx->enforcement_count), put(ctx->keep_tests);
    put(ctx->trav.ll), put(ctx->trav.stride), put(ctx->trav.queues), put(ctx->trav.chunk);
    ptr(ctx->s), ptr(ctx->s2), ptr(ctx->s3), ptr(ctx->s2b), ptr(ctx->s3b), ptr(ctx->ssum);
    for (Buf* b : context_bufs(ctx)) ptr(b->p);
    return k;
}

void graph_drop(mcgpu_context ctx)
{
    for (const GraphOp& op : ctx->prog)
        if (op.exec) (void)hipGraphExecDestroy(op.exec);
    ctx->prog.clear();
    ctx->prog_key.clear();
}

// Enqueues the validation and front end of a single pair or a batch: eagerly,
// or (graphs on) by replaying the context's program, recorded first when the
// dispatch's shape differs from the one it was recorded for.
mcgpu_result enqueue_dispatch(mcgpu_context ctx)
{
    ctx->state_copied = false;
    if (!ctx->graphs || ctx->stage_timing || ctx->kernel_timing || !ctx->own_streams) {
        if (ctx->batch) enqueue_validate_batch(ctx);
        else enqueue_validate(ctx);
        enqueue_frontend(ctx);
        return enqueue_state_copy(ctx);
    }
    std::vector<uint64_t> key = graph_key(ctx);
    if (key != ctx->prog_key || ctx->prog.empty()) {
        graph_drop(ctx);
        hipStream_t sts[4] = {ctx->s, ctx->batch ? ctx->s2b : ctx->s2, ctx->batch ? ctx->s3b : ctx->s3, ctx->ssum};
        ctx->rec_error = hipSuccess;
        for (hipStream_t st : sts) CK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        ctx->recording = true;
        if (ctx->batch) enqueue_validate_batch(ctx);
        else enqueue_validate(ctx);
        enqueue_frontend(ctx);
        for (hipStream_t st : sts) { // the streams' last segments, then the capture ends
            graph_cut(ctx, st);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g); // empty: every stream ends on an event record
            if (e != hipSuccess && ctx->rec_error == hipSuccess) ctx->rec_error = e;
        }
        ctx->recording = false;
        const hipError_t le = hipGetLastError();
        if (ctx->rec_error != hipSuccess || le != hipSuccess) {
            ctx->last_error = ctx->rec_error != hipSuccess ? ctx->rec_error : le;
            graph_drop(ctx);
            return MCGPU_ERROR_DEVICE;
        }
        ctx->prog_key.swap(key);
        ++ctx->recordings;
    }
    for (const GraphOp& op : ctx->prog) {
        switch (op.kind) {
        case GraphOp::LAUNCH: CK(hipGraphLaunch(op.exec, op.st)); break;
        case GraphOp::RECORD: CK(hipEventRecord(ctx->ev[op.ev], op.st)); break;
        case GraphOp::WAIT: CK(hipStreamWaitEvent(op.st, ctx->ev[op.ev], 0)); break;
        }
    }
    ++ctx->replays;
    return enqueue_state_copy(ctx);
}

mcgpu_result upload_mesh(mcgpu_context ctx, int m, const mcgpu_mesh* mesh, bool on_device)
{
    ctx->nv[m] = mesh->num_vertices;
    ctx->nf[m] = mesh->num_faces;
    if (on_device) {
        ctx->dv[m] = mesh->vertices;
        ctx->didx[m] = mesh->face_indices;
        ctx->dsz[m] = mesh->face_sizes;
        if (mesh->face_sizes) {
            std::vector<uint32_t> h(mesh->num_faces);
            CK(hipMemcpy(h.data(), mesh->face_sizes, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
            uint64_t k = 0;
            for (uint32_t x : h) k += x;
            ctx->ncorner[m] = k;
        } else {
            ctx->ncorner[m] = 3ull * mesh->num_faces;
        }
        return MCGPU_SUCCESS;
    }
    uint64_t k = 3ull * mesh->num_faces;
    if (mesh->face_sizes) {
        k = 0;
        for (uint32_t f = 0; f < mesh->num_faces; ++f) {
            if (mesh->face_sizes[f] < 3) return MCGPU_ERROR_INVALID_VALUE;
            k += mesh->face_sizes[f];
        }
    }
    ctx->ncorner[m] = k;
    double* v;
    uint32_t* ix;
    CK(ensure(ctx->in_v[m], 3ull * mesh->num_vertices, &v));
    CK(ensure(ctx->in_idx[m], k, &ix));
    CK(hipMemcpyAsync(v, mesh->vertices, 24ull * mesh->num_vertices, hipMemcpyHostToDevice, ctx->s));
    CK(hipMemcpyAsync(ix, mesh->face_indices, 4ull * k, hipMemcpyHostToDevice, ctx->s));
    ctx->dv[m] = v;
    ctx->didx[m] = ix;
    ctx->dsz[m] = nullptr;
    if (mesh->face_sizes) {
        uint32_t* sz;
        CK(ensure(ctx->in_sz[m], mesh->num_faces, &sz));
        CK(hipMemcpyAsync(sz, mesh->face_sizes, 4ull * mesh->num_faces, hipMemcpyHostToDevice, ctx->s));
        ctx->dsz[m] = sz;
    }
    return MCGPU_SUCCESS;
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

MCGPU_API const char* mcgpu_result_string(mcgpu_result r)
{
    switch (r) {
    case MCGPU_SUCCESS: return "success";
    case MCGPU_ERROR_INVALID_VALUE: return "invalid value";
    case MCGPU_ERROR_DEVICE: return "HIP runtime error";
    case MCGPU_ERROR_OUT_OF_MEMORY: return "out of device memory";
    case MCGPU_ERROR_UNSUPPORTED: return "unsupported input";
    case MCGPU_ERROR_NO_RESULT: return "no dispatch result";
    }
    return "unknown";
}

// Streams B and C of a dispatch: disjoint CU sets outside the first kSumCUs,
// B taking `share` of them; a device too small to partition gets no masks.
static hipError_t create_split_streams(int ncu, double share, hipStream_t* b, hipStream_t* c)
{
    const int nw = (ncu + 31) / 32;
    std::vector<uint32_t> mb(nw, 0u), mc(nw, 0u);
    const int nb = kSumCUs + (int)(share * (ncu - kSumCUs) + 0.5);
    for (int i = kSumCUs; i < ncu; ++i) (i < nb ? mb : mc)[i / 32] |= 1u << (i % 32);
    if (ncu <= 4 * kSumCUs) {
        std::fill(mb.begin(), mb.end(), ~0u);
        std::fill(mc.begin(), mc.end(), ~0u);
    }
    hipError_t e = hipExtStreamCreateWithCUMask(b, nw, mb.data());
    if (e != hipSuccess) return e;
    if ((e = hipExtStreamCreateWithCUMask(c, nw, mc.data())) != hipSuccess) {
        (void)hipStreamDestroy(*b);
        *b = nullptr;
    }
    return e;
}

// percent of the CUs outside the sums' reservation that the dispatch's
// broad-phase stream holds (100: one stream for everything)
static int split_pct(mcgpu_context ctx)
{
    if (!ctx->own_streams) return 100;
    return ctx->batch ? ctx->batch_share_pct : (int)(kBroadShare * 100.0 + 0.5);
}

static mcgpu_tuning tuning_report(const TravConfig& t, int pct)
{
    mcgpu_tuning r{};
    r.wave_leaves = 1u << t.ll;
    r.tree_stride = (uint32_t)t.stride;
    r.queues = (uint32_t)t.queues;
    r.chunk = (uint32_t)t.chunk;
    r.broad_share_pct = (uint32_t)pct;
    return r;
}

// The traversal shape of the dispatch about to run (sizes set): the
// caller's override, else the autotuned entry of its workload class, else the
// default.  Batches get their stream pair on the context's first batch.
static mcgpu_result apply_tuning(mcgpu_context ctx)
{
    if (ctx->batch && ctx->own_streams && !ctx->s2b)
        CK(create_split_streams(ctx->ncu, ctx->batch_share_pct / 100.0, &ctx->s2b, &ctx->s3b));
    TravConfig t;
    const TuneEntry* hit = nullptr;
    if (ctx->forced) {
        t = ctx->forced_trav;
    } else {
        const int bs = bits_for((uint64_t)ctx->nf[0] + 1), bc = bits_for((uint64_t)ctx->nf[1] + 1);
        for (const TuneEntry& e : ctx->tuned)
            if (e.batch == ctx->batch && e.bs == bs && e.bc == bc) {
                hit = &e;
                t = e.trav;
                break;
            }
    }
    ctx->trav = t;
    ctx->used = hit ? hit->report : tuning_report(t, split_pct(ctx));
    return MCGPU_SUCCESS;
}

// The last dispatch once more with the same inputs and capacities (which it
// sized): its device time in *ms.
static mcgpu_result rerun_dispatch(mcgpu_context ctx, float* ms)
{
    ctx->stage_mode = 0;
    mcgpu_result r = enqueue_dispatch(ctx);
    if (r != MCGPU_SUCCESS) return r;
    bool overflow = false;
    r = finish(ctx, nullptr, &overflow);
    if (r != MCGPU_SUCCESS) return r;
    if (overflow) return MCGPU_ERROR_DEVICE; // the same input cannot outgrow the lists it sized
    if (ctx->batch && (r = finish_batch(ctx, nullptr)) != MCGPU_SUCCESS) return r;
    *ms = ctx->total_ms;
    return MCGPU_SUCCESS;
}

// The traversal of the last dispatch's trees with launch shape t, `reps`
// times on its broad-phase stream, into the tuner's own counters: the median
// kernel time (from the launch's dispatch timestamps) in *ms, and whether
// every run found the dispatch's raw pair count with no overflow.
static mcgpu_result time_traversal(mcgpu_context ctx, const TravConfig& t, int reps, float* ms, bool* same)
{
    BroadView B = broad_view(ctx);
    B.trav = t;
    Tree T[2];
    RefitTree R[2];
    trees_of(B, ctx->mface_alt, T, R);
    Counters* C;
    CK(ensure(ctx->tune_c, 1, &C));
    hipStream_t st = ctx->batch ? ctx->s2b : ctx->s2;
    hipEvent_t e[2];
    CK(hipEventCreate(&e[0]));
    if (hipEventCreate(&e[1]) != hipSuccess) {
        (void)hipEventDestroy(e[0]);
        return MCGPU_ERROR_DEVICE;
    }
    std::vector<float> v;
    mcgpu_result r = MCGPU_SUCCESS;
    *same = true;
    for (int i = 0; i < reps && r == MCGPU_SUCCESS; ++i) {
        fill_u32(C, 0u, sizeof(Counters) / 4, st);
        launch_traversal(t, B.cus, T[0], T[1], B.raw_pairs, B.cap_pairs, C, st, e[0], e[1]);
        Counters h;
        float x = 0.f;
        if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess ||
            hipEventElapsedTime(&x, e[0], e[1]) != hipSuccess ||
            hipMemcpy(&h, C, sizeof(Counters), hipMemcpyDeviceToHost) != hipSuccess) {
            r = MCGPU_ERROR_DEVICE;
            break;
        }
        if (h.n_raw != ctx->C.n_raw || h.ov_pairs) *same = false;
        v.push_back(x);
    }
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    if (r != MCGPU_SUCCESS) return r;
    std::sort(v.begin(), v.end());
    *ms = v[v.size() / 2];
    return MCGPU_SUCCESS;
}

// The autotuner for the last dispatch's workload class: every traversal
// launch shape of the candidate lists (leaves per block, levels per step,
// queue mode) is timed alone on the dispatch's own trees (a median of three)
// and must find the same raw pairs; the fastest wins.  The whole dispatch is
// then re-run with the default and with the chosen shape (best of two device
// times each) for the report.  The choice is kept for the class (single pair
// or batch, bit lengths of the two face counts) and used by its later
// dispatches.  The CU split between the broad-phase and edge-numbering
// streams stays at its defaults: the streams of a context keep their masks
// for its lifetime.
static mcgpu_result autotune(mcgpu_context ctx, mcgpu_tuning* out)
{
    const TravConfig def{};
    float best = 3.4e38f, def_trav = 0.f;
    TravConfig bt = def;
    uint32_t n = 0;
    mcgpu_result r;
    for (int ll : kTuneLL)
        for (int st : kTuneStride)
            for (const auto& q : kTuneQueue) {
                TravConfig t;
                t.ll = ll;
                t.stride = st;
                t.queues = q[0];
                t.chunk = q[1];
                float ms = 0.f;
                bool same = true;
                if ((r = time_traversal(ctx, t, 3, &ms, &same)) != MCGPU_SUCCESS) return r;
                if (!same) return MCGPU_ERROR_DEVICE; // every shape finds the same pairs
                ++n;
                if (t.ll == def.ll && t.stride == def.stride && t.queues == def.queues && t.chunk == def.chunk) def_trav = ms;
                if (ms < best) {
                    best = ms;
                    bt = t;
                }
            }
    auto timed = [&](const TravConfig& t, float* ms) -> mcgpu_result {
        ctx->trav = t;
        float a = 0.f, b = 0.f;
        mcgpu_result e = rerun_dispatch(ctx, &a);
        if (e == MCGPU_SUCCESS) e = rerun_dispatch(ctx, &b);
        *ms = std::min(a, b);
        return e;
    };
    float def_disp = 0.f, disp = 0.f;
    if ((r = timed(def, &def_disp)) != MCGPU_SUCCESS || (r = timed(bt, &disp)) != MCGPU_SUCCESS) return r;
    TuneEntry e;
    e.batch = ctx->batch;
    e.bs = bits_for((uint64_t)ctx->nf[0] + 1);
    e.bc = bits_for((uint64_t)ctx->nf[1] + 1);
    e.trav = bt;
    e.report = tuning_report(bt, split_pct(ctx));
    e.report.candidates = n;
    e.report.traversal_ms = best;
    e.report.default_traversal_ms = def_trav;
    e.report.dispatch_ms = disp;
    e.report.default_dispatch_ms = def_disp;
    bool put = false;
    for (TuneEntry& x : ctx->tuned)
        if (x.batch == e.batch && x.bs == e.bs && x.bc == e.bc) {
            x = e;
            put = true;
        }
    if (!put) ctx->tuned.push_back(e);
    ctx->used = e.report;
    if (out) *out = e.report;
    return MCGPU_SUCCESS;
}

// after a dispatch that succeeded: tune its class once when the context asks
static mcgpu_result autotune_first(mcgpu_context ctx)
{
    if (!ctx->autotune || ctx->forced || !ctx->own_streams || ctx->used.candidates) return MCGPU_SUCCESS;
    return autotune(ctx, nullptr);
}

MCGPU_API mcgpu_result mcgpu_create_context(int device, mcgpu_context* out)
{
    if (!out) return MCGPU_ERROR_INVALID_VALUE;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MCGPU_ERROR_DEVICE;
    mcgpu_context ctx = new mcgpu_context_t();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->s, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MCGPU_ERROR_DEVICE;
    }
    // CU masks: the first kSumCUs compute units for the serial sums, the rest
    // for the streams that run beside them
    const int ncu = prop.multiProcessorCount;
    ctx->ncu = ncu;
    const int nw = (ncu + 31) / 32;
    std::vector<uint32_t> msum(nw, 0u);
    for (int c = 0; c < kSumCUs && c < ncu; ++c) msum[c / 32] |= 1u << (c % 32);
    if (hipExtStreamCreateWithCUMask(&ctx->ssum, nw, msum.data()) != hipSuccess ||
        create_split_streams(ncu, kBroadShare, &ctx->s2, &ctx->s3) != hipSuccess) {
        for (hipStream_t st : {ctx->ssum, ctx->s2, ctx->s3})
            if (st) (void)hipStreamDestroy(st);
        (void)hipStreamDestroy(ctx->s);
        delete ctx;
        return MCGPU_ERROR_DEVICE;
    }
    for (int i = 0; i < EV_COUNT; ++i) (void)hipEventCreate(&ctx->ev[i]);
    *out = ctx;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_release_context(mcgpu_context ctx)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->s);
    if (ctx->own_streams) {
        for (hipStream_t st : {ctx->s2, ctx->s3, ctx->s2b, ctx->s3b, ctx->ssum})
            if (st) (void)hipStreamSynchronize(st);
    }
    for (const GraphOp& op : ctx->prog)
        if (op.exec) (void)hipGraphExecDestroy(op.exec);
    for (Buf* b : context_bufs(ctx))
        if (b->p) (void)hipFree(b->p);
    if (ctx->hstate) (void)hipHostFree(ctx->hstate);
    for (int i = 0; i < EV_COUNT; ++i) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->ev_user) (void)hipEventDestroy(ctx->ev_user);
    if (ctx->own_streams) {
        for (hipStream_t st : {ctx->ssum, ctx->s3, ctx->s2, ctx->s3b, ctx->s2b})
            if (st) (void)hipStreamDestroy(st);
    }
    (void)hipStreamDestroy(ctx->s);
    delete ctx;
    return MCGPU_SUCCESS;
}

// A single pair's dispatch up to its first enqueue: checks, uploads, buffers,
// the traversal shape and the list capacities.
static mcgpu_result frontend_prepare(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                     const mcgpu_dispatch_desc* desc)
{
    if (!ctx || !desc) return MCGPU_ERROR_INVALID_VALUE;
    if (ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch has not been waited for
    mcgpu_result r;
    if ((r = check_mesh(src)) != MCGPU_SUCCESS || (r = check_mesh(cut)) != MCGPU_SUCCESS) return r;
    if (!(desc->gp_constant > 0.0)) return MCGPU_ERROR_INVALID_VALUE; // frontend.cpp:391-397
    CK(hipSetDevice(ctx->device));
    ctx->have = false;
    ctx->broad_only = false;
    ctx->batch = false;
    ctx->shifted_bp = false;
    ctx->allow_loose = true;
    ctx->cc_check = (desc->flags & MCGPU_DISPATCH_CHECK_COMPONENTS) != 0;
    ctx->sort_bits = 22;
    ctx->flags = desc->flags;
    ctx->gp_constant = desc->gp_constant;
    ctx->enforcement_count = desc->enforcement_count;
    ctx->keep_tests = desc->keep_tests ? 1 : 0;
    if ((r = upload_mesh(ctx, 0, src, desc->input_on_device != 0)) != MCGPU_SUCCESS) return r;
    if ((r = upload_mesh(ctx, 1, cut, desc->input_on_device != 0)) != MCGPU_SUCCESS) return r;
    const uint64_t F = (uint64_t)ctx->nf[0] + ctx->nf[1];
    ctx->gbits = bits_for(F + 1);
    ctx->kbits = ctx->gbits + bits_for(ctx->ncorner[0] + ctx->ncorner[1] + 1);
    if (ctx->kbits > 64) return MCGPU_ERROR_UNSUPPORTED;
    if ((r = alloc_mesh_buffers(ctx)) != MCGPU_SUCCESS) return r;
    if ((r = apply_tuning(ctx)) != MCGPU_SUCCESS) return r;
    if (ctx->cap_pairs < 4 * F) ctx->cap_pairs = std::max<uint64_t>(1 << 14, 4 * F);
    if (ctx->cap_tests < 2 * ctx->cap_pairs) ctx->cap_tests = 2 * ctx->cap_pairs;
    if (ctx->cap_esc < (1u << 14)) ctx->cap_esc = std::max<uint64_t>(1 << 14, ctx->cap_tests / 16);
    // the registry starts within the two-launch merge sort's range and grows
    // past it (to the radix path) only when a dispatch overflows it
    if (ctx->cap_reg < (1u << 14))
        ctx->cap_reg = std::min<uint64_t>(std::max<uint64_t>(1 << 14, ctx->cap_tests / 8), kRegMergeTiles * kRegTile);
    if (ctx->cap_big < 1024) ctx->cap_big = 1024;
    if (ctx->cap_trip < (1u << 16)) ctx->cap_trip = 1u << 16;
    return MCGPU_SUCCESS;
}

// Completes an enqueued single-pair dispatch: waits, reads its counters, and
// re-runs it with grown lists while one overflowed.
static mcgpu_result frontend_complete(mcgpu_context ctx, mcgpu_frontend_info* info, int attempt)
{
    mcgpu_result r;
    for (;; ++attempt) {
        bool overflow = false;
        if ((r = finish(ctx, info, &overflow)) != MCGPU_SUCCESS) return r;
        if (!overflow) return autotune_first(ctx);
        if (attempt + 1 >= 6) return MCGPU_ERROR_OUT_OF_MEMORY;
        if ((r = alloc_list_buffers(ctx)) != MCGPU_SUCCESS) return r;
        ctx->stage_mode = 0;
        if ((r = enqueue_dispatch(ctx)) != MCGPU_SUCCESS) return r;
    }
}

MCGPU_API mcgpu_result mcgpu_dispatch_frontend(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                               const mcgpu_dispatch_desc* desc, mcgpu_frontend_info* info)
{
    mcgpu_result r;
    if ((r = frontend_prepare(ctx, src, cut, desc)) != MCGPU_SUCCESS) return r;
    if ((r = alloc_list_buffers(ctx)) != MCGPU_SUCCESS) return r;
    ctx->stage_mode = 0;
    if ((r = enqueue_dispatch(ctx)) != MCGPU_SUCCESS) return r;
    return frontend_complete(ctx, info, 0);
}

MCGPU_API mcgpu_result mcgpu_dispatch_frontend_async(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                                     const mcgpu_dispatch_desc* desc, void* stream)
{
    mcgpu_result r;
    if ((r = frontend_prepare(ctx, src, cut, desc)) != MCGPU_SUCCESS) return r;
    if ((r = alloc_list_buffers(ctx)) != MCGPU_SUCCESS) return r;
    if (stream) { // the dispatch starts after the work queued so far on the caller's stream
        if (!ctx->ev_user) CK(hipEventCreateWithFlags(&ctx->ev_user, hipEventDisableTiming));
        CK(hipEventRecord(ctx->ev_user, (hipStream_t)stream));
        CK(hipStreamWaitEvent(ctx->s, ctx->ev_user, 0));
    }
    ctx->stage_mode = 0;
    if ((r = enqueue_dispatch(ctx)) != MCGPU_SUCCESS) return r;
    ctx->pending = true;
    return MCGPU_SUCCESS;
}

MCGPU_API int mcgpu_dispatch_query(mcgpu_context ctx)
{
    if (!ctx || !ctx->pending) return 1;
    return hipEventQuery(ctx->ev[EV_A_END]) == hipSuccess ? 1 : 0;
}

MCGPU_API mcgpu_result mcgpu_dispatch_wait(mcgpu_context ctx, mcgpu_frontend_info* info, void* stream)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->pending) return MCGPU_ERROR_NO_RESULT;
    CK(hipSetDevice(ctx->device));
    ctx->pending = false;
    const mcgpu_result r = frontend_complete(ctx, info, 0);
    if (r == MCGPU_SUCCESS && stream) // work queued on the caller's stream from now on sees the results
        CK(hipStreamWaitEvent((hipStream_t)stream, ctx->ev[EV_A_END], 0));
    return r;
}

MCGPU_API mcgpu_result mcgpu_dispatch_batch(mcgpu_context ctx, const mcgpu_batch* batch, const mcgpu_dispatch_desc* desc,
                                            mcgpu_frontend_info* info)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx || !batch || !desc || batch->num_pairs == 0) return MCGPU_ERROR_INVALID_VALUE;
    const uint32_t np = batch->num_pairs;
    const uint32_t* vo[2] = {batch->src_vertex_offsets, batch->cut_vertex_offsets};
    const uint32_t* fo[2] = {batch->src_face_offsets, batch->cut_face_offsets};
    const mcgpu_mesh* mesh[2] = {&batch->src, &batch->cut};
    mcgpu_result r;
    for (int m = 0; m < 2; ++m) {
        if ((r = check_mesh(mesh[m])) != MCGPU_SUCCESS) return r;
        if (!vo[m] || !fo[m] || vo[m][0] != 0 || fo[m][0] != 0 || vo[m][np] != mesh[m]->num_vertices ||
            fo[m][np] != mesh[m]->num_faces)
            return MCGPU_ERROR_INVALID_VALUE;
        for (uint32_t p = 0; p < np; ++p) // every pair: a mesh of at least 3 vertices and one face
            if (vo[m][p + 1] < vo[m][p] + 3 || fo[m][p + 1] < fo[m][p] + 1) return MCGPU_ERROR_INVALID_VALUE;
    }
    if (!(desc->gp_constant > 0.0)) return MCGPU_ERROR_INVALID_VALUE;
    const int pb = bits_for(np);
    if (pb > 28) return MCGPU_ERROR_UNSUPPORTED;
    CK(hipSetDevice(ctx->device));
    ctx->have = false;
    ctx->broad_only = false;
    ctx->batch = true;
    ctx->allow_loose = false;
    ctx->cc_check = (desc->flags & MCGPU_DISPATCH_CHECK_COMPONENTS) != 0;
    ctx->npairs = np;
    ctx->batch_mb = std::min(7, (31 - pb) / 3);
    ctx->sort_bits = 32;
    ctx->flags = desc->flags;
    ctx->gp_constant = desc->gp_constant;
    ctx->enforcement_count = desc->enforcement_count;
    ctx->keep_tests = 0;
    for (int m = 0; m < 2; ++m) {
        if ((r = upload_mesh(ctx, m, mesh[m], desc->input_on_device != 0)) != MCGPU_SUCCESS) return r;
        ctx->bin_idx[m] = ctx->didx[m];
    }
    // pair tables
    ctx->hpairs.assign(np, BatchPair{});
    ctx->hstarts.resize(4ull * np);
    std::vector<uint32_t> len(6ull * np);
    for (uint32_t p = 0; p < np; ++p) {
        BatchPair& b = ctx->hpairs[p];
        b.fs0 = fo[0][p], b.nfs = fo[0][p + 1] - fo[0][p];
        b.fc0 = fo[1][p], b.nfc = fo[1][p + 1] - fo[1][p];
        b.vs0 = vo[0][p], b.nvs = vo[0][p + 1] - vo[0][p];
        b.vc0 = vo[1][p], b.nvc = vo[1][p + 1] - vo[1][p];
        ctx->hstarts[p] = b.vs0;
        ctx->hstarts[np + p] = b.vc0;
        ctx->hstarts[2ull * np + p] = b.fs0;
        ctx->hstarts[3ull * np + p] = b.fc0;
        for (int j = 0; j < 6; ++j) len[6ull * p + j] = j < 3 ? b.nvs : b.nvc;
    }
    // the packed edge numbering with one table region per pair, when every
    // pair's vertex ids and corner ranks fit its word (corner counts are
    // known on the host for triangle meshes and host face sizes)
    ctx->batch_packed = true;
    for (int m = 0; m < 2 && ctx->batch_packed; ++m) {
        const bool tri = mesh[m]->face_sizes == nullptr;
        if (!tri && desc->input_on_device) {
            ctx->batch_packed = false;
            break;
        }
        for (uint32_t p = 0; p < np; ++p) {
            uint64_t k = 0;
            if (tri) {
                k = 3ull * (fo[m][p + 1] - fo[m][p]);
            } else {
                for (uint32_t f = fo[m][p]; f < fo[m][p + 1]; ++f) k += mesh[m]->face_sizes[f];
            }
            if (vo[m][p + 1] - vo[m][p] >= (1u << kPackedVertexBits) || k >= (1ull << kPackedRankBits)) {
                ctx->batch_packed = false;
                break;
            }
        }
    }
    // serial-sum chains, longest first
    ctx->horder.resize(6ull * np);
    for (uint32_t c = 0; c < 6 * np; ++c) ctx->horder[c] = c;
    std::stable_sort(ctx->horder.begin(), ctx->horder.end(), [&](uint32_t a, uint32_t b) { return len[a] > len[b]; });
    const uint64_t F = (uint64_t)ctx->nf[0] + ctx->nf[1];
    const uint64_t V = (uint64_t)ctx->nv[0] + ctx->nv[1];
    ctx->gbits = bits_for(F + 1);
    ctx->kbits = ctx->gbits + bits_for(ctx->ncorner[0] + ctx->ncorner[1] + 1);
    if (ctx->kbits > 64) return MCGPU_ERROR_UNSUPPORTED;
    if ((r = alloc_mesh_buffers(ctx)) != MCGPU_SUCCESS) return r;
    if ((r = apply_tuning(ctx)) != MCGPU_SUCCESS) return r;
    {
        BatchPair* bp;
        uint32_t* u;
        Params* pp;
        unsigned long long* ull;
        CK(ensure(ctx->bpairs, np, &bp));
        CK(ensure(ctx->bstarts, 4ull * np, &u));
        CK(ensure(ctx->border, 6ull * np, &u));
        CK(ensure(ctx->pflags, np, &u));
        CK(ensure(ctx->rstart, np, &u));
        CK(ensure(ctx->rend, np, &u));
        CK(ensure(ctx->fpair, F, &u));
        CK(ensure(ctx->vpair, V, &u));
        CK(ensure(ctx->bparams, np, &pp));
        CK(ensure(ctx->brbox, 12ull * np, &ull));
        for (int m = 0; m < 2; ++m) {
            CK(ensure(ctx->bidx[m], ctx->ncorner[m], &u));
            ctx->didx[m] = P_<uint32_t>(ctx->bidx[m]);
        }
        CK(hipMemcpyAsync(ctx->bpairs.p, ctx->hpairs.data(), sizeof(BatchPair) * np, hipMemcpyHostToDevice, ctx->s));
        CK(hipMemcpyAsync(ctx->bstarts.p, ctx->hstarts.data(), 16ull * np, hipMemcpyHostToDevice, ctx->s));
        CK(hipMemcpyAsync(ctx->border.p, ctx->horder.data(), 24ull * np, hipMemcpyHostToDevice, ctx->s));
    }
    if (ctx->cap_pairs < 4 * F) ctx->cap_pairs = std::max<uint64_t>(1 << 14, 4 * F);
    if (ctx->cap_tests < 2 * ctx->cap_pairs) ctx->cap_tests = 2 * ctx->cap_pairs;
    if (ctx->cap_esc < (1u << 14)) ctx->cap_esc = std::max<uint64_t>(1 << 14, ctx->cap_tests / 16);
    if (ctx->cap_reg < ctx->cap_tests / 32) ctx->cap_reg = std::max<uint64_t>(1 << 14, ctx->cap_tests / 32);
    if (ctx->cap_big < 1024) ctx->cap_big = 1024;
    if (ctx->cap_trip < (1u << 16)) ctx->cap_trip = 1u << 16;
    for (int attempt = 0; attempt < 6; ++attempt) {
        if ((r = alloc_list_buffers(ctx)) != MCGPU_SUCCESS) return r;
        ctx->stage_mode = 0;
        if ((r = enqueue_dispatch(ctx)) != MCGPU_SUCCESS) return r;
        bool overflow = false;
        if ((r = finish(ctx, info, &overflow)) != MCGPU_SUCCESS) return r;
        if (!overflow) {
            if ((r = finish_batch(ctx, info)) != MCGPU_SUCCESS) return r;
            return autotune_first(ctx);
        }
    }
    return MCGPU_ERROR_OUT_OF_MEMORY;
}

MCGPU_API mcgpu_result mcgpu_get_batch_status(mcgpu_context ctx, int32_t* status, uint64_t* registry_offsets,
                                              uint64_t capacity)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || !ctx->batch) return MCGPU_ERROR_NO_RESULT;
    const uint32_t np = ctx->npairs;
    if (capacity < np) return MCGPU_ERROR_INVALID_VALUE;
    uint64_t at = 0;
    for (uint32_t p = 0; p < np; ++p) {
        if (status) status[p] = ctx->hstatus[p];
        if (registry_offsets) registry_offsets[p] = at;
        at += ctx->hrend[p] - ctx->hrstart[p];
    }
    if (registry_offsets && capacity > np) registry_offsets[np] = at;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_batch_partial_cuts(mcgpu_context ctx, uint32_t* partial_cut, uint64_t capacity)
{
    if (!ctx || !partial_cut) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || !ctx->batch) return MCGPU_ERROR_NO_RESULT;
    if (capacity < ctx->npairs) return MCGPU_ERROR_INVALID_VALUE;
    std::memcpy(partial_cut, ctx->hpartial.data(), 4ull * ctx->npairs);
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_batch_candidate_pairs(mcgpu_context ctx, uint32_t* triples, uint64_t capacity)
{
    if (!ctx || !triples) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || !ctx->batch) return MCGPU_ERROR_NO_RESULT;
    const uint64_t n = ctx->C.n_pairs;
    std::vector<unsigned long long> h(n);
    if (n) CK(hipMemcpy(h.data(), ctx->pairs.p, 8 * n, hipMemcpyDeviceToHost));
    std::sort(h.begin(), h.end()); // source faces ascend with the pair
    const uint64_t m = std::min<uint64_t>(n, capacity);
    uint32_t p = 0;
    for (uint64_t i = 0; i < m; ++i) {
        const uint32_t sf = (uint32_t)(h[i] >> 32), cf = (uint32_t)h[i];
        while (p + 1 < ctx->npairs && ctx->hpairs[p + 1].fs0 <= sf) ++p;
        triples[3 * i] = p;
        triples[3 * i + 1] = sf - ctx->hpairs[p].fs0;
        triples[3 * i + 2] = cf - ctx->hpairs[p].fc0;
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_broad_phase(mcgpu_context ctx, const double* src_boxes, uint32_t num_src,
                                         const double* cut_boxes, uint32_t num_cut, uint32_t input_on_device,
                                         uint64_t* num_pairs)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx || !src_boxes || !cut_boxes || num_src == 0 || num_cut == 0) return MCGPU_ERROR_INVALID_VALUE;
    CK(hipSetDevice(ctx->device));
    ctx->have = false;
    ctx->broad_only = true;
    ctx->batch = false;
    ctx->cc_check = false;
    ctx->sort_bits = 22;
    ctx->trav = ctx->forced ? ctx->forced_trav : TravConfig{};
    ctx->nf[0] = num_src;
    ctx->nf[1] = num_cut;
    ctx->nv[0] = ctx->nv[1] = 0;
    ctx->ncorner[0] = ctx->ncorner[1] = 0;
    const uint64_t F = (uint64_t)num_src + num_cut;
    hipStream_t s = ctx->s;
    DevState* st;
    double* dp;
    float* fp;
    uint32_t* u;
    unsigned long long* ull;
    CK(ensure(ctx->state, 1, &st));
    CK(ensure(ctx->box, 6 * F, &dp));
    CK(ensure(ctx->rawbox, 6 * F, &fp));
    CK(ensure(ctx->mcode, F, &u));
    CK(ensure(ctx->mface, F, &u));
    CK(ensure(ctx->mcode2, F, &u));
    CK(ensure(ctx->mface2, F, &u));
    CK(ensure(ctx->cand, F + 8, &u));
    CK(ensure(ctx->cand_list, F + 8, &u));
    CK(ensure(ctx->hist, radix_hist_elems(F), &u));
    CK(ensure(ctx->scan_tmp, std::max(scan_tmp_elems(radix_hist_elems(F)), scan_tmp_elems(F + 8)) + 1024, &u));
    for (int m = 0; m < 2; ++m) {
        int D = 0;
        while ((1ull << D) < ctx->nf[m]) ++D;
        ctx->depth[m] = D;
        CK(ensure(ctx->nodes[m], 6 * ((2ull << D) - 1), &fp));
    }
    if (ctx->cap_pairs < 4 * F) ctx->cap_pairs = std::max<uint64_t>(1 << 14, 4 * F);
    const hipMemcpyKind kind = input_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    for (int attempt = 0; attempt < 6; ++attempt) {
        CK(ensure(ctx->raw_pairs, ctx->cap_pairs, &ull));
        CK(ensure(ctx->pairs, ctx->cap_pairs, &ull));
        CK(ensure(ctx->phash, next_pow2(2 * ctx->cap_pairs), &ull));
        Counters* C = &st->C;
        (void)hipEventRecord(ctx->ev[EV_A0], s);
        fill_u32(C, 0u, sizeof(Counters) / 4, s);
        CK(hipMemcpyAsync(dp, src_boxes, 48ull * num_src, kind, s));
        CK(hipMemcpyAsync(dp + 6ull * num_src, cut_boxes, 48ull * num_cut, kind, s));
        for (int m = 0; m < 2; ++m) {
            const uint64_t o = m ? num_src : 0;
            fill_u64(C->rbox[m], ~0ull, 3, s);
            k_box_bounds<<<grid_for(ctx->nf[m], kBlock, 64), kBlock, 0, s>>>(dp + 6 * o, ctx->nf[m], C->rbox[m]);
            k_morton_boxes<<<grid_for(ctx->nf[m]), kBlock, 0, s>>>(dp + 6 * o, ctx->nf[m], C->rbox[m], P_<uint32_t>(ctx->mcode) + o,
                                                                  P_<uint32_t>(ctx->mface) + o, P_<float>(ctx->rawbox) + 6 * o,
                                                                  m ? (1u << 21) : 0u);
        }
        enqueue_bvh_traverse(ctx, s, C, false);
        const uint64_t table = next_pow2(2 * ctx->cap_pairs);
        k_pair_hash_init<<<grid_for(table, kBlock, 1024), kBlock, 0, s>>>(P_<unsigned long long>(ctx->phash), ctx->cap_pairs,
                                                                          table, C);
        fill_u32(ctx->cand.p, 0u, F, s);
        k_filter_pairs<<<grid_for(ctx->cap_pairs, kFilterTile, 2048), kStageBlock, 0, s>>>(
            dp, P_<unsigned long long>(ctx->raw_pairs), ctx->cap_pairs, num_src, nullptr, P_<unsigned long long>(ctx->pairs),
            P_<unsigned long long>(ctx->phash), P_<uint32_t>(ctx->cand), P_<uint32_t>(ctx->cand_list), 0, C);
        (void)hipEventRecord(ctx->ev[EV_A_END], s);
        CK(hipStreamSynchronize(s));
        CK(hipGetLastError());
        CK(hipMemcpy(&ctx->C, C, sizeof(Counters), hipMemcpyDeviceToHost));
        if (ctx->C.ov_pairs) {
            ctx->cap_pairs = std::max<uint64_t>(2 * ctx->cap_pairs, ctx->C.n_raw + ctx->C.n_raw / 4);
            continue;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[EV_A0], ctx->ev[EV_A_END]) != hipSuccess) ms = 0.f;
        ctx->total_ms = ms;
        ctx->have = true;
        if (num_pairs) *num_pairs = ctx->C.n_pairs;
        return MCGPU_SUCCESS;
    }
    return MCGPU_ERROR_OUT_OF_MEMORY;
}

MCGPU_API mcgpu_result mcgpu_redispatch_perturbed(mcgpu_context ctx, const double* perturbation,
                                                  uint32_t enforcement_count, mcgpu_frontend_info* info)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx || !perturbation) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only || ctx->batch) return MCGPU_ERROR_NO_RESULT;
    CK(hipSetDevice(ctx->device));
    ctx->enforcement_count = enforcement_count;
    for (int attempt = 0; attempt < 6; ++attempt) {
        mcgpu_result r;
        if ((r = alloc_list_buffers(ctx)) != MCGPU_SUCCESS) return r;
        hipStream_t s = ctx->s;
        const uint64_t V = (uint64_t)ctx->nv[0] + ctx->nv[1];
        Counters* C = &P_<DevState>(ctx->state)->C;
        ctx->stage_mode = 1;
        ctx->last_timed = ctx->stage_timing || ctx->kernel_timing;
    ctx->last_kernel_only = !ctx->stage_timing && ctx->kernel_timing;
        (void)hipEventRecord(ctx->ev[EV_A0], s);
        ctx->state_copied = false; // finish() copies the state after its wait
        k_reset_narrow<<<1, 1, 0, s>>>(C);
        k_set_perturbation<<<1, 1, 0, s>>>(&P_<DevState>(ctx->state)->P, 1, perturbation[0], perturbation[1], perturbation[2]);
        stage_event(ctx, EV_A_PARAMS, s);
        k_shift_vertices<<<grid_for(3 * V), kBlock, 0, s>>>(ctx->dv[0], ctx->nv[0], ctx->dv[1], ctx->nv[1],
                                                            &P_<DevState>(ctx->state)->P, P_<double>(ctx->ps_v), 1);
        stage_event(ctx, EV_A_SHIFT, s);
        stage_event(ctx, EV_A_JOIN, s);
        stage_event(ctx, EV_A_FILTER, s);
        enqueue_narrow(ctx, make_soup(ctx), false);
        bool overflow = false;
        if ((r = finish(ctx, info, &overflow)) != MCGPU_SUCCESS) return r;
        if (!overflow) return MCGPU_SUCCESS;
        // a list overflowed: the pairs are intact, so only the narrow part reruns
    }
    return MCGPU_ERROR_OUT_OF_MEMORY;
}

MCGPU_API mcgpu_result mcgpu_predicate_stage(mcgpu_context ctx, uint64_t* counts, int8_t* qsign, int8_t* rsign,
                                             uint64_t capacity, float* device_ms)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx || !counts) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only || ctx->batch) return MCGPU_ERROR_NO_RESULT;
    CK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->s;
    const uint64_t n = ctx->C.n_tests;
    int8_t* q = nullptr;
    int8_t* r = nullptr;
    unsigned long long* cnt;
    CK(ensure(ctx->pcounts, 6, &cnt));
    const bool signs = qsign || rsign;
    if (signs) {
        CK(ensure(ctx->pqs, n, &q));
        CK(ensure(ctx->prs, n, &r));
    }
    fill_u64(cnt, 0ull, 6, s);
    (void)hipEventRecord(ctx->ev[EV_A0], s);
    k_predicates<<<grid_for(n, 256, 8192), 256, 0, s>>>(make_soup(ctx), P_<Plane>(ctx->planes), P_<uint32_t>(ctx->te),
                                                        P_<uint32_t>(ctx->tg), ctx->cap_tests,
                                                        &P_<DevState>(ctx->state)->C, q, r, cnt);
    (void)hipEventRecord(ctx->ev[EV_A_END], s);
    CK(hipStreamSynchronize(s));
    CK(hipGetLastError());
    unsigned long long h[6];
    CK(hipMemcpy(h, cnt, sizeof(h), hipMemcpyDeviceToHost));
    for (int k = 0; k < 6; ++k) counts[k] = h[k];
    if (device_ms) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[EV_A0], ctx->ev[EV_A_END]) != hipSuccess) ms = 0.f;
        *device_ms = ms;
    }
    ctx->stages_ready = false; // the events now time this kernel; stage times are gone
    if (signs && n) { // in the order of mcgpu_get_tests: sorted by (edge, face)
        std::vector<uint32_t> e(n), g(n);
        std::vector<int8_t> hq(n), hr(n);
        CK(hipMemcpy(e.data(), ctx->te.p, 4 * n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(g.data(), ctx->tg.p, 4 * n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(hq.data(), q, n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(hr.data(), r, n, hipMemcpyDeviceToHost));
        std::vector<uint32_t> perm(n);
        for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
        std::sort(perm.begin(), perm.end(),
                  [&](uint32_t a, uint32_t b) { return ((uint64_t)e[a] << 32 | g[a]) < ((uint64_t)e[b] << 32 | g[b]); });
        const uint64_t m = std::min<uint64_t>(n, capacity);
        for (uint64_t i = 0; i < m; ++i) {
            if (qsign) qsign[i] = hq[perm[i]];
            if (rsign) rsign[i] = hr[perm[i]];
        }
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_dispatch_enforced(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                               const mcgpu_dispatch_desc* desc, uint32_t attempts,
                                               mcgpu_frontend_info* info, double* perturbation, uint32_t* perturbations)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx || !desc || !info) return MCGPU_ERROR_INVALID_VALUE;
    mcgpu_dispatch_desc d = *desc;
    d.enforcement_count = 0;
    mcgpu_result r = mcgpu_dispatch_frontend(ctx, src, cut, &d, info);
    if (perturbations) *perturbations = 0;
    if (perturbation) perturbation[0] = perturbation[1] = perturbation[2] = 0.0;
    if (r != MCGPU_SUCCESS) return r;
    const bool enforce =
        (desc->flags & (MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION | MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION_ABSOLUTE)) != 0;
    if (!enforce) return MCGPU_SUCCESS;
    // preproc.cpp:2588-2631: on a violation, stop once `attempts` perturbations
    // were made, else draw the next vector (scaled by the same diag * constant
    // that sets eps, preproc.cpp:2667-2676) and rerun the narrow phase on the
    // kept candidates with the count of earlier perturbations
    uint32_t count = 0;
    while (info->status == MCGPU_STATUS_GENERAL_POSITION_VIOLATION && count < attempts) {
        const double scale = info->eps;
        double p[3];
        for (int i = 0; i < 3; ++i) p[i] = ctx->gp_uniform(ctx->gp_twister) * scale;
        if ((r = mcgpu_redispatch_perturbed(ctx, p, count, info)) != MCGPU_SUCCESS) return r;
        info->eps = scale;
        ++count;
        if (perturbation) perturbation[0] = p[0], perturbation[1] = p[1], perturbation[2] = p[2];
        if (perturbations) *perturbations = count;
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_registry(mcgpu_context ctx, uint32_t* edges, uint32_t* faces, double* points,
                                          uint64_t capacity)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only) return MCGPU_ERROR_NO_RESULT;
    const uint64_t n = std::min<uint64_t>(ctx->C.n_reg, capacity);
    const Buf& be = ctx->batch ? ctx->rle : ctx->re; // batches: pair-local ids, pairs in order
    const Buf& bg = ctx->batch ? ctx->rlg : ctx->rg;
    // ordered on the context's own stream (a server slot's stream synchronises with the null stream)
    if (edges) CK(hipMemcpyAsync(edges, be.p, 4 * n, hipMemcpyDeviceToHost, ctx->s));
    if (faces) CK(hipMemcpyAsync(faces, bg.p, 4 * n, hipMemcpyDeviceToHost, ctx->s));
    if (points) CK(hipMemcpyAsync(points, ctx->rpts.p, 24 * n, hipMemcpyDeviceToHost, ctx->s));
    CK(hipStreamSynchronize(ctx->s));
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_registry_device(mcgpu_context ctx, const uint32_t** edges, const uint32_t** faces,
                                             const double** points, uint64_t* count)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only) return MCGPU_ERROR_NO_RESULT;
    if (edges) *edges = P_<uint32_t>(ctx->batch ? ctx->rle : ctx->re);
    if (faces) *faces = P_<uint32_t>(ctx->batch ? ctx->rlg : ctx->rg);
    if (points) *points = P_<double>(ctx->rpts);
    if (count) *count = ctx->C.n_reg;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_candidate_pairs(mcgpu_context ctx, uint32_t* pairs, uint64_t capacity)
{
    if (!ctx || !pairs) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have) return MCGPU_ERROR_NO_RESULT;
    const uint64_t n = ctx->C.n_pairs;
    std::vector<unsigned long long> h(n);
    if (n) CK(hipMemcpy(h.data(), ctx->pairs.p, 8 * n, hipMemcpyDeviceToHost));
    std::sort(h.begin(), h.end()); // listing order only; the set is the device's
    const uint64_t m = std::min<uint64_t>(n, capacity);
    for (uint64_t i = 0; i < m; ++i) {
        pairs[2 * i] = (uint32_t)(h[i] >> 32);
        pairs[2 * i + 1] = (uint32_t)h[i];
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_edges(mcgpu_context ctx, uint32_t* v0, uint32_t* v1, int32_t* f0, int32_t* f1,
                                       uint64_t capacity)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only) return MCGPU_ERROR_NO_RESULT;
    const uint64_t n = std::min<uint64_t>(ctx->C.n_edges_src + ctx->C.n_edges_cut, capacity);
    if (v0) CK(hipMemcpy(v0, ctx->ev0.p, 4 * n, hipMemcpyDeviceToHost));
    if (v1) CK(hipMemcpy(v1, ctx->ev1.p, 4 * n, hipMemcpyDeviceToHost));
    if (f0) CK(hipMemcpy(f0, ctx->ef0.p, 4 * n, hipMemcpyDeviceToHost));
    if (f1) CK(hipMemcpy(f1, ctx->ef1.p, 4 * n, hipMemcpyDeviceToHost));
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_tests(mcgpu_context ctx, uint32_t* edges, uint32_t* faces, int8_t* qsign,
                                       int8_t* rsign, char* type, char* pip, uint64_t capacity)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only || !ctx->keep_tests) return MCGPU_ERROR_NO_RESULT;
    const uint64_t n = ctx->C.n_tests;
    std::vector<uint32_t> e(n), g(n);
    std::vector<int8_t> qs(n), rs(n);
    std::vector<char> ty(n), pp(n);
    if (n) {
        CK(hipMemcpy(e.data(), ctx->te.p, 4 * n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(g.data(), ctx->tg.p, 4 * n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(qs.data(), ctx->tqs.p, n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(rs.data(), ctx->trs.p, n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(ty.data(), ctx->ttype.p, n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(pp.data(), ctx->tpip.p, n, hipMemcpyDeviceToHost));
    }
    std::vector<uint64_t> order(n);
    for (uint64_t i = 0; i < n; ++i) order[i] = ((uint64_t)e[i] << 32 | g[i]) << 0;
    std::vector<uint32_t> perm(n);
    for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    std::sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return order[a] < order[b]; });
    const uint64_t m = std::min<uint64_t>(n, capacity);
    for (uint64_t i = 0; i < m; ++i) {
        const uint32_t j = perm[i];
        if (edges) edges[i] = e[j];
        if (faces) faces[i] = g[j];
        if (qsign) qsign[i] = qs[j];
        if (rsign) rsign[i] = rs[j];
        if (type) type[i] = ty[j];
        if (pip) pip[i] = pp[j];
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_face_planes(mcgpu_context ctx, uint32_t* faces, double* normals, double* d,
                                             int32_t* max_comp, int32_t* triple, uint64_t capacity)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only) return MCGPU_ERROR_NO_RESULT;
    const uint64_t F = (uint64_t)ctx->nf[0] + ctx->nf[1];
    std::vector<uint32_t> cand(F);
    std::vector<Plane> pl(F);
    CK(hipMemcpy(cand.data(), ctx->cand.p, 4 * F, hipMemcpyDeviceToHost));
    CK(hipMemcpy(pl.data(), ctx->planes.p, F * sizeof(Plane), hipMemcpyDeviceToHost));
    uint64_t k = 0;
    for (uint64_t f = 0; f < F && k < capacity; ++f) {
        if (!cand[f]) continue;
        if (faces) faces[k] = (uint32_t)f;
        for (int j = 0; j < 3; ++j) {
            if (normals) normals[3 * k + j] = pl[f].n[j];
            if (triple) triple[3 * k + j] = pl[f].tri[0] == kCollinear ? -1 : pl[f].tri[j];
        }
        if (d) d[k] = pl[f].d;
        if (max_comp) max_comp[k] = pl[f].k;
        ++k;
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_ps_vertices(mcgpu_context ctx, double* xyz, uint64_t capacity)
{
    if (!ctx || !xyz) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only) return MCGPU_ERROR_NO_RESULT;
    const uint64_t n = std::min<uint64_t>((uint64_t)ctx->nv[0] + ctx->nv[1], capacity);
    CK(hipMemcpy(xyz, ctx->ps_v.p, 24 * n, hipMemcpyDeviceToHost));
    return MCGPU_SUCCESS;
}

MCGPU_API int mcgpu_stage_times(mcgpu_context ctx, float* ms, int capacity)
{
    if (!ctx || !ms || !ctx->have || ctx->broad_only || !ctx->last_timed) return 0;
    if (!ctx->stages_ready) compute_stage_times(ctx);
    int n = std::min(capacity, (int)ST_COUNT);
    for (int i = 0; i < n; ++i) ms[i] = ctx->stage_ms[i];
    return n;
}

MCGPU_API mcgpu_result mcgpu_set_stage_timing(mcgpu_context ctx, int enable)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (enable < 0 || enable > 2) return MCGPU_ERROR_INVALID_VALUE;
    ctx->stage_timing = enable == 1;
    ctx->kernel_timing = enable == 2;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_autotune(mcgpu_context ctx, mcgpu_tuning* out)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only || ctx->stage_mode != 0 || ctx->res) return MCGPU_ERROR_NO_RESULT;
    CK(hipSetDevice(ctx->device));
    return autotune(ctx, out);
}

MCGPU_API mcgpu_result mcgpu_cutpath_edges(mcgpu_context ctx, const uint32_t* order, mcgpu_cutpath_info* info)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->have || ctx->broad_only || ctx->batch || ctx->res) return MCGPU_ERROR_NO_RESULT;
    CK(hipSetDevice(ctx->device));
    ctx->cp_have = false;
    const uint64_t n = ctx->C.n_reg, n2 = 2 * n;
    const uint64_t F = (uint64_t)ctx->nf[0] + ctx->nf[1];
    const int fbits = bits_for(F + 1);
    if (n == 0 || 2 * fbits > 62 || n2 > 0xffffffffull) {
        if (n == 0) {
            ctx->cp_counts = CutCounts{};
            ctx->cp_have = true;
            if (info) std::memset(info, 0, sizeof(*info));
            return MCGPU_SUCCESS;
        }
        return MCGPU_ERROR_UNSUPPORTED;
    }
    hipStream_t s = ctx->s;
    uint32_t *ord = nullptr, *u;
    unsigned long long* ull;
    uint8_t* b8;
    Trip* tp;
    CutCounts* cc;
    if (order) { // the caller's registry order: rank k -> registry index, a permutation
        std::vector<uint8_t> seen(n, 0);
        for (uint64_t k = 0; k < n; ++k) {
            if (order[k] >= n || seen[order[k]]) return MCGPU_ERROR_INVALID_VALUE;
            seen[order[k]] = 1;
        }
        CK(ensure(ctx->cp_order, n, &ord));
        CK(hipMemcpyAsync(ord, order, 4 * n, hipMemcpyHostToDevice, s));
    }
    CK(ensure(ctx->cp_keys, n2, &ull));
    CK(ensure(ctx->cp_keys2, n2, &ull));
    for (Buf* b : {&ctx->cp_vals, &ctx->cp_vals2, &ctx->cp_flag, &ctx->cp_gid, &ctx->cp_gstart, &ctx->cp_ncand,
                   &ctx->cp_coff, &ctx->cp_src, &ctx->cp_tgt, &ctx->cp_fa, &ctx->cp_fb, &ctx->cp_group, &ctx->cp_keep,
                   &ctx->cp_pos, &ctx->cp_osrc, &ctx->cp_otgt, &ctx->cp_ofa, &ctx->cp_ofb, &ctx->cp_ogroup})
        CK(ensure(*b, n2 + 8, &u));
    for (Buf* b : {&ctx->cp_kind, &ctx->cp_acc, &ctx->cp_okind}) CK(ensure(*b, n2 + 8, &b8));
    CK(ensure(ctx->cp_trip, n2 + 8, &tp));
    const uint64_t tslots = next_pow2(2 * n2 + 2);
    CK(ensure(ctx->cp_table, 2 * tslots, &ull));
    CK(ensure(ctx->cp_cnt, 1, &cc));
    CK(ensure(ctx->hist, radix_hist_elems(n2), &u));
    CK(ensure(ctx->scan_tmp, std::max<uint64_t>(scan_tmp_elems(n2 + 8), scan_tmp_elems(radix_hist_elems(n2))) + 1024, &u));
    fill_u32(cc, 0u, sizeof(CutCounts) / 4, s);
    const uint32_t* re = P_<uint32_t>(ctx->re);
    const uint32_t* rg = P_<uint32_t>(ctx->rg);
    const double* rp = P_<double>(ctx->rpts);
    const int32_t* ef0 = P_<int32_t>(ctx->ef0);
    const int32_t* ef1 = P_<int32_t>(ctx->ef1);
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    (void)hipEventRecord(e0, s);
    // groups: both keys of every point, stably sorted by (source face, cut face)
    k_cut_items<<<grid_for(n), kBlock, 0, s>>>(n, ord, re, rg, ef0, ef1, ctx->nf[0], fbits, P_<unsigned long long>(ctx->cp_keys),
                                             P_<uint32_t>(ctx->cp_vals));
    RadixTemp rt{P_<uint32_t>(ctx->hist), P_<uint32_t>(ctx->scan_tmp)};
    const bool alt = radix_sort<unsigned long long>(P_<unsigned long long>(ctx->cp_keys), P_<uint32_t>(ctx->cp_vals),
                                                    P_<unsigned long long>(ctx->cp_keys2), P_<uint32_t>(ctx->cp_vals2),
                                                    nullptr, n2, n2, 2 * fbits, rt, s);
    const unsigned long long* keys = P_<unsigned long long>(alt ? ctx->cp_keys2 : ctx->cp_keys);
    const uint32_t* vals = P_<uint32_t>(alt ? ctx->cp_vals2 : ctx->cp_vals);
    uint32_t* flag = P_<uint32_t>(ctx->cp_flag);
    uint32_t* gid = P_<uint32_t>(ctx->cp_gid);
    uint32_t* gstart = P_<uint32_t>(ctx->cp_gstart);
    uint32_t* ncand = P_<uint32_t>(ctx->cp_ncand);
    uint32_t* coff = P_<uint32_t>(ctx->cp_coff);
    uint32_t* stmp = P_<uint32_t>(ctx->scan_tmp);
    k_cut_flags<<<grid_for(n2), kBlock, 0, s>>>(keys, n2, flag, cc);
    scan_exclusive(flag, gid, n2, stmp, &cc->groups, s);
    k_cut_starts<<<grid_for(n2), kBlock, 0, s>>>(flag, gid, n2, gstart);
    fill_u32(ncand, 0u, n2, s);
    k_cut_sizes<<<grid_for(n2), kBlock, 0, s>>>(gstart, cc, ncand);
    scan_exclusive(ncand, coff, n2, stmp, &cc->candidates, s);
    CutCands cd{P_<uint32_t>(ctx->cp_src), P_<uint32_t>(ctx->cp_tgt), P_<uint32_t>(ctx->cp_fa), P_<uint32_t>(ctx->cp_fb),
                P_<uint32_t>(ctx->cp_group), P_<uint8_t>(ctx->cp_kind), P_<uint8_t>(ctx->cp_acc)};
    k_cut_groups<<<grid_for(n2, 64), 64, 0, s>>>(gstart, cc, keys, vals, ord, rp, coff, fbits, tp, cd);
    const Soup S = make_soup(ctx);
    k_cut_midpoints<<<grid_for(n2), kBlock, 0, s>>>(cc, cd, ord, re, rg, ef0, ef1, rp, P_<Plane>(ctx->planes),
                                                   P_<uint32_t>(ctx->cand), S, cc);
    // first accepted candidate of each edge, then the kept ones compacted in candidate order
    unsigned long long* table = P_<unsigned long long>(ctx->cp_table);
    fill_u64(table, kCutNone, 2 * tslots, s);
    k_cut_first<<<grid_for(n2), kBlock, 0, s>>>(cc, cd, table, tslots - 1);
    uint32_t* keep = P_<uint32_t>(ctx->cp_keep);
    uint32_t* pos = P_<uint32_t>(ctx->cp_pos);
    fill_u32(keep, 0u, n2, s);
    k_cut_keep<<<grid_for(n2), kBlock, 0, s>>>(cc, cd, table, tslots - 1, keep);
    scan_exclusive(keep, pos, n2, stmp, &cc->edges, s);
    CutEdges oe{P_<uint32_t>(ctx->cp_osrc), P_<uint32_t>(ctx->cp_otgt), P_<uint32_t>(ctx->cp_ofa), P_<uint32_t>(ctx->cp_ofb),
                P_<uint32_t>(ctx->cp_ogroup), P_<uint8_t>(ctx->cp_okind)};
    k_cut_emit<<<grid_for(n2), kBlock, 0, s>>>(cc, cd, keep, pos, oe);
    (void)hipEventRecord(e1, s);
    hipError_t err = hipStreamSynchronize(s);
    if (err == hipSuccess) err = hipGetLastError();
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (err == hipSuccess) err = hipMemcpy(&ctx->cp_counts, cc, sizeof(CutCounts), hipMemcpyDeviceToHost);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    CK(err);
    ctx->cp_have = true;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->num_edges = ctx->cp_counts.edges;
        info->num_groups = ctx->cp_counts.groups;
        info->num_candidates = ctx->cp_counts.candidates;
        info->short_group = ctx->cp_counts.short_group;
        info->unsupported = ctx->cp_counts.unsupported;
        info->device_ms = ms;
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_cutpath_edges(mcgpu_context ctx, uint32_t* src, uint32_t* tgt, uint32_t* faces,
                                               uint32_t* groups, uint8_t* kinds, uint64_t capacity)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->cp_have) return MCGPU_ERROR_NO_RESULT;
    const uint64_t n = std::min<uint64_t>(ctx->cp_counts.edges, capacity);
    if (n == 0) return MCGPU_SUCCESS;
    if (src) CK(hipMemcpy(src, ctx->cp_osrc.p, 4 * n, hipMemcpyDeviceToHost));
    if (tgt) CK(hipMemcpy(tgt, ctx->cp_otgt.p, 4 * n, hipMemcpyDeviceToHost));
    if (groups) CK(hipMemcpy(groups, ctx->cp_ogroup.p, 4 * n, hipMemcpyDeviceToHost));
    if (kinds) CK(hipMemcpy(kinds, ctx->cp_okind.p, n, hipMemcpyDeviceToHost));
    if (faces) {
        std::vector<uint32_t> a(n), b(n);
        CK(hipMemcpy(a.data(), ctx->cp_ofa.p, 4 * n, hipMemcpyDeviceToHost));
        CK(hipMemcpy(b.data(), ctx->cp_ofb.p, 4 * n, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < n; ++i) {
            faces[2 * i] = a[i];
            faces[2 * i + 1] = b[i];
        }
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_set_graphs(mcgpu_context ctx, int enable)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->own_streams) return MCGPU_ERROR_UNSUPPORTED;
    ctx->graphs = enable != 0;
    if (!ctx->graphs) {
        CK(hipSetDevice(ctx->device));
        graph_drop(ctx);
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_graph_stats(mcgpu_context ctx, uint64_t* recordings, uint64_t* replays,
                                         uint32_t* graph_launches)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    uint32_t n = 0;
    for (const GraphOp& op : ctx->prog) n += op.kind == GraphOp::LAUNCH;
    if (recordings) *recordings = ctx->recordings;
    if (replays) *replays = ctx->replays;
    if (graph_launches) *graph_launches = n;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_set_batch_split(mcgpu_context ctx, uint32_t broad_share_pct)
{
    if (!ctx || broad_share_pct < 10 || broad_share_pct > 90) return MCGPU_ERROR_INVALID_VALUE;
    if (!ctx->own_streams) return MCGPU_ERROR_UNSUPPORTED;
    if (ctx->s2b) return MCGPU_ERROR_INVALID_VALUE; // the batch streams exist: their masks are fixed
    ctx->batch_share_pct = (int)broad_share_pct;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_set_autotune(mcgpu_context ctx, int enable)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    ctx->autotune = enable != 0;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_set_tuning(mcgpu_context ctx, const mcgpu_tuning* t)
{
    if (!ctx) return MCGPU_ERROR_INVALID_VALUE;
    if (!t) {
        ctx->forced = false;
        return MCGPU_SUCCESS;
    }
    TravConfig c;
    int ll = 0;
    while (ll < 8 && (1u << ll) < t->wave_leaves) ++ll;
    c.ll = ll;
    c.stride = (int)t->tree_stride;
    c.queues = (int)t->queues;
    c.chunk = (int)t->chunk;
    if ((1u << ll) != t->wave_leaves || !trav_config_valid(c) || t->broad_share_pct != 0)
        return MCGPU_ERROR_INVALID_VALUE;
    ctx->forced = true;
    ctx->forced_trav = c;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_get_tuning(mcgpu_context ctx, mcgpu_tuning* out)
{
    if (!ctx || !out) return MCGPU_ERROR_INVALID_VALUE;
    *out = ctx->used;
    return MCGPU_SUCCESS;
}

MCGPU_API const char* mcgpu_stage_name(int stage)
{
    return (stage >= 0 && stage < ST_COUNT) ? kStageNames[stage] : "";
}

// ===========================================================================
// Serving: one resident source mesh, cut requests as they arrive
// ===========================================================================

} // extern "C"

struct mcgpu_server_t {
    int device = 0;
    mcgpu_context master = nullptr; // its buffers hold the resident source
    Resident res;
    mcgpu_server_limits lim{};
    mcgpu_dispatch_desc desc{};
    std::vector<mcgpu_context> slots;
    std::vector<hipEvent_t> done;
    std::vector<std::vector<uint32_t>> reg_e, reg_g;
    std::vector<std::vector<double>> reg_p;
    uint64_t slot_bytes = 0;
};

namespace {

uint64_t ctx_bytes(mcgpu_context c)
{
    uint64_t t = 0;
    const Buf* all[] = {&c->in_v[1], &c->in_idx[1], &c->in_sz[1], &c->off[1], &c->state, &c->ps_v, &c->box,
                        &c->hkeys[1], &c->hfirst[1], &c->cface[1], &c->eflag[0], &c->eid[0], &c->eflag[1],
                        &c->eid[1], &c->cedge[1],
                        &c->etotal, &c->ev0, &c->ev1, &c->ef0, &c->ef1, &c->mcode, &c->mface, &c->mcode2,
                        &c->mface2, &c->nodes[1], &c->rawbox, &c->raw_pairs, &c->pairs, &c->phash, &c->cand,
                        &c->cand_list, &c->planes, &c->big_list, &c->big_off, &c->big_scratch, &c->te, &c->tg, &c->esc,
                        &c->rkey, &c->rval, &c->rkey2, &c->rval2, &c->rp, &c->re, &c->rg, &c->rpts, &c->rcount,
                        &c->re_in, &c->rg_in, &c->shkeys, &c->shcnt, &c->scan_tmp, &c->scan_tmp_c, &c->hist,
                        &c->sbsum, &c->sitems, &c->slo, &c->shi, &c->snitem, &c->slist, &c->sckpt,
                        &c->sflags};
    for (const Buf* b : all) t += b->cap;
    return t;
}

// HBM a request touches: the per-face, per-corner and per-vertex arrays of
// its polygon soup (source and cut) and the lists sized from them
uint64_t request_footprint(const Resident& r, const mcgpu_mesh* cut, uint64_t cut_corners)
{
    const uint64_t F = (uint64_t)r.nf + cut->num_faces, V = (uint64_t)r.nv + cut->num_vertices;
    const uint64_t K = r.ncorner + cut_corners;
    return 400 * F + 100 * K + 48 * V;
}

uint64_t mesh_corners(const mcgpu_mesh* m)
{
    if (!m->face_sizes) return 3ull * m->num_faces;
    uint64_t k = 0;
    for (uint32_t f = 0; f < m->num_faces; ++f) k += m->face_sizes[f];
    return k;
}

mcgpu_result slot_create(int device, mcgpu_context* out)
{
    mcgpu_context ctx = new mcgpu_context_t();
    ctx->device = device;
    ctx->own_streams = false;
    // A slot's stream carries a CU mask (every CU), which gives it a hardware
    // queue of its own: plain streams share the process's few queues (four
    // by default), and in a program that had made and released other
    // contexts first, the slots' plain streams shared them, so that slots
    // ran one after another (1,900 requests/s served at 3,000 offered; a
    // fresh process, or sixteen queues, served all 3,000).  Such a stream
    // synchronises with the null stream, so a slot's copies are ordered on
    // its own stream (finish, mcgpu_get_registry).
    int ncu = 0;
    if (hipSetDevice(device) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu < 1) {
        delete ctx;
        return MCGPU_ERROR_DEVICE;
    }
    std::vector<uint32_t> all((ncu + 31) / 32, ~0u);
    if (hipExtStreamCreateWithCUMask(&ctx->s, (uint32_t)all.size(), all.data()) != hipSuccess) {
        delete ctx;
        return MCGPU_ERROR_DEVICE;
    }
    ctx->s2 = ctx->s3 = ctx->s2b = ctx->s3b = ctx->ssum = ctx->s;
    ctx->ncu = ncu;
    for (int i = 0; i < EV_COUNT; ++i) (void)hipEventCreate(&ctx->ev[i]);
    *out = ctx;
    return MCGPU_SUCCESS;
}

// a slot's view of one request: the resident source as mesh 0, the cut as mesh 1
mcgpu_result slot_bind(mcgpu_server srv, mcgpu_context ctx, const mcgpu_mesh* cut, bool on_device)
{
    const Resident& r = srv->res;
    ctx->res = &srv->res;
    ctx->have = false;
    ctx->broad_only = false;
    ctx->batch = false;
    ctx->cc_check = (srv->desc.flags & MCGPU_DISPATCH_CHECK_COMPONENTS) != 0;
    ctx->sort_bits = 22;
    ctx->flags = srv->desc.flags;
    ctx->gp_constant = srv->desc.gp_constant;
    ctx->enforcement_count = 0;
    ctx->keep_tests = 0;
    ctx->nv[0] = r.nv;
    ctx->nf[0] = r.nf;
    ctx->ncorner[0] = r.ncorner;
    ctx->dv[0] = r.dv;
    ctx->didx[0] = r.didx;
    ctx->dsz[0] = r.dsz;
    mcgpu_result rr;
    if ((rr = upload_mesh(ctx, 1, cut, on_device)) != MCGPU_SUCCESS) return rr;
    const uint64_t F = (uint64_t)ctx->nf[0] + ctx->nf[1];
    ctx->gbits = bits_for(F + 1);
    ctx->kbits = ctx->gbits + bits_for(ctx->ncorner[0] + ctx->ncorner[1] + 1);
    if (ctx->kbits > 64) return MCGPU_ERROR_UNSUPPORTED;
    if ((rr = alloc_mesh_buffers(ctx)) != MCGPU_SUCCESS) return rr;
    if (ctx->cap_pairs < 4 * F) ctx->cap_pairs = std::max<uint64_t>(1 << 14, 4 * F);
    if (ctx->cap_tests < 2 * ctx->cap_pairs) ctx->cap_tests = 2 * ctx->cap_pairs;
    if (ctx->cap_esc < (1u << 14)) ctx->cap_esc = std::max<uint64_t>(1 << 14, ctx->cap_tests / 16);
    if (ctx->cap_reg < (1u << 14))
        ctx->cap_reg = std::min<uint64_t>(std::max<uint64_t>(1 << 14, ctx->cap_tests / 8), kRegMergeTiles * kRegTile);
    if (ctx->cap_big < 1024) ctx->cap_big = 1024;
    if (ctx->cap_trip < (1u << 16)) ctx->cap_trip = 1u << 16;
    return alloc_list_buffers(ctx);
}

void slot_enqueue(mcgpu_context ctx, hipEvent_t done)
{
    ctx->stage_mode = 0;
    ctx->state_copied = false;
    enqueue_validate(ctx);
    enqueue_frontend(ctx);
    (void)enqueue_state_copy(ctx); // on failure finish() copies after its wait
    (void)hipEventRecord(done, ctx->s);
}

} // namespace

extern "C" {

MCGPU_API mcgpu_result mcgpu_server_create(int device, const mcgpu_mesh* src, const mcgpu_dispatch_desc* desc,
                                           const mcgpu_server_limits* limits, mcgpu_server* out)
{
    if (!out || !src || !desc || !limits || limits->max_slots == 0 || limits->max_cut_faces == 0 ||
        limits->max_cut_vertices < 3)
        return MCGPU_ERROR_INVALID_VALUE;
    *out = nullptr;
    mcgpu_result r;
    if ((r = check_mesh(src)) != MCGPU_SUCCESS) return r;
    mcgpu_server srv = new mcgpu_server_t();
    srv->device = device;
    srv->lim = *limits;
    srv->desc = *desc;
    srv->desc.keep_tests = 0;
    // the resident source: one dispatch of it against a far-away triangle
    // leaves every source-only product in the master context's buffers
    if ((r = mcgpu_create_context(device, &srv->master)) != MCGPU_SUCCESS) {
        delete srv;
        return r;
    }
    const double tri[9] = {1e6, 1e6, 1e6, 1e6 + 1.0, 1e6, 1e6, 1e6, 1e6 + 1.0, 1e6};
    const uint32_t tf[3] = {0, 1, 2};
    mcgpu_context m = srv->master;
    {
        mcgpu_mesh dummy{tri, tf, nullptr, 3, 1};
        mcgpu_dispatch_desc d = srv->desc;
        double* dv = nullptr;
        uint32_t* di = nullptr;
        if (desc->input_on_device) { // the dummy follows the source's residency
            if (hipMalloc(&dv, sizeof(tri)) != hipSuccess || hipMalloc(&di, sizeof(tf)) != hipSuccess ||
                hipMemcpy(dv, tri, sizeof(tri), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(di, tf, sizeof(tf), hipMemcpyHostToDevice) != hipSuccess) {
                if (dv) (void)hipFree(dv);
                if (di) (void)hipFree(di);
                mcgpu_release_context(m);
                delete srv;
                return MCGPU_ERROR_OUT_OF_MEMORY;
            }
            dummy.vertices = dv;
            dummy.face_indices = di;
        }
        mcgpu_frontend_info info;
        r = mcgpu_dispatch_frontend(m, src, &dummy, &d, &info);
        if (dv) (void)hipFree(dv);
        if (di) (void)hipFree(di);
        if (r != MCGPU_SUCCESS) {
            mcgpu_release_context(m);
            delete srv;
            return r;
        }
    }
    Resident& R = srv->res;
    R.nv = m->nv[0];
    R.nf = m->nf[0];
    R.ncorner = m->ncorner[0];
    R.dv = m->dv[0];
    R.didx = m->didx[0];
    R.dsz = m->dsz[0];
    R.off = P_<uint32_t>(m->off[0]);
    R.cedge = P_<uint32_t>(m->cedge[0]);
    R.ev0 = P_<uint32_t>(m->ev0);
    R.ev1 = P_<uint32_t>(m->ev1);
    R.ef0 = P_<int32_t>(m->ef0);
    R.ef1 = P_<int32_t>(m->ef1);
    R.ne = (uint32_t)m->C.n_edges_src;
    R.rawbox = P_<float>(m->rawbox);
    R.sface = P_<uint32_t>(m->mface_alt ? m->mface2 : m->mface);
    R.nodes = P_<float>(m->nodes[0]);
    R.depth = m->depth[0];
    ResidentGpu h;
    for (int j = 0; j < 3; ++j) h.sum[j] = m->P.sum[0][j];
    for (int j = 0; j < 6; ++j) h.rbox[j] = m->C.rbox[0][j];
    h.ne = m->C.n_edges_src;
    if (hipMalloc(&R.dres, sizeof(ResidentGpu)) != hipSuccess ||
        hipMemcpy(R.dres, &h, sizeof(ResidentGpu), hipMemcpyHostToDevice) != hipSuccess) {
        mcgpu_server_release(srv);
        return MCGPU_ERROR_DEVICE;
    }
    // request slots, each sized once for the largest request it may hold
    std::vector<double> mv(3ull * limits->max_cut_vertices, 0.0);
    const uint64_t mk = limits->max_cut_corners ? limits->max_cut_corners : 3ull * limits->max_cut_faces;
    for (uint32_t i = 0; i < limits->max_slots; ++i) {
        mcgpu_context c;
        if ((r = slot_create(device, &c)) != MCGPU_SUCCESS) {
            mcgpu_server_release(srv);
            return r;
        }
        srv->slots.push_back(c);
        hipEvent_t e;
        (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
        srv->done.push_back(e);
        c->res = &srv->res;
        c->nv[0] = R.nv, c->nf[0] = R.nf, c->ncorner[0] = R.ncorner;
        c->nv[1] = limits->max_cut_vertices, c->nf[1] = limits->max_cut_faces, c->ncorner[1] = mk;
        double* dp;
        uint32_t* u;
        const uint64_t F = (uint64_t)R.nf + limits->max_cut_faces;
        if (ensure(c->in_v[1], 3ull * limits->max_cut_vertices, &dp) != hipSuccess ||
            ensure(c->in_idx[1], mk, &u) != hipSuccess || ensure(c->in_sz[1], limits->max_cut_faces, &u) != hipSuccess ||
            alloc_mesh_buffers(c) != MCGPU_SUCCESS) {
            mcgpu_server_release(srv);
            return MCGPU_ERROR_OUT_OF_MEMORY;
        }
        c->cap_pairs = std::max<uint64_t>(1 << 14, 4 * F);
        c->cap_tests = 2 * c->cap_pairs;
        c->cap_esc = std::max<uint64_t>(1 << 14, c->cap_tests / 16);
        c->cap_reg = std::min<uint64_t>(std::max<uint64_t>(1 << 14, c->cap_tests / 8), kRegMergeTiles * kRegTile);
        c->cap_big = 1024;
        c->cap_trip = 1u << 16;
        if (alloc_list_buffers(c) != MCGPU_SUCCESS) {
            mcgpu_server_release(srv);
            return MCGPU_ERROR_OUT_OF_MEMORY;
        }
        srv->slot_bytes = std::max(srv->slot_bytes, ctx_bytes(c));
    }
    *out = srv;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_server_release(mcgpu_server srv)
{
    if (!srv) return MCGPU_ERROR_INVALID_VALUE;
    (void)hipSetDevice(srv->device);
    for (mcgpu_context c : srv->slots) mcgpu_release_context(c);
    for (hipEvent_t e : srv->done) (void)hipEventDestroy(e);
    if (srv->res.dres) (void)hipFree(srv->res.dres);
    if (srv->master) mcgpu_release_context(srv->master);
    delete srv;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_server_info(mcgpu_server srv, uint32_t* slots, uint64_t* slot_bytes, uint32_t* source_edges)
{
    if (!srv) return MCGPU_ERROR_INVALID_VALUE;
    if (slots) *slots = (uint32_t)srv->slots.size();
    if (slot_bytes) *slot_bytes = srv->slot_bytes;
    if (source_edges) *source_edges = srv->res.ne;
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_server_run(mcgpu_server srv, const mcgpu_mesh* cuts, const double* arrival_s, uint32_t n,
                                        uint32_t keep_registries, mcgpu_request_result* out)
{
    if (!srv || !cuts || !arrival_s || !out) return MCGPU_ERROR_INVALID_VALUE;
    for (uint32_t i = 0; i < n; ++i) {
        mcgpu_result r;
        if ((r = check_mesh(&cuts[i])) != MCGPU_SUCCESS) return r;
        if (cuts[i].num_faces > srv->lim.max_cut_faces || cuts[i].num_vertices > srv->lim.max_cut_vertices)
            return MCGPU_ERROR_INVALID_VALUE;
    }
    if (hipSetDevice(srv->device) != hipSuccess) return MCGPU_ERROR_DEVICE;
    const bool on_device = srv->desc.input_on_device != 0;
    srv->reg_e.assign(keep_registries ? n : 0, {});
    srv->reg_g.assign(keep_registries ? n : 0, {});
    srv->reg_p.assign(keep_registries ? n : 0, {});
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return arrival_s[a] < arrival_s[b]; });
    std::vector<uint64_t> est(n);
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = mcgpu_request_result{};
        out[i].arrival_s = arrival_s[i];
        est[i] = request_footprint(srv->res, &cuts[i], on_device ? 3ull * cuts[i].num_faces : mesh_corners(&cuts[i]));
    }
    const size_t S = srv->slots.size();
    std::vector<int64_t> busy(S, -1);
    std::vector<uint64_t> held(S, 0);
    std::vector<int> reruns(S, 0);
    uint64_t inflight = 0;
    std::deque<uint32_t> waiting;
    uint32_t next = 0, finished = 0;
    const auto t0 = std::chrono::steady_clock::now();
    auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    while (finished < n) {
        bool progressed = false;
        // completions: a finished slot is read back and freed
        for (size_t s = 0; s < S; ++s) {
            if (busy[s] < 0 || hipEventQuery(srv->done[s]) != hipSuccess) continue;
            const uint32_t i = (uint32_t)busy[s];
            mcgpu_context c = srv->slots[s];
            mcgpu_frontend_info info;
            bool overflow = false;
            mcgpu_result r = finish(c, &info, &overflow);
            if (r == MCGPU_SUCCESS && overflow) { // a list outgrew its capacity: grow and run it again
                if (++reruns[s] > 5) r = MCGPU_ERROR_OUT_OF_MEMORY;
                else if ((r = alloc_list_buffers(c)) == MCGPU_SUCCESS) {
                    slot_enqueue(c, srv->done[s]);
                    continue;
                }
            }
            out[i].done_s = now();
            if (r != MCGPU_SUCCESS) {
                out[i].status = (int32_t)r; // negative: an mcgpu_result
            } else {
                out[i].status = info.status;
                out[i].partial_cut = info.partial_cut;
                out[i].num_registry = info.num_registry;
                out[i].num_candidate_pairs = info.num_candidate_pairs;
                if (keep_registries && info.num_registry) {
                    const uint64_t m = info.num_registry;
                    srv->reg_e[i].resize(m);
                    srv->reg_g[i].resize(m);
                    srv->reg_p[i].resize(3 * m);
                    mcgpu_get_registry(c, srv->reg_e[i].data(), srv->reg_g[i].data(), srv->reg_p[i].data(), m);
                }
            }
            inflight -= held[s];
            held[s] = 0;
            busy[s] = -1;
            ++finished;
            progressed = true;
        }
        // arrivals
        const double t = now();
        while (next < n && arrival_s[order[next]] <= t) waiting.push_back(order[next++]);
        // admissions: oldest first, while a slot is free and the in-flight
        // footprint stays within the budget (an idle server always admits)
        while (!waiting.empty()) {
            size_t s = 0;
            while (s < S && busy[s] >= 0) ++s;
            if (s == S) break;
            const uint32_t i = waiting.front();
            if (inflight > 0 && inflight + est[i] > srv->lim.hbm_budget) break;
            waiting.pop_front();
            out[i].start_s = now();
            mcgpu_context c = srv->slots[s];
            const mcgpu_result r = slot_bind(srv, c, &cuts[i], on_device);
            if (r != MCGPU_SUCCESS) {
                out[i].status = (int32_t)r;
                out[i].done_s = now();
                ++finished;
                continue;
            }
            slot_enqueue(c, srv->done[s]);
            reruns[s] = 0;
            busy[s] = i;
            held[s] = est[i];
            inflight += est[i];
            progressed = true;
        }
        if (!progressed) {
            bool any = false;
            for (size_t s = 0; s < S; ++s) any = any || busy[s] >= 0;
            if (!any && waiting.empty() && next < n) {
                const double wait = arrival_s[order[next]] - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(10));
            }
        }
    }
    return MCGPU_SUCCESS;
}

MCGPU_API mcgpu_result mcgpu_server_get_registry(mcgpu_server srv, uint32_t request, uint32_t* edges, uint32_t* faces,
                                                 double* points, uint64_t capacity)
{
    if (!srv) return MCGPU_ERROR_INVALID_VALUE;
    if (request >= srv->reg_e.size()) return MCGPU_ERROR_NO_RESULT;
    const uint64_t m = std::min<uint64_t>(srv->reg_e[request].size(), capacity);
    if (edges) std::copy(srv->reg_e[request].begin(), srv->reg_e[request].begin() + m, edges);
    if (faces) std::copy(srv->reg_g[request].begin(), srv->reg_g[request].begin() + m, faces);
    if (points) std::copy(srv->reg_p[request].begin(), srv->reg_p[request].begin() + 3 * m, points);
    return MCGPU_SUCCESS;
}

} // extern "C"

// ===========================================================================
// Broad phase without device allocation (mcgpu_broad_phase_noalloc)
// ===========================================================================
// Every array of the broad phase lives in one file-scope device array,
// `scratch`, carved into the regions below for the largest input the entry
// point takes: MCGPU_NOALLOC_MAX_FACES faces per mesh and
// MCGPU_NOALLOC_MAX_PAIRS raw-space pairs.  Each device holds its own copy of
// the array (from module load until the process ends); its address is looked
// up once per device with that device current and kept in g_scratch_ptr.
// Every call on a device shares that device's array, and each call clears
// what must start empty (counters, candidate-face flags, pair hash, the
// sort's group totals): calls that overlap on one device race.
namespace {

constexpr uint64_t kNaFaces = MCGPU_NOALLOC_MAX_FACES;         // per mesh
constexpr uint64_t kNaF = 2 * kNaFaces;                        // both meshes
constexpr uint64_t kNaPairs = MCGPU_NOALLOC_MAX_PAIRS;         // raw-space pairs
constexpr uint64_t kNaTable = 2 * kNaPairs;                    // pair hash slots (a power of two)
constexpr int kNaDepth = 16;                                   // leaf level of a tree of kNaFaces leaves
static_assert((1ull << kNaDepth) == kNaFaces, "tree depth of the largest mesh");
static_assert((kNaTable & (kNaTable - 1)) == 0, "pair hash size is a power of two");

constexpr uint64_t na_align(uint64_t x) { return (x + 255) & ~255ull; }

struct NaLayout {
    uint64_t state, box, rawbox, mcode, mface, mcode2, mface2, cand, cand_list, hist, scan, nodes0, nodes1, raw, pairs,
        phash, total;
};

constexpr NaLayout na_layout()
{
    NaLayout L{};
    uint64_t at = 0;
    auto take = [&at](uint64_t bytes) {
        const uint64_t o = at;
        at = na_align(at + bytes);
        return o;
    };
    L.state = take(sizeof(DevState));
    L.box = take(48 * kNaF);
    L.rawbox = take(24 * kNaF);
    L.mcode = take(4 * kNaF);
    L.mface = take(4 * kNaF);
    L.mcode2 = take(4 * kNaF);
    L.mface2 = take(4 * kNaF);
    L.cand = take(4 * (kNaF + 8));
    L.cand_list = take(4 * (kNaF + 8));
    L.hist = take(4 * radix_hist_elems(kNaF));
    const uint64_t scan_a = scan_tmp_elems(radix_hist_elems(kNaF)), scan_b = scan_tmp_elems(kNaF + 8);
    L.scan = take(4 * ((scan_a > scan_b ? scan_a : scan_b) + 1024));
    L.nodes0 = take(4 * 6 * ((2ull << kNaDepth) - 1));
    L.nodes1 = take(4 * 6 * ((2ull << kNaDepth) - 1));
    L.raw = take(8 * kNaPairs);
    L.pairs = take(8 * kNaPairs);
    L.phash = take(8 * kNaTable);
    L.total = at;
    return L;
}

constexpr NaLayout kNa = na_layout();
static_assert(kNa.total == MCGPU_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

} // namespace

static __device__ __align__(256) unsigned char scratch[kNa.total];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kNaMaxDevices] = {};

// the device address of `scratch` on the current device (set up once per
// device; the array's name is a host-side stand-in, never a device address)
unsigned char* scratch_on(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch_ptr[device] = (unsigned char*)ptr;
    }
    return g_scratch_ptr[device];
}

} // namespace

extern "C" {

MCGPU_API mcgpu_result mcgpu_broad_phase_noalloc(int device, const double* src_boxes, uint32_t num_src,
                                                 const double* cut_boxes, uint32_t num_cut, uint64_t* pairs,
                                                 uint64_t capacity, uint64_t* num_pairs, void* stream)
{
    if (!src_boxes || !cut_boxes || !pairs || !num_pairs || num_src == 0 || num_cut == 0) return MCGPU_ERROR_INVALID_VALUE;
    if (num_src > kNaFaces || num_cut > kNaFaces) return MCGPU_ERROR_INVALID_VALUE; // beyond the scratch's maximum
    *num_pairs = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kNaMaxDevices)
        return MCGPU_ERROR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return MCGPU_ERROR_DEVICE;
    unsigned char* base = scratch_on(device);
    if (!base) return MCGPU_ERROR_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    const uint64_t F = (uint64_t)num_src + num_cut;
    DevState* st = (DevState*)(base + kNa.state);
    Counters* C = &st->C;
    double* box = (double*)(base + kNa.box);
    float* rawbox = (float*)(base + kNa.rawbox);
    uint32_t* mcode = (uint32_t*)(base + kNa.mcode);
    uint32_t* mface = (uint32_t*)(base + kNa.mface);
    uint32_t* cand = (uint32_t*)(base + kNa.cand);
    unsigned long long* raw = (unsigned long long*)(base + kNa.raw);
    unsigned long long* found = (unsigned long long*)(base + kNa.pairs);
    unsigned long long* phash = (unsigned long long*)(base + kNa.phash);

    fill_u32(C, 0u, sizeof(Counters) / 4, s);
    if (hipMemcpyAsync(box, src_boxes, 48ull * num_src, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(box + 6ull * num_src, cut_boxes, 48ull * num_cut, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return MCGPU_ERROR_DEVICE;
    const uint32_t nf[2] = {num_src, num_cut};
    for (int m = 0; m < 2; ++m) {
        const uint64_t o = m ? num_src : 0;
        fill_u64(C->rbox[m], ~0ull, 3, s);
        k_box_bounds<<<grid_for(nf[m], kBlock, 64), kBlock, 0, s>>>(box + 6 * o, nf[m], C->rbox[m]);
        k_morton_boxes<<<grid_for(nf[m]), kBlock, 0, s>>>(box + 6 * o, nf[m], C->rbox[m], mcode + o, mface + o,
                                                         rawbox + 6 * o, m ? (1u << 21) : 0u);
    }
    BroadView B{};
    B.mcode = mcode;
    B.mface = mface;
    B.mcode2 = (uint32_t*)(base + kNa.mcode2);
    B.mface2 = (uint32_t*)(base + kNa.mface2);
    B.rawbox = rawbox;
    B.nodes[0] = (float*)(base + kNa.nodes0);
    B.nodes[1] = (float*)(base + kNa.nodes1);
    for (int m = 0; m < 2; ++m) {
        int D = 0;
        while ((1ull << D) < nf[m]) ++D;
        B.nf[m] = nf[m];
        B.depth[m] = D;
    }
    B.rt = RadixTemp{(uint32_t*)(base + kNa.hist), (uint32_t*)(base + kNa.scan)};
    B.raw_pairs = raw;
    B.cap_pairs = kNaPairs;
    B.sort_bits = 22;
    enqueue_trees_traverse(B, s, C, false, nullptr, nullptr);
    k_pair_hash_init<<<grid_for(kNaTable, kBlock, 1024), kBlock, 0, s>>>(phash, kNaPairs, kNaTable, C);
    fill_u32(cand, 0u, F, s);
    k_filter_pairs<<<grid_for(kNaPairs, kFilterTile, 2048), kStageBlock, 0, s>>>(
        box, raw, kNaPairs, num_src, nullptr, found, phash, cand, (uint32_t*)(base + kNa.cand_list), 0, C);
    Counters h;
    if (hipMemcpyAsync(&h, C, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
        return MCGPU_ERROR_DEVICE;
    if (h.ov_pairs) { // more raw-space pairs than the scratch holds
        *num_pairs = h.n_raw;
        return MCGPU_ERROR_OUT_OF_MEMORY;
    }
    *num_pairs = h.n_pairs;
    if (h.n_pairs > capacity) return MCGPU_ERROR_OUT_OF_MEMORY;
    if (h.n_pairs && (hipMemcpyAsync(pairs, found, 8ull * h.n_pairs, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                      hipStreamSynchronize(s) != hipSuccess))
        return MCGPU_ERROR_DEVICE;
    return MCGPU_SUCCESS;
}

} // extern "C"

extern "C" {

MCGPU_API mcgpu_result mcgpu_dispatch_narrow(mcgpu_context ctx, const mcgpu_mesh* src, const mcgpu_mesh* cut,
                                             const double* src_boxes, const double* cut_boxes, const uint32_t* pairs,
                                             uint64_t num_pairs, const mcgpu_soup_topology* topology,
                                             const mcgpu_dispatch_desc* desc, mcgpu_frontend_info* info)
{
    if (ctx && ctx->pending) return MCGPU_ERROR_INVALID_VALUE; // an enqueued dispatch not yet waited for
    if (!ctx || !desc || !src_boxes || !cut_boxes || (num_pairs && !pairs)) return MCGPU_ERROR_INVALID_VALUE;
    if (topology && (!topology->corner_edges[0] || !topology->corner_edges[1] || !topology->edge_v0 ||
                     !topology->edge_v1 || !topology->edge_f0 || !topology->edge_f1 || topology->num_edges == 0))
        return MCGPU_ERROR_INVALID_VALUE;
    if (desc->input_on_device) return MCGPU_ERROR_INVALID_VALUE; // host arrays, as dispatch() holds them
    mcgpu_result r;
    if ((r = check_mesh(src)) != MCGPU_SUCCESS || (r = check_mesh(cut)) != MCGPU_SUCCESS) return r;
    for (uint64_t i = 0; i < num_pairs; ++i)
        if (pairs[2 * i] >= src->num_faces || pairs[2 * i + 1] >= cut->num_faces) return MCGPU_ERROR_INVALID_VALUE;
    CK(hipSetDevice(ctx->device));
    ctx->have = false;
    ctx->broad_only = false;
    ctx->batch = false;
    ctx->shifted_bp = false;
    ctx->allow_loose = false; // the caller's pairs: no broad phase runs here
    ctx->cc_check = false;
    ctx->sort_bits = 22;
    ctx->flags = desc->flags;
    ctx->gp_constant = desc->gp_constant > 0.0 ? desc->gp_constant : 1e-4;
    ctx->enforcement_count = desc->enforcement_count;
    ctx->keep_tests = desc->keep_tests ? 1 : 0;
    if ((r = upload_mesh(ctx, 0, src, false)) != MCGPU_SUCCESS) return r;
    if ((r = upload_mesh(ctx, 1, cut, false)) != MCGPU_SUCCESS) return r;
    const uint32_t Fs = ctx->nf[0];
    const uint64_t F = (uint64_t)Fs + ctx->nf[1];
    const uint64_t K = ctx->ncorner[0] + ctx->ncorner[1];
    const uint64_t NE = topology ? topology->num_edges : K;
    if (topology) { // every corner names an edge of the table, every edge soup vertices and faces
        const uint64_t nvs = (uint64_t)ctx->nv[0] + ctx->nv[1];
        for (int m = 0; m < 2; ++m)
            for (uint64_t c = 0; c < ctx->ncorner[m]; ++c)
                if (topology->corner_edges[m][c] >= NE) return MCGPU_ERROR_INVALID_VALUE;
        for (uint64_t e = 0; e < NE; ++e)
            if (topology->edge_v0[e] >= nvs || topology->edge_v1[e] >= nvs || topology->edge_f0[e] >= (int64_t)F ||
                topology->edge_f1[e] >= (int64_t)F || topology->edge_f0[e] < -1 || topology->edge_f1[e] < -1)
                return MCGPU_ERROR_INVALID_VALUE;
    }
    ctx->gbits = bits_for(F + 1);
    ctx->kbits = ctx->gbits + bits_for(std::max<uint64_t>(K, NE) + 1);
    if (ctx->kbits > 64) return MCGPU_ERROR_UNSUPPORTED;
    if ((r = alloc_mesh_buffers(ctx)) != MCGPU_SUCCESS) return r;
    if (topology) {
        uint32_t* u;
        int32_t* i32;
        CK(ensure(ctx->ev0, NE, &u));
        CK(ensure(ctx->ev1, NE, &u));
        CK(ensure(ctx->ef0, NE, &i32));
        CK(ensure(ctx->ef1, NE, &i32));
    }
    const uint64_t np = std::max<uint64_t>(num_pairs, 1);
    if (ctx->cap_pairs < np) ctx->cap_pairs = np + np / 4;
    if (ctx->cap_pairs < (1u << 14)) ctx->cap_pairs = 1u << 14;
    if (ctx->cap_tests < 2 * ctx->cap_pairs) ctx->cap_tests = 2 * ctx->cap_pairs;
    if (ctx->cap_esc < (1u << 14)) ctx->cap_esc = std::max<uint64_t>(1 << 14, ctx->cap_tests / 16);
    if (ctx->cap_reg < (1u << 14))
        ctx->cap_reg = std::min<uint64_t>(std::max<uint64_t>(1 << 14, ctx->cap_tests / 8), kRegMergeTiles * kRegTile);
    if (ctx->cap_big < 1024) ctx->cap_big = 1024;
    if (ctx->cap_trip < (1u << 16)) ctx->cap_trip = 1u << 16;
    std::vector<unsigned long long> keys(num_pairs);
    for (uint64_t i = 0; i < num_pairs; ++i) keys[i] = ((unsigned long long)pairs[2 * i] << 32) | pairs[2 * i + 1];
    for (int attempt = 0; attempt < 6; ++attempt) {
        if ((r = alloc_list_buffers(ctx)) != MCGPU_SUCCESS) return r;
        hipStream_t s = ctx->s;
        ctx->stage_mode = 1;
        enqueue_validate(ctx); // counters, face offsets, face-list validation
        DevState* st = P_<DevState>(ctx->state);
        Counters* C = &st->C;
        fill_u32(&st->P, 0u, sizeof(Params) / 4, s); // no shift: the coordinates are the soup's own
        // the soup's coordinates, face boxes and candidate pairs as given
        CK(hipMemcpyAsync(ctx->ps_v.p, ctx->dv[0], 24ull * ctx->nv[0], hipMemcpyDeviceToDevice, s));
        CK(hipMemcpyAsync(P_<double>(ctx->ps_v) + 3ull * ctx->nv[0], ctx->dv[1], 24ull * ctx->nv[1],
                          hipMemcpyDeviceToDevice, s));
        CK(hipMemcpyAsync(ctx->box.p, src_boxes, 48ull * Fs, hipMemcpyHostToDevice, s));
        CK(hipMemcpyAsync(P_<double>(ctx->box) + 6ull * Fs, cut_boxes, 48ull * ctx->nf[1], hipMemcpyHostToDevice, s));
        if (num_pairs)
            CK(hipMemcpyAsync(ctx->raw_pairs.p, keys.data(), 8ull * num_pairs, hipMemcpyHostToDevice, s));
        fill_u64(&C->n_raw, (unsigned long long)num_pairs, 1, s);
        if (topology) { // the caller's edge numbering
            for (int m = 0; m < 2; ++m)
                CK(hipMemcpyAsync(ctx->cedge[m].p, topology->corner_edges[m], 4ull * ctx->ncorner[m],
                                  hipMemcpyHostToDevice, s));
            CK(hipMemcpyAsync(ctx->ev0.p, topology->edge_v0, 4ull * NE, hipMemcpyHostToDevice, s));
            CK(hipMemcpyAsync(ctx->ev1.p, topology->edge_v1, 4ull * NE, hipMemcpyHostToDevice, s));
            CK(hipMemcpyAsync(ctx->ef0.p, topology->edge_f0, 4ull * NE, hipMemcpyHostToDevice, s));
            CK(hipMemcpyAsync(ctx->ef1.p, topology->edge_f1, 4ull * NE, hipMemcpyHostToDevice, s));
            fill_u64(&C->n_edges_src, (unsigned long long)NE, 1, s);
            fill_u64(&C->n_edges_cut, 0ull, 1, s);
            fill_u32(ctx->cand.p, 0u, F, s);
        } else {
            enqueue_topology(ctx, s);
        }
        // pair hash and candidate faces; the exact box test keeps every given pair
        const uint64_t table = next_pow2(2 * ctx->cap_pairs);
        k_pair_hash_init<<<grid_for(table, kBlock, 1024), kBlock, 0, s>>>(P_<unsigned long long>(ctx->phash),
                                                                          ctx->cap_pairs, table, C);
        k_filter_pairs<<<grid_for(ctx->cap_pairs, kFilterTile, 2048), kStageBlock, 0, s>>>(
            P_<double>(ctx->box), P_<unsigned long long>(ctx->raw_pairs), ctx->cap_pairs, Fs, nullptr,
            P_<unsigned long long>(ctx->pairs), P_<unsigned long long>(ctx->phash), P_<uint32_t>(ctx->cand),
            P_<uint32_t>(ctx->cand_list), 0, C);
        stage_event(ctx, EV_A_FILTER, s);
        enqueue_narrow(ctx, make_soup(ctx), false);
        bool overflow = false;
        if ((r = finish(ctx, info, &overflow)) != MCGPU_SUCCESS) return r;
        if (!overflow) return MCGPU_SUCCESS;
    }
    return MCGPU_ERROR_OUT_OF_MEMORY;
}

} // extern "C"
