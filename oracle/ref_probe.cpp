// Test-only probe into the reference MCUT library compiled from its own
// sources (oracle/Makefile builds it into oracle/_ref/).  This file is test
// infrastructure: it is never linked into the product library.
//
// It exposes, through a plain C ABI for the Python tests:
//   * the public API end to end (mcCreateContext[WithHelpers] + mcDispatch +
//     mcGetConnectedComponents/Data, include/mcut/mcut.h:567-1378), keeping
//     every connected component's vertices, faces, type and seam vertices;
//   * the reference's broad phase on its own (calculate_vertex_parameters,
//     preproc.cpp:2124; build_oibvh, bvh.cpp:219; intersectOIBVHs, bvh.cpp:638)
//     applied to halfedge meshes built with the reference's hmesh_t;
//   * the reference's narrow-phase building blocks one call at a time
//     (math.cpp:99-117,130-239,249-287,391-427,851-902; shewchuk.c:1695,2367).
#include "mcut/mcut.h"
#include "mcut/internal/bvh.h"
#include "mcut/internal/hmesh.h"
#include "mcut/internal/math.h"
#include "mcut/internal/tpool.h"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

// preproc.cpp:2124 (not declared in any reference header)
bool calculate_vertex_parameters(double& quantization_multiplier,
                                 vec3_<double>& pre_quantization_translation,
                                 vec3_<double>& srcmesh_cutmesh_com,
                                 vec3_<double>& srcmesh_bboxmin,
                                 vec3_<double>& srcmesh_bboxmax,
                                 vec3_<double>& cutmesh_bboxmin,
                                 vec3_<double>& cutmesh_bboxmax,
                                 McFlags dispatchFlags,
                                 const void* pSrcMeshVertices,
                                 McUint32 numSrcMeshVertices,
                                 const void* pCutMeshVertices,
                                 McUint32 numCutMeshVertices);

// math.cpp:289 (not declared in any reference header)
bool determine_three_noncollinear_vertices(int& i, int& j, int& k, const std::vector<vec3>& polygon_vertices,
                                           const vec3& polygon_normal, const int polygon_normal_largest_component,
                                           const double multiplier);

namespace {

struct cc_data {
    uint32_t type = 0;
    std::vector<double> vertices;
    std::vector<uint32_t> faces;
    std::vector<uint32_t> face_sizes;
    std::vector<uint32_t> seam;
    std::vector<uint32_t> vmap;
    double perturbation[3] = {0, 0, 0};
};

struct dispatch_result {
    int status = 0;
    double ms = 0;
    std::vector<cc_data> ccs;
};

dispatch_result g_dispatch;

struct broad_result {
    double com[3], t[3], eps;
    std::vector<uint32_t> pairs; // (src face, cut face) flattened, map order
    std::vector<double> src_face_aabb, cut_face_aabb; // per face: min xyz, max xyz
};

broad_result g_broad;

template <typename T>
static void query(McContext ctx, McConnectedComponent cc, McFlags what, std::vector<T>& out)
{
    McSize n = 0;
    out.clear();
    if (mcGetConnectedComponentData(ctx, cc, what, 0, NULL, &n) != MC_NO_ERROR || n == 0)
        return;
    out.resize(n / sizeof(T));
    mcGetConnectedComponentData(ctx, cc, what, n, out.data(), NULL);
}

static void build_hmesh(hmesh_t& m, const double* v, const uint32_t* f, const uint32_t* sz, uint32_t nv,
                        uint32_t nf, const vec3_<double>& com, const vec3_<double>& t)
{
    for (uint32_t i = 0; i < nv; ++i) {
        // client_input_arrays_to_hmesh, preproc.cpp:244-276 (unperturbed: "+ 0.")
        const double x = (v[3 * i + 0] - com[0]) + t[0];
        const double y = (v[3 * i + 1] - com[1]) + t[1];
        const double z = (v[3 * i + 2] - com[2]) + t[2];
        m.add_vertex(double(x) + double(0.), double(y) + double(0.), double(z) + double(0.));
    }
    uint32_t off = 0;
    std::vector<vd_t> fv;
    for (uint32_t i = 0; i < nf; ++i) {
        const uint32_t k = sz ? sz[i] : 3;
        fv.clear();
        for (uint32_t j = 0; j < k; ++j)
            fv.push_back(vd_t(f[off + j]));
        m.add_face(fv);
        off += k;
    }
}

} // namespace

// --- public API, end to end ------------------------------------------------

// the kernel's reports to the debug callback, one a line: preproc.cpp:2947-2955
// passes "<status_t name> : <reason>" before it throws, and the
// general-position loop reports each violation (preproc.cpp:2596-2610)
std::string g_kernel_message;

static void MCAPI_PTR on_debug_message(McDebugSource source, McDebugType, unsigned int, McDebugSeverity, size_t,
                                       const char* message, const McVoid*)
{
    if (source == MC_DEBUG_SOURCE_KERNEL && message)
        g_kernel_message += std::string(message) + "\n";
}

static int dispatch_impl(uint32_t flags, int helpers, bool debug,
                         const double* sv, const uint32_t* sf, const uint32_t* ssz, uint32_t nsv, uint32_t nsf,
                         const double* cv, const uint32_t* cf, const uint32_t* csz, uint32_t ncv, uint32_t ncf)
{
    g_dispatch = dispatch_result();
    g_kernel_message.clear();
    McContext ctx = MC_NULL_HANDLE;
    const McFlags cflags = debug ? MC_DEBUG : MC_NULL_HANDLE;
    McResult r = helpers > 0 ? mcCreateContextWithHelpers(&ctx, cflags, (uint32_t)helpers)
                             : mcCreateContext(&ctx, cflags);
    if (r != MC_NO_ERROR)
        return -1000 - (int)r;
    if (debug) {
        mcDebugMessageCallback(ctx, on_debug_message, nullptr);
        mcDebugMessageControl(ctx, MC_DEBUG_SOURCE_ALL, MC_DEBUG_TYPE_ALL, MC_DEBUG_SEVERITY_ALL, true);
    }
    auto t0 = std::chrono::steady_clock::now();
    r = mcDispatch(ctx, flags, sv, sf, ssz, nsv, nsf, cv, cf, csz, ncv, ncf);
    auto t1 = std::chrono::steady_clock::now();
    g_dispatch.ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    g_dispatch.status = (int)r;
    if (r == MC_NO_ERROR) {
        uint32_t n = 0;
        mcGetConnectedComponents(ctx, MC_CONNECTED_COMPONENT_TYPE_ALL, 0, NULL, &n);
        std::vector<McConnectedComponent> ccs(n);
        if (n)
            mcGetConnectedComponents(ctx, MC_CONNECTED_COMPONENT_TYPE_ALL, n, ccs.data(), NULL);
        for (uint32_t i = 0; i < n; ++i) {
            cc_data d;
            McSize nb = 0;
            mcGetConnectedComponentData(ctx, ccs[i], MC_CONNECTED_COMPONENT_DATA_TYPE, sizeof(uint32_t), &d.type, &nb);
            query(ctx, ccs[i], MC_CONNECTED_COMPONENT_DATA_VERTEX_DOUBLE, d.vertices);
            query(ctx, ccs[i], MC_CONNECTED_COMPONENT_DATA_FACE, d.faces);
            query(ctx, ccs[i], MC_CONNECTED_COMPONENT_DATA_FACE_SIZE, d.face_sizes);
            if (d.type != MC_CONNECTED_COMPONENT_TYPE_INPUT)
                query(ctx, ccs[i], MC_CONNECTED_COMPONENT_DATA_SEAM_VERTEX, d.seam);
            if (flags & MC_DISPATCH_INCLUDE_VERTEX_MAP)
                query(ctx, ccs[i], MC_CONNECTED_COMPONENT_DATA_VERTEX_MAP, d.vmap);
            mcGetConnectedComponentData(ctx, ccs[i], MC_CONNECTED_COMPONENT_DATA_DISPATCH_PERTURBATION_VECTOR,
                                        3 * sizeof(double), d.perturbation, NULL);
            g_dispatch.ccs.push_back(std::move(d));
        }
        if (n)
            mcReleaseConnectedComponents(ctx, n, ccs.data());
    }
    mcReleaseContext(ctx);
    return g_dispatch.status;
}

extern "C" {

// --- public API, end to end ------------------------------------------------


int ref_dispatch(uint32_t flags, int helpers,
                 const double* sv, const uint32_t* sf, const uint32_t* ssz, uint32_t nsv, uint32_t nsf,
                 const double* cv, const uint32_t* cf, const uint32_t* csz, uint32_t ncv, uint32_t ncf)
{
    return dispatch_impl(flags, helpers, false, sv, sf, ssz, nsv, nsf, cv, cf, csz, ncv, ncf);
}

// mcDispatch in a context created with MC_DEBUG and a debug callback that
// keeps the kernel's failure report (ref_kernel_message)
int ref_dispatch_debug(uint32_t flags,
                       const double* sv, const uint32_t* sf, const uint32_t* ssz, uint32_t nsv, uint32_t nsf,
                       const double* cv, const uint32_t* cf, const uint32_t* csz, uint32_t ncv, uint32_t ncf)
{
    return dispatch_impl(flags, 0, true, sv, sf, ssz, nsv, nsf, cv, cf, csz, ncv, ncf);
}

const char* ref_kernel_message() { return g_kernel_message.c_str(); }

double ref_dispatch_ms() { return g_dispatch.ms; }
uint32_t ref_cc_count() { return (uint32_t)g_dispatch.ccs.size(); }

// sizes[0..4] = #vertex doubles, #face indices, #faces, #seam vertices, #vmap entries; returns type
uint32_t ref_cc_sizes(uint32_t i, uint64_t* sizes)
{
    const cc_data& d = g_dispatch.ccs.at(i);
    sizes[0] = d.vertices.size();
    sizes[1] = d.faces.size();
    sizes[2] = d.face_sizes.size();
    sizes[3] = d.seam.size();
    sizes[4] = d.vmap.size();
    return d.type;
}

void ref_cc_copy(uint32_t i, double* vertices, uint32_t* faces, uint32_t* face_sizes, uint32_t* seam,
                 uint32_t* vmap, double* perturbation)
{
    const cc_data& d = g_dispatch.ccs.at(i);
    if (vertices) std::memcpy(vertices, d.vertices.data(), d.vertices.size() * sizeof(double));
    if (faces) std::memcpy(faces, d.faces.data(), d.faces.size() * sizeof(uint32_t));
    if (face_sizes) std::memcpy(face_sizes, d.face_sizes.data(), d.face_sizes.size() * sizeof(uint32_t));
    if (seam) std::memcpy(seam, d.seam.data(), d.seam.size() * sizeof(uint32_t));
    if (vmap) std::memcpy(vmap, d.vmap.data(), d.vmap.size() * sizeof(uint32_t));
    if (perturbation) std::memcpy(perturbation, d.perturbation, sizeof(d.perturbation));
}

// --- broad phase only --------------------------------------------------------

// Runs the reference's shift computation, hmesh construction, both OIBVH builds
// (cut mesh enlarged by eps = |cut bbox diagonal| * gp_constant, preproc.cpp:2667-2693)
// and the BVH x BVH traversal.  Returns the number of (src, cut) candidate pairs.
// timing_ms (optional, 3 entries): build src, build cut, traversal.
int64_t ref_broad_phase(int helpers, double gp_constant,
                        const double* sv, const uint32_t* sf, const uint32_t* ssz, uint32_t nsv, uint32_t nsf,
                        const double* cv, const uint32_t* cf, const uint32_t* csz, uint32_t ncv, uint32_t ncf,
                        double* timing_ms)
{
    g_broad = broad_result();
    double multiplier = 1;
    vec3_<double> t(0.0), com(0.0), sbmin, sbmax, cbmin, cbmax;
    if (!calculate_vertex_parameters(multiplier, t, com, sbmin, sbmax, cbmin, cbmax, MC_DISPATCH_VERTEX_ARRAY_DOUBLE,
                                     sv, nsv, cv, ncv))
        return -1;
    hmesh_t sm, cm;
    build_hmesh(sm, sv, sf, ssz, nsv, nsf, com, t);
    build_hmesh(cm, cv, cf, csz, ncv, ncf, com, t);
    const double diag = length(cbmax - cbmin, 1); // preproc.cpp:2518
    const double eps = diag * gp_constant;        // preproc.cpp:2667-2676

    thread_pool pool((uint32_t)helpers, 1);
    std::vector<bounding_box_t<vec3_<double>>> s_nodes, c_nodes, s_faces, c_faces;
    std::vector<fd_t> s_leaves, c_leaves;
    auto t0 = std::chrono::steady_clock::now();
    build_oibvh(pool, sm, s_nodes, s_leaves, s_faces, 0.0, multiplier);
    auto t1 = std::chrono::steady_clock::now();
    build_oibvh(pool, cm, c_nodes, c_leaves, c_faces, eps, multiplier);
    auto t2 = std::chrono::steady_clock::now();
    std::map<fd_t, std::vector<fd_t>> cand;
    intersectOIBVHs(cand, s_nodes, s_leaves, c_nodes, c_leaves);
    auto t3 = std::chrono::steady_clock::now();
    if (timing_ms) {
        timing_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        timing_ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
        timing_ms[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
    }
    for (int j = 0; j < 3; ++j) {
        g_broad.com[j] = com[j];
        g_broad.t[j] = t[j];
    }
    g_broad.eps = eps;
    for (auto& kv : cand) {
        if ((uint32_t)kv.first >= nsf)
            continue; // keys of the cut mesh repeat the same pairs mirrored
        for (fd_t c : kv.second) {
            g_broad.pairs.push_back((uint32_t)kv.first);
            g_broad.pairs.push_back((uint32_t)c - nsf);
        }
    }
    for (auto* pr : {&s_faces, &c_faces}) {
        std::vector<double>& dst = (pr == &s_faces) ? g_broad.src_face_aabb : g_broad.cut_face_aabb;
        for (auto& b : *pr) {
            for (int j = 0; j < 3; ++j) dst.push_back(b.minimum()[j]);
            for (int j = 0; j < 3; ++j) dst.push_back(b.maximum()[j]);
        }
    }
    return (int64_t)(g_broad.pairs.size() / 2);
}

void ref_broad_copy(double* com, double* t, double* eps, uint32_t* pairs, double* src_face_aabb,
                    double* cut_face_aabb)
{
    if (com) std::memcpy(com, g_broad.com, sizeof(g_broad.com));
    if (t) std::memcpy(t, g_broad.t, sizeof(g_broad.t));
    if (eps) *eps = g_broad.eps;
    if (pairs) std::memcpy(pairs, g_broad.pairs.data(), g_broad.pairs.size() * sizeof(uint32_t));
    if (src_face_aabb)
        std::memcpy(src_face_aabb, g_broad.src_face_aabb.data(), g_broad.src_face_aabb.size() * sizeof(double));
    if (cut_face_aabb)
        std::memcpy(cut_face_aabb, g_broad.cut_face_aabb.data(), g_broad.cut_face_aabb.size() * sizeof(double));
}

// --- narrow-phase building blocks ------------------------------------------

static std::vector<vec3> to_vec3(const double* p, int n)
{
    std::vector<vec3> out;
    for (int i = 0; i < n; ++i)
        out.push_back(vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
    return out;
}

int ref_plane(const double* verts, int n, double* normal, double* d)
{
    std::vector<vec3> v = to_vec3(verts, n);
    vec3 nn(0.0);
    double dd = 0;
    int k = compute_polygon_plane_coefficients(nn, dd, v.data(), n, 1.0);
    for (int j = 0; j < 3; ++j) normal[j] = nn[j];
    *d = dd;
    return k;
}

int ref_segment_type(const double* q, const double* r, const double* verts, int n, const double* normal, int k)
{
    std::vector<vec3> v = to_vec3(verts, n);
    return compute_segment_plane_intersection_type(vec3(q[0], q[1], q[2]), vec3(r[0], r[1], r[2]), v,
                                                   vec3(normal[0], normal[1], normal[2]), k, 1.0);
}

int ref_segment_point(const double* normal, double d, const double* q, const double* r, double* p)
{
    vec3 pp(0., 0., 0.);
    char c = compute_segment_plane_intersection(pp, vec3(normal[0], normal[1], normal[2]), d,
                                                vec3(q[0], q[1], q[2]), vec3(r[0], r[1], r[2]));
    for (int j = 0; j < 3; ++j) p[j] = pp[j];
    return c;
}

int ref_point_in_polygon(const double* p, const double* verts, int n, const double* normal, int k)
{
    std::vector<vec3> v = to_vec3(verts, n);
    return compute_point_in_polygon_test(vec3(p[0], p[1], p[2]), v, vec3(normal[0], normal[1], normal[2]), k, 1.0);
}

double ref_orient3d(const double* a, const double* b, const double* c, const double* d)
{
    return ::orient3d(a, b, c, d);
}

double ref_orient2d(const double* a, const double* b, const double* c) { return ::orient2d(a, b, c); }

// Polygon soup as dispatch() assembles it (kernel.cpp:1593-1732): a copy of
// the source hmesh, then the cut mesh's vertices and, face by face, the cut
// hmesh's vertex lists offset by the source vertex count.  Writes per edge
// the endpoints of halfedge 0 and the faces of halfedges 0 and 1 (-1 = none);
// returns the edge count (-1 if buffers are too small).
static std::vector<int32_t> g_ps_edges;

int64_t ref_ps_edges(const double* sv, const uint32_t* sf, const uint32_t* ssz, uint32_t nsv, uint32_t nsf,
                     const double* cv, const uint32_t* cf, const uint32_t* csz, uint32_t ncv, uint32_t ncf)
{
    hmesh_t sm, cs;
    const vec3_<double> zero(0.0);
    build_hmesh(sm, sv, sf, ssz, nsv, nsf, zero, zero);
    build_hmesh(cs, cv, cf, csz, ncv, ncf, zero, zero);
    hmesh_t ps = sm;
    for (auto i = cs.vertices_begin(); i != cs.vertices_end(); ++i) ps.add_vertex(cs.vertex(*i));
    std::vector<vd_t> fv;
    for (auto i = cs.faces_begin(); i != cs.faces_end(); ++i) {
        cs.get_vertices_around_face(fv, *i, nsv);
        ps.add_face(fv);
    }
    g_ps_edges.clear();
    for (auto e = ps.edges_begin(); e != ps.edges_end(); ++e) {
        const hd_t h0 = ps.halfedge(*e, 0), h1 = ps.halfedge(*e, 1);
        const fd_t f0 = ps.face(h0), f1 = ps.face(h1);
        g_ps_edges.push_back((int32_t)ps.source(h0));
        g_ps_edges.push_back((int32_t)ps.target(h0));
        g_ps_edges.push_back(f0 == hmesh_t::null_face() ? -1 : (int32_t)f0);
        g_ps_edges.push_back(f1 == hmesh_t::null_face() ? -1 : (int32_t)f1);
    }
    return (int64_t)(g_ps_edges.size() / 4);
}

void ref_ps_edges_copy(int32_t* out) { std::memcpy(out, g_ps_edges.data(), g_ps_edges.size() * sizeof(int32_t)); }

// determine_three_noncollinear_vertices, math.cpp:289-389
int ref_best_triple(const double* verts, int n, const double* normal, int k, int* ijk)
{
    std::vector<vec3> v = to_vec3(verts, n);
    int i = 0, j = 1, kk = 2;
    bool ok = determine_three_noncollinear_vertices(i, j, kk, v, vec3(normal[0], normal[1], normal[2]), k, 1.0);
    ijk[0] = i, ijk[1] = j, ijk[2] = kk;
    return ok ? 1 : 0;
}

} // extern "C"
