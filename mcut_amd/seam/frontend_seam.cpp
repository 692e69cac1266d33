// MCUT's intersection front end inside dispatch(), on libmcgpu (see
// frontend_seam.h).  Compiled with MCUT's own sources into the seam build.
#include "frontend_seam.h"

#include <algorithm>
#include <cstdint>
#include <future>
#include <mutex>
#include <stdexcept>
#include <string>

#include "mcgpu.h"
#include "mcut/internal/tpool.h"

#if !defined(MCUT_WITH_COMPUTE_HELPER_THREADPOOL)
#error "the seam follows the registry order of MCUT's thread-pool build"
#endif

namespace {

// One libmcgpu context for the process: MCUT runs each context's dispatches on
// that context's API thread, and dispatches of several MCUT contexts take
// turns on it (a libmcgpu context is not thread-safe).
std::mutex g_mutex;
mcgpu_context g_ctx = nullptr;

mcgpu_context seam_context()
{
    if (!g_ctx) {
        const mcgpu_result r = mcgpu_create_context(0, &g_ctx);
        if (r != MCGPU_SUCCESS) throw std::runtime_error(std::string("mcgpu_create_context: ") + mcgpu_result_string(r));
    }
    return g_ctx;
}

McgpuSeamCutpath& seam_cutpath()
{
    thread_local McgpuSeamCutpath cp;
    return cp;
}

void check(mcgpu_result r, const char* what)
{
    if (r != MCGPU_SUCCESS) throw std::runtime_error(std::string(what) + ": " + mcgpu_result_string(r));
}

// One mesh of the polygon soup as libmcgpu takes it: coordinates by vertex,
// each face's vertices in the order a caller of mcDispatch would have listed
// them, and the soup's own edge of every corner.  The soup lists a source
// face from its second input vertex (hmesh.cpp:705-733) and a cut face from
// its third (the cut mesh's list, inserted again, kernel.cpp:1667-1727), so the
// input order is the soup's list rotated back by one or two places; corner i
// (input vertex i to i + 1) is the face's halfedge with that source and
// target, which the seam checks.
struct MeshArrays {
    std::vector<double> v;
    std::vector<uint32_t> f, sz, cedge;
};

void soup_mesh(hmesh_t& ps, int f_begin, int f_end, int v_begin, int v_end, int rotate, MeshArrays& a)
{
    a.v.resize(3 * (size_t)(v_end - v_begin));
    for (int i = v_begin; i < v_end; ++i) {
        const vec3& p = ps.vertex(vd_t((vd_t::index_type)i));
        const size_t k = (size_t)(i - v_begin);
        a.v[3 * k] = p[0];
        a.v[3 * k + 1] = p[1];
        a.v[3 * k + 2] = p[2];
    }
    std::vector<vd_t> L;
    for (int fi = f_begin; fi < f_end; ++fi) {
        const fd_t f((fd_t::index_type)fi);
        ps.get_vertices_around_face(L, f);
        const std::vector<hd_t>& hs = ps.get_halfedges_around_face(f);
        const size_t n = L.size();
        a.sz.push_back((uint32_t)n);
        for (size_t i = 0; i < n; ++i) {
            const vd_t in = L[(i + n - rotate) % n], next = L[(i + 1 + n - rotate) % n];
            a.f.push_back((uint32_t)in - (uint32_t)v_begin);
            const hd_t h = hs[(i + n + 1 - rotate) % n];
            if (ps.source(h) != in || ps.target(h) != next)
                throw std::logic_error("libmcgpu seam: a face's halfedges do not follow its vertex list");
            a.cedge.push_back((uint32_t)ps.edge(h));
        }
    }
}

void flatten(const std::vector<bounding_box_t<vec3_<double>>>& boxes, std::vector<double>& out)
{
    out.resize(6 * boxes.size());
    for (size_t i = 0; i < boxes.size(); ++i)
        for (int j = 0; j < 3; ++j) {
            out[6 * i + j] = boxes[i].minimum()[j];
            out[6 * i + 3 + j] = boxes[i].maximum()[j];
        }
}

} // namespace

bool mcgpu_seam_frontend(const input_t& input, output_t& output, const hmesh_t& sm, const hmesh_t& cs,
                         const hmesh_t& ps_const, hmesh_t& m0,
                         std::unordered_map<fd_t, vec3>& face_normal,
                         std::unordered_map<fd_t, scalar_t>& face_d,
                         std::unordered_map<fd_t, int>& face_max_comp,
                         std::unordered_map<fd_t, std::vector<vec3>>& face_vertices,
                         std::vector<std::pair<ed_t, fd_t>>& registry,
                         std::vector<vd_t>& cm_border_reentrant,
                         std::unordered_map<ed_t, std::vector<vd_t>>& intersecting_edges,
                         std::map<pair<fd_t>, std::vector<vd_t>>& cutpath_edges,
                         std::unordered_map<fd_t, std::vector<vd_t>>& face_ivertices,
                         bool& partial_cut)
{
    seam_cutpath() = McgpuSeamCutpath(); // nothing from an earlier dispatch
    hmesh_t& ps = const_cast<hmesh_t&>(ps_const); // is_border(edge) is not a const member
    logger_t& lg = output.logger;
    const int sm_vtx_cnt = sm.number_of_vertices();
    const int sm_face_count = sm.number_of_faces();
    const int ps_vtx_cnt = ps.number_of_vertices();

    const int cm_face_count = cs.number_of_faces();
    if (ps.number_of_vertices_removed() != 0 || ps.number_of_faces_removed() != 0 ||
        ps.number_of_faces() != sm_face_count + cm_face_count)
        throw std::logic_error("libmcgpu seam: a polygon soup with removed elements");
    MeshArrays A, B;
    soup_mesh(ps, 0, sm_face_count, 0, sm_vtx_cnt, 1, A);
    soup_mesh(ps, sm_face_count, sm_face_count + cm_face_count, sm_vtx_cnt, ps_vtx_cnt, 2, B);
    // the soup's edges: halfedge-0 endpoints and faces (ids past the last
    // live edge, if any were removed, stay unreferenced)
    int ne = 0;
    for (edge_array_iterator_t e = ps.edges_begin(); e != ps.edges_end(); ++e) ne = std::max(ne, (int)*e + 1);
    std::vector<uint32_t> ev0(ne, 0u), ev1(ne, 0u);
    std::vector<int32_t> ef0(ne, -1), ef1(ne, -1);
    for (edge_array_iterator_t e = ps.edges_begin(); e != ps.edges_end(); ++e) {
        const hd_t h0 = ps.halfedge(*e, 0), h1 = ps.halfedge(*e, 1);
        const size_t k = (size_t)*e;
        ev0[k] = (uint32_t)ps.source(h0);
        ev1[k] = (uint32_t)ps.target(h0);
        const fd_t f0 = ps.face(h0), f1 = ps.face(h1);
        ef0[k] = f0 == hmesh_t::null_face() ? -1 : (int32_t)f0;
        ef1[k] = f1 == hmesh_t::null_face() ? -1 : (int32_t)f1;
    }
    mcgpu_soup_topology topo{};
    topo.corner_edges[0] = A.cedge.data();
    topo.corner_edges[1] = B.cedge.data();
    topo.edge_v0 = ev0.data();
    topo.edge_v1 = ev1.data();
    topo.edge_f0 = ef0.data();
    topo.edge_f1 = ef1.data();
    topo.num_edges = (uint32_t)ne;
    std::vector<double> sbox, cbox;
    flatten(*input.source_hmesh_face_aabb_array_ptr, sbox);
    flatten(*input.cut_hmesh_face_aabb_array_ptr, cbox);
    std::vector<uint32_t> pairs;
    for (const auto& kv : *input.ps_face_to_potentially_intersecting_others) {
        if ((int)kv.first >= sm_face_count) break; // keys ascend: the source faces come first
        for (const fd_t& c : kv.second) {
            pairs.push_back((uint32_t)kv.first);
            pairs.push_back((uint32_t)c - (uint32_t)sm_face_count);
        }
    }
    const mcgpu_mesh ms{A.v.data(), A.f.data(), A.sz.data(), (uint32_t)sm_vtx_cnt, (uint32_t)A.sz.size()};
    const mcgpu_mesh mc{B.v.data(), B.f.data(), B.sz.data(), (uint32_t)(ps_vtx_cnt - sm_vtx_cnt), (uint32_t)B.sz.size()};
    mcgpu_dispatch_desc desc{};
    desc.flags = (input.enforce_general_position ? MCGPU_DISPATCH_ENFORCE_GENERAL_POSITION : 0u) |
                 (input.require_looped_cutpaths ? MCGPU_DISPATCH_REQUIRE_THROUGH_CUTS : 0u);
    desc.gp_constant = 1e-4; // not read: the face boxes come enlarged
    desc.enforcement_count = (uint32_t)input.general_position_enforcement_count;

    mcgpu_frontend_info info;
    std::vector<uint32_t> re, rf, pf, v0;
    std::vector<double> rp, pn, pd;
    std::vector<int32_t> pk;
    // the context is held until the cut-path edges of this dispatch are read
    std::unique_lock<std::mutex> lock(g_mutex);
    mcgpu_context ctx = seam_context();
    {
        check(mcgpu_dispatch_narrow(ctx, &ms, &mc, sbox.data(), cbox.data(), pairs.data(), pairs.size() / 2, &topo, &desc,
                                    &info),
              "mcgpu_dispatch_narrow");
        const uint64_t n = info.num_registry, nc = info.num_candidate_faces, nedges = info.num_edges;
        re.resize(n), rf.resize(n), rp.resize(3 * n);
        if (n) check(mcgpu_get_registry(ctx, re.data(), rf.data(), rp.data(), n), "mcgpu_get_registry");
        pf.resize(nc), pn.resize(3 * nc), pd.resize(nc), pk.resize(nc);
        if (nc) check(mcgpu_get_face_planes(ctx, pf.data(), pn.data(), pd.data(), pk.data(), nullptr, nc), "mcgpu_get_face_planes");
        v0.resize(nedges);
        if (nedges) check(mcgpu_get_edges(ctx, v0.data(), nullptr, nullptr, nullptr, nedges), "mcgpu_get_edges");
    }

    // statuses decided inside the replaced block; INVALID_MESH_INTERSECTION
    // and the partial-cut exit are decided by dispatch() itself further on,
    // from the containers filled below
    if (info.status == MCGPU_STATUS_INVALID_SRC_MESH || info.status == MCGPU_STATUS_INVALID_CUT_MESH) {
        lg.set_reason_for_failure(std::string("a candidate face of the ") +
                                  (info.status == MCGPU_STATUS_INVALID_CUT_MESH ? "cut" : "source") +
                                  " mesh is degenerate (has zero area)");
        output.status = (status_t)info.status;
        return true;
    }
    if (info.gp_violation) {
        lg.set_reason_for_failure("general position violation in the narrow phase");
        output.status = status_t::GENERAL_POSITION_VIOLATION;
        return true;
    }

    // candidate-face properties (kernel.cpp:2184-2356)
    std::vector<vd_t> fv;
    for (size_t c = 0; c < pf.size(); ++c) {
        const fd_t f((fd_t::index_type)pf[c]);
        face_normal[f] = vec3(pn[3 * c], pn[3 * c + 1], pn[3 * c + 2]);
        face_d[f] = pd[c];
        face_max_comp[f] = pk[c];
        ps.get_vertices_around_face(fv, f);
        std::vector<vec3>& verts = face_vertices[f];
        for (const vd_t& v : fv) verts.push_back(ps.vertex(v));
    }

    // The order of the registry.  MCUT numbers intersection points in the
    // order its narrow phase finds them, and its cutting stage takes some
    // decisions in that order (which side of a floating polygon is split,
    // how seam vertices are shared), so the same order is rebuilt here with
    // MCUT's own containers and scheduler:
    //   - the edge keys of ps_edge_face_intersection_pairs, an unordered_map
    //     filled block by block from the candidate faces' halfedges and merged
    //     (kernel.cpp:1779-1983), whose iteration order the same insertions
    //     into the same container reproduce;
    //   - the narrow phase walks that map in parallel_for blocks, the last
    //     block on the calling thread, and appends the calling thread's points
    //     first and then the other blocks' in order (kernel.cpp:2415-2878);
    //     within an edge, faces ascend.
    std::vector<size_t> order;
    order.reserve(re.size());
    {
        typedef std::unordered_map<ed_t, std::vector<fd_t>> EdgeMap;
        typedef std::map<fd_t, std::vector<fd_t>>::const_iterator FaceIter;
        EdgeMap edge_map;
        std::vector<std::future<EdgeMap>> futures;
        auto fill = [&](FaceIter block_start, FaceIter block_end) {
            EdgeMap local;
            for (FaceIter it = block_start; it != block_end; ++it)
                for (const hd_t& h : ps.get_halfedges_around_face(it->first)) local[ps.edge(h)];
            return local;
        };
        const std::map<fd_t, std::vector<fd_t>>& cand = *input.ps_face_to_potentially_intersecting_others;
        parallel_for(*input.scheduler, cand.cbegin(), cand.cend(), fill, edge_map, futures);
        for (std::future<EdgeMap>& f : futures) {
            const EdgeMap r = f.get();
            for (const auto& kv : r)
                if (edge_map.find(kv.first) == edge_map.cend()) edge_map[kv.first] = kv.second;
        }
        // registry entries of each edge: a run of the (edge, face)-sorted registry
        std::unordered_map<uint32_t, std::pair<size_t, size_t>> runs;
        for (size_t i = 0; i < re.size();) {
            size_t j = i;
            while (j < re.size() && re[j] == re[i]) ++j;
            runs[re[i]] = std::make_pair(i, j);
            i = j;
        }
        typedef EdgeMap::const_iterator EdgeIter;
        std::vector<std::future<std::vector<size_t>>> blocks;
        std::vector<size_t> master;
        auto collect = [&](EdgeIter block_start, EdgeIter block_end) {
            std::vector<size_t> out;
            for (EdgeIter it = block_start; it != block_end; ++it) {
                const auto r = runs.find((uint32_t)it->first);
                if (r == runs.end()) continue;
                for (size_t k = r->second.first; k < r->second.second; ++k) out.push_back(k);
            }
            return out;
        };
        if (!edge_map.empty()) parallel_for(*input.scheduler, edge_map.cbegin(), edge_map.cend(), collect, master, blocks);
        order.insert(order.end(), master.begin(), master.end());
        for (std::future<std::vector<size_t>>& b : blocks) {
            const std::vector<size_t> r = b.get();
            order.insert(order.end(), r.begin(), r.end());
        }
        if (order.size() != re.size()) throw std::logic_error("libmcgpu seam: a registry edge outside the tested edges");
    }

    // the cut-path edges of this registry in MCUT's point order (kernel.cpp:3427-3615),
    // applied to m0 by the copy of dispatch() (cutpath_block.inc)
    {
        McgpuSeamCutpath& cp = seam_cutpath();
        cp = McgpuSeamCutpath();
        if (!order.empty()) {
            std::vector<uint32_t> ord(order.begin(), order.end());
            mcgpu_cutpath_info ci;
            check(mcgpu_cutpath_edges(ctx, ord.data(), &ci), "mcgpu_cutpath_edges");
            if (ci.unsupported) throw std::runtime_error("libmcgpu seam: a cut-path midpoint test names a face without a plane");
            cp.short_group = ci.short_group != 0;
            const size_t ne = ci.num_edges;
            cp.src.resize(ne), cp.tgt.resize(ne), cp.kind.resize(ne);
            std::vector<uint32_t> faces(2 * ne);
            if (ne)
                check(mcgpu_get_cutpath_edges(ctx, cp.src.data(), cp.tgt.data(), faces.data(), nullptr, cp.kind.data(), ne),
                      "mcgpu_get_cutpath_edges");
            cp.fa.resize(ne), cp.fb.resize(ne);
            for (size_t i = 0; i < ne; ++i) cp.fa[i] = faces[2 * i], cp.fb[i] = faces[2 * i + 1];
        }
    }
    lock.unlock();

    // the registry and its side maps (kernel.cpp:2603-2654): intersection
    // point k (in that order) is vertex ps_vtx_cnt + k of m0
    for (size_t k = 0; k < order.size(); ++k) {
        const size_t i = order[k];
        const ed_t e((ed_t::index_type)re[i]);
        const fd_t g((fd_t::index_type)rf[i]);
        const hd_t h0 = ps.halfedge(e, 0), h1 = ps.halfedge(e, 1);
        if ((uint32_t)ps.source(h0) != v0[re[i]])
            throw std::logic_error("libmcgpu seam: polygon-soup edge numbering differs from the halfedge mesh's");
        const vd_t v = m0.add_vertex(vec3(rp[3 * i], rp[3 * i + 1], rp[3 * i + 2]));
        if ((int)v != ps_vtx_cnt + (int)k) throw std::logic_error("libmcgpu seam: unexpected intersection vertex id");
        registry.emplace_back(e, g);
        intersecting_edges[e].push_back(v);
        const fd_t f0 = ps.face(h0), f1 = ps.face(h1);
        const fd_t ef = f0 != hmesh_t::null_face() ? f0 : f1;
        if ((int)ef >= sm_face_count) { // an edge of the cut mesh through source face g
            cutpath_edges[make_pair(g, ef)].push_back(v);
            const fd_t other = ef == f0 ? f1 : hmesh_t::null_face();
            if (other != hmesh_t::null_face()) cutpath_edges[make_pair(g, other)].push_back(v);
            if (ps.is_border(e)) cm_border_reentrant.push_back(v);
        } else { // an edge of the source mesh through cut face g
            cutpath_edges[make_pair(ef, g)].push_back(v);
            const fd_t other = ef == f0 ? f1 : f0;
            if (other != hmesh_t::null_face()) cutpath_edges[make_pair(other, g)].push_back(v);
        }
        face_ivertices[g].push_back(v);
        if (f0 != hmesh_t::null_face()) face_ivertices[f0].push_back(v);
        if (f1 != hmesh_t::null_face()) face_ivertices[f1].push_back(v);
        if (!partial_cut)
            partial_cut = (int)ps.source(h0) >= sm_vtx_cnt && (f0 == hmesh_t::null_face() || f1 == hmesh_t::null_face());
    }

    // empty registry (kernel.cpp:3211-3229)
    if (registry.empty()) {
        lg.set_reason_for_failure("no face intersection found");
        output.status = (input.enforce_general_position && input.general_position_enforcement_count > 0)
                            ? status_t::GENERAL_POSITION_VIOLATION
                            : status_t::SUCCESS;
        return true;
    }
    return false;
}

const McgpuSeamCutpath& mcgpu_seam_cutpath_result() { return seam_cutpath(); }
