// The record of the last dispatch's cut-path edges (cutpath_probe.h).
#include "cutpath_probe.h"

#include <algorithm>
#include <cstdint>
#include <mutex>

namespace {

struct Record {
    std::vector<uint32_t> reg_edge, reg_face; // registry in the dispatch's point order
    std::vector<uint32_t> src, tgt;           // cut-path edges: points (rank = vertex - ps vertex count)
    std::vector<uint32_t> fe;                 // (face, edge) of every face's edge list, faces ascending
    uint64_t dispatches = 0;
};

std::mutex g_mutex;
Record g_rec;

} // namespace

void mcut_probe_record_cutpath(const hmesh_t& m0, int ps_vtx_cnt, const std::vector<std::pair<ed_t, fd_t>>& registry,
                               const std::vector<ed_t>& cutpath_edges,
                               const std::unordered_map<fd_t, std::vector<ed_t>>& face_edges)
{
    Record r;
    for (const std::pair<ed_t, fd_t>& x : registry) {
        r.reg_edge.push_back((uint32_t)x.first);
        r.reg_face.push_back((uint32_t)x.second);
    }
    for (const ed_t& e : cutpath_edges) {
        const hd_t h = m0.halfedge(e, 0);
        r.src.push_back((uint32_t)((int)m0.source(h) - ps_vtx_cnt));
        r.tgt.push_back((uint32_t)((int)m0.target(h) - ps_vtx_cnt));
    }
    std::vector<uint32_t> faces;
    for (const auto& kv : face_edges) faces.push_back((uint32_t)kv.first);
    std::sort(faces.begin(), faces.end());
    for (uint32_t f : faces)
        for (const ed_t& e : face_edges.at(fd_t(f))) {
            r.fe.push_back(f);
            r.fe.push_back((uint32_t)e);
        }
    std::lock_guard<std::mutex> lock(g_mutex);
    r.dispatches = g_rec.dispatches + 1;
    g_rec = std::move(r);
}

extern "C" {

// sizes of the record: points, edges, (face, edge) entries; returns the number
// of dispatches recorded so far
uint64_t mcut_probe_cutpath_sizes(uint64_t* points, uint64_t* edges, uint64_t* face_edges)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (points) *points = g_rec.reg_edge.size();
    if (edges) *edges = g_rec.src.size();
    if (face_edges) *face_edges = g_rec.fe.size() / 2;
    return g_rec.dispatches;
}

void mcut_probe_cutpath_copy(uint32_t* reg_edge, uint32_t* reg_face, uint32_t* src, uint32_t* tgt, uint32_t* fe)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (reg_edge) std::copy(g_rec.reg_edge.begin(), g_rec.reg_edge.end(), reg_edge);
    if (reg_face) std::copy(g_rec.reg_face.begin(), g_rec.reg_face.end(), reg_face);
    if (src) std::copy(g_rec.src.begin(), g_rec.src.end(), src);
    if (tgt) std::copy(g_rec.tgt.begin(), g_rec.tgt.end(), tgt);
    if (fe) std::copy(g_rec.fe.begin(), g_rec.fe.end(), fe);
}

} // extern "C"
