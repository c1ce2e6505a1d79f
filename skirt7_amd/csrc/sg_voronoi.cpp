// sg_voronoi.cpp -- Voronoi neighbour lists for generator sites in a box (the setup step the
// reference delegates to Voro++ in VoronoiMesh::buildMesh, VoronoiMesh.cpp:297-405).
//
// Each cell starts as the domain box (faces tagged with the walls -1..-6: xmin, xmax, ymin, ymax,
// zmin, zmax) and is clipped by the bisector planes of nearby sites, visited shell by shell over
// a uniform bucket grid, until no unvisited site can be closer than twice the cell's current
// radius. The faces that remain name the neighbours. A neighbour list that contains every true
// neighbour gives the walk (VoronoiMesh.cpp:778-873) its exact path, because the walk exits
// through the nearest bisector plane ahead, which always belongs to a true neighbour.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/skirt7_amd.h"
#include "sg_host.hpp"

namespace sg {
namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Convex polyhedron as faces over a shared vertex pool: face k is fv[fo[k] .. fo[k+1]) with tag[k].
// All storage is reused from clip to clip (no allocation once the buffers have grown).
struct Cell {
    std::vector<V3> v, nv2;
    std::vector<int> fv, fo, tag;
    // clip work buffers
    std::vector<double> s;
    std::vector<int> nfv, nfo, ntag, loop, remap;
    std::vector<std::pair<uint64_t, int>> cut;   // edge -> new vertex
    std::vector<std::pair<int, int>> nxt;        // new face edges: from -> to along the cut

    void box(const double* e)
    {
        v.clear(); fv.clear(); fo.clear(); tag.clear();
        for (int q = 0; q < 8; q++)
            v.push_back({(q & 1) ? e[3] : e[0], (q & 2) ? e[4] : e[1], (q & 4) ? e[5] : e[2]});
        // faces ordered counter-clockwise seen from outside; tags are the wall ids
        static const int faces[6][4] = {{0, 4, 6, 2}, {1, 3, 7, 5}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 2, 3, 1}, {4, 5, 7, 6}};
        fo.push_back(0);
        for (int k = 0; k < 6; k++) {
            for (int q = 0; q < 4; q++) fv.push_back(faces[k][q]);
            fo.push_back((int)fv.size());
            tag.push_back(-1 - k);
        }
    }

    double radius2(V3 p) const
    {
        double r = 0;
        for (const auto& q : v) r = std::max(r, dot(q - p, q - p));
        return r;
    }

    // Keeps the part with dot(n, x) <= d. Returns true when the plane cut the cell.
    bool clip(V3 n, double d, int t, double eps)
    {
        const int nv = (int)v.size();
        s.resize(nv);
        bool any_out = false;
        for (int i = 0; i < nv; i++) {
            s[i] = dot(n, v[i]) - d;
            if (s[i] > eps) any_out = true;
        }
        if (!any_out) return false;
        cut.clear();
        auto edge_point = [&](int a, int b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
            for (const auto& c : cut)
                if (c.first == key) return c.second;
            double ta = s[a] / (s[a] - s[b]);
            V3 p{v[a].x + ta * (v[b].x - v[a].x), v[a].y + ta * (v[b].y - v[a].y), v[a].z + ta * (v[b].z - v[a].z)};
            int id = (int)v.size();
            v.push_back(p);
            cut.push_back({key, id});
            return id;
        };
        nfv.clear(); nfo.clear(); ntag.clear(); nxt.clear();
        nfo.push_back(0);
        const int nf = (int)tag.size();
        for (int k = 0; k < nf; k++) {
            const int f0 = fo[k], m = fo[k + 1] - f0;
            const int start = (int)nfv.size();
            int enter = -1, leave = -1;
            for (int i = 0; i < m; i++) {
                int a = fv[f0 + i], b = fv[f0 + (i + 1) % m];
                bool ina = s[a] <= eps, inb = s[b] <= eps;
                if (ina) nfv.push_back(a);
                if (ina && !inb) {
                    int c = (s[a] < -eps) ? edge_point(a, b) : a;
                    if (c != a) nfv.push_back(c);
                    leave = c;
                } else if (!ina && inb) {
                    int c = (s[b] < -eps) ? edge_point(a, b) : b;
                    if (c != b) nfv.push_back(c);
                    enter = c;
                }
            }
            if ((int)nfv.size() - start >= 3) {
                nfo.push_back((int)nfv.size());
                ntag.push_back(tag[k]);
                // the cut runs along this face from `leave` to `enter`; the new face traverses it backwards
                if (enter >= 0 && leave >= 0 && enter != leave) {
                    bool set = false;
                    for (auto& e : nxt)
                        if (e.first == enter) { e.second = leave; set = true; break; }
                    if (!set) nxt.push_back({enter, leave});
                }
            } else {
                nfv.resize(start);
            }
        }
        // assemble the new face from the chained cut edges
        if (!nxt.empty()) {
            loop.clear();
            int start = nxt[0].first, cur = start;
            for (size_t guard = 0; guard <= nxt.size() + 1; guard++) {
                loop.push_back(cur);
                int to = -1;
                for (const auto& e : nxt)
                    if (e.first == cur) { to = e.second; break; }
                if (to < 0) break;
                cur = to;
                if (cur == start) break;
            }
            if (loop.size() >= 3) {
                nfv.insert(nfv.end(), loop.begin(), loop.end());
                nfo.push_back((int)nfv.size());
                ntag.push_back(t);
            }
        }
        // compact the vertex pool
        remap.assign(v.size(), -1);
        nv2.clear();
        for (int& i : nfv) {
            if (remap[i] < 0) { remap[i] = (int)nv2.size(); nv2.push_back(v[i]); }
            i = remap[i];
        }
        v.swap(nv2);
        fv.swap(nfv);
        fo.swap(nfo);
        tag.swap(ntag);
        return true;
    }
};

} // namespace

// Builds neighbour lists (CSR) for n sites in ext. Returns false with a message on failure.
bool voronoi_neighbors(const double ext[6], const double* sites, int64_t n, std::vector<int64_t>& off,
                       std::vector<int32_t>& ids)
{
    const double* e = ext;
    int nb = std::max(1, std::min(512, (int)std::cbrt((double)n / 2.0)));
    double w[3] = {(e[3] - e[0]) / nb, (e[4] - e[1]) / nb, (e[5] - e[2]) / nb};
    double wmin = std::min(w[0], std::min(w[1], w[2]));
    auto bidx = [&](double x, int a) {
        int i = (int)(nb * (x - e[a]) / (e[a + 3] - e[a]));
        return std::max(0, std::min(nb - 1, i));
    };
    std::vector<int32_t> cnt((size_t)nb * nb * nb + 1, 0), bucket(n), bid(n);
    for (int64_t i = 0; i < n; i++) {
        int q = (bidx(sites[3 * i], 0) * nb + bidx(sites[3 * i + 1], 1)) * nb + bidx(sites[3 * i + 2], 2);
        bucket[i] = q;
        cnt[q + 1]++;
    }
    for (size_t q = 1; q < cnt.size(); q++) cnt[q] += cnt[q - 1];
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < n; i++) bid[fill[bucket[i]]++] = (int32_t)i;

    double diag = std::sqrt((e[3] - e[0]) * (e[3] - e[0]) + (e[4] - e[1]) * (e[4] - e[1]) + (e[5] - e[2]) * (e[5] - e[2]));
    const double eps = 1e-13 * diag;
    std::vector<std::vector<int32_t>> lists(n);
    std::atomic<int64_t> next(0);
    auto work = [&]() {
        Cell cell;
        std::vector<std::pair<double, int32_t>> cand;
        for (;;) {
            int64_t i0 = next.fetch_add(64);
            if (i0 >= n) break;
            for (int64_t i = i0; i < std::min(n, i0 + 64); i++) {
                V3 p{sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]};
                cell.box(e);
                int ci = bidx(p.x, 0), cj = bidx(p.y, 1), ck = bidx(p.z, 2);
                double R2 = cell.radius2(p);
                for (int ring = 0; ring <= nb; ring++) {
                    cand.clear();
                    for (int a = std::max(0, ci - ring); a <= std::min(nb - 1, ci + ring); a++)
                        for (int b = std::max(0, cj - ring); b <= std::min(nb - 1, cj + ring); b++)
                            for (int c = std::max(0, ck - ring); c <= std::min(nb - 1, ck + ring); c++) {
                                if (std::max(std::abs(a - ci), std::max(std::abs(b - cj), std::abs(c - ck))) != ring)
                                    continue;
                                // a bucket whose box lies beyond twice the radius holds no candidate
                                const double bx = std::max(0.0, std::max(e[0] + a * w[0] - p.x, p.x - (e[0] + (a + 1) * w[0])));
                                const double by = std::max(0.0, std::max(e[1] + b * w[1] - p.y, p.y - (e[1] + (b + 1) * w[1])));
                                const double bz = std::max(0.0, std::max(e[2] + c * w[2] - p.z, p.z - (e[2] + (c + 1) * w[2])));
                                if (bx * bx + by * by + bz * bz > 4.0 * R2 * (1.0 + 1e-9)) continue;
                                int q = (a * nb + b) * nb + c;
                                for (int t = cnt[q]; t < cnt[q + 1]; t++) {
                                    int32_t j = bid[t];
                                    if (j == i) continue;
                                    V3 pj{sites[3 * j], sites[3 * j + 1], sites[3 * j + 2]};
                                    // beyond twice the current radius a site cannot clip (the loop
                                    // below stops there, and the radius only shrinks)
                                    const double dj = dot(pj - p, pj - p);
                                    if (dj <= 4.0 * R2) cand.push_back({dj, j});
                                }
                            }
                    std::sort(cand.begin(), cand.end());
                    for (auto& cj2 : cand) {
                        if (cj2.first > 4.0 * R2) break;
                        if (cj2.first <= eps * eps) continue;   // coincident site: no bisector
                        int32_t j = cj2.second;
                        V3 pj{sites[3 * j], sites[3 * j + 1], sites[3 * j + 2]};
                        V3 nrm = pj - p;
                        double d = 0.5 * (dot(pj, nrm) + dot(p, nrm));
                        if (cell.clip(nrm, d, j, eps * std::sqrt(dot(nrm, nrm)))) R2 = cell.radius2(p);
                    }
                    double reach = ring * wmin;
                    if (reach * reach > 4.0 * R2 * (1.0 + 1e-9)) break;
                }
                lists[i].assign(cell.tag.begin(), cell.tag.end());
            }
        }
    };
    int T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    off.assign(n + 1, 0);
    for (int64_t i = 0; i < n; i++) off[i + 1] = off[i] + (int64_t)lists[i].size();
    ids.clear();
    ids.reserve(off[n]);
    for (auto& l : lists) ids.insert(ids.end(), l.begin(), l.end());
    return true;
}

} // namespace sg

extern "C" {

// Voronoi neighbour lists of n sites in the box ext (VoronoiMesh::buildMesh, VoronoiMesh.cpp:297-405):
// offsets[n+1] and ids (at most cap; walls -1..-6). *nids receives the total count; returns -5 when
// cap is too small (call again with a larger buffer).
int sg_voronoi_neighbors(const double ext[6], const double* sites, int64_t n, int64_t* offsets, int32_t* ids,
                         int64_t cap, int64_t* nids)
{
    std::vector<int64_t> off;
    std::vector<int32_t> v;
    if (n < 1) { sg::set_error("need at least one site"); return -1; }
    sg::voronoi_neighbors(ext, sites, n, off, v);
    *nids = (int64_t)v.size();
    std::memcpy(offsets, off.data(), off.size() * sizeof(int64_t));
    if ((int64_t)v.size() > cap) { sg::set_error("neighbour buffer too small"); return -5; }
    std::memcpy(ids, v.data(), v.size() * sizeof(int32_t));
    return 0;
}

} // extern "C"
