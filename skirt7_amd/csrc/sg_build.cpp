// sg_build.cpp -- host-side setup: tree topology and neighbour lists, Sersic tables,
// deterministic tree subdivision and cell-density sampling for the synthetic workloads.
//
// The reference builds these in its own setup phase (TreeDustGrid::setupSelfBefore,
// DustSystem::setupSelfAfter, SersicFunction). Where the transport result depends on the exact
// outcome (node numbering, child boxes, neighbour order), the steps follow the reference; the
// random sampling is replaced by counter-based streams so that setup is reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/skirt7_amd.h"
#include "sg_host.hpp"

namespace sg {

// ----------------------------------------------------------------------- tree topology

bool HostTree::build(int ar, const double e[6], const uint8_t* flags, int64_t nnodes)
{
    arity = ar;
    std::memcpy(ext, e, sizeof ext);
    double wx = e[3] - e[0], wy = e[4] - e[1], wz = e[5] - e[2];
    eps = 1e-12 * std::sqrt(wx * wx + wy * wy + wz * wz);   // TreeDustGrid.cpp:74
    child0.assign(1, -1);
    dir.assign(1, -1);
    father.assign(1, -1);
    bounds.assign(e, e + 6);
    level.assign(1, 0);
    level.reserve(nnodes);
    child0.reserve(nnodes); dir.reserve(nnodes); father.reserve(nnodes); bounds.reserve(6 * nnodes);
    for (int64_t l = 0; l < (int64_t)child0.size(); l++) {
        if (l >= nnodes) { set_error("subdivision flags end before the tree does"); return false; }
        if (!flags[l]) continue;
        int32_t id = (int32_t)child0.size();
        child0[l] = id;
        double b[6];
        std::memcpy(b, &bounds[6 * l], sizeof b);
        int lev = level[l] + 1;
        if (arity == 8) {
            // children in the order of OctTreeNode.cpp:38-49, split at Box::center()
            double c[3] = {0.5 * (b[0] + b[3]), 0.5 * (b[1] + b[4]), 0.5 * (b[2] + b[5])};
            for (int q = 0; q < 8; q++) {
                for (int a = 0; a < 3; a++) {
                    bool hi = (q >> a) & 1;
                    bounds.push_back(hi ? c[a] : b[a]);
                }
                for (int a = 0; a < 3; a++) {
                    bool hi = (q >> a) & 1;
                    bounds.push_back(hi ? b[a + 3] : c[a]);
                }
                child0.push_back(-1); dir.push_back(-1); father.push_back((int32_t)l); level.push_back((int16_t)lev);
            }
        } else {
            // BinTreeNode.cpp:40-77, alternating direction level % 3
            int a = level[l] % 3;
            dir[l] = (int8_t)a;
            double c = 0.5 * (b[a] + b[a + 3]);
            for (int q = 0; q < 2; q++) {
                double nb[6];
                std::memcpy(nb, b, sizeof nb);
                if (q == 0) nb[a + 3] = c;
                else nb[a] = c;
                bounds.insert(bounds.end(), nb, nb + 6);
                child0.push_back(-1); dir.push_back(-1); father.push_back((int32_t)l); level.push_back((int16_t)lev);
            }
        }
    }
    if ((int64_t)child0.size() != nnodes) { set_error("subdivision flags do not match the tree size"); return false; }
    // TreeDustGrid.cpp:110-121: cell numbers are the ranks of the leaves in id order
    cellnumber.assign(nnodes, -1);
    idv.clear();
    for (int64_t l = 0; l < nnodes; l++)
        if (child0[l] < 0) { cellnumber[l] = (int32_t)idv.size(); idv.push_back((int32_t)l); }
    return true;
}

// Ropes from the root down, in id order (children after their parents): a child's wall inside its
// parent leads to the sibling across it; a wall on the parent's boundary leads through the parent's
// rope R: to R itself when R is a leaf or coarser than the parent, else to R's child adjacent to the
// child's wall (R is then at the parent's level and, in an alternating binary tree, split along the
// same axis).
void HostTree::build_ropes(std::vector<int32_t>& rope) const
{
    const int64_t n = (int64_t)child0.size();
    rope.assign(6 * (size_t)n, -1);
    for (int64_t p = 0; p < n; p++) {
        const int32_t c0 = child0[p];
        if (c0 < 0) continue;
        const int nch = arity == 8 ? 8 : 2;
        for (int q = 0; q < nch; q++) {
            const int64_t c = c0 + q;
            for (int w = 0; w < 6; w++) {
                const int a = w >> 1, side = w & 1;
                int32_t r;
                // the child's position along the wall's axis: 0 lower half, 1 upper half, -1 whole extent
                const int pos = arity == 8 ? ((q >> a) & 1) : (dir[p] == a ? q : -1);
                if (pos >= 0 && pos != side) {
                    r = arity == 8 ? c0 + (q ^ (1 << a)) : c0 + (1 - q);   // sibling across an inner wall
                } else {
                    const int32_t R = rope[6 * (size_t)p + w];
                    if (R < 0 || child0[R] < 0 || level[R] != level[p]) r = R;
                    else if (arity == 8) r = child0[R] + (q ^ (1 << a));
                    else if (dir[R] == a) r = child0[R] + (1 - side);
                    else if (dir[R] == dir[p]) r = child0[R] + q;
                    else r = R;
                }
                rope[6 * (size_t)c + w] = r;
            }
        }
    }
}

// Neighbour lists, built by the same incremental procedure as OctTreeNode::addneighbors /
// BinTreeNode::addneighbors (called for every node in id order, TreeDustGrid.cpp:156-161) and
// ordered by decreasing overlap area with std::sort (TreeNode.cpp:97-158), so that the first
// neighbour containing a point is the one the reference picks.
void HostTree::build_neighbors()
{
    static const int comp[6] = {1, 0, 3, 2, 5, 4};
    size_t nn = child0.size();
    nb.assign(6 * nn, {});
    auto make = [&](int w1, int32_t n1, int32_t n2) {
        nb[6 * (size_t)n1 + w1].push_back(n2);
        nb[6 * (size_t)n2 + comp[w1]].push_back(n1);
    };
    auto drop = [&](int32_t l, int w, int32_t n) {
        auto& v = nb[6 * (size_t)l + w];
        auto it = std::find(v.begin(), v.end(), n);
        if (it != v.end()) v.erase(it);
    };
    for (size_t l = 0; l < nn; l++) {
        int32_t c0 = child0[l];
        if (c0 < 0) continue;
        if (arity == 8) {
            static const int inner[12][3] = {{1, 0, 1}, {3, 0, 2}, {5, 0, 4}, {3, 1, 3}, {5, 1, 5}, {1, 2, 3},
                                             {5, 2, 6}, {5, 3, 7}, {1, 4, 5}, {3, 4, 6}, {3, 5, 7}, {1, 6, 7}};
            for (auto& t : inner) make(t[0], c0 + t[1], c0 + t[2]);
            const double* cb = &bounds[6 * (size_t)c0];
            double c[3] = {cb[3], cb[4], cb[5]};
            for (int w = 0; w < 6; w++) {
                int a = w / 2, side = w % 2;
                int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
                std::vector<int32_t> lst = nb[6 * l + w];
                for (int32_t n : lst) {
                    drop(n, comp[w], (int32_t)l);
                    const double* q = &bounds[6 * (size_t)n];
                    for (int kv = 0; kv < 2; kv++)
                        for (int ku = 0; ku < 2; ku++) {
                            bool okU = ku ? q[u + 3] >= c[u] : q[u] <= c[u];
                            bool okV = kv ? q[v + 3] >= c[v] : q[v] <= c[v];
                            if (okU && okV) make(comp[w], n, c0 + ((side << a) | (ku << u) | (kv << v)));
                        }
                }
            }
        } else {
            int a = dir[l];
            double c = bounds[6 * (size_t)c0 + a + 3];
            make(2 * a + 1, c0, c0 + 1);
            for (int w = 0; w < 6; w++) {
                std::vector<int32_t> lst = nb[6 * l + w];
                for (int32_t n : lst) {
                    drop(n, comp[w], (int32_t)l);
                    if (w / 2 == a) make(comp[w], n, c0 + (w % 2));
                    else {
                        const double* q = &bounds[6 * (size_t)n];
                        if (q[a] <= c) make(comp[w], n, c0);
                        if (q[a + 3] >= c) make(comp[w], n, c0 + 1);
                    }
                }
            }
        }
    }
    for (size_t l = 0; l < nn; l++) {
        const double* b = &bounds[6 * l];
        for (int w = 0; w < 6; w++) {
            int a = w / 2;
            int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
            auto ov = [&](int32_t n) {
                const double* q = &bounds[6 * (size_t)n];
                return std::max(std::min(b[u + 3], q[u + 3]) - std::max(b[u], q[u]), 0.) *
                       std::max(std::min(b[v + 3], q[v + 3]) - std::max(b[v], q[v]), 0.);
            };
            auto& lst = nb[6 * l + w];
            std::sort(lst.begin(), lst.end(), [&](int32_t n1, int32_t n2) { return ov(n1) > ov(n2); });
        }
    }
}

// ----------------------------------------------------------------------- special functions

// Lanczos series for ln Gamma (Numerical Recipes gammln), as SpecialFunctions::lngamma
// (SpecialFunctions.cpp:16-33) evaluates it.
double lngamma(double a)
{
    static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                  -1.231739572450155, 0.1208650973866179e-2, -0.5395239384953e-5};
    double y = a;
    double t = a + 5.5;
    t -= (a + 0.5) * std::log(t);
    double ser = 1.000000000190015;
    for (double c : cof) ser += c / ++y;
    return -t + std::log(2.5066282746310005 * ser / a);
}

// Counter-based uniform in (0,1) for setup sampling.
static inline void philox_host(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0; c[1] = (uint32_t)p1; c[2] = n2; c[3] = (uint32_t)p0;
    }
}
static inline void sample3(uint32_t seed, uint64_t item, uint32_t j, double u[3])
{
    uint32_t c[4] = {(uint32_t)item, (uint32_t)(item >> 32), j, 0x5e7u};
    philox_host(c, seed, 0x9d2c5680u);
    for (int i = 0; i < 3; i++) u[i] = ((double)c[i] + 0.5) * (1.0 / 4294967296.0);
}

// Normalised geometry density for a record laid out like a stellar component
// (ExpDiskGeometry.cpp:25-40,103-114 with SpiralStructureGeometryDecorator.cpp:165-172,223-229).
struct GeomDensity {
    double p[16];
    double rho0 = 0, tanp = 1, CN = 1;
    explicit GeomDensity(const double* q)
    {
        std::memcpy(p, q, sizeof p);
        double hR = p[1], hz = p[2], Rmax = p[3], zmax = p[4], Rmin = p[5];
        double intphi = 2.0 * M_PI;
        double intz = (zmax > 0) ? -2.0 * hz * std::expm1(-zmax / hz) : 2.0 * hz;
        double tmin = (Rmin > 0) ? std::exp(-Rmin / hR) * (1.0 + Rmin / hR) : 1.0;
        double tmax = (Rmax > 0) ? std::exp(-Rmax / hR) * (1.0 + Rmax / hR) : 0.0;
        double intR = hR * hR * (tmin - tmax);
        rho0 = 1.0 / (intR * intphi * intz);
        if (p[6] != 0) {
            tanp = std::tan(p[8]);
            CN = std::sqrt(M_PI) * std::exp(lngamma(p[12] + 1.0)) / std::exp(lngamma(p[12] + 0.5));
        }
    }
    double operator()(double x, double y, double z) const
    {
        double R = std::sqrt(x * x + y * y), absz = std::fabs(z);
        double hR = p[1], hz = p[2], Rmax = p[3], zmax = p[4], Rmin = p[5];
        if (Rmax > 0.0 && R > Rmax) return 0.0;
        if (zmax > 0.0 && absz > zmax) return 0.0;
        if (R < Rmin) return 0.0;
        double rho = rho0 * std::exp(-R / hR) * std::exp(-absz / hz);
        if (p[6] != 0) {
            double m = p[7], R0 = p[9], phi0 = p[10], w = p[11];
            int N = (int)p[12];
            double phi = std::atan2(y, x);
            double gamma = std::log(R / R0) / tanp + phi0 + 0.5 * M_PI / m;
            rho *= (1.0 - w) + w * CN * std::pow(std::sin(0.5 * m * (gamma - phi)), 2 * N);
        }
        return rho;
    }
};

void disc_density_constants(const double* geom16, double* rho0, double* tanp, double* CN)
{
    GeomDensity d(geom16);
    *rho0 = d.rho0;
    *tanp = d.tanp;
    *CN = d.CN;
}

template <class F> static void parallel_for(int64_t n, F f)
{
    int T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int64_t> next(0);
    auto work = [&]() {
        for (;;) {
            int64_t i0 = next.fetch_add(1024);
            if (i0 >= n) break;
            int64_t i1 = std::min(n, i0 + 1024);
            for (int64_t i = i0; i < i1; i++) f(i);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

} // namespace sg

using namespace sg;

namespace sg {

// The stopping criteria of TreeDustGrid::subdivide (TreeDustGrid.cpp:194-220): a node below maxlevel is
// split when no criterion is set, or when at least one set criterion is not met -- its mass reaches
// maxmassfraction of the total, its optical depth kappaV M / V^(2/3) (TreeNodeSampleDensityCalculator.cpp,
// kappaV = 2600 m2/kg, Units.cpp:29) reaches maxopticaldepth, or the dispersion (max - min) / max of its
// density samples reaches maxdensdisp.
struct Criteria {
    double maxmassfraction = 0, maxopticaldepth = 0, maxdensdisp = 0;
};
// A node's mass and, for sampled densities, the smallest and largest sample (negative when not known).
struct NodeStats {
    double mass, minrho = -1, maxrho = -1;
};
static bool need_division(const Criteria& c, const NodeStats& st, const double* q, double total_mass)
{
    if (c.maxmassfraction == 0 && c.maxopticaldepth == 0 && c.maxdensdisp == 0) return true;
    if (c.maxmassfraction > 0 && st.mass / total_mass >= c.maxmassfraction) return true;
    if (c.maxopticaldepth > 0) {
        double vol = (q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]);
        if (2600.0 * st.mass / std::pow(vol, 2. / 3.) >= c.maxopticaldepth) return true;
    }
    if (c.maxdensdisp > 0) {
        double disp = st.maxrho > 0 ? (st.maxrho - st.minrho) / st.maxrho : 0;
        if (disp >= c.maxdensdisp) return true;
    }
    return false;
}

// TreeDustGrid::setupSelfBefore/subdivide (TreeDustGrid.cpp:49-230): breadth-first subdivision with the
// reference's node numbering and child boxes; a node is split while its level is at most minlevel, or
// when below maxlevel the criteria above ask for it (stats(box, node) gives its mass and samples).
template <class StatFn>
static int subdivide_tree(int arity, const double ext[6], int minlevel, int maxlevel, const Criteria& crit,
                          double total_mass, StatFn stats, uint8_t* flags, int64_t cap, int64_t* nnodes)
{
    std::vector<double> b(ext, ext + 6);
    std::vector<int16_t> level(1, 0);
    std::vector<uint8_t> f;
    int64_t begin = 0;
    while (begin < (int64_t)level.size()) {
        int64_t end = (int64_t)level.size();
        std::vector<uint8_t> split(end - begin, 0);
        parallel_for(end - begin, [&](int64_t i) {
            int64_t l = begin + i;
            int lev = level[l];
            if (lev <= minlevel) { split[i] = 1; return; }
            if (lev >= maxlevel) return;
            if (need_division(crit, stats(&b[6 * l], l), &b[6 * l], total_mass)) split[i] = 1;
        });
        for (int64_t l = begin; l < end; l++) {
            f.push_back(split[l - begin]);
            if (!split[l - begin]) continue;
            double q[6];
            std::memcpy(q, &b[6 * l], sizeof q);
            int lev = level[l] + 1;
            if (arity == 8) {
                double c[3] = {0.5 * (q[0] + q[3]), 0.5 * (q[1] + q[4]), 0.5 * (q[2] + q[5])};
                for (int k = 0; k < 8; k++) {
                    for (int a = 0; a < 3; a++) b.push_back(((k >> a) & 1) ? c[a] : q[a]);
                    for (int a = 0; a < 3; a++) b.push_back(((k >> a) & 1) ? q[a + 3] : c[a]);
                    level.push_back((int16_t)lev);
                }
            } else {
                int a = level[l] % 3;
                double c = 0.5 * (q[a] + q[a + 3]);
                for (int k = 0; k < 2; k++) {
                    double nb[6];
                    std::memcpy(nb, q, sizeof nb);
                    if (k == 0) nb[a + 3] = c;
                    else nb[a] = c;
                    b.insert(b.end(), nb, nb + 6);
                    level.push_back((int16_t)lev);
                }
            }
        }
        begin = end;
        if ((int64_t)level.size() > cap) { set_error("flag buffer too small"); *nnodes = (int64_t)level.size(); return -5; }
    }
    *nnodes = (int64_t)f.size();
    if ((int64_t)f.size() > cap) { set_error("flag buffer too small"); return -5; }
    std::memcpy(flags, f.data(), f.size());
    return 0;
}

// A density given per voxel of a regular grid over a box (value [i][j][k] at i*ny*nz + j*nz + k), constant
// inside each voxel. The mass in any axis-aligned box is exact: per axis the box covers at most a partial
// voxel, a run of whole voxels and a partial voxel, and each of the up to 27 combinations is a
// summed-volume lookup times its edge lengths.
struct VoxelField {
    double e[6];
    int n[3];
    double w[3];
    std::vector<double> P;   // prefix sums, (nx+1)(ny+1)(nz+1)
    double total = 0;
    VoxelField(const double ext[6], const double* v, int nx, int ny, int nz)
    {
        std::memcpy(e, ext, sizeof e);
        n[0] = nx; n[1] = ny; n[2] = nz;
        for (int a = 0; a < 3; a++) w[a] = (e[a + 3] - e[a]) / n[a];
        const size_t sy = (size_t)nz + 1, sx = (size_t)(ny + 1) * sy;
        P.assign((size_t)(nx + 1) * sx, 0.0);
        for (int i = 1; i <= nx; i++)
            for (int j = 1; j <= ny; j++)
                for (int k = 1; k <= nz; k++)
                    P[i * sx + j * sy + k] = v[((size_t)(i - 1) * ny + (j - 1)) * nz + (k - 1)] + P[(i - 1) * sx + j * sy + k] +
                                             P[i * sx + (j - 1) * sy + k] + P[i * sx + j * sy + k - 1] -
                                             P[(i - 1) * sx + (j - 1) * sy + k] - P[(i - 1) * sx + j * sy + k - 1] -
                                             P[i * sx + (j - 1) * sy + k - 1] + P[(i - 1) * sx + (j - 1) * sy + k - 1];
        total = P.back() * w[0] * w[1] * w[2];
    }
    double S(int i0, int i1, int j0, int j1, int k0, int k1) const
    {
        const size_t sy = (size_t)n[2] + 1, sx = (size_t)(n[1] + 1) * sy;
        auto at = [&](int i, int j, int k) { return P[i * sx + j * sy + k]; };
        return at(i1, j1, k1) - at(i0, j1, k1) - at(i1, j0, k1) - at(i1, j1, k0) + at(i0, j0, k1) + at(i0, j1, k0) +
               at(i1, j0, k0) - at(i0, j0, k0);
    }
    // voxel index ranges [i0, i1) and the edge length inside each voxel of the range, for [lo, hi] on axis a
    int segments(int a, double lo, double hi, int* i0, int* i1, double* len) const
    {
        double u0 = std::max(0.0, std::min((double)n[a], (lo - e[a]) / w[a]));
        double u1 = std::max(0.0, std::min((double)n[a], (hi - e[a]) / w[a]));
        int f0 = std::min(n[a] - 1, (int)std::floor(u0)), f1 = std::min(n[a] - 1, (int)std::floor(u1));
        int c = 0;
        if (f0 == f1) { i0[c] = f0; i1[c] = f0 + 1; len[c++] = (u1 - u0) * w[a]; return c; }
        if (f0 + 1 > u0) { i0[c] = f0; i1[c] = f0 + 1; len[c++] = (f0 + 1 - u0) * w[a]; }
        if (f1 > f0 + 1) { i0[c] = f0 + 1; i1[c] = f1; len[c++] = w[a]; }
        if (u1 > f1) { i0[c] = f1; i1[c] = f1 + 1; len[c++] = (u1 - f1) * w[a]; }
        return c;
    }
    double mass(const double* q) const
    {
        int a0[3], a1[3], b0[3], b1[3], c0[3], c1[3];
        double la[3], lb[3], lc[3];
        int na = segments(0, q[0], q[3], a0, a1, la);
        int nb = segments(1, q[1], q[4], b0, b1, lb);
        int nc = segments(2, q[2], q[5], c0, c1, lc);
        double m = 0;
        for (int x = 0; x < na; x++)
            for (int y = 0; y < nb; y++)
                for (int z = 0; z < nc; z++)
                    m += la[x] * lb[y] * lc[z] * S(a0[x], a1[x], b0[y], b1[y], c0[z], c1[z]);
        return m;
    }
};

} // namespace sg

extern "C" {

// SersicFunction (SersicFunction.cpp:15-68): 101 log-spaced radii s in [1e-6, 1e4] (units of reff)
// and the normalised cumulative mass M(s) of the deprojected Sersic profile.
int sg_sersic_table(double n, double* s101, double* M101)
{
    if (!(n >= 0.5 && n <= 10.0)) { set_error("the Sersic index must be in [0.5, 10]"); return -1; }
    const double b = 2.0 * n - 1.0 / 3.0 + 4.0 / 405.0 / n + 46.0 / 25515.0 / (n * n) + 131.0 / 1148175.0 / (n * n * n);
    const double I0 = std::pow(b, 2.0 * n) / (M_PI * std::exp(lngamma(2.0 * n + 1)));
    const int Ns = 101, Nu = 10000;
    std::vector<double> S(Ns);
    const double dlogs = (4.0 - -6.0) / (Ns - 1.0);
    const double umax = std::sqrt((100.0 + 1.0) * (100.0 - 1.0)), du = umax / Nu;
    for (int i = 0; i < Ns; i++) {
        double s = std::pow(10.0, -6.0 + i * dlogs);
        s101[i] = s;
        double alpha = b * std::pow(s, 1.0 / n);
        double sum = 0.0;
        for (int j = 0; j <= Nu; j++) {
            double u = j * du, u2 = u * u;
            double w = u > 1e-3 ? (std::pow(1.0 + u2, 2.0 * n) - 1.0) / u2
                                : 2.0 * n + n * (2.0 * n - 1.0) * u2 + 2.0 / 3.0 * n * (2.0 * n - 1.0) * (n - 1.0) * u2 * u2;
            sum += ((j == 0 || j == Nu) ? 0.5 : 1.0) * (2.0 * std::exp(-alpha * (1.0 + u2)) / std::sqrt(w));
        }
        S[i] = I0 * std::pow(b, n) * std::pow(alpha, 1.0 - n) / M_PI * du * sum;
    }
    auto Sof = [&](double s) {
        if (s <= s101[0]) return S[0];
        if (s >= s101[Ns - 1]) return S[Ns - 1];
        int lo = 0, hi = Ns - 1;
        while (hi - lo > 1) { int mid = (lo + hi) / 2; if (s < s101[mid]) hi = mid; else lo = mid; }
        double x = std::log10(s), x1 = std::log10(s101[lo]), x2 = std::log10(s101[lo + 1]);
        double f1 = S[lo], f2 = S[lo + 1];
        bool lf = f1 > 0 && f2 > 0;
        if (lf) { f1 = std::log10(f1); f2 = std::log10(f2); }
        double fx = f1 + ((x - x1) / (x2 - x1)) * (f2 - f1);
        return lf ? std::pow(10, fx) : fx;
    };
    M101[0] = 0.0;
    for (int i = 1; i < Ns; i++) {
        double sum = 0.0;
        for (int j = 0; j <= 32; j++) {
            double ds = (s101[i] - s101[i - 1]) / 32.0;
            double s = s101[i - 1] + j * ds;
            sum += ((j == 0 || j == 32) ? 0.5 : 1.0) * Sof(s) * s * s * ds;
        }
        M101[i] = M101[i - 1] + 4.0 * M_PI * sum;
    }
    double tot = M101[Ns - 1];
    for (int i = 0; i < Ns; i++) M101[i] /= tot;
    return 0;
}

// Normalised density of an exponential-disc geometry record (see SG_STAR_NPARAM layout) at n
// positions.
int sg_geometry_density(const double* geom16, int64_t n, const double* xyz, double* out)
{
    if ((int)geom16[0] != 1) { set_error("only the exponential disc density is available"); return -1; }
    GeomDensity d(geom16);
    for (int64_t i = 0; i < n; i++) out[i] = d(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return 0;
}

// TreeDustGrid::setupSelfBefore / subdivide (TreeDustGrid.cpp:49-230) with the mass-fraction
// criterion: nodes at level <= minlevel always split; below maxlevel a node splits when the mass
// estimated from nsamples density samples (TreeNodeSampleDensityCalculator.cpp:26-44) reaches
// maxmassfraction of total_mass. The samples come from a counter-based stream keyed by seed and
// node id, so the tree does not depend on threading. Writes one flag per node (breadth-first id
// order) into flags[0..cap) and the node count into *nnodes.
int sg_tree_subdivide(int arity, const double ext[6], const double* geom16, double scale, double total_mass,
                      int minlevel, int maxlevel, double maxmassfraction, int nsamples, uint32_t seed,
                      uint8_t* flags, int64_t cap, int64_t* nnodes)
{
    return sg_tree_subdivide_multi(arity, ext, 1, geom16, &scale, total_mass, minlevel, maxlevel, maxmassfraction,
                                   0.0, 0.0, nsamples, seed, flags, cap, nnodes);
}

// The same with the density of a CompDustDistribution: the sum over components of scale(h) times the
// geometry's density (CompDustDistribution::density), sampled at the same positions.
int sg_tree_subdivide_multi(int arity, const double ext[6], int ncomp, const double* geoms16, const double* scales,
                            double total_mass, int minlevel, int maxlevel, double maxmassfraction,
                            double maxopticaldepth, double maxdensdisp, int nsamples, uint32_t seed, uint8_t* flags,
                            int64_t cap, int64_t* nnodes)
{
    if (!(arity == 8 || arity == 2)) { set_error("arity must be 8 or 2"); return -1; }
    if (maxlevel < 2 || maxlevel <= minlevel) { set_error("bad tree levels"); return -1; }
    if (ncomp < 1 || !geoms16 || !scales) { set_error("at least one dust component is needed"); return -1; }
    if (nsamples < 1) { set_error("at least one density sample per node is needed"); return -1; }
    std::vector<GeomDensity> dens;
    for (int h = 0; h < ncomp; h++) dens.emplace_back(geoms16 + 16 * h);
    Criteria crit;
    crit.maxmassfraction = maxmassfraction;
    crit.maxopticaldepth = maxopticaldepth;
    crit.maxdensdisp = maxdensdisp;
    // TreeNodeSampleDensityCalculator: nsamples counter-based density samples in the node's box
    auto stats = [&](const double* q, int64_t l) {
        double vol = (q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]);
        double sum = 0, lo = 0, hi = 0;
        for (int j = 0; j < nsamples; j++) {
            double u[3];
            sample3(seed, (uint64_t)l, (uint32_t)j, u);
            double x = q[0] + u[0] * (q[3] - q[0]), y = q[1] + u[1] * (q[4] - q[1]), z = q[2] + u[2] * (q[5] - q[2]);
            double d = 0;
            for (int h = 0; h < ncomp; h++) d += scales[h] * dens[h](x, y, z);
            sum += d;
            lo = j ? std::min(lo, d) : d;
            hi = j ? std::max(hi, d) : d;
        }
        NodeStats st;
        st.mass = sum / nsamples * vol;
        st.minrho = lo;
        st.maxrho = hi;
        return st;
    };
    return subdivide_tree(arity, ext, minlevel, maxlevel, crit, total_mass, stats, flags, cap, nnodes);
}

// The same subdivision with node masses from a voxel density field (exact mass in the node's box).
int sg_tree_subdivide_field(int arity, const double ext[6], const double* field, int nx, int ny, int nz,
                            int minlevel, int maxlevel, double maxmassfraction, uint8_t* flags, int64_t cap,
                            int64_t* nnodes)
{
    return sg_tree_subdivide_field_od(arity, ext, field, nx, ny, nz, minlevel, maxlevel, maxmassfraction, 0.0, flags,
                                      cap, nnodes);
}

// With the optical-depth criterion too; TreeNodeBoxDensityCalculator gives masses only, so there is no
// density-dispersion criterion here (the reference throws "Calculation is not supported").
int sg_tree_subdivide_field_od(int arity, const double ext[6], const double* field, int nx, int ny, int nz,
                               int minlevel, int maxlevel, double maxmassfraction, double maxopticaldepth,
                               uint8_t* flags, int64_t cap, int64_t* nnodes)
{
    if (!(arity == 8 || arity == 2)) { set_error("arity must be 8 or 2"); return -1; }
    if (maxlevel < 2 || maxlevel <= minlevel) { set_error("bad tree levels"); return -1; }
    if (nx < 1 || ny < 1 || nz < 1) { set_error("empty density field"); return -1; }
    VoxelField vf(ext, field, nx, ny, nz);
    if (!(vf.total > 0)) { set_error("the density field has no mass"); return -1; }
    Criteria crit;
    crit.maxmassfraction = maxmassfraction;
    crit.maxopticaldepth = maxopticaldepth;
    auto stats = [&](const double* q, int64_t) {
        NodeStats st;
        st.mass = vf.mass(q);
        return st;
    };
    return subdivide_tree(arity, ext, minlevel, maxlevel, crit, vf.total, stats, flags, cap, nnodes);
}

// Cell densities from a voxel density field: the mass in each cell's box over its volume (the exact
// mass-in-box densities of TreeDustGrid.cpp:647-660). Cartesian and tree grids.
int sg_grid_field_density(sg_grid* g, const double* field, int nx, int ny, int nz, double* rho_out)
{
    if (g->kind == GK_VORO) { set_error("Voronoi cells take the density at their site"); return -1; }
    if (nx < 1 || ny < 1 || nz < 1) { set_error("empty density field"); return -1; }
    double ext[6];
    std::vector<double> hx, hy, hz;
    int Nx = 0, Ny = 0, Nz = 0;
    if (g->kind == GK_CART) {
        Nx = g->cart.Nx; Ny = g->cart.Ny; Nz = g->cart.Nz;
        hx.resize(Nx + 1); hy.resize(Ny + 1); hz.resize(Nz + 1);
        (void)hipMemcpy(hx.data(), g->cart.xv, hx.size() * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(hy.data(), g->cart.yv, hy.size() * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(hz.data(), g->cart.zv, hz.size() * 8, hipMemcpyDeviceToHost);
        ext[0] = hx[0]; ext[1] = hy[0]; ext[2] = hz[0]; ext[3] = hx[Nx]; ext[4] = hy[Ny]; ext[5] = hz[Nz];
    } else {
        std::memcpy(ext, g->tree.ext, sizeof ext);
    }
    VoxelField vf(ext, field, nx, ny, nz);
    parallel_for(g->ncells, [&](int64_t m) {
        double q[6];
        if (g->kind == GK_CART) {
            int i = (int)(m / ((int64_t)Nz * Ny)), j = (int)((m / Nz) % Ny), k = (int)(m % Nz);
            q[0] = hx[i]; q[1] = hy[j]; q[2] = hz[k]; q[3] = hx[i + 1]; q[4] = hy[j + 1]; q[5] = hz[k + 1];
        } else {
            const TreeRec64& r = g->hrec64[m];
            for (int a = 0; a < 3; a++) { q[a] = r.lo[a]; q[a + 3] = r.hi[a]; }
        }
        double vol = (q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]);
        rho_out[m] = vf.mass(q) / vol;
    });
    return 0;
}

// DustSystem::setSampleDensityBody (DustSystem.cpp:145-172): the density of each cell is the
// mean of nsamples geometry densities at positions drawn uniformly in the cell's box
// (Box::fracpos), times scale. Cartesian and tree grids.
int sg_grid_sample_density(sg_grid* g, const double* geom16, double scale, int nsamples, uint32_t seed,
                           double* rho_out)
{
    GeomDensity dens(geom16);
    if (g->kind == GK_VORO) { set_error("Voronoi cells take the density at their site"); return -1; }
    std::vector<double> hx, hy, hz;
    int Nx = 0, Ny = 0, Nz = 0;
    if (g->kind == GK_CART) {
        Nx = g->cart.Nx; Ny = g->cart.Ny; Nz = g->cart.Nz;
        hx.resize(Nx + 1); hy.resize(Ny + 1); hz.resize(Nz + 1);
        (void)hipMemcpy(hx.data(), g->cart.xv, hx.size() * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(hy.data(), g->cart.yv, hy.size() * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(hz.data(), g->cart.zv, hz.size() * 8, hipMemcpyDeviceToHost);
    }
    parallel_for(g->ncells, [&](int64_t m) {
        double q[6];
        if (g->kind == GK_CART) {
            int i = (int)(m / ((int64_t)Nz * Ny)), j = (int)((m / Nz) % Ny), k = (int)(m % Nz);
            q[0] = hx[i]; q[1] = hy[j]; q[2] = hz[k]; q[3] = hx[i + 1]; q[4] = hy[j + 1]; q[5] = hz[k + 1];
        } else {
            const TreeRec64& r = g->hrec64[m];
            for (int a = 0; a < 3; a++) { q[a] = r.lo[a]; q[a + 3] = r.hi[a]; }
        }
        double sum = 0;
        for (int j = 0; j < nsamples; j++) {
            double u[3];
            sample3(seed, (uint64_t)m, (uint32_t)j, u);
            sum += dens(q[0] + u[0] * (q[3] - q[0]), q[1] + u[1] * (q[4] - q[1]), q[2] + u[2] * (q[5] - q[2]));
        }
        rho_out[m] = scale * sum / nsamples;
    });
    return 0;
}

} // extern "C"
