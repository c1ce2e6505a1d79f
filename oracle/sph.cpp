// sph.cpp -- CPU restatement of SKIRT7's SPH dust medium and of TreeDustGrid's subdivision over it.
//
// TEST INFRASTRUCTURE ONLY: the checker for the device tree builder in skirt7_amd/csrc/sg_treebuild.hip.
// Only tests/ load it (through liboracle.so); the product never does. Each routine cites the SKIRT7
// file:line it restates (paths relative to the SKIRT7 source tree).
//
//   SPHGasParticle (SKIRTcore/SPHGasParticle.cpp:14-110): the sampled erf table, the Gaussian mass in
//     a box, the cubic-spline metal density;
//   SPHGasParticleGrid (SKIRTcore/SPHGasParticleGrid.cpp:14-190): the 20^3 "smart" grid over the
//     particles and each cell's list of overlapping particles;
//   SPHDustDistribution (SKIRTcore/SPHDustDistribution.cpp:38-100,192-256): density, mass in a box and
//     total mass, with the temperature cut and the negative-mass guard (positions, smoothing lengths and
//     masses in the caller's units: the reference's pc and Msun factors are left to the caller);
//   TreeDustGrid::setupSelfBefore / subdivide (SKIRTcore/TreeDustGrid.cpp:49-230) with the box density
//     calculator (TreeNodeBoxDensityCalculator.cpp:18-60), children as OctTreeNode.cpp:38-56 and
//     BinTreeNode.cpp:40-77 (alternating directions) create them; TreeDustGrid::density
//     (TreeDustGrid.cpp:655-660).
//
// Parity pins (tests/test_sph_pins.py): the erf table against math.erf; a box reaching past every
// particle's support holds the total mass; a half space through a particle's centre holds half its
// mass; the truncated Gaussian against math.erf; the cubic-spline density integrates to the metal
// mass and peaks at 8/pi M Z / h^3; the smart grid's borders split the particles evenly; each cell's
// list against a brute-force sphere-box overlap; every split decision of the subdivision against the
// criteria evaluated on the node's box.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

// SPHGasParticle.cpp:14-45: erf sampled at 5001 points over [0, 2.5]
const int kErfN = 5000;
const double kErfXmax = 2.5;
struct ErfTable {
    std::vector<double> v;
    ErfTable()
    {
        v.resize(kErfN + 1);
        for (int i = 0; i <= kErfN; i++) v[i] = std::erf((i * kErfXmax) / kErfN);
    }
};
const ErfTable& erf_table()
{
    static ErfTable t;
    return t;
}
inline double myerf(double x)
{
    const std::vector<double>& e = erf_table().v;
    if (x >= kErfXmax) return 1.;
    if (x >= 0) return e[static_cast<int>(x * (kErfN / kErfXmax))];
    if (x > -kErfXmax) return -e[static_cast<int>(-x * (kErfN / kErfXmax))];
    return -1;
}

// SPHGasParticle.cpp:48-110
struct Particle {
    double rc[3], h, M, Z;
    double norm, rho0, rho2, s, MZ8;
    Particle(const double* r, double h_, double M_, double Z_) : h(h_), M(M_), Z(Z_)
    {
        for (int a = 0; a < 3; a++) rc[a] = r[a];
        norm = 1 / (h * h);
        rho0 = 8.0 / M_PI * M / (h * h * h) * Z;
        rho2 = rho0 * 2.0;
        s = 2.42 / h;
        MZ8 = M * Z / 8.;
    }
    double metalMass() const { return M * Z; }
    // box: xmin ymin zmin xmax ymax zmax
    double metalMassInBox(const double* b) const
    {
        double r1[3], r2[3];
        for (int a = 0; a < 3; a++) {
            r1[a] = s * (b[a] - rc[a]);
            r2[a] = s * (b[a + 3] - rc[a]);
        }
        return MZ8 * (myerf(r2[0]) - myerf(r1[0])) * (myerf(r2[1]) - myerf(r1[1])) * (myerf(r2[2]) - myerf(r1[2]));
    }
    double metalDensity(const double* r) const
    {
        double dx = r[0] - rc[0], dy = r[1] - rc[1], dz = r[2] - rc[2];
        double u2 = norm * (dx * dx + dy * dy + dz * dz);
        if (u2 < 1.0) {
            double u = std::sqrt(u2);
            double u1m = 1.0 - u;
            if (u < 0.5)
                return rho0 * (1.0 - 6.0 * u2 * u1m);
            else
                return rho2 * u1m * u1m * u1m;
        }
        return 0.0;
    }
};

// NR::locate_clip (Fundamentals/NR.hpp:96-151)
int locate_clip(const std::vector<double>& xv, double x)
{
    int n = (int)xv.size();
    if (x < xv[0]) return 0;
    int jl = -1, ju = n - 1;
    while (ju - jl > 1) {
        int jm = (ju + jl) >> 1;
        if (x < xv[jm]) ju = jm;
        else jl = jm;
    }
    return jl;
}

// SPHGasParticleGrid.cpp:24-65: grid borders along one axis from the particles' binned centres
void makegrid(const std::vector<Particle>& pv, int dir, int gridsize, std::vector<double>& grid)
{
    int n = (int)pv.size();
    double cmin = +std::numeric_limits<double>::infinity();
    double cmax = -std::numeric_limits<double>::infinity();
    for (int p = 0; p < n; p++) {
        cmin = std::min(cmin, pv[p].rc[dir] - pv[p].h);
        cmax = std::max(cmax, pv[p].rc[dir] + pv[p].h);
    }
    int nbins = gridsize * 100;
    double binwidth = (cmax - cmin) / nbins;
    std::vector<int> bins(nbins);
    for (int p = 0; p < n; p++) bins[std::min(nbins - 1, int((pv[p].rc[dir] - cmin) / binwidth))] += 1;
    grid.assign(gridsize + 1, 0.0);
    grid[0] = -std::numeric_limits<double>::infinity();
    int percell = n / gridsize;
    int cumul = 0;
    int gridindex = 1;
    for (int binindex = 0; binindex < nbins; binindex++) {
        cumul += bins[binindex];
        if (cumul > percell * gridindex) {
            grid[gridindex] = cmin + (binindex + 1) * binwidth;
            gridindex += 1;
            if (gridindex >= gridsize) break;
        }
    }
    grid[gridsize] = std::numeric_limits<double>::infinity();
}

inline double square(double v) { return v * v; }

// SPHGasParticleGrid.cpp:67-90 (Arvo's box-sphere test)
bool intersects(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax, double xc, double yc,
                double zc, double r)
{
    double squaredist = square(r);
    if (xc < xmin) squaredist -= square(xc - xmin);
    else if (xc > xmax) squaredist -= square(xc - xmax);
    if (yc < ymin) squaredist -= square(yc - ymin);
    else if (yc > ymax) squaredist -= square(yc - ymax);
    if (zc < zmin) squaredist -= square(zc - zmin);
    else if (zc > zmax) squaredist -= square(zc - zmax);
    return squaredist > 0.;
}

struct Medium {
    std::vector<Particle> pv;
    double fdust = 0;
    int m = 20;
    std::vector<double> g[3];
    std::vector<std::vector<int>> lists;

    int index(int i, int j, int k) const { return ((i * m) + j) * m + k; }

    // SPHGasParticleGrid.cpp:94-132
    void build_grid()
    {
        for (int a = 0; a < 3; a++) makegrid(pv, a, m, g[a]);
        lists.assign((size_t)m * m * m, {});
        int n = (int)pv.size();
        for (int p = 0; p < n; p++) {
            const double* rc = pv[p].rc;
            double h = pv[p].h;
            int i1 = locate_clip(g[0], rc[0] - h), i2 = locate_clip(g[0], rc[0] + h);
            int j1 = locate_clip(g[1], rc[1] - h), j2 = locate_clip(g[1], rc[1] + h);
            int k1 = locate_clip(g[2], rc[2] - h), k2 = locate_clip(g[2], rc[2] + h);
            for (int i = i1; i <= i2; i++)
                for (int j = j1; j <= j2; j++)
                    for (int k = k1; k <= k2; k++)
                        if (intersects(g[0][i], g[0][i + 1], g[1][j], g[1][j + 1], g[2][k], g[2][k + 1], rc[0], rc[1],
                                       rc[2], h))
                            lists[index(i, j, k)].push_back(p);
        }
    }
    // SPHGasParticleGrid.cpp:160-190: one cell's list, or the union in particle order (a std::set of
    // pointers into the particle vector)
    std::vector<int> particlesFor(const double* b) const
    {
        int i1 = locate_clip(g[0], b[0]), j1 = locate_clip(g[1], b[1]), k1 = locate_clip(g[2], b[2]);
        int i2 = locate_clip(g[0], b[3]), j2 = locate_clip(g[1], b[4]), k2 = locate_clip(g[2], b[5]);
        if (i1 == i2 && j1 == j2 && k1 == k2) return lists[index(i1, j1, k1)];
        std::vector<int> joined;
        for (int i = i1; i <= i2; i++)
            for (int j = j1; j <= j2; j++)
                for (int k = k1; k <= k2; k++) {
                    const auto& l = lists[index(i, j, k)];
                    joined.insert(joined.end(), l.begin(), l.end());
                }
        std::sort(joined.begin(), joined.end());
        joined.erase(std::unique(joined.begin(), joined.end()), joined.end());
        return joined;
    }
    // SPHDustDistribution.cpp:192-202
    double density(const double* r) const
    {
        int i = locate_clip(g[0], r[0]), j = locate_clip(g[1], r[1]), k = locate_clip(g[2], r[2]);
        const auto& l = lists[index(i, j, k)];
        double sum = 0.0;
        for (int p : l) sum += pv[p].metalDensity(r);
        sum *= fdust;
        return std::max(sum, 0.);
    }
    // SPHDustDistribution.cpp:226-236
    double massInBox(const double* b) const
    {
        std::vector<int> l = particlesFor(b);
        double sum = 0.0;
        for (int p : l) sum += pv[p].metalMassInBox(b);
        sum *= fdust;
        return std::max(sum, 0.);
    }
    // SPHDustDistribution.cpp:248-256
    double mass() const
    {
        double sum = 0.0;
        for (const auto& p : pv) sum += p.metalMass();
        sum *= fdust;
        return std::max(sum, 0.);
    }
};

std::string g_sph_err;

} // namespace

extern "C" {

const char* or_sph_last_error(void) { return g_sph_err.c_str(); }

// SPHDustDistribution::setupSelfBefore (SPHDustDistribution.cpp:38-100): rows of x y z h M Z T; a particle
// hotter than Tmax is ignored (both positive); a negative total gas or metal mass suppresses all dust.
void* or_sph_create(const double* rows, int64_t n, double fdust, double Tmax)
{
    auto* md = new Medium();
    md->fdust = fdust;
    double Mtot = 0, Mmetal = 0;
    for (int64_t i = 0; i < n; i++) {
        const double* r = rows + 7 * i;
        double T = r[6];
        if (T > 0 && Tmax > 0 && T > Tmax) continue;
        md->pv.emplace_back(r, r[3], r[4], r[5]);
        Mtot += r[4];
        Mmetal += r[4] * r[5];
    }
    if (Mtot < 0 || Mmetal < 0) md->pv.clear();
    md->build_grid();
    return md;
}

void or_sph_destroy(void* s) { delete (Medium*)s; }
int64_t or_sph_count(void* s) { return (int64_t)((Medium*)s)->pv.size(); }
double or_sph_mass(void* s) { return ((Medium*)s)->mass(); }
double or_sph_myerf(double x) { return myerf(x); }

void or_sph_grid(void* s, double* borders)
{
    auto* md = (Medium*)s;
    for (int a = 0; a < 3; a++)
        for (int i = 0; i <= md->m; i++) borders[a * (md->m + 1) + i] = md->g[a][i];
}

// the list of cell (i, j, k): its length, and its entries in out[0 .. min(length, cap))
int64_t or_sph_cell_list(void* s, int i, int j, int k, int32_t* out, int64_t cap)
{
    auto* md = (Medium*)s;
    const auto& l = md->lists[md->index(i, j, k)];
    for (int64_t t = 0; t < (int64_t)l.size() && t < cap; t++) out[t] = l[t];
    return (int64_t)l.size();
}

void or_sph_mass_in_box(void* s, const double* boxes, int64_t n, double* out)
{
    for (int64_t i = 0; i < n; i++) out[i] = ((Medium*)s)->massInBox(boxes + 6 * i);
}

void or_sph_density(void* s, const double* pos, int64_t n, double* out)
{
    for (int64_t i = 0; i < n; i++) out[i] = ((Medium*)s)->density(pos + 3 * i);
}

// TreeDustGrid::setupSelfBefore / subdivide with TreeNodeBoxDensityCalculator: nodes in creation order,
// one flag per node (1: subdivided); masses[l] receives the node's mass where the criteria evaluated it
// (-1 elsewhere) when masses is not null. Returns 0, or -5 when cap is too small (*nnodes: the count so far).
int or_sph_subdivide(void* s, int arity, const double ext[6], int minlevel, int maxlevel, double maxmassfraction,
                     double maxopticaldepth, uint8_t* flags, double* masses, int64_t cap, int64_t* nnodes)
{
    auto* md = (Medium*)s;
    if (!(arity == 8 || arity == 2)) { g_sph_err = "arity must be 8 or 2"; return -1; }
    const double kappaV = 2600.0;   // Units::kappaV (Units.cpp:29)
    const double total = md->mass();
    std::vector<double> box(ext, ext + 6);
    std::vector<int> level(1, 0);
    for (int64_t l = 0; l < (int64_t)level.size(); l++) {
        if (l >= cap) { *nnodes = (int64_t)level.size(); g_sph_err = "flag buffer too small"; return -5; }
        const double q[6] = {box[6 * l], box[6 * l + 1], box[6 * l + 2], box[6 * l + 3], box[6 * l + 4], box[6 * l + 5]};
        const int lev = level[l];
        bool split = false;
        if (masses) masses[l] = -1;
        if (lev <= minlevel) split = true;
        else if (lev < maxlevel) {
            const double mass = md->massInBox(q);
            if (masses) masses[l] = mass;
            bool need = (maxopticaldepth == 0 && maxmassfraction == 0);
            if (!need && maxmassfraction > 0 && mass / total >= maxmassfraction) need = true;
            if (!need && maxopticaldepth > 0) {
                double vol = (q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]);
                if (kappaV * mass / std::pow(vol, 2. / 3.) >= maxopticaldepth) need = true;
            }
            split = need;
        }
        flags[l] = split ? 1 : 0;
        if (!split) continue;
        if (arity == 8) {
            // OctTreeNode::createchildren_splitpoint at the centre: child index bits x, y, z
            const double c[3] = {0.5 * (q[0] + q[3]), 0.5 * (q[1] + q[4]), 0.5 * (q[2] + q[5])};
            for (int k = 0; k < 8; k++) {
                double nb[6];
                for (int a = 0; a < 3; a++) {
                    bool upper = (k >> a) & 1;
                    nb[a] = upper ? c[a] : q[a];
                    nb[a + 3] = upper ? q[a + 3] : c[a];
                }
                box.insert(box.end(), nb, nb + 6);
                level.push_back(lev + 1);
            }
        } else {
            const int a = lev % 3;
            const double c = 0.5 * (q[a] + q[a + 3]);
            double lo[6], hi[6];
            std::memcpy(lo, q, sizeof lo);
            std::memcpy(hi, q, sizeof hi);
            lo[a + 3] = c;
            hi[a] = c;
            box.insert(box.end(), lo, lo + 6);
            level.push_back(lev + 1);
            box.insert(box.end(), hi, hi + 6);
            level.push_back(lev + 1);
        }
    }
    *nnodes = (int64_t)level.size();
    return 0;
}

} // extern "C"
