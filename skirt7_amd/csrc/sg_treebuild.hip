// sg_treebuild.hip -- tree grids built on the device, and the SPH particle medium they are built over.
//
// TreeDustGrid::setupSelfBefore / subdivide (TreeDustGrid.cpp:49-230) creates nodes breadth first: node l's
// children are appended when l is visited, so the nodes of one level form one contiguous run and their
// children the next. The device builder takes one level at a time: a kernel decides every node of the level
// (one thread per node), an exclusive scan of the decisions numbers the subdivided nodes, and a second kernel
// writes their children's boxes, which are the next level. The flags come out in the reference's node order.
// A node's mass comes from
//   - an SPH particle set's mass in the node's box (TreeNodeBoxDensityCalculator with
//     SPHDustDistribution::massInBox, SPHDustDistribution.cpp:218-238), or
//   - density samples in the node's box (TreeNodeSampleDensityCalculator) of the SPH particles' density or
//     of disc geometries, at the positions of the host builder's counter-based stream keyed by node id.
// The particle set mirrors SPHDustDistribution (SPHDustDistribution.cpp:38-100,192-256), SPHGasParticle
// (SPHGasParticle.cpp) and SPHGasParticleGrid (SPHGasParticleGrid.cpp): the grid and its lists are built on
// the host, the particles are evaluated on the device.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/skirt7_amd.h"
#include "sg_host.hpp"
#include "sg_geom.hpp"

namespace sg {
namespace {

constexpr int kPG = 20;           // SPHDustDistribution.cpp:89: the particle grid is 20 x 20 x 20
constexpr int kErfN = 5000;       // SPHGasParticle.cpp:19-24: erf sampled at 5001 points over [0, 2.5]
constexpr double kErfXmax = 2.5;
constexpr int kScanBlock = 256, kScanItems = 8, kScanTile = kScanBlock * kScanItems;

// The particle set on the device, structure of arrays in particle order.
struct SphView {
    const double *cx, *cy, *cz, *h, *s, *mz8, *norm, *rho0;
    const uint32_t* range;   // each particle's cell range in the grid: i1, i2, j1, j2, k1, k2 at 5 bits each
    const int64_t* coff;     // kPG^3 + 1 list offsets
    const int32_t* cids;     // list entries (particle indices, ascending within a list)
    const double* erf;       // kErfN + 1 sampled values
    const double* grid;      // 3 x (kPG + 1) borders, infinite at both ends
    double fdust;
    int64_t n;               // particles
};

__host__ __device__ inline int cell_index(int i, int j, int k) { return (i * kPG + j) * kPG + k; }

// NR::locate_clip (NR.hpp:96-151) over the kPG + 1 borders of one axis
__device__ __forceinline__ int locate_grid(const double* g, double x)
{
    if (x < g[0]) return 0;
    int jl = -1, ju = kPG;
    while (ju - jl > 1) {
        const int jm = (ju + jl) >> 1;
        if (x < g[jm]) ju = jm;
        else jl = jm;
    }
    return jl;
}

// SPHGasParticle.cpp:34-44: the sampled erf
__device__ __forceinline__ double myerf(const double* e, double x)
{
    if (x >= kErfXmax) return 1.;
    if (x >= 0) return e[static_cast<int>(x * (kErfN / kErfXmax))];
    if (x > -kErfXmax) return -e[static_cast<int>(-x * (kErfN / kErfXmax))];
    return -1;
}

// SPHGasParticle::metalMassInBox (SPHGasParticle.cpp:88-96). A factor whose two ends saturate on one side is
// exactly zero, and a zero term leaves the caller's sum unchanged, so such particles return early.
__device__ __forceinline__ double part_mass_in_box(const SphView& v, int p, const double* q)
{
    const double s = v.s[p], cx = v.cx[p], cy = v.cy[p], cz = v.cz[p];
    const double r1x = s * (q[0] - cx), r2x = s * (q[3] - cx);
    const double r1y = s * (q[1] - cy), r2y = s * (q[4] - cy);
    const double r1z = s * (q[2] - cz), r2z = s * (q[5] - cz);
    if (r1x >= kErfXmax || r2x <= -kErfXmax || r1y >= kErfXmax || r2y <= -kErfXmax || r1z >= kErfXmax ||
        r2z <= -kErfXmax)
        return 0.0;
    return v.mz8[p] * (myerf(v.erf, r2x) - myerf(v.erf, r1x)) * (myerf(v.erf, r2y) - myerf(v.erf, r1y)) *
           (myerf(v.erf, r2z) - myerf(v.erf, r1z));
}

__host__ __device__ inline double square(double x) { return x * x; }

// SPHGasParticleGrid.cpp:67-90: Arvo's test of a box against a sphere
__host__ __device__ inline bool intersects(double xmin, double xmax, double ymin, double ymax, double zmin,
                                           double zmax, double xc, double yc, double zc, double r)
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

// Whether grid cell c is the first cell (in index order) of the range [i1..i2] x [j1..j2] x [k1..k2] whose
// list holds particle p: p's own range limits the search, and a cell's list holds p where the sphere meets it.
__device__ bool first_holder(const SphView& v, int p, int c, int i1, int i2, int j1, int j2, int k1, int k2)
{
    const uint32_t r = v.range[p];
    const int a1 = max(i1, (int)(r & 31)), a2 = min(i2, (int)((r >> 5) & 31));
    const int b1 = max(j1, (int)((r >> 10) & 31)), b2 = min(j2, (int)((r >> 15) & 31));
    const int c1 = max(k1, (int)((r >> 20) & 31)), c2 = min(k2, (int)((r >> 25) & 31));
    const double *gx = v.grid, *gy = v.grid + kPG + 1, *gz = v.grid + 2 * (kPG + 1);
    const double x = v.cx[p], y = v.cy[p], z = v.cz[p], h = v.h[p];
    for (int i = a1; i <= a2; i++)
        for (int j = b1; j <= b2; j++)
            for (int k = c1; k <= c2; k++)
                if (intersects(gx[i], gx[i + 1], gy[j], gy[j + 1], gz[k], gz[k + 1], x, y, z, h))
                    return cell_index(i, j, k) == c;
    return false;
}

// SPHDustDistribution::massInBox (SPHDustDistribution.cpp:226-236) over SPHGasParticleGrid::particlesFor(box)
// (SPHGasParticleGrid.cpp:160-190): a box within one cell takes that cell's list; a box over several cells the
// union of their lists in particle order. Up to eight cells, the lists are merged in particle order, so the
// sum's terms come in the reference's order; over more cells (large nodes near the root), each cell's list
// is taken in turn, a particle counted at the first cell holding it, and the terms come in cell order.
__device__ double sph_mass_in_box(const SphView& v, const double* q)
{
    const double *gx = v.grid, *gy = v.grid + kPG + 1, *gz = v.grid + 2 * (kPG + 1);
    const int i1 = locate_grid(gx, q[0]), j1 = locate_grid(gy, q[1]), k1 = locate_grid(gz, q[2]);
    const int i2 = locate_grid(gx, q[3]), j2 = locate_grid(gy, q[4]), k2 = locate_grid(gz, q[5]);
    const int ni = i2 - i1 + 1, nj = j2 - j1 + 1, nk = k2 - k1 + 1;
    const int K = ni * nj * nk;
    double sum = 0.0;
    if (K <= 8) {
        int64_t pos[8], end[8];
        int32_t head[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            pos[t] = 0;
            end[t] = 0;
            if (t < K) {
                const int c = cell_index(i1 + t / (nj * nk), j1 + (t / nk) % nj, k1 + t % nk);
                pos[t] = v.coff[c];
                end[t] = v.coff[c + 1];
            }
            head[t] = pos[t] < end[t] ? v.cids[pos[t]] : INT32_MAX;
        }
        for (;;) {
            int32_t p = INT32_MAX;
#pragma unroll
            for (int t = 0; t < 8; t++) p = min(p, head[t]);
            if (p == INT32_MAX) break;
            sum += part_mass_in_box(v, p, q);
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (head[t] == p) {
                    pos[t]++;
                    head[t] = pos[t] < end[t] ? v.cids[pos[t]] : INT32_MAX;
                }
        }
    } else {
        for (int i = i1; i <= i2; i++)
            for (int j = j1; j <= j2; j++)
                for (int k = k1; k <= k2; k++) {
                    const int c = cell_index(i, j, k);
                    for (int64_t t = v.coff[c]; t < v.coff[c + 1]; t++) {
                        const int p = v.cids[t];
                        if (first_holder(v, p, c, i1, i2, j1, j2, k1, k2)) sum += part_mass_in_box(v, p, q);
                    }
                }
    }
    sum *= v.fdust;
    return sum < 0. ? 0. : sum;
}

// SPHDustDistribution::density (SPHDustDistribution.cpp:192-202) with SPHGasParticle::metalDensity
// (SPHGasParticle.cpp:98-110) over the list of the cell holding the position.
__device__ double sph_density(const SphView& v, double x, double y, double z)
{
    const double *gx = v.grid, *gy = v.grid + kPG + 1, *gz = v.grid + 2 * (kPG + 1);
    const int c = cell_index(locate_grid(gx, x), locate_grid(gy, y), locate_grid(gz, z));
    double sum = 0.0;
    for (int64_t t = v.coff[c]; t < v.coff[c + 1]; t++) {
        const int p = v.cids[t];
        const double dx = x - v.cx[p], dy = y - v.cy[p], dz = z - v.cz[p];
        const double u2 = v.norm[p] * (dx * dx + dy * dy + dz * dz);
        if (u2 < 1.0) {
            const double u = sqrt(u2);
            const double u1m = 1.0 - u;
            if (u < 0.5) sum += v.rho0[p] * (1.0 - 6.0 * u2 * u1m);
            else sum += v.rho0[p] * 2.0 * u1m * u1m * u1m;
        }
    }
    sum *= v.fdust;
    return sum < 0. ? 0. : sum;
}

__global__ void k_sph_mass_in_box(const SphView v, const double* boxes, int64_t n, double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = sph_mass_in_box(v, boxes + 6 * i);
}

__global__ void k_sph_density(const SphView v, const double* pos, int64_t n, double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = sph_density(v, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

// TreeDustGrid::density (TreeDustGrid.cpp:655-660): the mass in each leaf's box over its volume (Box::volume)
__global__ void k_leaf_density(const SphView v, const double* boxes, int64_t n, double* rho)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* q = boxes + 6 * i;
    rho[i] = sph_mass_in_box(v, q) / ((q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]));
}

// ------------------------------------------------------------------------- level-by-level subdivision

enum NodeMass { NM_SPH_BOX = 0, NM_SPH_SAMPLES = 1, NM_DISC_SAMPLES = 2 };

struct SplitArgs {
    int mode;                 // NodeMass
    SphView sph;
    const DiscDev* discs;     // NM_DISC_SAMPLES: ncomp discs and their scales
    const double* scales;
    int ncomp, nsamples;
    uint32_t seed;
    int minlevel, maxlevel;
    double total, maxmassfraction, maxopticaldepth, maxdensdisp;
};

// TreeDustGrid::subdivide (TreeDustGrid.cpp:166-230) for the nodes of one level: a node at or below minlevel
// is subdivided, one at maxlevel is not, one in between when no criterion is set or when a set criterion is
// not met (mass fraction, optical depth kappaV M / V^(2/3) with kappaV = 2600 m2/kg, density dispersion
// (max - min) / max of the samples). Writes split[i] and the node's mass (-1 where not evaluated).
__global__ void __launch_bounds__(256) k_node_split(const SplitArgs a, const double* boxes, int64_t n, int level,
                                                     int64_t first_id, uint8_t* split, double* mass_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* q = boxes + 6 * i;
    if (level <= a.minlevel || level >= a.maxlevel) {
        split[i] = level <= a.minlevel ? 1 : 0;
        mass_out[i] = -1;
        return;
    }
    const double vol = (q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]);
    double mass, lo = -1, hi = -1;
    if (a.mode == NM_SPH_BOX) {
        mass = sph_mass_in_box(a.sph, q);
    } else {
        double sum = 0;
        for (int j = 0; j < a.nsamples; j++) {
            double u[3];
            setup_sample3(a.seed, (uint64_t)(first_id + i), (uint32_t)j, u);
            const double x = q[0] + u[0] * (q[3] - q[0]), y = q[1] + u[1] * (q[4] - q[1]), z = q[2] + u[2] * (q[5] - q[2]);
            double d = 0;
            if (a.mode == NM_SPH_SAMPLES) d = sph_density(a.sph, x, y, z);
            else
                for (int h = 0; h < a.ncomp; h++) d += a.scales[h] * disc_density(a.discs[h], x, y, z);
            sum += d;
            lo = j ? (d < lo ? d : lo) : d;
            hi = j ? (hi < d ? d : hi) : d;
        }
        mass = sum / a.nsamples * vol;
    }
    mass_out[i] = mass;
    bool need = a.maxmassfraction == 0 && a.maxopticaldepth == 0 && a.maxdensdisp == 0;
    if (!need && a.maxmassfraction > 0 && mass / a.total >= a.maxmassfraction) need = true;
    if (!need && a.maxopticaldepth > 0 && 2600.0 * mass / pow(vol, 2. / 3.) >= a.maxopticaldepth) need = true;
    if (!need && a.maxdensdisp > 0) {
        const double disp = hi > 0 ? (hi - lo) / hi : 0;
        if (disp >= a.maxdensdisp) need = true;
    }
    split[i] = need ? 1 : 0;
}

// Exclusive scan of the split flags, first pass: each tile's local ranks and its total.
__global__ void __launch_bounds__(kScanBlock) k_scan_tiles(const uint8_t* split, int64_t n, int32_t* rank,
                                                            int64_t* tile_sum)
{
    __shared__ int32_t s_part[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems], run = 0;
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
        v[t] = (base + t < n) ? split[base + t] : 0;
        run += v[t];
    }
    s_part[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const int32_t add = threadIdx.x >= (unsigned)off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += add;
        __syncthreads();
    }
    int32_t acc = s_part[threadIdx.x] - run;   // exclusive prefix of this thread's items within the tile
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
        if (base + t < n) rank[base + t] = acc;
        acc += v[t];
    }
    if (threadIdx.x == kScanBlock - 1) tile_sum[blockIdx.x] = s_part[threadIdx.x];
}

// The children of the subdivided nodes of one level, in the reference's order: OctTreeNode::createchildren
// (OctTreeNode.cpp:38-56, split at the centre; child index bits x, y, z) and BinTreeNode::createchildren
// (BinTreeNode.cpp:40-77, split at the middle of the axis level % 3).
__global__ void k_children(const double* boxes, const uint8_t* split, const int32_t* rank, const int64_t* tile_off,
                           int64_t n, int arity, int level, double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !split[i]) return;
    const int64_t r = tile_off[i / kScanTile] + rank[i];
    const double* q = boxes + 6 * i;
    double* o = out + 6 * arity * r;
    if (arity == 8) {
        const double c[3] = {0.5 * (q[0] + q[3]), 0.5 * (q[1] + q[4]), 0.5 * (q[2] + q[5])};
        for (int k = 0; k < 8; k++)
            for (int ax = 0; ax < 3; ax++) {
                const bool upper = (k >> ax) & 1;
                o[6 * k + ax] = upper ? c[ax] : q[ax];
                o[6 * k + ax + 3] = upper ? q[ax + 3] : c[ax];
            }
    } else {
        const int ax = level % 3;
        const double c = 0.5 * (q[ax] + q[ax + 3]);
        for (int t = 0; t < 6; t++) o[t] = o[6 + t] = q[t];
        o[ax + 3] = c;
        o[6 + ax] = c;
    }
}

// ------------------------------------------------------------------------- subdivision over particle lists
//
// Each node of a level carries the list, in particle order, of the particles that add a nonzero term to its
// mass in box: those in the union of the grid lists over the node's range of grid cells (the reference's
// candidates) whose Gaussian reaches the node's box on every axis (the others add exactly zero, which leaves
// the sum as it is). A child's list is the part of its parent's list that passes the same test for the
// child's box: the child's box and range lie within the parent's, so no particle is missed, and the order is
// kept. A node's mass is then the sum over its list in order: the reference's sum, term for term.

// The grid cell range of a box (NR::locate_clip of its corners)
struct CellRange {
    int i1, i2, j1, j2, k1, k2;
};
__device__ __forceinline__ CellRange box_range(const SphView& v, const double* q)
{
    const double *gx = v.grid, *gy = v.grid + kPG + 1, *gz = v.grid + 2 * (kPG + 1);
    return CellRange{locate_grid(gx, q[0]), locate_grid(gx, q[3]), locate_grid(gy, q[1]),
                     locate_grid(gy, q[4]), locate_grid(gz, q[2]), locate_grid(gz, q[5])};
}

// Whether particle p adds a nonzero term to the mass in box q with grid range R: its Gaussian reaches q on
// every axis (the negation of part_mass_in_box's early return) and a grid list of a cell in R holds it.
__device__ __forceinline__ bool contributes(const SphView& v, int p, const double* q, const CellRange& R)
{
    const double s = v.s[p], cx = v.cx[p], cy = v.cy[p], cz = v.cz[p];
    if (s * (q[0] - cx) >= kErfXmax || s * (q[3] - cx) <= -kErfXmax || s * (q[1] - cy) >= kErfXmax ||
        s * (q[4] - cy) <= -kErfXmax || s * (q[2] - cz) >= kErfXmax || s * (q[5] - cz) <= -kErfXmax)
        return false;
    const uint32_t r = v.range[p];
    const int a1 = max(R.i1, (int)(r & 31)), a2 = min(R.i2, (int)((r >> 5) & 31));
    const int b1 = max(R.j1, (int)((r >> 10) & 31)), b2 = min(R.j2, (int)((r >> 15) & 31));
    const int c1 = max(R.k1, (int)((r >> 20) & 31)), c2 = min(R.k2, (int)((r >> 25) & 31));
    const double *gx = v.grid, *gy = v.grid + kPG + 1, *gz = v.grid + 2 * (kPG + 1);
    const double h = v.h[p];
    for (int i = a1; i <= a2; i++)
        for (int j = b1; j <= b2; j++)
            for (int k = c1; k <= c2; k++)
                if (intersects(gx[i], gx[i + 1], gy[j], gy[j + 1], gz[k], gz[k + 1], cx, cy, cz, h)) return true;
    return false;
}

// One wavefront per segment of a parent's list: the part of the segment that belongs to a child's list,
// counted (out null: cnt[segment]) or written in order at out + off[segment] by ballots. A long list is cut
// into segments of kSeg entries, one wavefront each, so that the few large nodes near the root keep the whole
// device busy; seg_child null means one segment per child, the whole list. The root's parent list is every
// particle (parent null).
constexpr int64_t kSeg = 4096;
__global__ void __launch_bounds__(256) k_child_lists(const SphView v, const double* cboxes, int64_t nseg,
                                                      const int32_t* seg_child, const int64_t* seg_first,
                                                      const int32_t* parent, const int64_t* poff,
                                                      const int32_t* pids, int64_t* cnt, const int64_t* off,
                                                      int32_t* out)
{
    const int64_t sg = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (sg >= nseg) return;   // the whole wavefront leaves together
    const int lane = threadIdx.x & 63;
    const int64_t w = seg_child ? seg_child[sg] : sg;
    const double* q = cboxes + 6 * w;
    const CellRange R = box_range(v, q);
    int64_t start = 0, len = v.n;
    if (parent) {
        const int32_t par = parent[w];
        start = poff[par];
        len = poff[par + 1] - start;
    }
    if (seg_child) {
        start += seg_first[sg];
        len = len - seg_first[sg] < kSeg ? len - seg_first[sg] : kSeg;
    }
    int64_t at = out ? off[sg] : 0, c = 0;
    for (int64_t base = 0; base < len; base += 64) {
        const int64_t idx = base + lane;
        int32_t p = 0;
        bool keep = false;
        if (idx < len) {
            p = pids ? pids[start + idx] : (int32_t)(start + idx);
            keep = contributes(v, p, q, R);
        }
        const uint64_t mask = __ballot(keep);
        if (out && keep) out[at + c + __popcll(mask & ((1ull << lane) - 1))] = p;
        c += __popcll(mask);
    }
    if (!out && lane == 0) cnt[sg] = c;
}

// Each child's list offset: the offset of its first segment (child_seg0 null: segment = child).
__global__ void k_child_offsets(const int64_t* segoff, const int64_t* child_seg0, int64_t nn, int64_t* off)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w <= nn) off[w] = segoff[child_seg0[w]];
}

// The nodes of one level with their lists: the mass (the ordered sum; every node beyond minlevel, those at
// maxlevel included, whose masses give the cells' densities) and the decision of TreeDustGrid::subdivide
// with the box calculator (mass fraction and optical-depth criteria).
__global__ void __launch_bounds__(256) k_list_split(const SplitArgs a, const double* boxes, int64_t n, int level,
                                                     const int64_t* off, const int32_t* ids, uint8_t* split,
                                                     double* mass_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (level <= a.minlevel) {
        split[i] = 1;
        mass_out[i] = -1;
        return;
    }
    const double* q = boxes + 6 * i;
    double sum = 0.0;
    for (int64_t t = off[i]; t < off[i + 1]; t++) sum += part_mass_in_box(a.sph, ids[t], q);
    sum *= a.sph.fdust;
    const double mass = sum < 0. ? 0. : sum;
    mass_out[i] = mass;
    if (level >= a.maxlevel) {
        split[i] = 0;
        return;
    }
    const double vol = (q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]);
    bool need = a.maxmassfraction == 0 && a.maxopticaldepth == 0;
    if (!need && a.maxmassfraction > 0 && mass / a.total >= a.maxmassfraction) need = true;
    if (!need && a.maxopticaldepth > 0 && 2600.0 * mass / pow(vol, 2. / 3.) >= a.maxopticaldepth) need = true;
    split[i] = need ? 1 : 0;
}

// The same with one wavefront per node, for levels whose lists are long: the lanes evaluate 64
// consecutive terms at once and every lane then adds them in list order, so the sum is the same.
__global__ void __launch_bounds__(256) k_list_split_wave(const SplitArgs a, const double* boxes, int64_t n,
                                                          int level, const int64_t* off, const int32_t* ids,
                                                          uint8_t* split, double* mass_out)
{
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;   // the whole wavefront leaves together
    const int lane = threadIdx.x & 63;
    if (level <= a.minlevel) {
        if (lane == 0) {
            split[i] = 1;
            mass_out[i] = -1;
        }
        return;
    }
    const double* q = boxes + 6 * i;
    const int64_t start = off[i], len = off[i + 1] - start;
    double sum = 0.0;
    for (int64_t base = 0; base < len; base += 64) {
        const int64_t idx = base + lane;
        const double t = idx < len ? part_mass_in_box(a.sph, ids[start + idx], q) : 0.0;
        const int m = len - base < 64 ? (int)(len - base) : 64;
        for (int j = 0; j < m; j++) sum += __shfl(t, j);
    }
    if (lane != 0) return;
    sum *= a.sph.fdust;
    const double mass = sum < 0. ? 0. : sum;
    mass_out[i] = mass;
    if (level >= a.maxlevel) {
        split[i] = 0;
        return;
    }
    const double vol = (q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]);
    bool need = a.maxmassfraction == 0 && a.maxopticaldepth == 0;
    if (!need && a.maxmassfraction > 0 && mass / a.total >= a.maxmassfraction) need = true;
    if (!need && a.maxopticaldepth > 0 && 2600.0 * mass / pow(vol, 2. / 3.) >= a.maxopticaldepth) need = true;
    split[i] = need ? 1 : 0;
}

// Exclusive scan of 64-bit counts, first pass: each tile's local offsets and its total.
__global__ void __launch_bounds__(kScanBlock) k_scan_tiles64(const int64_t* in, int64_t n, int64_t* out,
                                                              int64_t* tile_sum)
{
    __shared__ int64_t s_part[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems], run = 0;
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
        v[t] = (base + t < n) ? in[base + t] : 0;
        run += v[t];
    }
    s_part[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const int64_t add = threadIdx.x >= (unsigned)off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t acc = s_part[threadIdx.x] - run;
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
        if (base + t < n) out[base + t] = acc;
        acc += v[t];
    }
    if (threadIdx.x == kScanBlock - 1) tile_sum[blockIdx.x] = s_part[threadIdx.x];
}

__global__ void k_add_tile_offsets(int64_t* out, int64_t n, const int64_t* tile_off)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tile_off[i / kScanTile];
}

// The parent (index in its level) of each child of the next level.
__global__ void k_parents(const uint8_t* split, const int32_t* rank, const int64_t* tile_off, int64_t n, int arity,
                          int32_t* parent)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !split[i]) return;
    const int64_t r = tile_off[i / kScanTile] + rank[i];
    for (int k = 0; k < arity; k++) parent[arity * r + k] = (int32_t)i;
}

// DustSystem::setSampleDensityBody (DustSystem.cpp:145-172) on the device: each cell's density is scale times
// the mean of nsamples disc densities at the host sampler's counter-based positions in the cell's box.
__global__ void k_sample_boxes(const double* boxes, int64_t n, const DiscDev d, double scale, int nsamples,
                               uint32_t seed, double* rho)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double* q = boxes + 6 * m;
    double sum = 0;
    for (int j = 0; j < nsamples; j++) {
        double u[3];
        setup_sample3(seed, (uint64_t)m, (uint32_t)j, u);
        sum += disc_density(d, q[0] + u[0] * (q[3] - q[0]), q[1] + u[1] * (q[4] - q[1]), q[2] + u[2] * (q[5] - q[2]));
    }
    rho[m] = scale * sum / nsamples;
}

// ------------------------------------------------------------------------- k nearest neighbours

// Distance from each point to its k-th nearest other point, over a bucket grid: rings of buckets around the
// point's own are searched until the k-th smallest distance found is within the rings' reach.
__global__ void __launch_bounds__(256) k_knn(const double* pos, int64_t n, const int32_t* order, const int64_t* boff,
                                             int nb, double x0, double y0, double z0, double wx, double wy, double wz,
                                             int kth, double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    const int bi = min(nb - 1, max(0, (int)((px - x0) / wx)));
    const int bj = min(nb - 1, max(0, (int)((py - y0) / wy)));
    const int bk = min(nb - 1, max(0, (int)((pz - z0) / wz)));
    const double wmin = fmin(wx, fmin(wy, wz));
    double best[64];   // the kth smallest squared distances so far, ascending
    int cnt = 0;
    for (int ring = 0; ring <= nb; ring++) {
        for (int a = max(0, bi - ring); a <= min(nb - 1, bi + ring); a++)
            for (int b = max(0, bj - ring); b <= min(nb - 1, bj + ring); b++)
                for (int c = max(0, bk - ring); c <= min(nb - 1, bk + ring); c++) {
                    if (max(abs(a - bi), max(abs(b - bj), abs(c - bk))) != ring) continue;
                    const int64_t q = ((int64_t)a * nb + b) * nb + c;
                    for (int64_t t = boff[q]; t < boff[q + 1]; t++) {
                        const int32_t j = order[t];
                        if (j == i) continue;
                        const double dx = pos[3 * j] - px, dy = pos[3 * j + 1] - py, dz = pos[3 * j + 2] - pz;
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        if (cnt == kth && !(d2 < best[kth - 1])) continue;
                        int at = cnt < kth ? cnt++ : kth - 1;
                        while (at > 0 && best[at - 1] > d2) {
                            best[at] = best[at - 1];
                            at--;
                        }
                        best[at] = d2;
                    }
                }
        const double reach = ring * wmin;
        if (cnt == kth && best[kth - 1] <= reach * reach) break;
    }
    out[i] = cnt == kth ? sqrt(best[kth - 1]) : -1.0;
}

int64_t blocks(int64_t n, int b) { return (n + b - 1) / b; }

void ensure_device()
{
    int d = -1;
    if (hipGetDevice(&d) != hipSuccess || d < 0) {
        set_error("no HIP device");
        throw Fail{-3};
    }
}

} // namespace
} // namespace sg

using namespace sg;

// The particle set: host copies for the grid and the lists, device arrays for evaluation.
struct sg_sph {
    int64_t n = 0;
    double fdust = 0, mass = 0;
    std::vector<double> grid;       // 3 x (kPG + 1)
    std::vector<int64_t> coff;
    std::vector<int32_t> cids;
    DevBuf cx, cy, cz, h, s, mz8, norm, rho0, range, dcoff, dcids, erf, dgrid;
    SphView view() const
    {
        SphView v;
        v.cx = cx.as<double>(); v.cy = cy.as<double>(); v.cz = cz.as<double>(); v.h = h.as<double>();
        v.s = s.as<double>(); v.mz8 = mz8.as<double>(); v.norm = norm.as<double>(); v.rho0 = rho0.as<double>();
        v.range = range.as<uint32_t>();
        v.coff = dcoff.as<int64_t>();
        v.cids = dcids.as<int32_t>();
        v.erf = erf.as<double>();
        v.grid = dgrid.as<double>();
        v.fdust = fdust;
        v.n = n;
        return v;
    }
};

namespace sg {
namespace {

// NR::locate_clip on the host
int locate_grid_host(const double* g, double x)
{
    if (x < g[0]) return 0;
    int jl = -1, ju = kPG;
    while (ju - jl > 1) {
        const int jm = (ju + jl) >> 1;
        if (x < g[jm]) ju = jm;
        else jl = jm;
    }
    return jl;
}

// SPHGasParticleGrid.cpp:24-65 (makegrid): borders along one axis placed where the cumulative count of the
// particles' centres, binned at 100 bins per cell over the range of their spheres, passes each multiple of
// n / kPG; infinite at both ends. A centre on the top of the range goes into the last bin.
void make_axis(const std::vector<double>& c, const std::vector<double>& hv, double* grid)
{
    const int64_t n = (int64_t)c.size();
    double cmin = +std::numeric_limits<double>::infinity();
    double cmax = -std::numeric_limits<double>::infinity();
    for (int64_t p = 0; p < n; p++) {
        cmin = std::min(cmin, c[p] - hv[p]);
        cmax = std::max(cmax, c[p] + hv[p]);
    }
    const int nbins = kPG * 100;
    const double binwidth = (cmax - cmin) / nbins;
    std::vector<int> bins(nbins);
    for (int64_t p = 0; p < n; p++) bins[std::min(nbins - 1, int((c[p] - cmin) / binwidth))] += 1;
    for (int i = 0; i <= kPG; i++) grid[i] = 0.0;
    grid[0] = -std::numeric_limits<double>::infinity();
    const int percell = (int)(n / kPG);
    int cumul = 0, gridindex = 1;
    for (int binindex = 0; binindex < nbins; binindex++) {
        cumul += bins[binindex];
        if (cumul > percell * gridindex) {
            grid[gridindex] = cmin + (binindex + 1) * binwidth;
            gridindex += 1;
            if (gridindex >= kPG) break;
        }
    }
    grid[kPG] = std::numeric_limits<double>::infinity();
}

const std::vector<double>& erf_table()
{
    static std::vector<double> t = [] {
        std::vector<double> v(kErfN + 1);
        for (int i = 0; i <= kErfN; i++) v[i] = std::erf((i * kErfXmax) / kErfN);
        return v;
    }();
    return t;
}

// Runs the level-by-level subdivision with node decisions by k_node_split. When the nodes outgrow cap, the
// build runs on without writing and returns -5 with the full count in *nnodes.
int subdivide_device(int arity, const double ext[6], const SplitArgs& a, uint8_t* flags, double* masses, int64_t cap,
                     int64_t* nnodes)
{
    DevBuf cur, nxt, split, rank, tsum, toff, mass;
    cur.upload(ext, 6);
    int64_t begin = 0, n = 1;
    int level = 0;
    std::vector<int64_t> ts;
    bool overflow = false;
    while (n > 0) {
        if (begin + n > cap) overflow = true;
        split.alloc(n);
        mass.alloc(n * sizeof(double));
        hipLaunchKernelGGL(k_node_split, dim3(blocks(n, 256)), dim3(256), 0, 0, a, cur.as<double>(), n, level, begin,
                           split.as<uint8_t>(), mass.as<double>());
        SG_HIP(hipGetLastError());
        const int64_t tiles = blocks(n, kScanTile);
        rank.alloc(n * sizeof(int32_t));
        tsum.alloc(tiles * sizeof(int64_t));
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanBlock), 0, 0, split.as<uint8_t>(), n,
                           rank.as<int32_t>(), tsum.as<int64_t>());
        SG_HIP(hipGetLastError());
        if (!overflow) {
            SG_HIP(hipMemcpy(flags + begin, split.p, n, hipMemcpyDeviceToHost));
            if (masses) SG_HIP(hipMemcpy(masses + begin, mass.p, n * sizeof(double), hipMemcpyDeviceToHost));
        }
        ts.resize(tiles);
        SG_HIP(hipMemcpy(ts.data(), tsum.p, tiles * sizeof(int64_t), hipMemcpyDeviceToHost));
        int64_t nsplit = 0;
        for (int64_t t = 0; t < tiles; t++) {
            const int64_t v = ts[t];
            ts[t] = nsplit;
            nsplit += v;
        }
        const int64_t nn = nsplit * arity;
        if (nn > 0) {
            toff.upload(ts);
            nxt.alloc(nn * 6 * sizeof(double));
            hipLaunchKernelGGL(k_children, dim3(blocks(n, 256)), dim3(256), 0, 0, cur.as<double>(), split.as<uint8_t>(),
                               rank.as<int32_t>(), toff.as<int64_t>(), n, arity, level, nxt.as<double>());
            SG_HIP(hipGetLastError());
            std::swap(cur.p, nxt.p);
            std::swap(cur.bytes, nxt.bytes);
        }
        begin += n;
        n = nn;
        level++;
    }
    SG_HIP(hipDeviceSynchronize());
    *nnodes = begin;
    if (overflow) {
        set_error("flag buffer too small");
        return -5;
    }
    return 0;
}

// Host-side exclusive scan of 64-bit counts in place over n entries; returns the total.
int64_t scan64(DevBuf& vals, int64_t n, DevBuf& tmp)
{
    const int64_t tiles = blocks(n, kScanTile);
    DevBuf tsum, toff;
    tsum.alloc(tiles * sizeof(int64_t));
    tmp.alloc(n * sizeof(int64_t));
    hipLaunchKernelGGL(k_scan_tiles64, dim3(tiles), dim3(kScanBlock), 0, 0, vals.as<int64_t>(), n, tmp.as<int64_t>(),
                       tsum.as<int64_t>());
    SG_HIP(hipGetLastError());
    std::vector<int64_t> ts(tiles);
    SG_HIP(hipMemcpy(ts.data(), tsum.p, tiles * sizeof(int64_t), hipMemcpyDeviceToHost));
    int64_t total = 0;
    for (int64_t t = 0; t < tiles; t++) {
        const int64_t v = ts[t];
        ts[t] = total;
        total += v;
    }
    toff.upload(ts);
    hipLaunchKernelGGL(k_add_tile_offsets, dim3(blocks(n, 256)), dim3(256), 0, 0, tmp.as<int64_t>(), n,
                       toff.as<int64_t>());
    SG_HIP(hipGetLastError());
    std::swap(vals.p, tmp.p);
    std::swap(vals.bytes, tmp.bytes);
    return total;
}

// The subdivision with masses in box, level by level over the nodes' particle lists (see k_child_lists).
// Every node beyond minlevel gets its mass in masses (when not null); the others -1. When the nodes outgrow
// cap, the build runs on without writing and returns -5 with the full count in *nnodes.
int subdivide_lists(int arity, const double ext[6], const SplitArgs& a, uint8_t* flags, double* masses, int64_t cap,
                    int64_t* nnodes)
{
    const int64_t kWaveList = 256;   // levels whose lists average more entries sum with a wavefront per node
    DevBuf cur, nxt, split, rank, tsum, toff, mass, off, ids, noff, nids, par, tmp, segc, segf, seg0, segoff;
    std::vector<uint8_t> hsplit;
    std::vector<int64_t> hpoff;
    cur.upload(ext, 6);
    // Children's lists of one level: segmented (host-built table) where lists are long, else one wavefront
    // per child. plen(w) gives the length of child w's parent list.
    auto make_lists = [&](const double* cboxes, int64_t nn, const int32_t* parent, const int64_t* poff,
                          const int32_t* pids, const std::vector<int64_t>& plen) -> int64_t {
        std::vector<int32_t> hsc;
        std::vector<int64_t> hsf, hs0(nn + 1);
        bool segmented = false;
        for (int64_t w = 0; w < nn && !segmented; w++) segmented = plen[w] > kSeg;
        int64_t nseg = nn;
        if (segmented) {
            for (int64_t w = 0; w < nn; w++) {
                hs0[w] = (int64_t)hsc.size();
                for (int64_t f = 0; f == 0 || f < plen[w]; f += kSeg) {
                    hsc.push_back((int32_t)w);
                    hsf.push_back(f);
                }
            }
            nseg = (int64_t)hsc.size();
            hs0[nn] = nseg;
            segc.upload(hsc);
            segf.upload(hsf);
            seg0.upload(hs0);
        }
        segoff.alloc((nseg + 1) * sizeof(int64_t));
        const int64_t zero = 0;
        SG_HIP(hipMemcpy(segoff.as<int64_t>() + nseg, &zero, sizeof zero, hipMemcpyHostToDevice));
        const int32_t* sc = segmented ? segc.as<int32_t>() : nullptr;
        const int64_t* sf = segmented ? segf.as<int64_t>() : nullptr;
        hipLaunchKernelGGL(k_child_lists, dim3(blocks(nseg, 4)), dim3(256), 0, 0, a.sph, cboxes, nseg, sc, sf, parent,
                           poff, pids, segoff.as<int64_t>(), nullptr, nullptr);
        SG_HIP(hipGetLastError());
        const int64_t total = scan64(segoff, nseg + 1, tmp);   // entry nseg was 0: it is the total
        nids.alloc(std::max<int64_t>(total, 1) * sizeof(int32_t));
        hipLaunchKernelGGL(k_child_lists, dim3(blocks(nseg, 4)), dim3(256), 0, 0, a.sph, cboxes, nseg, sc, sf, parent,
                           poff, pids, nullptr, segoff.as<int64_t>(), nids.as<int32_t>());
        SG_HIP(hipGetLastError());
        noff.alloc((nn + 1) * sizeof(int64_t));
        if (segmented) {
            hipLaunchKernelGGL(k_child_offsets, dim3(blocks(nn + 1, 256)), dim3(256), 0, 0, segoff.as<int64_t>(),
                               seg0.as<int64_t>(), nn, noff.as<int64_t>());
            SG_HIP(hipGetLastError());
        } else {
            std::swap(noff.p, segoff.p);
            std::swap(noff.bytes, segoff.bytes);
        }
        return total;
    };
    // the root's list, from all particles
    int64_t entries = make_lists(cur.as<double>(), 1, nullptr, nullptr, nullptr, std::vector<int64_t>{a.sph.n});
    std::swap(off.p, noff.p);
    std::swap(off.bytes, noff.bytes);
    std::swap(ids.p, nids.p);
    std::swap(ids.bytes, nids.bytes);
    int64_t begin = 0, n = 1;
    int level = 0;
    bool overflow = false;
    std::vector<int64_t> ts;
    while (n > 0) {
        if (begin + n > cap) overflow = true;
        split.alloc(n);
        mass.alloc(n * sizeof(double));
        if (entries > kWaveList * n)
            hipLaunchKernelGGL(k_list_split_wave, dim3(blocks(n, 4)), dim3(256), 0, 0, a, cur.as<double>(), n, level,
                               off.as<int64_t>(), ids.as<int32_t>(), split.as<uint8_t>(), mass.as<double>());
        else
            hipLaunchKernelGGL(k_list_split, dim3(blocks(n, 256)), dim3(256), 0, 0, a, cur.as<double>(), n, level,
                               off.as<int64_t>(), ids.as<int32_t>(), split.as<uint8_t>(), mass.as<double>());
        SG_HIP(hipGetLastError());
        const int64_t tiles = blocks(n, kScanTile);
        rank.alloc(n * sizeof(int32_t));
        tsum.alloc(tiles * sizeof(int64_t));
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanBlock), 0, 0, split.as<uint8_t>(), n,
                           rank.as<int32_t>(), tsum.as<int64_t>());
        SG_HIP(hipGetLastError());
        hsplit.resize(n);
        SG_HIP(hipMemcpy(hsplit.data(), split.p, n, hipMemcpyDeviceToHost));
        if (!overflow) {
            std::memcpy(flags + begin, hsplit.data(), n);
            if (masses) SG_HIP(hipMemcpy(masses + begin, mass.p, n * sizeof(double), hipMemcpyDeviceToHost));
        }
        ts.resize(tiles);
        SG_HIP(hipMemcpy(ts.data(), tsum.p, tiles * sizeof(int64_t), hipMemcpyDeviceToHost));
        int64_t nsplit = 0;
        for (int64_t t = 0; t < tiles; t++) {
            const int64_t v = ts[t];
            ts[t] = nsplit;
            nsplit += v;
        }
        const int64_t nn = nsplit * arity;
        if (nn > 0) {
            SG_REQUIRE(n < ((int64_t)1 << 31), "too many nodes in one tree level");
            toff.upload(ts);
            nxt.alloc(nn * 6 * sizeof(double));
            hipLaunchKernelGGL(k_children, dim3(blocks(n, 256)), dim3(256), 0, 0, cur.as<double>(), split.as<uint8_t>(),
                               rank.as<int32_t>(), toff.as<int64_t>(), n, arity, level, nxt.as<double>());
            SG_HIP(hipGetLastError());
            par.alloc(nn * sizeof(int32_t));
            hipLaunchKernelGGL(k_parents, dim3(blocks(n, 256)), dim3(256), 0, 0, split.as<uint8_t>(), rank.as<int32_t>(),
                               toff.as<int64_t>(), n, arity, par.as<int32_t>());
            SG_HIP(hipGetLastError());
            // the parents' list lengths, per child
            hpoff.resize(n + 1);
            SG_HIP(hipMemcpy(hpoff.data(), off.p, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
            std::vector<int64_t> plen;
            plen.reserve(nn);
            for (int64_t i = 0; i < n; i++)
                if (hsplit[i])
                    for (int k = 0; k < arity; k++) plen.push_back(hpoff[i + 1] - hpoff[i]);
            entries = make_lists(nxt.as<double>(), nn, par.as<int32_t>(), off.as<int64_t>(), ids.as<int32_t>(), plen);
            std::swap(cur.p, nxt.p);
            std::swap(cur.bytes, nxt.bytes);
            std::swap(off.p, noff.p);
            std::swap(off.bytes, noff.bytes);
            std::swap(ids.p, nids.p);
            std::swap(ids.bytes, nids.bytes);
        }
        begin += n;
        n = nn;
        level++;
    }
    SG_HIP(hipDeviceSynchronize());
    *nnodes = begin;
    if (overflow) {
        set_error("flag buffer too small");
        return -5;
    }
    return 0;
}

void check_tree_args(int arity, int minlevel, int maxlevel, double mmf, double mod, double mdd)
{
    SG_REQUIRE(arity == 8 || arity == 2, "arity must be 8 (octree) or 2 (binary tree)");
    // TreeDustGrid::setupSelfBefore (TreeDustGrid.cpp:55-62)
    SG_REQUIRE(minlevel >= 0, "The minimum tree level should be at least 0");
    SG_REQUIRE(maxlevel >= 2, "The maximum tree level should be at least 2");
    SG_REQUIRE(maxlevel > minlevel, "Maximum tree level should be larger than minimum tree level");
    SG_REQUIRE(mod >= 0.0, "The maximum mean optical depth should be positive");
    SG_REQUIRE(mmf >= 0.0, "The maximum mass fraction should be positive");
    SG_REQUIRE(mdd >= 0.0, "The maximum density dispersion fraction should be positive");
}

} // namespace
} // namespace sg

#define SG_TRY_ try {
#define SG_CATCH_                                                                                    \
    }                                                                                                \
    catch (const sg::Fail& f) { return f.code; }                                                     \
    catch (const std::exception& e) { sg::set_error(e.what()); return -4; }                          \
    return 0;

extern "C" {

int sg_sph_create(const double* rows, int64_t n, double fdust, double Tmax, sg_sph** out)
{
    SG_TRY_
    ensure_device();
    // SPHDustDistribution::setupSelfBefore (SPHDustDistribution.cpp:38-100)
    SG_REQUIRE(fdust > 0.0, "The dust fraction should be positive");
    SG_REQUIRE(n >= 0 && (rows || n == 0), "bad particle table");
    std::vector<double> x, y, z, hv, M, Z;
    double Mtot = 0, Mmetal = 0;
    for (int64_t i = 0; i < n; i++) {
        const double* r = rows + 7 * i;
        const double T = r[6];
        if (T > 0 && Tmax > 0 && T > Tmax) continue;   // too hot: ignored
        SG_REQUIRE(r[3] > 0 && std::isfinite(r[3]), "smoothing lengths must be positive");
        x.push_back(r[0]); y.push_back(r[1]); z.push_back(r[2]); hv.push_back(r[3]); M.push_back(r[4]); Z.push_back(r[5]);
        Mtot += r[4];
        Mmetal += r[4] * r[5];
    }
    if (Mtot < 0 || Mmetal < 0) { x.clear(); y.clear(); z.clear(); hv.clear(); M.clear(); Z.clear(); }
    const int64_t np = (int64_t)x.size();
    SG_REQUIRE(np >= 1, "no SPH particles hold dust");
    SG_REQUIRE(np < (int64_t)1 << 31, "too many particles");
    auto* s = new sg_sph();
    try {
        s->n = np;
        s->fdust = fdust;
        // SPHGasParticle.cpp:48-56 and SPHDustDistribution::mass (SPHDustDistribution.cpp:248-256)
        std::vector<double> sv(np), mz8(np), norm(np), rho0(np);
        double sum = 0.0;
        for (int64_t p = 0; p < np; p++) {
            const double h = hv[p];
            norm[p] = 1 / (h * h);
            rho0[p] = 8.0 / M_PI * M[p] / (h * h * h) * Z[p];
            sv[p] = 2.42 / h;
            mz8[p] = M[p] * Z[p] / 8.;
            sum += M[p] * Z[p];
        }
        sum *= fdust;
        s->mass = std::max(sum, 0.);
        // SPHGasParticleGrid.cpp:94-132: the grid and each cell's list of the particles whose sphere meets it
        s->grid.assign(3 * (kPG + 1), 0.0);
        make_axis(x, hv, &s->grid[0]);
        make_axis(y, hv, &s->grid[kPG + 1]);
        make_axis(z, hv, &s->grid[2 * (kPG + 1)]);
        const double *gx = &s->grid[0], *gy = &s->grid[kPG + 1], *gz = &s->grid[2 * (kPG + 1)];
        std::vector<uint32_t> range(np);
        std::vector<int64_t> cnt(kPG * kPG * kPG + 1, 0);
        for (int pass = 0; pass < 2; pass++) {
            std::vector<int64_t> fill;
            if (pass == 1) {
                for (size_t c = 1; c < cnt.size(); c++) cnt[c] += cnt[c - 1];
                s->cids.resize(cnt.back());
                fill.assign(cnt.begin(), cnt.end() - 1);
            }
            for (int64_t p = 0; p < np; p++) {
                const double h = hv[p];
                const int i1 = locate_grid_host(gx, x[p] - h), i2 = locate_grid_host(gx, x[p] + h);
                const int j1 = locate_grid_host(gy, y[p] - h), j2 = locate_grid_host(gy, y[p] + h);
                const int k1 = locate_grid_host(gz, z[p] - h), k2 = locate_grid_host(gz, z[p] + h);
                range[p] = (uint32_t)i1 | (uint32_t)i2 << 5 | (uint32_t)j1 << 10 | (uint32_t)j2 << 15 |
                           (uint32_t)k1 << 20 | (uint32_t)k2 << 25;
                for (int i = i1; i <= i2; i++)
                    for (int j = j1; j <= j2; j++)
                        for (int k = k1; k <= k2; k++)
                            if (intersects(gx[i], gx[i + 1], gy[j], gy[j + 1], gz[k], gz[k + 1], x[p], y[p], z[p], h)) {
                                const int c = cell_index(i, j, k);
                                if (pass == 0) cnt[c + 1]++;
                                else s->cids[fill[c]++] = (int32_t)p;
                            }
            }
        }
        s->coff = cnt;
        s->cx.upload(x); s->cy.upload(y); s->cz.upload(z); s->h.upload(hv);
        s->s.upload(sv); s->mz8.upload(mz8); s->norm.upload(norm); s->rho0.upload(rho0);
        s->range.upload(range);
        s->dcoff.upload(s->coff);
        s->dcids.upload(s->cids);
        s->erf.upload(erf_table());
        s->dgrid.upload(s->grid);
    } catch (...) {
        delete s;
        throw;
    }
    *out = s;
    SG_CATCH_
}

int sg_sph_destroy(sg_sph* s)
{
    delete s;
    return 0;
}

int sg_sph_info(const sg_sph* s, int64_t* count, double* mass, double* borders, int64_t* list_entries)
{
    SG_TRY_
    SG_REQUIRE(s, "null particle set");
    if (count) *count = s->n;
    if (mass) *mass = s->mass;
    if (borders) std::memcpy(borders, s->grid.data(), s->grid.size() * sizeof(double));
    if (list_entries) *list_entries = (int64_t)s->cids.size();
    SG_CATCH_
}

int sg_sph_cell_list(const sg_sph* s, int i, int j, int k, int32_t* out, int64_t cap, int64_t* count)
{
    SG_TRY_
    SG_REQUIRE(s, "null particle set");
    SG_REQUIRE(i >= 0 && i < kPG && j >= 0 && j < kPG && k >= 0 && k < kPG, "grid cell index out of range");
    const int c = cell_index(i, j, k);
    const int64_t len = s->coff[c + 1] - s->coff[c];
    *count = len;
    if (out) std::memcpy(out, s->cids.data() + s->coff[c], std::min(len, cap) * sizeof(int32_t));
    SG_CATCH_
}

int sg_sph_mass_in_box(const sg_sph* s, const double* boxes, int64_t n, double* out)
{
    SG_TRY_
    ensure_device();
    SG_REQUIRE(s, "null particle set");
    if (n <= 0) return 0;
    DevBuf b, o;
    b.upload(boxes, 6 * n);
    o.alloc(n * sizeof(double));
    hipLaunchKernelGGL(k_sph_mass_in_box, dim3(blocks(n, 256)), dim3(256), 0, 0, s->view(), b.as<double>(), n,
                       o.as<double>());
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpy(out, o.p, n * sizeof(double), hipMemcpyDeviceToHost));
    SG_CATCH_
}

int sg_sph_density(const sg_sph* s, const double* pos, int64_t n, double* out)
{
    SG_TRY_
    ensure_device();
    SG_REQUIRE(s, "null particle set");
    if (n <= 0) return 0;
    DevBuf b, o;
    b.upload(pos, 3 * n);
    o.alloc(n * sizeof(double));
    hipLaunchKernelGGL(k_sph_density, dim3(blocks(n, 256)), dim3(256), 0, 0, s->view(), b.as<double>(), n,
                       o.as<double>());
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpy(out, o.p, n * sizeof(double), hipMemcpyDeviceToHost));
    SG_CATCH_
}

int sg_tree_subdivide_sph(int arity, const double ext[6], const sg_sph* s, int minlevel, int maxlevel,
                          double maxmassfraction, double maxopticaldepth, double maxdensdisp, int nsamples,
                          uint32_t seed, uint8_t* flags, double* masses, int64_t cap, int64_t* nnodes)
{
    SG_TRY_
    ensure_device();
    SG_REQUIRE(s, "null particle set");
    check_tree_args(arity, minlevel, maxlevel, maxmassfraction, maxopticaldepth, maxdensdisp);
    SG_REQUIRE(s->mass > 0, "the particles hold no dust");
    SplitArgs a{};
    // TreeDustGrid.cpp:70: the mass-in-box interface serves unless the dispersion criterion needs samples
    a.mode = maxdensdisp > 0 ? NM_SPH_SAMPLES : NM_SPH_BOX;
    SG_REQUIRE(a.mode == NM_SPH_BOX || nsamples >= 1, "Number of random samples must be at least 1");
    a.sph = s->view();
    a.nsamples = nsamples;
    a.seed = seed;
    a.minlevel = minlevel;
    a.maxlevel = maxlevel;
    a.total = s->mass;
    a.maxmassfraction = maxmassfraction;
    a.maxopticaldepth = maxopticaldepth;
    a.maxdensdisp = maxdensdisp;
    if (a.mode == NM_SPH_BOX) return subdivide_lists(arity, ext, a, flags, masses, cap, nnodes);
    return subdivide_device(arity, ext, a, flags, masses, cap, nnodes);
    SG_CATCH_
}

int sg_tree_subdivide_device(int arity, const double ext[6], int ncomp, const double* geoms16, const double* scales,
                             double total_mass, int minlevel, int maxlevel, double maxmassfraction,
                             double maxopticaldepth, double maxdensdisp, int nsamples, uint32_t seed, uint8_t* flags,
                             int64_t cap, int64_t* nnodes)
{
    SG_TRY_
    ensure_device();
    check_tree_args(arity, minlevel, maxlevel, maxmassfraction, maxopticaldepth, maxdensdisp);
    SG_REQUIRE(ncomp >= 1 && geoms16 && scales, "at least one dust component is needed");
    SG_REQUIRE(nsamples >= 1, "Number of random samples must be at least 1");
    std::vector<DiscDev> discs;
    for (int h = 0; h < ncomp; h++) {
        const double* g = geoms16 + 16 * h;
        SG_REQUIRE((int)g[0] == 1 && g[1] > 0 && g[2] > 0, "the dust geometries must be exponential discs");
        discs.push_back(disc_dev(g));
    }
    DevBuf dd, ds;
    dd.upload(discs);
    ds.upload(scales, ncomp);
    SplitArgs a{};
    a.mode = NM_DISC_SAMPLES;
    a.discs = dd.as<DiscDev>();
    a.scales = ds.as<double>();
    a.ncomp = ncomp;
    a.nsamples = nsamples;
    a.seed = seed;
    a.minlevel = minlevel;
    a.maxlevel = maxlevel;
    a.total = total_mass;
    a.maxmassfraction = maxmassfraction;
    a.maxopticaldepth = maxopticaldepth;
    a.maxdensdisp = maxdensdisp;
    return subdivide_device(arity, ext, a, flags, nullptr, cap, nnodes);
    SG_CATCH_
}

int sg_grid_sph_density(sg_grid* g, const sg_sph* s, const double* node_masses, int64_t nnodes, double* rho_out)
{
    SG_TRY_
    ensure_device();
    SG_REQUIRE(g && s, "null grid or particle set");
    SG_REQUIRE(g->kind == GK_TREE, "mass-in-box cell densities need a tree grid (TreeDustGrid::density)");
    const HostTree& t = g->ht;
    const int64_t nc = (int64_t)t.idv.size();
    if (node_masses) {
        // the masses sg_tree_subdivide_sph evaluated for this tree's nodes
        SG_REQUIRE(nnodes == (int64_t)t.child0.size(), "node masses do not belong to this tree");
        for (int64_t m = 0; m < nc; m++) {
            const int32_t l = t.idv[m];
            const double* q = &t.bounds[6 * (size_t)l];
            SG_REQUIRE(node_masses[l] >= 0, "a cell's mass was not evaluated");
            rho_out[m] = node_masses[l] / ((q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]));
        }
        return 0;
    }
    std::vector<double> boxes(6 * nc);
    for (int64_t m = 0; m < nc; m++) std::memcpy(&boxes[6 * m], &t.bounds[6 * (size_t)t.idv[m]], 6 * sizeof(double));
    DevBuf b, o;
    b.upload(boxes);
    o.alloc(nc * sizeof(double));
    hipLaunchKernelGGL(k_leaf_density, dim3(blocks(nc, 256)), dim3(256), 0, 0, s->view(), b.as<double>(), nc,
                       o.as<double>());
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpy(rho_out, o.p, nc * sizeof(double), hipMemcpyDeviceToHost));
    SG_CATCH_
}

int sg_grid_sample_density_device(sg_grid* g, const double* geom16, double scale, int nsamples, uint32_t seed,
                                  double* rho_out)
{
    SG_TRY_
    ensure_device();
    SG_REQUIRE(g && geom16, "null grid or geometry record");
    SG_REQUIRE(g->kind == GK_CART || g->kind == GK_TREE, "Voronoi cells take the density at their site");
    SG_REQUIRE((int)geom16[0] == 1 && geom16[1] > 0 && geom16[2] > 0, "the dust geometry must be an exponential disc");
    SG_REQUIRE(nsamples >= 1, "Number of random samples must be at least 1");
    const int64_t nc = g->ncells;
    std::vector<double> boxes(6 * nc);
    if (g->kind == GK_CART) {
        const int Nx = g->cart.Nx, Ny = g->cart.Ny, Nz = g->cart.Nz;
        std::vector<double> hx(Nx + 1), hy(Ny + 1), hz(Nz + 1);
        SG_HIP(hipMemcpy(hx.data(), g->cart.xv, hx.size() * 8, hipMemcpyDeviceToHost));
        SG_HIP(hipMemcpy(hy.data(), g->cart.yv, hy.size() * 8, hipMemcpyDeviceToHost));
        SG_HIP(hipMemcpy(hz.data(), g->cart.zv, hz.size() * 8, hipMemcpyDeviceToHost));
        for (int64_t m = 0; m < nc; m++) {
            const int i = (int)(m / ((int64_t)Nz * Ny)), j = (int)((m / Nz) % Ny), k = (int)(m % Nz);
            double* q = &boxes[6 * m];
            q[0] = hx[i]; q[1] = hy[j]; q[2] = hz[k]; q[3] = hx[i + 1]; q[4] = hy[j + 1]; q[5] = hz[k + 1];
        }
    } else {
        for (int64_t m = 0; m < nc; m++)
            for (int a = 0; a < 3; a++) {
                boxes[6 * m + a] = g->hrec64[m].lo[a];
                boxes[6 * m + a + 3] = g->hrec64[m].hi[a];
            }
    }
    DevBuf b, o;
    b.upload(boxes);
    o.alloc(nc * sizeof(double));
    hipLaunchKernelGGL(k_sample_boxes, dim3(blocks(nc, 256)), dim3(256), 0, 0, b.as<double>(), nc, disc_dev(geom16),
                       scale, nsamples, seed, o.as<double>());
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpy(rho_out, o.p, nc * sizeof(double), hipMemcpyDeviceToHost));
    SG_CATCH_
}

int sg_knn_distance(const double* pos, int64_t n, int k, double* out)
{
    SG_TRY_
    ensure_device();
    SG_REQUIRE(k >= 1 && k <= 64, "k must be between 1 and 64");
    SG_REQUIRE(n > k && n < (int64_t)1 << 31, "need more points than neighbours (and fewer than 2^31)");
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) { lo[a] = pos[a]; hi[a] = pos[a]; }
    for (int64_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], pos[3 * i + a]);
            hi[a] = std::max(hi[a], pos[3 * i + a]);
        }
    const int nb = std::max(1, std::min(256, (int)std::cbrt((double)n / 2.0)));
    double w[3];
    for (int a = 0; a < 3; a++) w[a] = std::max((hi[a] - lo[a]) / nb, 1e-300);
    std::vector<int64_t> boff((size_t)nb * nb * nb + 1, 0);
    std::vector<int64_t> bucket(n);
    for (int64_t i = 0; i < n; i++) {
        int b[3];
        for (int a = 0; a < 3; a++) b[a] = std::min(nb - 1, std::max(0, (int)((pos[3 * i + a] - lo[a]) / w[a])));
        bucket[i] = ((int64_t)b[0] * nb + b[1]) * nb + b[2];
        boff[bucket[i] + 1]++;
    }
    for (size_t q = 1; q < boff.size(); q++) boff[q] += boff[q - 1];
    std::vector<int32_t> order(n);
    {
        std::vector<int64_t> fill(boff.begin(), boff.end() - 1);
        for (int64_t i = 0; i < n; i++) order[fill[bucket[i]]++] = (int32_t)i;
    }
    DevBuf dp, dord, dboff, dout;
    dp.upload(pos, 3 * n);
    dord.upload(order);
    dboff.upload(boff);
    dout.alloc(n * sizeof(double));
    hipLaunchKernelGGL(k_knn, dim3(blocks(n, 256)), dim3(256), 0, 0, dp.as<double>(), n, dord.as<int32_t>(),
                       dboff.as<int64_t>(), nb, lo[0], lo[1], lo[2], w[0], w[1], w[2], k, dout.as<double>());
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpy(out, dout.p, n * sizeof(double), hipMemcpyDeviceToHost));
    SG_CATCH_
}

} // extern "C"
