// sg_kernels.hip -- kernels and the C-ABI of include/skirt7_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/skirt7_amd.h"
#include "sg_host.hpp"
#include "sg_mc.hpp"
#include "sg_split.hpp"
#include "sg_pan.hpp"
#include "sg_geom.hpp"

namespace sg {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

static int g_device = -1;
static int g_cus = 256;
// compaction threshold 0: each pipeline's own default (32 lanes for the lane pipeline, 24 for the walk kernel)
static int g_tune_block = 256, g_tune_waves = 0, g_tune_compact = 0;
static int64_t g_tune_slots = 0;   // in-flight packet slots of the event/walk pipeline (0: its default)
static int g_pipeline = 1;   // 0: persistent lane pipeline (k_mc); 1: event and walk kernels (k_events/k_walks)
static uint64_t g_batch_alone = (uint64_t)1 << 21;   // batched members this large run alone (0: never)
static int g_topdown_descent = 0;   // 1: mixed walks of TopDown trees descend from the root at every step
static int g_walk_order = 0;        // 1: the walk kernel takes each pass's walks in coherence-key order; 2: so,
                                    // claimed by XCD segment
static int g_slot_groups = 2;       // slot groups of the event/walk pipeline (2: alternating on two streams)
static uint64_t g_emit_batch = 0;   // packets emitted at once (0: each pipeline's default)

// Mixed-precision tree walks follow the ropes (the same cells as any search method) unless TopDown
// descents were asked for a TopDown tree.
static bool rope_walk(const sg_grid* g) { return g->tree.search == SG_NEIGHBOR || !g_topdown_descent; }

static void ensure_init()
{
    if (g_device >= 0) return;
    if (sg_init(0) != 0) throw Fail{-3};
}

// ------------------------------------------------------------------------- probe kernels

// DustGrid::path for a batch of rays: the segment list of each ray.
template <class W>
__global__ void __launch_bounds__(256) k_paths(const typename W::View g, int64_t n, const double* rk, int maxseg,
                                                int32_t* counts, int32_t* mo, double* dso)
{
    __shared__ int32_t s_nodes[SG_LDS_NODES];
    int nlds;
    stage_tree_nodes(g, s_nodes, nlds);
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* q = rk + 6 * i;
    W w;
    w.nrecompute = 0; w.nstuck = 0;
    w.init(g, s_nodes, nlds, q[0], q[1], q[2], q[3], q[4], q[5]);
    int c = 0, m;
    typename W::Real ds, rho;
    while (w.next(g, s_nodes, nlds, m, ds, rho)) {
        if (c < maxseg) { mo[i * maxseg + c] = m >= 0 ? W::cell_of(g, m) : m; dso[i * maxseg + c] = (double)ds; }
        c++;
    }
    counts[i] = c <= maxseg ? c : -1;
}

// DustSystem::opticaldepth (mode 0) or DustGridPath::pathlength (mode 1) for a batch of rays.
template <class W>
__global__ void __launch_bounds__(256) k_tau(const typename W::View g, int64_t n, const double* rk, double kext,
                                              const double* taus, int mode, double* out)
{
    using R = typename W::Real;
    __shared__ int32_t s_nodes[SG_LDS_NODES];
    int nlds;
    stage_tree_nodes(g, s_nodes, nlds);
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* q = rk + 6 * i;
    W w;
    w.nrecompute = 0; w.nstuck = 0;
    w.init(g, s_nodes, nlds, q[0], q[1], q[2], q[3], q[4], q[5]);
    int m;
    R ds, rho;
    R tau = 0, s = 0, tprev = 0, sprev = 0;
    int nseg = 0;
    if (mode == 0) {
        while (w.next(g, s_nodes, nlds, m, ds, rho)) tau += ((R)kext * rho) * ds;
        out[i] = (double)tau;
        return;
    }
    R target = (R)taus[i];
    bool found = false;
    R res = 0;
    while (w.next(g, s_nodes, nlds, m, ds, rho)) {
        tprev = tau; sprev = s;
        s += ds;
        tau += ((R)kext * rho) * ds;
        nseg++;
        if (target > (R)0 && tau > target) {
            res = nseg == 1 ? interp_linlin(target, (R)0, tau, (R)0, s) : interp_linlin(target, tprev, tau, sprev, s);
            found = true;
            break;
        }
    }
    if (!found) {
        if (nseg == 0 || !(target > (R)0)) res = 0;
        else if (target > tau) res = s;
        else if (nseg == 1) res = s;
        else res = interp_linlin(target, tprev, tau, sprev, s);
    }
    out[i] = (double)res;
}

// DustSystem::setSampleDensityBody (DustSystem.cpp:145-172) for a Cartesian grid: cell m's density is scale
// times the mean of nsamples disc densities at positions in the cell's box drawn from the counter-based
// stream (seed, m, sample) that the host sampler uses; written as the grid's FP64 and FP32 densities.
__global__ void k_resample_cart(const double* xv, const double* yv, const double* zv, int Nx, int Ny, int Nz,
                                DiscDev d, double scale, int nsamples, uint32_t seed, double* rho, float* rhof)
{
    int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (int64_t)Nx * Ny * Nz) return;
    int i = (int)(m / ((int64_t)Nz * Ny)), j = (int)((m / Nz) % Ny), k = (int)(m % Nz);
    double x0 = xv[i], y0 = yv[j], z0 = zv[k], x1 = xv[i + 1], y1 = yv[j + 1], z1 = zv[k + 1];
    double sum = 0;
    for (int q = 0; q < nsamples; q++) {
        double u[3];
        setup_sample3(seed, (uint64_t)m, (uint32_t)q, u);
        sum += disc_density(d, x0 + u[0] * (x1 - x0), y0 + u[1] * (y1 - y0), z0 + u[2] * (z1 - z0));
    }
    double r = scale * sum / nsamples;
    rho[m] = r;
    rhof[m] = (float)r;
}

__global__ void k_whichcell_cart(const CartView g, int64_t n, const double* r, int32_t* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = r[3 * i], y = r[3 * i + 1], z = r[3 * i + 2];
    // CartesianDustGrid::whichcell (CartesianDustGrid.cpp:109-118) with NR::locate_fail
    auto fail = [](const double* v, int size, double q) {
        if (q > v[size - 1]) return -1;
        int jl = -1, ju = size - 1;
        while (ju - jl > 1) {
            int jm = (ju + jl) >> 1;
            if (q < v[jm]) ju = jm;
            else jl = jm;
        }
        return jl;
    };
    int a = fail(g.xv, g.Nx + 1, x), b = fail(g.yv, g.Ny + 1, y), c = fail(g.zv, g.Nz + 1, z);
    out[i] = (a < 0 || b < 0 || c < 0) ? -1 : c + g.Nz * b + g.Nz * g.Ny * a;
}
__global__ void k_whichcell_tree(const TreeView t, int64_t n, const double* r, int32_t* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double lo[3], hi[3];
    out[i] = descend64(t, t.node, 0, r[3 * i], r[3 * i + 1], r[3 * i + 2], lo, hi);
}
__global__ void k_whichcell_voro(const VoroView g, int64_t n, const double* r, int32_t* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = voro_locate(g, r[3 * i], r[3 * i + 1], r[3 * i + 2], -1);
}

} // namespace sg

using namespace sg;

static void split_resources();
static void static_scratch_setup(int device);

// ========================================================================= C-ABI

#define SG_TRY try {
#define SG_CATCH                                                                                     \
    }                                                                                                \
    catch (const sg::Fail& f) { return f.code; }                                                     \
    catch (const std::exception& e) { sg::set_error(e.what()); return -4; }                          \
    return 0;

extern "C" {

const char* sg_last_error(void) { return sg::g_err.c_str(); }

int sg_init(int device)
{
    SG_TRY
    int n = 0;
    SG_HIP(hipGetDeviceCount(&n));
    SG_REQUIRE(n > device && device >= 0, "no such HIP device");
    SG_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    SG_HIP(hipGetDeviceProperties(&p, device));
    SG_REQUIRE(std::string(p.gcnArchName).find("gfx950") != std::string::npos,
               std::string("this library is built for gfx950, found ") + p.gcnArchName);
    g_cus = p.multiProcessorCount;
    g_device = device;
    // the transport streams, the pinned progress buffer and this device's static work array are set up
    // here, so that sg_sim_run_static allocates nothing
    split_resources();
    static_scratch_setup(device);
    SG_CATCH
}

int sg_device_name(char* buf, int len)
{
    SG_TRY
    ensure_init();
    hipDeviceProp_t p;
    SG_HIP(hipGetDeviceProperties(&p, g_device));
    std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    SG_CATCH
}

int sg_set_pipeline(int mode)
{
    if (mode != 0 && mode != 1) {
        sg::set_error("pipeline mode must be 0 (lane pipeline) or 1 (event and walk kernels)");
        return -1;
    }
    g_pipeline = mode;
    return 0;
}

int sg_set_batch_alone(uint64_t min_packets)
{
    g_batch_alone = min_packets;
    return 0;
}

int sg_set_mixed_topdown(int descend)
{
    g_topdown_descent = descend ? 1 : 0;
    return 0;
}

int sg_set_emit_batch(uint64_t packets)
{
    if (packets != 0 && packets < 64) {
        sg::set_error("emission batch must be 0 (default) or at least 64 packets");
        return -1;
    }
    g_emit_batch = packets;
    return 0;
}

int sg_set_slot_groups(int groups)
{
    if (groups != 1 && groups != 2) {
        sg::set_error("slot groups must be 1 or 2");
        return -1;
    }
    g_slot_groups = groups;
    return 0;
}

int sg_set_walk_order(int order)
{
    if (order < 0 || order > 2) {
        sg::set_error("walk order must be 0 (as appended), 1 (sorted) or 2 (sorted, claimed by XCD segment)");
        return -1;
    }
    g_walk_order = order;
    return 0;
}

int sg_set_tuning(int block, int waves_per_cu, int compact_lanes)
{
    if (block == 64 || block == 128 || block == 256) g_tune_block = block;
    if (block < 0) g_tune_block = 256;
    if (waves_per_cu > 0) g_tune_waves = waves_per_cu;
    if (waves_per_cu < 0) g_tune_waves = 0;
    if (compact_lanes > 0 && compact_lanes <= 64) g_tune_compact = compact_lanes;
    if (compact_lanes < 0) g_tune_compact = 0;
    return 0;
}

// ---------------------------------------------------------------- grids

int sg_grid_create_cartesian(const double ext[6], const double* xv, int nxv, const double* yv, int nyv,
                             const double* zv, int nzv, sg_grid** out)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(nxv >= 2 && nyv >= 2 && nzv >= 2, "a Cartesian grid needs at least one bin per axis");
    SG_REQUIRE((int64_t)(nxv - 1) * (nyv - 1) * (nzv - 1) < (int64_t)1 << 31, "too many cells");
    auto* g = new sg_grid();
    g->kind = GK_CART;
    g->ncells = (int64_t)(nxv - 1) * (nyv - 1) * (nzv - 1);
    std::vector<float> f;
    g->xv.upload(xv, nxv); g->yv.upload(yv, nyv); g->zv.upload(zv, nzv);
    f.assign(xv, xv + nxv); g->xvf.upload(f);
    f.assign(yv, yv + nyv); g->yvf.upload(f);
    f.assign(zv, zv + nzv); g->zvf.upload(f);
    g->rho.alloc(g->ncells * sizeof(double)); g->rho.zero();
    g->rhof.alloc(g->ncells * sizeof(float)); g->rhof.zero();
    CartView& c = g->cart;
    c.xv = g->xv.as<double>(); c.yv = g->yv.as<double>(); c.zv = g->zv.as<double>();
    c.xvf = g->xvf.as<float>(); c.yvf = g->yvf.as<float>(); c.zvf = g->zvf.as<float>();
    c.rho = g->rho.as<double>(); c.rhof = g->rhof.as<float>();
    c.Nx = nxv - 1; c.Ny = nyv - 1; c.Nz = nzv - 1;
    c.xmin = ext[0]; c.ymin = ext[1]; c.zmin = ext[2]; c.xmax = ext[3]; c.ymax = ext[4]; c.zmax = ext[5];
    double wmin = DBL_MAX;
    for (int i = 0; i + 1 < nxv; i++) wmin = std::min(wmin, xv[i + 1] - xv[i]);
    for (int i = 0; i + 1 < nyv; i++) wmin = std::min(wmin, yv[i + 1] - yv[i]);
    for (int i = 0; i + 1 < nzv; i++) wmin = std::min(wmin, zv[i + 1] - zv[i]);
    c.wminf = (float)wmin;
    *out = g;
    SG_CATCH
}

int sg_grid_create_tree(int arity, const double ext[6], const uint8_t* subdivided, int64_t nnodes, int search,
                        sg_grid** out)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(arity == 8 || arity == 2, "arity must be 8 (octree) or 2 (binary tree)");
    SG_REQUIRE(search == SG_TOPDOWN || search == SG_NEIGHBOR || search == SG_BOOKKEEPING,
               "search method must be TopDown, Neighbor or Bookkeeping");
    // BinTreeDustGrid::setupSelfBefore (BinTreeDustGrid.cpp:23-24)
    SG_REQUIRE(!(search == SG_BOOKKEEPING && arity != 8), "Bookkeeping method is not compatible with binary tree");
    SG_REQUIRE(nnodes >= 1 && nnodes < (int64_t)1 << 31, "bad node count");
    auto* g = new sg_grid();
    g->kind = GK_TREE;
    HostTree& h = g->ht;
    if (!h.build(arity, ext, subdivided, nnodes)) { delete g; throw Fail{-1}; }
    if (search == SG_NEIGHBOR) h.build_neighbors();   // TopDown needs no neighbour lists
    const bool lists = !h.nb.empty();
    int64_t nn = (int64_t)h.child0.size(), nc = (int64_t)h.idv.size();
    g->ncells = nc;
    // Device node numbering: the breadth-first prefix of whole sibling groups that the LDS stage holds
    // keeps its ids; every deeper sibling group follows in depth-first order of its parent, so that
    // leaves that are neighbours in space, which a walk meets one after another, lie near each other
    // in memory (a node record fetched at random brings a whole cache line in). Sibling groups stay
    // contiguous, so a node's children are still its first child + q. Host-side ids stay breadth-first.
    // Below that prefix, sibling groups that hold an internal node come first, then the groups made of
    // leaves only, each part in depth-first order: the walks read leaf records, and a cache line then holds
    // leaves of one small subtree instead of leaves mixed with internal nodes (half the nodes of a binary
    // tree), which a walk reaches only when a rope leads to a finer region.
    std::vector<int32_t> nid(nn, -1);
    const int64_t top = nn <= SG_LDS_NODES ? nn : 1 + arity * ((SG_LDS_NODES - 1) / arity);
    int64_t ndev = nn;   // device ids in use, padding included
    for (int64_t l = 0; l < top; l++) nid[l] = (int32_t)l;
    {
        std::vector<int32_t> groups;   // first child of each sibling group below the prefix, depth-first
        std::vector<int32_t> stack{0};
        while (!stack.empty()) {
            const int32_t l = stack.back();
            stack.pop_back();
            const int32_t c0 = h.child0[l];
            if (c0 < 0) continue;
            if (nid[c0] < 0) groups.push_back(c0);
            for (int q = arity - 1; q >= 0; q--) stack.push_back(c0 + q);
        }
        int64_t next = top, used = top;
        for (int pass = 0; pass < 2; pass++) {
            // the leaf groups start on a 128-byte line (four 32-byte records), so that every group of two
            // or eight lies within whole lines; the skipped ids hold no node
            if (pass == 1) next = (next + 3) / 4 * 4;
            for (int32_t c0 : groups) {
                bool leaves = true;
                for (int q = 0; q < arity; q++) leaves = leaves && h.child0[c0 + q] < 0;
                if (leaves != (pass == 1)) continue;
                for (int q = 0; q < arity; q++) nid[c0 + q] = (int32_t)next++;
                used += arity;
            }
        }
        SG_REQUIRE(used == nn, "tree renumbering missed nodes");
        ndev = next;
    }
    SG_REQUIRE(ndev < (int64_t)1 << 31, "too many tree nodes");
    std::vector<int32_t> words(ndev, 0);
    for (int64_t l = 0; l < nn; l++) words[nid[l]] = h.child0[l] >= 0 ? nid[h.child0[l]] : -(h.cellnumber[l] + 1);
    g->node.upload(words);
    g->hrec64.assign(nc, TreeRec64{});
    g->hnode32.assign(ndev, TreeNode32{});
    std::vector<int32_t> off(nc * 6 + 1, 0), ids;
    for (int64_t c = 0; c < nc; c++) {
        int l = h.idv[c];
        TreeRec64& r = g->hrec64[c];
        for (int a = 0; a < 3; a++) { r.lo[a] = h.bounds[6 * l + a]; r.hi[a] = h.bounds[6 * l + 3 + a]; }
        r.rho = 0; r.node = nid[l];
        for (int w = 0; w < 6; w++) {
            off[c * 6 + w] = (int32_t)ids.size();
            if (!lists) continue;
            for (int32_t nbn : h.nb[6 * (size_t)l + w]) ids.push_back(h.cellnumber[nbn]);
        }
    }
    // mixed-precision records: subdivision counts per axis (a child adds one along each axis of an
    // octree, or along its parent's split axis of a binary tree), density or first child, ropes
    std::vector<int32_t> cellof(ndev, -1);
    std::vector<uint64_t> kpack(ndev, 0);
    std::vector<int32_t> ropes;
    h.build_ropes(ropes);
    std::vector<uint32_t> nsub(nn, 0);   // nx | ny << 5 | nz << 10
    for (int64_t l = 0; l < nn; l++) {
        const int32_t c0 = h.child0[l];
        if (c0 >= 0) {
            uint32_t add = arity == 8 ? (1u | 1u << 5 | 1u << 10) : (1u << (5 * h.dir[l]));
            for (int q = 0; q < arity; q++) nsub[c0 + q] = nsub[l] + add;
        }
    }
    // integer indices of every node along x, y, z at its own subdivision (children in the order of
    // OctTreeNode.cpp:38-49: bit a of the child number is the upper half along axis a)
    std::vector<int32_t> kidx(3 * (size_t)nn, 0);
    for (int64_t l = 0; l < nn; l++) {
        const int32_t c0 = h.child0[l];
        if (c0 < 0) continue;
        for (int q = 0; q < arity; q++)
            for (int a = 0; a < 3; a++) {
                const int32_t k = kidx[3 * (size_t)l + a];
                const bool split = arity == 8 || h.dir[l] == a;
                const int bit = arity == 8 ? (q >> a) & 1 : q;
                kidx[3 * (size_t)(c0 + q) + a] = split ? 2 * k + bit : k;
            }
    }
    std::array<float, 3> extf3, widthf3;
    for (int a = 0; a < 3; a++) { extf3[a] = (float)ext[a]; widthf3[a] = (float)(ext[a + 3] - ext[a]); }
    // split plane of an internal node along axis a, exactly as the walker computes walls: one fused
    // multiply-add, so that every wall is the same float wherever it is computed
    auto split_plane = [&](int64_t l, int a) {
        const int n = (int)((nsub[l] >> (5 * a)) & 31u) + 1;
        return std::fmaf((float)(2 * kidx[3 * (size_t)l + a] + 1), std::ldexp(widthf3[a], -n), extf3[a]);
    };
    for (int64_t l = 0; l < nn; l++) {
        uint32_t nx = nsub[l] & 31u, ny = (nsub[l] >> 5) & 31u, nz = (nsub[l] >> 10) & 31u;
        SG_REQUIRE(nx <= 21 && ny <= 21 && nz <= 21,
                   "tree too deep for the mixed-precision records (21 subdivisions per axis)");
        TreeNode32& f = g->hnode32[nid[l]];
        const bool leaf = h.child0[l] < 0;
        f.code = nsub[l] | (leaf ? kTreeLeaf : 0u) | ((uint32_t)(leaf ? 0 : h.dir[l]) << 16);
        f.val = leaf ? 0u : (uint32_t)nid[h.child0[l]];
        cellof[nid[l]] = h.cellnumber[l];
        kpack[nid[l]] = (uint64_t)kidx[3 * (size_t)l] | (uint64_t)kidx[3 * (size_t)l + 1] << 21 |
                        (uint64_t)kidx[3 * (size_t)l + 2] << 42;
        if (!leaf) {
            // an internal node holds its split planes where a leaf holds its ropes: x, y, z for an
            // octree node, the one plane of its split axis for a binary node
            float pl[3] = {0.0f, 0.0f, 0.0f};
            if (arity == 8) for (int a = 0; a < 3; a++) pl[a] = split_plane(l, a);
            else pl[0] = split_plane(l, h.dir[l]);
            for (int w = 0; w < 6; w++) {
                float v = w < 3 ? pl[w] : 0.0f;
                std::memcpy(&f.rope[w], &v, sizeof v);
            }
            continue;
        }
        for (int w = 0; w < 6; w++) {
            const int32_t rope = ropes[6 * (size_t)l + w];
            f.rope[w] = rope >= 0 ? nid[rope] : -1;
            // the walker steps into a rope's indices by shifting them down per axis
            if (rope >= 0)
                for (int a = 0; a < 3; a++)
                    SG_REQUIRE(((nsub[rope] >> (5 * a)) & 31u) <= ((nsub[l] >> (5 * a)) & 31u),
                               "a tree rope is finer than its leaf");
            if (lists) {
                // consistency with the neighbour lists: the first (largest-overlap) neighbour, raised to
                // this node's level when finer, is the same rope
                int32_t r = h.nb[6 * (size_t)l + w].empty() ? -1 : h.nb[6 * (size_t)l + w][0];
                while (r >= 0 && h.level[r] > h.level[l]) r = h.father[r];
                SG_REQUIRE(r == rope, "tree ropes disagree with the neighbour lists");
            }
        }
    }
    g->cellof.upload(cellof);
    g->kpack.upload(kpack);
    off[nc * 6] = (int32_t)ids.size();
    SG_REQUIRE(ids.size() < ((size_t)1 << 31), "neighbour lists too long");
    g->nboff.upload(off);
    if (ids.empty()) ids.push_back(-1);
    g->nbids.upload(ids);
    g->rec64.upload(g->hrec64);
    g->node32.upload(g->hnode32);
    TreeView& t = g->tree;
    t.node = g->node.as<int32_t>();
    t.rec64 = g->rec64.as<TreeRec64>();
    t.node32 = g->node32.as<TreeNode32>();
    t.cellof = g->cellof.as<int32_t>();
    t.kpack = g->kpack.as<uint64_t>();
    t.nboff = g->nboff.as<int32_t>();
    t.nbids = g->nbids.as<int32_t>();
    t.nnodes = (int32_t)ndev; t.ncells = (int32_t)nc;
    t.arity = arity; t.search = search;
    if (search == SG_BOOKKEEPING) {
        // the node arrays of the Bookkeeping walk, in the reference's breadth-first order
        std::vector<int32_t> fa(nn), cn(nn), c0(nn);
        for (int64_t l = 0; l < nn; l++) { fa[l] = h.father[l]; cn[l] = h.cellnumber[l]; c0[l] = h.child0[l]; }
        g->bkb.upload(h.bounds.data(), 6 * (size_t)nn);
        g->bkchild0.upload(c0);
        g->bkfather.upload(fa);
        g->bkcell.upload(cn);
        t.bkb = g->bkb.as<double>();
        t.bkchild0 = g->bkchild0.as<int32_t>();
        t.bkfather = g->bkfather.as<int32_t>();
        t.bkcell = g->bkcell.as<int32_t>();
    }
    for (int i = 0; i < 6; i++) { t.ext[i] = ext[i]; t.extf[i] = (float)ext[i]; }
    for (int a = 0; a < 3; a++) t.widthf[a] = (float)(ext[a + 3] - ext[a]);
    t.eps = h.eps;
    // the neighbour lists are only needed on the host while building records
    std::vector<std::vector<int32_t>>().swap(h.nb);
    *out = g;
    SG_CATCH
}

int sg_grid_create_voronoi(const double ext[6], const double* sites, int64_t n, const int64_t* nb_offsets,
                           const int32_t* nb_ids, sg_grid** out)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(n >= 1 && n < (int64_t)1 << 31, "bad site count");
    SG_REQUIRE(nb_offsets[n] < (int64_t)1 << 31, "neighbour lists too long");
    auto* g = new sg_grid();
    g->kind = GK_VORO;
    g->ncells = n;
    g->hsite.resize(n);
    for (int64_t i = 0; i < n; i++) g->hsite[i] = VoroSite{sites[3 * i], sites[3 * i + 1], sites[3 * i + 2], 0.0};
    g->site.upload(g->hsite);
    SG_REQUIRE(nb_offsets[0] == 0, "neighbour offsets must start at 0");
    for (int64_t i = 0; i < n; i++)
        SG_REQUIRE(nb_offsets[i + 1] > nb_offsets[i], "every Voronoi cell needs at least one neighbour or wall");
    for (int64_t t = 0; t < nb_offsets[n]; t++)
        SG_REQUIRE(nb_ids[t] >= -6 && nb_ids[t] < n && nb_ids[t] != -99, "neighbour id out of range");
    // Device lists: per cell a header record (its own site and density) followed by its neighbour
    // entries; the cells' blocks in Morton order of their sites, so that cells a walk crosses one
    // after another, which are neighbours in space, lie near each other in memory. voff holds each
    // cell's (first record, entry count).
    std::vector<uint64_t> key(n);
    for (int64_t i = 0; i < n; i++) {
        uint64_t k = 0;
        for (int a = 0; a < 3; a++) {
            const double u = (sites[3 * i + a] - ext[a]) / (ext[a + 3] - ext[a]);
            const uint64_t c = (uint64_t)std::min(2097151.0, std::max(0.0, u * 2097152.0));
            for (int b = 0; b < 21; b++) k |= ((c >> b) & 1ull) << (3 * b + a);
        }
        key[i] = k;
    }
    std::vector<int32_t> order(n);
    for (int64_t i = 0; i < n; i++) order[i] = (int32_t)i;
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
    std::vector<int32_t> be(2 * n);
    g->hvnb.clear();
    g->hvnb.reserve(n + nb_offsets[n]);
    for (int32_t m : order) {
        be[2 * m] = (int32_t)g->hvnb.size();
        be[2 * m + 1] = (int32_t)(nb_offsets[m + 1] - nb_offsets[m]);
        VoroNbr head{g->hsite[m].x, g->hsite[m].y, g->hsite[m].z, 0, 0};   // density set by set_rho
        g->hvnb.push_back(head);
        for (int64_t t = nb_offsets[m]; t < nb_offsets[m + 1]; t++) {
            const int32_t v = nb_ids[t];
            g->hvnb.push_back(v >= 0 ? VoroNbr{g->hsite[v].x, g->hsite[v].y, g->hsite[v].z, v, 0}
                                     : VoroNbr{0.0, 0.0, 0.0, v, 0});
        }
    }
    SG_REQUIRE(g->hvnb.size() < ((size_t)1 << 31), "neighbour lists too long");
    g->voff.upload(be);
    g->vnb.upload(g->hvnb);
    // bucket grid for the nearest-site search: about 2 sites per bucket
    int nbk = std::max(1, std::min(256, (int)std::cbrt((double)n / 2.0)));
    std::vector<int32_t> cnt((size_t)nbk * nbk * nbk + 1, 0), bucket(n);
    auto bidx = [&](double x, int a) {
        int i = (int)(nbk * (x - ext[a]) / (ext[a + 3] - ext[a]));
        return std::max(0, std::min(nbk - 1, i));
    };
    for (int64_t i = 0; i < n; i++) {
        int q = (bidx(g->hsite[i].x, 0) * nbk + bidx(g->hsite[i].y, 1)) * nbk + bidx(g->hsite[i].z, 2);
        bucket[i] = q;
        cnt[q + 1]++;
    }
    for (size_t q = 1; q < cnt.size(); q++) cnt[q] += cnt[q - 1];
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1), bid(n);
    for (int64_t i = 0; i < n; i++) bid[fill[bucket[i]]++] = (int32_t)i;
    // start sites of the greedy location: in a bucket with sites the one nearest its centre; an empty
    // bucket takes the start site of the nearest bucket with sites (breadth-first over face neighbours)
    const size_t nq = (size_t)nbk * nbk * nbk;
    std::vector<int32_t> rep(nq, -1), frontier, next;
    for (size_t q = 0; q < nq; q++) {
        if (cnt[q] == cnt[q + 1]) continue;
        const int qi = (int)(q / ((size_t)nbk * nbk)), qj = (int)(q / nbk % nbk), qk = (int)(q % nbk);
        const double cx = ext[0] + (qi + 0.5) * (ext[3] - ext[0]) / nbk;
        const double cy = ext[1] + (qj + 0.5) * (ext[4] - ext[1]) / nbk;
        const double cz = ext[2] + (qk + 0.5) * (ext[5] - ext[2]) / nbk;
        double bd = DBL_MAX;
        for (int32_t t = cnt[q]; t < cnt[q + 1]; t++) {
            const VoroSite& s = g->hsite[bid[t]];
            const double d = (s.x - cx) * (s.x - cx) + (s.y - cy) * (s.y - cy) + (s.z - cz) * (s.z - cz);
            if (d < bd) { bd = d; rep[q] = bid[t]; }
        }
        frontier.push_back((int32_t)q);
    }
    while (!frontier.empty()) {
        next.clear();
        for (int32_t q : frontier) {
            const int qi = q / (nbk * nbk), qj = q / nbk % nbk, qk = q % nbk;
            const int nbq[6][3] = {{qi - 1, qj, qk}, {qi + 1, qj, qk}, {qi, qj - 1, qk},
                                   {qi, qj + 1, qk}, {qi, qj, qk - 1}, {qi, qj, qk + 1}};
            for (const auto& c : nbq) {
                if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= nbk || c[1] >= nbk || c[2] >= nbk) continue;
                const int32_t r = (c[0] * nbk + c[1]) * nbk + c[2];
                if (rep[r] < 0) { rep[r] = rep[q]; next.push_back(r); }
            }
        }
        frontier.swap(next);
    }
    g->boff.upload(cnt);
    g->bids.upload(bid);
    g->brep.upload(rep);
    VoroView& v = g->voro;
    v.site = g->site.as<VoroSite>();
    v.off = g->voff.as<int2>();
    v.nbr = g->vnb.as<VoroNbr>();
    v.boff = g->boff.as<int32_t>();
    v.bids = g->bids.as<int32_t>();
    v.brep = g->brep.as<int32_t>();
    v.nsites = (int32_t)n;
    v.nbk = nbk;
    for (int i = 0; i < 6; i++) v.ext[i] = ext[i];
    double wx = ext[3] - ext[0], wy = ext[4] - ext[1], wz = ext[5] - ext[2];
    v.eps = 1e-12 * std::sqrt(wx * wx + wy * wy + wz * wz);
    *out = g;
    SG_CATCH
}

int sg_grid_destroy(sg_grid* g)
{
    delete g;
    return 0;
}

int sg_grid_num_cells(const sg_grid* g, int64_t* n)
{
    *n = g->ncells;
    return 0;
}

void sg_grid::set_rho(const double* r)
{
    if (kind == GK_CART) {
        SG_HIP(hipMemcpy(rho.p, r, ncells * sizeof(double), hipMemcpyHostToDevice));
        std::vector<float> f(r, r + ncells);
        SG_HIP(hipMemcpy(rhof.p, f.data(), ncells * sizeof(float), hipMemcpyHostToDevice));
    } else if (kind == GK_TREE) {
        for (int64_t c = 0; c < ncells; c++) {
            hrec64[c].rho = r[c];
            float f = (float)r[c];
            std::memcpy(&hnode32[hrec64[c].node].val, &f, sizeof f);
        }
        SG_HIP(hipMemcpy(rec64.p, hrec64.data(), ncells * sizeof(TreeRec64), hipMemcpyHostToDevice));
        SG_HIP(hipMemcpy(node32.p, hnode32.data(), hnode32.size() * sizeof(TreeNode32), hipMemcpyHostToDevice));
    } else {
        for (int64_t c = 0; c < ncells; c++) hsite[c].rho = r[c];
        SG_HIP(hipMemcpy(site.p, hsite.data(), ncells * sizeof(VoroSite), hipMemcpyHostToDevice));
        // the density in each cell's list header (the double in place of the entry's id and pad)
        std::vector<int32_t> be(2 * ncells);
        SG_HIP(hipMemcpy(be.data(), voff.p, be.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        static_assert(sizeof(VoroNbr) == sizeof(VoroSite) && offsetof(VoroSite, rho) == offsetof(VoroNbr, id),
                      "a list header is a site record in a list entry's place");
        for (int64_t c = 0; c < ncells; c++)
            std::memcpy(reinterpret_cast<char*>(&hvnb[be[2 * c]]) + offsetof(VoroSite, rho), &r[c], sizeof(double));
        SG_HIP(hipMemcpy(vnb.p, hvnb.data(), hvnb.size() * sizeof(VoroNbr), hipMemcpyHostToDevice));
    }
}

// A dust system of several components at wavelength ell: the grid's densities become the combined
// kappa*rho of each cell (KappaRho, DustSystem.cpp:459-485: the sum over components of kappaext(h) rho(m, h),
// in component order from zero, so that walks multiply it by 1 and give the reference's optical depths),
// and albm each cell's albedo, ksca/kext with both sums taken over the components as
// simulateescapeandabsorption takes them (MonteCarloSimulation.cpp:478-490; 0 where kext is 0). Nothing
// to do when the system was last prepared for ell.
static void dust_prepare(sg_dust* d, int ell)
{
    if (d->ncomp < 2 || d->prepared == ell) return;
    const int64_t nc = d->grid->ncells;
    const int nl = d->nlambda, H = d->ncomp;
    std::vector<double> kr(nc), alb(nc);
    for (int64_t m = 0; m < nc; m++) {
        double result = 0, ksca = 0, kext = 0;
        for (int h = 0; h < H; h++) {
            double rho = d->rhoc[(size_t)h * nc + m];
            result += d->kextc[(size_t)h * nl + ell] * rho;
            ksca += rho * d->kscac[(size_t)h * nl + ell];
            kext += rho * d->kextc[(size_t)h * nl + ell];
        }
        kr[m] = result;
        alb[m] = (kext > 0.0) ? ksca / kext : 0.0;
    }
    d->grid->set_rho(kr.data());
    SG_HIP(hipMemcpy(d->albm.p, alb.data(), nc * sizeof(double), hipMemcpyHostToDevice));
    d->prepared = ell;
}

// Host helper: copies a batch of n records to the device and back.
struct Batch {
    DevBuf in, in2, o1, o2, o3;
};

static int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

int sg_grid_whichcell(sg_grid* g, int64_t n, const double* r, int32_t* m)
{
    SG_TRY
    ensure_init();
    if (n <= 0) return 0;
    Batch b;
    b.in.upload(r, 3 * n);
    b.o1.alloc(n * sizeof(int32_t));
    int nb = blocks_for(n, 256);
    if (g->kind == GK_CART) hipLaunchKernelGGL(k_whichcell_cart, dim3(nb), dim3(256), 0, 0, g->cart, n, b.in.as<double>(), b.o1.as<int32_t>());
    else if (g->kind == GK_TREE) hipLaunchKernelGGL(k_whichcell_tree, dim3(nb), dim3(256), 0, 0, g->tree, n, b.in.as<double>(), b.o1.as<int32_t>());
    else hipLaunchKernelGGL(k_whichcell_voro, dim3(nb), dim3(256), 0, 0, g->voro, n, b.in.as<double>(), b.o1.as<int32_t>());
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpy(m, b.o1.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    SG_CATCH
}

} // extern "C"
template <class W>
static void launch_paths(const typename W::View& v, int64_t n, const double* rk, int maxseg, int32_t* c, int32_t* m,
                         double* ds)
{
    hipLaunchKernelGGL(k_paths<W>, dim3(blocks_for(n, 256)), dim3(256), 0, 0, v, n, rk, maxseg, c, m, ds);
    SG_HIP(hipGetLastError());
}
extern "C" {

int sg_grid_paths(sg_grid* g, int64_t n, const double* rk, int max_seg, int precision, int32_t* counts, int32_t* m,
                  double* ds)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(max_seg > 0, "max_seg must be positive");
    if (n <= 0) return 0;
    SG_REQUIRE((double)n * max_seg * 12.0 < 64e9, "output too large");
    Batch b;
    b.in.upload(rk, 6 * n);
    b.o1.alloc(n * sizeof(int32_t));
    b.o2.alloc((size_t)n * max_seg * sizeof(int32_t));
    b.o3.alloc((size_t)n * max_seg * sizeof(double));
    const double* d = b.in.as<double>();
    int32_t* oc = b.o1.as<int32_t>();
    int32_t* om = b.o2.as<int32_t>();
    double* od = b.o3.as<double>();
    bool mixed = precision == SG_MIXED;
    if (g->kind == GK_CART) {
        if (mixed) launch_paths<CartW32>(g->cart, n, d, max_seg, oc, om, od);
        else launch_paths<CartW64>(g->cart, n, d, max_seg, oc, om, od);
    } else if (g->kind == GK_TREE) {
        if (!mixed && g->tree.search == SG_BOOKKEEPING) launch_paths<TreeW64Book>(g->tree, n, d, max_seg, oc, om, od);
        else if (!mixed) launch_paths<TreeW64>(g->tree, n, d, max_seg, oc, om, od);
        else if (rope_walk(g)) launch_paths<TreeW32>(g->tree, n, d, max_seg, oc, om, od);
        else launch_paths<TreeW32TopDown>(g->tree, n, d, max_seg, oc, om, od);
    } else {
        if (mixed) launch_paths<VoroW<float>>(g->voro, n, d, max_seg, oc, om, od);
        else launch_paths<VoroW<double>>(g->voro, n, d, max_seg, oc, om, od);
    }
    SG_HIP(hipDeviceSynchronize());
    SG_HIP(hipMemcpy(counts, oc, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    SG_HIP(hipMemcpy(m, om, (size_t)n * max_seg * sizeof(int32_t), hipMemcpyDeviceToHost));
    SG_HIP(hipMemcpy(ds, od, (size_t)n * max_seg * sizeof(double), hipMemcpyDeviceToHost));
    SG_CATCH
}

// ---------------------------------------------------------------- dust

int sg_dust_create(sg_grid* g, const double* rho, int64_t ncells, int nlambda, const double* kappaext,
                   const double* albedo, const double* asymmpar, int store_absorption, sg_dust** out)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(ncells == g->ncells, "density array size differs from the number of cells");
    SG_REQUIRE(nlambda >= 1, "need at least one wavelength");
    for (int64_t i = 0; i < ncells; i++)
        SG_REQUIRE(rho[i] >= 0 && std::isfinite(rho[i]), "dust density must be finite and non-negative");
    auto* d = new sg_dust();
    d->grid = g;
    d->nlambda = nlambda;
    d->kext.assign(kappaext, kappaext + nlambda);
    d->albedo.assign(albedo, albedo + nlambda);
    d->g.assign(asymmpar, asymmpar + nlambda);
    d->store_abs = store_absorption != 0;
    g->set_rho(rho);
    if (d->store_abs) {
        d->labs.alloc((size_t)nlambda * ncells * sizeof(double));
        d->labs.zero();
    }
    *out = d;
    SG_CATCH
}

int sg_dust_create_multi(sg_grid* g, int ncomp, const double* rho, int64_t ncells, int nlambda, const double* kappaext,
                         const double* kappasca, const double* albedo, const double* asymmpar, int store_absorption,
                         sg_dust** out)
{
    if (ncomp == 1) return sg_dust_create(g, rho, ncells, nlambda, kappaext, albedo, asymmpar, store_absorption, out);
    SG_TRY
    ensure_init();
    SG_REQUIRE(ncomp >= 2 && ncomp <= SG_MAX_DUST, "between 1 and 8 dust components are supported");
    SG_REQUIRE(ncells == g->ncells, "density array size differs from the number of cells");
    SG_REQUIRE(nlambda >= 1, "need at least one wavelength");
    const size_t n = (size_t)ncomp * ncells;
    for (size_t i = 0; i < n; i++)
        SG_REQUIRE(rho[i] >= 0 && std::isfinite(rho[i]), "dust density must be finite and non-negative");
    std::unique_ptr<sg_dust> d(new sg_dust());
    d->grid = g;
    d->nlambda = nlambda;
    d->ncomp = ncomp;
    d->kext.assign(nlambda, 1.0);
    d->albedo.assign(nlambda, 0.0);
    d->g.assign(nlambda, 0.0);
    d->rhoc.assign(rho, rho + n);
    d->kextc.assign(kappaext, kappaext + (size_t)ncomp * nlambda);
    d->kscac.assign(kappasca, kappasca + (size_t)ncomp * nlambda);
    d->gc.assign(asymmpar, asymmpar + (size_t)ncomp * nlambda);
    d->store_abs = store_absorption != 0;
    d->drhoc.upload(d->rhoc);
    d->albm.alloc((size_t)ncells * sizeof(double));
    dust_prepare(d.get(), 0);
    if (d->store_abs) {
        d->labs.alloc((size_t)nlambda * ncells * sizeof(double));
        d->labs.zero();
    }
    *out = d.release();
    SG_CATCH
}

int sg_dust_destroy(sg_dust* d)
{
    delete d;
    return 0;
}

int sg_grid_resample_density(sg_grid* g, const double* geom16, double scale, int nsamples, uint32_t seed)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(g && geom16, "null grid or geometry record");
    SG_REQUIRE(g->kind == GK_CART, "device density sampling needs a Cartesian grid (tree and Voronoi grids take "
                                   "their structure from the density: a new density means a new grid)");
    SG_REQUIRE((int)geom16[0] == 1 && geom16[1] > 0 && geom16[2] > 0, "the dust geometry must be an exponential disc");
    SG_REQUIRE(nsamples >= 1 && scale >= 0 && std::isfinite(scale), "need at least one sample and a finite scale");
    const DiscDev d = disc_dev(geom16);
    const CartView& c = g->cart;
    hipLaunchKernelGGL(k_resample_cart, dim3(blocks_for(g->ncells, 256)), dim3(256), 0, 0, c.xv, c.yv, c.zv, c.Nx,
                       c.Ny, c.Nz, d, scale, nsamples, seed, g->rho.as<double>(), g->rhof.as<float>());
    SG_HIP(hipGetLastError());
    SG_CATCH
}


} // extern "C"
template <class W>
static void launch_tau(const typename W::View& v, int64_t n, const double* rk, double kext, const double* taus,
                       int mode, double* out)
{
    hipLaunchKernelGGL(k_tau<W>, dim3(blocks_for(n, 256)), dim3(256), 0, 0, v, n, rk, kext, taus, mode, out);
    SG_HIP(hipGetLastError());
}

static int tau_common(sg_dust* d, int ell, int64_t n, const double* rk, const double* taus, int precision, int mode,
                      double* out)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(ell >= 0 && ell < d->nlambda, "wavelength index out of range");
    if (n <= 0) return 0;
    Batch b;
    b.in.upload(rk, 6 * n);
    if (taus) b.in2.upload(taus, n);
    b.o1.alloc(n * sizeof(double));
    const double* ri = b.in.as<double>();
    const double* ti = taus ? b.in2.as<double>() : nullptr;
    double* o = b.o1.as<double>();
    sg_grid* g = d->grid;
    bool mixed = precision == SG_MIXED;
    dust_prepare(d, ell);
    double ke = d->kext[ell];
    if (g->kind == GK_CART) {
        if (mixed) launch_tau<CartW32>(g->cart, n, ri, ke, ti, mode, o);
        else launch_tau<CartW64>(g->cart, n, ri, ke, ti, mode, o);
    } else if (g->kind == GK_TREE) {
        if (!mixed && g->tree.search == SG_BOOKKEEPING) launch_tau<TreeW64Book>(g->tree, n, ri, ke, ti, mode, o);
        else if (!mixed) launch_tau<TreeW64>(g->tree, n, ri, ke, ti, mode, o);
        else if (rope_walk(g)) launch_tau<TreeW32>(g->tree, n, ri, ke, ti, mode, o);
        else launch_tau<TreeW32TopDown>(g->tree, n, ri, ke, ti, mode, o);
    } else {
        if (mixed) launch_tau<VoroW<float>>(g->voro, n, ri, ke, ti, mode, o);
        else launch_tau<VoroW<double>>(g->voro, n, ri, ke, ti, mode, o);
    }
    SG_HIP(hipDeviceSynchronize());
    SG_HIP(hipMemcpy(out, o, n * sizeof(double), hipMemcpyDeviceToHost));
    SG_CATCH
}
extern "C" {

int sg_dust_optical_depths(sg_dust* d, int ell, int64_t n, const double* rk, int precision, double* tau)
{
    return tau_common(d, ell, n, rk, nullptr, precision, 0, tau);
}

int sg_dust_path_lengths(sg_dust* d, int ell, int64_t n, const double* rk, const double* taus, int precision,
                         double* s)
{
    return tau_common(d, ell, n, rk, taus, precision, 1, s);
}

int sg_dust_read_absorption(sg_dust* d, double* labs)
{
    SG_TRY
    SG_REQUIRE(d->store_abs, "absorption rates are not stored for this dust system");
    int64_t nc = d->grid->ncells;
    std::vector<double> t((size_t)d->nlambda * nc);
    SG_HIP(hipMemcpy(t.data(), d->labs.p, t.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t m = 0; m < nc; m++)
        for (int ell = 0; ell < d->nlambda; ell++) labs[m * d->nlambda + ell] = t[(size_t)ell * nc + m];
    SG_CATCH
}

int sg_dust_clear_absorption(sg_dust* d)
{
    SG_TRY
    if (d->store_abs) d->labs.zero();
    SG_CATCH
}

// ---------------------------------------------------------------- stars

// StellarSystem and its components' setup from the parameter records (the Sersic tables to the device).
static void stars_fill(sg_stars* s, int ncomp, const double* params, const double* lum, const double* sersic_s,
                       const double* sersic_M, int ntable, double emission_bias)
{
    SG_REQUIRE(ncomp >= 1 && ncomp <= SG_MAX_STARS, "between 1 and 4 stellar components are supported");
    SG_REQUIRE(emission_bias >= 0 && emission_bias <= 1, "emission bias must be in [0,1]");
    for (int c = 0; c < ncomp; c++) {
        int kind = (int)params[c * SG_STAR_NPARAM];
        SG_REQUIRE(kind == 1 || kind == 2, "stellar geometry must be 1 (exponential disc) or 2 (Sersic)");
        if (kind == 2) SG_REQUIRE(sersic_s && sersic_M && ntable >= 2, "Sersic component needs its mass table");
    }
    s->params.assign(params, params + ncomp * SG_STAR_NPARAM);
    s->lum.assign(lum, lum + (size_t)ncomp * s->nlambda);
    s->bias = emission_bias;
    while ((int)s->tables.size() < ncomp) s->tables.push_back(new DevBuf());
    int nser = 0;
    for (int c = 0; c < ncomp; c++) {
        if ((int)params[c * SG_STAR_NPARAM] != 2) continue;
        std::vector<double> both;
        both.insert(both.end(), sersic_s + nser * ntable, sersic_s + (nser + 1) * ntable);
        both.insert(both.end(), sersic_M + nser * ntable, sersic_M + (nser + 1) * ntable);
        DevBuf* tb = s->tables[c];
        if (tb->bytes == both.size() * sizeof(double))
            SG_HIP(hipMemcpy(tb->p, both.data(), tb->bytes, hipMemcpyHostToDevice));
        else
            tb->upload(both);
        nser++;
    }
    s->ts.resize(ncomp);
}

int sg_stars_create(int ncomp, const double* params, int nlambda, const double* lum, const double* sersic_s,
                    const double* sersic_M, int ntable, double emission_bias, sg_stars** out)
{
    SG_TRY
    ensure_init();
    std::unique_ptr<sg_stars> s(new sg_stars());
    s->nlambda = nlambda;
    stars_fill(s.get(), ncomp, params, lum, sersic_s, sersic_M, ntable, emission_bias);
    *out = s.release();
    SG_CATCH
}

int sg_stars_update(sg_stars* s, int ncomp, const double* params, const double* lum, const double* sersic_s,
                    const double* sersic_M, int ntable, double emission_bias)
{
    SG_TRY
    SG_REQUIRE(s && params && lum, "null stellar system or parameters");
    SG_REQUIRE(ncomp == (int)(s->params.size() / SG_STAR_NPARAM), "a stellar system update keeps the number of components");
    stars_fill(s, ncomp, params, lum, sersic_s, sersic_M, ntable, emission_bias);
    SG_CATCH
}

int sg_stars_destroy(sg_stars* s)
{
    delete s;
    return 0;
}

// ---------------------------------------------------------------- instruments

// DistantInstrument::setupSelfBefore (DistantInstrument.cpp:27-50) and SingleFrameInstrument::setupSelfBefore
// (SingleFrameInstrument.cpp:26-42) for one parameter record; the frame pointer is left null.
static InstrDev instrument_setup(const double* p)
{
    InstrDev d{};
    double incl = p[1], azim = p[2], pa = p[3];
    d.costheta = std::cos(incl); d.sintheta = std::sin(incl);
    d.cosphi = std::cos(azim); d.sinphi = std::sin(azim);
    d.cospa = std::cos(pa); d.sinpa = std::sin(pa);
    const double eps = 1e-8;
    if (incl <= eps) { d.kobs[0] = 0; d.kobs[1] = 0; d.kobs[2] = 1; }
    else if (incl >= M_PI - eps) { d.kobs[0] = 0; d.kobs[1] = 0; d.kobs[2] = -1; }
    else {
        double st = std::sin(incl);
        d.kobs[0] = st * std::cos(azim); d.kobs[1] = st * std::sin(azim); d.kobs[2] = std::cos(incl);
    }
    d.Nx = (int)p[4]; d.Ny = (int)p[5];
    SG_REQUIRE(d.Nx > 0 && d.Ny > 0 && p[6] > 0 && p[7] > 0, "instrument needs pixels and a field of view");
    d.xpmin = p[8] - 0.5 * p[6]; d.xpsiz = p[6] / d.Nx;
    d.ypmin = p[9] - 0.5 * p[7]; d.ypsiz = p[7] / d.Ny;
    return d;
}

int sg_instruments_create(int n, const double* params, int nlambda, sg_instruments** out)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(n >= 0 && n <= SG_MAX_INSTR, "at most 8 instruments per simulation");
    auto* in = new sg_instruments();
    in->n = n;
    in->nlambda = nlambda;
    for (int i = 0; i < n; i++) {
        InstrDev d = instrument_setup(params + i * SG_INSTR_NPARAM);
        auto* f = new DevBuf();
        f->alloc((size_t)nlambda * d.Nx * d.Ny * sizeof(double));
        f->zero();
        d.frame = f->as<double>();
        in->frames.push_back(f);
        in->dev.push_back(d);
    }
    *out = in;
    SG_CATCH
}

int sg_instruments_update(sg_instruments* in, const double* params)
{
    SG_TRY
    SG_REQUIRE(in && params, "null instrument system or parameters");
    std::vector<InstrDev> next(in->n);
    for (int i = 0; i < in->n; i++) {
        next[i] = instrument_setup(params + i * SG_INSTR_NPARAM);
        SG_REQUIRE(next[i].Nx == in->dev[i].Nx && next[i].Ny == in->dev[i].Ny,
                   "an instrument update keeps the number of pixels");
        next[i].frame = in->dev[i].frame;
    }
    in->dev = next;
    SG_CATCH
}

int sg_instruments_destroy(sg_instruments* in)
{
    delete in;
    return 0;
}

int sg_instruments_read(sg_instruments* in, int i, double* frame)
{
    SG_TRY
    SG_REQUIRE(i >= 0 && i < in->n, "instrument index out of range");
    SG_HIP(hipMemcpy(frame, in->frames[i]->p, in->frames[i]->bytes, hipMemcpyDeviceToHost));
    SG_CATCH
}

int sg_instruments_clear(sg_instruments* in)
{
    SG_TRY
    for (auto* f : in->frames) f->zero();
    SG_CATCH
}

// ---------------------------------------------------------------- simulation

int sg_sim_create(sg_dust* dust, sg_stars* stars, sg_instruments* instr, const double* mc, sg_sim** out)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(stars && instr, "a simulation needs a stellar system and an instrument system");
    SG_REQUIRE(mc[1] > 0, "minWeightReduction must be positive");
    SG_REQUIRE(mc[3] >= 0 && mc[3] <= 1, "scattBias must be in [0,1]");
    if (dust) SG_REQUIRE(dust->nlambda == stars->nlambda && dust->nlambda == instr->nlambda,
                         "dust, stars and instruments disagree on the number of wavelengths");
    auto* s = new sg_sim();
    s->dust = dust; s->stars = stars; s->instr = instr;
    std::memcpy(s->mc, mc, sizeof s->mc);
    s->counters.alloc(SG_NCOUNTERS * sizeof(unsigned long long));
    s->next.alloc(sizeof(unsigned long long));
    s->err.alloc(sizeof(int));
    s->scene.alloc(sizeof(SceneDev));
    *out = s;
    SG_CATCH
}

int sg_sim_destroy(sg_sim* s)
{
    delete s;
    return 0;
}
} // extern "C"

// Device time of a series of transport launches: an event pair around each launch on the default
// stream, summed once the stream has drained.
// Device time of a run's transport: the span of its passes (start/stop on stream 0, or a span
// measured elsewhere) and, per launch of the walk kernel, that launch's own duration between HIP
// events recorded on its stream just before and after it. The per-launch events come from a
// process-wide pool that is reused from call to call (one host thread drives the library).
struct LaunchTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    static std::vector<std::pair<hipEvent_t, hipEvent_t>>& kpool()
    {
        static std::vector<std::pair<hipEvent_t, hipEvent_t>> p;
        return p;
    }
    size_t kused = 0;
    void start()
    {
        hipEvent_t a, b;
        SG_HIP(hipEventCreate(&a));
        SG_HIP(hipEventCreate(&b));
        ev.push_back({a, b});
        SG_HIP(hipEventRecord(a, 0));
    }
    void stop() { SG_HIP(hipEventRecord(ev.back().second, 0)); }
    // a span measured elsewhere that covered `launches` launches
    void add(double ms, int launches) { extra_ms += ms; extra_launches += launches; }
    // brackets one walk-kernel launch on stream s
    void kbefore(hipStream_t s)
    {
        auto& p = kpool();
        if (kused == p.size()) {
            hipEvent_t a, b;
            SG_HIP(hipEventCreate(&a));
            SG_HIP(hipEventCreate(&b));
            p.push_back({a, b});
        }
        SG_HIP(hipEventRecord(p[kused].first, s));
    }
    void kafter(hipStream_t s) { SG_HIP(hipEventRecord(kpool()[kused++].second, s)); }
    double extra_ms = 0;
    int extra_launches = 0;
    int launches() const { return (int)ev.size() + extra_launches; }
    double total_ms()
    {
        double t = extra_ms;
        for (auto& p : ev) {
            float ms = 0;
            SG_HIP(hipEventSynchronize(p.second));
            SG_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            t += ms;
        }
        return t;
    }
    // summed per-launch durations: the bracketed launches and the start/stop spans (one kernel each)
    double kernel_ms()
    {
        double t = total_ms() - extra_ms;
        auto& p = kpool();
        for (size_t i = 0; i < kused; i++) {
            float ms = 0;
            SG_HIP(hipEventSynchronize(p[i].second));
            SG_HIP(hipEventElapsedTime(&ms, p[i].first, p[i].second));
            t += ms;
        }
        return t;
    }
    ~LaunchTimer()
    {
        for (auto& p : ev) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
    }
};

static void fill_args(sg_sim* s, int ell, uint64_t first, uint64_t count, uint64_t npp, McArgs& a)
{
    std::memset(&a, 0, sizeof a);
    SceneDev sd;
    std::memset(&sd, 0, sizeof sd);
    sg_stars* st = s->stars;
    int ns = (int)(st->params.size() / SG_STAR_NPARAM);
    SG_REQUIRE(ell >= 0 && ell < st->nlambda, "wavelength index out of range");
    double Ltot = 0;
    for (int c = 0; c < ns; c++) Ltot += st->lum[(size_t)c * st->nlambda + ell];
    a.first = first; a.count = count;
    a.Lpkt = Ltot / (double)npp;
    a.Lthreshold = a.Lpkt / s->mc[1];
    a.ell = ell; a.nlambda = st->nlambda;
    a.minfs = (int)s->mc[2];
    a.xi = s->mc[3];
    a.seed = (uint32_t)s->mc[4];
    a.member = (uint32_t)s->mc[6];
    a.tau_cut = s->mc[7];
    a.has_dust = s->dust != nullptr;
    if (s->dust) {
        dust_prepare(s->dust, ell);
        a.kext = s->dust->kext[ell]; a.albedo = s->dust->albedo[ell]; a.g = s->dust->g[ell];
        a.ncomp = s->dust->ncomp;
        if (a.ncomp > 1) {
            a.rhoc = s->dust->drhoc.as<double>();
            a.albm = s->dust->albm.as<double>();
            for (int h = 0; h < a.ncomp; h++) {
                sd.ksca[h] = s->dust->kscac[(size_t)h * s->dust->nlambda + ell];
                sd.gc[h] = s->dust->gc[(size_t)h * s->dust->nlambda + ell];
            }
        }
        a.store_abs = s->dust->store_abs;
        a.labs = s->dust->store_abs ? s->dust->labs.as<double>() : nullptr;
        a.ncells = s->dust->grid->ncells;
        // the raster of start points for the walk order: the grid's box
        const sg_grid* gr = s->dust->grid;
        double box[6];
        if (gr->kind == GK_CART) {
            const CartView& c = gr->cart;
            box[0] = c.xmin; box[1] = c.ymin; box[2] = c.zmin; box[3] = c.xmax; box[4] = c.ymax; box[5] = c.zmax;
        } else if (gr->kind == GK_TREE) for (int i = 0; i < 6; i++) box[i] = gr->tree.ext[i];
        else for (int i = 0; i < 6; i++) box[i] = gr->voro.ext[i];
        for (int i = 0; i < 3; i++) {
            a.key_lo[i] = box[i];
            a.key_scale[i] = box[i + 3] > box[i] ? 16.0 / (box[i + 3] - box[i]) : 0.0;
        }
    }
    a.ni = s->instr->n;
    for (int i = 0; i < a.ni; i++) sd.instr[i] = s->instr->dev[i];
    a.ns = ns;
    a.emission_bias = st->bias;
    // StellarSystem setup: normalised cumulative luminosity (NR::cdf)
    sd.Xcum[0] = 0;
    for (int c = 0; c < ns; c++) sd.Xcum[c + 1] = sd.Xcum[c] + st->lum[(size_t)c * st->nlambda + ell];
    for (int c = 0; c <= ns; c++) sd.Xcum[c] /= sd.Xcum[ns];
    a.Lmean = Ltot / ns;
    for (int c = 0; c < ns; c++) {
        const double* p = &st->params[c * SG_STAR_NPARAM];
        StarDev& d = sd.star[c];
        d.kind = (int)p[0];
        for (int i = 0; i < 5; i++) d.p[i] = p[1 + i];
        if (d.kind == 2) { d.p[0] = p[1]; d.p[1] = p[2]; }
        d.spiral = p[6] != 0;
        for (int i = 0; i < 6; i++) d.sp[i] = p[7 + i];
        if (d.spiral) {
            // SpiralStructureGeometryDecorator::setupSelfBefore (SpiralStructureGeometryDecorator.cpp:23-40)
            d.tanp = std::tan(p[8]);
            double N = p[12];
            d.CN = std::sqrt(M_PI) * std::exp(lngamma(N + 1.0)) / std::exp(lngamma(N + 0.5));
        }
        if (d.kind == 2) {
            DevBuf* tb = st->tables[c];
            int nt = (int)(tb->bytes / sizeof(double) / 2);
            d.ts = tb->as<double>(); d.tM = tb->as<double>() + nt; d.nt = nt;
        }
        d.lum = st->lum[(size_t)c * st->nlambda + ell];
    }
    a.counters = s->counters.as<unsigned long long>();
    a.next_packet = s->next.as<unsigned long long>();
    a.error = s->err.as<int>();
    a.compact = g_tune_compact > 0 ? g_tune_compact : (g_pipeline == 1 ? 24 : 32);
    SG_HIP(hipMemcpy(s->scene.p, &sd, sizeof sd, hipMemcpyHostToDevice));
    a.scene = s->scene.as<SceneDev>();
}

template <class W, bool ABS, bool MULTI = false>
static void launch_mc(const typename W::View& v, const McArgs& a, int& blocks, DevBuf& cold, DevBuf& args,
                      int nlds = 0)
{
    using R = typename W::Real;
    int bs = g_tune_block;
    // LDS: the hot-state parking area (only when lanes are re-armed out of line), then the top tree nodes
    size_t shm = (SG_PARK ? hot_words<W>() * 4 * (size_t)bs : 0) + (size_t)nlds * sizeof(int32_t);
    int per = 0;
    SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_mc<W, ABS, MULTI>, bs, shm));
    if (per < 1) per = 1;
    if (g_tune_waves > 0) per = std::max(1, std::min(per, g_tune_waves * 64 / bs));
    blocks = per * g_cus;
    int nslots = blocks * bs;
    size_t need = cold_bytes<R>(nslots);
    if (cold.bytes < need) cold.alloc(need);
    R* cr = cold.as<R>();
    uint32_t* cu = reinterpret_cast<uint32_t*>(cold.as<unsigned char>() + (size_t)nslots * CR_NR * sizeof(R));
    // device copies of the view and arguments for the re-arming function
    const size_t off = (sizeof(v) + 255) / 256 * 256;
    if (args.bytes < off + sizeof(McArgs)) args.alloc(off + sizeof(McArgs));
    SG_HIP(hipMemcpy(args.p, &v, sizeof(v), hipMemcpyHostToDevice));
    SG_HIP(hipMemcpy(args.as<unsigned char>() + off, &a, sizeof(McArgs), hipMemcpyHostToDevice));
    const auto* vd = reinterpret_cast<const typename W::View*>(args.p);
    const auto* ad = reinterpret_cast<const McArgs*>(args.as<unsigned char>() + off);
    hipLaunchKernelGGL((k_mc<W, ABS, MULTI>), dim3(blocks), dim3(bs), shm, 0, v, a, nlds, cr, cu, nslots, vd, ad);
    SG_HIP(hipGetLastError());
}

template <bool ABS, bool MULTI = false> static void dispatch_mc(sg_sim* s, const McArgs& a, int precision, int& blocks)
{
    bool mixed = precision == SG_MIXED;
    sg_grid* g = s->dust ? s->dust->grid : nullptr;
    if (!g) {
        // no dust system: peel-offs need no walk; any grid kind serves as a placeholder
        CartView v{};
        launch_mc<CartW64, false>(v, a, blocks, s->cold, s->args);
        return;
    }
    if (g->kind == GK_CART) {
        if (mixed) launch_mc<CartW32, ABS, MULTI>(g->cart, a, blocks, s->cold, s->args);
        else launch_mc<CartW64, ABS, MULTI>(g->cart, a, blocks, s->cold, s->args);
    } else if (g->kind == GK_TREE) {
        if (!mixed && g->tree.search == SG_BOOKKEEPING)
            launch_mc<TreeW64Book, ABS, MULTI>(g->tree, a, blocks, s->cold, s->args, 0);
        else if (!mixed) launch_mc<TreeW64, ABS, MULTI>(g->tree, a, blocks, s->cold, s->args, SG_LDS_NODES);
        else if (rope_walk(g)) launch_mc<TreeW32, ABS, MULTI>(g->tree, a, blocks, s->cold, s->args, SG_LDS_NODES);
        else launch_mc<TreeW32TopDown, ABS, MULTI>(g->tree, a, blocks, s->cold, s->args, SG_LDS_NODES);
    } else {
        if (mixed) launch_mc<VoroW<float>, ABS, MULTI>(g->voro, a, blocks, s->cold, s->args);
        else launch_mc<VoroW<double>, ABS, MULTI>(g->voro, a, blocks, s->cold, s->args);
    }
}

// The lane pipeline's kernel for a run's arguments: absorption stored or not, one or several dust components.
static void dispatch_mc_any(sg_sim* s, const McArgs& a, int precision, int& blocks)
{
    if (a.ncomp > 1) {
        if (a.store_abs) dispatch_mc<true, true>(s, a, precision, blocks);
        else dispatch_mc<false, true>(s, a, precision, blocks);
    } else {
        if (a.store_abs) dispatch_mc<true>(s, a, precision, blocks);
        else dispatch_mc<false>(s, a, precision, blocks);
    }
}
// ---------------------------------------------------------------- split pipeline

__global__ void k_revive(uint32_t* state, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && state[i] == PS_DEAD) state[i] = PS_FRESH;
}

static const int64_t kSplitSlots = (int64_t)1 << 22;
static const uint64_t kSplitEmit = (uint64_t)1 << 25;   // packets emitted at once
static const int kSplitCheck = 4;                       // passes between progress checks

// Two non-blocking streams of default priority, one per slot group, and a pinned progress buffer.
static hipStream_t g_split_stream[2] = {nullptr, nullptr};
static unsigned long long* g_split_host = nullptr;
// Work buffers of the split pipeline and of the emission, shared by every simulation of the process
// (runs are synchronous and driven by one host thread): slot groups, walk queues, progress counters,
// launch states. Allocated on first use, grown as needed, kept until the process ends.
struct SplitPool {
    DevBuf slots[2], queue[2], prog[2], emitted, emitdraw;
    DevBuf bins[2];   // per group: bin counts (zero between passes: cleared once, then by each pass's scan), offsets
};
static SplitPool* g_pool = nullptr;
static SplitPool& pool()
{
    if (!g_pool) g_pool = new SplitPool();
    return *g_pool;
}

static void split_resources()
{
    if (g_split_stream[0]) return;
    for (int g = 0; g < 2; g++) SG_HIP(hipStreamCreateWithFlags(&g_split_stream[g], hipStreamNonBlocking));
    SG_HIP(hipHostMalloc(reinterpret_cast<void**>(&g_split_host), 8 * sizeof(unsigned long long), hipHostMallocDefault));
}

// Packets [first, first+count) through event and walk passes over two slot groups that alternate on two
// streams, so that one group's event pass and the drain of its walk pass overlap the other group's
// walks. Up to 2^25 packets are emitted at once; the host reads the progress every few passes and
// stops when every packet has been claimed and neither group queued a walk in its last pass.
template <class W, bool ABS>
static void run_split(sg_sim* s, const typename W::View& v, const McArgs& a, uint64_t first, uint64_t count, int nlds,
                      LaunchTimer& lt, int& blocks)
{
    using R = typename W::Real;
    split_resources();
    hipStream_t* st = g_split_stream;
    const int64_t slots = g_tune_slots > 0 ? g_tune_slots : kSplitSlots;
    const int64_t P = (int64_t)std::max<uint64_t>(1, std::min<uint64_t>(count, (uint64_t)slots));
    const int G = (P >= 2 && g_slot_groups == 2) ? 2 : 1;
    const int64_t Pg[2] = {G == 2 ? (P + 1) / 2 : P, G == 2 ? P / 2 : 0};
    Slots<R> sl[2];
    WalkQueue<R> q[2];
    unsigned long long* pc[2];   // per group: [1] queue tail, [3] requests claimed
    for (int g = 0; g < G; g++) {
        const int64_t cap = Pg[g] * (a.ni + 1);
        SplitPool& pl = pool();
        const bool sorted = g_walk_order != 0;
        if (pl.slots[g].bytes < slot_bytes<R>(Pg[g])) pl.slots[g].alloc(slot_bytes<R>(Pg[g]));
        if (pl.queue[g].bytes < queue_bytes<R>(cap, sorted)) pl.queue[g].alloc(queue_bytes<R>(cap, sorted));
        if (sorted && pl.bins[g].bytes == 0) {
            pl.bins[g].alloc(2 * (size_t)kKeyBins * sizeof(uint32_t));
            SG_HIP(hipMemsetAsync(pl.bins[g].p, 0, kKeyBins * sizeof(uint32_t), st[g]));
        }
        // [1] queue tail, [3] requests claimed, [4..11] requests claimed per XCD segment
        if (pl.prog[g].bytes < 12 * sizeof(unsigned long long)) pl.prog[g].alloc(12 * sizeof(unsigned long long));
        pc[g] = pl.prog[g].as<unsigned long long>();
        sl[g] = Slots<R>{pl.slots[g].as<R>(),
                         reinterpret_cast<uint32_t*>(pl.slots[g].as<unsigned char>() + (size_t)Pg[g] * SR_NR * sizeof(R)),
                         Pg[g]};
        // requests, then (sorted order) the sorted requests, keys, ranks, bin counts and offsets
        unsigned char* qb = pl.queue[g].as<unsigned char>();
        q[g] = WalkQueue<R>{};
        q[g].req = reinterpret_cast<WalkReq<R>*>(qb);
        q[g].cap = cap;
        q[g].tail = pc[g] + 1;
        if (sorted) {
            unsigned char* p = qb + (size_t)cap * sizeof(WalkReq<R>);
            q[g].sorted = reinterpret_cast<WalkReq<R>*>(p);
            p += (size_t)cap * sizeof(WalkReq<R>);
            q[g].key = reinterpret_cast<uint32_t*>(p);
            q[g].rank = q[g].key + cap;
            q[g].hist = pl.bins[g].as<uint32_t>();
            q[g].off = q[g].hist + kKeyBins;
        }
        // slot states PS_FRESH and slot counts zero (adjacent arrays)
        SG_HIP(hipMemsetAsync(sl[g].u + (size_t)SU_STATE * Pg[g], 0, (size_t)Pg[g] * (SU_NU - SU_STATE) * sizeof(uint32_t),
                              st[g]));
    }
    const int bs = g_tune_block;
    const size_t shm = (size_t)nlds * sizeof(int32_t);
    int per = 0;
    SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_walks<W, ABS>, bs, shm));
    blocks = std::max(1, per) * g_cus;
    // event workgroups stage up to blockDim * (ni + 1) requests in LDS
    int ebs = 256;
    while (ebs > 64 && staged_bytes<R>(ebs, a.ni) > 64 * 1024) ebs /= 2;
    const size_t eshm = staged_bytes<R>(ebs, a.ni);
    const uint64_t batch = std::min<uint64_t>(count, g_emit_batch ? g_emit_batch : kSplitEmit);
    const size_t rsz = sizeof(R);
    DevBuf& emitted = pool().emitted;
    DevBuf& emitdraw = pool().emitdraw;
    if (emitted.bytes < 7 * batch * rsz) emitted.alloc(7 * batch * rsz);
    if (emitdraw.bytes < batch * sizeof(uint32_t)) emitdraw.alloc(batch * sizeof(uint32_t));
    unsigned long long* claims = s->next.as<unsigned long long>();
    hipEvent_t ev_start, ev_join, ev_end[2];
    SG_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    SG_HIP(hipEventCreate(&ev_start));
    for (int g = 0; g < 2; g++) SG_HIP(hipEventCreate(&ev_end[g]));
    int walk_launches = 0;
    for (uint64_t b = 0; b < count; b += batch) {
        McArgs ab = a;
        ab.first = first + b;
        ab.count = std::min(batch, count - b);
        ab.emitted = emitted.p;
        ab.emit_draw = emitdraw.as<uint32_t>();
        ab.next_packet = claims;
        // the previous batch is finished on both streams: emit on stream 0, then both groups start
        if (G == 2) {
            SG_HIP(hipEventRecord(ev_join, st[1]));
            SG_HIP(hipStreamWaitEvent(st[0], ev_join, 0));
        }
        SG_HIP(hipMemsetAsync(claims, 0, sizeof(unsigned long long), st[0]));
        int emb = (int)std::min<uint64_t>((ab.count + 255) / 256, (uint64_t)g_cus * 16);
        hipLaunchKernelGGL(k_emit<R>, dim3(emb), dim3(256), 0, st[0], ab, emitted.as<R>(), emitdraw.as<uint32_t>());
        SG_HIP(hipGetLastError());
        for (int g = 0; g < G && b > 0; g++)
            hipLaunchKernelGGL(k_revive, dim3((Pg[g] + 255) / 256), dim3(256), 0, st[0], sl[g].u + (size_t)SU_STATE * Pg[g],
                               Pg[g]);
        if (b == 0) SG_HIP(hipEventRecord(ev_start, st[0]));
        if (G == 2) {
            SG_HIP(hipEventRecord(ev_join, st[0]));
            SG_HIP(hipStreamWaitEvent(st[1], ev_join, 0));
        }
        for (;;) {
            for (int p = 0; p < kSplitCheck; p++)
                for (int g = 0; g < G; g++) {
                    SG_HIP(hipMemsetAsync(pc[g] + 1, 0, sizeof(unsigned long long), st[g]));
                    hipLaunchKernelGGL(k_events<R>, dim3((Pg[g] + ebs - 1) / ebs), dim3(ebs), eshm, st[g], ab, sl[g], q[g]);
                    SG_HIP(hipGetLastError());
                    if (q[g].sorted) {
                        hipLaunchKernelGGL(k_key_scan, dim3(1), dim3(1024), 0, st[g], q[g].hist, q[g].off);
                        hipLaunchKernelGGL(k_key_scatter<R>, dim3(g_cus * 8), dim3(256), 0, st[g], q[g], ab.error);
                        SG_HIP(hipGetLastError());
                    }
                    SG_HIP(hipMemsetAsync(pc[g] + 3, 0, (1 + kXcdSegs) * sizeof(unsigned long long), st[g]));
                    lt.kbefore(st[g]);
                    hipLaunchKernelGGL((k_walks<W, ABS>), dim3(blocks), dim3(bs), shm, st[g], v, ab, q[g], sl[g], pc[g] + 3,
                                       nlds, g_walk_order == 2 ? kXcdSegs : 0);
                    SG_HIP(hipGetLastError());
                    lt.kafter(st[g]);
                    walk_launches++;
                }
            // progress after the last pass of each group: packets claimed, walks each group queued
            if (G == 2) {
                SG_HIP(hipEventRecord(ev_join, st[1]));
                SG_HIP(hipStreamWaitEvent(st[0], ev_join, 0));
            }
            SG_HIP(hipMemcpyAsync(g_split_host, claims, sizeof(unsigned long long), hipMemcpyDeviceToHost, st[0]));
            for (int g = 0; g < G; g++)
                SG_HIP(hipMemcpyAsync(g_split_host + 1 + g, pc[g] + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                      st[0]));
            SG_HIP(hipStreamSynchronize(st[0]));
            bool queued = false;
            for (int g = 0; g < G; g++) {
                SG_REQUIRE(g_split_host[1 + g] <= (unsigned long long)(Pg[g] * (a.ni + 1)), "walk queue overflow");
                queued = queued || g_split_host[1 + g] > 0;
            }
            if (g_split_host[0] >= ab.count && !queued) break;
        }
    }
    for (int g = 0; g < G; g++)
        hipLaunchKernelGGL(k_slot_counts, dim3(std::min<int64_t>((Pg[g] + 255) / 256, 1024)), dim3(256), 0, st[0], sl[g].u,
                           Pg[g], a.counters);
    SG_HIP(hipGetLastError());
    for (int g = 0; g < G; g++) SG_HIP(hipEventRecord(ev_end[g], st[g]));
    SG_HIP(hipStreamSynchronize(st[0]));
    if (G == 2) SG_HIP(hipStreamSynchronize(st[1]));
    float ms = 0, ms1 = 0;
    SG_HIP(hipEventElapsedTime(&ms, ev_start, ev_end[0]));
    if (G == 2) SG_HIP(hipEventElapsedTime(&ms1, ev_start, ev_end[1]));
    lt.add(std::max(ms, ms1), walk_launches);
    (void)hipEventDestroy(ev_start);
    (void)hipEventDestroy(ev_join);
    for (int g = 0; g < 2; g++) (void)hipEventDestroy(ev_end[g]);
}

template <bool ABS>
static void dispatch_split(sg_sim* s, const McArgs& a, int precision, uint64_t first, uint64_t count, LaunchTimer& lt,
                           int& blocks)
{
    const bool mixed = precision == SG_MIXED;
    sg_grid* g = s->dust ? s->dust->grid : nullptr;
    if (!g) {
        CartView v{};
        if (mixed) run_split<CartW32, false>(s, v, a, first, count, 0, lt, blocks);
        else run_split<CartW64, false>(s, v, a, first, count, 0, lt, blocks);
        return;
    }
    if (g->kind == GK_CART) {
        if (mixed) run_split<CartW32, ABS>(s, g->cart, a, first, count, 0, lt, blocks);
        else run_split<CartW64, ABS>(s, g->cart, a, first, count, 0, lt, blocks);
    } else if (g->kind == GK_TREE) {
        if (!mixed && g->tree.search == SG_BOOKKEEPING) run_split<TreeW64Book, ABS>(s, g->tree, a, first, count, 0, lt, blocks);
        else if (!mixed) run_split<TreeW64, ABS>(s, g->tree, a, first, count, SG_LDS_NODES, lt, blocks);
        else if (rope_walk(g)) run_split<TreeW32, ABS>(s, g->tree, a, first, count, SG_LDS_NODES, lt, blocks);
        else run_split<TreeW32TopDown, ABS>(s, g->tree, a, first, count, SG_LDS_NODES, lt, blocks);
    } else {
        if (mixed) run_split<VoroW<float>, ABS>(s, g->voro, a, first, count, 0, lt, blocks);
        else run_split<VoroW<double>, ABS>(s, g->voro, a, first, count, 0, lt, blocks);
    }
}

extern "C" {

int sg_sim_run(sg_sim* s, int ell, uint64_t first, uint64_t count, uint64_t npp, int precision, uint64_t* counters)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(npp > 0, "npp must be positive");
    McArgs a;
    fill_args(s, ell, first, count, npp, a);
    s->counters.zero();
    s->err.zero();
    const bool mixed = precision == SG_MIXED;
    const size_t rsz = mixed ? sizeof(float) : sizeof(double);
    const uint64_t batch = std::min<uint64_t>(count, g_emit_batch ? g_emit_batch : (uint64_t)1 << 22);
    // several dust components run in the lane pipeline
    const int pipeline = a.ncomp > 1 ? 0 : g_pipeline;
    if (pipeline != g_pipeline && g_tune_compact <= 0) a.compact = 32;
    if (pipeline == 0) {
        if (s->emitted.bytes < 7 * batch * rsz) s->emitted.alloc(7 * batch * rsz);
        if (s->emitdraw.bytes < batch * sizeof(uint32_t)) s->emitdraw.alloc(batch * sizeof(uint32_t));
    }
    hipEvent_t e0, e1;
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    SG_HIP(hipEventRecord(e0, 0));
    int blocks = 0;
    LaunchTimer lt;
    if (pipeline == 1 && count > 0) {
        if (a.store_abs) dispatch_split<true>(s, a, precision, first, count, lt, blocks);
        else dispatch_split<false>(s, a, precision, first, count, lt, blocks);
    }
    // the lane pipeline runs in batches: emission kernel, then the persistent transport kernel
    for (uint64_t b = 0; pipeline == 0 && b < count; b += batch) {
        McArgs ab = a;
        ab.first = first + b;
        ab.count = std::min(batch, count - b);
        ab.emitted = s->emitted.p;
        ab.emit_draw = s->emitdraw.as<uint32_t>();
        s->next.zero();
        int eb = (int)std::min<uint64_t>((ab.count + 255) / 256, (uint64_t)g_cus * 16);
        if (mixed) hipLaunchKernelGGL(k_emit<float>, dim3(eb), dim3(256), 0, 0, ab, s->emitted.as<float>(), s->emitdraw.as<uint32_t>());
        else hipLaunchKernelGGL(k_emit<double>, dim3(eb), dim3(256), 0, 0, ab, s->emitted.as<double>(), s->emitdraw.as<uint32_t>());
        SG_HIP(hipGetLastError());
        lt.start();
        dispatch_mc_any(s, ab, precision, blocks);
        lt.stop();
    }
    SG_HIP(hipEventRecord(e1, 0));
    SG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    unsigned long long c[SG_NCOUNTERS];
    SG_HIP(hipMemcpy(c, s->counters.p, sizeof c, hipMemcpyDeviceToHost));
    int err = 0;
    SG_HIP(hipMemcpy(&err, s->err.p, sizeof err, hipMemcpyDeviceToHost));
    if (counters) {
        for (int i = 0; i < SG_NCOUNTERS; i++) counters[i] = c[i];
        counters[6] = npp;
        counters[10] = (uint64_t)(ms * 1000.0);
        counters[11] = blocks;
        counters[12] = (uint64_t)(lt.total_ms() * 1000.0);
        counters[13] = lt.launches();
        counters[14] = (uint64_t)(lt.kernel_ms() * 1000.0);
    }
    SG_REQUIRE(!(err & 2), "walk order: a request's sorted place lies beyond the pass's requests");
    SG_REQUIRE(!(err & 4), "walk queue: a request names no instrument pixel or packet slot");
    SG_REQUIRE(err == 0, "The optical depth along the path is not a positive number");
    SG_CATCH
}

} // extern "C"

// ---------------------------------------------------------------- panchromatic dust emission

static CellBoxes cell_boxes(const sg_grid* g)
{
    CellBoxes cb{};
    cb.kind = g->kind;
    if (g->kind == GK_CART) {
        cb.xv = g->cart.xv; cb.yv = g->cart.yv; cb.zv = g->cart.zv;
        cb.Nx = g->cart.Nx; cb.Ny = g->cart.Ny; cb.Nz = g->cart.Nz;
    } else {
        cb.rec = g->tree.rec64;
    }
    return cb;
}

extern "C" {

int sg_dust_emission_spectra(sg_dust* d, const double* rho, const double* labs_stellar, const double* labs_dust,
                             int npop, const double* sigmaabs, double mu, const double* kappaabs, const double* lambdas,
                             const double* dlambdas, int nT, const double* Tv, const double* planckabs,
                             double length_unit, double* dustlum, double* labsbol)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(d && rho && labs_stellar && sigmaabs && kappaabs && lambdas && dlambdas && Tv && planckabs && dustlum &&
               labsbol, "null argument");
    SG_REQUIRE(d->ncomp == 1, "dust emission supports one dust component");
    SG_REQUIRE(d->grid->kind == GK_CART || d->grid->kind == GK_TREE,
               "dust emission needs cells that are boxes (Cartesian or tree grids)");
    SG_REQUIRE(npop >= 1 && npop <= 8 && nT >= 2 && mu > 0, "between 1 and 8 dust populations, a Planck table and a "
                                                            "positive mass per hydrogen atom are needed");
    const int64_t nc = d->grid->ncells;
    const int nl = d->nlambda;
    DevBuf b_rho, b_ls, b_ld, b_sig, b_ka, b_lam, b_dl, b_T, b_P, b_out, b_bol;
    b_rho.upload(rho, nc);
    b_ls.upload(labs_stellar, (size_t)nc * nl);
    if (labs_dust) b_ld.upload(labs_dust, (size_t)nc * nl);
    b_sig.upload(sigmaabs, (size_t)npop * nl);
    b_ka.upload(kappaabs, nl);
    b_lam.upload(lambdas, nl);
    b_dl.upload(dlambdas, nl);
    b_T.upload(Tv, nT + 1);
    b_P.upload(planckabs, (size_t)npop * (nT + 1));
    b_out.alloc((size_t)nc * nl * sizeof(double));
    b_bol.alloc((size_t)nc * sizeof(double));
    SpectraArgs a{};
    a.nl = nl; a.npop = npop; a.nT = nT;
    a.lambda = b_lam.as<double>(); a.dlambda = b_dl.as<double>();
    a.sigmaabs = b_sig.as<double>(); a.kappaabs = b_ka.as<double>(); a.mu = mu;
    a.Tv = b_T.as<double>(); a.planck = b_P.as<double>();
    a.labs_s = b_ls.as<double>(); a.labs_d = labs_dust ? b_ld.as<double>() : nullptr;
    a.rho = b_rho.as<double>(); a.len2 = length_unit * length_unit; a.ncells = nc;
    a.dustlum = b_out.as<double>(); a.labsbol = b_bol.as<double>();
    hipLaunchKernelGGL(k_dust_spectra, dim3(blocks_for(nc, 256)), dim3(256), 0, 0, a, cell_boxes(d->grid));
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpy(dustlum, b_out.p, (size_t)nc * nl * sizeof(double), hipMemcpyDeviceToHost));
    SG_HIP(hipMemcpy(labsbol, b_bol.p, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost));
    SG_CATCH
}

int sg_sim_run_dust(sg_sim* s, int ell, uint64_t first, uint64_t count, uint64_t npp, const double* Lv, double bias,
                    int peeloffs, int store_absorption, uint32_t stream, int precision, uint64_t* counters)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(s && s->dust && Lv, "dust emission needs a simulation with a dust system and the cells' luminosities");
    SG_REQUIRE(npp > 0, "npp must be positive");
    SG_REQUIRE(bias >= 0 && bias < 1, "the dust emission bias must be in [0, 1)");
    SG_REQUIRE(s->dust->ncomp == 1, "dust emission supports one dust component");
    const sg_grid* g = s->dust->grid;
    SG_REQUIRE(g->kind == GK_CART || g->kind == GK_TREE,
               "dust emission needs cells that are boxes (Cartesian or tree grids)");
    const int64_t nc = g->ncells;
    if (counters)
        for (int i = 0; i < SG_NCOUNTERS; i++) counters[i] = 0;
    // NR::cdf of the cells' luminosities, in the reference's order (PanMonteCarloSimulation.cpp:249-255,281-282)
    std::vector<double> cum(nc + 1, 0.0);
    for (int64_t m = 0; m < nc; m++) {
        SG_REQUIRE(Lv[m] >= 0 && std::isfinite(Lv[m]), "cell luminosities must be finite and non-negative");
        cum[m + 1] = cum[m] + Lv[m];
    }
    double Ltot = 0;
    for (int64_t m = 0; m < nc; m++) Ltot += Lv[m];
    if (!(Ltot > 0) || count == 0) return 0;
    const double norm = cum[nc];
    for (int64_t m = 0; m <= nc; m++) cum[m] /= norm;
    McArgs a;
    fill_args(s, ell, first, count, npp, a);
    a.member = a.member ^ (stream * 0x9E3779B9u);
    a.Lpkt = Ltot / (double)npp;
    a.Lthreshold = a.Lpkt / s->mc[1];
    a.store_abs = store_absorption && s->dust->store_abs;
    if (!a.store_abs) a.labs = nullptr;
    if (!peeloffs) a.ni = 0;
    if (g_tune_compact <= 0) a.compact = 32;
    DevBuf b_cum, b_lv;
    b_cum.upload(cum);
    b_lv.upload(Lv, nc);
    DustLaunch dl{};
    dl.cum = b_cum.as<double>(); dl.Lv = b_lv.as<double>(); dl.ncells = nc;
    dl.xi = bias; dl.Lmean = Ltot / nc; dl.Lem = Ltot / (double)npp;
    dl.cb = cell_boxes(g);
    s->counters.zero();
    s->err.zero();
    const bool mixed = precision == SG_MIXED;
    const size_t rsz = mixed ? sizeof(float) : sizeof(double);
    const uint64_t batch = std::min<uint64_t>(count, g_emit_batch ? g_emit_batch : (uint64_t)1 << 22);
    if (s->emitted.bytes < 7 * batch * rsz) s->emitted.alloc(7 * batch * rsz);
    if (s->emitdraw.bytes < batch * sizeof(uint32_t)) s->emitdraw.alloc(batch * sizeof(uint32_t));
    int blocks = 0;
    for (uint64_t b = 0; b < count; b += batch) {
        McArgs ab = a;
        ab.first = first + b;
        ab.count = std::min(batch, count - b);
        ab.emitted = s->emitted.p;
        ab.emit_draw = s->emitdraw.as<uint32_t>();
        s->next.zero();
        int eb = (int)std::min<uint64_t>((ab.count + 255) / 256, (uint64_t)g_cus * 16);
        if (mixed) hipLaunchKernelGGL(k_emit_dust<float>, dim3(eb), dim3(256), 0, 0, ab, dl, s->emitted.as<float>(), s->emitdraw.as<uint32_t>());
        else hipLaunchKernelGGL(k_emit_dust<double>, dim3(eb), dim3(256), 0, 0, ab, dl, s->emitted.as<double>(), s->emitdraw.as<uint32_t>());
        SG_HIP(hipGetLastError());
        dispatch_mc_any(s, ab, precision, blocks);
    }
    SG_HIP(hipDeviceSynchronize());
    unsigned long long c[SG_NCOUNTERS];
    SG_HIP(hipMemcpy(c, s->counters.p, sizeof c, hipMemcpyDeviceToHost));
    int err = 0;
    SG_HIP(hipMemcpy(&err, s->err.p, sizeof err, hipMemcpyDeviceToHost));
    if (counters) {
        for (int i = 0; i < SG_NCOUNTERS; i++) counters[i] = c[i];
        counters[6] = npp;
        counters[11] = blocks;
    }
    SG_REQUIRE(err == 0, "The optical depth along the path is not a positive number");
    SG_CATCH
}

} // extern "C"

// ---------------------------------------------------------------- transport without device allocation
//
// sg_sim_run_static runs the event/walk pipeline on one slot group and one stream with all of its work
// memory (packet slots, walk queue, progress counters, the launch states of its packets) in `scratch`,
// a file-scope device array of kStaticScratchBytes that every device the library runs on holds until
// the process ends. Host code never uses the array's name as an address: sg_init takes this device's
// address of it with hipGetSymbolAddress and keeps it in g_scratch. Counters and flags are cleared by
// a fill kernel, not by hipMemsetAsync. Calls that overlap on one device race, since they share it.
constexpr size_t kStaticScratchBytes = (size_t)64 << 20;
constexpr uint64_t kStaticMaxPackets = (uint64_t)1 << 19;
constexpr int64_t kStaticSlots = (int64_t)1 << 15;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
static void* g_scratch[64] = {};

static void static_scratch_setup(int device)
{
    SG_REQUIRE(device < 64, "device index beyond the static work array table");
    if (g_scratch[device]) return;
    void* p = nullptr;
    SG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SG_REQUIRE(p != nullptr, "no device address for the static work array");
    g_scratch[device] = p;
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
static void fill_u32(void* p, int64_t n, hipStream_t st)
{
    const int64_t b = std::min<int64_t>((n + 255) / 256, (int64_t)g_cus * 4);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::max<int64_t>(b, 1)), dim3(256), 0, st, static_cast<uint32_t*>(p), n,
                       0u);
    SG_HIP(hipGetLastError());
}

static size_t align256(size_t n) { return (n + 255) / 256 * 256; }

template <class W, bool ABS>
static void run_static(sg_sim* s, const typename W::View& v, const McArgs& a0, int nlds, int& blocks)
{
    using R = typename W::Real;
    hipStream_t st = g_split_stream[0];
    const int64_t P = (int64_t)std::max<uint64_t>(1, std::min<uint64_t>(a0.count, (uint64_t)kStaticSlots));
    const int64_t cap = P * (a0.ni + 1);
    // layout of the static array: slots, walk requests, progress counters, launch states, draw counts
    const size_t o_q = align256(slot_bytes<R>(P));
    const size_t o_pc = o_q + align256((size_t)cap * sizeof(WalkReq<R>));
    const size_t o_em = o_pc + 256;
    const size_t o_ed = o_em + align256(7 * a0.count * sizeof(R));
    const size_t need = o_ed + align256(a0.count * sizeof(uint32_t));
    SG_REQUIRE(need <= kStaticScratchBytes, "sg_sim_run_static: the run needs " + std::to_string(need) +
                                                " bytes of work memory, more than the static array holds");
    unsigned char* base = static_cast<unsigned char*>(g_scratch[g_device]);
    SG_REQUIRE(base != nullptr, "sg_sim_run_static: the static work array of this device is not set up");
    Slots<R> sl{reinterpret_cast<R*>(base), reinterpret_cast<uint32_t*>(base + (size_t)P * SR_NR * sizeof(R)), P};
    unsigned long long* pc = reinterpret_cast<unsigned long long*>(base + o_pc);   // [1] queue tail, [3] claims
    WalkQueue<R> q{};
    q.req = reinterpret_cast<WalkReq<R>*>(base + o_q);
    q.cap = cap;
    q.tail = pc + 1;
    McArgs a = a0;
    a.emitted = base + o_em;
    a.emit_draw = reinterpret_cast<uint32_t*>(base + o_ed);
    a.next_packet = s->next.as<unsigned long long>();
    // cleared by kernels: slot states (PS_FRESH) and counts, progress, the packet claims, counters, error flag
    fill_u32(sl.u + (size_t)SU_STATE * P, P * (SU_NU - SU_STATE), st);
    fill_u32(pc, 8, st);
    fill_u32(s->next.p, 2, st);
    fill_u32(s->counters.p, 2 * SG_NCOUNTERS, st);
    fill_u32(s->err.p, 1, st);
    const int bs = g_tune_block;
    const size_t shm = (size_t)nlds * sizeof(int32_t);
    int per = 0;
    SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_walks<W, ABS>, bs, shm));
    blocks = std::max(1, per) * g_cus;
    int ebs = 256;
    while (ebs > 64 && staged_bytes<R>(ebs, a.ni) > 64 * 1024) ebs /= 2;
    const size_t eshm = staged_bytes<R>(ebs, a.ni);
    if (a.count > 0) {
        const int emb = (int)std::min<uint64_t>((a.count + 255) / 256, (uint64_t)g_cus * 16);
        hipLaunchKernelGGL(k_emit<R>, dim3(emb), dim3(256), 0, st, a, static_cast<R*>(const_cast<void*>(a.emitted)),
                           const_cast<uint32_t*>(a.emit_draw));
        SG_HIP(hipGetLastError());
    }
    for (;;) {
        for (int p = 0; p < kSplitCheck; p++) {
            fill_u32(pc + 1, 2, st);
            hipLaunchKernelGGL(k_events<R>, dim3((P + ebs - 1) / ebs), dim3(ebs), eshm, st, a, sl, q);
            SG_HIP(hipGetLastError());
            fill_u32(pc + 3, 2, st);
            hipLaunchKernelGGL((k_walks<W, ABS>), dim3(blocks), dim3(bs), shm, st, v, a, q, sl, pc + 3, nlds, 0);
            SG_HIP(hipGetLastError());
        }
        SG_HIP(hipMemcpyAsync(g_split_host, a.next_packet, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        SG_HIP(hipMemcpyAsync(g_split_host + 1, pc + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        SG_HIP(hipStreamSynchronize(st));
        SG_REQUIRE(g_split_host[1] <= (unsigned long long)cap, "walk queue overflow");
        if (g_split_host[0] >= a.count && g_split_host[1] == 0) break;
    }
    hipLaunchKernelGGL(k_slot_counts, dim3(std::min<int64_t>((P + 255) / 256, 1024)), dim3(256), 0, st, sl.u, P, a.counters);
    SG_HIP(hipGetLastError());
    SG_HIP(hipStreamSynchronize(st));
}

template <bool ABS>
static void dispatch_static(sg_sim* s, const McArgs& a, int precision, int& blocks)
{
    const bool mixed = precision == SG_MIXED;
    sg_grid* g = s->dust ? s->dust->grid : nullptr;
    if (!g) {
        CartView v{};
        if (mixed) run_static<CartW32, false>(s, v, a, 0, blocks);
        else run_static<CartW64, false>(s, v, a, 0, blocks);
    } else if (g->kind == GK_CART) {
        if (mixed) run_static<CartW32, ABS>(s, g->cart, a, 0, blocks);
        else run_static<CartW64, ABS>(s, g->cart, a, 0, blocks);
    } else if (g->kind == GK_TREE) {
        if (!mixed && g->tree.search == SG_BOOKKEEPING) run_static<TreeW64Book, ABS>(s, g->tree, a, 0, blocks);
        else if (!mixed) run_static<TreeW64, ABS>(s, g->tree, a, SG_LDS_NODES, blocks);
        else if (rope_walk(g)) run_static<TreeW32, ABS>(s, g->tree, a, SG_LDS_NODES, blocks);
        else run_static<TreeW32TopDown, ABS>(s, g->tree, a, SG_LDS_NODES, blocks);
    } else {
        if (mixed) run_static<VoroW<float>, ABS>(s, g->voro, a, 0, blocks);
        else run_static<VoroW<double>, ABS>(s, g->voro, a, 0, blocks);
    }
}

extern "C" {

int sg_sim_run_static(sg_sim* s, int ell, uint64_t first, uint64_t count, uint64_t npp, int precision,
                      uint64_t* counters)
{
    SG_TRY
    SG_REQUIRE(g_device >= 0, "sg_sim_run_static: call sg_init first (the static work array is set up there)");
    SG_REQUIRE(npp > 0, "npp must be positive");
    SG_REQUIRE(precision == SG_MIXED || precision == SG_FP64, "precision must be SG_FP64 or SG_MIXED");
    SG_REQUIRE(count <= kStaticMaxPackets, "sg_sim_run_static: at most " + std::to_string(kStaticMaxPackets) +
                                               " packets per call (the static work array holds their launch states); "
                                               "split the packet range over several calls");
    SG_REQUIRE(!s->dust || s->dust->ncomp == 1,
               "sg_sim_run_static: a dust system of several components needs sg_sim_run (lane pipeline)");
    McArgs a;
    fill_args(s, ell, first, count, npp, a);
    int blocks = 0;
    if (a.store_abs) dispatch_static<true>(s, a, precision, blocks);
    else dispatch_static<false>(s, a, precision, blocks);
    unsigned long long c[SG_NCOUNTERS];
    SG_HIP(hipMemcpy(c, s->counters.p, sizeof c, hipMemcpyDeviceToHost));
    int err = 0;
    SG_HIP(hipMemcpy(&err, s->err.p, sizeof err, hipMemcpyDeviceToHost));
    if (counters) {
        for (int i = 0; i < SG_NCOUNTERS; i++) counters[i] = c[i];
        counters[6] = npp;
        counters[11] = blocks;
    }
    SG_REQUIRE(!(err & 4), "walk queue: a request names no instrument pixel or packet slot");
    SG_REQUIRE(err == 0, "The optical depth along the path is not a positive number");
    SG_CATCH
}

int sg_set_slots(int64_t slots)
{
    SG_TRY
    SG_REQUIRE(slots == 0 || (slots >= 1024 && slots <= ((int64_t)1 << 24)), "slots must be 0 or 1024..2^24");
    g_tune_slots = slots;
    SG_CATCH
}

// Coordinate search over the launch parameters of the current pipeline, each candidate timed on the
// same sample of packets (the histories do not depend on these parameters, only the speed does):
// the compaction threshold, then the in-flight packet slots (event/walk pipeline) or the waves per
// CU (lane pipeline), then the block size. Every candidate runs twice and the faster run counts, so
// that work-buffer growth stays out of the comparison. The sample's tallies are taken back out: the
// frames and the absorption table hold what they held before.
int sg_autotune(sg_sim* s, int ell, uint64_t first, uint64_t sample, int precision, int choose_pipeline,
                int64_t* best, double* rate)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(s != nullptr && sample > 0, "autotune needs a simulation and a positive sample");
    if (s->dust && s->dust->ncomp > 1) {
        // several dust components always run in the lane pipeline: tune that
        choose_pipeline = 0;
        g_pipeline = 0;
    }
    std::vector<std::pair<DevBuf*, void*>> keep;
    auto save = [&](void* p, size_t bytes) {
        auto* c = new DevBuf();
        c->alloc(bytes);
        SG_HIP(hipMemcpy(c->p, p, bytes, hipMemcpyDeviceToDevice));
        keep.push_back({c, p});
    };
    for (DevBuf* f : s->instr->frames) save(f->p, f->bytes);
    if (s->dust && s->dust->store_abs) save(s->dust->labs.p, s->dust->labs.bytes);
    auto trial = [&]() {
        double r = 0;
        for (int rep = 0; rep < 2; rep++) {
            uint64_t c[32] = {0};
            if (sg_sim_run(s, ell, first, sample, sample, precision, c) != 0) throw std::runtime_error(sg_last_error());
            r = std::max(r, (double)c[0] / std::max(1e-9, (double)c[10] * 1e-6));
        }
        return r;
    };
    const int pipeline0 = g_pipeline;
    const int order0 = g_walk_order;
    const int block0 = g_tune_block, compact0 = g_tune_compact, waves0 = g_tune_waves;
    const int64_t slots0 = g_tune_slots;
    double top = 0;
    if (choose_pipeline) {
        // the transport organisation first, each with its default launch parameters
        try {
            double r[2] = {0, 0};
            for (int p = 0; p < 2; p++) {
                g_pipeline = p;
                g_tune_block = 256; g_tune_compact = 0; g_tune_waves = 0; g_tune_slots = 0;
                r[p] = trial();
            }
            g_pipeline = r[1] >= r[0] ? 1 : 0;
        } catch (...) {
            g_pipeline = pipeline0; g_tune_block = block0; g_tune_compact = compact0; g_tune_waves = waves0;
            g_tune_slots = slots0; g_walk_order = order0;
            for (auto& k : keep) {
                (void)hipMemcpy(k.second, k.first->p, k.first->bytes, hipMemcpyDeviceToDevice);
                delete k.first;
            }
            throw;
        }
    }
    int block = g_tune_block;
    int compact = g_tune_compact > 0 ? g_tune_compact : (g_pipeline == 1 ? 24 : 32);
    int64_t slots = g_tune_slots > 0 ? g_tune_slots : kSplitSlots;
    int waves = g_tune_waves;
    int order = g_walk_order;
    g_tune_compact = compact;
    try {
        top = trial();
        for (int c : {8, 16, 24, 32, 40, 48}) {
            if (c == compact) continue;
            g_tune_compact = c;
            const double r = trial();
            if (r > top) { top = r; compact = c; }
        }
        g_tune_compact = compact;
        if (g_pipeline == 1) {
            for (int64_t p : {(int64_t)1 << 20, (int64_t)1 << 21, (int64_t)1 << 22, (int64_t)1 << 23, (int64_t)1 << 24}) {
                if (p == slots || (uint64_t)p > sample) continue;   // beyond the sample: all slots alike
                g_tune_slots = p;
                const double r = trial();
                if (r > top) { top = r; slots = p; }
            }
            g_tune_slots = slots;
            // the walk order: as appended, sorted by coherence key (which pays where the walks' grid data
            // does not fit the caches), or sorted and claimed by XCD segment
            for (int o : {0, 1, 2}) {
                if (o == order) continue;
                g_walk_order = o;
                const double r = trial();
                if (r > top) { top = r; order = o; }
            }
            g_walk_order = order;
        } else {
            for (int w : {4, 8, 12, 16}) {
                if (w == waves) continue;
                g_tune_waves = w;
                const double r = trial();
                if (r > top) { top = r; waves = w; }
            }
            g_tune_waves = waves;
        }
        for (int b : {64, 128, 256}) {
            if (b == block) continue;
            g_tune_block = b;
            const double r = trial();
            if (r > top) { top = r; block = b; }
        }
        g_tune_block = block;
    } catch (...) {
        g_tune_block = block; g_tune_compact = compact; g_tune_slots = slots; g_tune_waves = waves;
        g_walk_order = order;
        for (auto& k : keep) {
            (void)hipMemcpy(k.second, k.first->p, k.first->bytes, hipMemcpyDeviceToDevice);
            delete k.first;
        }
        throw;
    }
    for (auto& k : keep) {
        SG_HIP(hipMemcpy(k.second, k.first->p, k.first->bytes, hipMemcpyDeviceToDevice));
        delete k.first;
    }
    if (best) {
        best[0] = block; best[1] = g_pipeline == 1 ? slots : waves; best[2] = compact; best[3] = g_pipeline;
        best[4] = g_pipeline == 1 ? order : 0;
    }
    if (rate) *rate = top;
    SG_CATCH
}

} // extern "C"

// One round of a batched run: emission for every range, then the batched transport kernel. The
// tables (run arguments, range offsets, grid views) go to the first member's work buffer.
template <class W, bool ABS>
static void launch_batch(sg_sim* s0, const std::vector<McArgs>& rg, const std::vector<sg_sim*>& owner, int& blocks,
                         LaunchTimer& lt)
{
    using R = typename W::Real;
    using V = typename W::View;
    const size_t nr = rg.size();
    std::vector<V> views(nr);
    for (size_t k = 0; k < nr; k++) {
        if constexpr (std::is_same<V, CartView>::value) views[k] = owner[k]->dust ? owner[k]->dust->grid->cart : CartView{};
        else if constexpr (std::is_same<V, TreeView>::value) views[k] = owner[k]->dust->grid->tree;
        else views[k] = owner[k]->dust->grid->voro;
    }
    std::vector<uint64_t> roff(nr + 1, 0);
    for (size_t k = 0; k < nr; k++) roff[k + 1] = roff[k] + rg[k].count;
    const size_t o_off = (nr * sizeof(McArgs) + 255) / 256 * 256;
    const size_t o_view = o_off + ((nr + 1) * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t need = o_view + nr * sizeof(V);
    if (s0->btab.bytes < need) s0->btab.alloc(need);
    unsigned char* base = s0->btab.as<unsigned char>();
    SG_HIP(hipMemcpy(base, rg.data(), nr * sizeof(McArgs), hipMemcpyHostToDevice));
    SG_HIP(hipMemcpy(base + o_off, roff.data(), (nr + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
    SG_HIP(hipMemcpy(base + o_view, views.data(), nr * sizeof(V), hipMemcpyHostToDevice));
    const auto* atab = reinterpret_cast<const McArgs*>(base);
    const auto* otab = reinterpret_cast<const uint64_t*>(base + o_off);
    const auto* vtab = reinterpret_cast<const V*>(base + o_view);

    int eb = (int)std::min<uint64_t>((roff[nr] + 255) / 256, (uint64_t)g_cus * 16);
    hipLaunchKernelGGL(k_emit_batch<R>, dim3(eb), dim3(256), 0, 0, atab, otab, (int)nr);
    SG_HIP(hipGetLastError());

    int bs = g_tune_block;
    size_t shm = SG_PARK ? hot_words<W>() * 4 * (size_t)bs : 0;
    int per = 0;
    SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_mc_batch<W, ABS>, bs, shm));
    if (per < 1) per = 1;
    if (g_tune_waves > 0) per = std::max(1, std::min(per, g_tune_waves * 64 / bs));
    blocks = per * g_cus;
    int nslots = blocks * bs;
    size_t cneed = cold_bytes<R>(nslots);
    if (s0->cold.bytes < cneed) s0->cold.alloc(cneed);
    R* cr = s0->cold.as<R>();
    uint32_t* cu = reinterpret_cast<uint32_t*>(s0->cold.as<unsigned char>() + (size_t)nslots * CR_NR * sizeof(R));
    lt.start();
    hipLaunchKernelGGL((k_mc_batch<W, ABS>), dim3(blocks), dim3(bs), shm, 0, vtab, atab, (int)nr, cr, cu, nslots);
    SG_HIP(hipGetLastError());
    lt.stop();
}

template <bool ABS>
static void dispatch_batch(sg_sim* s0, const std::vector<McArgs>& rg, const std::vector<sg_sim*>& owner,
                           int precision, int& blocks, LaunchTimer& lt)
{
    const bool mixed = precision == SG_MIXED;
    sg_grid* g = s0->dust ? s0->dust->grid : nullptr;
    if (!g) { launch_batch<CartW64, false>(s0, rg, owner, blocks, lt); return; }
    if (g->kind == GK_CART) {
        if (mixed) launch_batch<CartW32, ABS>(s0, rg, owner, blocks, lt);
        else launch_batch<CartW64, ABS>(s0, rg, owner, blocks, lt);
    } else if (g->kind == GK_TREE) {
        if (!mixed && g->tree.search == SG_BOOKKEEPING) launch_batch<TreeW64Book, ABS>(s0, rg, owner, blocks, lt);
        else if (!mixed) launch_batch<TreeW64, ABS>(s0, rg, owner, blocks, lt);
        else if (rope_walk(g)) launch_batch<TreeW32, ABS>(s0, rg, owner, blocks, lt);
        else launch_batch<TreeW32TopDown, ABS>(s0, rg, owner, blocks, lt);
    } else {
        if (mixed) launch_batch<VoroW<float>, ABS>(s0, rg, owner, blocks, lt);
        else launch_batch<VoroW<double>, ABS>(s0, rg, owner, blocks, lt);
    }
}

extern "C" {

int sg_sim_run_batch(int nmembers, sg_sim* const* sims, const int* ells, const uint64_t* counts, int precision,
                     uint64_t* counters)
{
    SG_TRY
    ensure_init();
    SG_REQUIRE(nmembers >= 1 && sims && ells && counts, "a batched run needs at least one member");
    sg_sim* s0 = sims[0];
    auto kind_of = [](const sg_sim* s) { return s->dust ? (int)s->dust->grid->kind : 0; };
    auto search_of = [](const sg_sim* s) { return s->dust && s->dust->grid->kind == GK_TREE ? s->dust->grid->tree.search : -1; };
    auto abs_of = [](const sg_sim* s) { return s->dust ? s->dust->store_abs : 0; };
    std::vector<const sg_sim*> seen;
    for (int i = 0; i < nmembers; i++) {
        SG_REQUIRE(sims[i] != nullptr, "null member simulation");
        SG_REQUIRE(kind_of(sims[i]) == kind_of(s0) && search_of(sims[i]) == search_of(s0),
                   "the members of a batched run must have dust grids of one kind (and one tree search method)");
        SG_REQUIRE(abs_of(sims[i]) == abs_of(s0), "the members of a batched run must agree on storing absorption");
        SG_REQUIRE(std::find(seen.begin(), seen.end(), sims[i]) == seen.end(),
                   "a simulation may appear only once in a batched run");
        seen.push_back(sims[i]);
    }
    std::vector<McArgs> base(nmembers);
    uint64_t total = 0;
    for (int i = 0; i < nmembers; i++) {
        fill_args(sims[i], ells[i], 0, counts[i], std::max<uint64_t>(counts[i], 1), base[i]);
        sims[i]->counters.zero();
        sims[i]->err.zero();
        total += counts[i];
    }
    const bool mixed = precision == SG_MIXED;
    const size_t rsz = mixed ? sizeof(float) : sizeof(double);
    // packets per round: the emission buffer holds one round's launch states
    const uint64_t B = std::max<uint64_t>(1, std::min<uint64_t>(total, (uint64_t)1 << 23));
    if (s0->emitted.bytes < 7 * B * rsz) s0->emitted.alloc(7 * B * rsz);
    if (s0->emitdraw.bytes < B * sizeof(uint32_t)) s0->emitdraw.alloc(B * sizeof(uint32_t));
    if (s0->bnext.bytes < (size_t)nmembers * sizeof(unsigned long long))
        s0->bnext.alloc((size_t)nmembers * sizeof(unsigned long long));
    hipEvent_t e0, e1;
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    SG_HIP(hipEventRecord(e0, 0));
    int blocks = 0;
    LaunchTimer lt;
    std::vector<uint64_t> done(nmembers, 0);
    // members with several dust components: one at a time through the lane pipeline
    for (int i = 0; i < nmembers; i++) {
        if (base[i].ncomp < 2 || counts[i] == 0) continue;
        for (uint64_t b = 0; b < counts[i]; b += B) {
            McArgs ab = base[i];
            ab.first = b;
            ab.count = std::min(B, counts[i] - b);
            ab.emitted = s0->emitted.p;
            ab.emit_draw = s0->emitdraw.as<uint32_t>();
            if (g_tune_compact <= 0) ab.compact = 32;
            sims[i]->next.zero();
            int eb = (int)std::min<uint64_t>((ab.count + 255) / 256, (uint64_t)g_cus * 16);
            if (mixed) hipLaunchKernelGGL(k_emit<float>, dim3(eb), dim3(256), 0, 0, ab, s0->emitted.as<float>(), s0->emitdraw.as<uint32_t>());
            else hipLaunchKernelGGL(k_emit<double>, dim3(eb), dim3(256), 0, 0, ab, s0->emitted.as<double>(), s0->emitdraw.as<uint32_t>());
            SG_HIP(hipGetLastError());
            lt.start();
            dispatch_mc_any(sims[i], ab, precision, blocks);
            lt.stop();
        }
        done[i] = counts[i];
    }
    // large members: their own event/walk passes, which keep the walk kernel full on their own
    for (int i = 0; g_pipeline == 1 && g_batch_alone > 0 && i < nmembers; i++) {
        if (done[i] == counts[i]) continue;
        if (counts[i] < g_batch_alone) continue;
        if (base[i].store_abs) dispatch_split<true>(sims[i], base[i], precision, 0, counts[i], lt, blocks);
        else dispatch_split<false>(sims[i], base[i], precision, 0, counts[i], lt, blocks);
        done[i] = counts[i];
    }
    int j = 0;
    for (;;) {
        // the next round: the members' remaining packets in order, split where the round is full
        std::vector<McArgs> rg;
        std::vector<sg_sim*> owner;
        uint64_t used = 0;
        while (j < nmembers && used < B) {
            uint64_t left = counts[j] - done[j];
            if (left == 0) { j++; continue; }
            uint64_t take = std::min(left, B - used);
            McArgs a = base[j];
            a.first = done[j];
            a.count = take;
            a.emitted = s0->emitted.as<unsigned char>() + 7 * used * rsz;
            a.emit_draw = s0->emitdraw.as<uint32_t>() + used;
            a.next_packet = s0->bnext.as<unsigned long long>() + rg.size();
            rg.push_back(a);
            owner.push_back(sims[j]);
            used += take;
            done[j] += take;
            if (done[j] == counts[j]) j++;
        }
        if (rg.empty()) break;
        SG_HIP(hipMemset(s0->bnext.p, 0, rg.size() * sizeof(unsigned long long)));
        if (abs_of(s0)) dispatch_batch<true>(s0, rg, owner, precision, blocks, lt);
        else dispatch_batch<false>(s0, rg, owner, precision, blocks, lt);
    }
    SG_HIP(hipEventRecord(e1, 0));
    SG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    unsigned long long sum[SG_NCOUNTERS] = {0};
    int err = 0;
    for (int i = 0; i < nmembers; i++) {
        unsigned long long c[SG_NCOUNTERS];
        SG_HIP(hipMemcpy(c, sims[i]->counters.p, sizeof c, hipMemcpyDeviceToHost));
        for (int k = 0; k < SG_NCOUNTERS; k++) sum[k] += c[k];
        int e = 0;
        SG_HIP(hipMemcpy(&e, sims[i]->err.p, sizeof e, hipMemcpyDeviceToHost));
        err |= e;
    }
    if (counters) {
        for (int k = 0; k < SG_NCOUNTERS; k++) counters[k] = sum[k];
        counters[6] = total;
        counters[10] = (uint64_t)(ms * 1000.0);
        counters[11] = blocks;
        counters[12] = (uint64_t)(lt.total_ms() * 1000.0);
        counters[13] = lt.launches();
        counters[14] = (uint64_t)(lt.kernel_ms() * 1000.0);
    }
    SG_REQUIRE(!(err & 2), "walk order: a request's sorted place lies beyond the pass's requests");
    SG_REQUIRE(!(err & 4), "walk queue: a request names no instrument pixel or packet slot");
    SG_REQUIRE(err == 0, "The optical depth along the path is not a positive number");
    SG_CATCH
}

} // extern "C"
