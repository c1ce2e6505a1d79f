// sg_walk.hpp -- DustGrid::path as per-lane segment generators.
//
// Every walker yields the (m, ds) segments of DustGridPath one at a time, including the m = -1
// segments that move an outside start point onto the grid, and drops ds <= 0 segments like
// DustGridPath::addSegment (DustGridPath.cpp:46-53). No segment list is stored: callers fuse
// the optical-depth sums, absorption and the tau -> s inversion into the stream.
//
// The FP64 walkers repeat the reference's arithmetic operation for operation (compile with
// -ffp-contract=off) so the cell sequences are bit-exact; the FP32 walkers are the mixed-precision
// variants: FP32 positions, reciprocal directions, snapped exit coordinates, and an FP64
// recompute of the wall distances when the smallest two are within 1e-5 of each other or the
// smallest is within 1e-5 of zero (relative to the second).
#pragma once
#include <type_traits>

#include "sg_common.hpp"

namespace sg {

// Field-by-field parking of lane state in LDS (structure-of-arrays, one slot per lane). LOAD
// selects the direction. Walkers list their fields in park(); no address of the state is taken,
// so it stays in registers outside the parking points.
template <bool LOAD> struct Park {
    uint32_t* base;
    int nl, lane, i;
    __device__ __forceinline__ void w(uint32_t& v)
    {
        if (LOAD) v = base[i * nl + lane];
        else base[i * nl + lane] = v;
        i++;
    }
    __device__ __forceinline__ void io(uint32_t& v) { w(v); }
    __device__ __forceinline__ void io(int& v)
    {
        uint32_t t = (uint32_t)v;
        w(t);
        v = (int)t;
    }
    __device__ __forceinline__ void io(bool& v)
    {
        uint32_t t = v ? 1u : 0u;
        w(t);
        v = t != 0;
    }
    __device__ __forceinline__ void io(float& v)
    {
        uint32_t t = __float_as_uint(v);
        w(t);
        v = __uint_as_float(t);
    }
    __device__ __forceinline__ void io(double& v)
    {
        unsigned long long b = (unsigned long long)__double_as_longlong(v);
        uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
        w(lo);
        w(hi);
        v = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
};

// Stages the first breadth-first tree node words (the top levels) in LDS.
__device__ __forceinline__ void stage_tree_nodes(const TreeView& t, int32_t* lds, int& n)
{
    n = min((int)t.nnodes, SG_LDS_NODES);
    for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = t.node[i];
}
__device__ __forceinline__ void stage_tree_nodes(const CartView&, int32_t*, int& n) { n = 0; }
__device__ __forceinline__ void stage_tree_nodes(const VoroView&, int32_t*, int& n) { n = 0; }

// ====================================================================== Cartesian

// CartesianDustGrid.cpp:136-283, exact FP64.
struct CartW64 {
    using Real = double;
    using View = CartView;
    double x, y, z, kx, ky, kz;
    int i, j, k;
    int ne;
    double e0, e1, e2;
    bool done;
    uint32_t nrecompute, nstuck;
    static constexpr int kParkWords = 25;
    __device__ __forceinline__ static int cell_of(const View&, int n) { return n; }
    static constexpr bool kHint = false;
    __device__ __forceinline__ int entries() const { return ne; }
    static constexpr bool kStep = false;   // next() only
    template <class P> __device__ __forceinline__ void park(P& p)
    {
        p.io(x); p.io(y); p.io(z); p.io(kx); p.io(ky); p.io(kz);
        p.io(i); p.io(j); p.io(k); p.io(ne); p.io(e0); p.io(e1); p.io(e2);
        p.io(done); p.io(nrecompute); p.io(nstuck);
    }

    __device__ __forceinline__ void push(double ds)
    {
        if (ds > 0) {
            // value selects, not an indexed store: keeps the walker out of scratch memory
            e0 = ne == 0 ? ds : e0;
            e1 = ne == 1 ? ds : e1;
            e2 = ne >= 2 ? ds : e2;
            ne++;
        }
    }
    __device__ __forceinline__ void init(const View& g, const int32_t*, int, double rx, double ry, double rz, double dx, double dy,
                         double dz)
    {
        x = rx; y = ry; z = rz; kx = dx; ky = dy; kz = dz;
        ne = 0; done = false;
        double ds;
        if (x < g.xmin) {
            if (kx <= 0.0) { ne = 0; done = true; return; }
            ds = (g.xmin - x) / kx; push(ds);
            x = g.xmin + 1e-8 * (g.xv[1] - g.xv[0]); y += ky * ds; z += kz * ds;
        } else if (x > g.xmax) {
            if (kx >= 0.0) { ne = 0; done = true; return; }
            ds = (g.xmax - x) / kx; push(ds);
            x = g.xmax - 1e-8 * (g.xv[g.Nx] - g.xv[g.Nx - 1]); y += ky * ds; z += kz * ds;
        }
        if (y < g.ymin) {
            if (ky <= 0.0) { ne = 0; done = true; return; }
            ds = (g.ymin - y) / ky; push(ds);
            x += kx * ds; y = g.ymin + 1e-8 * (g.yv[1] - g.yv[0]); z += kz * ds;
        } else if (y > g.ymax) {
            if (ky >= 0.0) { ne = 0; done = true; return; }
            ds = (g.ymax - y) / ky; push(ds);
            x += kx * ds; y = g.ymax - 1e-8 * (g.yv[g.Ny] - g.yv[g.Ny - 1]); z += kz * ds;
        }
        if (z < g.zmin) {
            if (kz <= 0.0) { ne = 0; done = true; return; }
            ds = (g.zmin - z) / kz; push(ds);
            x += kx * ds; y += ky * ds; z = g.zmin + 1e-8 * (g.zv[1] - g.zv[0]);
        } else if (z > g.zmax) {
            if (kz >= 0.0) { ne = 0; done = true; return; }
            ds = (g.zmax - z) / kz; push(ds);
            x += kx * ds; y += ky * ds; z = g.zmax - 1e-8 * (g.zv[g.Nz] - g.zv[g.Nz - 1]);
        }
        if (x < g.xmin || x > g.xmax || y < g.ymin || y > g.ymax || z < g.zmin || z > g.zmax) {
            ne = 0; done = true; return;
        }
        i = locate_clip(g.xv, g.Nx + 1, x);
        j = locate_clip(g.yv, g.Ny + 1, y);
        k = locate_clip(g.zv, g.Nz + 1, z);
    }
    __device__ __forceinline__ bool next(const View& g, const int32_t*, int, int& m, double& ds, double& rho)
    {
        if (ne) {
            ds = e0; e0 = e1; e1 = e2; ne--; m = -1; rho = 0.0;
            return true;
        }
        while (!done) {
            int mm = k + g.Nz * j + g.Nz * g.Ny * i;
            double xE = (kx < 0.0) ? g.xv[i] : g.xv[i + 1];
            double yE = (ky < 0.0) ? g.yv[j] : g.yv[j + 1];
            double zE = (kz < 0.0) ? g.zv[k] : g.zv[k + 1];
            double dsx = (fabs(kx) > 1e-15) ? (xE - x) / kx : DBL_MAX;
            double dsy = (fabs(ky) > 1e-15) ? (yE - y) / ky : DBL_MAX;
            double dsz = (fabs(kz) > 1e-15) ? (zE - z) / kz : DBL_MAX;
            double d;
            if (dsx <= dsy && dsx <= dsz) {
                d = dsx; i += (kx < 0.0) ? -1 : 1;
                if (i >= g.Nx || i < 0) done = true;
                else { x = xE; y += ky * d; z += kz * d; }
            } else if (dsy < dsx && dsy <= dsz) {
                d = dsy; j += (ky < 0.0) ? -1 : 1;
                if (j >= g.Ny || j < 0) done = true;
                else { x += kx * d; y = yE; z += kz * d; }
            } else if (dsz < dsx && dsz < dsy) {
                d = dsz; k += (kz < 0.0) ? -1 : 1;
                if (k >= g.Nz || k < 0) done = true;
                else { x += kx * d; y += ky * d; z = zE; }
            } else {
                // only reachable with NaN input; the reference would not terminate
                done = true;
                return false;
            }
            if (d > 0) { m = mm; ds = d; rho = g.rho[mm]; return true; }
        }
        return false;
    }
};

// Chooses the exit wall among three distances with the reference tie rules
// (CartesianDustGrid.cpp:243-269: x on <=, y on < x and <= z, z strictly smallest).
__device__ __forceinline__ int choose_axis_cart(double dx, double dy, double dz)
{
    if (dx <= dy && dx <= dz) return 0;
    if (dy < dx && dy <= dz) return 1;
    return 2;
}
__device__ __forceinline__ int choose_axis_cart(float dx, float dy, float dz)
{
    if (dx <= dy && dx <= dz) return 0;
    if (dy < dx && dy <= dz) return 1;
    return 2;
}
// Tree tie rules (TreeDustGrid.cpp:451-466: x, then y, then z, all on <=).
template <typename T> __device__ __forceinline__ int choose_axis_tree(T dx, T dy, T dz)
{
    if (dx <= dy && dx <= dz) return 0;
    if (dy <= dx && dy <= dz) return 1;
    return 2;
}

// One of three per-axis values by flags (fy: y, fz: z, neither: x). The values pass through an empty
// register constraint first: a select between two loads of one structure is otherwise folded into a
// load through a selected address, which keeps the whole walker in private (scratch) memory.
template <typename T> __device__ __forceinline__ T pick3(bool fy, bool fz, T vx, T vy, T vz)
{
    asm("" : "+v"(vx), "+v"(vy), "+v"(vz));
    return fz ? vz : fy ? vy : vx;
}

// True when the FP32 exit decision is ambiguous: the two smallest wall distances agree to 1e-5
// relative (a corner or edge crossing), or the chosen distance d is within 1e-5 of zero relative
// to the cell, measured along the exit axis (ka: direction component, wa: cell width on that axis).
// Distances of directions parallel to a wall are FLT_MAX and never count as ambiguous.
__device__ __forceinline__ bool ambiguous(float dx, float dy, float dz, float d, float ka, float wa)
{
    float mid = fmaxf(fminf(dx, dy), fminf(fmaxf(dx, dy), dz));
    return ((mid < FLT_MAX) & ((mid - d) <= 1e-5f * mid)) | (d * fabsf(ka) <= 1e-5f * wa);
}

// Mixed-precision Cartesian walk.
struct CartW32 {
    using Real = float;
    using View = CartView;
    float x, y, z, kx, ky, kz, ix, iy, iz;
    int i, j, k;
    int ne;
    float e0, e1, e2;
    bool done;
    uint32_t nrecompute, nstuck;
    static constexpr int kParkWords = 19;
    __device__ __forceinline__ static int cell_of(const View&, int n) { return n; }
    static constexpr bool kHint = false;
    __device__ __forceinline__ int entries() const { return ne; }
    static constexpr bool kStep = false;   // next() only
    template <class P> __device__ __forceinline__ void park(P& p)
    {
        p.io(x); p.io(y); p.io(z); p.io(kx); p.io(ky); p.io(kz); p.io(ix); p.io(iy); p.io(iz);
        p.io(i); p.io(j); p.io(k); p.io(ne); p.io(e0); p.io(e1); p.io(e2);
        p.io(done); p.io(nrecompute); p.io(nstuck);
    }

    __device__ __forceinline__ void push(float ds)
    {
        if (ds > 0) {
            // value selects, not an indexed store: keeps the walker out of scratch memory
            e0 = ne == 0 ? ds : e0;
            e1 = ne == 1 ? ds : e1;
            e2 = ne >= 2 ? ds : e2;
            ne++;
        }
    }
    __device__ __forceinline__ void init(const View& g, const int32_t*, int, double rxd, double ryd, double rzd, double dxd,
                         double dyd, double dzd)
    {
        // entry in FP64, state in FP32
        double x0 = rxd, y0 = ryd, z0 = rzd;
        ne = 0; done = false;
        double ds;
        if (x0 < g.xmin) {
            if (dxd <= 0.0) { done = true; return; }
            ds = (g.xmin - x0) / dxd; push((float)ds); x0 = g.xmin; y0 += dyd * ds; z0 += dzd * ds;
        } else if (x0 > g.xmax) {
            if (dxd >= 0.0) { done = true; return; }
            ds = (g.xmax - x0) / dxd; push((float)ds); x0 = g.xmax; y0 += dyd * ds; z0 += dzd * ds;
        }
        if (y0 < g.ymin) {
            if (dyd <= 0.0) { ne = 0; done = true; return; }
            ds = (g.ymin - y0) / dyd; push((float)ds); x0 += dxd * ds; y0 = g.ymin; z0 += dzd * ds;
        } else if (y0 > g.ymax) {
            if (dyd >= 0.0) { ne = 0; done = true; return; }
            ds = (g.ymax - y0) / dyd; push((float)ds); x0 += dxd * ds; y0 = g.ymax; z0 += dzd * ds;
        }
        if (z0 < g.zmin) {
            if (dzd <= 0.0) { ne = 0; done = true; return; }
            ds = (g.zmin - z0) / dzd; push((float)ds); x0 += dxd * ds; y0 += dyd * ds; z0 = g.zmin;
        } else if (z0 > g.zmax) {
            if (dzd >= 0.0) { ne = 0; done = true; return; }
            ds = (g.zmax - z0) / dzd; push((float)ds); x0 += dxd * ds; y0 += dyd * ds; z0 = g.zmax;
        }
        if (x0 < g.xmin || x0 > g.xmax || y0 < g.ymin || y0 > g.ymax || z0 < g.zmin || z0 > g.zmax) {
            ne = 0; done = true; return;
        }
        x = (float)x0; y = (float)y0; z = (float)z0;
        kx = (float)dxd; ky = (float)dyd; kz = (float)dzd;
        ix = fabsf(kx) > 1e-15f ? 1.0f / kx : 0.0f;
        iy = fabsf(ky) > 1e-15f ? 1.0f / ky : 0.0f;
        iz = fabsf(kz) > 1e-15f ? 1.0f / kz : 0.0f;
        i = locate_clip(g.xvf, g.Nx + 1, x);
        j = locate_clip(g.yvf, g.Ny + 1, y);
        k = locate_clip(g.zvf, g.Nz + 1, z);
    }
    __device__ __forceinline__ bool next(const View& g, const int32_t*, int, int& m, float& ds, float& rho)
    {
        if (ne) {
            ds = e0; e0 = e1; e1 = e2; ne--; m = -1; rho = 0.0f;
            return true;
        }
        while (!done) {
            int mm = k + g.Nz * j + g.Nz * g.Ny * i;
            int ie = (kx < 0.0f) ? i : i + 1, je = (ky < 0.0f) ? j : j + 1, ke = (kz < 0.0f) ? k : k + 1;
            float xE = g.xvf[ie], yE = g.yvf[je], zE = g.zvf[ke];
            float dsx = ix != 0.0f ? (xE - x) * ix : FLT_MAX;
            float dsy = iy != 0.0f ? (yE - y) * iy : FLT_MAX;
            float dsz = iz != 0.0f ? (zE - z) * iz : FLT_MAX;
            int a = choose_axis_cart(dsx, dsy, dsz);
            float d = a == 0 ? dsx : a == 1 ? dsy : dsz;
            if (ambiguous(dsx, dsy, dsz, d, a == 0 ? kx : a == 1 ? ky : kz, g.wminf)) {
                nrecompute++;
                double X = x, Y = y, Z = z;
                double ddx = ix != 0.0f ? (g.xv[ie] - X) / (double)kx : DBL_MAX;
                double ddy = iy != 0.0f ? (g.yv[je] - Y) / (double)ky : DBL_MAX;
                double ddz = iz != 0.0f ? (g.zv[ke] - Z) / (double)kz : DBL_MAX;
                a = choose_axis_cart(ddx, ddy, ddz);
                d = (float)(a == 0 ? ddx : a == 1 ? ddy : ddz);
            }
            if (a == 0) {
                i += (kx < 0.0f) ? -1 : 1;
                if (i >= g.Nx || i < 0) done = true;
                else { x = xE; y = fmaf(ky, d, y); z = fmaf(kz, d, z); }
            } else if (a == 1) {
                j += (ky < 0.0f) ? -1 : 1;
                if (j >= g.Ny || j < 0) done = true;
                else { x = fmaf(kx, d, x); y = yE; z = fmaf(kz, d, z); }
            } else {
                k += (kz < 0.0f) ? -1 : 1;
                if (k >= g.Nz || k < 0) done = true;
                else { x = fmaf(kx, d, x); y = fmaf(ky, d, y); z = zE; }
            }
            if (d > 0) { m = mm; ds = d; rho = g.rhof[mm]; return true; }
        }
        return false;
    }
};

// ====================================================================== trees

__device__ __forceinline__ int32_t node_word(const TreeView& t, const int32_t* lds, int nlds, int id)
{
    return id < nlds ? lds[id] : t.node[id];
}

// TreeNode::whichnode from the root (TreeNode.cpp:70-80) with the child choice of
// OctTreeNode.cpp:184-189 / BinTreeNode.cpp:326-335. Node bounds are not stored: each child box
// is recomputed from its parent's with the split at 0.5*(min+max) exactly as the reference
// creates it (OctTreeNode.cpp:38-49, BinTreeNode.cpp:40-70), so only one node word per level is
// read. Returns the leaf's cell number or -1 when the point is outside the root box.
__device__ __forceinline__ int descend64(const TreeView& t, const int32_t* lds, int nlds, double x, double y,
                                         double z, double* lo, double* hi)
{
    if (!box_contains(t.ext, x, y, z)) return -1;
    lo[0] = t.ext[0]; lo[1] = t.ext[1]; lo[2] = t.ext[2];
    hi[0] = t.ext[3]; hi[1] = t.ext[4]; hi[2] = t.ext[5];
    int32_t w = node_word(t, lds, nlds, 0);
    int level = 0;
    while (w >= 0) {
        if (t.arity == 8) {
            double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
            int q = 0;
            if (x < cx) hi[0] = cx; else { lo[0] = cx; q += 1; }
            if (y < cy) hi[1] = cy; else { lo[1] = cy; q += 2; }
            if (z < cz) hi[2] = cz; else { lo[2] = cz; q += 4; }
            w = node_word(t, lds, nlds, w + q);
        } else {
            int a = level % 3;
            double p = a == 0 ? x : a == 1 ? y : z;
            double l = a == 0 ? lo[0] : a == 1 ? lo[1] : lo[2];
            double h = a == 0 ? hi[0] : a == 1 ? hi[1] : hi[2];
            double c = 0.5 * (l + h);
            int q = (p < c) ? 0 : 1;
            if (q == 0) { if (a == 0) hi[0] = c; else if (a == 1) hi[1] = c; else hi[2] = c; }
            else { if (a == 0) lo[0] = c; else if (a == 1) lo[1] = c; else lo[2] = c; }
            w = node_word(t, lds, nlds, w + q);
        }
        level++;
    }
    return -w - 1;
}

// DustGridPath::moveInside (DustGridPath.cpp:57-150). Returns false for the OUTSIDE result.
struct Entry {
    int ne;
    double e0, e1, e2;
    __device__ __forceinline__ void push(double ds)
    {
        if (ds > 0) {
            // value selects, not an indexed store: keeps the walker out of scratch memory
            e0 = ne == 0 ? ds : e0;
            e1 = ne == 1 ? ds : e1;
            e2 = ne >= 2 ? ds : e2;
            ne++;
        }
    }
    __device__ __forceinline__ double pop()
    {
        double d = e0; e0 = e1; e1 = e2; ne--;
        return d;
    }
    __device__ __forceinline__ bool move_inside(const double* b, double eps, double& rx, double& ry, double& rz, double kx,
                                double ky, double kz)
    {
        ne = 0;
        if (rx <= b[0]) {
            if (kx <= 0.0) return false;
            double ds = (b[0] - rx) / kx; push(ds);
            rx = b[0] + eps; ry += ky * ds; rz += kz * ds;
        } else if (rx >= b[3]) {
            if (kx >= 0.0) return false;
            double ds = (b[3] - rx) / kx; push(ds);
            rx = b[3] - eps; ry += ky * ds; rz += kz * ds;
        }
        if (ry <= b[1]) {
            if (ky <= 0.0) return false;
            double ds = (b[1] - ry) / ky; push(ds);
            rx += kx * ds; ry = b[1] + eps; rz += kz * ds;
        } else if (ry >= b[4]) {
            if (ky >= 0.0) return false;
            double ds = (b[4] - ry) / ky; push(ds);
            rx += kx * ds; ry = b[4] - eps; rz += kz * ds;
        }
        if (rz <= b[2]) {
            if (kz <= 0.0) return false;
            double ds = (b[2] - rz) / kz; push(ds);
            rx += kx * ds; ry += ky * ds; rz = b[2] + eps;
        } else if (rz >= b[5]) {
            if (kz >= 0.0) return false;
            double ds = (b[5] - rz) / kz; push(ds);
            rx += kx * ds; ry += ky * ds; rz = b[5] - eps;
        }
        return true;
    }
};

// TreeDustGrid::path, TopDown (:393-437) and Neighbor (:441-502) methods, exact FP64.
struct TreeW64 {
    using Real = double;
    using View = TreeView;
    double x, y, z, kx, ky, kz;
    double lo[3], hi[3];
    double rho;
    int m;
    Entry en;
    bool done;
    uint32_t steps, nrecompute, nstuck;
    static constexpr int kParkWords = 38;
    __device__ __forceinline__ static int cell_of(const View&, int n) { return n; }
    static constexpr bool kHint = false;
    __device__ __forceinline__ int entries() const { return en.ne; }
    static constexpr bool kStep = false;   // next() only
    template <class P> __device__ __forceinline__ void park(P& p)
    {
        p.io(x); p.io(y); p.io(z); p.io(kx); p.io(ky); p.io(kz);
        p.io(lo[0]); p.io(lo[1]); p.io(lo[2]); p.io(hi[0]); p.io(hi[1]); p.io(hi[2]);
        p.io(rho); p.io(m); p.io(en.ne); p.io(en.e0); p.io(en.e1); p.io(en.e2);
        p.io(done); p.io(steps); p.io(nrecompute); p.io(nstuck);
    }

    __device__ __forceinline__ void load_cell(const View& t, int c)
    {
        const TreeRec64& r = t.rec64[c];
        lo[0] = r.lo[0]; lo[1] = r.lo[1]; lo[2] = r.lo[2];
        hi[0] = r.hi[0]; hi[1] = r.hi[1]; hi[2] = r.hi[2];
        rho = r.rho;
        m = c;
    }
    __device__ __forceinline__ bool descend(const View& t, const int32_t* lds, int nlds)
    {
        double l[3], h[3];
        int c = descend64(t, lds, nlds, x, y, z, l, h);
        if (c < 0) return false;
        lo[0] = l[0]; lo[1] = l[1]; lo[2] = l[2];
        hi[0] = h[0]; hi[1] = h[1]; hi[2] = h[2];
        m = c;
        rho = t.rec64[c].rho;
        return true;
    }
    __device__ __forceinline__ void init(const View& t, const int32_t* lds, int nlds, double rx, double ry, double rz, double dx,
                         double dy, double dz)
    {
        kx = dx; ky = dy; kz = dz;
        steps = 0; done = false;
        if (!en.move_inside(t.ext, t.eps, rx, ry, rz, kx, ky, kz)) {
            // OUTSIDE sentinel: no node contains it, path cleared
            en.ne = 0; done = true; return;
        }
        x = rx; y = ry; z = rz;
        if (!descend(t, lds, nlds)) { en.ne = 0; done = true; }
    }
    __device__ __forceinline__ bool next(const View& t, const int32_t* lds, int nlds, int& mo, double& ds,
                                         double& rh)
    {
        if (en.ne) { ds = en.pop(); mo = -1; rh = 0.0; return true; }
        while (!done) {
            double xnext = (kx < 0.0) ? lo[0] : hi[0];
            double ynext = (ky < 0.0) ? lo[1] : hi[1];
            double znext = (kz < 0.0) ? lo[2] : hi[2];
            double dsx = (fabs(kx) > 1e-15) ? (xnext - x) / kx : DBL_MAX;
            double dsy = (fabs(ky) > 1e-15) ? (ynext - y) / ky : DBL_MAX;
            double dsz = (fabs(kz) > 1e-15) ? (znext - z) / kz : DBL_MAX;
            int a = choose_axis_tree(dsx, dsy, dsz);
            double d = a == 0 ? dsx : a == 1 ? dsy : dsz;
            int wall = 2 * a + ((a == 0 ? kx : a == 1 ? ky : kz) < 0.0 ? 0 : 1);
            int segm = m;
            double segrho = rho;
            x += (d + t.eps) * kx;
            y += (d + t.eps) * ky;
            z += (d + t.eps) * kz;
            int oldm = m;
            bool found = false;
            if (t.search == 1) {
                int i0 = t.nboff[oldm * 6 + wall], i1 = t.nboff[oldm * 6 + wall + 1];
                for (int q = i0; q < i1; q++) {
                    int c = t.nbids[q];
                    const TreeRec64& r = t.rec64[c];
                    if (x >= r.lo[0] && x <= r.hi[0] && y >= r.lo[1] && y <= r.hi[1] && z >= r.lo[2] &&
                        z <= r.hi[2]) {
                        load_cell(t, c);
                        found = true;
                        break;
                    }
                }
            }
            if (!found) found = descend(t, lds, nlds);
            if (!found) done = true;
            else if (m == oldm) {
                nstuck++;
                x = nextafter(x, (kx < 0.0) ? -DBL_MAX : DBL_MAX);
                y = nextafter(y, (ky < 0.0) ? -DBL_MAX : DBL_MAX);
                z = nextafter(z, (kz < 0.0) ? -DBL_MAX : DBL_MAX);
                if (!descend(t, lds, nlds)) done = true;
                else if (m == oldm) done = true;
            }
            if (++steps > SG_MAX_STEPS) done = true;
            if (d > 0) { mo = segm; ds = d; rh = segrho; return true; }
        }
        return false;
    }
};

// TreeDustGrid::path, Bookkeeping method (TreeDustGrid.cpp:508-640), exact FP64, octrees only. The nodes are
// addressed in the reference's breadth-first creation order, in which every sibling group of eight starts
// at an index 1 + 8k, so a node's octant is (l - 1) % 8: the next cell is found by climbing to the first
// ancestor that has a sibling across the exit wall, stepping to that sibling (l -+ 1, 2 or 4) and
// descending to the leaf by the upper planes of child 0's box (a point on a plane goes to the lower child,
// the reference's <=). The tie rules are this method's own: x on <= with both, y on < x and <= z, z on <
// both; the exit coordinate is set to the wall exactly and no small step is added.
struct TreeW64Book {
    using Real = double;
    using View = TreeView;
    double x, y, z, kx, ky, kz;
    int l;                   // current node, breadth-first index
    Entry en;
    bool done;
    uint32_t steps, nrecompute, nstuck;
    static constexpr int kParkWords = 25;
    // segments carry cell numbers
    __device__ __forceinline__ static int cell_of(const View&, int n) { return n; }
    static constexpr bool kHint = false;
    __device__ __forceinline__ int entries() const { return en.ne; }
    static constexpr bool kStep = false;   // next() only
    template <class P> __device__ __forceinline__ void park(P& p)
    {
        p.io(x); p.io(y); p.io(z); p.io(kx); p.io(ky); p.io(kz);
        p.io(l); p.io(en.ne); p.io(en.e0); p.io(en.e1); p.io(en.e2);
        p.io(done); p.io(steps); p.io(nrecompute); p.io(nstuck);
    }
    __device__ __forceinline__ void init(const View& t, const int32_t*, int, double rx, double ry, double rz, double dx,
                                         double dy, double dz)
    {
        kx = dx; ky = dy; kz = dz;
        steps = 0; done = false; l = 0;
        if (!en.move_inside(t.ext, t.eps, rx, ry, rz, kx, ky, kz)) { en.ne = 0; done = true; return; }
        x = rx; y = ry; z = rz;
        // TreeNode::whichnode from the root (TreeNode.cpp:70-80): inside the closed root box, then
        // OctTreeNode::child by "<" against child 0's upper corner (OctTreeNode.cpp:184-189)
        const double* b = t.bkb;
        if (!(x >= b[0] && x <= b[3] && y >= b[1] && y <= b[4] && z >= b[2] && z <= b[5])) {
            en.ne = 0; done = true; return;
        }
        int n = 0;
        for (int level = 0; t.bkchild0[n] >= 0 && level < 64; level++) {
            const int c0 = t.bkchild0[n];
            const double* cb = t.bkb + 6 * (size_t)c0;
            n = c0 + (x < cb[3] ? 0 : 1) + (y < cb[4] ? 0 : 2) + (z < cb[5] ? 0 : 4);
        }
        l = n;
    }
    // child q of node n
    __device__ __forceinline__ static int child(const View& t, int n, int q) { return t.bkchild0[n] + q; }
    __device__ __forceinline__ bool next(const View& t, const int32_t*, int, int& mo, double& ds, double& rh)
    {
        if (en.ne) { ds = en.pop(); mo = -1; rh = 0.0; return true; }
        while (!done) {
            const double* nb = t.bkb + 6 * (size_t)l;
            const double xnext = (kx < 0.0) ? nb[0] : nb[3];
            const double ynext = (ky < 0.0) ? nb[1] : nb[4];
            const double znext = (kz < 0.0) ? nb[2] : nb[5];
            const double dsx = (fabs(kx) > 1e-15) ? (xnext - x) / kx : DBL_MAX;
            const double dsy = (fabs(ky) > 1e-15) ? (ynext - y) / ky : DBL_MAX;
            const double dsz = (fabs(kz) > 1e-15) ? (znext - z) / kz : DBL_MAX;
            const int cell = t.bkcell[l];
            double d;
            if (dsx <= dsy && dsx <= dsz) {
                d = dsx;
                x = xnext; y += ky * dsx; z += kz * dsx;
                // climb while the node lies on the side of its father the ray leaves through
                for (;;) {
                    const int oct = ((l - 1) % 8) + 1;
                    if (!((kx < 0.0) ? (oct % 2 == 1) : (oct % 2 == 0))) break;
                    l = t.bkfather[l];
                    if (l == 0) { done = true; break; }
                }
                if (!done) {
                    l += (kx < 0.0) ? -1 : 1;
                    while (t.bkcell[l] == -1) {
                        const double* cb = t.bkb + 6 * (size_t)child(t, l, 0);
                        const double yM = cb[4], zM = cb[5];
                        if (kx < 0.0) l = (y <= yM) ? child(t, l, (z <= zM) ? 1 : 5) : child(t, l, (z <= zM) ? 3 : 7);
                        else l = (y <= yM) ? child(t, l, (z <= zM) ? 0 : 4) : child(t, l, (z <= zM) ? 2 : 6);
                    }
                }
            } else if (dsy < dsx && dsy <= dsz) {
                d = dsy;
                x += kx * dsy; y = ynext; z += kz * dsy;
                for (;;) {
                    if (!((ky < 0.0) ? ((l - 1) % 4 < 2) : ((l - 1) % 4 > 1))) break;
                    l = t.bkfather[l];
                    if (l == 0) { done = true; break; }
                }
                if (!done) {
                    l += (ky < 0.0) ? -2 : 2;
                    while (t.bkcell[l] == -1) {
                        const double* cb = t.bkb + 6 * (size_t)child(t, l, 0);
                        const double xM = cb[3], zM = cb[5];
                        if (ky < 0.0) l = (x <= xM) ? child(t, l, (z <= zM) ? 2 : 6) : child(t, l, (z <= zM) ? 3 : 7);
                        else l = (x <= xM) ? child(t, l, (z <= zM) ? 0 : 4) : child(t, l, (z <= zM) ? 1 : 5);
                    }
                }
            } else if (dsz < dsx && dsz < dsy) {
                d = dsz;
                x += kx * dsz; y += ky * dsz; z = znext;
                for (;;) {
                    const int oct = ((l - 1) % 8) + 1;
                    if (!((kz < 0.0) ? (oct < 5) : (oct > 4))) break;
                    l = t.bkfather[l];
                    if (l == 0) { done = true; break; }
                }
                if (!done) {
                    l += (kz < 0.0) ? -4 : 4;
                    while (t.bkcell[l] == -1) {
                        const double* cb = t.bkb + 6 * (size_t)child(t, l, 0);
                        const double xM = cb[3], yM = cb[4];
                        if (kz < 0.0) l = (x <= xM) ? child(t, l, (y <= yM) ? 4 : 6) : child(t, l, (y <= yM) ? 5 : 7);
                        else l = (x <= xM) ? child(t, l, (y <= yM) ? 0 : 2) : child(t, l, (y <= yM) ? 1 : 3);
                    }
                }
            } else {
                // only with NaN input, where the reference would not terminate
                done = true;
                return false;
            }
            if (++steps > SG_MAX_STEPS) done = true;
            // DustGridPath::addSegment keeps only segments of positive length
            if (d > 0) { mo = cell; ds = d; rh = t.rec64[cell].rho; return true; }
        }
        return false;
    }
};

// Mixed-precision tree walk (Neighbor semantics). Every step reads one 32-byte node record (two
// 16-byte loads issued together). The walker carries the current leaf's integer indices along x, y, z
// and its subdivision counts; its exit planes (the walls on the sides the ray heads to, each
// ext + k * width / 2^n as one fused multiply-add, the same float wherever it is computed) and its
// ropes across them are computed when the leaf is taken over, so a step is: three guarded FP32
// distances, the smallest (FP64 decides when the FP32 decision is ambiguous: the two smallest agree
// to 1e-5 relative, an edge or corner crossing, or the smallest is below 1e-5 of the second, a
// position on or just past a wall after rounding), the position snapped onto the exit wall, and the
// next leaf: the rope's node, or a descent from it when the rope is coarser. A directed root
// descent covers the rare case of a point that left the rope node's face through an edge or corner
// by rounding.
struct TreeW32 {
    using Real = float;
    using View = TreeView;
    float x, y, z, kx, ky, kz, ix, iy, iz;
    float gx, gy, gz;        // 0, or FLT_MAX along an axis the direction is parallel to (|k| < 1e-15)
    float px, py, pz;        // exit planes of the current leaf
    int32_t ci, cj, ck;      // its integer indices along x, y, z
    uint32_t nsub;           // its subdivision counts, nx | ny << 5 | nz << 10
    int32_t rx, ry, rz;      // its ropes across the exit walls
    float rho;
    int m;                   // its node id
    int lm;                  // node of the last segment returned
    int ne;
    float e0, e1, e2;
    bool done;
    uint32_t steps, nrecompute, nstuck;
    static constexpr int kParkWords = 34;
    static constexpr bool kHint = true;
    template <class P> __device__ __forceinline__ void park(P& p)
    {
        p.io(x); p.io(y); p.io(z); p.io(kx); p.io(ky); p.io(kz); p.io(ix); p.io(iy); p.io(iz);
        p.io(gx); p.io(gy); p.io(gz); p.io(px); p.io(py); p.io(pz); p.io(ci); p.io(cj); p.io(ck); p.io(nsub);
        p.io(rx); p.io(ry); p.io(rz); p.io(rho); p.io(m); p.io(lm); p.io(ne); p.io(e0); p.io(e1); p.io(e2);
        p.io(done); p.io(steps); p.io(nrecompute); p.io(nstuck);
    }
    // cell number of a walk segment's node
    __device__ __forceinline__ static int cell_of(const View& t, int n) { return t.cellof[n]; }
    __device__ __forceinline__ int entries() const { return ne; }

    // Node record n as two 16-byte loads issued together: (code, val, rope0, rope1), (rope2..rope5).
    __device__ __forceinline__ static void load(const View& t, int n, uint4& a, uint4& b)
    {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4* p = reinterpret_cast<const u32x4*>(t.node32 + n);
        u32x4 va = p[0], vb = p[1];
        // both halves at once: otherwise the compiler loads the first word alone and sinks the rest
        // of the record into the branch that uses it, a second dependent memory access per step
        asm("" : "+v"(va), "+v"(vb));
        a = make_uint4(va.x, va.y, va.z, va.w);
        b = make_uint4(vb.x, vb.y, vb.z, vb.w);
    }
    // Takes over leaf n (record a, b; indices i, j, k at its subdivision): exit planes and ropes on the
    // sides the ray heads to.
    __device__ __forceinline__ void take(const View& t, uint4 a, uint4 b, int n, int32_t i, int32_t j, int32_t k)
    {
        const bool nx = kx < 0.0f, ny = ky < 0.0f, nz = kz < 0.0f;
        const uint32_t ns = a.x & 0x7fffu;
        const float wx = ldexpf(t.widthf[0], -(int)(ns & 31u));
        const float wy = ldexpf(t.widthf[1], -(int)((ns >> 5) & 31u));
        const float wz = ldexpf(t.widthf[2], -(int)((ns >> 10) & 31u));
        px = fmaf((float)(nx ? i : i + 1), wx, t.extf[0]);
        py = fmaf((float)(ny ? j : j + 1), wy, t.extf[1]);
        pz = fmaf((float)(nz ? k : k + 1), wz, t.extf[2]);
        ci = i; cj = j; ck = k;
        nsub = ns;
        rho = __uint_as_float(a.y);
        rx = (int32_t)(nx ? a.z : a.w);
        ry = (int32_t)(ny ? b.x : b.y);
        rz = (int32_t)(nz ? b.z : b.w);
        m = n;
    }
    // Descends from node n (record a, b already loaded; indices i, j, k at its subdivision) to the leaf
    // containing the current position and takes it over. Child choice by "<" against the split point
    // (OctTreeNode.cpp:184-189, BinTreeNode.cpp:326-335), except that a point lying exactly on a split
    // plane goes to the child the ray is heading into (a walk never steps back into a cell behind it;
    // the reference instead steps a small distance past every wall). Per-axis choices are selects on
    // flags, not indexed arrays, so that nothing goes to private memory.
    __device__ __forceinline__ void descend_loaded(const View& t, uint4 a, uint4 b, int n, int32_t i, int32_t j,
                                                   int32_t k)
    {
        // upper child: beyond the plane, or on it while heading up (bitwise forms stay in mask registers)
        const bool ux = !(kx < 0.0f), uy = !(ky < 0.0f), uz = !(kz < 0.0f);
        int level = 0;
        while (!(a.x & kTreeLeaf)) {
            if (++level > 128) { done = true; return; }   // no tree the records describe is this deep
            const int c0 = (int)a.y;
            if (t.arity == 8) {
                // split planes ext + (2k + 1) * width / 2^(n + 1), held by the internal node's record
                const float cx = __uint_as_float(a.z), cy = __uint_as_float(a.w), cz = __uint_as_float(b.x);
                const bool qx = (x > cx) | ((x == cx) & ux);
                const bool qy = (y > cy) | ((y == cy) & uy);
                const bool qz = (z > cz) | ((z == cz) & uz);
                n = c0 + (qx ? 1 : 0) + (qy ? 2 : 0) + (qz ? 4 : 0);
                i = 2 * i + (qx ? 1 : 0); j = 2 * j + (qy ? 1 : 0); k = 2 * k + (qz ? 1 : 0);
            } else {
                // binary node: split at the midpoint of its recorded axis (bits 16-17: 0 x, 1 y, 2 z)
                const bool fy = (a.x >> 16) & 1u, fz = (a.x >> 17) & 1u;
                const float p = pick3(fy, fz, x, y, z);
                const bool up = fz ? uz : fy ? uy : ux;
                const float c = __uint_as_float(a.z);   // the split plane, held by the record
                const int q = ((p > c) | ((p == c) & up)) ? 1 : 0;
                i = (fy || fz) ? i : 2 * i + q;
                j = fy ? 2 * j + q : j;
                k = fz ? 2 * k + q : k;
                n = c0 + q;
            }
            load(t, n, a, b);
        }
        take(t, a, b, n, i, j, k);
    }
    __device__ __forceinline__ void descend_from(const View& t, int n, int32_t i, int32_t j, int32_t k)
    {
        uint4 a, b;
        load(t, n, a, b);
        descend_loaded(t, a, b, n, i, j, k);
    }
    // Root descent in FP32 over the node words (the top levels from LDS), with every split plane
    // computed by the same single-rounding formula as the records hold. At the start of a walk (xa < 0)
    // by "<" as the reference; after a step through a wall on axis xa, every axis is compared in the
    // travel direction, so a point lying exactly on a wall goes to the cell the ray is heading into.
    // Returns the leaf's node id (and its indices) or -1.
    __device__ __forceinline__ int descend_root(const View& t, const int32_t* lds, int nlds, int xa, int xdir,
                                                int32_t& i, int32_t& j, int32_t& k) const
    {
        const float* e = t.extf;
        // inside the grid: closed box, except half-open on the side of the wall just crossed
        const bool ax = xa == 0, ay = xa == 1, az = xa == 2, up = xdir > 0;
        const bool inx = x >= e[0] && x <= e[3] && !(ax && (up ? x >= e[3] : x <= e[0]));
        const bool iny = y >= e[1] && y <= e[4] && !(ay && (up ? y >= e[4] : y <= e[1]));
        const bool inz = z >= e[2] && z <= e[5] && !(az && (up ? z >= e[5] : z <= e[2]));
        if (!(inx && iny && inz)) return -1;
        const bool dir = xa >= 0;
        const bool nx = dir && kx < 0.0f, ny = dir && ky < 0.0f, nz = dir && kz < 0.0f;
        int32_t ki = 0, kj = 0, kk = 0;
        int id = 0;
        int32_t w = node_word(t, lds, nlds, 0);
        int level = 0;
        int axis = 0;   // binary tree: split axis of this level (x, y, z in turn)
        while (w >= 0) {
            int q = 0;
            if (t.arity == 8) {
                const float cx = fmaf((float)(2 * ki + 1), ldexpf(t.widthf[0], -(level + 1)), e[0]);
                const float cy = fmaf((float)(2 * kj + 1), ldexpf(t.widthf[1], -(level + 1)), e[1]);
                const float cz = fmaf((float)(2 * kk + 1), ldexpf(t.widthf[2], -(level + 1)), e[2]);
                const int qx = (nx ? x <= cx : x < cx) ? 0 : 1;
                const int qy = (ny ? y <= cy : y < cy) ? 0 : 1;
                const int qz = (nz ? z <= cz : z < cz) ? 0 : 1;
                ki = 2 * ki + qx; kj = 2 * kj + qy; kk = 2 * kk + qz;
                q = qx + 2 * qy + 4 * qz;
            } else {
                const bool fy = axis == 1, fz = axis == 2;
                const float ea = pick3(fy, fz, e[0], e[1], e[2]);
                const float wa = pick3(fy, fz, t.widthf[0], t.widthf[1], t.widthf[2]);
                const int32_t ka = fz ? kk : fy ? kj : ki;
                const float p = pick3(fy, fz, x, y, z);
                const bool neg = fz ? nz : fy ? ny : nx;
                const float c = fmaf((float)(2 * ka + 1), ldexpf(wa, -(level / 3 + 1)), ea);
                q = (neg ? p <= c : p < c) ? 0 : 1;
                ki = (fy || fz) ? ki : 2 * ki + q;
                kj = fy ? 2 * kj + q : kj;
                kk = fz ? 2 * kk + q : kk;
                axis = fz ? 0 : axis + 1;
            }
            id = w + q;
            w = node_word(t, lds, nlds, id);
            level++;
        }
        i = ki; j = kj; k = kk;
        return id;
    }
    __device__ __forceinline__ void push(float ds)
    {
        if (ds > 0) {
            // value selects, not an indexed store: keeps the walker out of scratch memory
            e0 = ne == 0 ? ds : e0;
            e1 = ne == 1 ? ds : e1;
            e2 = ne >= 2 ? ds : e2;
            ne++;
        }
    }
    __device__ __forceinline__ void set_direction(double dx, double dy, double dz)
    {
        kx = (float)dx; ky = (float)dy; kz = (float)dz;
        const bool ox = fabsf(kx) > 1e-15f, oy = fabsf(ky) > 1e-15f, oz = fabsf(kz) > 1e-15f;
        ix = ox ? 1.0f / kx : 0.0f;
        iy = oy ? 1.0f / ky : 0.0f;
        iz = oz ? 1.0f / kz : 0.0f;
        gx = ox ? 0.0f : FLT_MAX;
        gy = oy ? 0.0f : FLT_MAX;
        gz = oz ? 0.0f : FLT_MAX;
    }
    __device__ __forceinline__ void init(const View& t, const int32_t* lds, int nlds, double rx_, double ry_, double rz_,
                                         double dx, double dy, double dz)
    {
        steps = 0; done = false; m = -1; lm = -1;
        Entry en;
        if (!en.move_inside(t.ext, t.eps, rx_, ry_, rz_, dx, dy, dz)) { ne = 0; done = true; return; }
        ne = 0;
        for (int q = 0; q < en.ne; q++) push((float)(q == 0 ? en.e0 : q == 1 ? en.e1 : en.e2));
        x = (float)rx_; y = (float)ry_; z = (float)rz_;
        set_direction(dx, dy, dz);
        int32_t i, j, k;
        const int n = descend_root(t, lds, nlds, -1, 0, i, j, k);
        if (n < 0) { ne = 0; done = true; return; }
        descend_from(t, n, i, j, k);
    }
    // The leaf of the last segment returned (-1 before the first one): after a walk that stopped
    // inside a segment, the leaf holding the stopping point.
    __device__ __forceinline__ int hint() const { return lm; }
    // init for a start point strictly inside the grid with a candidate leaf hn: the leaf is taken when
    // it contains the point in the half-open sense lo <= p < hi (closed at the grid's upper walls),
    // which is exactly the leaf the root descent selects; otherwise init.
    __device__ __forceinline__ void init_hinted(const View& t, const int32_t* lds, int nlds, double rx_, double ry_,
                                                double rz_, double dx, double dy, double dz, int hn)
    {
        const double* e = t.ext;
        const bool inside = rx_ > e[0] && rx_ < e[3] && ry_ > e[1] && ry_ < e[4] && rz_ > e[2] && rz_ < e[5];
        if (!(inside && hn >= 0 && hn < t.nnodes)) { init(t, lds, nlds, rx_, ry_, rz_, dx, dy, dz); return; }
        steps = 0; done = false; m = -1; lm = -1; ne = 0;
        x = (float)rx_; y = (float)ry_; z = (float)rz_;
        set_direction(dx, dy, dz);
        uint4 a, b;
        load(t, hn, a, b);
        const uint64_t kp = t.kpack[hn];
        const int32_t hi = (int32_t)(kp & 0x1fffffu), hj = (int32_t)((kp >> 21) & 0x1fffffu),
                      hk = (int32_t)((kp >> 42) & 0x1fffffu);
        bool ok = (a.x & kTreeLeaf) != 0;
        if (ok) {
            const uint32_t ns = a.x & 0x7fffu;
            const float p[3] = {x, y, z};
            const int32_t kk[3] = {hi, hj, hk};
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int n = (int)((ns >> (5 * q)) & 31u);
                const float w = ldexpf(t.widthf[q], -n);
                const float lo = fmaf((float)kk[q], w, t.extf[q]), up = fmaf((float)(kk[q] + 1), w, t.extf[q]);
                const bool top = kk[q] + 1 == (1 << n);
                ok = ok && p[q] >= lo && (p[q] < up || (top && p[q] <= up));
            }
        }
        if (ok) { take(t, a, b, hn, hi, hj, hk); return; }
        int32_t i, j, k;
        const int n = descend_root(t, lds, nlds, -1, 0, i, j, k);
        if (n < 0) { done = true; return; }
        descend_from(t, n, i, j, k);
    }
    // Chooses the exit wall and moves the position onto it; returns the segment length and the exit
    // axis as flags (fx: x, fy: y, neither: z). The choices are selects on these flags, never an axis
    // number, which the compiler would turn into branches or an indexed private array.
    __device__ __forceinline__ float exit_wall(bool& fx, bool& fy)
    {
        // guarded distances: an axis parallel to the direction gives FLT_MAX
        const float dsx = fmaf(px - x, ix, gx);
        const float dsy = fmaf(py - y, iy, gy);
        const float dsz = fmaf(pz - z, iz, gz);
        float d = fminf(fminf(dsx, dsy), dsz);
        const float mid = __builtin_amdgcn_fmed3f(dsx, dsy, dsz);
        // tree tie rule (TreeDustGrid.cpp:451-466): x on ties with y or z, then y on ties with z
        fx = dsx == d;
        fy = !fx && dsy == d;
        const float tol = 1e-5f * mid;
        if ((mid < FLT_MAX) & (((mid - d) <= tol) | (d <= tol))) {
            nrecompute++;
            const double X = x, Y = y, Z = z;
            const double ddx = ix != 0.0f ? ((double)px - X) / (double)kx : DBL_MAX;
            const double ddy = iy != 0.0f ? ((double)py - Y) / (double)ky : DBL_MAX;
            const double ddz = iz != 0.0f ? ((double)pz - Z) / (double)kz : DBL_MAX;
            fx = ddx <= ddy && ddx <= ddz;
            fy = !fx && ddy <= ddz;
            d = (float)(fx ? ddx : fy ? ddy : ddz);
        }
        const float dd = d > 0.0f ? d : 0.0f;   // a select: fmaxf would first canonicalize d
        const bool fz = !fx && !fy;
        x = fx ? px : fmaf(kx, dd, x);
        y = fy ? py : fmaf(ky, dd, y);
        z = fz ? pz : fmaf(kz, dd, z);
        return d;
    }
    // One wall step: an entry segment, or the segment from the position to the next wall, whose length
    // is zero (or below) when the position lies on an edge or corner of the cell; false at the end of
    // the path. Transport loops take one step per iteration and skip segments with ds <= 0, so that no
    // loop nests inside theirs.
    static constexpr bool kStep = true;
    __device__ __forceinline__ bool step(const View& t, const int32_t* lds, int nlds, int& mo, float& ds, float& rh)
    {
        if (ne) {
            ds = e0; e0 = e1; e1 = e2; ne--; mo = -1; rh = 0.0f;
            return true;
        }
        if (done) return false;
        const int segm = m;
        const float segrho = rho;
        bool fx, fy;
        const float d = exit_wall(fx, fy);
        const bool fz = !fx && !fy;
        int32_t r0 = rx, r1 = ry, r2 = rz;
        asm("" : "+v"(r0), "+v"(r1), "+v"(r2));
        const int r = fx ? r0 : fy ? r1 : r2;
        if (r < 0) done = true;   // a wall of the grid: the path leaves the grid
        else {
            uint4 a, b;
            load(t, r, a, b);
            // the position lies on the exit wall and, having moved away from the entry walls, at or
            // beyond the leaf's lower walls in the travel direction: it is on the leaf's face (and so in
            // the rope node's box, which covers the whole face) exactly when it has not passed the exit
            // planes of the two other axes, which rounding can do by an ulp near an edge. Not passed:
            // (p - x) k >= 0 (zero on the exit axis, where the position lies on the plane)
            const float on = fminf(fminf((px - x) * kx, (py - y) * ky), (pz - z) * kz);
            // indices across the exit wall at the rope node's subdivision (a rope is never finer than
            // its leaf along any axis, so the packed per-axis differences do not borrow)
            const uint32_t dn = nsub - (a.x & 0x7fffu);
            const int32_t i = (ci + (fx ? (kx < 0.0f ? -1 : 1) : 0)) >> (dn & 31u);
            const int32_t j = (cj + (fy ? (ky < 0.0f ? -1 : 1) : 0)) >> ((dn >> 5) & 31u);
            const int32_t k = (ck + (fz ? (kz < 0.0f ? -1 : 1) : 0)) >> ((dn >> 10) & 31u);
            if (on >= 0.0f) descend_loaded(t, a, b, r, i, j, k);
            else {
                const int xa = fx ? 0 : fy ? 1 : 2;
                const float ka = fx ? kx : fy ? ky : kz;
                int32_t i2, j2, k2;
                const int n = descend_root(t, lds, nlds, xa, ka < 0.0f ? -1 : 1, i2, j2, k2);
                if (n < 0) done = true;
                else descend_from(t, n, i2, j2, k2);
            }
            if (!done && m == segm) { nstuck++; done = true; }
        }
        if (++steps > SG_MAX_STEPS) done = true;
        mo = segm; ds = d; rh = segrho;
        lm = d > 0.0f ? segm : lm;
        return true;
    }
    // The next segment of positive length.
    __device__ __forceinline__ bool next(const View& t, const int32_t* lds, int nlds, int& mo, float& ds, float& rh)
    {
        while (step(t, lds, nlds, mo, ds, rh))
            if (ds > 0) return true;
        return false;
    }
};

// Mixed-precision tree walk with TopDown semantics (every next cell by directed root descent).
struct TreeW32TopDown : TreeW32 {
    __device__ __forceinline__ bool step(const View& t, const int32_t* lds, int nlds, int& mo, float& ds,
                                         float& rh)
    {
        if (ne) {
            ds = e0; e0 = e1; e1 = e2; ne--; mo = -1; rh = 0.0f;
            return true;
        }
        if (done) return false;
        const int segm = m;
        const float segrho = rho;
        bool fx, fy;
        const float d = exit_wall(fx, fy);
        const int xa = fx ? 0 : fy ? 1 : 2;
        const float ka = fx ? kx : fy ? ky : kz;
        int32_t i, j, k;
        const int n = descend_root(t, lds, nlds, xa, ka < 0.0f ? -1 : 1, i, j, k);
        if (n < 0) done = true;
        else {
            descend_from(t, n, i, j, k);
            if (m == segm) { nstuck++; done = true; }
        }
        if (++steps > SG_MAX_STEPS) done = true;
        mo = segm; ds = d; rh = segrho;
        lm = d > 0.0f ? segm : lm;
        return true;
    }
    __device__ __forceinline__ bool next(const View& t, const int32_t* lds, int nlds, int& mo, float& ds,
                                         float& rh)
    {
        while (step(t, lds, nlds, mo, ds, rh))
            if (ds > 0) return true;
        return false;
    }
};

// ====================================================================== Voronoi

// VoronoiMesh::cellIndex (VoronoiMesh.cpp:541-570): the nearest site, ties to the lower index,
// found over a uniform bucket grid by growing shells until no closer site can remain.
__device__ int voro_nearest(const VoroView& g, double x, double y, double z)
{
    if (!box_contains(g.ext, x, y, z)) return -1;
    int n = g.nbk;
    double w[3] = {(g.ext[3] - g.ext[0]) / n, (g.ext[4] - g.ext[1]) / n, (g.ext[5] - g.ext[2]) / n};
    int ci = min(n - 1, max(0, (int)(n * (x - g.ext[0]) / (g.ext[3] - g.ext[0]))));
    int cj = min(n - 1, max(0, (int)(n * (y - g.ext[1]) / (g.ext[4] - g.ext[1]))));
    int ck = min(n - 1, max(0, (int)(n * (z - g.ext[2]) / (g.ext[5] - g.ext[2]))));
    double wmin = fmin(w[0], fmin(w[1], w[2]));
    double bd = DBL_MAX;
    int best = -1;
    for (int ring = 0; ring <= n; ring++) {
        for (int i = max(0, ci - ring); i <= min(n - 1, ci + ring); i++)
            for (int j = max(0, cj - ring); j <= min(n - 1, cj + ring); j++)
                for (int k = max(0, ck - ring); k <= min(n - 1, ck + ring); k++) {
                    int cheb = max(abs(i - ci), max(abs(j - cj), abs(k - ck)));
                    if (cheb != ring) continue;
                    int q = (i * n + j) * n + k;
                    for (int t = g.boff[q]; t < g.boff[q + 1]; t++) {
                        int s = g.bids[t];
                        VoroSite p = g.site[s];
                        double dx = x - p.x, dy = y - p.y, dz = z - p.z;
                        double d = dx * dx + dy * dy + dz * dz;
                        if (d < bd || (d == bd && s < best)) { bd = d; best = s; }
                    }
                }
        double reach = ring * wmin * (1.0 - 1e-12);
        if (best >= 0 && reach * reach > bd) break;
    }
    return best;
}

// VoronoiMesh::cellIndex by a greedy walk over the neighbour lists from a start site (a known
// nearby cell, else the start site of the point's bucket): move to the neighbour nearest the point
// while one is nearer than the current site (ties to the lower index). A point of the box nearer to
// site h than to all of h's neighbours lies in h's cell (the segment from h to the point leaves the
// cell through a face shared with a nearer neighbour), so the walk ends at the nearest site. When
// some neighbour is within 1e-9 (relative, squared distances) as near as the final site, the point
// may sit on a bisector, where rounding or a tie decides: the exact search settles it. Each move
// strictly decreases (distance, index), so the walk ends.
__device__ int voro_locate(const VoroView& g, double x, double y, double z, int start)
{
    if (!box_contains(g.ext, x, y, z)) return -1;
    int cur = start;
    if (cur < 0 || cur >= g.nsites) {
        const int n = g.nbk;
        const int ci = min(n - 1, max(0, (int)(n * (x - g.ext[0]) / (g.ext[3] - g.ext[0]))));
        const int cj = min(n - 1, max(0, (int)(n * (y - g.ext[1]) / (g.ext[4] - g.ext[1]))));
        const int ck = min(n - 1, max(0, (int)(n * (z - g.ext[2]) / (g.ext[5] - g.ext[2]))));
        cur = g.brep[(ci * n + cj) * n + ck];
    }
    VoroSite s = g.site[cur];
    double dc = (x - s.x) * (x - s.x) + (y - s.y) * (y - s.y) + (z - s.z) * (z - s.z);
    bool thin = false;
    for (int it = 0; it < g.nsites; it++) {
        int best = cur;
        double db = dc;
        thin = false;
        const int2 be = g.off[cur];
        const int t0 = be.x + 1, t1 = be.x + 1 + be.y;
        // entries in groups of eight loaded together; past the end a group repeats the last entry,
        // whose second test changes nothing (same distance and index)
        for (int t = t0; t < t1; t += 8) {
            const int last = t1 - 1;
            VoroNbr q[8];
#pragma unroll
            for (int u = 0; u < 8; u++) q[u] = g.nbr[min(t + u, last)];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int mi = q[u].id;
                const double d = (x - q[u].x) * (x - q[u].x) + (y - q[u].y) * (y - q[u].y) + (z - q[u].z) * (z - q[u].z);
                const bool site = mi >= 0;
                thin = thin || (site && d - dc <= 1e-9 * d);
                if (site && (d < db || (d == db && mi < best))) { db = d; best = mi; }
            }
        }
        if (best == cur) break;
        cur = best;
        dc = db;
    }
    return thin ? voro_nearest(g, x, y, z) : cur;
}

// VoronoiMesh::path (VoronoiMesh.cpp:778-873): the bisector-plane tests stay FP64 in both
// precision modes; Real only sets the type of the emitted segment lengths.
// The distance along the ray from (x, y, z) to the bisector plane of the sites p (the cell's) and q (a
// neighbour's), when the plane lies ahead (n.k > 0, n = q - p), else 0 (VoronoiMesh.cpp:806-858).
__device__ __forceinline__ double voro_site_si(double x, double y, double z, double kx, double ky, double kz, double px,
                                               double py, double pz, double qx, double qy, double qz)
{
    const double nx = qx - px, ny = qy - py, nz = qz - pz;
    const double ndotk = nx * kx + ny * ky + nz * kz;
    if (!(ndotk > 0)) return 0;
    const double mx = 0.5 * (qx + px), my = 0.5 * (qy + py), mz = 0.5 * (qz + pz);
    return (nx * (mx - x) + ny * (my - y) + nz * (mz - z)) / ndotk;
}
// The distance to the box wall of entry code mi (-1..-6: -x +x -y +y -z +z).
__device__ __forceinline__ double voro_wall_si(const VoroView& g, double x, double y, double z, double kx, double ky,
                                               double kz, int mi)
{
    switch (mi) {
    case -1: return (g.ext[0] - x) / kx;
    case -2: return (g.ext[3] - x) / kx;
    case -3: return (g.ext[1] - y) / ky;
    case -4: return (g.ext[4] - y) / ky;
    case -5: return (g.ext[2] - z) / kz;
    case -6: return (g.ext[5] - z) / kz;
    default: return 0;
    }
}

#ifndef SG_VORO_GROUP
#define SG_VORO_GROUP 16
#endif
// Voronoi walk (VoronoiMesh.cpp:778-873) as single steps: a step tests one group of entries (16, loaded
// together) of the current cell's list in list order (the reference's loop, cut into pieces), and the
// step that ends the list moves the position and returns the cell's segment; the other steps return no
// segment. A wave then runs as many plane tests as its lanes need group by group, instead of running
// every cell's loop to the length of the wave's longest list. The cell's list offset is loaded as soon
// as the cell is entered.
template <typename R> struct VoroW {
    using Real = R;
    using View = VoroView;
    static constexpr int NO_INDEX = -99;
    static constexpr int kGroup = SG_VORO_GROUP;   // entries tested per step
    double x, y, z, kx, ky, kz;
    double sqc;              // the current cell's nearest plane ahead so far, and its entry (mqc)
    int mr, lm, mqc;
    int hb, tcur, tend;      // the current cell's header record, next entry to test, end of its entries
    Entry en;
    bool done, segout;
    uint32_t steps, nrecompute, nstuck;
    static constexpr int kParkWords = 31;
    __device__ __forceinline__ static int cell_of(const View&, int n) { return n; }
    static constexpr bool kHint = true;
    __device__ __forceinline__ int entries() const { return en.ne; }
    static constexpr bool kStep = true;
    template <class P> __device__ __forceinline__ void park(P& p)
    {
        p.io(x); p.io(y); p.io(z); p.io(kx); p.io(ky); p.io(kz); p.io(sqc);
        p.io(mr); p.io(lm); p.io(mqc); p.io(hb); p.io(tcur); p.io(tend);
        p.io(en.ne); p.io(en.e0); p.io(en.e1); p.io(en.e2);
        p.io(done); p.io(steps); p.io(nrecompute); p.io(nstuck);
    }

    __device__ __forceinline__ void init(const View& g, const int32_t* lds, int nlds, double rx, double ry, double rz,
                                         double dx, double dy, double dz)
    {
        init_hinted(g, lds, nlds, rx, ry, rz, dx, dy, dz, -1);
    }
    // The cell of the last segment returned (-1 before the first one): after a walk that stopped
    // inside a segment, the cell holding the stopping point.
    __device__ __forceinline__ int hint() const { return lm; }
    // Enters cell mr: its list's place, and no plane found yet.
    __device__ __forceinline__ void enter(const View& g)
    {
        const int2 be = g.off[mr];
        hb = be.x; tcur = be.x + 1; tend = be.x + 1 + be.y;
        sqc = DBL_MAX; mqc = NO_INDEX;
    }
    // init with a candidate cell hn (or -1): the greedy location starts there; the cell found is the
    // nearest site whatever the start
    __device__ __forceinline__ void init_hinted(const View& g, const int32_t*, int, double rx, double ry, double rz,
                                                double dx, double dy, double dz, int hn)
    {
        kx = dx; ky = dy; kz = dz;
        steps = 0; done = false; lm = -1;
        hb = 0; tcur = 0; tend = 0; sqc = DBL_MAX; mqc = NO_INDEX;
        if (!en.move_inside(g.ext, g.eps, rx, ry, rz, kx, ky, kz)) { en.ne = 0; done = true; return; }
        x = rx; y = ry; z = rz;
        // a start point moved onto the box from outside is not where the candidate was found
        mr = voro_locate(g, x, y, z, en.ne ? -1 : hn);
        if (mr < 0) { en.ne = 0; done = true; return; }
        enter(g);
    }
    // The distance along the ray to the bisector plane (or wall) of one neighbour entry of the cell with
    // site pr, taken when it lies ahead and is the nearest so far (VoronoiMesh.cpp:806-858).
    __device__ __forceinline__ void plane(const View& g, const VoroSite& pr, const VoroNbr& pi, double& sq,
                                          int& mq) const
    {
        const int mi = pi.id;
        const double si = mi >= 0 ? voro_site_si(x, y, z, kx, ky, kz, pr.x, pr.y, pr.z, pi.x, pi.y, pi.z)
                                  : voro_wall_si(g, x, y, z, kx, ky, kz, mi);
        if (si > 0 && si < sq) { sq = si; mq = mi; }
    }
    // One step: an entry segment; or kGroup more entries of the cell's list (no segment, segout false);
    // or, when they end the list, the cell's segment (the position moves onto the next cell). False at
    // the end of the path.
    __device__ __forceinline__ bool step(const View& g, const int32_t*, int, int& mo, R& ds, R& rh)
    {
        segout = true;
        if (en.ne) { ds = (R)en.pop(); mo = -1; rh = (R)0; return true; }
        if (done) return false;
        // the cell's header record (its site and density), then its entries, contiguous; every cell has at
        // least one entry (checked at creation), and past the end a group repeats the last entry, whose
        // second test changes nothing: the nearest plane ahead is taken on a strict "<"
        const VoroSite pr = *reinterpret_cast<const VoroSite*>(g.nbr + hb);
        const int t = tcur, last = tend - 1;
        VoroNbr q[kGroup];
#pragma unroll
        for (int i = 0; i < kGroup; i++) q[i] = g.nbr[min(t + i, last)];
        double sq = sqc;
        int mq = mqc;
#pragma unroll
        for (int i = 0; i < kGroup; i++) plane(g, pr, q[i], sq, mq);
        sqc = sq; mqc = mq;
        tcur = t + kGroup;
        if (tcur < tend) { segout = false; mo = -1; ds = (R)0; rh = (R)0; return true; }
        if (++steps > SG_MAX_STEPS) { done = true; return false; }
        if (mq == NO_INDEX) {
            // no plane ahead: move on by eps and locate the point again
            nstuck++;
            x += kx * g.eps; y += ky * g.eps; z += kz * g.eps;
            mr = voro_locate(g, x, y, z, mr);
            if (mr < 0) { done = true; return false; }
            enter(g);
            segout = false; mo = -1; ds = (R)0; rh = (R)0;
            return true;
        }
        const int segm = mr;
        x += (sq + g.eps) * kx; y += (sq + g.eps) * ky; z += (sq + g.eps) * kz;
        mr = mq;
        if (mr < 0) done = true;
        else enter(g);
        mo = segm; ds = (R)sq; rh = (R)pr.rho;
        lm = segm;
        return true;
    }
    // The next segment (the reference's DustGridPath, every segment the cell loop returns).
    __device__ __forceinline__ bool next(const View& g, const int32_t* lds, int nlds, int& mo, R& ds, R& rh)
    {
        while (step(g, lds, nlds, mo, ds, rh))
            if (segout) return true;
        return false;
    }
};

} // namespace sg
