// sg_host.hpp -- host-side objects behind the C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "sg_common.hpp"

namespace sg {

void set_error(const std::string& msg);
double lngamma(double a);
// Normalisation and spiral constants of an exponential-disc geometry record (rho0, tan(pitch), C_N).
void disc_density_constants(const double* geom16, double* rho0, double* tanp, double* CN);
struct Fail {
    int code;
};
#define SG_HIP(expr)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            sg::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                        \
            throw sg::Fail{-2};                                                                      \
        }                                                                                            \
    } while (0)
#define SG_REQUIRE(cond, msg)                                                                        \
    do {                                                                                             \
        if (!(cond)) {                                                                               \
            sg::set_error(msg);                                                                      \
            throw sg::Fail{-1};                                                                      \
        }                                                                                            \
    } while (0)

// device buffer owned by a host object
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() {}
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t n)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = n;
        if (n) SG_HIP(hipMalloc(&p, n));
    }
    template <class T> void upload(const T* h, size_t count)
    {
        alloc(count * sizeof(T));
        if (count) SG_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
    }
    template <class T> void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
    void zero()
    {
        if (bytes) SG_HIP(hipMemset(p, 0, bytes));
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

enum GridKind { GK_CART = 1, GK_TREE = 2, GK_VORO = 3 };

// Breadth-first tree topology rebuilt on the host from the subdivision flags, with the
// reference's numbering (TreeDustGrid.cpp:83-121).
struct HostTree {
    int arity = 8;
    double ext[6];
    double eps;
    std::vector<int32_t> child0;       // first child id, -1 for leaves
    std::vector<int8_t> dir;           // binary tree split axis
    std::vector<int32_t> father;
    std::vector<int16_t> level;
    std::vector<double> bounds;        // 6 per node
    std::vector<int32_t> cellnumber;   // per node, -1 for internal nodes
    std::vector<int32_t> idv;          // cell -> node
    std::vector<std::vector<int32_t>> nb;   // 6 lists per node (node ids), built on request
    bool build(int arity, const double ext[6], const uint8_t* flags, int64_t nnodes);
    void build_neighbors();
    // Ropes of every node (6 per node, walls in the order -x +x -y +y -z +z): the node across the
    // wall at the same level, or the coarser leaf covering the whole wall, or -1 on the grid boundary.
    void build_ropes(std::vector<int32_t>& rope) const;
};

} // namespace sg

struct sg_grid {
    sg::GridKind kind;
    int64_t ncells = 0;
    // Cartesian
    sg::DevBuf xv, yv, zv, xvf, yvf, zvf, rho, rhof;
    sg::CartView cart{};
    // tree
    sg::HostTree ht;
    sg::DevBuf node, rec64, node32, cellof, kpack, nboff, nbids, bkb, bkchild0, bkfather, bkcell;
    std::vector<sg::TreeRec64> hrec64;
    std::vector<sg::TreeNode32> hnode32;
    sg::TreeView tree{};
    // Voronoi
    sg::DevBuf site, voff, vnb, boff, bids, brep;
    std::vector<sg::VoroSite> hsite;
    std::vector<sg::VoroNbr> hvnb;     // device list records (headers and neighbour entries)
    sg::VoroView voro{};
    void set_rho(const double* r);
};

struct sg_dust {
    sg_grid* grid;
    int nlambda;
    std::vector<double> kext, albedo, g;
    bool store_abs;
    sg::DevBuf labs;   // [ell][m]
    // several dust components: each one's density ([h][m]) and mix per wavelength ([h][ell]); the grid
    // then holds the combined kappa*rho of the wavelength `prepared` and albm each cell's albedo there
    int ncomp = 1;
    std::vector<double> rhoc, kextc, kscac, gc;
    sg::DevBuf drhoc, albm;
    int prepared = -1;
};

struct sg_stars {
    int nlambda;
    std::vector<double> params;   // ncomp * 16
    std::vector<double> lum;      // ncomp * nlambda
    std::vector<sg::DevBuf*> tables;
    std::vector<std::vector<double>> ts, tM;
    double bias;
    ~sg_stars() { for (auto* t : tables) delete t; }
};

struct sg_instruments {
    int n, nlambda;
    std::vector<sg::InstrDev> dev;
    std::vector<sg::DevBuf*> frames;
    ~sg_instruments() { for (auto* f : frames) delete f; }
};

struct sg_sim {
    sg_dust* dust;
    sg_stars* stars;
    sg_instruments* instr;
    double mc[8];
    sg::DevBuf counters, next, err, scene, cold, args, emitted, emitdraw;
    // work buffers of a batched run, held by the batch's first member: range counters and the
    // tables of grid views and run arguments
    sg::DevBuf bnext, btab;

};
