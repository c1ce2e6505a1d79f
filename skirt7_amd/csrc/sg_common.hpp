// sg_common.hpp -- device-side data views and small numerics shared by all kernels.
//
// Layouts in HBM (all read-only during transport except the tallies):
//   Cartesian : three edge arrays (FP64 and an FP32 copy) and rho[m], m = k + Nz*j + Nz*Ny*i.
//   Tree      : node words in breadth-first id order (>= 0 first child id, < 0 leaf: -(m+1));
//               one 64-byte FP64 record per cell (bounds, rho) with the full neighbour lists per
//               (cell, wall) in CSR form; one 32-byte mixed-precision record per node (subdivision
//               counts, density or first child, six ropes or split planes).
//   Voronoi   : one 32-byte record per site (x, y, z, rho); per cell a 32-byte header (its site and
//               density) followed by its neighbour entries (walls -1..-6), each a 32-byte record
//               holding the neighbour's site, the cells' blocks in Morton order of their sites; a
//               uniform bucket grid over the sites with one start site per bucket for cellIndex.
//   Tallies   : FP64, instrument frames [ell][pixel], absorption [ell][m].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#define SG_MAX_INSTR 8
#define SG_MAX_STARS 4
#define SG_MAX_DUST 8
#define SG_LDS_NODES 4096
#define SG_MAX_STEPS 4000000

namespace sg {

enum { WALL_BACK = 0, WALL_FRONT, WALL_LEFT, WALL_RIGHT, WALL_BOTTOM, WALL_TOP };

// ------------------------------------------------------------------------- grid views

struct CartView {
    const double* xv; const double* yv; const double* zv;
    const float* xvf; const float* yvf; const float* zvf;
    const double* rho; const float* rhof;
    int Nx, Ny, Nz;
    double xmin, ymin, zmin, xmax, ymax, zmax;
    float wminf;                     // smallest cell width over the three axes
};

// Mixed-precision walk record, one per tree NODE (device id), 32 bytes, read as two 16-byte loads:
//   code : subdivisions of the root along x, y, z (5 bits each: nx | ny << 5 | nz << 10), leaf flag
//          (bit 15) and the split axis of an internal binary-tree node (bits 16-17);
//   val  : a leaf's density (float bits) or an internal node's first child;
//   rope : for a leaf, the node across each wall that covers the whole wall (same level or coarser
//          along every axis; -1 on the grid boundary): the next cell is found by descending from it.
//          For an internal node the first words hold its split planes as floats (x, y, z for an
//          octree node, the plane of its split axis for a binary node), so a descent compares without
//          computing them.
// Node bounds are not stored: along axis b a node with integer index k_b spans
// ext_b + [k_b, k_b + 1] * width_b / 2^n_b, each wall one fused multiply-add (a single rounding of
// the exact value, so that neighbouring leaves share their wall floats exactly and a parent's split
// plane equals its children's common wall); the walker carries the current leaf's k_b. A node's
// indices are also kept packed (21 bits per axis) for the walks that start in a given leaf.
constexpr uint32_t kTreeLeaf = 1u << 15;
struct alignas(32) TreeNode32 {
    uint32_t code;
    uint32_t val;
    int32_t rope[6];
};
struct alignas(16) TreeRec64 {       // 64 bytes
    double lo[3], hi[3];
    double rho;
    int32_t node;
    int32_t pad;
};

struct TreeView {
    const int32_t* node;             // breadth-first node words
    const TreeRec64* rec64;
    const TreeNode32* node32;        // mixed-precision walk records
    const int32_t* cellof;           // cell number per node (-1 for internal nodes)
    const uint64_t* kpack;           // integer indices per node, i | j << 21 | k << 42
    // Bookkeeping search (octrees): per node in the reference's breadth-first order, its box (6 doubles),
    // first child (-1 for a leaf), father (-1 for the root) and cell number (-1 for an internal node)
    const double* bkb;
    const int32_t* bkchild0;
    const int32_t* bkfather;
    const int32_t* bkcell;
    const int32_t* nboff;            // CSR offsets, index cell*6 + wall
    const int32_t* nbids;            // neighbour cell numbers
    int32_t nnodes, ncells;
    int arity;                       // 8 octree, 2 alternating binary tree
    int search;                      // 0 TopDown, 1 Neighbor, 2 Bookkeeping
    double ext[6];
    float extf[6];
    float widthf[3];                 // root box widths
    double eps;
};

struct alignas(16) VoroSite { double x, y, z, rho; };
// One entry of a cell's neighbour list with the neighbour's site inline (walls: id -1..-6, no site),
// so that a walk step reads its cell's list as one contiguous stream instead of gathering sites.
struct alignas(32) VoroNbr { double x, y, z; int32_t id, pad; };

struct VoroView {
    const VoroSite* site;
    const int2* off;                 // per cell: its header record in nbr, and its entry count
    const VoroNbr* nbr;              // per cell a header (site, density) then its entries, walls -1..-6
    const int32_t* boff;             // bucket offsets, nbk^3 + 1
    const int32_t* bids;             // site ids per bucket
    const int32_t* brep;             // start site per bucket for the greedy location, nbk^3
    int32_t nsites;
    int nbk;
    double ext[6];
    double eps;
};

// ------------------------------------------------------------------------- scene records

struct InstrDev {
    double kobs[3];
    double costheta, sintheta, cosphi, sinphi, cospa, sinpa;
    double xpmin, xpsiz, ypmin, ypsiz;
    int Nx, Ny;
    double* frame;                   // [nlambda][Ny*Nx]
};

struct StarDev {
    int kind;                        // 1 exponential disc, 2 Sersic
    int spiral;
    double p[6];                     // disc: hR hz Rmax zmax Rmin | Sersic: n reff
    double sp[6];                    // arms pitch R0 phi0 weight index
    double tanp, CN;
    const double* ts;                // Sersic table radii (101)
    const double* tM;                // Sersic table masses (101)
    int nt;
    double lum;                      // luminosity at the run's wavelength
};

// Per-run tables indexed per lane (component, instrument): kept in device memory rather than in
// the kernel arguments, where per-lane indexing would pull whole arrays into registers.
struct SceneDev {
    InstrDev instr[SG_MAX_INSTR];
    StarDev star[SG_MAX_STARS];
    double Xcum[SG_MAX_STARS + 1];   // normalised cumulative luminosity
    // dust components at the run's wavelength (used when there are two or more): scattering
    // coefficient and asymmetry parameter of each component's mix
    double ksca[SG_MAX_DUST], gc[SG_MAX_DUST];
};

struct McArgs {
    uint64_t first, count;
    double Lpkt, Lthreshold;
    int ell, nlambda;
    double kext, albedo, g;
    int minfs;
    double xi;
    uint32_t seed, member;
    double tau_cut;
    int store_abs;
    double* labs;                    // [ell][m] when stored
    int64_t ncells;
    int has_dust;
    int ni;
    int ns;
    const struct SceneDev* scene;    // instruments and stellar components (device memory)
    double emission_bias;
    double Lmean;
    unsigned long long* counters;    // SG_NCOUNTERS
    unsigned long long* next_packet; // work counter
    int* error;
    int compact;
    // launch states written by the emission kernel for packets [first, first+count):
    // position, direction and luminosity as 7 arrays of `count` values of the run's real type,
    // and the number of random draws the emission consumed
    const void* emitted;
    const uint32_t* emit_draw;
    // coherence keys of walk requests (event/walk pipeline): start points are binned on a raster of
    // the grid box, t = (r - key_lo) * key_scale in [0, 16)
    double key_lo[3], key_scale[3];
    // several dust components (DustSystem with Ncomp > 1): the walks see the combined kappa*rho per cell
    // (kext is then 1); rhoc holds each component's density ([h][m], ncells apart) for the peel-off
    // weights and the choice of the scattering mix, albm each cell's albedo at the run's wavelength
    int ncomp;
    const double* rhoc;
    const double* albm;
};

// ------------------------------------------------------------------------- Philox4x32-10

__device__ __forceinline__ void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                         uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
}

// Per-packet stream: key (seed, member), counter (packet lo, packet hi, draw/4, ell); each
// Philox block yields four consecutive draws.
struct PacketRng {
    uint32_t k0, k1, plo, phi, ell, draw;
    uint32_t w0, w1, w2, w3;
    __device__ void start(uint32_t seed, uint32_t member, uint64_t pkt, int l)
    {
        k0 = seed; k1 = member; plo = (uint32_t)pkt; phi = (uint32_t)(pkt >> 32); ell = (uint32_t)l; draw = 0;
    }
    __device__ __forceinline__ uint32_t word()
    {
        uint32_t q = draw & 3u;
        if (q == 0) {
            w0 = plo; w1 = phi; w2 = draw >> 2; w3 = ell;
            philox10(w0, w1, w2, w3, k0, k1);
        }
        draw++;
        return q == 0 ? w0 : q == 1 ? w1 : q == 2 ? w2 : w3;
    }
    __device__ __forceinline__ double ud() { return ((double)word() + 0.5) * 2.3283064365386962890625e-10; }
    __device__ __forceinline__ float uf() { return ((float)(word() >> 9) + 0.5f) * 1.1920928955078125e-07f; }
};

// ------------------------------------------------------------------------- small numerics

__device__ __forceinline__ bool box_contains(const double* b, double x, double y, double z)
{
    return x >= b[0] && x <= b[3] && y >= b[1] && y <= b[4] && z >= b[2] && z <= b[5];
}

__device__ __forceinline__ double interp_linlin(double x, double x1, double x2, double f1, double f2)
{
    return f1 + ((x - x1) / (x2 - x1)) * (f2 - f1);
}
__device__ __forceinline__ float interp_linlin(float x, float x1, float x2, float f1, float f2)
{
    return f1 + ((x - x1) / (x2 - x1)) * (f2 - f1);
}

// Binary search of NR::locate_clip semantics (x below the first border -> bin 0).
template <typename T> __device__ __forceinline__ int locate_clip(const T* xv, int size, T x)
{
    if (x < xv[0]) return 0;
    int jl = -1, ju = size - 1;
    while (ju - jl > 1) {
        int jm = (ju + jl) >> 1;
        if (x < xv[jm]) ju = jm;
        else jl = jm;
    }
    return jl;
}

} // namespace sg
