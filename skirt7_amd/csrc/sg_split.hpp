// sg_split.hpp -- the photon-packet life cycle (MonteCarloSimulation::dostellaremissionchunk,
// MonteCarloSimulation.cpp:259-297) as two alternating kernels over device-resident queues:
//
//   k_events: one thread per in-flight packet slot runs the packet's state machine up to its next
//             walk -- launch, emission peel-offs, escape/absorption and the stop test, biased
//             propagation, scattering peel-offs and the new direction -- and appends the walks it
//             needs (peel-off, own path, propagation re-walk) to the walk queue;
//   k_walks:  persistent lanes take walk requests from the queue and run the cell walks, with the
//             same ballot-compacted refill as the lane pipeline; a peel-off walk ends with its
//             detection (FrameInstrument::detect), an own-path or propagation walk with its result
//             written back to the packet slot.
//
// The walk kernel carries only the walker and the running sums, so it keeps several times more
// waves in flight than a kernel that also holds the event logic; the event kernel's FP64 and libm
// work runs at full width over the slots. Each packet still draws from its own Philox stream in the
// same order, and every walk repeats the same arithmetic, so results equal the lane pipeline's up
// to the order of the tally additions.
#pragma once
#include "sg_mc.hpp"

namespace sg {

enum SlotR { SR_RX = 0, SR_RY, SR_RZ, SR_KX, SR_KY, SR_KZ, SR_L, SR_RES, SR_NR };
// The last four slot words count the slot's packets, scatterings, detected peel-offs and started
// walks over a run; they are summed once at the end (no per-pass atomics).
// SU_HINT: the leaf (tree node, or Voronoi cell) the packet's position lies in, as the last
// propagation walk left it (-1: unknown), handed to the walks that start there.
enum SlotU { SU_PLO = 0, SU_PHI, SU_DRAW, SU_W0, SU_W1, SU_W2, SU_W3, SU_NSCATT, SU_HINT,
             SU_STATE, SU_CPK, SU_CSCAT, SU_CPEEL, SU_CWALK, SU_NU };
enum PacketState : uint32_t { PS_FRESH = 0, PS_OWN, PS_PROP, PS_DEAD };
enum WalkKind : uint32_t { WK_PEEL = 0, WK_OWN = 1, WK_PROP = 2 };

// Packet slots: structure of arrays, n slots.
template <typename R> struct Slots {
    R* r;
    uint32_t* u;
    int64_t n;
    __device__ __forceinline__ R& fr(int f, int64_t i) const { return r[f * n + i]; }
    __device__ __forceinline__ uint32_t& fu(int f, int64_t i) const { return u[f * n + i]; }
};
template <typename R> __host__ __device__ constexpr size_t slot_bytes(int64_t n)
{
    return (size_t)n * (SR_NR * sizeof(R) + SU_NU * sizeof(uint32_t));
}

// Walk requests: origin (3), direction (3, unused for peel-offs: the instrument's direction is
// taken from its record) and a value (peel-off luminosity, own-path luminosity, or propagation
// target optical depth) in the run's real type; a tag (kind | instrument << 2), an index (packet
// slot, or pixel for a peel-off) and the start leaf hint of the walker. In the LDS staging area
// they are structure-of-arrays (QR_*, QU_* fields); in the queues one record per request, so that
// a request moved to its sorted place, or read by a lane, is one contiguous block.
enum QueueR { QR_OX = 0, QR_OY, QR_OZ, QR_DX, QR_DY, QR_DZ, QR_VAL, QR_NR };
enum QueueU { QU_TAG = 0, QU_IDX, QU_HINT, QU_NU };
template <typename R> struct alignas(16) WalkReq {
    R o[3], d[3], val;
    uint32_t tag, idx, hint;
};
static_assert(sizeof(WalkReq<float>) == 48 && sizeof(WalkReq<double>) == 80, "walk request records");

// Walk order. The walks of a pass are sorted by a coherence key before the walk kernel takes them:
// peel-offs by instrument and the Morton order of their pixel (rays toward one instrument through
// neighbouring pixels cross the same cells), own-path and propagation walks by the Morton order of
// their start point. Lanes and waves that run at the same time then touch a small part of the grid,
// which the L2 cache of their XCD holds. A counting sort: the event kernel counts each request into
// its key's bin and keeps its rank there, a one-block scan turns the counts into bin offsets (and
// clears them), and a scatter moves each request to offset + rank. Histories do not depend on the
// order (only the order of the tally additions changes).
constexpr int kKeyBits = 16;
constexpr uint32_t kKeyBins = 1u << kKeyBits;
template <typename R> struct WalkQueue {
    WalkReq<R>* req;            // requests in the order the event kernel appended them
    WalkReq<R>* sorted;         // the same in key order (null: the walk kernel reads req)
    uint32_t* key;              // per request: its bin, and its rank within the bin
    uint32_t* rank;
    uint32_t* hist;             // kKeyBins counts (zero between passes; a buffer of its own, cleared once)
    uint32_t* off;              // kKeyBins bin offsets
    int64_t cap;
    unsigned long long* tail;   // requests appended by the event kernel
};
// requests, sorted requests, keys and ranks (the bin counts and offsets are held apart)
template <typename R> __host__ __device__ constexpr size_t queue_bytes(int64_t cap, bool sort)
{
    return (size_t)cap * sizeof(WalkReq<R>) * (sort ? 2 : 1) + (sort ? (size_t)cap * 8 : 0);
}

// 2-D and 3-D Morton codes (bit interleaving) of small integer coordinates.
__device__ __forceinline__ uint32_t spread2(uint32_t v)   // 8 bits -> every second bit
{
    v &= 0xffu;
    v = (v | (v << 4)) & 0x0f0fu;
    v = (v | (v << 2)) & 0x3333u;
    v = (v | (v << 1)) & 0x5555u;
    return v;
}
__device__ __forceinline__ uint32_t spread3(uint32_t v)   // 5 bits -> every third bit
{
    v &= 0x1fu;
    v = (v | (v << 8)) & 0x100fu;
    v = (v | (v << 4)) & 0x10c3u;
    v = (v | (v << 2)) & 0x1249u;
    return v;
}
// The coherence key of a request: 4 bits of class (peel-off instrument 0-7, own path 8, propagation
// 9) above 12 bits of Morton order (a peel-off's pixel on a 64 x 64 raster of its frame; a start
// point on a 16 x 16 x 16 raster of the grid box).
template <typename R>
__device__ __forceinline__ uint32_t walk_key(const McArgs& a, const InstrDev* instr, uint32_t tag, uint32_t idx, R ox,
                                             R oy, R oz)
{
    const uint32_t kind = tag & 3u;
    if (kind == WK_PEEL) {
        const InstrDev& in = instr[tag >> 2];
        const uint32_t i = idx % (uint32_t)in.Nx, j = idx / (uint32_t)in.Nx;
        const uint32_t qi = (uint32_t)(((uint64_t)i << 6) / (uint32_t)in.Nx);
        const uint32_t qj = (uint32_t)(((uint64_t)j << 6) / (uint32_t)in.Ny);
        return ((tag >> 2) << 12) | spread2(qi) | (spread2(qj) << 1);
    }
    auto q = [](double v, double lo, double sc) {
        const double t = (v - lo) * sc;
        return (uint32_t)(!(t > 0.0) ? 0.0 : t > 15.0 ? 15.0 : t);   // NaN to 0
    };
    const uint32_t qx = q((double)ox, a.key_lo[0], a.key_scale[0]);
    const uint32_t qy = q((double)oy, a.key_lo[1], a.key_scale[1]);
    const uint32_t qz = q((double)oz, a.key_lo[2], a.key_scale[2]);
    return ((kind == WK_OWN ? 8u : 9u) << 12) | spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
}

// Appends one request per active lane (one atomic per wave); returns this lane's queue position.
__device__ __forceinline__ uint64_t wave_append(unsigned long long* tail)
{
    uint64_t mask = __ballot(1);
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(tail, (unsigned long long)__popcll(mask));
    uint32_t lo = __shfl((uint32_t)base, leader), hi = __shfl((uint32_t)(base >> 32), leader);
    base = ((uint64_t)hi << 32) | lo;
    uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    return base + (uint64_t)__popcll(below);
}

// Walk requests of one workgroup staged in LDS (structure of arrays) in two parts: peel-offs (origin,
// value, tag, index, hint: a peel-off walks toward its instrument, so it carries no direction) in capp
// entries, and the own-path and propagation walks (with their direction) in capw entries, each part with
// its own count. A lane stages at most ni peel-offs and one other walk per pass, so capp = ni and capw = 1
// entries per lane suffice, 28 + 40 / ni bytes per request slot in FP32 instead of 40, which leaves room
// for more workgroups per CU. The workgroup appends them to the walk queue as one block with a single
// global atomic.
enum StagedPeelR { QP_OX = 0, QP_OY, QP_OZ, QP_VAL, QP_NR };
template <typename R> struct StagedQueue {
    R* pr;                      // peel-offs: QP_NR reals and QU_NU words per entry
    uint32_t* pu;
    int capp;
    unsigned* cntp;
    R* wr;                      // other walks: QR_NR reals and QU_NU words per entry
    uint32_t* wu;
    int capw;
    unsigned* cntw;
    __device__ __forceinline__ void push(uint32_t kind, uint32_t inst, uint32_t idx, R ox, R oy, R oz, R dx, R dy,
                                         R dz, R val, int32_t hint) const
    {
        const bool peel = kind == WK_PEEL;   // the same for every lane of a call
        uint64_t mask = __ballot(1);
        int lane = __lane_id();
        int leader = __ffsll((unsigned long long)mask) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(peel ? cntp : cntw, (unsigned)__popcll(mask));
        base = __shfl(base, leader);
        uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
        const int i = (int)(base + (unsigned)__popcll(below));
        if (peel) {
            pr[QP_OX * capp + i] = ox; pr[QP_OY * capp + i] = oy; pr[QP_OZ * capp + i] = oz;
            pr[QP_VAL * capp + i] = val;
            pu[QU_TAG * capp + i] = kind | (inst << 2);
            pu[QU_IDX * capp + i] = idx;
            pu[QU_HINT * capp + i] = (uint32_t)hint;
        } else {
            wr[QR_OX * capw + i] = ox; wr[QR_OY * capw + i] = oy; wr[QR_OZ * capw + i] = oz;
            wr[QR_DX * capw + i] = dx; wr[QR_DY * capw + i] = dy; wr[QR_DZ * capw + i] = dz;
            wr[QR_VAL * capw + i] = val;
            wu[QU_TAG * capw + i] = kind | (inst << 2);
            wu[QU_IDX * capw + i] = idx;
            wu[QU_HINT * capw + i] = (uint32_t)hint;
        }
    }
    // Request j of the workgroup's block: the peel-offs first, then the other walks.
    __device__ __forceinline__ WalkReq<R> get(unsigned j, unsigned np) const
    {
        WalkReq<R> w;
        if (j < np) {
            w.o[0] = pr[QP_OX * capp + j]; w.o[1] = pr[QP_OY * capp + j]; w.o[2] = pr[QP_OZ * capp + j];
            w.d[0] = (R)0; w.d[1] = (R)0; w.d[2] = (R)0;
            w.val = pr[QP_VAL * capp + j];
            w.tag = pu[QU_TAG * capp + j]; w.idx = pu[QU_IDX * capp + j]; w.hint = pu[QU_HINT * capp + j];
        } else {
            const unsigned k = j - np;
            w.o[0] = wr[QR_OX * capw + k]; w.o[1] = wr[QR_OY * capw + k]; w.o[2] = wr[QR_OZ * capw + k];
            w.d[0] = wr[QR_DX * capw + k]; w.d[1] = wr[QR_DY * capw + k]; w.d[2] = wr[QR_DZ * capw + k];
            w.val = wr[QR_VAL * capw + k];
            w.tag = wu[QU_TAG * capw + k]; w.idx = wu[QU_IDX * capw + k]; w.hint = wu[QU_HINT * capw + k];
        }
        return w;
    }
};
// LDS of a workgroup's staging area for block lanes and ni instruments (the peel-off part first; its
// size is a multiple of 8 bytes, so that the FP64 reals of the other part stay aligned)
template <typename R> __host__ __device__ constexpr size_t staged_peel_bytes(int block, int ni)
{
    return (((size_t)block * ni * (QP_NR * sizeof(R) + QU_NU * sizeof(uint32_t))) + 7) / 8 * 8;
}
template <typename R> __host__ __device__ constexpr size_t staged_bytes(int block, int ni)
{
    return staged_peel_bytes<R>(block, ni) + (size_t)block * (QR_NR * sizeof(R) + QU_NU * sizeof(uint32_t));
}

// Peel-offs toward every instrument (MonteCarloSimulation::peeloffemission/peeloffscattering,
// MonteCarloSimulation.cpp:299-357): pixel, phase-function weight (scattering only), and either the
// walk request or, without a dust system, the detection itself.
template <typename R, class Q>
__device__ __forceinline__ uint32_t peel_offs(const McArgs& a, const InstrDev* instr, const Q& q,
                                              bool scattering, R rx, R ry, R rz, R kx, R ky, R kz, R L,
                                              int32_t hint)
{
    const R gg = (R)a.g;
    uint32_t n = 0;
    for (int i = 0; i < a.ni; i++) {
        const InstrDev& in = instr[i];
        int pix = pixel_on_detector<R>(in, rx, ry, rz);
        if (pix < 0) continue;
        R Lp = L;
        if (scattering) {
            // DustMix::phaseFunctionValue, Henyey-Greenstein (DustMix.cpp:648-653)
            R cosalpha = kx * (R)in.kobs[0] + ky * (R)in.kobs[1] + kz * (R)in.kobs[2];
            R t = (R)1.0 + gg * gg - (R)2 * gg * cosalpha;
            Lp = L * (((R)1.0 - gg) * ((R)1.0 + gg) / sqrt(t * t * t));
        }
        n++;
        if (!a.has_dust) unsafeAtomicAdd(in.frame + (size_t)a.ell * in.Nx * in.Ny + pix, (double)Lp);
        else q.push(WK_PEEL, (uint32_t)i, (uint32_t)pix, rx, ry, rz, (R)0, (R)0, (R)0, Lp, hint);
    }
    return n;
}

// Henyey-Greenstein scattering direction (DustMix.cpp:596-601) and Random::direction(k, cos)
// (Random.cpp:189-223).
template <typename R>
__device__ __forceinline__ void hg_scatter(PacketRng& rng, R gg, R& kx, R& ky, R& kz)
{
    R nkx, nky, nkz;
    if (fabs((double)gg) < 1e-6) {
        double a0, a1, a2;
        iso_direction(rng, a0, a1, a2);
        nkx = (R)a0; nky = (R)a1; nkz = (R)a2;
    } else {
        R f = (((R)1.0 - gg) * ((R)1.0 + gg)) / ((R)1.0 - gg + (R)2.0 * gg * uni<R>(rng));
        R costheta = ((R)1.0 + gg * gg - f * f) / ((R)2.0 * gg);
        R phi = (R)(2.0 * M_PI) * uni<R>(rng);
        R cosphi = cos(phi), sinphi = sin(phi);
        R sintheta = sqrt(fabs(((R)1.0 - costheta) * ((R)1.0 + costheta)));
        if (kz > (R)0.99999) { nkx = cosphi * sintheta; nky = sinphi * sintheta; nkz = costheta; }
        else if (kz < (R)-0.99999) { nkx = cosphi * sintheta; nky = sinphi * sintheta; nkz = -costheta; }
        else {
            R root = sqrt(((R)1.0 - kz) * ((R)1.0 + kz));
            nkx = sintheta / root * (-kx * kz * cosphi + ky * sinphi) + kx * costheta;
            nky = -sintheta / root * (ky * kz * cosphi + kx * sinphi) + ky * costheta;
            nkz = root * sintheta * cosphi + kz * costheta;
        }
    }
    kx = nkx; ky = nky; kz = nkz;
}

// Event pass: every slot advances its packet to its next walk (or takes a new packet). Dynamic LDS:
// the workgroup's staged requests, blockDim.x * a.ni peel-offs and blockDim.x other walks (the most one
// pass can append: a lane's pass ends with the walk it requests, after at most one set of peel-offs).
template <typename R>
__global__ void __launch_bounds__(256) k_events(const McArgs a, Slots<R> sl, WalkQueue<R> q)
{
    __shared__ InstrDev s_instr[SG_MAX_INSTR];
    __shared__ unsigned s_cntp, s_cntw;
    __shared__ unsigned long long s_base;
    extern __shared__ __align__(16) unsigned char smem[];
    const int capp = (int)blockDim.x * a.ni, capw = (int)blockDim.x;
    unsigned char* wbase = smem + staged_peel_bytes<R>((int)blockDim.x, a.ni);
    StagedQueue<R> sq{reinterpret_cast<R*>(smem), reinterpret_cast<uint32_t*>(smem + (size_t)capp * QP_NR * sizeof(R)),
                      capp, &s_cntp, reinterpret_cast<R*>(wbase),
                      reinterpret_cast<uint32_t*>(wbase + (size_t)capw * QR_NR * sizeof(R)), capw, &s_cntw};
    if (threadIdx.x < SG_MAX_INSTR) s_instr[threadIdx.x] = a.scene->instr[threadIdx.x];
    if (threadIdx.x == 0) { s_cntp = 0; s_cntw = 0; }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // every lane stays to the end: the counts below are summed over the whole wave
    uint32_t state = i < sl.n ? sl.fu(SU_STATE, i) : (uint32_t)PS_DEAD;
    uint32_t c_pk = 0, c_scat = 0, c_npeel = 0, c_walks = 0;
    if (state != PS_DEAD) {
        const R albedo = (R)a.albedo, gg = (R)a.g;
        R rx = sl.fr(SR_RX, i), ry = sl.fr(SR_RY, i), rz = sl.fr(SR_RZ, i);
        R kx = sl.fr(SR_KX, i), ky = sl.fr(SR_KY, i), kz = sl.fr(SR_KZ, i);
        R L = sl.fr(SR_L, i);
        PacketRng rng;
        rng.k0 = a.seed; rng.k1 = a.member; rng.ell = (uint32_t)a.ell;
        rng.plo = sl.fu(SU_PLO, i); rng.phi = sl.fu(SU_PHI, i); rng.draw = sl.fu(SU_DRAW, i);
        rng.w0 = sl.fu(SU_W0, i); rng.w1 = sl.fu(SU_W1, i); rng.w2 = sl.fu(SU_W2, i); rng.w3 = sl.fu(SU_W3, i);
        int nscatt = (int)sl.fu(SU_NSCATT, i);
        int32_t hint = (int32_t)sl.fu(SU_HINT, i);
        // at most a few transitions per pass: a finished packet is replaced by a new one at once; a slot
        // without a dust system detects its peel-offs directly and takes up to four packets per pass
        for (int step = 0; step < 4; step++) {
            if (state == PS_FRESH) {
                uint64_t p;
                if (!wave_claim(a.next_packet, a.count, p)) { state = PS_DEAD; break; }
                // launch state from the emission kernel; continue the packet's random stream
                const R* e = reinterpret_cast<const R*>(a.emitted);
                const size_t n = a.count;
                c_pk++;
                R Le = e[6 * n + p];
                if (!(Le > (R)0)) continue;
                rx = e[p]; ry = e[n + p]; rz = e[2 * n + p];
                kx = e[3 * n + p]; ky = e[4 * n + p]; kz = e[5 * n + p];
                L = Le;
                rng.start(a.seed, a.member, a.first + p, a.ell);
                rng.draw = a.emit_draw[p];
                if (rng.draw & 3u) {
                    rng.w0 = rng.plo; rng.w1 = rng.phi; rng.w2 = rng.draw >> 2; rng.w3 = rng.ell;
                    philox10(rng.w0, rng.w1, rng.w2, rng.w3, rng.k0, rng.k1);
                }
                nscatt = 0;
                hint = -1;   // a new position: the leaf is not known yet
                uint32_t np = peel_offs<R>(a, s_instr, sq, false, rx, ry, rz, kx, ky, kz, L, hint);
                c_npeel += np;
                if (!a.has_dust) continue;
                c_walks += np + 1;
                sq.push(WK_OWN, 0, (uint32_t)i, rx, ry, rz, kx, ky, kz, L, hint);
                state = PS_OWN;
                break;
            }
            if (state == PS_OWN) {
                // DustSystem::fillOpticalDepth check (DustSystem.cpp:952-955)
                R taupath = sl.fr(SR_RES, i);
                if (taupath < (R)0 || isnan(taupath) || isinf(taupath)) {
                    atomicOr(a.error, 1);
                    state = PS_FRESH;
                    continue;
                }
                // simulateescapeandabsorption, single component (MonteCarloSimulation.cpp:432-461)
                R expfactor = -rexpm1<R>(-taupath);
                L = L * albedo * expfactor;
                if (L <= (R)0 || (L <= (R)a.Lthreshold && nscatt >= a.minfs)) { state = PS_FRESH; continue; }
                if (taupath != (R)0) {
                    // simulatepropagation (MonteCarloSimulation.cpp:513-531): biased free path, then the
                    // re-walk to the interaction point
                    R tt;
                    if (a.xi == 0.0) tt = exponcutoff(rng, taupath);
                    else {
                        R X = uni<R>(rng);
                        tt = (X < (R)a.xi) ? uni<R>(rng) * taupath : exponcutoff(rng, taupath);
                        R pp = -rexp<R>(-tt) / rexpm1<R>(-taupath);
                        R qq = ((R)1.0 - (R)a.xi) * pp + (R)a.xi / taupath;
                        L = L * (pp / qq);
                    }
                    sq.push(WK_PROP, 0, (uint32_t)i, rx, ry, rz, kx, ky, kz, tt, hint);
                    state = PS_PROP;
                    break;
                }
                // zero optical depth along the path: scattering in place
                state = PS_PROP;
                sl.fr(SR_RES, i) = (R)0;
            }
            if (state == PS_PROP) {
                // PhotonPackage::propagate (PhotonPackage.cpp:93-96) by pathlength(tau)
                R d = sl.fr(SR_RES, i);
                rx = rx + d * kx; ry = ry + d * ky; rz = rz + d * kz;
                uint32_t np = peel_offs<R>(a, s_instr, sq, true, rx, ry, rz, kx, ky, kz, L, hint);
                c_npeel += np;
                hg_scatter<R>(rng, gg, kx, ky, kz);
                nscatt++;
                c_scat++;
                c_walks += np + 1;
                sq.push(WK_OWN, 0, (uint32_t)i, rx, ry, rz, kx, ky, kz, L, hint);
                state = PS_OWN;
                break;
            }
        }
        sl.fr(SR_RX, i) = rx; sl.fr(SR_RY, i) = ry; sl.fr(SR_RZ, i) = rz;
        sl.fr(SR_KX, i) = kx; sl.fr(SR_KY, i) = ky; sl.fr(SR_KZ, i) = kz;
        sl.fr(SR_L, i) = L;
        sl.fu(SU_PLO, i) = rng.plo; sl.fu(SU_PHI, i) = rng.phi; sl.fu(SU_DRAW, i) = rng.draw;
        sl.fu(SU_W0, i) = rng.w0; sl.fu(SU_W1, i) = rng.w1; sl.fu(SU_W2, i) = rng.w2; sl.fu(SU_W3, i) = rng.w3;
        sl.fu(SU_NSCATT, i) = (uint32_t)nscatt;
        sl.fu(SU_HINT, i) = (uint32_t)hint;
        sl.fu(SU_CPK, i) += c_pk; sl.fu(SU_CSCAT, i) += c_scat; sl.fu(SU_CPEEL, i) += c_npeel;
        sl.fu(SU_CWALK, i) += c_walks;
    }
    if (i < sl.n) sl.fu(SU_STATE, i) = state;
    // append the workgroup's requests to the walk queue
    __syncthreads();
    const unsigned np = s_cntp, n = s_cntp + s_cntw;
    if (threadIdx.x == 0) s_base = n ? atomicAdd(q.tail, (unsigned long long)n) : 0ull;
    __syncthreads();
    const uint64_t base = s_base;
    for (unsigned j = threadIdx.x; j < n; j += blockDim.x) {
        const WalkReq<R> w = sq.get(j, np);
        q.req[base + j] = w;
        if (q.sorted) {
            const uint32_t k = walk_key<R>(a, s_instr, w.tag, w.idx, w.o[0], w.o[1], w.o[2]);
            q.key[base + j] = k;
            q.rank[base + j] = atomicAdd(q.hist + k, 1u);
        }
    }
}

// Bin offsets of the walk order: exclusive prefix sum of the bin counts, which are cleared for the
// next pass. One block of 1024 threads; each thread sums kKeyBins / 1024 consecutive bins.
__global__ void __launch_bounds__(1024) k_key_scan(uint32_t* hist, uint32_t* off)
{
    constexpr int per = (int)(kKeyBins / 1024);
    __shared__ uint32_t s_wave[16];
    const int t = threadIdx.x;
    uint32_t v[per];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < per; k++) { v[k] = hist[t * per + k]; sum += v[k]; }
    // inclusive scan of the per-thread sums: within each wave, then across the 16 waves
    uint32_t x = sum;
    const int lane = __lane_id(), w = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    if (t < 64) {
        uint32_t z = t < 16 ? s_wave[t] : 0u;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const uint32_t y = __shfl_up(z, o);
            if (lane >= o) z += y;
        }
        if (t < 16) s_wave[t] = z;
    }
    __syncthreads();
    uint32_t run = x - sum + (w ? s_wave[w - 1] : 0u);
#pragma unroll
    for (int k = 0; k < per; k++) {
        off[t * per + k] = run;
        run += v[k];
        hist[t * per + k] = 0u;
    }
}

// Moves each request of the pass to its place in the walk order. A place beyond the pass's requests
// (bin counts that were not zero when the pass began) is reported through the error flag (bit 2)
// and not written.
template <typename R>
__global__ void __launch_bounds__(256) k_key_scatter(WalkQueue<R> q, int* error)
{
    const uint64_t n = *q.tail;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = q.key[j];
        const uint64_t pos = (uint64_t)q.off[k < kKeyBins ? k : 0] + q.rank[j];
        if (k < kKeyBins && pos < n) q.sorted[pos] = q.req[j];
        else atomicOr(error, 2);
    }
}

// End of a run: the slots' packet, scattering, peel-off and walk counts into the run counters.
__global__ void __launch_bounds__(256) k_slot_counts(const uint32_t* u, int64_t n, unsigned long long* counters)
{
    __shared__ unsigned long long part[4][4];
    unsigned long long v[4] = {0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        v[0] += u[SU_CPK * n + i]; v[1] += u[SU_CSCAT * n + i]; v[2] += u[SU_CPEEL * n + i];
        v[3] += u[SU_CWALK * n + i];
    }
    for (int k = 0; k < 4; k++)
        for (int off = 32; off > 0; off >>= 1) {
            uint32_t lo = __shfl_down((uint32_t)v[k], off), hi = __shfl_down((uint32_t)(v[k] >> 32), off);
            v[k] += ((unsigned long long)hi << 32) | lo;
        }
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0)
        for (int k = 0; k < 4; k++) part[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int dst[4] = {0, 3, 4, 7};
        for (int k = 0; k < 4; k++) {
            unsigned long long t = 0;
            for (int j = 0; j < (int)(blockDim.x >> 6); j++) t += part[k][j];
            if (t) atomicAdd(counters + dst[k], t);
        }
    }
}

// Request claims of one wave: chunks of kClaimChunk requests per global atomic, handed out in order to
// the lanes that ask. Called by the whole wave; cur/end is the wave's current chunk (equal in every
// lane). Returns whether this lane got request r.
constexpr uint64_t kClaimChunk = 128;
__device__ __forceinline__ bool chunk_claim(unsigned long long* ctr, uint64_t nq, bool want, uint64_t& cur,
                                            uint64_t& end, uint64_t& r)
{
    const uint64_t mask = __ballot(want);
    const uint64_t k = (uint64_t)__popcll(mask);
    if (k == 0) return false;
    const int lane = __lane_id();
    const uint64_t rank = (uint64_t)__popcll(mask & (lane ? (~0ull >> (64 - lane)) : 0ull));
    const uint64_t avail = end - cur;
    uint64_t nb = 0, ne = 0;
    if (avail < k) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(ctr, (unsigned long long)kClaimChunk);
        const uint32_t lo = __shfl((uint32_t)b, 0), hi = __shfl((uint32_t)(b >> 32), 0);
        nb = ((uint64_t)hi << 32) | lo;
        ne = nb >= nq ? nb : min(nb + kClaimChunk, nq);
    }
    bool ok;
    if (rank < avail) { r = cur + rank; ok = true; }
    else { r = nb + (rank - avail); ok = r < ne; }
    if (avail >= k) cur += k;
    else { cur = min(nb + (k - avail), ne); end = ne; }
    return ok && want;
}

// Request claims by XCD segment: the requests [0, nq) of a pass, in walk order, are cut into kXcdSegs
// contiguous segments; a wave takes chunks from the segment of its workgroup's XCD (workgroups are dispatched
// to the XCDs in turn, so blockIdx.x mod 8 names the XCD) and moves on to the following segments once that one
// is exhausted. Walks that are neighbours in the walk order (start points or pixels close together) thus meet
// in the same XCD's L2. ctr holds one counter per segment. Lanes the wave's chunk does not serve stay idle
// until the next claim; exhausted (wave-uniform) becomes true when every segment is exhausted.
constexpr int kXcdSegs = 8;
__device__ __forceinline__ bool seg_claim(unsigned long long* ctr, uint64_t nq, bool want, int& seg, int& tried,
                                          uint64_t& cur, uint64_t& end, uint64_t& r, bool& exhausted)
{
    const uint64_t mask = __ballot(want);
    const uint64_t k = (uint64_t)__popcll(mask);
    if (k == 0 || exhausted) return false;
    const int lane = __lane_id();
    const uint64_t rank = (uint64_t)__popcll(mask & (lane ? (~0ull >> (64 - lane)) : 0ull));
    if (end == cur) {
        while (tried < kXcdSegs) {
            const uint64_t s0 = nq * (uint64_t)seg / kXcdSegs, s1 = nq * (uint64_t)(seg + 1) / kXcdSegs;
            unsigned long long b = 0;
            if (lane == 0) b = atomicAdd(ctr + seg, (unsigned long long)kClaimChunk);
            const uint32_t lo = __shfl((uint32_t)b, 0), hi = __shfl((uint32_t)(b >> 32), 0);
            const uint64_t nb = s0 + (((uint64_t)hi << 32) | lo);
            if (nb < s1) {
                cur = nb;
                end = min(nb + kClaimChunk, s1);
                break;
            }
            seg = (seg + 1) % kXcdSegs;
            tried++;
        }
        if (end == cur) {
            exhausted = true;
            return false;
        }
    }
    const uint64_t take = min(k, end - cur);
    const bool ok = rank < take;
    if (ok) r = cur + rank;
    cur += take;
    return ok && want;
}

// Occupancy target of the walk kernel (waves per SIMD): four for the mixed tree walk, which then
// takes 108 registers and spills nothing (five fit in 96 registers with 44 bytes of spills per lane
// and ran 1-2% slower, six slower still); three for the exact tree walk, whose FP64 state needs about 200 registers (two
// waves do not spill but hide less of the root descents' latency, four spill more); four for the
// others.
#ifndef SG_WALK_WAVES_PER_EU
#define SG_WALK_WAVES_PER_EU 4
#endif
template <class W> struct WalkWaves { static constexpr int value = SG_WALK_WAVES_PER_EU; };
#ifndef SG_W32_WAVES
#define SG_W32_WAVES 4
#endif
template <> struct WalkWaves<TreeW32> { static constexpr int value = SG_W32_WAVES; };
#ifndef SG_W64_WAVES
#define SG_W64_WAVES 3
#endif
template <> struct WalkWaves<TreeW64> { static constexpr int value = SG_W64_WAVES; };
template <> struct WalkWaves<VoroW<double>> { static constexpr int value = 3; };   // 4 would spill
template <> struct WalkWaves<VoroW<float>> { static constexpr int value = 3; };

// Walk pass: persistent lanes over the requests [0, *q.tail).
template <class W, bool ABS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WalkWaves<W>::value, 8))) k_walks(const typename W::View g, const McArgs a, WalkQueue<typename W::Real> q,
                                              Slots<typename W::Real> sl, unsigned long long* claim, int nlds_req,
                                              int xcd_segments)
{
    using R = typename W::Real;
    __shared__ InstrDev s_instr[SG_MAX_INSTR];
    extern __shared__ __align__(16) unsigned char smem[];
    int32_t* s_nodes = reinterpret_cast<int32_t*>(smem);
    int nlds = 0;
    if (nlds_req > 0) {
        stage_tree_nodes(g, s_nodes, nlds);
        nlds = min(nlds, nlds_req);
    }
    if (threadIdx.x < SG_MAX_INSTR) s_instr[threadIdx.x] = a.scene->instr[threadIdx.x];
    __syncthreads();
    const uint64_t nq = *q.tail;
    const WalkReq<R>* reqs = q.sorted ? q.sorted : q.req;
    Hot<W> h;
    h.stage = ST_NEWPKT;   // idle: needs a request
    h.nseg = 0; h.nent = 0;
    h.w.nrecompute = 0; h.w.nstuck = 0;
    uint32_t idx = 0, inst = 0;
    uint32_t n_own = 0, n_peel = 0, n_re = 0;   // the lane's crossings by walk kind
    uint64_t ccur = 0, cend = 0;                 // the wave's chunk of requests
    int seg = blockIdx.x % kXcdSegs, tried = 0;  // XCD segments: the wave's segment, segments left behind
    bool exhausted = false;
    const int thr = a.compact;
    for (;;) {
        if (h.stage == ST_PEELW_DONE) {
            // FrameInstrument::detect (FrameInstrument.cpp:33-47)
            const InstrDev& in = s_instr[inst];
            unsafeAtomicAdd(in.frame + (size_t)a.ell * in.Nx * in.Ny + idx, (double)(h.L0 * rexp<R>(-h.tau)));
        } else if (h.stage == ST_OWN_DONE) {
            sl.fr(SR_RES, idx) = h.tau;
        } else if (h.stage == ST_PROP_DONE) {
            sl.fr(SR_RES, idx) = prop_length(h);
            // the leaf holding the interaction point: where the packet's next walks start
            if constexpr (W::kHint) sl.fu(SU_HINT, idx) = (uint32_t)h.w.hint();
        }
        if (finished_stage(h.stage)) {
            count_walk(h, n_own, n_peel, n_re);
            h.stage = ST_NEWPKT;
        }
        const bool want = h.stage == ST_NEWPKT;
        uint64_t r = 0;
        const bool got = xcd_segments ? seg_claim(claim + 1, nq, want, seg, tried, ccur, cend, r, exhausted)
                                      : chunk_claim(claim, nq, want, ccur, cend, r);
        if (want && !got && xcd_segments && !exhausted) {
            // not served by this claim: the lane stays idle and claims again
        } else if (want) {
            const WalkReq<R> w = got ? reqs[r] : WalkReq<R>{};
            const uint32_t kind = w.tag & 3u;
            const uint32_t wi = w.tag >> 2;
            // a request that names no pixel of an instrument or no packet slot is reported (error bit 2)
            // and dropped, never followed; the lane stays idle and takes the next one
            const uint32_t npix = wi < (uint32_t)a.ni ? (uint32_t)(s_instr[wi].Nx * s_instr[wi].Ny) : 0u;
            const bool valid = kind == WK_PEEL ? w.idx < npix : kind <= WK_PROP && (int64_t)w.idx < sl.n;
            if (!got) h.stage = ST_DEAD;
            else if (!valid) atomicOr(a.error, 4);
            else {
                idx = w.idx;
                inst = wi;
                const R ox = w.o[0], oy = w.o[1], oz = w.o[2];
                const R val = w.val;
                double dx, dy, dz;
                if (kind == WK_PEEL) {
                    const InstrDev& in = s_instr[inst];
                    dx = in.kobs[0]; dy = in.kobs[1]; dz = in.kobs[2];
                } else {
                    dx = w.d[0]; dy = w.d[1]; dz = w.d[2];
                }
                if constexpr (W::kHint)
                    h.w.init_hinted(g, s_nodes, nlds, ox, oy, oz, dx, dy, dz, (int)w.hint);
                else
                    h.w.init(g, s_nodes, nlds, ox, oy, oz, dx, dy, dz);
                // the value: a peel-off's or own path's luminosity, or the propagation target depth
                h.L0 = val;
                if (kind == WK_PEEL) h.start_walk(ST_PEELW, (R)a.tau_cut);
                else if (kind == WK_OWN) h.start_walk(ST_OWN, (R)0);
                else h.start_walk(ST_PROP, val);
            }
        }
        const uint64_t alive = __ballot(h.stage != ST_DEAD);
        if (!alive) break;
        // wave-uniform, kept in a scalar register
        const int need_n = __builtin_amdgcn_readfirstlane(max(1, min(thr, __popcll(alive) >> 1)));
        bool walk = walking_stage(h.stage);
        for (;;) {
            if (walk) walk = crossing<W, ABS>(h, g, s_nodes, nlds, a);
            // lanes alive and not walking are idle (alive is fixed while the wave walks)
            const uint64_t wm = __ballot(walk);
            if (__popcll(alive & ~wm) >= need_n || !wm) break;
        }
    }
    add_wave(a.counters + 1, n_own);
    add_wave(a.counters + 2, n_peel);
    add_wave(a.counters + 8, n_re);
    add_wave(a.counters + 5, h.w.nstuck);
    add_wave(a.counters + 9, h.w.nrecompute);
}

} // namespace sg
