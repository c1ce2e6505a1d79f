// sg_mc.hpp -- the photon-packet life cycle of MonteCarloSimulation::dostellaremissionchunk
// (MonteCarloSimulation.cpp:259-297) as a persistent, event-driven lane pipeline.
//
// Each lane owns one packet at a time and moves it through the stages
//   EMIT -> PEEL(emission) -> OWN walk -> [stop?] -> PROP re-walk -> PEEL(scattering) -> SCATTER -> OWN ...
// Every stage that needs a grid walk starts a segment generator; the wave then runs the shared
// crossing loop until a compaction threshold of its lanes has finished its walk, and only then
// re-arms those lanes together (wavefront ballot compaction). Packet indices are claimed with
// one atomic per wave (ballot + prefix count), and each packet's random numbers come from its
// own Philox stream, so the result does not depend on which lane or wave ran it.
#pragma once
#include "sg_walk.hpp"

namespace sg {

// Each walking stage is followed by its finished stage (a walk ends by stepping its stage by one).
enum Stage : int { ST_NEWPKT = 0, ST_PEEL, ST_PEELW, ST_PEELW_DONE, ST_OWN_START, ST_OWN, ST_OWN_DONE, ST_PROP,
                   ST_PROP_DONE, ST_SCATTER, ST_DEAD };
static_assert(ST_PEELW_DONE == ST_PEELW + 1 && ST_OWN_DONE == ST_OWN + 1 && ST_PROP_DONE == ST_PROP + 1,
              "a finished stage follows its walking stage");

// ------------------------------------------------------------------------- emission (FP64)

// Lambert W_{-1} on [-1/e, 0): branch-point series start (Corless et al. 1996, series in
// p = -sqrt(2(1 + e z))) refined with Halley steps; the same stopping rule as
// SpecialFunctions::LambertW1 (SpecialFunctions.cpp:579-620).
__device__ double lambertw_m1(double z)
{
    const double em1 = 0.36787944117144232159552377016146;
    if (!(z >= -em1 && z < 0.0)) return z == 0.0 ? -DBL_MAX : NAN;
    const double a[12] = {-1.0, 1.0, -1.0 / 3.0, 11.0 / 72.0, -43.0 / 540.0, 769.0 / 17280.0, -221.0 / 8505.0,
                          680863.0 / 43545600.0, -1963.0 / 204120.0, 226287557.0 / 37623398400.0,
                          -5776369.0 / 1515591000.0, 169709463197.0 / 69528040243200.0};
    double q = z + em1;
    double p = -sqrt(2.0 * 2.718281828459045235360287 * q);   // p <= 0 on the W_{-1} branch
    double w0 = a[11];
    for (int i = 10; i >= 0; i--) w0 = a[i] + p * w0;
    if (q < 3.0e-3) return w0;
    double w;
    if (z < -1e-6) w = w0;
    else {
        double l1 = log(-z), l2 = log(-l1);
        w = l1 - l2 + l2 / l1;
    }
    for (int i = 0; i < 10; i++) {
        double e = exp(w);
        double t = w * e - z;
        double pp = w + 1.0;
        t /= e * pp - 0.5 * (pp + 1.0) * t / pp;
        w -= t;
        if (fabs(t) < 1.0e-12 * (1.0 + fabs(w))) return w;
    }
    return w;
}

// Direction(theta, phi) (Direction.cpp:12-36)
__device__ __forceinline__ void dir_from_angles(double theta, double phi, double& kx, double& ky, double& kz)
{
    const double eps = 1e-8;
    if (theta <= eps) { kx = 0; ky = 0; kz = 1; }
    else if (theta >= M_PI - eps) { kx = 0; ky = 0; kz = -1; }
    else {
        double st = sin(theta);
        kx = st * cos(phi); ky = st * sin(phi); kz = cos(theta);
    }
}

// Random::direction (Random.cpp:180-185)
__device__ __forceinline__ void iso_direction(PacketRng& rng, double& kx, double& ky, double& kz)
{
    double theta = acos(2.0 * rng.ud() - 1.0);
    double phi = 2.0 * M_PI * rng.ud();
    dir_from_angles(theta, phi, kx, ky, kz);
}

__device__ __forceinline__ double loglog(double x, double x1, double x2, double f1, double f2)
{
    x = log10(x); x1 = log10(x1); x2 = log10(x2);
    bool lf = f1 > 0 && f2 > 0;
    if (lf) { f1 = log10(f1); f2 = log10(f2); }
    double fx = f1 + ((x - x1) / (x2 - x1)) * (f2 - f1);
    return lf ? pow(10.0, fx) : fx;
}

// GeometricStellarComp::launch position: ExpDiskGeometry.cpp:117-146 / SepAxGeometry.cpp:21-30,
// SersicGeometry.cpp:85-90 with SersicFunction::inversemass (SersicFunction.cpp:112-124),
// SpiralStructureGeometryDecorator.cpp:176-191,223-229.
__device__ __forceinline__ void star_position(const StarDev& s, PacketRng& rng, double& x, double& y, double& z)
{
    if (s.kind == 1) {
        double hR = s.p[0], hz = s.p[1], Rmax = s.p[2], zmax = s.p[3], Rmin = s.p[4];
        double R, X;
        do {
            X = rng.ud();
            R = hR * (-1.0 - lambertw_m1((X - 1.0) / M_E));
        } while ((Rmax > 0.0 && R >= Rmax) || R <= Rmin);
        double phi = 2.0 * M_PI * rng.ud();
        double zz;
        do {
            X = rng.ud();
            zz = (X <= 0.5) ? hz * log(2.0 * X) : -hz * log(2.0 * (1.0 - X));
        } while (zmax > 0.0 && fabs(zz) >= zmax);
        x = R * cos(phi); y = R * sin(phi); z = zz;
    } else {
        double M = rng.ud();
        double sr;
        int n = s.nt;
        if (M <= s.tM[0]) sr = s.ts[0];
        else if (M >= s.tM[n - 1]) sr = s.ts[n - 1];
        else {
            int i = locate_clip(s.tM, n, M);
            sr = loglog(M, s.tM[i], s.tM[i + 1], s.ts[i], s.ts[i + 1]);
        }
        double r = s.p[1] * sr;
        double kx, ky, kz;
        iso_direction(rng, kx, ky, kz);
        x = r * kx; y = r * ky; z = r * kz;
    }
    if (s.spiral) {
        double R = sqrt(x * x + y * y);
        double m = s.sp[0], R0 = s.sp[2], phi0 = s.sp[3], w = s.sp[4];
        int N = (int)s.sp[5];
        double c = 1.0 + (s.CN - 1.0) * w;
        double phi, t;
        do {
            phi = 2.0 * M_PI * rng.ud();
            double gamma = log(R / R0) / s.tanp + phi0 + 0.5 * M_PI / m;
            double pert = (1.0 - w) + w * s.CN * pow(sin(0.5 * m * (gamma - phi)), 2 * N);
            t = rng.ud() * c / pert;
        } while (t > 1);
        x = R * cos(phi); y = R * sin(phi);
    }
}

// StellarSystem::launch (StellarSystem.cpp:116-158) with GeometricStellarComp::launch
// (GeometricStellarComp.cpp:75-81): luminosity weight, position, isotropic direction.
__device__ __forceinline__ double emit(const McArgs& a, PacketRng& rng, double& x, double& y, double& z, double& kx, double& ky,
                       double& kz)
{
    int N = a.ns;
    int h = 0;
    double L = a.Lpkt;
    if (N > 1) {
        double X = rng.ud();
        if (X < a.emission_bias) h = max(0, min(N - 1, (int)(N * X / a.emission_bias)));
        else h = locate_clip(a.scene->Xcum, N + 1, (X - a.emission_bias) / (1.0 - a.emission_bias));
        double Lh = a.scene->star[h].lum;
        if (!(Lh > 0)) return 0.0;
        L = a.Lpkt * (1.0 / (1.0 - a.emission_bias + a.emission_bias * a.Lmean / Lh));
    }
    star_position(a.scene->star[h], rng, x, y, z);
    iso_direction(rng, kx, ky, kz);
    return L;
}

// Emission stage of the pipeline: StellarSystem::launch for packets [first, first+count), one lane
// per packet, writing the launch state (structure of arrays in the run's real type) and the
// number of random draws consumed, so the transport kernel continues each packet's stream.
// Emission of packet i of the range of `a`: its launch state goes to `out` (structure of arrays of
// a.count values) and the number of random draws it consumed to draws[i].
template <typename R>
__device__ __forceinline__ void emit_one(const McArgs& a, R* out, uint32_t* draws, size_t i)
{
    const size_t n = a.count;
    PacketRng rng;
    rng.start(a.seed, a.member, a.first + i, a.ell);
    double x = 0, y = 0, z = 0, kx = 0, ky = 0, kz = 0;
    double L = emit(a, rng, x, y, z, kx, ky, kz);
    if (!(L > 0)) { L = 0; x = y = z = kx = ky = kz = 0; }
    out[i] = (R)x; out[n + i] = (R)y; out[2 * n + i] = (R)z;
    out[3 * n + i] = (R)kx; out[4 * n + i] = (R)ky; out[5 * n + i] = (R)kz;
    out[6 * n + i] = (R)L;
    draws[i] = rng.draw;
}

template <typename R>
__global__ void __launch_bounds__(256) k_emit(const McArgs a, R* out, uint32_t* draws)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.count; i += (size_t)gridDim.x * blockDim.x)
        emit_one<R>(a, out, draws, i);
}

// Emission for a table of packet ranges laid end to end (range k owns positions
// [roff[k], roff[k+1])); each range writes where its own `emitted` and `emit_draw` point.
template <typename R>
__global__ void __launch_bounds__(256) k_emit_batch(const McArgs* atab, const uint64_t* roff, int nranges)
{
    const uint64_t total = roff[nranges];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nranges - 1;   // the last range that starts at or before i
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (roff[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        const McArgs& a = atab[lo];
        emit_one<R>(a, static_cast<R*>(const_cast<void*>(a.emitted)), const_cast<uint32_t*>(a.emit_draw),
                    (size_t)(i - roff[lo]));
    }
}

// SingleFrameInstrument::pixelondetector (SingleFrameInstrument.cpp:130-147)
template <typename R> __device__ __forceinline__ int pixel_on_detector(const InstrDev& in, R x, R y, R z)
{
    R xpp = -(R)in.sinphi * x + (R)in.cosphi * y;
    R ypp = -(R)in.cosphi * (R)in.costheta * x - (R)in.sinphi * (R)in.costheta * y + (R)in.sintheta * z;
    R xp = (R)in.cospa * xpp - (R)in.sinpa * ypp;
    R yp = (R)in.sinpa * xpp + (R)in.cospa * ypp;
    R fi = floor((xp - (R)in.xpmin) / (R)in.xpsiz);
    R fj = floor((yp - (R)in.ypmin) / (R)in.ypsiz);
    if (!(fi >= (R)0 && fi < (R)in.Nx && fj >= (R)0 && fj < (R)in.Ny)) return -1;
    return (int)fi + in.Nx * (int)fj;
}

// ------------------------------------------------------------------------- lane state
//
// Hot state (touched on every crossing) stays in registers: the walker, the stage and the running
// optical depth / path length. Cold per-packet state (position, direction, luminosity, random
// stream, peel-off bookkeeping) lives in LDS as structure-of-arrays, one slot per lane, and is only
// read and written when a lane is re-armed. This keeps the crossing loop's register footprint low
// enough for several waves per SIMD.

// limit: the optical depth beyond which the walk ends (the propagation target, or the peel-off
// optical-depth cut; infinite for none); nent: the walk's entry segments (outside the grid).
template <class W> struct Hot {
    using R = typename W::Real;
    W w;
    R tau, s, tprev, sprev, limit, L0;
    int nseg, nent, stage;
    // several dust components only: the own walk's scattered luminosity, and the cell of the walk's
    // last segment (for a propagation re-walk, the cell of the interaction point)
    R lsca;
    int mcell;
    template <class P> __device__ __forceinline__ void park(P& p)
    {
        w.park(p);
        p.io(tau); p.io(s); p.io(tprev); p.io(sprev); p.io(limit); p.io(L0);
        p.io(nseg); p.io(nent); p.io(stage); p.io(lsca); p.io(mcell);
    }
    // After w.init: the running sums start at zero. stage: ST_PEELW, ST_OWN or ST_PROP; lim: the end
    // optical depth (> 0), anything else for none.
    __device__ __forceinline__ void start_walk(int st, R lim)
    {
        tau = 0; s = 0; tprev = 0; sprev = 0; nseg = 0;
        nent = w.entries();
        limit = lim > (R)0 ? lim : (R)INFINITY;
        stage = st;
    }
    // cell crossings of the finished walk (segments inside the grid)
    __device__ __forceinline__ uint32_t crossings() const { return (uint32_t)(nseg - nent); }
};
// DustGridPath::pathlength (DustGridPath.cpp:162-173, NR.hpp:76-95) at the end of a propagation
// re-walk: the distance at which the cumulative optical depth reaches the target, interpolated
// linearly in the last segment; the path's full length when the target lies beyond its end, and 0
// for an empty path or a target of zero.
template <class W> __device__ __forceinline__ typename W::Real prop_length(const Hot<W>& h)
{
    using R = typename W::Real;
    if (h.nseg == 0 || !(h.limit < (R)INFINITY)) return (R)0;
    if (h.tau > h.limit || (h.tau == h.limit && h.nseg > 1))
        return interp_linlin(h.limit, h.tprev, h.tau, h.sprev, h.s);
    return h.s;
}

template <class W> __host__ __device__ constexpr int hot_words()
{
    return W::kParkWords + (int)(7 * sizeof(typename W::Real) / 4) + 4;
}

// Hot state is parked in LDS (structure-of-arrays, one slot per lane) while a wave re-arms lanes,
// so that none of it is live in registers across the state-machine code.
template <class W> __device__ __forceinline__ void lds_store(Hot<W>& v, uint32_t* base, int nl, int lane)
{
    Park<false> p{base, nl, lane, 0};
    v.park(p);
}
template <class W> __device__ __forceinline__ void lds_load(Hot<W>& v, uint32_t* base, int nl, int lane)
{
    Park<true> p{base, nl, lane, 0};
    v.park(p);
}

enum ColdR { CR_RX = 0, CR_RY, CR_RZ, CR_KX, CR_KY, CR_KZ, CR_L, CR_LPEEL, CR_NR };
enum ColdU { CU_PLO = 0, CU_PHI, CU_DRAW, CU_W0, CU_W1, CU_W2, CU_W3, CU_NSCATT, CU_PEEL, CU_PEELKIND, CU_PIXEL,
             CU_MCELL, CU_PK, CU_SCAT, CU_NPEEL, CU_WALKS, CU_NU };

template <typename R> struct ColdSlot {
    R* r;          // CR_NR arrays of nl slots (device memory, one slot per launched lane)
    uint32_t* u;   // CU_NU arrays of nl slots
    int nl, lane;
    __device__ __forceinline__ R& fr(int f) const { return r[f * nl + lane]; }
    __device__ __forceinline__ uint32_t& fu(int f) const { return u[f * nl + lane]; }
};

template <typename R> __host__ __device__ constexpr size_t cold_bytes(int nl)
{
    return (size_t)nl * (CR_NR * sizeof(R) + CU_NU * sizeof(uint32_t));
}

__device__ __forceinline__ bool wave_claim(unsigned long long* ctr, uint64_t count, uint64_t& p)
{
    uint64_t mask = __ballot(1);
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(ctr, (unsigned long long)__popcll(mask));
    uint32_t lo = __shfl((uint32_t)base, leader), hi = __shfl((uint32_t)(base >> 32), leader);
    base = ((uint64_t)hi << 32) | lo;
    uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    p = base + (uint64_t)__popcll(below);
    return p < count;
}

// Random::exponcutoff (Random.cpp:163-176)
__device__ __forceinline__ double exponcutoff(PacketRng& rng, double xmax)
{
    if (xmax == 0.0) return 0.0;
    else if (xmax < 1e-10) return rng.ud() * xmax;
    double x = -log(1.0 - rng.ud() * (1.0 - exp(-xmax)));
    while (x > xmax) x = -log(1.0 - rng.ud() * (1.0 - exp(-xmax)));
    return x;
}
// FP32 form: -log1p(-U * (-expm1(-xmax))) keeps full relative accuracy at small optical depth.
__device__ __forceinline__ float exponcutoff(PacketRng& rng, float xmax)
{
    if (xmax == 0.0f) return 0.0f;
    else if (xmax < 1e-10f) return rng.uf() * xmax;
    float em = -expm1f(-xmax);
    float x = -log1pf(-rng.uf() * em);
    while (x > xmax) x = -log1pf(-rng.uf() * em);
    return x;
}

template <typename R> __device__ __forceinline__ R uni(PacketRng& rng);
template <> __device__ __forceinline__ double uni<double>(PacketRng& rng) { return rng.ud(); }
template <> __device__ __forceinline__ float uni<float>(PacketRng& rng) { return rng.uf(); }

template <typename R> __device__ __forceinline__ R rexp(R x);
template <> __device__ __forceinline__ double rexp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float rexp<float>(float x) { return __expf(x); }
template <typename R> __device__ __forceinline__ R rexpm1(R x);
template <> __device__ __forceinline__ double rexpm1<double>(double x) { return expm1(x); }
template <> __device__ __forceinline__ float rexpm1<float>(float x) { return expm1f(x); }

// ------------------------------------------------------------------------- dust components

// Weights of the components' phase functions at the scattering point, kappasca(h) rho(m, h)
// (MonteCarloSimulation.cpp:318-331): false when the point lies outside the grid or has no
// scattering dust, which aborts the scattering peel-off.
template <typename R> __device__ __forceinline__ bool peel_weights_ok(const McArgs& a, int m)
{
    if (m < 0) return false;
    double sum = 0;
    for (int k = 0; k < a.ncomp; k++) sum += a.scene->ksca[k] * a.rhoc[(size_t)k * a.ncells + m];
    return sum > 0;
}

// The peel-off's phase-function value at cos(alpha): each component's Henyey-Greenstein value
// (DustMix.cpp:648-653) weighted by its normalised kappasca(h) rho(m, h) (MonteCarloSimulation.cpp:318-343)
template <typename R> __device__ __forceinline__ R mixed_phase(const McArgs& a, int m, R cosalpha)
{
    double sum = 0;
    for (int k = 0; k < a.ncomp; k++) sum += a.scene->ksca[k] * a.rhoc[(size_t)k * a.ncells + m];
    R I = 0;
    for (int k = 0; k < a.ncomp; k++) {
        R w = (R)(a.scene->ksca[k] * a.rhoc[(size_t)k * a.ncells + m] / sum);
        R gk = (R)a.scene->gc[k];
        R t = (R)1.0 + gk * gk - (R)2 * gk * cosalpha;
        I += w * (((R)1.0 - gk) * ((R)1.0 + gk) / sqrt(t * t * t));
    }
    return I;
}

// DustSystem::randomMixForPosition (DustSystem.cpp:855-869): the component drawn with probability
// proportional to kappasca(h) rho(m, h) (NR::cdf, then NR::locate_clip on one uniform deviate); the first
// component, without a draw, outside the grid.
__device__ __forceinline__ int random_mix(const McArgs& a, int m, PacketRng& rng)
{
    if (m < 0) return 0;
    double X[SG_MAX_DUST + 1];
    X[0] = 0;
    for (int k = 0; k < a.ncomp; k++) X[k + 1] = X[k] + a.scene->ksca[k] * a.rhoc[(size_t)k * a.ncells + m];
    const double total = X[a.ncomp];
    for (int k = 0; k <= a.ncomp; k++) X[k] /= total;
    double u = rng.ud();
    int h = 0;
    for (int k = 1; k < a.ncomp; k++) h = (X[k] <= u) ? k : h;
    return h;
}

// ------------------------------------------------------------------------- stages

// Re-arms one lane: runs the packet state machine from its current stage until the lane has a
// walk to do (PEELW, OWN or PROP) or no packets are left (DEAD).
template <class W, bool ABS, bool MULTI = false>
__device__ __forceinline__ void service(Hot<W>& h, const ColdSlot<typename W::Real>& c, const typename W::View& g,
                                        const int32_t* lds, int nlds, const McArgs& a, const InstrDev* instr)
{
    using R = typename W::Real;
    const R albedo = (R)a.albedo;
    R gg = (R)a.g;
    R rx = c.fr(CR_RX), ry = c.fr(CR_RY), rz = c.fr(CR_RZ);
    R kx = c.fr(CR_KX), ky = c.fr(CR_KY), kz = c.fr(CR_KZ);
    R L = c.fr(CR_L), Lpeel = c.fr(CR_LPEEL);
    PacketRng rng;
    rng.k0 = a.seed; rng.k1 = a.member; rng.ell = (uint32_t)a.ell;
    rng.plo = c.fu(CU_PLO); rng.phi = c.fu(CU_PHI); rng.draw = c.fu(CU_DRAW);
    rng.w0 = c.fu(CU_W0); rng.w1 = c.fu(CU_W1); rng.w2 = c.fu(CU_W2); rng.w3 = c.fu(CU_W3);
    int nscatt = (int)c.fu(CU_NSCATT), peel = (int)c.fu(CU_PEEL), peelkind = (int)c.fu(CU_PEELKIND);
    int pixel = (int)c.fu(CU_PIXEL);
    int mcell = MULTI ? (int)c.fu(CU_MCELL) : -1;
    uint32_t c_pk = 0, c_scat = 0, c_npeel = 0, c_walks = 0;
    bool armed = false;
    while (!armed) {
        switch (h.stage) {
        case ST_NEWPKT: {
            uint64_t p;
            if (!wave_claim(a.next_packet, a.count, p)) { h.stage = ST_DEAD; armed = true; break; }
            // launch state from the emission kernel; continue the packet's random stream where
            // the emission left it
            const R* e = reinterpret_cast<const R*>(a.emitted);
            const size_t n = a.count;
            c_pk++;
            R Le = e[6 * n + p];
            if (!(Le > (R)0)) break;
            rx = e[p]; ry = e[n + p]; rz = e[2 * n + p];
            kx = e[3 * n + p]; ky = e[4 * n + p]; kz = e[5 * n + p];
            L = Le;
            rng.start(a.seed, a.member, a.first + p, a.ell);
            rng.draw = a.emit_draw[p];
            if (rng.draw & 3u) {
                rng.w0 = rng.plo; rng.w1 = rng.phi; rng.w2 = rng.draw >> 2; rng.w3 = rng.ell;
                philox10(rng.w0, rng.w1, rng.w2, rng.w3, rng.k0, rng.k1);
            }
            nscatt = 0; peel = 0; peelkind = 0;
            h.stage = ST_PEEL;
            break;
        }
        case ST_PEEL: {
            if (peel >= a.ni) {
                if (peelkind == 0) h.stage = a.has_dust ? ST_OWN_START : ST_NEWPKT;
                else h.stage = ST_SCATTER;
                break;
            }
            if (MULTI && peelkind && peel == 0 && !peel_weights_ok<R>(a, mcell)) {
                // peeloffscattering aborts when the scattering point has no scattering dust
                // (MonteCarloSimulation.cpp:321-327)
                peel = a.ni;
                break;
            }
            const InstrDev& in = instr[peel];
            int pix = pixel_on_detector<R>(in, rx, ry, rz);
            if (pix < 0) { peel++; break; }
            R Lp = L;
            if (peelkind) {
                // DustMix::phaseFunctionValue, Henyey-Greenstein (DustMix.cpp:648-653)
                R cosalpha = kx * (R)in.kobs[0] + ky * (R)in.kobs[1] + kz * (R)in.kobs[2];
                if (MULTI) Lp = L * mixed_phase<R>(a, mcell, cosalpha);
                else {
                    R t = (R)1.0 + gg * gg - (R)2 * gg * cosalpha;
                    Lp = L * (((R)1.0 - gg) * ((R)1.0 + gg) / sqrt(t * t * t));
                }
            }
            c_npeel++;
            if (!a.has_dust) {
                unsafeAtomicAdd(in.frame + (size_t)a.ell * in.Nx * in.Ny + pix, (double)Lp);
                peel++;
                break;
            }
            pixel = pix; Lpeel = Lp;
            h.w.init(g, lds, nlds, rx, ry, rz, in.kobs[0], in.kobs[1], in.kobs[2]);
            h.start_walk(ST_PEELW, (R)a.tau_cut);
            c_walks++;
            armed = true;
            break;
        }
        case ST_PEELW_DONE: {
            const InstrDev& in = instr[peel];
            // FrameInstrument::detect (FrameInstrument.cpp:33-47)
            unsafeAtomicAdd(in.frame + (size_t)a.ell * in.Nx * in.Ny + pixel, (double)(Lpeel * rexp<R>(-h.tau)));
            peel++;
            h.stage = ST_PEEL;
            break;
        }
        case ST_OWN_START: {
            h.L0 = L;
            h.lsca = (R)0;
            h.w.init(g, lds, nlds, rx, ry, rz, kx, ky, kz);
            h.start_walk(ST_OWN, (R)0);
            c_walks++;
            armed = true;
            break;
        }
        case ST_OWN_DONE: {
            // DustSystem::fillOpticalDepth check (DustSystem.cpp:952-955)
            R taupath = h.tau;
            if (taupath < (R)0 || isnan(taupath) || isinf(taupath)) {
                atomicOr(a.error, 1);
                h.stage = ST_NEWPKT;
                break;
            }
            // simulateescapeandabsorption: one component (MonteCarloSimulation.cpp:432-461), or the sum of
            // each cell's albedo times its interacting luminosity, gathered along the walk (:463-508)
            if (MULTI) L = h.lsca;
            else {
                R expfactor = -rexpm1<R>(-taupath);
                L = h.L0 * albedo * expfactor;
            }
            if (L <= (R)0 || (L <= (R)a.Lthreshold && nscatt >= a.minfs)) { h.stage = ST_NEWPKT; break; }
            // simulatepropagation (MonteCarloSimulation.cpp:513-531)
            if (taupath == (R)0) { peel = 0; peelkind = 1; mcell = -1; h.stage = ST_PEEL; break; }
            R tt;
            if (a.xi == 0.0) tt = exponcutoff(rng, taupath);
            else {
                R X = uni<R>(rng);
                tt = (X < (R)a.xi) ? uni<R>(rng) * taupath : exponcutoff(rng, taupath);
                R p = -rexp<R>(-tt) / rexpm1<R>(-taupath);
                R q = ((R)1.0 - (R)a.xi) * p + (R)a.xi / taupath;
                L = L * (p / q);
            }
            h.w.init(g, lds, nlds, rx, ry, rz, kx, ky, kz);
            h.start_walk(ST_PROP, tt);
            armed = true;
            break;
        }
        case ST_PROP_DONE: {
            // PhotonPackage::propagate (PhotonPackage.cpp:93-96) by pathlength(tau)
            R d = prop_length(h);
            rx = rx + d * kx; ry = ry + d * ky; rz = rz + d * kz;
            peel = 0; peelkind = 1;
            mcell = h.mcell;
            h.stage = ST_PEEL;
            break;
        }
        case ST_SCATTER: {
            // simulatescattering, Henyey-Greenstein (DustMix.cpp:596-601) and Random::direction(k, cos)
            // (Random.cpp:189-223); with several components, the mix drawn first
            // (DustSystem::randomMixForPosition, DustSystem.cpp:855-869)
            if (MULTI) gg = (R)a.scene->gc[random_mix(a, mcell, rng)];
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
            nscatt++;
            c_scat++;
            h.stage = ST_OWN_START;
            break;
        }
        default:
            armed = true;   // walking or dead
            break;
        }
    }
    c.fr(CR_RX) = rx; c.fr(CR_RY) = ry; c.fr(CR_RZ) = rz;
    c.fr(CR_KX) = kx; c.fr(CR_KY) = ky; c.fr(CR_KZ) = kz;
    c.fr(CR_L) = L; c.fr(CR_LPEEL) = Lpeel;
    c.fu(CU_PLO) = rng.plo; c.fu(CU_PHI) = rng.phi; c.fu(CU_DRAW) = rng.draw;
    c.fu(CU_W0) = rng.w0; c.fu(CU_W1) = rng.w1; c.fu(CU_W2) = rng.w2; c.fu(CU_W3) = rng.w3;
    c.fu(CU_NSCATT) = (uint32_t)nscatt; c.fu(CU_PEEL) = (uint32_t)peel; c.fu(CU_PEELKIND) = (uint32_t)peelkind;
    c.fu(CU_PIXEL) = (uint32_t)pixel;
    if (MULTI) c.fu(CU_MCELL) = (uint32_t)mcell;
    c.fu(CU_PK) += c_pk; c.fu(CU_SCAT) += c_scat; c.fu(CU_NPEEL) += c_npeel; c.fu(CU_WALKS) += c_walks;
}

// The re-arming step as a separate (not inlined) function: it reads the lane's parked hot state,
// runs the state machine and parks the result. Its register demand (FP64 emission and scattering
// math) then stays out of the crossing loop's allocation; any spilling it needs happens here only.
template <class W, bool ABS, bool MULTI = false>
__device__ __noinline__ void service_call(uint32_t* park, int nl, int lane, ColdSlot<typename W::Real> c,
                                          const typename W::View* g, const int32_t* lds, int nlds, const McArgs* a,
                                          const InstrDev* instr)
{
    Hot<W> q;
    lds_load(q, park, nl, lane);
    service<W, ABS, MULTI>(q, c, *g, lds, nlds, *a, instr);
    lds_store(q, park, nl, lane);
}

// One step of the current walk; marks the walk finished (its stage steps to the finished one) when
// the path ends, or when the optical depth passes the walk's limit (the propagation target, or the
// peel-off cut). Returns whether the walk goes on. Crossings are counted per walk when it ends
// (segments minus entry segments), so nothing here is counted across the wave.
template <class W, bool ABS, bool MULTI = false>
__device__ __forceinline__ bool crossing(Hot<W>& h, const typename W::View& g, const int32_t* lds, int nlds,
                                         const McArgs& a)
{
    using R = typename W::Real;
    int m;
    R ds, rho;
    bool got;
    if constexpr (W::kStep) got = h.w.step(g, lds, nlds, m, ds, rho);
    else got = h.w.next(g, lds, nlds, m, ds, rho);
    // a segment of positive length (a single-step walker also returns zero-length steps along
    // edges and through corners)
    const bool seg = got && ds > (R)0;
    const R dsv = seg ? ds : (R)0;
    const R dtau = seg ? ((R)a.kext * rho) * ds : (R)0;   // KappaRho (DustSystem.cpp:459-485)
    if constexpr (MULTI) {
        // several components: each cell's own albedo (MonteCarloSimulation.cpp:478-505), its scattered
        // part summed for the packet, its absorbed part into the table
        if (seg) {
            const int64_t cell = m >= 0 ? W::cell_of(g, m) : -1;
            if (h.stage == ST_OWN && cell >= 0) {
                R expfactorm = -rexpm1<R>(-dtau);
                R Lintm = h.L0 * rexp<R>(-h.tau) * expfactorm;
                R alb = (R)a.albm[cell];
                h.lsca += alb * Lintm;
                if (ABS) unsafeAtomicAdd(a.labs + (size_t)a.ell * a.ncells + cell, (double)(((R)1.0 - alb) * Lintm));
            }
            h.mcell = (int)cell;
        }
    } else if (ABS && h.stage == ST_OWN && seg && m >= 0) {
        // OligoDustSystem::absorb via simulateescapeandabsorption (MonteCarloSimulation.cpp:445-458)
        R expfactorm = -rexpm1<R>(-dtau);
        R Lintm = h.L0 * rexp<R>(-h.tau) * expfactorm;
        R Labsm = ((R)1.0 - (R)a.albedo) * Lintm;
        unsafeAtomicAdd(a.labs + (size_t)a.ell * a.ncells + W::cell_of(g, m), (double)Labsm);
    }
    // running sums, kept for every walk kind without branching (only the propagation re-walk reads
    // the path length and the previous segment's sums): cumulative s and tau per segment
    h.tprev = seg ? h.tau : h.tprev;
    h.sprev = seg ? h.s : h.sprev;
    h.s += dsv;
    h.tau += dtau;
    h.nseg += seg ? 1 : 0;
    // the walk ends at the end of the path, for a peel-off beyond the optical-depth cut, and for the
    // propagation re-walk at the first segment past the target optical depth (prop_length then
    // interpolates the interaction point)
    const bool end = !got || h.tau > h.limit;
    h.stage += end ? 1 : 0;
    return !end;
}

// Adds a finished walk's crossings to the lane's count of its kind (own path, peel-off, re-walk).
template <class W> __device__ __forceinline__ void count_walk(const Hot<W>& h, uint32_t& n_own, uint32_t& n_peel,
                                                               uint32_t& n_re)
{
    const uint32_t c = h.crossings();
    n_own += h.stage == ST_OWN_DONE ? c : 0u;
    n_peel += h.stage == ST_PEELW_DONE ? c : 0u;
    n_re += h.stage == ST_PROP_DONE ? c : 0u;
}

// ST_PEELW, ST_OWN or ST_PROP, as a bit test (a chain of comparisons compiles to branches)
__device__ __forceinline__ bool walking_stage(int st)
{
    static_assert(ST_PEELW < 32 && ST_OWN < 32 && ST_PROP < 32, "stage bits");
    return ((1u << ST_PEELW | 1u << ST_OWN | 1u << ST_PROP) >> (uint32_t)st) & 1u;
}

// ST_PEELW_DONE, ST_OWN_DONE or ST_PROP_DONE
__device__ __forceinline__ bool finished_stage(int st)
{
    return ((1u << ST_PEELW_DONE | 1u << ST_OWN_DONE | 1u << ST_PROP_DONE) >> (uint32_t)st) & 1u;
}

__device__ __forceinline__ void add_wave(unsigned long long* dst, unsigned long long x)
{
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t lo = __shfl_down((uint32_t)x, off), hi = __shfl_down((uint32_t)(x >> 32), off);
        x += ((unsigned long long)hi << 32) | lo;
    }
    if (__lane_id() == 0 && x) atomicAdd(dst, x);
}


// Occupancy target of the transport kernel (waves per SIMD) and whether lanes are re-armed by an
// out-of-line function with the hot state parked in LDS (1) or inline (0).
#ifndef SG_MC_WAVES_PER_EU
#define SG_MC_WAVES_PER_EU 2
#endif
#ifndef SG_MC_CART32_WAVES
#define SG_MC_CART32_WAVES 3
#endif
#ifndef SG_MC_TREE32_WAVES
#define SG_MC_TREE32_WAVES 4
#endif
// Per walker, as measured with the lane pipeline: three waves per SIMD for the mixed Cartesian walk
// (32% faster than two on the Cartesian disc workload, four 17% slower than three); four for the mixed
// tree walk (on the spiral octree 31% faster than two despite the spills, five collapses); two for the
// others, the Voronoi walk above all (its FP64 state spills at three, 26% slower; at one it spills
// nothing but hides too little latency, 29% slower).
template <class W> struct McWaves { static constexpr int value = SG_MC_WAVES_PER_EU; };
template <> struct McWaves<CartW32> { static constexpr int value = SG_MC_CART32_WAVES; };
template <> struct McWaves<TreeW32> { static constexpr int value = SG_MC_TREE32_WAVES; };
#ifndef SG_PARK
#define SG_PARK 0
#endif

// The lane pipeline of one wave over the packet range of `a`: lanes claim packets until the range
// is exhausted and every packet the wave claimed has finished; then the wave's event counts go
// to a.counters.
template <class W, bool ABS, bool MULTI = false>
__device__ __forceinline__ void drain(const typename W::View& g, const McArgs& a, uint32_t* park,
                                      const int32_t* s_nodes, int nlds, const ColdSlot<typename W::Real>& c,
                                      const InstrDev* s_instr, const typename W::View* gdev, const McArgs* adev)
{
    const int nl = blockDim.x, lane = threadIdx.x;
    (void)park; (void)nl; (void)lane; (void)gdev; (void)adev;
    c.fu(CU_PK) = 0; c.fu(CU_SCAT) = 0; c.fu(CU_NPEEL) = 0; c.fu(CU_WALKS) = 0;
    Hot<W> h;
    h.stage = ST_NEWPKT;
    h.nseg = 0; h.nent = 0;
    h.w.nrecompute = 0; h.w.nstuck = 0;
    uint32_t n_own = 0, n_peel = 0, n_re = 0;   // the lane's crossings by walk kind
    const int thr = a.compact;
    for (;;) {
        bool need = !walking_stage(h.stage) && h.stage != ST_DEAD;
        if (finished_stage(h.stage)) count_walk(h, n_own, n_peel, n_re);
#if SG_PARK
        if (__ballot(need)) {
            // park every lane's hot state, re-arm the idle lanes, reload
            lds_store(h, park, nl, lane);
            if (need) service_call<W, ABS, MULTI>(park, nl, lane, c, gdev, s_nodes, nlds, adev, s_instr);
            lds_load(h, park, nl, lane);
        }
#else
        if (need) service<W, ABS, MULTI>(h, c, g, s_nodes, nlds, a, s_instr);
#endif
        const uint64_t alive = __ballot(h.stage != ST_DEAD);
        if (!alive) break;
        const int need_n = max(1, min(thr, __popcll(alive) >> 1));
        bool walk = walking_stage(h.stage);
        for (;;) {
            if (walk) walk = crossing<W, ABS, MULTI>(h, g, s_nodes, nlds, a);
            // lanes alive and not walking are idle (alive is fixed while the wave walks)
            const uint64_t wm = __ballot(walk);
            if (__popcll(alive & ~wm) >= need_n || !wm) break;
        }
    }
    add_wave(a.counters + 1, n_own);
    add_wave(a.counters + 2, n_peel);
    add_wave(a.counters + 8, n_re);
    add_wave(a.counters + 0, c.fu(CU_PK));
    add_wave(a.counters + 3, c.fu(CU_SCAT));
    add_wave(a.counters + 4, c.fu(CU_NPEEL));
    add_wave(a.counters + 5, h.w.nstuck);
    add_wave(a.counters + 7, c.fu(CU_WALKS));
    add_wave(a.counters + 9, h.w.nrecompute);
}

// Persistent transport kernel. Dynamic LDS: the hot-state parking area (hot_words<W>() per lane),
// then (optionally) the top tree nodes. Cold lane state lives in device memory at cold_r/cold_u.
template <class W, bool ABS, bool MULTI = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(McWaves<W>::value, 8)))
k_mc(const typename W::View g, const McArgs a, int nlds_req, typename W::Real* cold_r, uint32_t* cold_u,
     int nslots, const typename W::View* gdev, const McArgs* adev)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ InstrDev s_instr[SG_MAX_INSTR];
    const int nl = blockDim.x;
    uint32_t* park = reinterpret_cast<uint32_t*>(smem);
    int32_t* s_nodes = reinterpret_cast<int32_t*>(smem + (SG_PARK ? hot_words<W>() * 4 * nl : 0));
    int nlds = 0;
    if (nlds_req > 0) {
        stage_tree_nodes(g, s_nodes, nlds);
        nlds = min(nlds, nlds_req);
    }
    if (threadIdx.x < SG_MAX_INSTR) s_instr[threadIdx.x] = a.scene->instr[threadIdx.x];
    ColdSlot<typename W::Real> c;
    c.r = cold_r; c.u = cold_u; c.nl = nslots;
    c.lane = blockIdx.x * blockDim.x + threadIdx.x;
    __syncthreads();
    drain<W, ABS, MULTI>(g, a, park, s_nodes, nlds, c, s_instr, gdev, adev);
}

// Batched transport over many packet ranges, each with its own grid view and run arguments (one
// range per member simulation of a population, or per slice of a large member). Each wave takes
// one range at a time, runs it dry with the lane pipeline above and moves on to the next range that
// still has packets, so that members too small to fill the device share it. A range whose work
// counter has reached its count is finished for good, and every visit to an unfinished range
// advances its counter, so the scan ends.
template <class W, bool ABS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(McWaves<W>::value, 8)))
k_mc_batch(const typename W::View* gtab, const McArgs* atab, int nranges, typename W::Real* cold_r,
           uint32_t* cold_u, int nslots)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ InstrDev s_instr[4 * SG_MAX_INSTR];   // one instrument table per wave of the block
    uint32_t* park = reinterpret_cast<uint32_t*>(smem);
    InstrDev* my_instr = s_instr + (threadIdx.x >> 6) * SG_MAX_INSTR;
    ColdSlot<typename W::Real> c;
    c.r = cold_r; c.u = cold_u; c.nl = nslots;
    c.lane = blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    int r = (int)(((long long)wave * nranges) / nwaves);
    for (int scanned = 0; scanned < nranges;) {
        r = __builtin_amdgcn_readfirstlane(r);
        const McArgs* ap = atab + r;
        unsigned long long taken = __hip_atomic_load(ap->next_packet, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (taken >= ap->count) {
            r = (r + 1 == nranges) ? 0 : r + 1;
            scanned++;
            continue;
        }
        __builtin_amdgcn_wave_barrier();
        if (__lane_id() < SG_MAX_INSTR) my_instr[__lane_id()] = ap->scene->instr[__lane_id()];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        drain<W, ABS>(gtab[r], *ap, park, nullptr, 0, c, my_instr, gtab + r, ap);
        scanned = 0;
    }
}

} // namespace sg
