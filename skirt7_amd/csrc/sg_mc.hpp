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

enum Stage : int { ST_NEWPKT = 0, ST_PEEL, ST_PEELW, ST_OWN_START, ST_OWN, ST_OWN_DONE, ST_PROP, ST_PROP_DONE,
                   ST_PEELW_DONE, ST_SCATTER, ST_DEAD };

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
        else h = locate_clip(a.Xcum, N + 1, (X - a.emission_bias) / (1.0 - a.emission_bias));
        double Lh = a.star[h].lum;
        if (!(Lh > 0)) return 0.0;
        L = a.Lpkt * (1.0 / (1.0 - a.emission_bias + a.emission_bias * a.Lmean / Lh));
    }
    star_position(a.star[h], rng, x, y, z);
    iso_direction(rng, kx, ky, kz);
    return L;
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

template <class W> struct Lane {
    using R = typename W::Real;
    W w;
    PacketRng rng;
    R L, L0, rx, ry, rz, kx, ky, kz;
    R tau, s, tprev, sprev, target, Lpeel;
    int nseg;
    int stage, peel, peelkind, pixel, nscatt;
    uint32_t c_pk, c_own, c_peel, c_scat, c_npeel, c_walks, c_re;
};

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

// ------------------------------------------------------------------------- stages

template <class W, bool ABS>
__device__ __forceinline__ void service(Lane<W>& s, const typename W::View& g, const int32_t* lds, int nlds, const McArgs& a,
                        const InstrDev* instr)
{
    using R = typename W::Real;
    const R albedo = (R)a.albedo, gg = (R)a.g;
    for (;;) {
        switch (s.stage) {
        case ST_NEWPKT: {
            uint64_t p;
            if (!wave_claim(a.next_packet, a.count, p)) { s.stage = ST_DEAD; return; }
            s.rng.start(a.seed, a.member, a.first + p, a.ell);
            double x, y, z, kx, ky, kz;
            double L = emit(a, s.rng, x, y, z, kx, ky, kz);
            s.c_pk++;
            if (!(L > 0)) continue;
            s.L = (R)L; s.rx = (R)x; s.ry = (R)y; s.rz = (R)z; s.kx = (R)kx; s.ky = (R)ky; s.kz = (R)kz;
            s.nscatt = 0; s.peel = 0; s.peelkind = 0;
            s.stage = ST_PEEL;
            continue;
        }
        case ST_PEEL: {
            if (s.peel >= a.ni) {
                if (s.peelkind == 0) s.stage = a.has_dust ? ST_OWN_START : ST_NEWPKT;
                else s.stage = ST_SCATTER;
                continue;
            }
            const InstrDev& in = instr[s.peel];
            int pix = pixel_on_detector<R>(in, s.rx, s.ry, s.rz);
            if (pix < 0) { s.peel++; continue; }
            R Lp = s.L;
            if (s.peelkind) {
                // DustMix::phaseFunctionValue, Henyey-Greenstein (DustMix.cpp:648-653)
                R cosalpha = s.kx * (R)in.kobs[0] + s.ky * (R)in.kobs[1] + s.kz * (R)in.kobs[2];
                R t = (R)1.0 + gg * gg - (R)2 * gg * cosalpha;
                Lp = s.L * (((R)1.0 - gg) * ((R)1.0 + gg) / sqrt(t * t * t));
            }
            s.c_npeel++;
            if (!a.has_dust) {
                unsafeAtomicAdd(in.frame + (size_t)a.ell * in.Nx * in.Ny + pix, (double)Lp);
                s.peel++;
                continue;
            }
            s.pixel = pix; s.Lpeel = Lp; s.tau = 0;
            s.w.init(g, lds, nlds, s.rx, s.ry, s.rz, in.kobs[0], in.kobs[1], in.kobs[2]);
            s.c_walks++;
            s.stage = ST_PEELW;
            return;
        }
        case ST_PEELW_DONE: {
            const InstrDev& in = instr[s.peel];
            // FrameInstrument::detect (FrameInstrument.cpp:33-47)
            unsafeAtomicAdd(in.frame + (size_t)a.ell * in.Nx * in.Ny + s.pixel, (double)(s.Lpeel * rexp<R>(-s.tau)));
            s.peel++;
            s.stage = ST_PEEL;
            continue;
        }
        case ST_OWN_START: {
            s.tau = 0; s.L0 = s.L;
            s.w.init(g, lds, nlds, s.rx, s.ry, s.rz, s.kx, s.ky, s.kz);
            s.c_walks++;
            s.stage = ST_OWN;
            return;
        }
        case ST_OWN_DONE: {
            // DustSystem::fillOpticalDepth check (DustSystem.cpp:952-955)
            R taupath = s.tau;
            if (taupath < (R)0 || isnan(taupath) || isinf(taupath)) {
                atomicOr(a.error, 1);
                s.stage = ST_NEWPKT;
                continue;
            }
            // simulateescapeandabsorption, single component (MonteCarloSimulation.cpp:432-461)
            R expfactor = -rexpm1<R>(-taupath);
            s.L = s.L0 * albedo * expfactor;
            if (s.L <= (R)0 || (s.L <= (R)a.Lthreshold && s.nscatt >= a.minfs)) { s.stage = ST_NEWPKT; continue; }
            // simulatepropagation (MonteCarloSimulation.cpp:513-531)
            if (taupath == (R)0) { s.peel = 0; s.peelkind = 1; s.stage = ST_PEEL; continue; }
            R tt;
            if (a.xi == 0.0) tt = exponcutoff(s.rng, taupath);
            else {
                R X = uni<R>(s.rng);
                tt = (X < (R)a.xi) ? uni<R>(s.rng) * taupath : exponcutoff(s.rng, taupath);
                R p = -rexp<R>(-tt) / rexpm1<R>(-taupath);
                R q = ((R)1.0 - (R)a.xi) * p + (R)a.xi / taupath;
                s.L = s.L * (p / q);
            }
            s.target = tt;
            s.tau = 0; s.s = 0; s.tprev = 0; s.sprev = 0; s.nseg = 0;
            s.w.init(g, lds, nlds, s.rx, s.ry, s.rz, s.kx, s.ky, s.kz);
            s.stage = ST_PROP;
            return;
        }
        case ST_PROP_DONE: {
            // PhotonPackage::propagate (PhotonPackage.cpp:93-96): s.s holds pathlength(tau)
            R d = s.s;
            s.rx = s.rx + d * s.kx; s.ry = s.ry + d * s.ky; s.rz = s.rz + d * s.kz;
            s.peel = 0; s.peelkind = 1;
            s.stage = ST_PEEL;
            continue;
        }
        case ST_SCATTER: {
            // simulatescattering, Henyey-Greenstein (DustMix.cpp:596-601) and Random::direction(k, cos)
            // (Random.cpp:189-223)
            R nkx, nky, nkz;
            if (fabs((double)gg) < 1e-6) {
                double a0, a1, a2;
                iso_direction(s.rng, a0, a1, a2);
                nkx = (R)a0; nky = (R)a1; nkz = (R)a2;
            } else {
                R f = (((R)1.0 - gg) * ((R)1.0 + gg)) / ((R)1.0 - gg + (R)2.0 * gg * uni<R>(s.rng));
                R costheta = ((R)1.0 + gg * gg - f * f) / ((R)2.0 * gg);
                R phi = (R)(2.0 * M_PI) * uni<R>(s.rng);
                R cosphi = cos(phi), sinphi = sin(phi);
                R sintheta = sqrt(fabs(((R)1.0 - costheta) * ((R)1.0 + costheta)));
                R kx = s.kx, ky = s.ky, kz = s.kz;
                if (kz > (R)0.99999) { nkx = cosphi * sintheta; nky = sinphi * sintheta; nkz = costheta; }
                else if (kz < (R)-0.99999) { nkx = cosphi * sintheta; nky = sinphi * sintheta; nkz = -costheta; }
                else {
                    R root = sqrt(((R)1.0 - kz) * ((R)1.0 + kz));
                    nkx = sintheta / root * (-kx * kz * cosphi + ky * sinphi) + kx * costheta;
                    nky = -sintheta / root * (ky * kz * cosphi + kx * sinphi) + ky * costheta;
                    nkz = root * sintheta * cosphi + kz * costheta;
                }
            }
            s.kx = nkx; s.ky = nky; s.kz = nkz;
            s.nscatt++;
            s.c_scat++;
            s.stage = ST_OWN_START;
            continue;
        }
        default:
            return;   // walking or dead
        }
    }
}

// One crossing of the current walk; marks the walk finished when the path ends (or, for the
// propagation re-walk, when the target optical depth is reached).
template <class W, bool ABS>
__device__ __forceinline__ void crossing(Lane<W>& s, const typename W::View& g, const int32_t* lds, int nlds,
                                         const McArgs& a)
{
    using R = typename W::Real;
    int m;
    R ds, rho;
    bool got = s.w.next(g, lds, nlds, m, ds, rho);
    if (!got) {
        if (s.stage == ST_PROP) {
            // DustGridPath::pathlength end cases (DustGridPath.cpp:162-173, NR.hpp:76-95)
            R res;
            if (s.nseg == 0 || !(s.target > (R)0)) res = 0;
            else if (s.target > s.tau) res = s.s;
            else if (s.nseg == 1) res = s.s;
            else res = interp_linlin(s.target, s.tprev, s.tau, s.sprev, s.s);
            s.s = res;
            s.stage = ST_PROP_DONE;
        } else if (s.stage == ST_OWN) s.stage = ST_OWN_DONE;
        else s.stage = ST_PEELW_DONE;
        return;
    }
    R kr = (R)a.kext * rho;        // KappaRho (DustSystem.cpp:459-485): 0 + kappaext*rho, 0 outside
    R dtau = kr * ds;
    if (s.stage == ST_PEELW) {
        s.tau += dtau;
        s.c_peel += (m >= 0);
        if (a.tau_cut > 0.0 && s.tau > (R)a.tau_cut) s.stage = ST_PEELW_DONE;
    } else if (s.stage == ST_OWN) {
        if (ABS && m >= 0) {
            // OligoDustSystem::absorb via simulateescapeandabsorption (MonteCarloSimulation.cpp:445-458)
            R expfactorm = -rexpm1<R>(-dtau);
            R Lintm = s.L0 * rexp<R>(-s.tau) * expfactorm;
            R Labsm = ((R)1.0 - (R)a.albedo) * Lintm;
            unsafeAtomicAdd(a.labs + (size_t)a.ell * a.ncells + m, (double)Labsm);
        }
        s.tau += dtau;
        s.c_own += (m >= 0);
    } else {
        // ST_PROP: cumulative s and tau per segment; stop at the first segment past the target
        s.tprev = s.tau; s.sprev = s.s;
        s.s += ds;
        s.tau += dtau;
        s.nseg++;
        s.c_re += (m >= 0);
        if (s.target > (R)0 && s.tau > s.target) {
            s.s = (s.nseg == 1) ? interp_linlin(s.target, (R)0, s.tau, (R)0, s.s)
                                : interp_linlin(s.target, s.tprev, s.tau, s.sprev, s.s);
            s.stage = ST_PROP_DONE;
        }
    }
}

__device__ __forceinline__ bool walking_stage(int st) { return st == ST_PEELW || st == ST_OWN || st == ST_PROP; }

template <class W, bool ABS>
__global__ void __launch_bounds__(256) k_mc(const typename W::View g, const McArgs a)
{
    __shared__ int32_t s_nodes[SG_LDS_NODES];
    __shared__ InstrDev s_instr[SG_MAX_INSTR];
    int nlds = 0;
    stage_tree_nodes(g, s_nodes, nlds);
    if (threadIdx.x < SG_MAX_INSTR) s_instr[threadIdx.x] = a.instr[threadIdx.x];
    __syncthreads();

    Lane<W> s;
    s.stage = ST_NEWPKT;
    s.c_pk = s.c_own = s.c_peel = s.c_scat = s.c_npeel = s.c_walks = s.c_re = 0;
    s.w.nrecompute = 0; s.w.nstuck = 0;
    const int thr = a.compact;
    for (;;) {
        if (!walking_stage(s.stage) && s.stage != ST_DEAD) service<W, ABS>(s, g, s_nodes, nlds, a, s_instr);
        uint64_t alive = __ballot(s.stage != ST_DEAD);
        if (!alive) break;
        int need = max(1, min(thr, __popcll(alive) >> 1));
        for (;;) {
            if (walking_stage(s.stage)) crossing<W, ABS>(s, g, s_nodes, nlds, a);
            bool walk = walking_stage(s.stage);
            uint64_t idle = __ballot(!walk && s.stage != ST_DEAD);
            if (__popcll(idle) >= need || !__ballot(walk)) break;
        }
    }
    // per-lane counters -> device totals
    unsigned long long v[10] = {s.c_pk, s.c_own, s.c_peel, s.c_scat, s.c_npeel, s.w.nstuck, 0, s.c_walks, s.c_re,
                                s.w.nrecompute};
#pragma unroll
    for (int i = 0; i < 10; i++) {
        unsigned long long x = v[i];
        for (int off = 32; off > 0; off >>= 1) {
            uint32_t lo = __shfl_down((uint32_t)x, off), hi = __shfl_down((uint32_t)(x >> 32), off);
            x += ((unsigned long long)hi << 32) | lo;
        }
        if (__lane_id() == 0 && i != 6 && x) atomicAdd(a.counters + i, x);
    }
}

} // namespace sg
