// sg_pan.hpp -- panchromatic dust emission: the dust's emission spectra from its absorbed luminosity, and the
// launch of dust photon packets, whose transport is the stellar packets' (same kernels).
#pragma once
#include "sg_common.hpp"
#include "sg_mc.hpp"

namespace sg {

// NR::locate_clip over xv[0..n-1] (Fundamentals/NR.hpp:96-151): the largest j in [0, n-2] with xv[j] <= x
__device__ __forceinline__ int locate_clip_dev(const double* xv, int64_t n, double x)
{
    if (x < xv[0]) return 0;
    int64_t jl = -1, ju = n - 1;
    while (ju - jl > 1) {
        int64_t jm = (ju + jl) >> 1;
        if (x < xv[jm]) ju = jm;
        else jl = jm;
    }
    return (int)jl;
}

// A cell's box: Cartesian cell m = k + Nz*j + Nz*Ny*i from the border arrays, or a tree leaf's record.
struct CellBoxes {
    int kind;                         // GK_CART or GK_TREE
    const double *xv, *yv, *zv;
    int Nx, Ny, Nz;
    const TreeRec64* rec;
    __device__ __forceinline__ void box(int64_t m, double q[6]) const
    {
        if (kind == GK_CART) {
            int i = (int)(m / ((int64_t)Nz * Ny)), j = (int)((m / Nz) % Ny), k = (int)(m % Nz);
            q[0] = xv[i]; q[1] = yv[j]; q[2] = zv[k]; q[3] = xv[i + 1]; q[4] = yv[j + 1]; q[5] = zv[k + 1];
        } else {
            const TreeRec64& r = rec[m];
            for (int a = 0; a < 3; a++) { q[a] = r.lo[a]; q[a + 3] = r.hi[a]; }
        }
    }
};

// ------------------------------------------------------------------------- emission spectra

// The inputs of PanDustSystem::calculatedustemission with a grey-body emissivity, one component.
struct SpectraArgs {
    int nl, npop, nT;
    const double* lambda;    // nl (m)
    const double* dlambda;   // nl (m)
    const double* sigmaabs;  // npop x nl, per population
    const double* kappaabs;  // nl, the mix's absorption coefficient
    double mu;               // the mix's total mass per hydrogen atom
    const double* Tv;        // nT + 1 temperatures of the Planck table
    const double* planck;    // npop x (nT + 1) Planck-integrated absorption cross sections
    const double* labs_s;    // ncells x nl absorbed stellar luminosity ([m][ell])
    const double* labs_d;    // the same for dust emission, or null
    const double* rho;       // ncells
    double len2;             // (metres per grid length unit)^2
    int64_t ncells;
    double* dustlum;         // out: ncells x nl normalised emission spectrum
    double* labsbol;         // out: ncells bolometric absorbed luminosity
};

// PlanckFunction::operator() (PlanckFunction.cpp:21-30) with the constants of Units.cpp:17-19
__device__ __forceinline__ double planck_b(double lambda, double T)
{
    const double h = 6.62606957e-34, c = 2.99792458e8, k = 1.3806488e-23;
    double x = h * c / (lambda * k * T);
    return 2.0 * h * c * c / pow(lambda, 5) / (exp(x) - 1.0);
}

// Per cell: the mean intensity (DustSystem::meanintensityv, DustSystem.cpp:911-931), each population's
// equilibrium temperature (DustMix::equilibrium and invplanckabs, DustMix.cpp:674-697), the grey-body
// emissivity (GreyBodyDustEmissivity.cpp:20-38), converted to luminosities per bin and normalised
// (DustLib.cpp, one component), and the bolometric absorbed luminosity (PanDustSystem::Labsbolv: each
// table's row sum, then their sum).
__global__ void __launch_bounds__(256) k_dust_spectra(const SpectraArgs s, const CellBoxes cb)
{
    int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= s.ncells) return;
    double q[6];
    cb.box(m, q);
    // the grid's densities give kappa*rho per grid length unit, so kappa rho V in SI is kappa rho V len2
    const double vol = (q[3] - q[0]) * (q[4] - q[1]) * (q[5] - q[2]);
    const double fac = 4.0 * M_PI * vol * s.len2;
    const double* ls = s.labs_s + (size_t)m * s.nl;
    const double* ld = s.labs_d ? s.labs_d + (size_t)m * s.nl : nullptr;
    double T[8];
    for (int c = 0; c < s.npop; c++) {
        double planckabs = 0.0;
        for (int ell = 0; ell < s.nl; ell++) {
            double labs = ls[ell] + (ld ? ld[ell] : 0.0);
            double kappaabsrho = 0.0 + s.kappaabs[ell] * s.rho[m];
            double J = labs / (kappaabsrho * fac) / s.dlambda[ell];
            J = isfinite(J) ? J : 0.0;
            planckabs += s.sigmaabs[(size_t)c * s.nl + ell] * J * s.dlambda[ell];
        }
        const double* P = s.planck + (size_t)c * (s.nT + 1);
        int p = locate_clip_dev(P, s.nT + 1, planckabs);
        T[c] = s.Tv[p] + ((planckabs - P[p]) / (P[p + 1] - P[p])) * (s.Tv[p + 1] - s.Tv[p]);
    }
    double* out = s.dustlum + (size_t)m * s.nl;
    double total = 0.0;
    for (int ell = 0; ell < s.nl; ell++) {
        double ev = 0.0;
        for (int c = 0; c < s.npop; c++) ev += s.sigmaabs[(size_t)c * s.nl + ell] * planck_b(s.lambda[ell], T[c]);
        ev /= s.mu;
        double L = ev * s.dlambda[ell];
        out[ell] = L;
        total += L;
    }
    if (total > 0)
        for (int ell = 0; ell < s.nl; ell++) out[ell] /= total;
    double bs = 0.0, bd = 0.0;
    for (int ell = 0; ell < s.nl; ell++) bs += ls[ell];
    if (ld)
        for (int ell = 0; ell < s.nl; ell++) bd += ld[ell];
    s.labsbol[m] = bs + bd;
}

// ------------------------------------------------------------------------- dust packet launch

// The launch of PanMonteCarloSimulation::dodustemissionchunk (PanMonteCarloSimulation.cpp:273-318): a cell
// chosen from a uniform distribution for a fraction xi of the packets and from the cells' luminosities for
// the rest (NR::locate_clip on their normalised cumulative distribution), the weight that corrects for it, a
// uniform position in the cell's box (Random::position, Random.cpp:227-235) and an isotropic direction
// (Random::direction, :180-185). With xi = 0 it is the launch of dodustselfabsorptionchunk (:170-187).
struct DustLaunch {
    const double* cum;   // ncells + 1
    const double* Lv;    // ncells
    int64_t ncells;
    double xi, Lmean, Lem;
    CellBoxes cb;
};

template <typename R>
__global__ void __launch_bounds__(256) k_emit_dust(const McArgs a, const DustLaunch d, R* out, uint32_t* draws)
{
    const size_t n = a.count;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        PacketRng rng;
        rng.start(a.seed, a.member, a.first + i, a.ell);
        double X = rng.ud();
        int64_t m;
        if (X < d.xi) {
            int64_t c = (int64_t)(d.ncells * X / d.xi);
            m = c < 0 ? 0 : (c > d.ncells - 1 ? d.ncells - 1 : c);
        } else {
            m = locate_clip_dev(d.cum, d.ncells + 1, (X - d.xi) / (1 - d.xi));
        }
        double weight = 1.0 / (1 - d.xi + d.xi * d.Lmean / d.Lv[m]);
        double q[6];
        d.cb.box(m, q);
        double fx = rng.ud();
        double fy = rng.ud();
        double fz = rng.ud();
        double x = q[0] + fx * (q[3] - q[0]), y = q[1] + fy * (q[4] - q[1]), z = q[2] + fz * (q[5] - q[2]);
        double kx, ky, kz;
        iso_direction(rng, kx, ky, kz);
        double L = d.Lem * weight;
        if (!(L > 0)) { L = 0; x = y = z = kx = ky = kz = 0; }
        out[i] = (R)x; out[n + i] = (R)y; out[2 * n + i] = (R)z;
        out[3 * n + i] = (R)kx; out[4 * n + i] = (R)ky; out[5 * n + i] = (R)kz;
        out[6 * n + i] = (R)L;
        draws[i] = rng.draw;
    }
}

} // namespace sg
