// sg_geom.hpp -- dust geometry densities evaluated on the device (grid setup: cell densities and tree
// subdivision by density samples).
#pragma once
#include "sg_common.hpp"
#include "sg_host.hpp"

namespace sg {

// An exponential disc, optionally spiral-decorated, as the host sampler evaluates it (ExpDiskGeometry.cpp:103-114,
// SpiralStructureGeometryDecorator.cpp:165-172): normalisation and spiral constants come from the host.
struct DiscDev {
    double hR, hz, Rmax, zmax, Rmin, rho0;
    double m, tanp, R0, phi0, w, CN;
    int spiral, N;
};

__device__ inline double disc_density(const DiscDev& d, double x, double y, double z)
{
    double R = sqrt(x * x + y * y), absz = fabs(z);
    if (d.Rmax > 0.0 && R > d.Rmax) return 0.0;
    if (d.zmax > 0.0 && absz > d.zmax) return 0.0;
    if (R < d.Rmin) return 0.0;
    double rho = d.rho0 * exp(-R / d.hR) * exp(-absz / d.hz);
    if (d.spiral) {
        double phi = atan2(y, x);
        double gamma = log(R / d.R0) / d.tanp + d.phi0 + 0.5 * M_PI / d.m;
        rho *= (1.0 - d.w) + d.w * d.CN * pow(sin(0.5 * d.m * (gamma - phi)), (double)(2 * d.N));
    }
    return rho;
}

// The device record of an exponential-disc geometry record (16 doubles as the stellar components lay them out).
inline DiscDev disc_dev(const double* geom16)
{
    DiscDev d{};
    d.hR = geom16[1]; d.hz = geom16[2]; d.Rmax = geom16[3]; d.zmax = geom16[4]; d.Rmin = geom16[5];
    d.spiral = geom16[6] != 0;
    d.m = geom16[7]; d.R0 = geom16[9]; d.phi0 = geom16[10]; d.w = geom16[11]; d.N = (int)geom16[12];
    disc_density_constants(geom16, &d.rho0, &d.tanp, &d.CN);
    return d;
}

// The setup samplers' counter-based uniform triple (seed, item, sample): the same numbers as the host
// builder's stream, so that both place a node's or a cell's samples at the same positions.
__device__ __forceinline__ void setup_sample3(uint32_t seed, uint64_t item, uint32_t j, double u[3])
{
    uint32_t c0 = (uint32_t)item, c1 = (uint32_t)(item >> 32), c2 = j, c3 = 0x5e7u;
    philox10(c0, c1, c2, c3, seed, 0x9d2c5680u);
    const double s32 = 1.0 / 4294967296.0;
    u[0] = ((double)c0 + 0.5) * s32;
    u[1] = ((double)c1 + 0.5) * s32;
    u[2] = ((double)c2 + 0.5) * s32;
}

} // namespace sg
