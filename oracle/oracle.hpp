// oracle.hpp -- C API of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP implementation in
// skirt7_amd/: a CPU restatement of SKIRT7's Monte Carlo photon-packet transport loop.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product
// never links, loads or calls anything declared here.
//
// Parameter-array layouts shared with the product's C-ABI (include/skirt7_amd.h):
//   stellar component  (16 doubles): [0] kind (1 = exponential disc, 2 = Sersic sphere)
//                                    [1..5] disc: hR, hz, Rmax, zmax, Rmin | Sersic: n, reff
//                                    [6] spiral flag, [7..12] arms, pitch(rad), R0, phi0(rad), weight, index
//   frame instrument   (10 doubles): distance, inclination, azimuth, position angle (rad),
//                                    Nx, Ny, fovx, fovy, xc, yc
//   Monte Carlo params  (8 doubles): packages, minWeightReduction, minScattEvents, scattBias,
//                                    seed, emissionBias, member, reserved
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void* or_scene_create(void);
void  or_scene_destroy(void* s);

// grids (frozen inputs shared with the product)
int or_scene_set_cartesian(void* s, const double ext[6], const double* xv, int nxv, const double* yv,
                           int nyv, const double* zv, int nzv);
int or_scene_set_tree(void* s, int arity, const double ext[6], const uint8_t* subdivided, int64_t nnodes,
                      int search /*0 TopDown, 1 Neighbor, 2 Bookkeeping*/);
int or_scene_set_voronoi(void* s, const double ext[6], const double* sites, int64_t nsites,
                         const int64_t* nb_offsets, const int32_t* nb_ids);
int64_t or_scene_num_cells(void* s);

// medium, sources, instruments
int or_scene_set_dust(void* s, const double* rho, int64_t ncells, int nlambda, const double* kappaext,
                      const double* albedo, const double* asymmpar, int store_absorption);
int or_scene_set_dust_multi(void* s, int ncomp, const double* rho, int64_t ncells, int nlambda, const double* kext,
                            const double* ksca, const double* albedo, const double* g, int store_abs);
int or_scene_add_star(void* s, const double* params16, const double* lum, const double* table_s,
                      const double* table_M, int ntable);
int or_scene_add_instrument(void* s, const double* params10);

// DustGrid::path on one ray; returns number of segments (or -1 if more than maxseg)
int or_scene_path(void* s, const double rk[6], int maxseg, int32_t* m, double* ds);
int or_scene_whichcell(void* s, const double r[3]);
double or_scatter_phase(void* s, int ell, const double r[3], const double k[3], const double kobs[3]);
int or_random_mix(void* s, int ell, const double r[3], double u);
// DustSystem::opticaldepth on one ray at wavelength ell
double or_scene_opticaldepth(void* s, const double rk[6], int ell);
// DustGridPath::pathlength after fillOpticalDepth along one ray, for each of ntau targets
void or_scene_pathlength(void* s, const double rk[6], int ell, int ntau, const double* taus, double* out);

// Full transport loop, reference semantics: MT19937 streams per thread, chunk rule of
// MonteCarloSimulation::setChunkParams, dynamic chunk claiming. Returns 0 on success.
//   frames: sum over instruments of nlambda*Nx*Ny doubles, instrument-major, then (ell, pixel)
//   labs:   ncells*nlambda doubles, (m, ell) row-major; may be null when not stored
//   counters: [0] packets, [1] own-path crossings, [2] peel-off crossings, [3] scattering events,
//             [4] peel-offs, [5] stuck-path escapes, [6] Npp, [7] walks
int or_run_reference(void* s, const double* mc8, int threads, double* frames, double* labs,
                     uint64_t* counters);
// Same loop with the counter-based Philox streams of the product: packets [first, first+count)
// at wavelength ell with normalisation Npp.
int or_run_packets(void* s, const double* mc8, int ell, uint64_t first, uint64_t count, uint64_t npp,
                   int threads, double* frames, double* labs, uint64_t* counters);

// building blocks exposed for closed-form tests
double or_mt_first_uniforms(uint32_t seed, int n, double* out);   // Random::uniform
void   or_mt_first_words(uint32_t seed, int n, uint32_t* out);
void   or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_lambertw1(double z);
int    or_nr_locate_clip(const double* xv, int n, double x);
int    or_nr_locate_fail(const double* xv, int n, double x);
int    or_nr_locate_tau(const double* taus, int n, double x);
double or_nr_interp_loglog(double x, double x1, double x2, double f1, double f2);
void   or_sersic_table(double n, double* s101, double* M101);
void   or_hg_sample(double g, const double k[3], double u1, double u2, double out[3]);
double or_hg_value(double g, const double k[3], const double kobs[3]);
const char* or_last_error(void);
// sampling building blocks on the per-packet Philox streams keyed (seed, 0), packets [first, first+n)
int    or_star_positions(void* s, int comp, uint32_t seed, uint64_t first, uint64_t n, double* xyz);
int    or_launch_sample(void* s, int ell, double bias, uint32_t seed, uint64_t first, uint64_t n, int32_t* comp,
                        double* weight);
void   or_exponcutoff_sample(double xmax, uint32_t seed, uint64_t first, uint64_t n, double* out);
void   or_pixelondetector(const double* params10, int64_t n, const double* xyz, int32_t* pix, double* kobs);

// FitSKIRT evaluation (one stellar component)
//   GaussianKernel::setupSelfBefore (GaussianKernel.cpp:19-46) + ConvolutionKernel::setupSelfAfter
//   (ConvolutionKernel.cpp:23-28): dimension^2 doubles, pixel (x, y) at x + dimension*y
void   or_gaussian_kernel(double fwhm, int dimension, double* out);
//   Convolution::nested_loop (Convolution.cpp:34-66): image Nx x Ny -> out, kernel kx x ky
void   or_convolve_nested(const double* image, int Nx, int Ny, const double* kernel, int kx, int ky, double* out);
//   GoldenSection::optimize (GoldenSection.cpp:57-124): frame is updated as the reference returns it
void   or_golden_section(const double* ref, double* frame, int npix, double minlum, double maxlum, double* lum,
                         double* chi2);

int or_dust_spectra(void* s, const double* labs_s, const double* labs_d, int npop, const double* sigmaabs, double mu,
                    const double* kappaabs, const double* lambda, const double* dlambda, int nT, const double* Tv,
                    const double* planck, double len2, double* dustlum, double* labsbol);
int or_run_dust_packets(void* s, const double* mc8, int ell, uint64_t first, uint64_t count, uint64_t npp,
                        const double* Lv, double xi, int peeloffs, int store_abs, uint32_t stream, int threads,
                        double* frames, double* labs, uint64_t* counters);

#ifdef __cplusplus
}
#endif
