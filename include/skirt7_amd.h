/* skirt7_amd.h -- C-ABI of the MI355X photon-packet transport library.
 *
 * Drop-in boundary for SKIRT7's Monte Carlo photon-packet transport loop. Every entry point
 * replaces one reference interface (file:line relative to the SKIRT7 source tree) with a
 * batch-oriented call: a per-packet round trip to the GPU would cost more than the work.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative status on failure; sg_last_error()
 *    returns a thread-local message. No C++ exception crosses this boundary (the reference
 *    throws FatalError, e.g. DustSystem.cpp:953-955; the host wrapper rethrows its own).
 *  - Pointers are host pointers; sizes are element counts. Positions in the caller's length
 *    unit, luminosities in the caller's power unit, angles in radians.
 *  - Objects are immutable after creation except for their tallies and the in-place updates that
 *    pooled fit members take (sg_stars_update, sg_instruments_update, sg_grid_resample_density);
 *    one host thread drives one device stream (the reference's Parallel threads become device lanes).
 *  - precision: SG_FP64 reproduces the reference arithmetic (cell-index sequences bit-exact);
 *    SG_MIXED runs positions, directions, optical depths and weights in FP32, with FP64
 *    wall-distance recompute near ties and FP64 tallies.
 */
#ifndef SKIRT7_AMD_H
#define SKIRT7_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sg_grid sg_grid;            /* DustGrid                        */
typedef struct sg_dust sg_dust;            /* DustSystem (one dust component) */
typedef struct sg_stars sg_stars;          /* StellarSystem                   */
typedef struct sg_instruments sg_instruments;  /* InstrumentSystem of FrameInstruments */
typedef struct sg_sim sg_sim;              /* MonteCarloSimulation            */
typedef struct sg_sph sg_sph;              /* SPHDustDistribution's particles */

enum { SG_FP64 = 0, SG_MIXED = 1 };
enum { SG_TOPDOWN = 0, SG_NEIGHBOR = 1, SG_BOOKKEEPING = 2 };   /* TreeDustGrid::SearchMethod (TreeDustGrid.hpp:149) */

/* Number of doubles in the parameter records below. */
#define SG_STAR_NPARAM 16
#define SG_INSTR_NPARAM 10
#define SG_MC_NPARAM 8
#define SG_NCOUNTERS 15

/* ---- library ---------------------------------------------------------------------------- */
const char* sg_last_error(void);
/* Selects the device and checks it is a gfx950; every other call implies sg_init(0). */
int sg_init(int device);
int sg_device_name(char* buf, int len);

/* ---- DustGrid ----------------------------------------------------------------------------- */
/* CartesianDustGrid (CartesianDustGrid.cpp:28-43): box extent {xmin,ymin,zmin,xmax,ymax,zmax}
 * and the Nx+1, Ny+1, Nz+1 cell borders; cell index m = k + Nz*j + Nz*Ny*i (:326-329). */
int sg_grid_create_cartesian(const double ext[6], const double* xv, int nxv, const double* yv, int nyv,
                             const double* zv, int nzv, sg_grid** out);
/* OctTreeDustGrid (arity 8) or BinTreeDustGrid with alternating directions (arity 2)
 * (TreeDustGrid.cpp:49-162): one subdivision flag per node in the reference's breadth-first
 * node numbering; cell numbers are the leaves' ranks in that order (:110-121). */
int sg_grid_create_tree(int arity, const double ext[6], const uint8_t* subdivided, int64_t nnodes, int search,
                        sg_grid** out);
/* VoronoiDustGrid / VoronoiMesh (VoronoiMesh.cpp:297-405): generator sites (x,y,z per site)
 * and per-site neighbour lists in CSR form; walls are the ids -1..-6 (xmin,xmax,ymin,ymax,zmin,zmax). */
int sg_grid_create_voronoi(const double ext[6], const double* sites, int64_t nsites, const int64_t* nb_offsets,
                           const int32_t* nb_ids, sg_grid** out);
int sg_grid_destroy(sg_grid* g);
int sg_grid_num_cells(const sg_grid* g, int64_t* ncells);
/* DustGrid::whichcell (DustGrid.hpp:89) for n positions (x,y,z); -1 outside. */
int sg_grid_whichcell(sg_grid* g, int64_t n, const double* r, int32_t* m);
/* DustGrid::path (DustGrid.hpp:106) for n rays (rx,ry,rz,kx,ky,kz): each ray's (m, ds)
 * segment list exactly as DustGridPath holds it, including the -1 entry segments. Ray i writes
 * its segments at [i*max_seg, ...); counts[i] = number of segments, or -1 if more than max_seg. */
int sg_grid_paths(sg_grid* g, int64_t n, const double* rk, int max_seg, int precision, int32_t* counts,
                  int32_t* m, double* ds);

/* ---- DustSystem ----------------------------------------------------------------------------- */
/* Dust density per cell (DustSystem::_rhovv, DustSystem.hpp:412) and the dust mix per
 * wavelength: kappaext, albedo and asymmetry parameter as DustMix derives them
 * (DustMix.cpp:44-91, SimpleOligoDustMix.cpp:22-51). store_absorption mirrors
 * OligoDustSystem::storeabsorptionrates (OligoDustSystem.cpp:61-64). */
int sg_dust_create(sg_grid* g, const double* rho, int64_t ncells, int nlambda, const double* kappaext,
                   const double* albedo, const double* asymmpar, int store_absorption, sg_dust** out);
int sg_dust_destroy(sg_dust* d);
/* A dust system of ncomp components (DustSystem with Ncomp > 1, DustSystem.hpp:412; CompDustDistribution):
 * rho holds each component's cell densities (ncomp x ncells, [h][m]); kappaext, kappasca, albedo and
 * asymmpar are ncomp x nlambda ([h][ell]) as each component's DustMix derives them (DustMix.cpp:44-91). Walks
 * see KappaRho, the sum over components of kappaext(h) rho(m, h) (DustSystem.cpp:459-485); escape and
 * absorption use each cell's own albedo (MonteCarloSimulation.cpp:463-508); scattering peel-offs weight the
 * components' phase functions by kappasca(h) rho(m, h) (:318-343) and scattering draws the component with
 * those weights (DustSystem::randomMixForPosition, DustSystem.cpp:855-869). Such systems run in the lane
 * pipeline (sg_sim_run and batched runs; sg_sim_run_static refuses them). ncomp == 1 is sg_dust_create. */
int sg_dust_create_multi(sg_grid* g, int ncomp, const double* rho, int64_t ncells, int nlambda, const double* kappaext,
                         const double* kappasca, const double* albedo, const double* asymmpar, int store_absorption,
                         sg_dust** out);
/* DustSystem::opticaldepth (DustSystem.cpp:960-976) along n rays at wavelength ell. */
int sg_dust_optical_depths(sg_dust* d, int ell, int64_t n, const double* rk, int precision, double* tau);
/* DustGridPath::pathlength after fillOpticalDepth (DustGridPath.cpp:162-173): for ray i and
 * optical depth taus[i], the path length s[i]. */
int sg_dust_path_lengths(sg_dust* d, int ell, int64_t n, const double* rk, const double* taus, int precision,
                         double* s);
/* OligoDustSystem::Labs (OligoDustSystem.cpp:82-85): the absorbed-luminosity table,
 * ncells x nlambda doubles in the reference's (m, ell) row-major order. */
int sg_dust_read_absorption(sg_dust* d, double* labs);
int sg_dust_clear_absorption(sg_dust* d);

/* ---- StellarSystem ------------------------------------------------------------------------ */
/* Stellar components (StellarSystem.cpp:116-158, GeometricStellarComp.cpp:75-81): params is
 * ncomp x SG_STAR_NPARAM doubles
 *   [0] geometry: 1 exponential disc (ExpDiskGeometry.cpp:117-146), 2 Sersic (SersicGeometry.cpp:85-90)
 *   [1..5] disc: hR, hz, Rmax, zmax, Rmin | Sersic: n, reff
 *   [6] spiral decorator flag, [7..12] arms, pitch, R0, phi0, weight, index
 *       (SpiralStructureGeometryDecorator.cpp:23-40,176-229)
 * lum is ncomp x nlambda luminosities; sersic tables (101 radii and normalised masses per
 * Sersic component, SersicFunction.cpp:15-68) are concatenated in component order. */
int sg_stars_create(int ncomp, const double* params, int nlambda, const double* lum, const double* sersic_s,
                    const double* sersic_M, int ntable, double emission_bias, sg_stars** out);
int sg_stars_destroy(sg_stars* s);
/* The same setup into an existing stellar system (same component count): FitSKIRT evaluates each individual
 * with its own stellar parameters (AdjustableSkirtSimulation::performWith, AdjustableSkirtSimulation.cpp:119-141,
 * rebuilds the whole simulation); here a pooled member takes the new records in place. */
int sg_stars_update(sg_stars* s, int ncomp, const double* params, const double* lum, const double* sersic_s,
                    const double* sersic_M, int ntable, double emission_bias);
/* SersicFunction table helper (SersicFunction.cpp:15-68): 101 radii s and masses M. */
int sg_sersic_table(double n, double* s101, double* M101);

/* ---- InstrumentSystem (FrameInstrument) ------------------------------------------------------ */
/* n x SG_INSTR_NPARAM: distance, inclination, azimuth, position angle, Nx, Ny, fovx, fovy, xc, yc
 * (DistantInstrument.cpp:27-50, SingleFrameInstrument.cpp:26-42). */
int sg_instruments_create(int n, const double* params, int nlambda, sg_instruments** out);
int sg_instruments_destroy(sg_instruments* in);
/* New parameter records for an existing instrument system, as sg_instruments_create reads them (the pixel
 * counts stay; the frames keep their contents): a pooled fit member's viewing angles (FitSKIRT, see
 * sg_stars_update). */
int sg_instruments_update(sg_instruments* in, const double* params);
/* FrameInstrument::_distftotv (FrameInstrument.cpp:33-47): instrument i's raw detected
 * luminosities, nlambda x Ny x Nx doubles ([ell][j][i] = reference's ell*Nframep + i + Nx*j). */
int sg_instruments_read(sg_instruments* in, int i, double* frame);
int sg_instruments_clear(sg_instruments* in);

/* ---- MonteCarloSimulation ----------------------------------------------------------------- */
/* mc is SG_MC_NPARAM doubles: packages, minWeightReduction, minScattEvents, scattBias, seed,
 * emissionBias (informational: the stellar system's own value is used), member, tau_cut
 * (MonteCarloSimulation.hpp:37-66). tau_cut > 0 ends a peel-off walk
 * once its optical depth exceeds tau_cut (its contribution is below exp(-tau_cut) of L); 0 keeps
 * the reference's full walk. dust may be null (no dust system). */
int sg_sim_create(sg_dust* dust, sg_stars* stars, sg_instruments* instr, const double* mc, sg_sim** out);
int sg_sim_destroy(sg_sim* s);
/* MonteCarloSimulation::dostellaremissionchunk (MonteCarloSimulation.cpp:259-297) for packets
 * [first, first+count) at wavelength ell, each with luminosity L(ell)/npp. Packet p draws its
 * random numbers from its own Philox4x32-10 stream keyed (seed, member) and counted by
 * (p, draw, ell), so results do not depend on scheduling. Accumulates into the instrument and
 * absorption tallies; counters (may be null) receive SG_NCOUNTERS values:
 *   [0] packets [1] own-path crossings [2] peel-off crossings [3] scattering events
 *   [4] peel-offs detected [5] stuck-path escapes [6] npp [7] walks [8] re-walk crossings
 *   [9] FP64 wall recomputes [10] device microseconds (emission and transport kernels)
 *   [11] transport workgroups per launch [12] transport microseconds (device span of the transport
 *   passes) [13] transport (walk-kernel) launches [14] walk-kernel microseconds: the sum over those
 *   launches of each launch's own duration (HIP events on its stream; launches on two streams overlap) */
int sg_sim_run(sg_sim* s, int ell, uint64_t first, uint64_t count, uint64_t npp, int precision,
               uint64_t* counters);
/* sg_sim_run that allocates no device memory. No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same packet histories and tallies as sg_sim_run (event/walk kernels on one slot group of up to
 * 2^15 in-flight packets, one stream), with all work memory in a static 64 MiB device array that each
 * GPU the library runs on holds until the process ends, set up by sg_init (call it first). At most
 * 2^19 packets per call (the array holds their launch states): larger ranges are refused with an
 * error, split them over several calls. Calls that overlap on one GPU race (they share the array);
 * each call clears what it needs. counters as sg_sim_run, without the timings [10], [12]-[14]. */
int sg_sim_run_static(sg_sim* s, int ell, uint64_t first, uint64_t count, uint64_t npp, int precision,
                      uint64_t* counters);
/* Batched population (one persistent launch over many members, each its own simulation on a
 * grid of the same kind): member i runs packets [0, counts[i]) at wavelength ells[i]. With the
 * event/walk pipeline selected, members of at least sg_set_batch_alone packets run one at a time
 * through it and the rest share the persistent launches; each member's result is its single run. */
int sg_sim_run_batch(int nmembers, sg_sim* const* sims, const int* ells, const uint64_t* counts, int precision,
                     uint64_t* counters);
/* ---- PanMonteCarloSimulation: dust emission (PanMonteCarloSimulation.cpp:106-340) ------------------- */
/* PanDustSystem::calculatedustemission for one dust component with a grey-body emissivity: per cell the
 * mean intensity from the absorbed luminosity (DustSystem::meanintensityv, DustSystem.cpp:911-931; rho is the
 * component's density, kappa*rho per grid length unit, length_unit the grid unit in metres), each population's
 * equilibrium temperature by inverting its Planck-integrated absorption table (DustMix::equilibrium and
 * invplanckabs, DustMix.cpp:674-697; Tv has nT+1 temperatures, planckabs npop x (nT+1) values), the emissivity
 * sum(sigmaabs B(lambda, T)) / mu (GreyBodyDustEmissivity.cpp:20-38) times the bin widths, normalised
 * (DustLib.cpp): dustlum is ncells x nlambda ([m][ell]). labsbol receives PanDustSystem::Labsbolv (the row sums
 * of labs_stellar and of labs_dust, which may be null). Tables are ncells x nlambda as sg_dust_read_absorption
 * returns them; wavelengths and widths in metres. Cartesian and tree grids. */
int sg_dust_emission_spectra(sg_dust* d, const double* rho, const double* labs_stellar, const double* labs_dust,
                             int npop, const double* sigmaabs, double mu, const double* kappaabs, const double* lambdas,
                             const double* dlambdas, int nT, const double* Tv, const double* planckabs,
                             double length_unit, double* dustlum, double* labsbol);
/* Dust photon packets [first, first+count) at wavelength ell emitted by the cells with luminosities Lv
 * (ncells values: Labsbol times the cell's spectrum at ell), each with luminosity sum(Lv)/npp times its bias
 * weight: PanMonteCarloSimulation::dodustemissionchunk (PanMonteCarloSimulation.cpp:246-340; the cell drawn
 * uniformly for a fraction `bias` of the packets, PanDustSystem::emissionBias) with emission and scattering
 * peel-offs, or dodustselfabsorptionchunk (:160-204: bias 0, peeloffs 0, store_absorption 1). The packets'
 * transport is the stellar packets' (lane pipeline). Their random streams are keyed by (seed, member ^
 * stream * 0x9E3779B9): give each emission phase or self-absorption cycle its own stream. One dust component,
 * Cartesian and tree grids. counters as sg_sim_run without the timings. */
int sg_sim_run_dust(sg_sim* s, int ell, uint64_t first, uint64_t count, uint64_t npp, const double* Lv, double bias,
                    int peeloffs, int store_absorption, uint32_t stream, int precision, uint64_t* counters);

/* ---- FitSKIRT evaluation of a population ------------------------------------------------------ */
/* GaussianKernel (GaussianKernel.cpp:19-46) normalised to unit sum (ConvolutionKernel.cpp:23-28):
 * dimension x dimension doubles, pixel (x, y) at x + dimension*y. */
int sg_gaussian_kernel(double fwhm, int dimension, double* out);
/* ReferenceImage::chi2 for one stellar component (ReferenceImage.cpp:95-114) over n members: frame
 * ell of instrument `instrument` of each member's instrument system (raw detected luminosities, as
 * sg_instruments_read returns them) is convolved with the kx x ky kernel (Convolution::nested_loop,
 * Convolution.cpp:34-66) and its luminosity scale in [minlum, maxlum] is fitted to the reference
 * image ref (Ny x Nx) by golden section (GoldenSection.cpp:57-124). lum[i] and chi2[i] receive the
 * fit; convolved (may be null) receives the n convolved frames with the pixels where ref is zero set
 * to zero, as GoldenSection::optimize returns them. */
int sg_fit_frames(int n, sg_instruments* const* instr, int instrument, int ell, const double* kernel, int kx,
                  int ky, const double* ref, double minlum, double maxlum, double* lum, double* chi2,
                  double* convolved);

/* ---- FitSKIRT's genetic algorithm (Optimization) --------------------------------------------- */
/* Optimization::setupSelfBefore (FitSKIRTcore/Optimization.cpp:42-90): GAlib's steady-state GA over one
 * real gene per parameter range [lower[k], upper[k]], population popsize, generations generations,
 * mutation and crossover probabilities pmut and pcross, minimising the score. seed is the GAlib seed
 * (OligoFitScheme's fixedSeed uses 4357, Optimization.cpp:156-163; 0 takes one from the clock). The
 * algorithm makes GAlib's random draws and choices, so a fit repeats the reference's sequence of
 * individuals for the same seed and the same scores. Host code; it allocates no device memory. */
typedef struct sg_ga sg_ga;
int sg_ga_create(int ngenes, const double* lower, const double* upper, int popsize, int generations, double pmut,
                 double pcross, uint32_t seed, sg_ga** out);
int sg_ga_destroy(sg_ga* ga);
/* Optimization::initialize and Optimization::step (Optimization.cpp:156-163,260-263) up to the evaluation
 * of the population: afterwards sg_ga_pending lists the individuals to evaluate. */
int sg_ga_initialize(sg_ga* ga);
int sg_ga_step(sg_ga* ga);
/* The individuals that Optimization::PopEvaluate would evaluate (Optimization.cpp:301-333), in its order:
 * *n of them, their genes (n x ngenes doubles) into genes when it is not null. */
int sg_ga_pending(sg_ga* ga, int* n, double* genes);
/* Their scores (chi2, Optimization.cpp:345-357), in the same order; completes initialize or step. */
int sg_ga_set_scores(sg_ga* ga, int n, const double* scores);
/* Generations completed and Optimization::done (Optimization.cpp:149-152). */
int sg_ga_state(sg_ga* ga, int* generation, int* done);
/* The best individual found (GAStatistics::bestIndividual) and its score. */
int sg_ga_best(sg_ga* ga, double* genes, double* score);
/* The population (popsize x ngenes genes and popsize scores) in GAlib's population order. */
int sg_ga_population(sg_ga* ga, double* genes, double* scores);

/* Transport organisation used by sg_sim_run (both give the same per-packet histories):
 * 1 (default) alternating kernels -- an event kernel over in-flight packet slots that runs the
 *   packet state machine and queues walks, and a persistent walk kernel over that queue;
 * 0 one persistent kernel whose lanes run the state machine and the walks in turn. */
int sg_set_pipeline(int mode);
/* Launch tuning used by sg_sim_run (0 = keep current, negative = back to the default): block size
 * (default 256), waves per CU (default: as many as fit), compaction threshold (lanes that must
 * finish their walk before the wave re-arms them; by default 32 for the lane pipeline and 24 for
 * the walk kernel). */
int sg_set_tuning(int block, int waves_per_cu, int compact_lanes);
/* In-flight packet slots of the event/walk pipeline, split over its two slot groups (0 = the
 * default 2^22; otherwise 1024..2^24). No reference counterpart: SKIRT7's chunk size
 * (MonteCarloSimulation.cpp:68-114) sets the CPU work granularity, this the GPU's. */
int sg_set_slots(int64_t slots);
/* Launch-parameter autotuner (no reference counterpart; the GPU analogue of the chunk rule of
 * MonteCarloSimulation::setChunkParams, MonteCarloSimulation.cpp:68-114): times sg_sim_run on the
 * sample of packets [first, first+sample) over the compaction threshold (8..48 lanes), then the
 * in-flight slots (2^20..2^24 up to the sample size, event/walk pipeline) or waves per CU (lane
 * pipeline), then the block
 * size (64, 128, 256), keeping the fastest of each in turn, and leaves them set for later runs.
 * With choose_pipeline nonzero it first times both transport organisations of sg_set_pipeline with
 * their default parameters and keeps the faster; the event/walk pipeline also tries the three walk
 * orders (sg_set_walk_order). best (may be null) receives {block, slots or waves per CU, compaction threshold,
 * pipeline, walk order}; rate (may be null) the sample's packets per second with
 * them. The results of a run do not depend on these choices. The tallies the sample adds are
 * removed again (frames and absorption as before). */
int sg_autotune(sg_sim* s, int ell, uint64_t first, uint64_t sample, int precision, int choose_pipeline,
                int64_t* best, double* rate);
/* Smallest member (packets) that sg_sim_run_batch runs alone through the event/walk pipeline
 * instead of in the shared persistent launch (0 = never; default 2^21). */
int sg_set_batch_alone(uint64_t min_packets);
/* Mixed-precision walks of trees built with the TopDown search method: 0 (default) follow the ropes
 * between leaves, as for Neighbor trees (the same cells: the search method only decides how the next
 * cell is found, TreeDustGrid.cpp:371-643); 1 descend from the root at every step as TopDown does.
 * FP64 walks always use the tree's own method. */
int sg_set_mixed_topdown(int descend);
/* Order in which the walk kernel of the event/walk pipeline takes a pass's walks: 1 sorted by a
 * coherence key (peel-offs by instrument and pixel, other walks by start point), 2 sorted and cut into
 * eight contiguous segments, each workgroup taking walks from its XCD's segment first (so that walks near
 * each other in the order share that XCD's L2), 0 (default) as appended. The order changes no packet
 * history, only the order of the tally additions. No counterpart in the reference, whose threads walk
 * one packet at a time (MonteCarloSimulation.cpp:259-297). */
int sg_set_walk_order(int order);
/* Slot groups of the event/walk pipeline: 2 (default) alternate on two streams, so that one group's
 * event pass and the drain of its walk pass overlap the other group's walks; 1 runs every pass on one
 * stream, one kernel at a time (each walk launch then has the device to itself, which is how its own
 * duration is measured). No counterpart in the reference. */
int sg_set_slot_groups(int groups);
/* Packets whose launch states sg_sim_run emits at once (0: the default, 2^25 for the event/walk pipeline and
 * 2^22 for the lane pipeline; otherwise at least 64): a longer range is transported batch by batch, with the
 * same packet histories. No counterpart in the reference. */
int sg_set_emit_batch(uint64_t packets);

/* ---- setup helpers (the reference's own setup phase, made reproducible) ---------------------- */
/* Normalised density of an exponential-disc geometry record laid out like a stellar component
 * (ExpDiskGeometry.cpp:25-40,103-114; spiral: SpiralStructureGeometryDecorator.cpp:165-172). */
int sg_geometry_density(const double* geom16, int64_t n, const double* xyz, double* out);
/* TreeDustGrid::setupSelfBefore/subdivide (TreeDustGrid.cpp:49-230), mass-fraction criterion,
 * with node masses from nsamples counter-based density samples: one flag per node in
 * breadth-first id order (at most cap), the node count in *nnodes. */
int sg_tree_subdivide(int arity, const double ext[6], const double* geom16, double scale, double total_mass,
                      int minlevel, int maxlevel, double maxmassfraction, int nsamples, uint32_t seed,
                      uint8_t* flags, int64_t cap, int64_t* nnodes);
/* The same for a CompDustDistribution of ncomp components (the density is the sum over components of
 * scales[h] times geometry h's density; geoms16 holds ncomp records of 16, CompDustDistribution::density),
 * with all three stopping criteria of TreeDustGrid::subdivide (TreeDustGrid.cpp:194-220): a node below
 * maxlevel is split when no criterion is set (all three 0), or when its mass fraction reaches
 * maxmassfraction, its optical depth kappaV M / V^(2/3) (kappaV 2600 m2/kg) reaches maxopticaldepth or its
 * samples' density dispersion (max - min) / max reaches maxdensdisp (criteria at 0 are off). */
int sg_tree_subdivide_multi(int arity, const double ext[6], int ncomp, const double* geoms16, const double* scales,
                            double total_mass, int minlevel, int maxlevel, double maxmassfraction,
                            double maxopticaldepth, double maxdensdisp, int nsamples, uint32_t seed, uint8_t* flags,
                            int64_t cap, int64_t* nnodes);
/* The same subdivision with node masses from a density field given per voxel of a regular
 * nx x ny x nz grid over ext (value [i][j][k] at i*ny*nz + j*nz + k, constant inside a voxel): exact mass in
 * each node's box, the total being the field's mass. */
int sg_tree_subdivide_field(int arity, const double ext[6], const double* field, int nx, int ny, int nz,
                            int minlevel, int maxlevel, double maxmassfraction, uint8_t* flags, int64_t cap,
                            int64_t* nnodes);
/* The same with the optical-depth criterion (TreeNodeBoxDensityCalculator::opticalDepth); a field's exact
 * node masses give no density dispersion, as in the reference. */
int sg_tree_subdivide_field_od(int arity, const double ext[6], const double* field, int nx, int ny, int nz,
                               int minlevel, int maxlevel, double maxmassfraction, double maxopticaldepth,
                               uint8_t* flags, int64_t cap, int64_t* nnodes);
/* Cell densities from such a voxel field: mass in the cell's box over its volume, the exact mass-in-box
 * densities of TreeDustGrid's particle path (TreeDustGrid.cpp:647-660). Cartesian and tree grids. */
int sg_grid_field_density(sg_grid* g, const double* field, int nx, int ny, int nz, double* rho_out);
/* DustSystem::setSampleDensityBody (DustSystem.cpp:145-172): mean of nsamples densities per cell
 * (Cartesian and tree grids), times scale. */
int sg_grid_sample_density(sg_grid* g, const double* geom16, double scale, int nsamples, uint32_t seed,
                           double* rho_out);
/* The same sampling on the device, stored as a Cartesian grid's densities in place of the current ones
 * (the positions are the same counter-based draws; densities agree with sg_grid_sample_density to rounding
 * of the device's exp/log/pow). Exponential-disc geometries (spiral-decorated or not); queued on the
 * default stream without waiting: a FitSKIRT member's new dust density without a host round trip. */
int sg_grid_resample_density(sg_grid* g, const double* geom16, double scale, int nsamples, uint32_t seed);
/* ---- device grid builders and the SPH particle medium ---------------------------------------- */
/* SPHDustDistribution::setupSelfBefore (SPHDustDistribution.cpp:38-100): n rows of x y z h M Z T (lengths
 * in the grid's unit, masses in the caller's mass unit: the reference's pc and Msun factors are the
 * caller's); a particle with T > Tmax (both positive) is ignored; a negative total gas or metal mass
 * suppresses all dust; dust fraction fdust > 0. Builds the 20^3 particle grid and each cell's list
 * (SPHGasParticleGrid.cpp:94-132) on the host and holds the particles on the device. */
int sg_sph_create(const double* rows, int64_t n, double fdust, double Tmax, sg_sph** out);
int sg_sph_destroy(sg_sph* s);
/* Particles kept, SPHDustDistribution::mass (SPHDustDistribution.cpp:248-256), the grid's 3 x 21 borders
 * (infinite at both ends; any pointer may be null) and the total number of list entries. */
int sg_sph_info(const sg_sph* s, int64_t* count, double* mass, double* borders, int64_t* list_entries);
/* Particle grid cell (i, j, k)'s list (particle indices in the order kept): its length in *count and up
 * to cap entries in out. */
int sg_sph_cell_list(const sg_sph* s, int i, int j, int k, int32_t* out, int64_t cap, int64_t* count);
/* SPHDustDistribution::massInBox (SPHDustDistribution.cpp:226-236) of n boxes (xmin ymin zmin xmax ymax
 * zmax), on the device. */
int sg_sph_mass_in_box(const sg_sph* s, const double* boxes, int64_t n, double* out);
/* SPHDustDistribution::density (SPHDustDistribution.cpp:192-202) at n positions, on the device. */
int sg_sph_density(const sg_sph* s, const double* pos, int64_t n, double* out);
/* TreeDustGrid::setupSelfBefore / subdivide (TreeDustGrid.cpp:49-230) over the particle set, built on the
 * device level by level: node masses from the mass in the node's box (TreeNodeBoxDensityCalculator), or,
 * when maxdensdisp > 0 (TreeDustGrid.cpp:70), from nsamples density samples per node at the counter-based
 * positions of sg_tree_subdivide_multi; the three stopping criteria as there. Flags as
 * sg_tree_subdivide; masses (may be null, else cap entries) receives each node's mass: with masses in box,
 * every node beyond minlevel (those at maxlevel too, for the cells' densities), the reference's sum term
 * for term; with samples, the nodes the criteria evaluated. -1 elsewhere. */
int sg_tree_subdivide_sph(int arity, const double ext[6], const sg_sph* s, int minlevel, int maxlevel,
                          double maxmassfraction, double maxopticaldepth, double maxdensdisp, int nsamples,
                          uint32_t seed, uint8_t* flags, double* masses, int64_t cap, int64_t* nnodes);
/* sg_tree_subdivide_multi on the device (exponential-disc geometries): the same sample positions, the
 * densities to the rounding of the device's exp/log/pow, the same flags. */
int sg_tree_subdivide_device(int arity, const double ext[6], int ncomp, const double* geoms16, const double* scales,
                             double total_mass, int minlevel, int maxlevel, double maxmassfraction,
                             double maxopticaldepth, double maxdensdisp, int nsamples, uint32_t seed, uint8_t* flags,
                             int64_t cap, int64_t* nnodes);
/* TreeDustGrid::density (TreeDustGrid.cpp:655-660) over the particle set: each leaf's mass in its box over
 * its volume, for every cell of a tree grid. node_masses (nnodes entries, may be null) are the masses that
 * sg_tree_subdivide_sph returned for this tree; without them the masses are evaluated on the device from
 * the particle grid's lists. */
int sg_grid_sph_density(sg_grid* g, const sg_sph* s, const double* node_masses, int64_t nnodes, double* rho_out);
/* sg_grid_sample_density on the device (exponential-disc geometries): the same positions, the densities to
 * the rounding of the device's exp/log/pow. Cartesian and tree grids. */
int sg_grid_sample_density_device(sg_grid* g, const double* geom16, double scale, int nsamples, uint32_t seed,
                                  double* rho_out);
/* The distance from each of n points to its k-th nearest other point (k <= 64), on the device: the
 * smoothing lengths of a particle set generated for a workload (no counterpart in the reference, which
 * reads them from the particle file). */
int sg_knn_distance(const double* pos, int64_t n, int k, double* out);
/* Voronoi neighbour lists of n sites in the box ext (VoronoiMesh::buildMesh, VoronoiMesh.cpp:297-405):
 * offsets[n+1] and at most cap ids (walls -1..-6); *nids receives the total; -5 if cap is too small. */
int sg_voronoi_neighbors(const double ext[6], const double* sites, int64_t n, int64_t* offsets, int32_t* ids,
                         int64_t cap, int64_t* nids);

/* ---- output --------------------------------------------------------------------------------- */
/* FITSInOut::write (FITSInOut.cpp:32-95): one primary image of 32-bit floats, nx x ny (x nz frames when
 * nz > 1) from data[x + nx*y + nx*ny*z], with BSCALE, BZERO, DATE, ORIGIN, BUNIT and the CRPIXn /
 * CRVALn / CDELTn / CTYPEn keywords (reference pixels at the image centre, increments incx, incy,
 * centre values xc, yc in xyunits). stamp: the DATE value ("yyyy-mm-ddThh:mm:ss"), or null for
 * the current UTC time. An existing file is replaced. */
int sg_fits_write(const char* path, const double* data, int nx, int ny, int nz, double incx, double incy, double xc,
                  double yc, const char* dataunits, const char* xyunits, const char* stamp);

#ifdef __cplusplus
}
#endif
#endif
