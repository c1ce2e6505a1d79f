"""Synthetic simulation setups of the benchmark configurations (all lengths in kpc, angles in radians).

disc_cartesian: plain FP64 reference case -- stellar exponential disc (hR 4, hz 0.4) plus Sersic n=4 bulge
         (reff 1, B/T 0.2), dust exponential disc (hR 6, hz 0.25, face-on tau_V 1), HG g=0.6 albedo 0.6,
         Cartesian 50^3 over +-20, three 256x256 frames over 40 kpc at inclinations 0/60/88 deg.
spiral_octree: the same galaxy with a two-arm logarithmic spiral (pitch 20 deg, weight 0.3) on stars and dust,
         face-on tau in {0.5, 2, 5}, octree to level 8 by dust mass fraction 1e-6, four 512x512 frames at
         0/45/75/89 deg.
fit_population: a population of small Cartesian galaxies run in one batched launch, each fitted to a
         Poisson-noised target by PSF convolution and golden section (FitSKIRT's evaluation step).
Sizes can be scaled down for tests; every array a run depends on is kept in the returned dict so that
the CPU oracle can be given the same inputs.
"""
import math

import numpy as np

from . import skirt as S
from .capi import SG_NEIGHBOR, SG_TOPDOWN

EXTENT = np.array([-20.0, -20.0, -20.0, 20.0, 20.0, 20.0])
DISTANCE = 1.0e4   # 10 Mpc in kpc


def _galaxy(spiral=False, dust_tau=1.0, bt=0.2, ltot=1.0e10, dust_hR=6.0, dust_hz=0.25):
    disk = S.ExpDiskGeometry(4.0, 0.4)
    dgeom = S.ExpDiskGeometry(dust_hR, dust_hz)
    if spiral:
        disk = S.SpiralStructureGeometryDecorator(disk, arms=2, pitch=math.radians(20.0), radius=4.0, phase=0.0,
                                                  perturbWeight=0.3, index=1)
        dgeom = S.SpiralStructureGeometryDecorator(dgeom, arms=2, pitch=math.radians(20.0), radius=4.0, phase=0.0,
                                                   perturbWeight=0.3, index=1)
    comps = [S.OligoStellarComp(disk, [ltot * (1.0 - bt)])]
    if bt > 0:
        comps.append(S.OligoStellarComp(S.SersicGeometry(4.0, 1.0), [ltot * bt]))
    mix = S.SimpleOligoDustMix([1.0], [0.6], [0.6])
    # face-on optical depth of the dust disc; with the spiral, that of the underlying disc (the perturbation
    # averages to one over azimuth). SKIRT7's FaceOnDustCompNormalization refuses the decorator itself, so a
    # ski file of this galaxy states the same normalisation as a dust mass.
    scale = S.faceon_normalization(dgeom.geometry if spiral else dgeom, dust_tau, mix.kappaext[0])
    return comps, dgeom, mix, scale


def _frames(incl_deg, pixels, nlambda=1):
    ins = [S.FrameInstrument(DISTANCE, math.radians(a), 0.0, 0.0, pixels, pixels, 40.0, 40.0) for a in incl_deg]
    return S.InstrumentSystem(ins, nlambda)


def _finish(d, packages, seed, tau_cut=0.0):
    d["stars"] = S.StellarSystem(d["comps"], emissionBias=0.5)
    d["dust"] = S.DustSystem(d["grid"], d["rho"], d["mix"], writeMeanIntensity=d.get("store_abs", False))
    d["sim"] = S.OligoMonteCarloSimulation(d["dust"], d["stars"], d["instr"], packages=packages,
                                           minWeightReduction=1e4, minScattEvents=0, scattBias=0.5, seed=seed,
                                           tauCut=tau_cut)
    d["packages"] = packages
    return d


def disc_cartesian(grid_n=50, pixels=256, packages=1e6, seed=4357, nsamples=100, tau=1.0):
    comps, dgeom, mix, scale = _galaxy(False, tau)
    grid = S.CartesianDustGrid.uniform(EXTENT, grid_n, grid_n, grid_n)
    rho = S.sample_density(grid, dgeom, scale, nsamples, seed)
    d = dict(name="disc_cartesian", kind="cartesian", grid=grid, rho=rho, mix=mix, comps=comps, dgeom=dgeom,
             instr=_frames((0.0, 60.0, 88.0), pixels))
    return _finish(d, packages, seed)


def spiral_octree(tau=2.0, maxlevel=8, pixels=512, packages=1e8, seed=4357, nsamples=100, search=SG_NEIGHBOR,
            maxmassfraction=1e-6, tau_cut=0.0, device_setup=True):
    """Octree over the spiral galaxy's dust; the subdivision and the cells' density samples on the device
    (device_setup, the default) or on the host, at the same sample positions."""
    comps, dgeom, mix, scale = _galaxy(True, tau)
    build = S.OctTreeDustGrid.subdivide_device if device_setup else S.OctTreeDustGrid.subdivide
    grid = build(EXTENT, dgeom, scale, total_mass=scale, minlevel=2, maxlevel=maxlevel,
                 maxmassfraction=maxmassfraction, nsamples=nsamples, seed=seed, search=search)
    rho = S.sample_density(grid, dgeom, scale, nsamples, seed, device=device_setup)
    d = dict(name="spiral_octree", kind="octree", grid=grid, rho=rho, mix=mix, comps=comps, dgeom=dgeom,
             instr=_frames((0.0, 45.0, 75.0, 89.0), pixels))
    return _finish(d, packages, seed, tau_cut)


def fit_population(members=256, packets=(1.0e5, 1.0e7), grid_n=(32, 128), pixels=(256, 128), seed=4357, nsamples=8,
            fwhm=3.0, kernel_dim=6, target_counts=1.0e7, store_target=True):
    """Batched population: `members` galaxies (Cartesian grids with N per axis drawn uniformly from grid_n,
    dust hR U(2,8), hz U(0.1,0.6), face-on tau U(0.1,5), inclination U(85,90) deg, B/T U(0,0.5), packets
    log-uniform over `packets`), one pixels[0] x pixels[1] frame each (40 x 20 kpc), plus one extra member
    drawn the same way that serves as the fit target. Member i runs with seed+1+i, the target with seed; the
    parameter draws use numpy's default generator seeded with seed. The Gaussian PSF (fwhm in pixels,
    kernel_dim^2) and the fit interval are included; the target image itself is made by target_image() after
    the target member has run."""
    rng = np.random.default_rng(seed)
    nx_pix, ny_pix = pixels
    out = []
    for i in range(members + 1):
        hR, hz = rng.uniform(2.0, 8.0), rng.uniform(0.1, 0.6)
        tau, incl, bt = rng.uniform(0.1, 5.0), rng.uniform(85.0, 90.0), rng.uniform(0.0, 0.5)
        nx, ny, nz = (int(v) for v in rng.integers(grid_n[0], grid_n[1] + 1, size=3))
        npk = int(round(math.exp(rng.uniform(math.log(packets[0]), math.log(packets[1])))))
        is_target = i == members
        mseed = seed if is_target else seed + 1 + i
        comps, dgeom, mix, scale = _galaxy(False, tau, bt=bt, dust_hR=hR, dust_hz=hz)
        grid = S.CartesianDustGrid.uniform(EXTENT, nx, ny, nz)
        rho = S.sample_density(grid, dgeom, scale, nsamples, mseed)
        instr = S.InstrumentSystem([S.FrameInstrument(DISTANCE, math.radians(incl), 0.0, 0.0, nx_pix, ny_pix, 40.0,
                                                      40.0 * ny_pix / nx_pix)])
        d = dict(name="fit_population-target" if is_target else "fit_population-member%d" % i, kind="cartesian", grid=grid,
                 rho=rho, mix=mix, comps=comps, dgeom=dgeom, instr=instr,
                 params=dict(hR=hR, hz=hz, tau=tau, inclination=incl, bt=bt, grid=(nx, ny, nz)))
        out.append(_finish(d, npk, mseed))
    target = out.pop()
    kernel = S.GaussianKernel(fwhm, kernel_dim)
    return dict(name="fit_population", members=out, target=target, kernel=kernel, target_counts=target_counts)


def target_image(frame, counts, seed=4357):
    """Synthetic observation: frame (Ny, Nx) scaled to `counts` in total and Poisson-noised; returns the image
    and the luminosity-to-counts gain it used."""
    total = float(np.sum(frame))
    gain = counts / total if total > 0 else 1.0
    rng = np.random.default_rng(seed)
    return rng.poisson(np.clip(frame, 0.0, None) * gain).astype(np.float64), gain


def sample_sites(geometry, n, extent=EXTENT, seed=4357):
    """n positions drawn from a (spiral-decorated) exponential-disc density inside the box: radius from the
    disc's R exp(-R/hR) law, height from exp(-|z|/hz), azimuth uniform, then the spiral perturbation by
    rejection against its maximum (the same density the grid samples)."""
    rng = np.random.default_rng(seed)
    base = geometry.geometry if isinstance(geometry, S.SpiralStructureGeometryDecorator) else geometry
    hR, hz = base.hR, base.hz
    out = np.zeros((0, 3))
    while len(out) < n:
        m = 2 * (n - len(out)) + 1024
        R = rng.gamma(2.0, hR, m)
        z = rng.laplace(0.0, hz, m)
        phi = rng.uniform(0.0, 2.0 * math.pi, m)
        p = np.stack([R * np.cos(phi), R * np.sin(phi), z], 1)
        inside = np.all((p > extent[:3]) & (p < extent[3:]), axis=1)
        p = p[inside]
        if isinstance(geometry, S.SpiralStructureGeometryDecorator):
            w = geometry.perturbWeight
            N = geometry.index
            cn = math.sqrt(math.pi) * math.exp(math.lgamma(N + 1.0) - math.lgamma(N + 0.5))
            ratio = S.geometry_density(geometry, p) / np.maximum(S.geometry_density(base, p), 1e-300)
            p = p[rng.uniform(0.0, (1.0 - w) + w * cn, len(p)) < ratio]
        out = np.concatenate([out, p])
    return out[:n]


def spiral_voronoi(nsites=1_000_000, nlambda=50, pixels=256, packages=2e7, seed=4357, tau_v=1.0, store_abs=True):
    """Voronoi: nsites generator sites drawn from the spiral_octree dust density, each cell with the analytic
    density at its site; nlambda wavelengths log-spaced over 0.1-10 micron with kappa(lambda) =
    kappa_V (lambda/0.55)^-1.3 normalised to a face-on tau_V, albedo 0.7 -> 0.4 and g 0.7 -> 0.3 linear in
    the wavelength index; the absorbed luminosity per cell and wavelength is stored. Stars as spiral_octree
    with a flat spectrum; two frames (0 and 75 deg) of pixels^2."""
    comps, dgeom, _, _ = _galaxy(True, 1.0)
    lam = np.logspace(-1.0, 1.0, nlambda)
    kappa = (lam / 0.55) ** -1.3
    idx = np.arange(nlambda) / max(1, nlambda - 1)
    mix = S.SimpleOligoDustMix(kappa, 0.7 - 0.3 * idx, 0.7 - 0.4 * idx)
    scale = S.faceon_normalization(dgeom.geometry, tau_v, 1.0)
    sites = S.voronoi_cells(EXTENT, sample_sites(dgeom, nsites, EXTENT, seed))
    grid = S.VoronoiDustGrid(EXTENT, sites)
    rho = scale * S.geometry_density(dgeom, sites)
    comps = [S.OligoStellarComp(c.geometry, [c.luminosities[0]] * nlambda) for c in comps]
    ins = [S.FrameInstrument(DISTANCE, math.radians(a), 0.0, 0.0, pixels, pixels, 40.0, 40.0) for a in (0.0, 75.0)]
    d = dict(name="spiral_voronoi", kind="voronoi", grid=grid, rho=rho, mix=mix, comps=comps, dgeom=dgeom,
             instr=S.InstrumentSystem(ins, nlambda), store_abs=store_abs, wavelengths=lam)
    return _finish(d, packages, seed)


def lognormal_field(n, sigma=1.5, slope=-11.0 / 3.0, seed=4357):
    """A periodic n^3 log-normal density field exp(g) with g Gaussian of standard deviation sigma and power
    spectrum P(k) ~ k^slope (k in cycles per box), normalised to unit mean."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n) * n
    kk = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + k1[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    amp[kk > 0] = kk[kk > 0] ** (0.5 * slope)
    g = np.fft.ifftn(np.fft.fftn(rng.standard_normal((n, n, n))) * amp).real
    g = (g - g.mean()) / g.std() * sigma
    rho = np.exp(g)
    return rho / rho.mean()


def lognormal_particles(nparticles, nvox=256, seed=4357, half=15.0, neighbours=32, total_mass=1.0):
    """An SPH particle table (rows x y z h M Z T) sampled from the log-normal field over a box of half-width
    half: each particle in a voxel drawn with probability proportional to the voxel's density, at a uniform
    position inside it; smoothing length h the distance to the neighbours-th nearest other particle; equal
    masses summing to total_mass; metallicity 1 and no temperature."""
    field = lognormal_field(nvox, seed=seed)
    rng = np.random.default_rng(seed + 1)
    n = int(nparticles)
    cdf = np.cumsum(field.ravel())
    idx = np.minimum(np.searchsorted(cdf, rng.random(n) * cdf[-1], side="right"), cdf.size - 1)
    ijk = np.stack(np.unravel_index(idx, field.shape), axis=1)
    pos = -half + (ijk + rng.random((n, 3))) * (2.0 * half / nvox)
    h = S.knn_distance(pos, neighbours)
    return np.column_stack([pos, h, np.full(n, total_mass / n), np.ones(n), np.zeros(n)])


def lognormal_kdtree(nvox=256, maxmassfraction=None, maxlevel=40, pixels=1024, packages=1e9, seed=4357, tau=1.0,
                     search=SG_TOPDOWN, half=15.0, medium="particles", particles=1e7, neighbours=32):
    """kd-tree over a log-normal dust medium (sigma_ln 1.5, P(k) ~ k^-11/3) in a 30-kpc box: by default an SPH
    particle set (SPHDustDistribution) of `particles` particles sampled from the field with smoothing lengths
    enclosing `neighbours` neighbours, the BinTreeDustGrid (alternating split axes) built over it on the device
    with node masses from the particles' mass in each node's box, and the cells' exact mass-in-box densities
    (TreeDustGrid.cpp:655-660); medium="voxels" takes the nvox^3 voxel field itself as the medium instead. The
    medium's mass gives a mean optical depth tau across the box height. maxmassfraction (default 9e-8) sets the
    leaf count: 1.615e7 for the particles and 1.61e7 for the voxel field at the defaults. Stars: an
    exponential disc (hR 4, hz 0.4). Eight frames of pixels^2 at inclinations 30 and 60 deg, azimuths 0, 90,
    180, 270 deg."""
    ext = np.array([-half, -half, -half, half, half, half])
    mix = S.SimpleOligoDustMix([1.0], [0.6], [0.6])
    d = dict(name="lognormal_kdtree", kind="bintree", mix=mix, dgeom=None)
    if medium == "voxels":
        field = lognormal_field(nvox, seed=seed)
        grid = S.BinTreeDustGrid.subdivide_field(ext, field, minlevel=3, maxlevel=maxlevel,
                                                 maxmassfraction=9e-8 if maxmassfraction is None else maxmassfraction,
                                                 search=search)
        rho = S.field_density(grid, field) * tau / (2.0 * half * mix.kappaext[0])
        d["field"] = field
    elif medium == "particles":
        # mean density tau / (2 half kappa) over the box volume (2 half)^3
        mass = tau * (2.0 * half) ** 2 / mix.kappaext[0]
        rows = lognormal_particles(particles, nvox, seed, half, neighbours, mass)
        sph = S.SPHDustDistribution(rows, 1.0)
        grid, masses = S.BinTreeDustGrid.subdivide_sph(ext, sph, minlevel=3, maxlevel=maxlevel,
                                                       maxmassfraction=9e-8 if maxmassfraction is None
                                                       else maxmassfraction, search=search, masses=True)
        rho = S.sph_density(grid, sph, masses)
        d["sph"] = sph
        d["particles"] = rows
    else:
        raise ValueError("medium is 'particles' or 'voxels'")
    comps = [S.OligoStellarComp(S.ExpDiskGeometry(4.0, 0.4), [1.0e10])]
    ins = [S.FrameInstrument(DISTANCE, math.radians(i), math.radians(a), 0.0, pixels, pixels, 2.0 * half, 2.0 * half)
           for i in (30.0, 60.0) for a in (0.0, 90.0, 180.0, 270.0)]
    d.update(grid=grid, rho=rho, comps=comps, instr=S.InstrumentSystem(ins))
    return _finish(d, packages, seed)


# a two-segment stand-in for the solar SED (lambda in micron, L in W/micron): OligoStellarComp luminosities are
# in solar units, and the fit's luminosity scale absorbs the conversion, so any positive SED serves
SUN_STANDIN = """# a two-segment stand-in for the solar SED (lambda in micron, L in W/micron)
3
0.4  1.0e26
0.6  2.0e26
1.0  1.0e26
"""

FIT_SKI = """<?xml version="1.0" encoding="UTF-8"?>
<skirt-simulation-hierarchy type="MonteCarloSimulation" format="6.1">
  <OligoMonteCarloSimulation packages="[packages:%(packages)g]" minWeightReduction="1e4" scattBias="0.5">
    <random type="Random"><Random seed="4357"/></random>
    <units type="Units"><ExtragalacticUnits fluxOutputStyle="Neutral"/></units>
    <instrumentSystem type="InstrumentSystem">
      <InstrumentSystem>
        <instruments type="Instrument">
          <FrameInstrument instrumentName="edgeon" distance="10 Mpc" inclination="[incl:88 deg]" azimuth="0 deg"
             positionAngle="0 deg" fieldOfViewX="40 kpc" pixelsX="%(nx)d" fieldOfViewY="%(fovy)g kpc" pixelsY="%(ny)d"/>
        </instruments>
      </InstrumentSystem>
    </instrumentSystem>
    <wavelengthGrid type="OligoWavelengthGrid"><OligoWavelengthGrid wavelengths="0.55 micron"/></wavelengthGrid>
    <stellarSystem type="StellarSystem">
      <StellarSystem emissionBias="0.5">
        <components type="StellarComp">
          <OligoStellarComp luminosities="1e10">
            <geometry type="Geometry"><ExpDiskGeometry radialScale="[stellar_hr:4000 pc]" axialScale="[stellar_hz:400 pc]"/></geometry>
          </OligoStellarComp>
        </components>
      </StellarSystem>
    </stellarSystem>
    <dustSystem type="DustSystem">
      <OligoDustSystem sampleCount="%(nsamples)d" writeMeanIntensity="false">
        <dustDistribution type="DustDistribution">
          <CompDustDistribution>
            <components type="DustComp">
              <DustComp>
                <geometry type="Geometry"><ExpDiskGeometry radialScale="[dust_hr:6000 pc]" axialScale="[dust_hz:250 pc]"/></geometry>
                <mix type="DustMix"><SimpleOligoDustMix opacities="1 m2/kg" albedos="0.6" asymmetryParameters="0.6"/></mix>
                <normalization type="DustCompNormalization">
                  <FaceOnDustCompNormalization wavelength="0.55 micron" opticalDepth="[tau:1]"/>
                </normalization>
              </DustComp>
            </components>
          </CompDustDistribution>
        </dustDistribution>
        <dustGrid type="DustGrid">
          <CartesianDustGrid minX="-20 kpc" maxX="20 kpc" minY="-20 kpc" maxY="20 kpc" minZ="-5 kpc" maxZ="5 kpc">
            <meshX type="MoveableMesh"><LinMesh numBins="%(gx)d"/></meshX>
            <meshY type="MoveableMesh"><LinMesh numBins="%(gx)d"/></meshY>
            <meshZ type="MoveableMesh"><LinMesh numBins="%(gz)d"/></meshZ>
          </CartesianDustGrid>
        </dustGrid>
      </OligoDustSystem>
    </dustSystem>
  </OligoMonteCarloSimulation>
</skirt-simulation-hierarchy>
"""


def fitskirt_edgeon(packages=1e6, pixels=(256, 128), grid=(100, 30), nsamples=8):
    """A FitSKIRT problem: an edge-on disc galaxy (stellar disc with a dust disc, Cartesian grid of
    grid[0]^2 x grid[1] cells over 40 x 40 x 10 kpc) whose ski file has placeholders for six parameters (the
    ranges of FitSKIRT's edge-on tutorial kind: stellar and dust scale lengths and heights, face-on optical
    depth, inclination), the parameter ranges and the true values of a synthetic observation. Lengths of the
    ranges in SI (m) and the inclination in radians, as FitSKIRT's ParameterRange holds them."""
    from . import fitskirt as F
    from .output import PC
    text = FIT_SKI % dict(packages=packages, nx=pixels[0], ny=pixels[1], fovy=40.0 * pixels[1] / pixels[0],
                          nsamples=nsamples, gx=grid[0], gz=grid[1])
    ranges = [F.ParameterRange("stellar_hr", "length", 2500 * PC, 5500 * PC),
              F.ParameterRange("stellar_hz", "length", 200 * PC, 700 * PC),
              F.ParameterRange("dust_hr", "length", 3000 * PC, 9000 * PC),
              F.ParameterRange("dust_hz", "length", 100 * PC, 500 * PC),
              F.ParameterRange("tau", "", 0.2, 3.0),
              F.ParameterRange("incl", "posangle", math.radians(85.0), math.radians(90.0))]
    truth = [3800 * PC, 420 * PC, 6500 * PC, 240 * PC, 1.2, math.radians(87.5)]
    return dict(ski=text, ranges=ranges, truth=truth)


def fitskirt_observation(problem, resources, counts=1e6, mult=10, seed=5):
    """The synthetic observation of fitskirt_edgeon's truth: its frame with `mult` times the packets, calibrated,
    convolved with the 3-pixel FWHM Gaussian PSF and Poisson-noised at `counts` counts in all (the chi2's
    sigma is Poisson). Returns (reference image, PSF kernel, gain)."""
    from scipy.ndimage import convolve

    from . import fitskirt as F
    from . import output, ski
    from .capi import SG_MIXED
    units = F._units_of(problem["ski"])
    repl = {r.label: (float(v), r.quantityString()) for r, v in zip(problem["ranges"], problem["truth"])}
    spec0 = ski.Spec(F.adjusted_ski_content(problem["ski"], repl, units), resources)
    repl["packages"] = (spec0.packages * mult, "")
    spec = ski.Spec(F.adjusted_ski_content(problem["ski"], repl, units), resources)
    sim = ski.Simulation(spec)
    sim.run(SG_MIXED)
    f = spec.frames[0]
    cal = output.calibrate_frame(sim.instr.frame(0), spec.lambdas, spec.dlambdas, f["distance"],
                                 f["fovX"] / f["pixelsX"], f["fovY"] / f["pixelsY"], spec.units)[0]
    kernel = S.GaussianKernel(3.0, 6)
    cal = convolve(cal, kernel.data, mode="constant")
    gain = counts / cal.sum()
    data = np.random.default_rng(seed).poisson(cal * gain).astype(np.float64)
    return data, kernel, gain


def two_dust_disc(kind="cartesian", grid_n=32, maxlevel=5, nsites=20000, pixels=64, packages=1e5, seed=4357,
                  nsamples=20, store_abs=True):
    """Two dust components (a CompDustDistribution of two DustComps): a thick diffuse disc (hR 6, hz 0.5,
    face-on tau 0.5 at the first wavelength, albedo 0.7, g 0.6) and a thin dense lane (hR 3, hz 0.12, face-on
    tau 2, albedo 0.3, g 0.1), each with its own two-wavelength mix (kappa falling by 0.6 and 0.8 to the
    second wavelength); the disc + bulge stars of disc_cartesian; frames at 0, 60 and 88 deg. kind: cartesian
    (grid_n^3), octree (subdivided on the lane's density to maxlevel) or voronoi (nsites drawn from the
    lane)."""
    comps, _, _, _ = _galaxy(False, 1.0)
    for c in comps:
        c.luminosities = np.array([c.luminosities[0], 0.7 * c.luminosities[0]])
    diffuse, lane = S.ExpDiskGeometry(6.0, 0.5), S.ExpDiskGeometry(3.0, 0.12)
    mixes = [S.SimpleOligoDustMix([1.0, 0.6], [0.7, 0.65], [0.6, 0.5]),
             S.SimpleOligoDustMix([1.0, 0.8], [0.3, 0.25], [0.1, 0.05])]
    scales = [S.faceon_normalization(diffuse, 0.5, 1.0), S.faceon_normalization(lane, 2.0, 1.0)]
    sites = None
    if kind == "cartesian":
        grid = S.CartesianDustGrid.uniform(EXTENT, grid_n, grid_n, grid_n)
    elif kind == "octree":
        grid = S.OctTreeDustGrid.subdivide(EXTENT, lane, scales[1], total_mass=scales[1], minlevel=2,
                                           maxlevel=maxlevel, maxmassfraction=1e-4, nsamples=nsamples, seed=seed)
    else:
        sites = S.voronoi_cells(EXTENT, sample_sites(lane, nsites, EXTENT, seed))
        grid = S.VoronoiDustGrid(EXTENT, sites)
    if sites is not None:
        rhos = [s * S.geometry_density(g, sites) for g, s in zip((diffuse, lane), scales)]
    else:
        rhos = [S.sample_density(grid, g, s, nsamples, seed) for g, s in zip((diffuse, lane), scales)]
    d = dict(name="two_dust_disc-" + kind, kind=kind, grid=grid, rho=rhos, mix=mixes, comps=comps, dgeom=lane,
             instr=_frames((0.0, 60.0, 88.0), pixels, nlambda=2), store_abs=store_abs)
    return _finish(d, packages, seed)


def pan_disc(kind="cartesian", nlambda=30, grid_n=24, maxlevel=5, pixels=48, packages=2e4, seed=4357, tau=1.0,
             nsamples=20, selfAbsorption=False, cycles=0, precision=None):
    """A panchromatic disc galaxy (PanMonteCarloSimulation): a stellar exponential disc (hR 4, hz 0.4) with a
    6000 K blackbody spectrum of 1e10 solar luminosities, a dust disc (hR 6, hz 0.3) of Benchmark1DDustMix
    dust with face-on optical depth `tau` at 0.55 micron, nlambda wavelengths log-spaced over 0.1-1000 micron,
    frames at 0 and 88 deg (nlambda planes each); grid: cartesian (grid_n^3) or octree (to maxlevel). Lengths
    in kpc, luminosities in W."""
    from . import panchromatic as P
    from .capi import SG_MIXED
    from .output import PC
    lam, dl = P.log_wavelength_grid(1e-7, 1e-3, nlambda)
    mix = P.Benchmark1DDustMix(lam, dl)
    sed = P.planck(lam, 6000.0) * dl
    lum = 1e10 * 3.846e26 * sed / sed.sum()
    comps = [S.OligoStellarComp(S.ExpDiskGeometry(4.0, 0.4), lum)]
    dgeom = S.ExpDiskGeometry(6.0, 0.3)
    ellv = int(np.argmin(np.abs(np.log(lam / 0.55e-6))))
    scale = S.faceon_normalization(dgeom, tau, mix.kappaext[ellv])
    if kind == "cartesian":
        grid = S.CartesianDustGrid.uniform(EXTENT, grid_n, grid_n, grid_n)
    else:
        grid = S.OctTreeDustGrid.subdivide(EXTENT, dgeom, scale, total_mass=scale, minlevel=2, maxlevel=maxlevel,
                                           maxmassfraction=1e-4, nsamples=nsamples, seed=seed)
    rho = S.sample_density(grid, dgeom, scale, nsamples, seed)
    dust = P.PanDustSystem(grid, rho, mix, selfAbsorption=selfAbsorption, cycles=cycles, length_unit=1000.0 * PC)
    instr = _frames((0.0, 88.0), pixels, nlambda=nlambda)
    stars = S.StellarSystem(comps, emissionBias=0.5)
    pan = P.PanMonteCarloSimulation(stars, dust, instr, packages=packages, seed=seed,
                                    precision=SG_MIXED if precision is None else precision)
    return dict(name="pan_disc-" + kind, kind=kind, grid=grid, rho=rho, mix=mix, comps=comps, dgeom=dgeom,
                instr=instr, stars=stars, dust=dust.ds, pan=pan, sim=pan.sim, packages=packages, lambdas=lam)
