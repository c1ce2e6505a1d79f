"""FitSKIRT on the GPU: a generation's individuals evaluated together by pooled member simulations.

- The device density sampler (sg_grid_resample_density) against the host sampler: optical depths along random
  rays through both grids agree to rounding, for a plain and a spiral-decorated disc.
- Pooled members updated in place (density resampled on the device, stellar and instrument parameters
  replaced) give the frames of simulations built from scratch for the same adjusted ski files, in FP64 and
  mixed precision, also when a member is reused for another individual.
- A fit run: the library's genetic algorithm asks, generation by generation, for the same individuals as
  the reference's GAlib (oracle/_ref/ga_ref) fed the same GPU chi2 scores; the best chi2 never rises and the
  individual nearest the target's parameters scores well below the first generation's median.
"""
import math

import numpy as np
import pytest

from skirt7_amd import fitskirt as F
from skirt7_amd import output, ski
from skirt7_amd import skirt as S
from skirt7_amd.capi import SG_FP64, SG_MIXED
from tests import galib_ref
from tests.helpers import random_rays

pytestmark = pytest.mark.gpu

SUN = """# a two-segment stand-in for the solar SED (lambda in micron, L in W/micron)
3
0.4  1.0e26
0.6  2.0e26
1.0  1.0e26
"""

FSKI = """<?xml version="1.0" encoding="UTF-8"?>
<skirt-simulation-hierarchy type="MonteCarloSimulation" format="6.1">
  <OligoMonteCarloSimulation packages="[packages:1e5]" minWeightReduction="1e4" scattBias="0.5">
    <random type="Random"><Random seed="4357"/></random>
    <units type="Units"><ExtragalacticUnits fluxOutputStyle="Neutral"/></units>
    <instrumentSystem type="InstrumentSystem">
      <InstrumentSystem>
        <instruments type="Instrument">
          <FrameInstrument instrumentName="edgeon" distance="10 Mpc" inclination="[incl:88 deg]" azimuth="0 deg"
             positionAngle="0 deg" fieldOfViewX="40 kpc" pixelsX="48" fieldOfViewY="20 kpc" pixelsY="24"/>
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
      <OligoDustSystem sampleCount="6" writeMeanIntensity="false">
        <dustDistribution type="DustDistribution">
          <CompDustDistribution>
            <components type="DustComp">
              <DustComp>
                <geometry type="Geometry"><ExpDiskGeometry radialScale="[dust_hr:6000 pc]" axialScale="250 pc"/></geometry>
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
            <meshX type="MoveableMesh"><LinMesh numBins="40"/></meshX>
            <meshY type="MoveableMesh"><LinMesh numBins="40"/></meshY>
            <meshZ type="MoveableMesh"><LinMesh numBins="20"/></meshZ>
          </CartesianDustGrid>
        </dustGrid>
      </OligoDustSystem>
    </dustSystem>
  </OligoMonteCarloSimulation>
</skirt-simulation-hierarchy>
"""

PC = output.PC
RANGES = [F.ParameterRange("stellar_hr", "length", 2500 * PC, 5500 * PC),
          F.ParameterRange("dust_hr", "length", 3000 * PC, 9000 * PC),
          F.ParameterRange("tau", "", 0.2, 3.0),
          F.ParameterRange("incl", "posangle", math.radians(85.0), math.radians(90.0))]


@pytest.fixture
def resources(tmp_path):
    d = tmp_path / "dat"
    (d / "SED" / "Sun").mkdir(parents=True)
    (d / "SED" / "Sun" / "SunSED.dat").write_text(SUN)
    return str(d)


def repl(values):
    return {r.label: (float(v), r.quantityString()) for r, v in zip(RANGES, values)}


@pytest.mark.parametrize("spiral", [False, True])
def test_device_density_equals_host(spiral):
    ext = np.array([-20.0, -20.0, -5.0, 20.0, 20.0, 5.0])
    disc = S.ExpDiskGeometry(5.0, 0.3)
    geom = disc
    if spiral:
        geom = S.SpiralStructureGeometryDecorator(disc, arms=2, pitch=math.radians(20.0), radius=4.0, phase=0.3,
                                                  perturbWeight=0.3, index=2)
    scale = S.faceon_normalization(disc, 1.5, 1.0)
    g_host = S.CartesianDustGrid.uniform(ext, 30, 26, 18)
    g_dev = S.CartesianDustGrid.uniform(ext, 30, 26, 18)
    rho = S.sample_density(g_host, geom, scale, 7, 123)
    mix = S.SimpleOligoDustMix([1.0], [0.6], [0.6])
    d_host = S.DustSystem(g_host, rho, mix)
    d_dev = S.DustSystem(g_dev, np.zeros_like(rho), mix)
    S.resample_density(g_dev, geom, scale, 7, 123)
    rays = random_rays(2000, 7, ext)
    t_host = d_host.opticaldepths(rays, 0, SG_FP64)
    t_dev = d_dev.opticaldepths(rays, 0, SG_FP64)
    assert t_host.max() > 0.5
    np.testing.assert_allclose(t_dev, t_host, rtol=1e-12, atol=0)
    np.testing.assert_allclose(d_dev.opticaldepths(rays, 0, SG_MIXED), d_host.opticaldepths(rays, 0, SG_MIXED),
                               rtol=1e-6, atol=0)


@pytest.mark.parametrize("precision", [SG_FP64, SG_MIXED])
def test_pooled_members_equal_fresh_simulations(resources, precision):
    sim = F.AdjustableSkirtSimulation(FSKI, resources=resources, precision=precision)
    sets = [[4000 * PC, 6000 * PC, 1.0, math.radians(88.0)],
            [3000 * PC, 8000 * PC, 2.5, math.radians(86.0)],
            [5000 * PC, 4000 * PC, 0.4, math.radians(89.5)]]
    first = sim.evaluate(sim.specs([repl(sets[0]), repl(sets[1])]))
    got = [m.sim.instr.frame(0).copy() for m in first]
    # reuse: member 0 takes individual 2, member 1 takes individual 0
    again = sim.evaluate(sim.specs([repl(sets[2]), repl(sets[0])]))
    got += [m.sim.instr.frame(0).copy() for m in again]
    assert sim.rebuilt == 0 and len(sim.pool) == 2
    for values, frame in zip([sets[0], sets[1], sets[2], sets[0]], got):
        fresh = ski.Simulation(ski.Spec(F.adjusted_ski_content(FSKI, repl(values), sim.units), resources))
        fresh.run(precision)
        ref = fresh.instr.frame(0)
        assert ref.sum() > 0
        assert abs(frame.sum() - ref.sum()) <= 1e-9 * ref.sum()
        hit = ref != 0
        close = np.abs(frame - ref)[hit] <= 1e-9 * np.abs(ref[hit])
        assert np.mean(close) > 0.999
    # the same individual in two different members gives the same frame
    assert np.allclose(got[0], got[3], rtol=1e-9, atol=0)


def test_fit_run_follows_galib(resources):
    if not galib_ref.available():
        pytest.skip("oracle/_ref/ga_ref is not built")
    truth = [3800 * PC, 6500 * PC, 1.2, math.radians(87.5)]
    sim = F.AdjustableSkirtSimulation(FSKI, resources=resources, precision=SG_MIXED)
    # the observation: the truth with 20x the packets, calibrated
    obs_spec = ski.Spec(F.adjusted_ski_content(FSKI, dict(repl(truth), packages=(2e6, "")), sim.units), resources)
    obs = ski.Simulation(obs_spec)
    obs.run(SG_MIXED)
    f = obs_spec.frames[0]
    cal = output.calibrate_frame(obs.instr.frame(0), obs_spec.lambdas, obs_spec.dlambdas, f["distance"],
                                 f["fovX"] / f["pixelsX"], f["fovY"] / f["pixelsY"], obs_spec.units)[0]
    # seen through the PSF, in counts (the chi2's Poisson sigma needs them): 2e5 counts in all, so the
    # luminosity scale is the gain
    kernel = S.GaussianKernel(2.0, 5)
    from scipy.ndimage import convolve
    cal = convolve(cal, kernel.data, mode="constant")
    gain = 2e5 / cal.sum()
    data = np.random.default_rng(5).poisson(cal * gain).astype(np.float64)
    image = F.ReferenceImage(data, kernel, [0.2 * gain], [5.0 * gain])
    optim = F.Optimization(popsize=12, generations=4, pmut=0.03, pcross=0.65)
    scheme = F.OligoFitScheme(sim, RANGES, [image], optim, fixedSeed=True)
    lo = [r.minimumValue for r in RANGES]
    hi = [r.maximumValue for r in RANGES]
    ref = galib_ref.GaRef(lo, hi, 12, 4, 0.03, 0.65, 4357)
    try:
        optim.setup(RANGES, 4357)
        bests = []
        first = True
        while True:
            kind, gen, payload, *rest = ref.request()
            if kind == "DONE":
                break
            if first:
                optim.initialize()
                first = False
            else:
                optim.step()
            genes = optim.pending()
            print("generation %d: %d individuals" % (gen, len(genes)), flush=True)
            assert np.array_equal(genes, payload), gen
            chi2, lums, chis = scheme.objective(genes)
            assert np.all(np.isfinite(chi2)) and np.all(chi2 > 0)
            assert np.all((lums > 0.2 * gain * (1 - 1e-12)) & (lums < 5.0 * gain * (1 + 1e-12)))
            if gen == 0:
                median0 = float(np.median(chi2))
            ref.scores(chi2)
            optim.set_scores(chi2)
            bests.append(optim.best()[1])
        assert optim.done() and gen == 4
        assert all(b2 <= b1 for b1, b2 in zip(bests, bests[1:]))
        bg, bs = optim.best()
        assert np.array_equal(bg, rest[0][0]) and bs == rest[0][1]
    finally:
        ref.close()
    # the truth itself scores far better than a typical random individual
    c_truth, l_truth, _ = scheme.objective(np.array([truth]))
    assert c_truth[0] < 0.5 * median0
    assert abs(l_truth[0, 0] / gain - 1.0) < 0.1
