"""Panchromatic dust emission on the GPU against the oracle.

- The dust's emission spectra and bolometric absorbed luminosity (sg_dust_emission_spectra) from a stellar
  run's absorption table, against the oracle's restatement on the same table.
- Dust photon packets (sg_sim_run_dust) packet by packet against the oracle on the same Philox streams: the
  dust emission launch with emission bias and peel-offs (frames), and the self-absorption launch without
  peel-offs (absorption table), on Cartesian and octree grids.
- A whole PanMonteCarloSimulation with self-absorption cycles: the dust re-emits what it absorbs, the frames
  hold dust emission in the far infrared, the cycles converge.
"""
import numpy as np
import pytest

from skirt7_amd import panchromatic as P
from skirt7_amd import workloads
from skirt7_amd.capi import SG_FP64, SG_MIXED
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["cartesian", "octree"])
def w(request):
    w = workloads.pan_disc(request.param, nlambda=24, packages=20000, precision=SG_FP64)
    w["pan"].runstellaremission()
    return w


def oracle_scene(w):
    sc = O.scene_from_workload(w)
    sc.ncells, sc.nlambda = len(w["rho"]), len(w["lambdas"])
    return sc


def test_spectra_against_oracle(w):
    pan = w["pan"]
    ls = pan.Labs_stellar
    assert ls.sum() > 0
    dl, bol = pan.dust.emission_spectra(ls)
    sc = oracle_scene(w)
    from skirt7_amd.output import PC
    odl, obol = sc.dust_spectra(w["mix"], ls, len2=(1000.0 * PC) ** 2)
    assert np.array_equal(bol, obol)
    hit = obol > 0
    assert hit.sum() > 100
    np.testing.assert_allclose(dl[hit], odl[hit], rtol=1e-9, atol=1e-300)
    # the luminosity-weighted dust spectrum peaks in the far infrared (cold dust)
    peak = w["lambdas"][np.argmax((bol[:, None] * dl).sum(axis=0))]
    assert 3e-5 < peak < 6e-4


def frames_of(w, ell):
    return [w["instr"].frame(i)[ell].copy() for i in range(len(w["instr"].instruments))]


def oracle_frames(w, ref, ell, nl):
    out, off = [], 0
    for ins in w["instr"].instruments:
        n = ins.Nx * ins.Ny
        out.append(ref[off + ell * n:off + (ell + 1) * n].reshape(ins.Ny, ins.Nx))
        off += nl * n
    return out


def test_dust_packets_against_oracle(w):
    from skirt7_amd import capi
    pan = w["pan"]
    dl, bol = pan.dust.emission_spectra(pan.Labs_stellar)
    sc = oracle_scene(w)
    nl = len(w["lambdas"])
    mc = w["sim"].mc
    n = 20000
    for ell in (nl // 2, nl - 6):
        Lv = np.ascontiguousarray(bol * dl[:, ell])
        assert Lv.sum() > 0
        # the dust emission launch: bias 0.5, peel-offs
        w["instr"].clear()
        c = np.zeros(capi.SG_NCOUNTERS, np.uint64)
        capi.check(capi.lib().sg_sim_run_dust(w["sim"]._h, ell, 0, n, n, capi.dptr(Lv), 0.5, 1, 0, 1000, SG_FP64,
                                               c.ctypes.data_as(capi._U64)))
        ref, rc = sc.run_dust_packets(mc, ell, 0, n, n, Lv, 0.5, True, False, 1000)
        assert int(c[0]) == int(rc[0]) == n
        for g, o in zip(frames_of(w, ell), oracle_frames(w, ref, ell, nl)):
            assert o.sum() > 0
            assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()
            hit = o != 0
            assert np.mean(np.abs(g - o)[hit] <= 1e-9 * np.abs(o[hit])) > 0.99
        # the self-absorption launch: no peel-offs, absorption stored
        w["dust"].clear()
        w["instr"].clear()
        capi.check(capi.lib().sg_sim_run_dust(w["sim"]._h, ell, 0, n, n, capi.dptr(Lv), 0.0, 0, 1, 7, SG_FP64,
                                               c.ctypes.data_as(capi._U64)))
        labs = w["dust"].Labs()[:, ell]
        assert all(f.sum() == 0 for f in frames_of(w, ell))
        sc.run_dust_packets(mc, ell, 0, n, n, Lv, 0.0, False, True, 7)
        rl = sc.labs.reshape(-1, nl)[:, ell]
        assert rl.sum() > 0
        assert abs(labs.sum() - rl.sum()) <= 1e-6 * rl.sum()
        close = np.abs(labs - rl) <= 1e-9 * np.abs(rl) + 1e-300
        assert np.mean(close) > 0.99


def test_whole_run_with_self_absorption():
    w = workloads.pan_disc("cartesian", nlambda=24, packages=50000, tau=3.0, selfAbsorption=True, cycles=2)
    pan = w["pan"]
    pan.runSelf()
    assert len(pan.log) == 6
    assert pan.Labs_dust.sum() > 0
    lam = w["lambdas"]
    # the dust re-emits all it absorbed: the cells' bolometric absorbed luminosity is stellar plus dust
    assert abs(pan.Labsbol.sum() / (pan.Labs_stellar.sum() + pan.Labs_dust.sum()) - 1) < 1e-12
    # far-infrared flux comes from the dust emission phase
    fir = [ell for ell in range(len(lam)) if lam[ell] > 5e-5]
    edge = w["instr"].frame(1)
    assert all(edge[ell].sum() > 0 for ell in fir)
    # the absorbed dust luminosity settles: the last cycles of the last stage change little
    last = [tot for (stage, cycle, tot, eps) in pan.log if stage == 2]
    assert abs(last[-1] / last[-2] - 1) < 0.05


def test_ski_file_panchromatic_run(tmp_path):
    """A PanMonteCarloSimulation ski file (LogWavelengthGrid, PanStellarComp with a black-body SED, PanDustSystem
    of Benchmark1DDustMix dust with self-absorption) runs as the same objects built through the Python API:
    the same frames at every wavelength (to the order of the tallies' additions), with the dust's far-infrared
    emission in them."""
    import math
    from skirt7_amd import ski
    from skirt7_amd import skirt as S
    from skirt7_amd.output import LSUN, PC
    from tests.test_ski_fits import SUN, pan_ski
    d = tmp_path / "dat"
    (d / "SED" / "Sun").mkdir(parents=True)
    (d / "SED" / "Sun" / "SunSED.dat").write_text(SUN)
    sim = ski.Simulation(ski.Spec(pan_ski(selfabs=True), str(d)))
    sim.run()
    got = [sim.instr.frame(i).copy() for i in range(2)]
    lam, dl = P.log_wavelength_grid(1e-7, 1e-3, 20)
    mix = P.Benchmark1DDustMix(lam, dl)
    ext = np.array([-2e4, -2e4, -2e4, 2e4, 2e4, 2e4])
    grid = S.CartesianDustGrid(ext, S.linmesh(-2e4, 2e4, 16), S.linmesh(-2e4, 2e4, 16), S.linmesh(-2e4, 2e4, 12))
    dgeom = S.ExpDiskGeometry(6000.0, 300.0)
    scale = S.faceon_normalization(dgeom, 1.0, mix.kappaext[P.nearest_wavelength(lam, 0.55e-6)])
    rho = S.sample_density(grid, dgeom, scale, 20, 4357)
    dust = P.PanDustSystem(grid, rho, mix, selfAbsorption=True, cycles=2, length_unit=PC)
    stars = S.StellarSystem([S.OligoStellarComp(S.ExpDiskGeometry(4000.0, 400.0),
                                                list(1e10 * LSUN * P.blackbody_sed(lam, 6000.0)))], emissionBias=0.5)
    instr = S.InstrumentSystem([S.FrameInstrument(1e7, 0.0, 0.0, 0.0, 24, 24, 4e4, 4e4),
                                S.FrameInstrument(1e7, math.radians(88.0), 0.0, 0.0, 24, 12, 4e4, 2e4)], 20)
    pan = P.PanMonteCarloSimulation(stars, dust, instr, packages=2e4, seed=4357)
    pan.runSelf()
    for i in range(2):
        # the same packets; the tallies' FP64 atomic additions come in a different order from run to run
        np.testing.assert_allclose(got[i], instr.frame(i), rtol=1e-11, atol=0)
    fir = int(np.argmin(np.abs(lam - 1e-4)))
    assert got[0][fir].sum() > 0
