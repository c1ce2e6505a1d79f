"""Dust systems of several components (a CompDustDistribution of two DustComps with their own mixes) on the
GPU against the oracle.

- FP64 packets on the Cartesian, octree and Voronoi grids at both wavelengths, packet by packet on the same
  Philox streams: frames and the absorption table (each cell's own albedo, MonteCarloSimulation.cpp:463-508),
  with scattering peel-offs weighted per component (:318-343) and the scattering mix drawn per cell
  (DustSystem.cpp:855-869).
- Optical depths through the combined KappaRho (DustSystem.cpp:459-485) bit for bit, at both wavelengths.
- Mixed precision against FP64 on the same streams; a batched run of multi- and single-component members
  against their single runs; sg_sim_run_static refuses such a system.
"""
import numpy as np
import pytest

from skirt7_amd import skirt as S
from skirt7_amd import workloads
from skirt7_amd.capi import SG_FP64, SG_MIXED
from tests import oracle_lib as O
from tests.helpers import random_rays

pytestmark = pytest.mark.gpu


def gpu_frames(w, ell):
    return [w["instr"].frame(i)[ell].copy() for i in range(len(w["instr"].instruments))]


def oracle_frames(w, ref, ell, nl):
    out, off = [], 0
    for ins in w["instr"].instruments:
        n = ins.Nx * ins.Ny
        out.append(ref[off + ell * n:off + (ell + 1) * n].reshape(ins.Ny, ins.Nx))
        off += nl * n
    return out


@pytest.fixture(scope="module", params=["cartesian", "octree", "voronoi"])
def w(request):
    return workloads.two_dust_disc(request.param, packages=20000)


def test_optical_depths_bit_exact(w):
    sc = O.scene_from_workload(w)
    rays = random_rays(400, 21, w["grid"].extent)
    for ell in (0, 1):
        got = w["dust"].opticaldepths(rays, ell, SG_FP64)
        ref = np.array([sc.opticaldepth(r[:3], r[3:], ell) for r in rays])
        assert ref.max() > 0.5
        assert np.array_equal(got, ref), ell


def test_packets_against_oracle(w):
    sc = O.scene_from_workload(w)
    n = 20000
    for ell in (0, 1):
        w["instr"].clear()
        w["dust"].clear()
        print("%s: %d FP64 packets at wavelength %d" % (w["name"], n, ell), flush=True)
        c = w["sim"].run_packets(ell, 0, n, n, SG_FP64)
        got = gpu_frames(w, ell)
        labs = w["dust"].Labs()[:, ell]
        ref, rc = sc.run_packets(w, ell, 0, n, n, threads=16)
        rlabs = sc.labs.reshape(-1, 2)[:, ell]
        assert c["packets"] == rc[0] == n and c["stuck"] == 0
        assert c["scatterings"] > 1000
        assert abs(int(c["scatterings"]) - int(rc[3])) <= 1e-4 * rc[3] + 2
        for g, o in zip(got, oracle_frames(w, ref, ell, 2)):
            assert o.sum() > 0
            assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()
            hit = o != 0
            assert np.mean(np.abs(g - o)[hit] <= 1e-9 * np.abs(o[hit])) > 0.99
        assert rlabs.sum() > 0
        assert abs(labs.sum() - rlabs.sum()) <= 1e-6 * rlabs.sum()
        close = np.abs(labs - rlabs) <= 1e-9 * np.abs(rlabs) + 1e-300
        assert np.mean(close) > 0.99


def test_mixed_against_fp64(w):
    n = 200000
    w["instr"].clear()
    w["sim"].run_packets(0, 0, n, n, SG_FP64)
    f64 = gpu_frames(w, 0)
    w["instr"].clear()
    cm = w["sim"].run_packets(0, 0, n, n, SG_MIXED)
    assert cm["stuck"] == 0
    for a, b in zip(gpu_frames(w, 0), f64):
        assert abs(a.sum() - b.sum()) <= 2e-3 * b.sum()
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 0.05


def test_batch_with_multi_component_members():
    members = [workloads.two_dust_disc("cartesian", grid_n=16, pixels=32, packages=5000, store_abs=False),
               workloads.disc_cartesian(grid_n=16, pixels=32, packages=5000),
               workloads.two_dust_disc("cartesian", grid_n=20, pixels=32, packages=7000, store_abs=False)]
    counts = [5000, 5000, 7000]
    for precision in (SG_FP64, SG_MIXED):
        singles = []
        for m, n in zip(members, counts):
            m["instr"].clear()
            m["sim"].run_packets(0, 0, n, n, precision)
            singles.append([m["instr"].frame(i)[0].copy() for i in range(len(m["instr"].instruments))])
        for m in members:
            m["instr"].clear()
        S.run_batch([m["sim"] for m in members], counts, precision=precision)
        for m, ref in zip(members, singles):
            for i, r in enumerate(ref):
                assert r.sum() > 0
                np.testing.assert_allclose(m["instr"].frame(i)[0], r, rtol=1e-12, atol=0)


def test_static_run_refuses_multi_component():
    w = workloads.two_dust_disc("cartesian", grid_n=8, pixels=16, packages=100)
    with pytest.raises(S.FatalError, match="several components"):
        w["sim"].run_packets_static(0, 0, 100, 100, SG_FP64)


@pytest.mark.parametrize("grid", ["cartesian", "octree"])
def test_ski_file_with_two_dust_components(tmp_path, grid):
    """A ski file whose CompDustDistribution holds two DustComps runs as the two-component system: FP64
    packets against the oracle built from the same objects."""
    from skirt7_amd import ski
    from tests.test_ski_fits import SUN, two_dust_ski
    d = tmp_path / "dat"
    (d / "SED" / "Sun").mkdir(parents=True)
    (d / "SED" / "Sun" / "SunSED.dat").write_text(SUN)
    sim = ski.Simulation(ski.Spec(two_dust_ski(grid), str(d)))
    assert sim.dust.ncomp == 2
    w = dict(kind="cartesian" if grid == "cartesian" else "octree", grid=sim.grid, rho=sim.rho, mix=sim.mix,
             comps=sim.stars.components, instr=sim.instr, sim=sim.sim)
    sc = O.scene_from_workload(w)
    n = 20000
    sim.instr.clear()
    c = sim.sim.run_packets(0, 0, n, n, SG_FP64)
    ref, rc = sc.run_packets(w, 0, 0, n, n, threads=16)
    assert c["packets"] == rc[0] == n and c["scatterings"] > 0
    for i, o in enumerate(oracle_frames(w, ref, 0, 1)):
        g = sim.instr.frame(i)[0]
        assert o.sum() > 0
        assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()
