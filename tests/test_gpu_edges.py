"""Edge cases of the transport boundary on the GPU: empty inputs, rays that miss the grid, packet indices
beyond 32 bits, limits and the reference's error paths.

The reference's own checks: DustSystem::fillOpticalDepth throws when the optical depth along a path is
negative, infinite or NaN (DustSystem.cpp:952-955); a path that misses the grid is empty
(CartesianDustGrid.cpp:136-283, DustGridPath::clear).
"""
import numpy as np
import pytest

from skirt7_amd import skirt as S
from skirt7_amd import workloads
from skirt7_amd.capi import SG_FP64, SG_MIXED, SG_NEIGHBOR
from tests import oracle_lib as O
from tests.helpers import EXT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0_tiny():
    return workloads.disc_cartesian(grid_n=10, pixels=24, packages=3000)


def frames(w):
    return [w["instr"].frame(i)[0].copy() for i in range(len(w["instr"].instruments))]


@pytest.mark.parametrize("precision", [SG_FP64, SG_MIXED])
def test_zero_packets_change_nothing(c0_tiny, precision):
    w = c0_tiny
    w["instr"].clear()
    w["sim"].run_packets(0, 0, 500, 500, precision)
    before = frames(w)
    c = w["sim"].run_packets(0, 500, 0, 500, precision)
    assert c["packets"] == 0 and c["own_crossings"] == 0 and c["peeloffs"] == 0
    for a, b in zip(before, frames(w)):
        assert np.array_equal(a, b)


def test_empty_ray_batches():
    g = S.CartesianDustGrid.uniform(EXT, 4, 4, 4)
    counts, m, ds = g.paths(np.zeros((0, 6)), 16, SG_FP64)
    assert len(counts) == 0
    assert len(g.whichcell(np.zeros((0, 3)))) == 0


@pytest.mark.parametrize("kind", ["cartesian", "octree", "voronoi"])
def test_rays_that_miss_the_grid_have_empty_paths(kind):
    """Starting outside and pointing away, or passing beside the box: no segment at all (also no entry
    segment), as the oracle says; optical depth zero."""
    if kind == "cartesian":
        g = S.CartesianDustGrid.uniform(EXT, 6, 5, 4)
        sc = O.Scene().cartesian(g.extent, g.xv, g.yv, g.zv)
    elif kind == "octree":
        flags = np.zeros(9, np.uint8)
        flags[0] = 1
        g = S.OctTreeDustGrid(EXT, flags, SG_NEIGHBOR)
        sc = O.Scene().tree(8, EXT, flags, SG_NEIGHBOR)
    else:
        rng = np.random.default_rng(3)
        sites = rng.uniform(-15, 15, (40, 3))
        g = S.VoronoiDustGrid(EXT, sites)
        sc = O.Scene().voronoi(EXT, g.sites, g.nb_offsets, g.nb_ids)
    rays = np.array([[30.0, 0.0, 0.0, 1.0, 0.0, 0.0],      # outside, moving away
                     [-30.0, 25.0, 0.0, 1.0, 0.0, 0.0],    # passes beside the box
                     [0.0, 0.0, 30.0, 0.0, 0.6, 0.8],      # above, moving up and sideways
                     [25.0, 25.0, 25.0, 0.0, 0.0, -1.0]])  # above a corner, moving down past it
    counts, _, _ = g.paths(rays, 64, SG_FP64)
    for i, rk in enumerate(rays):
        om, _ = sc.path(rk[:3], rk[3:])
        assert counts[i] == len(om) == 0
    d = S.DustSystem(g, np.ones(g.numCells()), S.SimpleOligoDustMix([1.0], [0.5], [0.5]))
    assert np.array_equal(d.opticaldepths(rays), np.zeros(len(rays)))


def test_packet_indices_beyond_32_bits(c0_tiny):
    """Philox counters carry the packet index in two words: packets past 2^32 follow the oracle's
    histories too."""
    w = c0_tiny
    first, n = (1 << 32) + 12345, 3000
    w["instr"].clear()
    c = w["sim"].run_packets(0, first, n, n, SG_FP64)
    gpu = np.concatenate([f.ravel() for f in frames(w)])
    sc = O.scene_from_workload(w)
    ref, rc = sc.run_packets(w, 0, first, n, n)
    assert c["packets"] == rc[0] == n
    assert abs(gpu.sum() - ref.sum()) <= 1e-9 * ref.sum()
    # and they are other packets than those at first - 2^32
    w["instr"].clear()
    w["sim"].run_packets(0, first - (1 << 32), n, n, SG_FP64)
    low = np.concatenate([f.ravel() for f in frames(w)])
    assert not np.array_equal(low, gpu)


@pytest.mark.parametrize("precision", [SG_FP64, SG_MIXED])
def test_nan_optical_depth_is_a_fatal_error(precision):
    """An opacity that makes the optical depth NaN: the run finishes and reports the reference's error."""
    grid = S.CartesianDustGrid.uniform(EXT, 4, 4, 4)
    mix = S.SimpleOligoDustMix([float("nan")], [0.6], [0.6])
    ds = S.DustSystem(grid, np.full(grid.numCells(), 1e-3), mix)
    comps = [S.OligoStellarComp(S.ExpDiskGeometry(4.0, 0.4), [1e10])]
    ss = S.StellarSystem(comps)
    ins = S.InstrumentSystem([S.FrameInstrument(1e4, 0.0, 0.0, 0.0, 8, 8, 40.0, 40.0)], 1)
    sim = S.OligoMonteCarloSimulation(ds, ss, ins, packages=1000)
    with pytest.raises(S.FatalError, match="optical depth"):
        sim.run_packets(0, 0, 1000, 1000, precision)


def test_invalid_inputs_are_rejected():
    grid = S.CartesianDustGrid.uniform(EXT, 4, 4, 4)
    rho = np.full(grid.numCells(), 1e-3)
    rho[5] = float("nan")
    with pytest.raises(S.FatalError, match="density"):
        S.DustSystem(grid, rho, S.SimpleOligoDustMix([1.0], [0.6], [0.6]))
    with pytest.raises(S.FatalError, match="instruments"):
        S.InstrumentSystem([S.FrameInstrument(1e4, 0.0, 0.0, 0.0, 8, 8, 40.0, 40.0)] * 9, 1)
    with pytest.raises(S.FatalError):
        S.set_slots(100)
