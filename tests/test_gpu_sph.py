"""The SPH particle medium and the device tree builders against the oracle (oracle/sph.cpp) and the host
builder.

- The particle grid's borders and every cell's list equal the oracle's.
- SPHDustDistribution::density at random positions: bit for bit.
- SPHDustDistribution::massInBox: bit for bit for boxes over at most eight grid cells (the lists merged in
  particle order, the reference's order of the terms); to 1e-13 relative for larger boxes, whose terms come
  in cell order.
- TreeDustGrid subdivision over the particles on the device (octree and alternating kd-tree, mass-fraction
  and optical-depth criteria): the same flags as the oracle and the same node masses bit for bit (each
  node's list of contributing particles in particle order); cell densities (TreeDustGrid::density) from
  those masses bit for bit, and evaluated anew from the grid lists to 1e-13.
- The dispersion criterion's sampled mode: each decision against the oracle's densities at the sample
  positions of the counter-based stream.
- The device subdivision by geometry density samples: the same flags as the host builder.
- k-th neighbour distances against scipy's k-d tree.
- Transport over a kd-tree built on a particle medium: FP64 packets against the oracle.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from skirt7_amd import skirt as S
from skirt7_amd import workloads
from skirt7_amd.capi import SG_FP64
from tests import oracle_lib as O
from tests.test_sph_pins import _tree_nodes, particles

pytestmark = pytest.mark.gpu

EXT = np.array([-10.0, -10, -10, 10, 10, 10])


@pytest.fixture(scope="module")
def medium():
    rows = particles(20000, seed=21)
    return rows, S.SPHDustDistribution(rows, 0.4), O.SphMedium(rows, 0.4)


def _ncells(g, boxes):
    lo = [np.searchsorted(g[a], boxes[:, a], side="right") - 1 for a in range(3)]
    hi = [np.searchsorted(g[a], boxes[:, a + 3], side="right") - 1 for a in range(3)]
    k = np.ones(len(boxes), int)
    for a in range(3):
        k *= np.clip(hi[a], 0, 19) - np.clip(lo[a], 0, 19) + 1
    return k


def test_grid_and_lists(medium):
    rows, dev, ora = medium
    assert dev.count == ora.count() == len(rows)
    assert np.array_equal(dev.grid_borders, ora.grid())
    assert dev.mass() == ora.mass()
    rng = np.random.default_rng(2)
    for (i, j, k) in [(0, 0, 0), (19, 19, 19)] + [tuple(rng.integers(0, 20, 3)) for _ in range(40)]:
        assert np.array_equal(dev.cell_list(i, j, k), ora.cell_list(i, j, k)), (i, j, k)


def test_density_bit_exact(medium):
    rows, dev, ora = medium
    rng = np.random.default_rng(3)
    pos = np.concatenate([rng.uniform(-10, 10, (20000, 3)), rows[:3000, :3] + rng.normal(0, 0.1, (3000, 3))])
    got, ref = dev.density(pos), ora.density(pos)
    assert np.count_nonzero(ref) > 5000
    assert np.array_equal(got, ref)


def test_mass_in_box(medium):
    rows, dev, ora = medium
    rng = np.random.default_rng(4)
    c = rng.uniform(-9, 9, (6000, 3))
    w = np.exp(rng.uniform(np.log(0.05), np.log(12.0), (6000, 3)))
    boxes = np.concatenate([c - w / 2, c + w / 2], axis=1)
    got, ref = dev.massInBox(boxes), ora.mass_in_box(boxes)
    k = _ncells(ora.grid(), boxes)
    small = k <= 8
    assert small.sum() > 1000 and (~small).sum() > 1000
    assert np.count_nonzero(ref[small]) > 500
    assert np.array_equal(got[small], ref[small])
    assert np.allclose(got[~small], ref[~small], rtol=1e-13, atol=0)


@pytest.mark.parametrize("arity,mmf,mod", [(8, 2e-4, 0.0), (2, 1e-4, 0.0), (8, 0.0, 8.0), (2, 1e-4, 8.0)])
def test_subdivision_against_oracle(medium, arity, mmf, mod):
    rows, dev, ora = medium
    minlevel, maxlevel = (1, 7) if arity == 8 else (3, 21)
    flags, masses = ora.subdivide(arity, EXT, minlevel, maxlevel, mmf, mod)
    cls = S.OctTreeDustGrid if arity == 8 else S.BinTreeDustGrid
    print("subdivision arity %d: %d nodes in the oracle" % (arity, len(flags)), flush=True)
    grid, dmass = cls.subdivide_sph(EXT, dev, minlevel=minlevel, maxlevel=maxlevel, maxmassfraction=mmf,
                                    maxopticaldepth=mod, search=S.SG_TOPDOWN, masses=True)
    assert len(flags) > 2000
    assert np.array_equal(grid.subdivided, flags)
    boxes, levels = _tree_nodes(flags, arity, EXT)
    # every node beyond minlevel has its mass, the reference's sum term for term: equal to the oracle's
    # where the criteria evaluated it, and to the oracle's mass in box at maxlevel
    assert np.array_equal(dmass >= 0, levels > minlevel)
    ev = masses >= 0
    assert np.array_equal(dmass[ev], masses[ev])
    top = levels == maxlevel
    assert np.array_equal(dmass[top], ora.mass_in_box(boxes[top]))
    # cell densities: each leaf's mass in its box over its volume, from the node masses or evaluated anew
    leaves = boxes[flags == 0]
    ref = ora.mass_in_box(leaves) / np.prod(leaves[:, 3:] - leaves[:, :3], axis=1)
    rho = S.sph_density(grid, dev, dmass)
    assert len(rho) == len(leaves)
    assert np.array_equal(rho, ref)
    assert np.allclose(S.sph_density(grid, dev), ref, rtol=1e-13, atol=0)


def test_subdivision_reports_its_size(medium):
    """A flag buffer too small: the build runs on, returns -5 and the full node count; the retry fits."""
    import ctypes
    from skirt7_amd import capi
    rows, dev, ora = medium
    flags, _ = ora.subdivide(2, EXT, 3, 21, 1e-4)
    e = np.ascontiguousarray(EXT)
    small = np.zeros(100, np.uint8)
    n = ctypes.c_int64()
    st = capi.lib().sg_tree_subdivide_sph(2, capi.dptr(e), dev._h, 3, 21, 1e-4, 0.0, 0.0, 100, 4357,
                                          small.ctypes.data_as(capi._U8), None, 100, ctypes.byref(n))
    assert st == -5 and n.value == len(flags)
    full = np.zeros(n.value, np.uint8)
    assert capi.lib().sg_tree_subdivide_sph(2, capi.dptr(e), dev._h, 3, 21, 1e-4, 0.0, 0.0, 100, 4357,
                                            full.ctypes.data_as(capi._U8), None, n.value, ctypes.byref(n)) == 0
    assert np.array_equal(full, flags)


def _sample_positions(boxes, ids, nsamples, seed):
    """The setup samplers' counter-based positions: Philox4x32-10 on (id lo, id hi, sample, 0x5e7), key
    (seed, 0x9d2c5680), uniforms (c + 1/2) / 2^32."""
    out = np.zeros((len(ids), nsamples, 3))
    key = np.array([seed, 0x9D2C5680], np.uint32)
    for a, (b, l) in enumerate(zip(boxes, ids)):
        for j in range(nsamples):
            ctr = np.array([l & 0xFFFFFFFF, l >> 32, j, 0x5E7], np.uint32)
            r = np.zeros(4, np.uint32)
            O.lib().or_philox4x32(ctr.ctypes.data_as(O._U32), key.ctypes.data_as(O._U32), r.ctypes.data_as(O._U32))
            u = (r[:3].astype(np.float64) + 0.5) / 4294967296.0
            out[a, j] = b[:3] + u * (b[3:] - b[:3])
    return out


def test_dispersion_criterion_samples(medium):
    rows, dev, ora = medium
    nsamples, seed, mdd = 20, 77, 0.9
    grid = S.OctTreeDustGrid.subdivide_sph(EXT, dev, minlevel=1, maxlevel=5, maxmassfraction=0.0,
                                           maxdensdispfraction=mdd, nsamples=nsamples, seed=seed,
                                           search=S.SG_TOPDOWN)
    flags = grid.subdivided
    boxes, levels = _tree_nodes(flags, 8, EXT)
    ev = np.flatnonzero((levels > 1) & (levels < 5))
    assert len(ev) > 100
    pos = _sample_positions(boxes[ev], ev, nsamples, seed)
    d = ora.density(pos.reshape(-1, 3)).reshape(len(ev), nsamples)
    hi, lo = d.max(axis=1), d.min(axis=1)
    disp = np.where(hi > 0, (hi - lo) / np.where(hi > 0, hi, 1), 0.0)
    assert np.array_equal(flags[ev].astype(bool), disp >= mdd)
    assert 0 < flags[ev].sum() < len(ev)


@pytest.mark.parametrize("arity,kw", [(8, dict(maxmassfraction=1e-4)),
                                      (8, dict(maxmassfraction=0.0, maxopticaldepth=3e2)),
                                      (2, dict(maxmassfraction=3e-5, maxdensdispfraction=0.6))])
def test_device_geometry_subdivision_equals_host(arity, kw):
    comps, dgeom, mix, scale = workloads._galaxy(True, 2.0)
    cls = S.OctTreeDustGrid if arity == 8 else S.BinTreeDustGrid
    levels = dict(minlevel=2, maxlevel=7) if arity == 8 else dict(minlevel=3, maxlevel=20)
    host = cls.subdivide(workloads.EXTENT, dgeom, scale, total_mass=scale, nsamples=50, seed=11, **levels, **kw)
    dev = cls.subdivide_device(workloads.EXTENT, dgeom, scale, total_mass=scale, nsamples=50, seed=11, **levels,
                               **kw)
    assert len(host.subdivided) > 5000
    assert np.array_equal(dev.subdivided, host.subdivided)


def test_full_size_octree_device_setup_equals_host():
    """The spiral octree at its full size (718509 leaves) built and sampled on the device: the host builder's
    flags, and its cell densities to the rounding of exp/log/pow."""
    host = workloads.spiral_octree(pixels=32, packages=1e4, device_setup=False)
    dev = workloads.spiral_octree(pixels=32, packages=1e4, device_setup=True)
    assert host["grid"].numCells() == 718509
    assert np.array_equal(dev["grid"].subdivided, host["grid"].subdivided)
    assert np.allclose(dev["rho"], host["rho"], rtol=1e-12, atol=0)


def test_knn_against_kdtree():
    rng = np.random.default_rng(9)
    pos = np.concatenate([rng.normal(0, 1, (60000, 3)), rng.uniform(-6, 6, (40000, 3))])
    pos[5] = pos[6]   # a duplicate point: distance 0 counts as a neighbour
    for k in (1, 8, 32):
        got = S.knn_distance(pos, k)
        ref = cKDTree(pos).query(pos, k + 1)[0][:, -1]
        # the query's own point is its first hit except for the duplicate pair
        ref5 = np.sort(np.sqrt(np.sum((pos - pos[5]) ** 2, axis=1)))[k]
        ref[5] = ref[6] = ref5
        assert np.allclose(got, ref, rtol=1e-14, atol=0), k


def test_particle_kdtree_transport_against_oracle():
    w = workloads.lognormal_kdtree(nvox=32, particles=30000, maxmassfraction=1e-3, pixels=32, packages=6000)
    assert w["grid"].numCells() > 1000
    n = 6000
    w["instr"].clear()
    c = w["sim"].run_packets(0, 0, n, n, SG_FP64)
    sc = O.scene_from_workload(w)
    ref, rc = sc.run_packets(w, 0, 0, n, n, threads=16)
    assert c["packets"] == rc[0] == n and c["scatterings"] > 100
    assert abs(int(c["scatterings"]) - int(rc[3])) <= 2
    off = 0
    for i, ins in enumerate(w["instr"].instruments):
        npx = ins.Nx * ins.Ny
        o = ref[off:off + npx].reshape(ins.Ny, ins.Nx)
        off += npx
        g = w["instr"].frame(i)[0]
        assert o.sum() > 0
        assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()


@pytest.mark.parametrize("grid", ["bintree", "octree"])
def test_ski_file_with_sph_distribution(tmp_path, grid):
    """A ski file with an SPHDustDistribution over a particle file and a tree grid: the tree equals the oracle's
    subdivision of the same particles (converted to the transport's units), its densities the oracle's mass in
    each cell over its volume, and FP64 packets the oracle's."""
    from skirt7_amd import ski
    from skirt7_amd.output import MSUN, PC
    from tests.test_ski_fits import SUN, sph_ski, write_particles
    d = tmp_path / "dat"
    (d / "SED" / "Sun").mkdir(parents=True)
    (d / "SED" / "Sun" / "SunSED.dat").write_text(SUN)
    rows = write_particles(tmp_path / "particles.dat")
    sim = ski.Simulation(ski.Spec(sph_ski(grid), str(d), str(tmp_path)))
    conv = rows.copy()
    conv[:, 4] *= MSUN / PC ** 2
    ora = O.SphMedium(conv, 0.4, 2e4)
    assert sim.sph.count == ora.count() < len(rows)
    ext = np.array([-1e4, -1e4, -1e4, 1e4, 1e4, 1e4])
    arity, lo, hi = (2, 3, 24) if grid == "bintree" else (8, 1, 7)
    flags, masses = ora.subdivide(arity, ext, lo, hi, 2e-3)
    assert len(flags) > 500 and np.array_equal(sim.grid.subdivided, flags)
    boxes, _ = _tree_nodes(flags, arity, ext)
    leaves = boxes[flags == 0]
    assert np.array_equal(sim.rho, ora.mass_in_box(leaves) / np.prod(leaves[:, 3:] - leaves[:, :3], axis=1))
    w = dict(kind="bintree" if grid == "bintree" else "octree", grid=sim.grid, rho=sim.rho, mix=sim.mix,
             comps=sim.stars.components, instr=sim.instr, sim=sim.sim)
    sc = O.scene_from_workload(w)
    n = 20000
    sim.instr.clear()
    c = sim.sim.run_packets(0, 0, n, n, SG_FP64)
    ref, rc = sc.run_packets(w, 0, 0, n, n, threads=16)
    assert c["packets"] == rc[0] == n and c["scatterings"] > 0
    off = 0
    for i, ins in enumerate(sim.instr.instruments):
        npx = ins.Nx * ins.Ny
        o = ref[off:off + npx].reshape(ins.Ny, ins.Nx)
        off += npx
        g = sim.instr.frame(i)[0]
        assert o.sum() > 0
        assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()
