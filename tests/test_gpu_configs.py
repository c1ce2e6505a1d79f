"""The benchmark workloads at their specified sizes, beyond the headline run.

- The spiral-galaxy octree workload at its other face-on optical depths (0.5 and 5) on the full 718509-leaf octree: FP64 packets
  followed packet by packet against the oracle, and mixed precision against FP64 on the same streams.
- The exponential-disc Cartesian workload at its real size (50^3, three 256x256 frames, 1e6 packets): every packet against the
  oracle on the same Philox streams.
- The Voronoi workload's 50 wavelengths on the full 1e6-site mesh: the absorbed-luminosity table (1e6 x 50) and
  the frames of FP64 runs at every wavelength against the oracle packet by packet.
- The particle kd-tree workload's full tree (1.61e7 leaves): FP64 walks bit for bit against the oracle, and the mixed rope walk
  against them.
- The Bookkeeping search method (octrees): FP64 packets against the oracle, mixed against FP64.
- A packet range emitted in several batches equals the same range emitted at once (both pipelines, both
  precisions).
"""
import numpy as np
import pytest

from skirt7_amd import skirt as S
from skirt7_amd import workloads
from skirt7_amd.capi import SG_BOOKKEEPING, SG_FP64, SG_MIXED
from tests import oracle_lib as O
from tests.helpers import random_rays
from tests.test_gpu_paths import assert_paths_equal
from tests.test_gpu_transport import frames, oracle_frames

pytestmark = pytest.mark.gpu

EXACT = ("packets", "own_crossings", "peel_crossings", "rewalk_crossings", "scatterings", "peeloffs", "walks")


def snapshot(w):
    return [f.copy() for f in frames(w)]


def assert_matches_oracle(w, sc, first, n, npp, match=0.99):
    """FP64 packets [first, first+n) at wavelength 0 against the oracle on the same Philox streams."""
    w["instr"].clear()
    c = w["sim"].run_packets(0, first, n, npp, SG_FP64)
    gpu = snapshot(w)
    ref, rc = sc.run_packets(w, 0, first, n, npp, threads=16)
    assert c["packets"] == rc[0] == n and c["stuck"] == 0
    # identical histories up to rare libm-rounding branch flips
    assert abs(int(c["scatterings"]) - int(rc[3])) <= 1e-4 * rc[3] + 2
    assert abs(int(c["own_crossings"]) - int(rc[1])) <= 1e-3 * rc[1] + 10
    for g, o in zip(gpu, oracle_frames(w, ref)):
        assert o.sum() > 0
        assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()
        hit = o != 0
        assert np.mean(np.abs(g - o)[hit] <= 1e-9 * np.abs(o[hit])) > match
    return c


@pytest.mark.parametrize("tau", [0.5, 5.0])
def test_octree_other_optical_depths(tau):
    w = workloads.spiral_octree(tau=tau, packages=1e6)
    assert w["grid"].numCells() == 718509
    sc = O.scene_from_workload(w)
    print("tau %g: 2000 packets FP64 against the oracle" % tau, flush=True)
    c = assert_matches_oracle(w, sc, 777777, 2000, 2000)
    assert c["scatterings"] > 0
    print("tau %g: mixed against FP64 on 300000 packets" % tau, flush=True)
    n = 300000
    w["instr"].clear()
    w["sim"].run_packets(0, 0, n, n, SG_FP64)
    f64 = snapshot(w)
    w["instr"].clear()
    cm = w["sim"].run_packets(0, 0, n, n, SG_MIXED)
    assert cm["stuck"] == 0
    for a, b in zip(snapshot(w), f64):
        assert abs(a.sum() - b.sum()) <= 1e-3 * b.sum()
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 0.03


def test_cartesian_disc_full_size_against_oracle():
    w = workloads.disc_cartesian(packages=1e6)
    assert w["grid"].numCells() == 125000
    assert [i.Nx for i in w["instr"].instruments] == [256, 256, 256]
    sc = O.scene_from_workload(w)
    print("Cartesian 50^3 disc: 1e6 packets FP64 against the oracle", flush=True)
    assert_matches_oracle(w, sc, 0, 1000000, 1000000, match=0.98)


def test_voronoi_fifty_wavelengths_absorption_table():
    w = workloads.spiral_voronoi(nlambda=50, pixels=32, packages=300)
    assert w["grid"].numCells() == 1000000
    sc = O.scene_from_workload(w)
    n = 300
    w["dust"].clear()
    w["instr"].clear()
    print("Voronoi 1e6 sites: %d packets at each of 50 wavelengths, FP64" % n, flush=True)
    for ell in range(50):
        w["sim"].run_packets(ell, 0, n, n, SG_FP64)
    labs = w["dust"].Labs()
    assert labs.shape == (1000000, 50)
    for ell in range(50):
        _, rc = sc.run_packets(w, ell, 0, n, n, threads=16)
        ref = sc.labs.reshape(-1, 50)[:, ell]
        got = labs[:, ell]
        assert rc[0] == n
        assert ref.sum() > 0, ell
        assert abs(got.sum() - ref.sum()) <= 1e-9 * ref.sum(), (ell, got.sum(), ref.sum())
        close = np.abs(got - ref) <= 1e-9 * np.abs(ref) + 1e-300
        assert np.mean(close) > 0.99, ell
    # absorption falls with wavelength (kappa ~ lambda^-1.3 and the albedo falls from 0.7 to 0.4)
    tot = labs.sum(axis=0)
    assert tot[0] > tot[-1] > 0


def test_kdtree_full_size_walks():
    w = workloads.lognormal_kdtree(pixels=64, packages=1e4)
    g = w["grid"]
    assert 1.6e7 < g.numCells() < 1.62e7
    sc = O.scene_from_workload(w)
    rays = random_rays(300, 51, g.extent)
    print("particle kd-tree: 300 FP64 walks on %d leaves" % g.numCells(), flush=True)
    counts = assert_paths_equal(g, sc, rays, maxseg=32768)
    assert counts.max() > 200
    # the mixed rope walk (TopDown tree: the mixed walk follows the ropes) against the exact walks: the same
    # cells except where a rounding flips an edge or corner crossing
    inside = rays[:, :3]
    lo, hi = g.extent[:3], g.extent[3:]
    sel = np.all((inside > lo) & (inside < hi), axis=1)
    c64, m64, d64 = g.paths(rays[sel], 32768, SG_FP64)
    c32, m32, d32 = g.paths(rays[sel], 32768, SG_MIXED)
    same = [np.array_equal(m64[i][:c64[i]], m32[i][:c32[i]]) for i in range(int(sel.sum()))]
    assert np.mean(same) > 0.9
    tot64 = np.array([d64[i][:c64[i]].sum() for i in range(int(sel.sum()))])
    tot32 = np.array([d32[i][:c32[i]].sum() for i in range(int(sel.sum()))])
    assert np.allclose(tot32, tot64, rtol=1e-4)


@pytest.mark.parametrize("pipeline", [1, 0])
@pytest.mark.parametrize("precision", [SG_FP64, SG_MIXED])
def test_emission_in_batches_equals_one_batch(pipeline, precision):
    w = workloads.spiral_octree(tau=2.0, maxlevel=5, pixels=48, packages=5000, nsamples=20)
    n = 5000
    S.set_pipeline(pipeline)
    try:
        w["instr"].clear()
        S.set_emit_batch(0)
        one = w["sim"].run_packets(0, 0, n, n, precision)
        f_one = snapshot(w)
        w["instr"].clear()
        S.set_emit_batch(1000)
        many = w["sim"].run_packets(0, 0, n, n, precision)
        f_many = snapshot(w)
    finally:
        S.set_emit_batch(0)
        S.set_pipeline(1)
    for k in EXACT:
        assert one[k] == many[k], k
    for a, b in zip(f_many, f_one):
        assert b.sum() > 0
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=0.0)


def test_bookkeeping_octree_transport():
    w = workloads.spiral_octree(tau=2.0, maxlevel=6, pixels=64, packages=20000, nsamples=20, search=SG_BOOKKEEPING)
    sc = O.scene_from_workload(w)
    print("bookkeeping octree: 20000 packets FP64 against the oracle", flush=True)
    assert_matches_oracle(w, sc, 0, 20000, 20000)
    n = 200000
    w["instr"].clear()
    w["sim"].run_packets(0, 0, n, n, SG_FP64)
    f64 = snapshot(w)
    w["instr"].clear()
    cm = w["sim"].run_packets(0, 0, n, n, SG_MIXED)
    assert cm["stuck"] == 0
    for a, b in zip(snapshot(w), f64):
        assert abs(a.sum() - b.sum()) <= 1e-3 * b.sum()
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 0.05
