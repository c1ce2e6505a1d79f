"""The benchmark's own workload at its full size: the spiral-galaxy octree with 718509 leaves (maximum level
8, face-on optical depth 2, Neighbor search) and four 512x512 frames.

At this size the oracle still checks in seconds what it can: exact FP64 walks on rays through the whole tree
(bit for bit), and a few thousand packets followed packet by packet on the same Philox streams. The rest is
checked through properties that do not depend on the size: every packet counted once and none stuck, a run
split into two packet ranges equal to the run in one piece (the streams are keyed by the packet index), and
mixed precision equal to FP64 on the same streams up to the rare packets whose history a rounding flips.
The Voronoi workload's full grid (1e6 sites drawn from the spiral disc's density) is checked the same way
for its walks and its point location.
"""
import numpy as np
import pytest

from skirt7_amd import workloads
from skirt7_amd.capi import SG_FP64, SG_MIXED
from tests import oracle_lib as O
from tests.helpers import random_rays
from tests.test_gpu_paths import assert_mixed_voronoi_exact, assert_paths_equal
from tests.test_gpu_transport import frames, oracle_frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    w = workloads.spiral_octree(packages=1e6)
    assert w["grid"].numCells() == 718509
    return w, O.scene_from_workload(w)


def snapshot(w):
    return [f.copy() for f in frames(w)]


def test_fp64_walks_bitexact_at_full_size(full):
    w, sc = full
    counts = assert_paths_equal(w["grid"], sc, random_rays(800, 31, w["grid"].extent), maxseg=8192)
    assert counts.max() > 100


def test_fp64_transport_matches_oracle_at_full_size(full):
    w, sc = full
    n, first = 3000, 123456789
    w["instr"].clear()
    c = w["sim"].run_packets(0, first, n, n, SG_FP64)
    gpu = snapshot(w)
    ref, rc = sc.run_packets(w, 0, first, n, n)
    assert c["packets"] == rc[0] == n and c["stuck"] == 0
    assert abs(int(c["scatterings"]) - int(rc[3])) <= 2
    assert abs(int(c["own_crossings"]) - int(rc[1])) <= 0.001 * rc[1]
    for g, o in zip(gpu, oracle_frames(w, ref)):
        assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()
        hit = o != 0
        assert np.mean(np.abs(g - o)[hit] <= 1e-9 * np.abs(o[hit])) > 0.99


@pytest.mark.parametrize("precision", [SG_FP64, SG_MIXED])
def test_split_ranges_equal_one_run(full, precision):
    w, _ = full
    n = 400000
    w["instr"].clear()
    c = w["sim"].run_packets(0, 0, n, n, precision)
    assert c["packets"] == n and c["stuck"] == 0
    whole = snapshot(w)
    w["instr"].clear()
    a = w["sim"].run_packets(0, 0, 150000, n, precision)
    b = w["sim"].run_packets(0, 150000, n - 150000, n, precision)
    assert a["packets"] + b["packets"] == n
    assert a["scatterings"] + b["scatterings"] == c["scatterings"]
    for x, y in zip(snapshot(w), whole):
        assert np.allclose(x, y, rtol=1e-12, atol=0.0)


def test_mixed_equals_fp64_on_the_same_streams(full):
    w, _ = full
    n = 400000
    w["instr"].clear()
    w["sim"].run_packets(0, 0, n, n, SG_FP64)
    f64 = snapshot(w)
    w["instr"].clear()
    c = w["sim"].run_packets(0, 0, n, n, SG_MIXED)
    assert c["stuck"] == 0
    f32 = snapshot(w)
    for a, b in zip(f32, f64):
        assert abs(a.sum() - b.sum()) <= 1e-3 * b.sum(), (a.sum(), b.sum())
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 0.03, rel


def test_voronoi_full_size_walks_and_location():
    w = workloads.spiral_voronoi(nlambda=1, pixels=32, packages=1e4)
    g = w["grid"]
    assert g.numCells() == 1000000
    sc = O.Scene().voronoi(g.extent, g.sites, g.nb_offsets, g.nb_ids)
    rng = np.random.default_rng(41)
    lo, hi = g.extent[:3], g.extent[3:]
    pts = np.concatenate([g.sites[rng.integers(0, len(g.sites), 500)] + rng.normal(0.0, 1e-3, (500, 3)),
                          lo + rng.random((500, 3)) * (hi - lo)])
    assert np.array_equal(g.whichcell(pts), np.array([sc.whichcell(p) for p in pts]))
    rays = random_rays(300, 42, g.extent, special=False)
    rays[:150, :3] = g.sites[rng.integers(0, len(g.sites), 150)]
    assert_paths_equal(g, sc, rays, maxseg=16384)
    more = random_rays(3000, 43, g.extent, special=False)
    more[:1500, :3] = g.sites[rng.integers(0, len(g.sites), 1500)]
    assert_mixed_voronoi_exact(g, more, maxseg=16384)
