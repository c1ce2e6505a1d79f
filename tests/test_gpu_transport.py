"""The full photon-packet life cycle on the GPU against the CPU oracle.

The oracle runs the reference loop (MonteCarloSimulation.cpp:259-297) with the same per-packet
Philox streams, so each packet follows the same history on both sides up to libm rounding:
images agree to near machine precision in FP64. Mixed precision is compared with FP64 on the
same streams (an image-error measurement free of Monte Carlo noise).
"""
import numpy as np
import pytest

from skirt7_amd import workloads
from skirt7_amd.capi import SG_FP64, SG_MIXED
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def frames(w):
    ins = w["instr"]
    return [ins.frame(i)[0] for i in range(len(ins.instruments))]


def oracle_frames(w, ref):
    out, off = [], 0
    for ins in w["instr"].instruments:
        n = ins.Nx * ins.Ny
        out.append(ref[off:off + n].reshape(ins.Ny, ins.Nx))
        off += n
    return out


@pytest.fixture(scope="module")
def c0_small():
    return workloads.disc_cartesian(grid_n=16, pixels=48, packages=20000)


def test_fp64_matches_oracle_packet_by_packet(c0_small):
    w = c0_small
    w["instr"].clear()
    n = 20000
    c = w["sim"].run_packets(0, 0, n, n, SG_FP64)
    gpu = frames(w)
    sc = O.scene_from_workload(w)
    ref, rc = sc.run_packets(w, 0, 0, n, n)
    orc = oracle_frames(w, ref)
    assert c["packets"] == rc[0] == n
    # identical histories: counters match exactly except for rare libm-rounding branch flips
    assert abs(int(c["scatterings"]) - int(rc[3])) <= 2
    assert abs(int(c["own_crossings"]) - int(rc[1])) <= 0.001 * rc[1]
    for g, o in zip(gpu, orc):
        assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()
        diff = np.abs(g - o)
        assert np.mean(diff <= 1e-9 * np.abs(o) + 1e-300) > 0.99


def test_mixed_matches_fp64_image(c0_small):
    w = c0_small
    n = 20000
    w["instr"].clear()
    w["sim"].run_packets(0, 0, n, n, SG_FP64)
    f64 = frames(w)
    w["instr"].clear()
    w["sim"].run_packets(0, 0, n, n, SG_MIXED)
    f32 = frames(w)
    for a, b in zip(f64, f32):
        assert abs(a.sum() - b.sum()) <= 1e-3 * a.sum()
        rms = np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(a ** 2))
        assert rms < 0.01, rms


def test_scheduling_independent(c0_small):
    """Two runs that split the packet range differently give the same image."""
    w = c0_small
    n = 6000
    w["instr"].clear()
    w["sim"].run_packets(0, 0, n, n, SG_FP64)
    a = frames(w)
    w["instr"].clear()
    w["sim"].run_packets(0, 0, 2500, n, SG_FP64)
    w["sim"].run_packets(0, 2500, n - 2500, n, SG_FP64)
    b = frames(w)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-300)


EXACT = ("packets", "own_crossings", "peel_crossings", "rewalk_crossings", "scatterings", "peeloffs", "walks")


@pytest.mark.parametrize("which", ["cartesian", "octree", "voronoi"])
def test_split_and_lane_pipelines_agree(c0_small, which):
    """The event/walk kernels and the single lane-pipeline kernel run the same packet histories: identical
    event and crossing counts, images equal up to the order of the tally additions."""
    from skirt7_amd import skirt as S
    if which == "cartesian":
        w = c0_small
    elif which == "octree":
        w = workloads.spiral_octree(maxlevel=5, pixels=32, packages=8000, nsamples=4)
    else:
        w = workloads.spiral_voronoi(nsites=3000, nlambda=1, pixels=32, packages=8000)
    n = 8000
    try:
        for prec in (SG_FP64, SG_MIXED):
            res = []
            for mode in (0, 1):
                S.set_pipeline(mode)
                w["instr"].clear()
                c = w["sim"].run_packets(0, 0, n, n, prec)
                res.append((c, frames(w)))
            (c0, f0), (c1, f1) = res
            for k in EXACT:
                assert c0[k] == c1[k], (prec, k, c0[k], c1[k])
            for a, b in zip(f0, f1):
                assert np.allclose(a, b, rtol=1e-10, atol=1e-300)
    finally:
        S.set_pipeline(1)


@pytest.mark.parametrize("which", ["octree", "voronoi"])
def test_walk_orders_agree(which):
    """The walk kernel taking a pass's walks as appended, sorted by coherence key, or sorted and claimed by
    XCD segment runs the same packet histories: identical counters, images equal up to the order of the tally
    additions (FP64 and mixed; a run large enough that passes hold many claim chunks per segment)."""
    from skirt7_amd import skirt as S
    if which == "octree":
        w = workloads.spiral_octree(maxlevel=6, pixels=64, packages=200000, nsamples=4)
    else:
        w = workloads.spiral_voronoi(nsites=20000, nlambda=1, pixels=48, packages=100000)
    n = int(w["packages"])
    try:
        S.set_pipeline(1)
        for prec in (SG_FP64, SG_MIXED):
            res = []
            for order in (0, 1, 2):
                S.set_walk_order(order)
                w["instr"].clear()
                c = w["sim"].run_packets(0, 0, n, n, prec)
                res.append((c, frames(w)))
            (c0, f0) = res[0]
            for c, f in res[1:]:
                for k in EXACT:
                    assert c[k] == c0[k], (prec, k, c[k], c0[k])
                for a, b in zip(f, f0):
                    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-300)
            assert c0["peel_crossings"] > 1e5
    finally:
        S.set_walk_order(0)


@pytest.fixture(scope="module")
def c4_small():
    return workloads.lognormal_kdtree(nvox=16, maxmassfraction=2e-4, pixels=32, packages=6000, medium="voxels")


def test_kdtree_fp64_matches_oracle(c4_small):
    """Alternating binary tree with TopDown search over a log-normal field: packet-by-packet parity."""
    w = c4_small
    n = 6000
    w["instr"].clear()
    c = w["sim"].run_packets(0, 0, n, n, SG_FP64)
    gpu = frames(w)
    sc = O.scene_from_workload(w)
    ref, rc = sc.run_packets(w, 0, 0, n, n)
    orc = oracle_frames(w, ref)
    assert c["packets"] == rc[0] == n
    assert abs(int(c["scatterings"]) - int(rc[3])) <= 2
    assert abs(int(c["own_crossings"]) - int(rc[1])) <= 0.001 * rc[1]
    for g, o in zip(gpu, orc):
        assert abs(g.sum() - o.sum()) <= 1e-6 * o.sum()


def test_kdtree_mixed_matches_fp64(c4_small):
    w = c4_small
    n = 6000
    out = []
    for prec in (SG_FP64, SG_MIXED):
        w["instr"].clear()
        w["sim"].run_packets(0, 0, n, n, prec)
        out.append(frames(w))
    for a, b in zip(*out):
        assert abs(a.sum() - b.sum()) <= 2e-3 * a.sum()


def test_autotune_keeps_tallies_and_results(c0_small):
    """The autotuner times the sample with several launch parameter sets and leaves the frames as they
    were; a run with the parameters it picked gives the image of the default parameters."""
    from skirt7_amd import skirt as S
    w = c0_small
    n = 6000
    try:
        w["instr"].clear()
        w["sim"].run_packets(0, 0, n, n, SG_FP64)
        before = [f.copy() for f in frames(w)]
        best = S.autotune(w["sim"], ell=0, first=0, sample=4000, precision=SG_FP64)
        assert best["block"] in (64, 128, 256)
        assert best["compact_lanes"] in (8, 16, 24, 32, 40, 48)
        assert best["slots"] in (1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24)
        assert best["pipeline"] == 1
        assert best["sample_packets_per_s"] > 0
        for a, b in zip(before, frames(w)):
            assert np.array_equal(a, b)
        # choosing the transport organisation too
        best = S.autotune(w["sim"], ell=0, first=0, sample=4000, precision=SG_FP64, choose_pipeline=True)
        assert best["pipeline"] in (0, 1) and S.get_pipeline() == best["pipeline"]
        for a, b in zip(before, frames(w)):
            assert np.array_equal(a, b)
        w["instr"].clear()
        w["sim"].run_packets(0, 0, n, n, SG_FP64)
        for a, b in zip(before, frames(w)):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300)
    finally:
        S.set_tuning(-1, -1, -1)
        S.set_slots(0)
        S.set_pipeline(1)
