"""sg_sim_run_static (no device allocation: work memory in the library's static device array) against
sg_sim_run on the same packets, through the C-ABI: equal counters and equal tallies (frames, and the
absorption table where stored) up to the order of the FP64 tally additions, on two calls in a row, the
second with fewer packets, and the refusal of a packet range larger than the array holds."""
import ctypes

import numpy as np
import pytest

from skirt7_amd import capi
from skirt7_amd import workloads
from skirt7_amd.capi import SG_FP64, SG_MIXED

pytestmark = pytest.mark.gpu

EXACT = ("packets", "own_crossings", "peel_crossings", "rewalk_crossings", "scatterings", "peeloffs", "walks",
         "stuck", "fp64_recomputes")
MAX_PACKETS = 1 << 19


def _tallies(w):
    ins = w["instr"]
    f = [ins.frame(i).copy() for i in range(len(ins.instruments))]
    a = w["dust"].Labs() if w.get("store_abs") else None
    return f, a


def _clear(w):
    w["instr"].clear()
    if w.get("store_abs"):
        w["dust"].clear()


def _run(w, static, ell, first, count, precision):
    c = np.zeros(capi.SG_NCOUNTERS, np.uint64)
    fn = capi.lib().sg_sim_run_static if static else capi.lib().sg_sim_run
    rc = fn(w["sim"]._h, ell, int(first), int(count), int(w["packages"]), precision,
            c.ctypes.data_as(capi._U64))
    assert rc == 0, capi.lib().sg_last_error().decode()
    return dict(zip(capi.COUNTER_NAMES, (int(x) for x in c)))


def _compare(w, ell, first, count, precision):
    _clear(w)
    ref = _run(w, False, ell, first, count, precision)
    f0, a0 = _tallies(w)
    _clear(w)
    got = _run(w, True, ell, first, count, precision)
    f1, a1 = _tallies(w)
    for k in EXACT:
        assert got[k] == ref[k], (k, got[k], ref[k])
    assert ref["packets"] == count
    for x, y in zip(f1, f0):
        assert y.sum() > 0
        np.testing.assert_allclose(x, y, rtol=1e-11, atol=1e-300)
    if a0 is not None:
        assert a0.sum() > 0
        np.testing.assert_allclose(a1, a0, rtol=1e-11, atol=1e-300)


@pytest.fixture(scope="module")
def octree():
    return workloads.spiral_octree(tau=2.0, maxlevel=6, pixels=96, packages=600000, nsamples=20)


@pytest.fixture(scope="module")
def cartesian():
    return workloads.disc_cartesian(grid_n=20, pixels=64, packages=300000, nsamples=20)


@pytest.fixture(scope="module")
def voronoi():
    return workloads.spiral_voronoi(nsites=20000, nlambda=3, pixels=64, packages=200000, store_abs=True)


@pytest.mark.parametrize("precision", [SG_MIXED, SG_FP64])
def test_static_equals_run_octree(octree, precision):
    print("octree: %d then %d packets, precision %d" % (MAX_PACKETS, 70000, precision), flush=True)
    _compare(octree, 0, 0, MAX_PACKETS, precision)
    _compare(octree, 0, MAX_PACKETS, 70000, precision)


def test_static_equals_run_cartesian(cartesian):
    print("cartesian: 200000 then 5000 packets, FP64", flush=True)
    _compare(cartesian, 0, 0, 200000, SG_FP64)
    _compare(cartesian, 0, 200000, 5000, SG_FP64)


def test_static_equals_run_voronoi_absorption(voronoi):
    print("voronoi with absorption: 150000 then 20000 packets, mixed", flush=True)
    _compare(voronoi, 1, 0, 150000, SG_MIXED)
    _compare(voronoi, 2, 150000, 20000, SG_MIXED)


def test_static_refuses_a_range_beyond_its_array(cartesian):
    c = np.zeros(capi.SG_NCOUNTERS, np.uint64)
    rc = capi.lib().sg_sim_run_static(cartesian["sim"]._h, 0, 0, MAX_PACKETS + 1, int(cartesian["packages"]),
                                      SG_FP64, c.ctypes.data_as(capi._U64))
    assert rc != 0
    assert "at most 524288 packets per call" in capi.lib().sg_last_error().decode()


def test_static_zero_packets(cartesian):
    _clear(cartesian)
    got = _run(cartesian, True, 0, 0, 0, SG_MIXED)
    assert got["packets"] == 0 and got["walks"] == 0
    f, _ = _tallies(cartesian)
    assert all(x.sum() == 0 for x in f)
