"""Batched population runs and the FitSKIRT evaluation step on the GPU.

A member's result in a batched run (one persistent launch over many members' packet ranges) equals its
own single run: each packet has its own Philox stream keyed by the member's seed, so only the order of the
tally additions differs. The fit (PSF convolution + golden-section luminosity) is compared with the oracle's
restatement of Convolution::nested_loop and GoldenSection::optimize on the same frames.
"""
import numpy as np
import pytest

from skirt7_amd import skirt as S
from skirt7_amd import workloads
from skirt7_amd.capi import SG_FP64, SG_MIXED
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

EXACT = ("packets", "own_crossings", "peel_crossings", "scatterings", "peeloffs", "walks")


@pytest.fixture(scope="module")
def pop():
    return workloads.fit_population(members=6, packets=(2.0e3, 3.0e4), grid_n=(8, 24), pixels=(64, 32), nsamples=2)


def _single_runs(ms, precision):
    frames, tot = [], {k: 0 for k in EXACT}
    for m in ms:
        m["instr"].clear()
        c = m["sim"].run_packets(0, 0, m["packages"], m["packages"], precision)
        frames.append(m["instr"].frame(0)[0].copy())
        for k in EXACT:
            tot[k] += c[k]
    return frames, tot


@pytest.mark.parametrize("precision", [SG_FP64, SG_MIXED])
def test_batch_equals_single_runs(pop, precision):
    ms = pop["members"]
    print("members", [(m["params"]["grid"], m["packages"]) for m in ms], flush=True)
    single, tot = _single_runs(ms, precision)
    for m in ms:
        m["instr"].clear()
    cb = S.run_batch([m["sim"] for m in ms], [m["packages"] for m in ms], precision=precision)
    for k in EXACT:
        assert cb[k] == tot[k], k
    assert cb["packets"] == sum(m["packages"] for m in ms)
    for m, f in zip(ms, single):
        g = m["instr"].frame(0)[0]
        assert f.sum() > 0
        assert np.allclose(g, f, rtol=1e-11, atol=1e-300)


def test_batch_splits_members_across_rounds(pop):
    """A member larger than one round is split into ranges; the result is still its single run."""
    ms = pop["members"][:2]
    counts = [9_000_000, 3]   # the first member spans two rounds of 2^23 packets
    m0 = ms[0]
    m0["instr"].clear()
    ms[1]["instr"].clear()
    S.set_batch_alone(0)      # every member in the shared launches
    try:
        S.run_batch([m["sim"] for m in ms], counts, precision=SG_MIXED)
    finally:
        S.set_batch_alone(1 << 21)
    g = m0["instr"].frame(0)[0].copy()
    m0["instr"].clear()
    m0["sim"].run_packets(0, 0, counts[0], counts[0], SG_MIXED)
    f = m0["instr"].frame(0)[0]
    assert np.allclose(g, f, rtol=1e-9, atol=1e-300)


@pytest.mark.parametrize("precision", [SG_FP64, SG_MIXED])
def test_batch_runs_large_members_alone(pop, precision):
    """Members at or above the threshold run alone through the event/walk kernels, the others share the
    batched launches: every member's frame and the summed counters are still those of single runs."""
    ms = pop["members"]
    counts = sorted(m["packages"] for m in ms)
    threshold = counts[len(counts) // 2]
    print("threshold", threshold, "counts", [m["packages"] for m in ms], flush=True)
    single, tot = _single_runs(ms, precision)
    for m in ms:
        m["instr"].clear()
    S.set_batch_alone(threshold)
    try:
        cb = S.run_batch([m["sim"] for m in ms], [m["packages"] for m in ms], precision=precision)
    finally:
        S.set_batch_alone(1 << 21)
    for k in EXACT:
        assert cb[k] == tot[k], k
    for m, f in zip(ms, single):
        assert np.allclose(m["instr"].frame(0)[0], f, rtol=1e-11, atol=1e-300)


def test_batch_rejects_mixed_grid_kinds(pop):
    w = workloads.spiral_octree(maxlevel=3, pixels=16, packages=100, nsamples=2)
    with pytest.raises(S.FatalError):
        S.run_batch([pop["members"][0]["sim"], w["sim"]], [10, 10])
    with pytest.raises(S.FatalError):
        S.run_batch([pop["members"][0]["sim"], pop["members"][0]["sim"]], [10, 10])


def test_fit_matches_oracle(pop):
    ms = pop["members"]
    for m in ms:
        m["instr"].clear()
    S.run_batch([m["sim"] for m in ms], [m["packages"] for m in ms], precision=SG_FP64)
    tw = pop["target"]
    tw["instr"].clear()
    tw["sim"].run_packets(0, 0, tw["packages"], tw["packages"], SG_FP64)
    ref, gain = workloads.target_image(tw["instr"].frame(0)[0], 1.0e6)
    ref[:2, :] = 0.0   # a masked band
    kern = pop["kernel"]
    lo, hi = gain * 1e-2, gain * 1e2
    lum, chi2, conv = S.fit_frames([m["instr"] for m in ms], ref, kern, lo, hi, return_frames=True)
    for i, m in enumerate(ms):
        frame = m["instr"].frame(0)[0]
        oc = O.convolve(frame, kern.data)
        ol, ochi, ofr = O.golden_section(ref, oc, lo, hi)
        # same terms summed in another order
        assert np.allclose(conv[i], ofr, rtol=1e-12, atol=1e-300)
        # chi2 sums in another order can flip the last golden-section comparisons, where both trial
        # points are within rounding of the same chi2: both results stay inside the final bracket
        assert abs(lum[i] - ol) <= 1e-6 * ol, (lum[i], ol)
        assert abs(chi2[i] - ochi) <= 1e-9 * ochi, (chi2[i], ochi)


def test_fit_full_frame_size():
    """256 x 128 frames and the 6-pixel FWHM-3 PSF of the population workload."""
    pop = workloads.fit_population(members=3, packets=(5.0e3, 1.0e4), grid_n=(16, 24), nsamples=2)
    ms = pop["members"]
    S.run_batch([m["sim"] for m in ms], [m["packages"] for m in ms], precision=SG_MIXED)
    frame0 = ms[0]["instr"].frame(0)[0]
    assert frame0.shape == (128, 256)
    ref, gain = workloads.target_image(frame0, 1.0e6, seed=11)
    lum, chi2 = S.fit_frames([m["instr"] for m in ms], ref, pop["kernel"], gain * 1e-2, gain * 1e2)
    ol, ochi, _ = O.golden_section(ref, O.convolve(frame0, pop["kernel"].data), gain * 1e-2, gain * 1e2)
    assert abs(lum[0] - ol) <= 1e-6 * ol
    assert abs(chi2[0] - ochi) <= 1e-9 * ochi
    # the target was made from member 0's own (unconvolved) frame: its fitted scale is near the gain
    assert abs(lum[0] - gain) < 0.2 * gain
