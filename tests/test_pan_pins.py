"""Panchromatic dust emission on the CPU: the wavelength grid, the Planck tables, and the oracle's emission
spectra and dust packet launch against closed forms.

- LogWavelengthGrid / PanWavelengthGrid: log-spaced wavelengths whose bins tile [lambda_min, lambda_max].
- PlanckFunction: its integral over a wide fine grid is sigma T^4 / pi (Stefan-Boltzmann, from h, c, k).
- The oracle's spectra (meanintensityv, equilibrium temperature, grey-body emissivity): a cell whose
  absorbed luminosity is that of a radiation field equal to B(T0) has its dust at T0, so its normalised
  spectrum is sigma_abs B(T0) dlambda over its sum; the bolometric absorbed luminosity is the table's row sums.
- The oracle's dust packet launch: with a thin medium, the biased launch (emission bias 0.5) conserves the
  emitted luminosity in the frames, and each pixel of a face-on frame aligned with the cells receives the
  luminosity of the cells projected onto it; the self-absorption launch makes no peel-offs.
"""
import math

import numpy as np

from skirt7_amd import panchromatic as P
from tests import oracle_lib as O


def test_log_wavelength_grid():
    lam, dl = P.log_wavelength_grid(1e-7, 1e-3, 41)
    r = lam[1:] / lam[:-1]
    assert np.allclose(r, 10 ** 0.1, rtol=1e-13)
    assert abs(lam[0] - 1e-7) < 1e-20 and abs(lam[-1] / 1e-3 - 1) < 1e-13
    assert abs(dl.sum() / (1e-3 - 1e-7) - 1) < 1e-12
    # inner bins extend from the geometric mean with the previous point to that with the next
    assert abs(dl[5] - (math.sqrt(lam[5] * lam[6]) - math.sqrt(lam[4] * lam[5]))) < 1e-25


def test_planck_stefan_boltzmann():
    h, c, k = 6.62606957e-34, 2.99792458e8, 1.3806488e-23
    sigma = 2 * math.pi ** 5 * k ** 4 / (15 * h ** 3 * c ** 2)
    lam, dl = P.log_wavelength_grid(1e-8, 1e-1, 4001)
    for T in (20.0, 300.0, 6000.0):
        tot = float(np.sum(P.planck(lam, T) * dl))
        assert abs(tot / (sigma * T ** 4 / math.pi) - 1) < 2e-3, T


def _cell_scene(nl, rho=1e-3):
    sc = O.Scene()
    e = np.linspace(-1.0, 1.0, 3)
    sc.cartesian(np.array([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0]), e, e, e)
    sc.dust(np.full(8, rho), np.ones(nl), np.full(nl, 0.5), np.zeros(nl))
    return sc


def test_oracle_spectra_equilibrium_temperature():
    lam, dl = P.log_wavelength_grid(1e-7, 1e-3, 61)
    mix = P.Benchmark1DDustMix(lam, dl)
    nl = len(lam)
    rho = 1e-3
    sc = _cell_scene(nl, rho)
    vol = 1.0                        # each of the 2 x 2 x 2 cells is a unit cube (grid units)
    len2 = O.PC ** 2 if hasattr(O, "PC") else 3.08567758e16 ** 2
    labs = np.zeros((8, nl))
    temps = [15.0, 40.0, 120.0, 15.0, 40.0, 120.0, 300.0, 1000.0]
    for m, T0 in enumerate(temps):
        labs[m] = P.planck(lam, T0) * (mix.kappaabs * rho) * 4 * math.pi * vol * len2 * dl
    dustlum, bol = sc.dust_spectra(mix, labs, len2=len2)
    assert np.allclose(bol, labs.sum(axis=1), rtol=1e-14)
    for m, T0 in enumerate(temps):
        expect = mix.sigmaabsvv[0] * P.planck(lam, T0) * dl
        expect /= expect.sum()
        big = expect > 1e-6 * expect.max()
        assert abs(dustlum[m].sum() - 1) < 1e-12
        assert np.max(np.abs(dustlum[m][big] / expect[big] - 1)) < 5e-3, (T0, np.max(np.abs(dustlum[m][big] / expect[big] - 1)))
    # the dust part of the absorbed luminosity counts the same as the stellar part
    d2, b2 = sc.dust_spectra(mix, 0.5 * labs, 0.5 * labs, len2=len2)
    assert np.allclose(b2, bol, rtol=1e-14) and np.allclose(d2, dustlum, rtol=1e-12)


def test_oracle_dust_launch():
    nl = 1
    n = 4
    sc = O.Scene()
    e = np.linspace(-2.0, 2.0, n + 1)
    sc.cartesian(np.array([-2.0, -2.0, -2.0, 2.0, 2.0, 2.0]), e, e, e)
    sc.dust(np.full(n ** 3, 1e-9), np.ones(nl), np.full(nl, 0.5), np.zeros(nl))
    # a face-on frame whose pixels are the cells' columns
    params = np.array([1e6, 0.0, 0.0, 0.0, n, n, 4.0, 4.0, 0.0, 0.0])
    sc.instrument(params)
    sc.nlambda, sc.ncells = nl, n ** 3
    rng = np.random.default_rng(4)
    Lv = rng.random(n ** 3) ** 2
    Lv[7] = 0.0
    N = 400000
    mc = np.array([N, 1e4, 0, 0.5, 4357, 0.5, 0, 0.0])
    frames, c = sc.run_dust_packets(mc, 0, 0, N, N, Lv, 0.5, True, False, 3)
    assert int(c[0]) == N
    img = frames.reshape(n, n)
    assert abs(img.sum() / Lv.sum() - 1) < 0.01
    # expected image: each cell's luminosity on the pixel its centre projects to
    centres = np.array([[(e[i] + e[i + 1]) / 2, (e[j] + e[j + 1]) / 2, (e[k] + e[k + 1]) / 2]
                        for i in range(n) for j in range(n) for k in range(n)])
    pix, _ = O.pixelondetector(params, centres)
    expect = np.bincount(pix, weights=Lv, minlength=n * n).reshape(n, n)
    assert np.allclose(img, expect, rtol=0.03, atol=0.01 * Lv.sum() / n ** 2)
    # the self-absorption launch: no peel-offs
    frames2, _ = sc.run_dust_packets(mc, 0, 0, 1000, 1000, Lv, 0.0, False, True, 4)
    assert frames2.sum() == 0.0
