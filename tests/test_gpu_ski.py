"""Ski files run end to end on the GPU: parse, build, transport, calibrate, write FITS.

The written cubes must equal the calibration of the instruments' raw frames (to float32), carry the
units system's keywords, and, without dust and with a truncated stellar disc inside the field of view,
conserve the flux: the surface brightness summed over the pixels' solid angles equals L / (4 pi d^2)
per unit wavelength.
"""
import math

import numpy as np
import pytest

from skirt7_amd import output as O
from skirt7_amd import ski
from skirt7_amd.capi import SG_FP64, SG_MIXED
from tests.test_ski_fits import SUN, ski_text

pytestmark = pytest.mark.gpu


def read_fits(path):
    """Primary image and header values of a FITS file (test-side reader: 80-byte cards, 2880-byte blocks)."""
    raw = open(path, "rb").read()
    hdr, pos = {}, 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        if card.startswith("END"):
            break
        if card[8:10] == "= ":
            v = card[10:].split(" /")[0].strip()
            hdr[card[:8].strip()] = v.strip("'").strip() if v.startswith("'") else v
    pos = (pos + 2879) // 2880 * 2880
    n = int(hdr["NAXIS"])
    dims = [int(hdr["NAXIS%d" % (a + 1)]) for a in range(n)]
    count = int(np.prod(dims))
    data = np.frombuffer(raw[pos:pos + 4 * count], dtype=">f4").astype(np.float64)
    return data.reshape(dims[::-1]), hdr


@pytest.fixture
def resources(tmp_path):
    d = tmp_path / "dat"
    (d / "SED" / "Sun").mkdir(parents=True)
    (d / "SED" / "Sun" / "SunSED.dat").write_text(SUN)
    return str(d)


@pytest.mark.parametrize("grid,precision", [("cartesian", SG_FP64), ("octree", SG_MIXED)])
def test_ski_run_writes_calibrated_frames(tmp_path, resources, grid, precision):
    p = tmp_path / ("model_%s.ski" % grid)
    p.write_text(ski_text(grid=grid, packages="2e4"))
    print("ski run", grid, precision, flush=True)
    s = ski.load(str(p), resources)
    c = s.run(precision)
    assert c["packets"] == 20000 and c["own_crossings"] > 0
    paths = s.write(str(tmp_path), stamp="2026-12-02T21:07:00")
    assert [str(tmp_path / ("model_%s_%s_total.fits" % (grid, n))) for n in ("i0", "i88")] == paths
    sp = s.spec
    for i, (path, f) in enumerate(zip(paths, sp.frames)):
        data, hdr = read_fits(path)
        assert data.shape == (f["pixelsY"], f["pixelsX"])
        assert hdr["BUNIT"] == "W/m2/arcsec2" and hdr["CTYPE1"] == "pc" and hdr["DATE"] == "2026-12-02T21:07:00"
        assert abs(float(hdr["CDELT1"]) - f["fovX"] / f["pixelsX"] / O.PC) < 1e-9
        want = O.calibrate_frame(s.instr.frame(i), sp.lambdas, sp.dlambdas, f["distance"], f["fovX"] / f["pixelsX"],
                                 f["fovY"] / f["pixelsY"], sp.units)[0]
        assert np.array_equal(data, want.astype(np.float32).astype(np.float64))
        assert data.sum() > 0


def test_ski_dust_free_flux_conservation(tmp_path, resources):
    t = ski_text(dust=False, units="SIUnits", style="Wavelength", packages="3e4")
    t = t.replace('<ExpDiskGeometry radialScale="4 kpc" axialScale="400 pc"/>',
                  '<ExpDiskGeometry radialScale="4 kpc" axialScale="400 pc" radialTrunc="9 kpc" axialTrunc="1 kpc"/>')
    t = t.replace('<OligoStellarComp luminosities="2e9">\n            <geometry type="Geometry"><SersicGeometry index="4" radius="1 kpc"/></geometry>\n          </OligoStellarComp>', "")
    assert "SersicGeometry" not in t
    p = tmp_path / "free.ski"
    p.write_text(t)
    print("ski run dust-free", flush=True)
    s = ski.load(str(p), resources)
    s.run(SG_FP64)
    sp = s.spec
    L = sp.stellar[0][1][0]                        # W in the 0.55 micron bin
    for i, path in enumerate(s.write(str(tmp_path))):
        f = sp.frames[i]
        data, hdr = read_fits(path)
        assert hdr["BUNIT"] == "W/m3/sr" and hdr["CTYPE1"] == "m"
        xps, yps = f["fovX"] / f["pixelsX"], f["fovY"] / f["pixelsY"]
        omega = (2.0 * math.atan(xps / (2.0 * f["distance"]))) * (2.0 * math.atan(yps / (2.0 * f["distance"])))
        F = data.sum() * omega
        want = L / sp.dlambdas[0] / (4.0 * math.pi * f["distance"] ** 2)
        assert abs(F / want - 1.0) < 1e-5, (i, F, want)
