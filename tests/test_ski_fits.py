"""Ski-file front end, units, frame calibration and FITS output (CPU).

The FITS writer is pinned against the reference's own vendored cfitsio (Cfitsio/, compiled from the
reference tree into oracle/_ref/libfitsref.so behind a shim that makes FITSInOut::write's calls): the
files must be byte-identical for the same data, keywords and DATE value, and cfitsio must read ours back.
The calibration is pinned by flux conservation (summing the surface brightness over the pixels' solid
angles recovers L / (4 pi d^2) per unit wavelength) and by the relations between the flux output styles.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from skirt7_amd import output as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FITSREF = os.path.join(ROOT, "oracle", "_ref", "libfitsref.so")
_D = ctypes.POINTER(ctypes.c_double)


def _fitsref():
    if not os.path.exists(FITSREF):
        pytest.skip("oracle/_ref/libfitsref.so not built (needs the reference tree)")
    L = ctypes.CDLL(FITSREF)
    L.ref_fits_write.argtypes = [ctypes.c_char_p, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p,
                                 ctypes.c_char_p]
    L.ref_fits_read.argtypes = [ctypes.c_char_p, _D, ctypes.c_long, ctypes.POINTER(ctypes.c_int)]
    L.ref_fits_key.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _D, ctypes.c_char_p, ctypes.c_int]
    return L


STAMP = "2026-12-02T21:07:00"


@pytest.mark.parametrize("shape,incx,incy,xc,yc,du,xu", [
    ((1, 5, 7), 0.123456789, 98765.4321, -1.5e-3, 2.5e20, "W/m2/arcsec2", "pc"),
    ((3, 4, 6), 1.0, 2.0, 0.0, 0.0, "MJy/sr", "m"),
    ((1, 128, 256), 156.25, 156.25, 0.0, -12.5, "W/m2/micron/arcsec2", "pc"),
    ((2, 1, 1), 1e-30, 7e100, 1.0 / 3.0, -0.0, "", "Mpc"),
])
def test_fits_bytes_match_reference_cfitsio(tmp_path, shape, incx, incy, xc, yc, du, xu):
    L = _fitsref()
    rng = np.random.default_rng(sum(shape))
    d = rng.lognormal(0.0, 6.0, shape) * 1e-15
    d.flat[0] = 0.0
    if d.size > 2:
        d.flat[1] = -2.5
        d.flat[2] = 3.0e38
    ref, ours = str(tmp_path / "r.fits"), str(tmp_path / "o.fits")
    assert L.ref_fits_write(ref.encode(), np.ascontiguousarray(d).ctypes.data_as(_D), shape[2], shape[1], shape[0],
                            incx, incy, xc, yc, du.encode(), xu.encode(), STAMP.encode()) == 0
    O.fits_write(ours, d, incx, incy, xc, yc, du, xu, STAMP)
    a, b = open(ref, "rb").read(), open(ours, "rb").read()
    assert len(a) == len(b) and len(a) % 2880 == 0
    assert a == b


def test_fits_read_back_by_reference_cfitsio(tmp_path):
    L = _fitsref()
    d = np.arange(2 * 3 * 5, dtype=np.float64).reshape(2, 3, 5) * 0.1 + 1e-7
    p = str(tmp_path / "x.fits")
    O.fits_write(p, d, 2.5, 3.5, 1.0, -1.0, "W/m2/arcsec2", "pc")      # current UTC date
    out = np.zeros(64)
    dims = (ctypes.c_int * 3)()
    assert L.ref_fits_read(p.encode(), out.ctypes.data_as(_D), 64, dims) == 0
    assert tuple(dims) == (5, 3, 2)
    assert np.array_equal(out[:30], d.ravel().astype(np.float32).astype(np.float64))
    v = ctypes.c_double()
    for key, want in (("CRPIX1", 3.0), ("CRPIX2", 2.0), ("CDELT1", 2.5), ("CDELT2", 3.5), ("CRVAL2", -1.0)):
        assert L.ref_fits_key(p.encode(), key.encode(), ctypes.byref(v), None, 0) == 0
        assert v.value == want, key
    buf = ctypes.create_string_buffer(80)
    assert L.ref_fits_key(p.encode(), b"BUNIT", None, buf, 80) == 0
    assert buf.value == b"W/m2/arcsec2"
    assert L.ref_fits_key(p.encode(), b"DATE", None, buf, 80) == 0
    assert len(buf.value) == 19 and buf.value[4:5] == b"-" and buf.value[10:11] == b"T"


def test_units_systems_and_flux_styles():
    u = O.Units("ExtragalacticUnits")
    assert u.unit("length") == "pc" and u.unit("distance") == "Mpc" and u.unit("wavelength") == "micron"
    assert u.unit("surfacebrightness") == "W/m2/arcsec2"
    assert O.Units("ExtragalacticUnits", "Frequency").unit("surfacebrightness") == "MJy/sr"
    assert O.Units("SIUnits", "Wavelength").unit("surfacebrightness") == "W/m3/sr"
    assert abs(u.in_("distance", 10.0, "Mpc") / (1e7 * O.PC) - 1.0) < 1e-15
    assert u.out("length", O.PC * 250.0) == 250.0
    assert u.in_("posangle", 90.0, "deg") == math.pi / 2
    # an arcsec is 1/206264.806... rad, so one sr holds 4.25e10 arcsec^2
    assert abs(1.0 / O.ARCSEC ** 2 / 4.254517029e10 - 1.0) < 1e-9
    # styles: lambda f_lambda (neutral) = nu f_nu, and f_nu = lambda^2 f_lambda / c, in SI
    si_l = O.Units("SIUnits", "Wavelength")
    si_n = O.Units("SIUnits", "Neutral")
    si_f = O.Units("SIUnits", "Frequency")
    lam, flam = 0.55e-6, 3.7e-5
    assert si_n.osurfacebrightness(lam, flam) == lam * si_l.osurfacebrightness(lam, flam)
    fnu = si_f.osurfacebrightness(lam, flam)
    nu = O.C / lam
    assert abs(nu * fnu / si_n.osurfacebrightness(lam, flam) - 1.0) < 1e-14
    with pytest.raises(O.FatalError):
        u.in_("length", 1.0, "furlong")


def test_calibration_conserves_flux():
    """Summing the calibrated f_lambda (SI) over the pixels' solid angles gives F_lambda = L_lambda / (4 pi d^2):
    the detected luminosity per bin over the bin width, the pixels' angular size at the distance, the
    inverse-square dilution."""
    rng = np.random.default_rng(2)
    nl, ny, nx = 2, 16, 24
    frame = rng.random((nl, ny, nx)) * 1e30                     # W per pixel per wavelength bin
    lam = np.array([0.5e-6, 2.2e-6])
    dl = 0.001 * lam
    d = 10.0e6 * O.PC
    xps, yps = 40.0 * O.PC, 25.0 * O.PC
    f = O.calibrate_frame(frame, lam, dl, d, xps, yps, O.Units("SIUnits", "Wavelength"))
    omega = (2.0 * math.atan(xps / (2.0 * d))) * (2.0 * math.atan(yps / (2.0 * d)))
    for ell in range(nl):
        F = f[ell].sum() * omega
        assert abs(F / (frame[ell].sum() / dl[ell] / (4.0 * math.pi * d * d)) - 1.0) < 1e-12
    # small-angle check of the pixel solid angle: (xps/d)(yps/d) up to terms in (xps/d)^2 / 12
    assert abs(omega / (xps * yps / d ** 2) - 1.0) < 1e-10
    # the extragalactic neutral style is lambda f_lambda in W/m2/arcsec2
    g = O.calibrate_frame(frame, lam, dl, d, xps, yps, O.Units("ExtragalacticUnits"))
    assert np.allclose(g[1], f[1] * lam[1] * O.ARCSEC ** 2, rtol=1e-14, atol=0)


SUN = """# a two-segment stand-in for the solar SED (lambda in micron, L in W/micron)
3
0.4  1.0e26
0.6  2.0e26
1.0  1.0e26
"""


def ski_text(grid="cartesian", packages="1e4", units="ExtragalacticUnits", dust=True, style="Neutral"):
    grids = {
        "cartesian": """<CartesianDustGrid minX="-20 kpc" maxX="20 kpc" minY="-20 kpc" maxY="20 kpc" minZ="-20 kpc" maxZ="20 kpc">
                  <meshX type="MoveableMesh"><LinMesh numBins="10"/></meshX>
                  <meshY type="MoveableMesh"><LinMesh numBins="12"/></meshY>
                  <meshZ type="MoveableMesh"><LinMesh numBins="8"/></meshZ>
                </CartesianDustGrid>""",
        "octree": """<OctTreeDustGrid minX="-20 kpc" maxX="20 kpc" minY="-20 kpc" maxY="20 kpc" minZ="-20 kpc" maxZ="20 kpc"
                  minLevel="2" maxLevel="4" searchMethod="Neighbor" sampleCount="20" maxMassFraction="1e-3"/>""",
    }
    dust_xml = ""
    if dust:
        dust_xml = """
        <dustSystem type="DustSystem">
          <OligoDustSystem sampleCount="10" writeMeanIntensity="false">
            <dustDistribution type="DustDistribution">
              <CompDustDistribution>
                <components type="DustComp">
                  <DustComp>
                    <geometry type="Geometry"><ExpDiskGeometry radialScale="6 kpc" axialScale="250 pc"/></geometry>
                    <mix type="DustMix"><SimpleOligoDustMix opacities="1 m2/kg" albedos="0.6" asymmetryParameters="0.6"/></mix>
                    <normalization type="DustCompNormalization">
                      <FaceOnDustCompNormalization wavelength="0.55 micron" opticalDepth="1"/>
                    </normalization>
                  </DustComp>
                </components>
              </CompDustDistribution>
            </dustDistribution>
            <dustGrid type="DustGrid">%s</dustGrid>
          </OligoDustSystem>
        </dustSystem>""" % grids[grid]
    return """<?xml version="1.0" encoding="UTF-8"?>
<skirt-simulation-hierarchy type="MonteCarloSimulation" format="6.1">
  <OligoMonteCarloSimulation packages="%s" minWeightReduction="1e4" scattBias="0.5">
    <random type="Random"><Random seed="4357"/></random>
    <units type="Units"><%s fluxOutputStyle="%s"/></units>
    <instrumentSystem type="InstrumentSystem">
      <InstrumentSystem>
        <instruments type="Instrument">
          <FrameInstrument instrumentName="i0" distance="10 Mpc" inclination="0 deg" azimuth="0 deg" positionAngle="0 deg"
             fieldOfViewX="40 kpc" pixelsX="32" centerX="0 pc" fieldOfViewY="40 kpc" pixelsY="24" centerY="0 pc"/>
          <FrameInstrument instrumentName="i88" distance="10 Mpc" inclination="88 deg"
             fieldOfViewX="4e4 pc" pixelsX="32" fieldOfViewY="2e4 pc" pixelsY="16"/>
        </instruments>
      </InstrumentSystem>
    </instrumentSystem>
    <wavelengthGrid type="OligoWavelengthGrid"><OligoWavelengthGrid wavelengths="0.55 micron"/></wavelengthGrid>
    <stellarSystem type="StellarSystem">
      <StellarSystem emissionBias="0.5">
        <components type="StellarComp">
          <OligoStellarComp luminosities="8e9">
            <geometry type="Geometry"><ExpDiskGeometry radialScale="4 kpc" axialScale="400 pc"/></geometry>
          </OligoStellarComp>
          <OligoStellarComp luminosities="2e9">
            <geometry type="Geometry"><SersicGeometry index="4" radius="1 kpc"/></geometry>
          </OligoStellarComp>
        </components>
      </StellarSystem>
    </stellarSystem>%s
  </OligoMonteCarloSimulation>
</skirt-simulation-hierarchy>
""" % (packages, units, style, dust_xml)


@pytest.fixture
def resources(tmp_path):
    d = tmp_path / "dat"
    (d / "SED" / "Sun").mkdir(parents=True)
    (d / "SED" / "Sun" / "SunSED.dat").write_text(SUN)
    return str(d)


def test_ski_spec(resources):
    from skirt7_amd import ski
    sp = ski.Spec(ski_text(), resources)
    assert sp.packages == 1e4 and sp.seed == 4357 and sp.scattBias == 0.5 and sp.minScattEvents == 0
    assert sp.units.system == "ExtragalacticUnits"
    assert np.allclose(sp.lambdas, [0.55e-6]) and np.allclose(sp.dlambdas, [0.55e-9])
    # solar L_lambda at 0.55 micron: 1.0e26 + 0.75 (2.0e26 - 1.0e26) W/micron = 1.75e32 W/m
    geom0, L0 = sp.stellar[0]
    assert abs(L0[0] / (8e9 * 1.75e32 * 0.55e-9) - 1.0) < 1e-12
    assert abs(geom0.hR - 4000.0) < 1e-9 and abs(geom0.hz - 400.0) < 1e-9     # pc
    geom1, L1 = sp.stellar[1]
    assert abs(L1[0] / L0[0] - 0.25) < 1e-12
    f0, f1 = sp.frames
    assert f0["name"] == "i0" and abs(f0["distance"] / (1e7 * O.PC) - 1.0) < 1e-15
    assert f1["inclination"] == 88.0 * math.pi / 180.0 and f1["pixelsY"] == 16
    assert abs(f1["fovX"] - 4e4 * O.PC) < 1e-3 * O.PC and f1["centerX"] == 0.0
    d = sp.dust
    assert d["grid"]["kind"] == "CartesianDustGrid" and d["grid"]["bins"] == [10, 12, 8]
    assert np.allclose(d["grid"]["extent"], [-2e4, -2e4, -2e4, 2e4, 2e4, 2e4])
    # face-on optical depth 1 through the dust disc: scale * kappa * SigmaZ = tau (lengths in pc)
    assert abs(d["scale"] * d["opacities"][0] * d["geometry"].SigmaZ() - 1.0) < 1e-12
    sp2 = ski.Spec(ski_text(grid="octree", units="SIUnits", style="Frequency"), resources)
    g = sp2.dust["grid"]
    assert g["kind"] == "OctTreeDustGrid" and g["maxLevel"] == 4 and g["maxMassFraction"] == 1e-3
    assert sp2.units.unit("surfacebrightness") == "W/m2/Hz/sr"


def test_ski_tree_search_methods(resources):
    """searchMethod TopDown, Neighbor or Bookkeeping for an octree (TreeDustGrid.hpp:47-50); Bookkeeping on a
    binary tree is the reference's fatal error (BinTreeDustGrid.cpp:23-24)."""
    from skirt7_amd import ski
    for m in ("TopDown", "Neighbor", "Bookkeeping"):
        t = ski_text(grid="octree").replace('searchMethod="Neighbor"', 'searchMethod="%s"' % m)
        assert ski.Spec(t, resources).dust["grid"]["search"] == m
    t = ski_text(grid="octree").replace('searchMethod="Neighbor"', 'searchMethod="Bookkeeping"')
    t = t.replace("<OctTreeDustGrid ", "<BinTreeDustGrid ")
    with pytest.raises(O.FatalError, match="Bookkeeping method is not compatible with binary tree"):
        ski.Spec(t, resources)


def test_ski_rejects_what_it_cannot_run(resources):
    from skirt7_amd import ski
    t = ski_text().replace("<SersicGeometry", "<PlummerGeometry")
    with pytest.raises(O.FatalError, match="PlummerGeometry"):
        ski.Spec(t, resources)
    t = ski_text().replace('wavelength="0.55 micron" opticalDepth', 'wavelength="0.6 micron" opticalDepth')
    with pytest.raises(O.FatalError, match="wavelength"):
        ski.Spec(t, resources)
    with pytest.raises(O.FatalError, match="SunSED"):
        ski.Spec(ski_text(), os.path.join(resources, "missing"))
    t = ski_text().replace('luminosities="8e9"', 'luminosities="8e9, 1e9"')
    with pytest.raises(O.FatalError, match="luminosities"):
        ski.Spec(t, resources)


REF_SUN = "/root/reference/dat/SED/Sun/SunSED.dat"


def test_reference_sun_sed(tmp_path):
    """The reference's own solar SED resource (dat/SED/Sun/SunSED.dat): read as OligoStellarComp does
    (micron -> m, W/micron -> W/m), its integral is the solar luminosity of Units.cpp (3.839e26 W), the
    lookup equals the reference's NR::locate_fail + NR::interpolate_linlin (compiled from the SKIRT7 tree)
    on and between the grid points, and a ski file run against that resource directory gets
    L = luminosity x L_sun,lambda x dlambda per bin."""
    import ctypes
    from skirt7_amd import ski
    if not os.path.exists(REF_SUN) or not os.path.exists(O_REF_SHIM()):
        pytest.skip("the SKIRT7 tree (and its compiled numerics) is not present")
    lv, Lv = ski.read_sun_sed(REF_SUN)
    assert len(lv) == 1099 and np.all(np.diff(lv) > 0)
    assert abs(np.trapezoid(Lv, lv) / 3.839e26 - 1.0) < 0.01
    R = ctypes.CDLL(O_REF_SHIM())
    D = ctypes.POINTER(ctypes.c_double)
    R.ref_locate_fail.restype = ctypes.c_int
    R.ref_locate_fail.argtypes = [D, ctypes.c_int, ctypes.c_double]
    R.ref_interpolate_linlin.restype = ctypes.c_double
    R.ref_interpolate_linlin.argtypes = [ctypes.c_double] * 5
    rng = np.random.default_rng(4)
    lams = np.concatenate([lv[1:-1:7], rng.uniform(lv[0], lv[-1], 400), [lv[0], lv[-1]]])
    for lam in lams:
        k = R.ref_locate_fail(lv.ctypes.data_as(D), len(lv), lam)
        assert k >= 0
        expect = R.ref_interpolate_linlin(lam, lv[k], lv[k + 1], Lv[k], Lv[k + 1])
        assert ski.sun_luminosity(lam, (lv, Lv)) == expect
    assert R.ref_locate_fail(lv.ctypes.data_as(D), len(lv), lv[0] * 0.99) < 0
    with pytest.raises(O.FatalError, match="does not emit"):
        ski.sun_luminosity(lv[0] * 0.99, (lv, Lv))
    res = tmp_path / "dat"
    res.mkdir()
    os.symlink(os.path.dirname(os.path.dirname(REF_SUN)), str(res / "SED"))
    sp = ski.Spec(ski_text(), str(res))
    geom0, L0 = sp.stellar[0]
    lam = 0.55e-6
    k = R.ref_locate_fail(lv.ctypes.data_as(D), len(lv), lam)
    Lsun = R.ref_interpolate_linlin(lam, lv[k], lv[k + 1], Lv[k], Lv[k + 1])
    assert abs(L0[0] / (8e9 * Lsun * 0.55e-9) - 1.0) < 1e-14
    # the Sun emits about 5e26 W per micron at 0.55 micron (4 pi AU^2 x 1.87 W/m2/nm)
    assert 4.5e32 < Lsun < 6.0e32


def O_REF_SHIM():
    from tests import oracle_lib
    return oracle_lib.REFSHIM


def test_faceon_normalization_needs_an_axisymmetric_geometry(resources):
    """FaceOnDustCompNormalization.cpp:67-74: a geometry that is not an AxGeometry (a spiral decorator, a
    GenGeometry, or a Sersic sphere) raises "Geometry is not axisymmetric"."""
    from skirt7_amd import ski
    disc = '<ExpDiskGeometry radialScale="6 kpc" axialScale="250 pc"/>'
    spiral = ('<SpiralStructureGeometryDecorator arms="2" pitch="20 deg" radius="4 kpc" perturbWeight="0.3">'
              '<geometry type="Geometry">%s</geometry></SpiralStructureGeometryDecorator>' % disc)
    t = ski_text().replace(disc, spiral)
    assert spiral in t
    with pytest.raises(O.FatalError, match="Geometry is not axisymmetric"):
        ski.Spec(t, resources)
    t = ski_text().replace(disc, '<SersicGeometry index="2" radius="3 kpc"/>')
    with pytest.raises(O.FatalError, match="Geometry is not axisymmetric"):
        ski.Spec(t, resources)
    # the same spiral dust normalised by its dust mass is accepted
    t = ski_text().replace(disc, spiral).replace('<FaceOnDustCompNormalization wavelength="0.55 micron" opticalDepth="1"/>',
                                                 '<DustMassDustCompNormalization dustMass="1e7 Msun"/>')
    assert ski.Spec(t, resources).dust["scale"] > 0


def two_dust_ski(grid="cartesian"):
    """ski_text with a second DustComp (a thin lane of its own mix and dust-mass normalisation)."""
    second = """
                  <DustComp>
                    <geometry type="Geometry"><ExpDiskGeometry radialScale="3 kpc" axialScale="120 pc"/></geometry>
                    <mix type="DustMix"><SimpleOligoDustMix opacities="2 m2/kg" albedos="0.3" asymmetryParameters="0.1"/></mix>
                    <normalization type="DustCompNormalization">
                      <DustMassDustCompNormalization dustMass="1e7 Msun"/>
                    </normalization>
                  </DustComp>
                </components>"""
    return ski_text(grid=grid).replace("</DustComp>\n                </components>", "</DustComp>" + second, 1)


def test_ski_two_dust_components(resources):
    from skirt7_amd import ski
    sp = ski.Spec(two_dust_ski(), resources)
    comps = sp.dust["components"]
    assert len(comps) == 2
    assert comps[0]["opacities"] == [1.0] and comps[1]["opacities"] == [2.0]
    assert comps[1]["albedos"] == [0.3] and comps[1]["asymmetry"] == [0.1]
    # dust mass normalisation: Msun in kg over pc^2 (the geometry's density is per pc^3)
    from skirt7_amd.output import MSUN, PC
    assert abs(comps[1]["scale"] - 1e7 * MSUN / PC ** 2) < 1e-9 * comps[1]["scale"]
    # the first component's entries stay at the top level
    assert sp.dust["scale"] == comps[0]["scale"] and sp.dust["opacities"] == [1.0]


def sph_ski(grid="bintree", criteria='maxMassFraction="2e-3"'):
    """ski_text with an SPHDustDistribution over particles.dat and a tree grid."""
    grids = {"bintree": '<BinTreeDustGrid minX="-10 kpc" maxX="10 kpc" minY="-10 kpc" maxY="10 kpc" minZ="-10 kpc" '
                        'maxZ="10 kpc" minLevel="3" maxLevel="24" searchMethod="TopDown" %s/>' % criteria,
             "octree": '<OctTreeDustGrid minX="-10 kpc" maxX="10 kpc" minY="-10 kpc" maxY="10 kpc" minZ="-10 kpc" '
                       'maxZ="10 kpc" minLevel="1" maxLevel="7" %s/>' % criteria}
    sph = """
            <dustDistribution type="DustDistribution">
              <SPHDustDistribution filename="particles.dat" dustFraction="0.4" maximumTemperature="2e4 K">
                <dustMix type="DustMix"><SimpleOligoDustMix opacities="1 m2/kg" albedos="0.6" asymmetryParameters="0.6"/></dustMix>
              </SPHDustDistribution>
            </dustDistribution>
            <dustGrid type="DustGrid">%s</dustGrid>""" % grids[grid]
    text = ski_text(grid="octree")
    start = text.index("<dustDistribution")
    end = text.index("</dustGrid>") + len("</dustGrid>")
    return text[:start] + sph.strip() + text[end:]


def write_particles(path, n=6000, seed=3):
    """A particle file (pc, pc, pc, pc, Msun, Z, K): clustered, with comments, a blank line and a few rows
    without temperature; a few particles hotter than the maximum temperature."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    c = rng.uniform(-6000, 6000, (8, 3))
    pos = np.clip(c[rng.integers(0, 8, n)] + rng.normal(0, 1500, (n, 3)), -9900, 9900)
    h = cKDTree(pos).query(pos, 33)[0][:, -1]
    M = rng.uniform(1e5, 3e5, n)
    Z = rng.uniform(0.01, 0.03, n)
    T = rng.uniform(5e3, 1.5e4, n)
    T[::97] = 5e4
    lines = ["# x y z h M Z T", ""]
    for i in range(n):
        if i % 50 == 7:
            lines.append("%.17g %.17g %.17g %.17g %.17g %.17g" % (pos[i, 0], pos[i, 1], pos[i, 2], h[i], M[i], Z[i]))
        else:
            lines.append("%.17g %.17g %.17g %.17g %.17g %.17g %.17g" % (pos[i, 0], pos[i, 1], pos[i, 2], h[i], M[i],
                                                                        Z[i], T[i]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    rows = np.column_stack([pos, h, M, Z, T])
    rows[np.arange(n) % 50 == 7, 6] = 0.0
    return rows


def test_ski_sph_distribution(tmp_path, resources):
    from skirt7_amd import ski
    rows = write_particles(tmp_path / "particles.dat")
    sp = ski.Spec(sph_ski(), resources, str(tmp_path))
    d = sp.dust
    assert d["sph"]["dustFraction"] == 0.4 and d["sph"]["maximumTemperature"] == 2e4
    assert np.array_equal(d["sph"]["rows"], rows)
    assert d["opacities"] == [1.0] and d["grid"]["kind"] == "BinTreeDustGrid"
    # defaults of SPHDustDistribution (SPHDustDistribution.hpp: dustFraction 0.3, maximumTemperature 75000 K)
    sp2 = ski.Spec(sph_ski().replace(' dustFraction="0.4" maximumTemperature="2e4 K"', ""), resources, str(tmp_path))
    assert sp2.dust["sph"]["dustFraction"] == 0.3 and sp2.dust["sph"]["maximumTemperature"] == 75000.0
    # a missing file, a missing column, a non-number, a Cartesian grid
    from skirt7_amd.output import FatalError
    with pytest.raises(FatalError, match="Could not open"):
        ski.Spec(sph_ski(), resources, str(tmp_path / "nowhere"))
    (tmp_path / "bad").mkdir()
    (tmp_path / "bad" / "particles.dat").write_text("1 2 3 4 5\n")
    with pytest.raises(FatalError, match="missing"):
        ski.Spec(sph_ski(), resources, str(tmp_path / "bad"))
    (tmp_path / "bad" / "particles.dat").write_text("1 2 3 4 5 x 7\n")
    with pytest.raises(FatalError, match="floating point"):
        ski.Spec(sph_ski(), resources, str(tmp_path / "bad"))
    cart = sph_ski().replace(sph_ski()[sph_ski().index("<BinTreeDustGrid"):sph_ski().index("/>", sph_ski().index(
        "<BinTreeDustGrid")) + 2], '<CartesianDustGrid minX="-1 kpc" maxX="1 kpc" minY="-1 kpc" maxY="1 kpc" '
        'minZ="-1 kpc" maxZ="1 kpc"><meshX type="MoveableMesh"><LinMesh numBins="2"/></meshX><meshY '
        'type="MoveableMesh"><LinMesh numBins="2"/></meshY><meshZ type="MoveableMesh"><LinMesh numBins="2"/></meshZ>'
        '</CartesianDustGrid>')
    with pytest.raises(FatalError, match="tree dust grids"):
        ski.Spec(cart, resources, str(tmp_path))


def pan_ski(emission=True, selfabs=False, sed='<BlackBodySED temperature="6000 K"/>', points=20, comps=1,
            minw="0.1 micron"):
    """A PanMonteCarloSimulation ski file: log wavelength grid, one PanStellarComp, a PanDustSystem of
    Benchmark1DDustMix dust on a Cartesian grid (grey-body emissivity and all-cells library when emission)."""
    emis = ('<dustEmissivity type="DustEmissivity"><GreyBodyDustEmissivity/></dustEmissivity>'
            '<dustLib type="DustLib"><AllCellsDustLib/></dustLib>') if emission else ""
    comp = """<DustComp>
                    <geometry type="Geometry"><ExpDiskGeometry radialScale="6 kpc" axialScale="300 pc"/></geometry>
                    <mix type="DustMix"><Benchmark1DDustMix/></mix>
                    <normalization type="DustCompNormalization">
                      <FaceOnDustCompNormalization wavelength="0.55 micron" opticalDepth="1"/>
                    </normalization>
                  </DustComp>"""
    return """<?xml version="1.0" encoding="UTF-8"?>
<skirt-simulation-hierarchy type="MonteCarloSimulation" format="6.1">
  <PanMonteCarloSimulation packages="2e4" minWeightReduction="1e4" scattBias="0.5">
    <random type="Random"><Random seed="4357"/></random>
    <units type="Units"><ExtragalacticUnits fluxOutputStyle="Neutral"/></units>
    <instrumentSystem type="InstrumentSystem">
      <InstrumentSystem>
        <instruments type="Instrument">
          <FrameInstrument instrumentName="faceon" distance="10 Mpc" inclination="0 deg" azimuth="0 deg"
             positionAngle="0 deg" fieldOfViewX="40 kpc" pixelsX="24" fieldOfViewY="40 kpc" pixelsY="24"/>
          <FrameInstrument instrumentName="edgeon" distance="10 Mpc" inclination="88 deg" azimuth="0 deg"
             positionAngle="0 deg" fieldOfViewX="40 kpc" pixelsX="24" fieldOfViewY="20 kpc" pixelsY="12"/>
        </instruments>
      </InstrumentSystem>
    </instrumentSystem>
    <wavelengthGrid type="PanWavelengthGrid"><LogWavelengthGrid minWavelength="%s" maxWavelength="1000 micron" points="%d"/></wavelengthGrid>
    <stellarSystem type="StellarSystem">
      <StellarSystem emissionBias="0.5">
        <components type="StellarComp">
          <PanStellarComp>
            <geometry type="Geometry"><ExpDiskGeometry radialScale="4 kpc" axialScale="400 pc"/></geometry>
            <sed type="StellarSED">%s</sed>
            <normalization type="StellarCompNormalization"><BolLuminosityStellarCompNormalization luminosity="1e10"/></normalization>
          </PanStellarComp>
        </components>
      </StellarSystem>
    </stellarSystem>
    <dustSystem type="PanDustSystem">
      <PanDustSystem sampleCount="20" emissionBias="0.5" emissionBoost="1" selfAbsorption="%s" cycles="2">
        <dustDistribution type="DustDistribution">
          <CompDustDistribution>
            <components type="DustComp">
              %s
            </components>
          </CompDustDistribution>
        </dustDistribution>
        <dustGrid type="DustGrid">
          <CartesianDustGrid minX="-20 kpc" maxX="20 kpc" minY="-20 kpc" maxY="20 kpc" minZ="-20 kpc" maxZ="20 kpc">
            <meshX type="MoveableMesh"><LinMesh numBins="16"/></meshX>
            <meshY type="MoveableMesh"><LinMesh numBins="16"/></meshY>
            <meshZ type="MoveableMesh"><LinMesh numBins="12"/></meshZ>
          </CartesianDustGrid>
        </dustGrid>
        %s
      </PanDustSystem>
    </dustSystem>
  </PanMonteCarloSimulation>
</skirt-simulation-hierarchy>
""" % (minw, points, sed, "true" if selfabs else "false", "\n".join([comp] * comps), emis)


def test_ski_panchromatic(resources):
    import math
    from skirt7_amd import panchromatic as P
    from skirt7_amd import ski
    from skirt7_amd.output import FatalError, LSUN
    sp = ski.Spec(pan_ski(), resources)
    assert sp.pan and sp.nlambda == 20
    lam, dl = P.log_wavelength_grid(1e-7, 1e-3, 20)
    assert np.allclose(sp.lambdas, lam, rtol=1e-14) and np.allclose(sp.dlambdas, dl, rtol=1e-14)
    geom, Lw = sp.stellar[0]
    Lw = np.array(Lw)
    assert abs(Lw.sum() / (1e10 * LSUN) - 1) < 1e-12
    # each bin holds the integral of the Planck function over the bin (quadrature), normalised
    from scipy.integrate import quad
    lo, hi = P.wavelength_bins(lam)
    bb = np.array([quad(lambda x: float(P.planck(np.array([x]), 6000.0)[0]), a, b, epsrel=1e-10)[0]
                   for a, b in zip(lo, hi)])
    bb /= bb.sum()
    main = bb > 1e-12
    # (the 101-point trapezoid in log lambda: 1.3e-4 where a bin spans the Wien tail)
    assert np.max(np.abs(Lw[main] / (1e10 * LSUN) / bb[main] - 1)) < 5e-4
    d = sp.dust
    assert d["dustEmission"] and not d["selfAbsorption"] and d["cycles"] == 2
    ell = P.nearest_wavelength(lam, 0.55e-6)
    assert abs(lam[ell] / 0.55e-6 - 1) < 0.5
    # face-on normalisation with the Benchmark1D mix's extinction at the nearest grid wavelength
    assert abs(d["scale"] - S_faceon(geom_d=d["geometry"], kappa=d["panmix"].kappaext[ell])) < 1e-12 * d["scale"]
    # a solar SED from the resource file: normalised, zero only beyond the table
    sun = ski.Spec(pan_ski(sed="<SunSED/>"), resources)
    Ls = np.array(sun.stellar[0][1])
    assert abs(Ls.sum() / (1e10 * LSUN) - 1) < 1e-12 and np.all(Ls >= 0)
    # the reference's errors
    with pytest.raises(FatalError, match="dust library"):
        ski.Spec(pan_ski().replace('<dustLib type="DustLib"><AllCellsDustLib/></dustLib>', ""), resources)
    with pytest.raises(FatalError, match="0.55 micron"):
        ski.Spec(pan_ski(minw="1 micron"), resources)
    with pytest.raises(FatalError, match="one dust component"):
        ski.Spec(pan_ski(comps=2), resources)
    assert not ski.Spec(pan_ski(emission=False), resources).dust["dustEmission"]


def S_faceon(geom_d, kappa):
    from skirt7_amd import skirt as S
    return S.faceon_normalization(geom_d, 1.0, kappa)
