"""Ski-file front end: runs an unchanged SKIRT7 parameter file (.ski) of the supported classes on the GPU
and writes the frame instruments' FITS files as SKIRT7 does.

A ski file is an XML hierarchy (doc/Part 1 - User Guide/SKIRT/ski files.txt): an element with a
capitalised name is an object of that class, its attributes are the object's scalar properties, and a
child element with a lower-case name is a compound property holding the object(s) nested in it. Scalar
values follow Discover's property handlers: a number with an optional unit, a unitless value of a
physical quantity taken in the units system's default unit (DoublePropertyHandler.cpp:169-199),
comma-separated lists (DoubleListPropertyHandler.cpp:96-108), booleans true/t/yes/y/1
(BoolPropertyHandler.cpp:78-83). Property names and defaults are those of the classes' Q_CLASSINFO
declarations.

Supported classes (anything else raises FatalError naming the class):
  OligoMonteCarloSimulation and PanMonteCarloSimulation (+ MonteCarloSimulation, Simulation properties),
  LogWavelengthGrid, PanStellarComp with BlackBodySED or SunSED and BolLuminosityStellarCompNormalization,
  PanDustSystem (one DustComp of Benchmark1DDustMix, GreyBodyDustEmissivity, AllCellsDustLib), SIUnits,
  ExtragalacticUnits, Random, OligoWavelengthGrid, StellarSystem, OligoStellarComp, ExpDiskGeometry,
  SersicGeometry, SpiralStructureGeometryDecorator, OligoDustSystem, CompDustDistribution (DustComps),
  SPHDustDistribution (particle file read from the ski file's directory; tree dust grids only),
  DustComp, SimpleOligoDustMix, FaceOnDustCompNormalization, DustMassDustCompNormalization,
  CartesianDustGrid with LinMesh, OctTreeDustGrid, BinTreeDustGrid (the three stopping criteria),
  VoronoiDustGrid (sites from the dust density or uniform), InstrumentSystem, FrameInstrument.

The transport works in parsecs (lengths) and watts (luminosities); outputs are calibrated by
output.calibrate_frame and written by output.write_frame as <prefix>_<instrument>_total.fits.
"""
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import output
from . import skirt as S
from .capi import SG_BOOKKEEPING, SG_MIXED, SG_NEIGHBOR, SG_TOPDOWN
from . import panchromatic as P
from .output import LSUN, MSUN, PC, FatalError, Units

# TreeDustGrid::SearchMethod names (TreeDustGrid.hpp:47-50)
_SEARCH = {"TopDown": SG_TOPDOWN, "Neighbor": SG_NEIGHBOR, "Bookkeeping": SG_BOOKKEEPING}


class Item:
    """One object of the hierarchy: class name, scalar attributes, compound properties."""

    def __init__(self, cls, attrs, props):
        self.cls, self.attrs, self.props = cls, attrs, props

    def one(self, name, default_cls=None):
        items = self.props.get(name)
        if not items:
            if default_cls is None:
                return None
            return Item(default_cls, {}, {})
        if len(items) != 1:
            raise FatalError("property %s of %s holds %d objects, expected one" % (name, self.cls, len(items)))
        return items[0]

    def many(self, name):
        return list(self.props.get(name, []))


def _parse(elem):
    props = {}
    for child in elem:
        if not isinstance(child.tag, str):
            continue
        if child.tag[:1].islower():
            props[child.tag] = [_parse(c) for c in child if isinstance(c.tag, str)]
        else:
            raise FatalError("element %s inside %s is not a property" % (child.tag, elem.tag))
    return Item(elem.tag, dict(elem.attrib), props)


def parse(text):
    """The simulation object of a ski document (text)."""
    root = ET.fromstring(text)
    if root.tag != "skirt-simulation-hierarchy":
        raise FatalError("not a ski file: root element is " + root.tag)
    sims = [c for c in root if isinstance(c.tag, str)]
    if len(sims) != 1:
        raise FatalError("a ski file holds exactly one simulation")
    return _parse(sims[0])


class Values:
    """Scalar property conversion in one units system."""

    def __init__(self, units):
        self.units = units

    def double(self, item, name, qty=None, default=None):
        s = item.attrs.get(name)
        if s is None:
            if default is None:
                raise FatalError("%s requires property %s" % (item.cls, name))
            s = default
        return self.convert(s, qty, item.cls, name)

    def convert(self, s, qty, cls="", name=""):
        seg = s.split()
        if not seg or len(seg) > 2:
            raise FatalError("invalid value '%s' for %s.%s" % (s, cls, name))
        try:
            v = float(seg[0])
        except ValueError:
            raise FatalError("invalid number '%s' for %s.%s" % (s, cls, name))
        if qty is None:
            if len(seg) != 1:
                raise FatalError("dimensionless %s.%s has a unit: '%s'" % (cls, name, s))
            return v
        return self.units.in_(qty, v, seg[1] if len(seg) == 2 else None)

    def doubles(self, item, name, qty=None, default=None):
        s = item.attrs.get(name, default)
        if s is None:
            raise FatalError("%s requires property %s" % (item.cls, name))
        return [self.convert(x, qty, item.cls, name) for x in s.split(",")]

    @staticmethod
    def integer(item, name, default=None):
        s = item.attrs.get(name, default)
        if s is None:
            raise FatalError("%s requires property %s" % (item.cls, name))
        return int(float(s))

    @staticmethod
    def boolean(item, name, default="false"):
        return item.attrs.get(name, default).strip().lower() in ("true", "t", "yes", "y", "1")

    @staticmethod
    def string(item, name, default=None):
        s = item.attrs.get(name, default)
        if s is None:
            raise FatalError("%s requires property %s" % (item.cls, name))
        return s


def _require(item, *classes):
    if item.cls not in classes:
        raise FatalError("class %s is not supported here (supported: %s)" % (item.cls, ", ".join(classes)))


def read_sun_sed(path):
    """The solar SED resource (SunSED.dat: a comment line, the count, then lambda in micron and L_lambda in
    W/micron per line), as lambda in m and L_lambda in W/m (OligoStellarComp.cpp:39-58)."""
    with open(path) as f:
        f.readline()
        tok = f.read().split()
    n = int(tok[0])
    vals = np.array([float(t) for t in tok[1:1 + 2 * n]]).reshape(n, 2)
    return vals[:, 0] / 1e6, vals[:, 1] * 1e6


def sun_luminosity(lam, sed):
    """Solar L_lambda (W/m) at lam (m): NR::locate_fail + linear interpolation (OligoStellarComp.cpp:63-72)."""
    lv, Lv = sed
    if lam < lv[0] or lam > lv[-1]:
        raise FatalError("The sun does not emit at the wavelength of the simulation (lambda = %g m)" % lam)
    k = int(np.searchsorted(lv, lam, side="right")) - 1
    k = min(k, len(lv) - 2)
    return Lv[k] + (lam - lv[k]) / (lv[k + 1] - lv[k]) * (Lv[k + 1] - Lv[k])


def _geometry(item, val):
    """A geometry, lengths in pc."""
    if item is None:
        raise FatalError("a geometry is required")
    if item.cls == "ExpDiskGeometry":
        return S.ExpDiskGeometry(val.double(item, "radialScale", "length") / PC,
                                 val.double(item, "axialScale", "length") / PC,
                                 val.double(item, "radialTrunc", "length", "0") / PC,
                                 val.double(item, "axialTrunc", "length", "0") / PC,
                                 val.double(item, "innerRadius", "length", "0") / PC)
    if item.cls == "SersicGeometry":
        return S.SersicGeometry(val.double(item, "index", None, "1"), val.double(item, "radius", "length") / PC)
    if item.cls == "SpiralStructureGeometryDecorator":
        inner = _geometry(item.one("geometry"), val)
        return S.SpiralStructureGeometryDecorator(
            inner, arms=val.integer(item, "arms", "1"), pitch=val.double(item, "pitch", "posangle", "10 deg"),
            radius=val.double(item, "radius", "length") / PC, phase=val.double(item, "phase", "posangle", "0 deg"),
            perturbWeight=val.double(item, "perturbWeight"), index=val.integer(item, "index", "1"))
    _require(item, "ExpDiskGeometry", "SersicGeometry", "SpiralStructureGeometryDecorator")


def _box(item, val):
    return np.array([val.double(item, k, "length") / PC for k in ("minX", "minY", "minZ", "maxX", "maxY", "maxZ")])


class Spec:
    """A ski file's parameters, converted and checked on the host: everything a run needs, with
    lengths in pc (transport) and SI (calibration), luminosities in W per wavelength bin."""

    def __init__(self, text, resources=None, inputs=None):
        self.inputs = inputs or "."
        sim = parse(text)
        _require(sim, "OligoMonteCarloSimulation", "PanMonteCarloSimulation")
        self.pan = sim.cls == "PanMonteCarloSimulation"
        uitem = sim.one("units", "ExtragalacticUnits")
        _require(uitem, "ExtragalacticUnits", "SIUnits")
        self.units = Units(uitem.cls, uitem.attrs.get("fluxOutputStyle", "Neutral"))
        val = self.val = Values(self.units)

        # MonteCarloSimulation.hpp / Simulation.hpp properties
        self.packages = val.double(sim, "packages", None, "1e6")
        self.minWeightReduction = val.double(sim, "minWeightReduction", None, "1e4")
        self.minScattEvents = val.integer(sim, "minScattEvents", "0")
        self.scattBias = val.double(sim, "scattBias", None, "0.5")
        if val.boolean(sim, "continuousScattering", "no"):
            raise FatalError("continuous scattering is not supported")
        rnd = sim.one("random", "Random")
        _require(rnd, "Random")
        self.seed = val.integer(rnd, "seed", "4357")

        if self.pan:
            # LogWavelengthGrid (LogWavelengthGrid.cpp:20-31) with PanWavelengthGrid's bins
            wg = sim.one("wavelengthGrid", "LogWavelengthGrid")
            _require(wg, "LogWavelengthGrid")
            self.lambdas, self.dlambdas = P.log_wavelength_grid(val.double(wg, "minWavelength", "wavelength"),
                                                                val.double(wg, "maxWavelength", "wavelength"),
                                                                val.integer(wg, "points", "25"))
        else:
            # OligoWavelengthGrid: sorted wavelengths, bin widths 0.001 lambda (OligoWavelengthGrid.cpp:20-36)
            wg = sim.one("wavelengthGrid", "OligoWavelengthGrid")
            _require(wg, "OligoWavelengthGrid")
            self.lambdas = np.sort(np.array(val.doubles(wg, "wavelengths", "wavelength")))
            self.dlambdas = 0.001 * self.lambdas
        nl = self.nlambda = len(self.lambdas)

        # StellarSystem of OligoStellarComps (luminosities in solar units at each wavelength) or PanStellarComps
        # (an SED and a bolometric luminosity): W per bin
        ss = sim.one("stellarSystem", "StellarSystem")
        _require(ss, "StellarSystem")
        self.emissionBias = val.double(ss, "emissionBias", None, "0.5")
        sed = None
        self.stellar = []
        for c in ss.many("components"):
            if self.pan:
                self.stellar.append((_geometry(c.one("geometry"), val), self._pan_luminosities(c, resources)))
                continue
            _require(c, "OligoStellarComp")
            lum = val.doubles(c, "luminosities")
            if len(lum) != nl:
                raise FatalError("The number of stellar component luminosities differs from the number of wavelengths")
            if sed is None:
                sed = read_sun_sed(_resource(resources, os.path.join("SED", "Sun", "SunSED.dat")))
            Lw = [lum[ell] * sun_luminosity(self.lambdas[ell], sed) * self.dlambdas[ell] for ell in range(nl)]
            self.stellar.append((_geometry(c.one("geometry"), val), Lw))
        if not self.stellar:
            raise FatalError("the stellar system has no components")

        # frame instruments (SI values)
        isys = sim.one("instrumentSystem", "InstrumentSystem")
        _require(isys, "InstrumentSystem")
        self.frames = []
        for ins in isys.many("instruments"):
            _require(ins, "FrameInstrument")
            self.frames.append(dict(
                name=val.string(ins, "instrumentName"), distance=val.double(ins, "distance", "distance"),
                inclination=val.double(ins, "inclination", "posangle", "0"),
                azimuth=val.double(ins, "azimuth", "posangle", "0"),
                positionAngle=val.double(ins, "positionAngle", "posangle", "0"),
                pixelsX=val.integer(ins, "pixelsX", "250"), pixelsY=val.integer(ins, "pixelsY", "250"),
                fovX=val.double(ins, "fieldOfViewX", "length"), fovY=val.double(ins, "fieldOfViewY", "length"),
                centerX=val.double(ins, "centerX", "length", "0"), centerY=val.double(ins, "centerY", "length", "0")))

        # dust system (optional)
        ds = sim.one("dustSystem")
        self.dust = None if ds is None else self._dust(ds, nl)

    def _pan_luminosities(self, c, resources):
        """PanStellarComp::setupSelfAfter (PanStellarComp.cpp:30-50): the normalisation's total luminosity times
        the normalised SED (BlackBodySED or SunSED) at each wavelength of the grid; W per bin."""
        val = self.val
        _require(c, "PanStellarComp")
        sedi = c.one("sed")
        if sedi is None:
            raise FatalError("PanStellarComp requires an SED")
        if sedi.cls == "BlackBodySED":
            sed = P.blackbody_sed(self.lambdas, val.double(sedi, "temperature", "temperature"))
        elif sedi.cls == "SunSED":
            sed = P.sun_sed(self.lambdas, self.dlambdas,
                            read_sun_sed(_resource(resources, os.path.join("SED", "Sun", "SunSED.dat"))))
        else:
            _require(sedi, "BlackBodySED", "SunSED")
        norm = c.one("normalization")
        if norm is None:
            raise FatalError("PanStellarComp requires a normalization")
        _require(norm, "BolLuminosityStellarCompNormalization")
        # BolLuminosityStellarCompNormalization.cpp:24-28: solar units converted with Units::Lsun
        L = val.double(norm, "luminosity")
        if not L > 0:
            raise FatalError("The bolometric luminosity should be positive")
        return list(L * LSUN * sed)

    def _dust(self, ds, nl):
        val = self.val
        if self.pan:
            return self._pan_dust(ds, nl)
        _require(ds, "OligoDustSystem")
        dd = ds.one("dustDistribution", "CompDustDistribution")
        _require(dd, "CompDustDistribution", "SPHDustDistribution")
        if dd.cls == "SPHDustDistribution":
            return self._sph(ds, dd, nl)
        comps = dd.many("components")
        if not comps:
            raise FatalError("the dust distribution has no components")
        d = dict(components=[self._dustcomp(dc, nl) for dc in comps])
        # the first component's entries at the top level, for single-component callers
        d.update({k: v for k, v in d["components"][0].items()})
        d["sampleCount"] = val.integer(ds, "sampleCount", "100")
        d["writeMeanIntensity"] = val.boolean(ds, "writeMeanIntensity", "no")
        d["grid"] = self._grid(ds.one("dustGrid", "OctTreeDustGrid"))
        return d

    def _pan_dust(self, ds, nl):
        """PanDustSystem (PanDustSystem.cpp:30-70) with one DustComp of Benchmark1DDustMix dust, the grey-body
        emissivity and the all-cells library (or no dust emission), emission bias and boost, self-absorption."""
        val = self.val
        _require(ds, "PanDustSystem")
        # PanDustSystem::setupSelfBefore runs before its components are set up (PanDustSystem.cpp:61-64)
        if P.nearest_wavelength(self.lambdas, 0.55e-6) < 0:
            raise FatalError("Wavelength range should include 0.55 micron for a panchromatic simulation with dust")
        dd = ds.one("dustDistribution", "CompDustDistribution")
        _require(dd, "CompDustDistribution")
        comps = dd.many("components")
        if len(comps) != 1:
            raise FatalError("a panchromatic dust system of one dust component is supported")
        dc = comps[0]
        _require(dc, "DustComp")
        d = dict(geometry=_geometry(dc.one("geometry"), val))
        mix = dc.one("mix")
        if mix is None:
            raise FatalError("DustComp requires a mix")
        _require(mix, "Benchmark1DDustMix")
        d["panmix"] = P.Benchmark1DDustMix(self.lambdas, self.dlambdas)
        norm = dc.one("normalization")
        if norm is None:
            raise FatalError("DustComp requires a normalization")
        if norm.cls == "FaceOnDustCompNormalization":
            ell = P.nearest_wavelength(self.lambdas, val.double(norm, "wavelength", "wavelength"))
            if ell < 0:
                raise FatalError("Optical properties are not defined for this wavelength")
            d["scale"] = S.faceon_normalization(d["geometry"], val.double(norm, "opticalDepth"),
                                                d["panmix"].kappaext[ell])
        elif norm.cls == "DustMassDustCompNormalization":
            d["scale"] = val.double(norm, "dustMass", "mass") / (PC * PC)
        else:
            _require(norm, "FaceOnDustCompNormalization", "DustMassDustCompNormalization")
        emis = ds.one("dustEmissivity")
        d["dustEmission"] = emis is not None
        if emis is not None:
            _require(emis, "GreyBodyDustEmissivity")
            lib_ = ds.one("dustLib")
            if lib_ is None:
                raise FatalError("There should be a dust library when dust emission is turned on")
            _require(lib_, "AllCellsDustLib")
        d["emissionBias"] = val.double(ds, "emissionBias", None, "0.5")
        d["emissionBoost"] = val.double(ds, "emissionBoost", None, "1")
        d["selfAbsorption"] = d["dustEmission"] and val.boolean(ds, "selfAbsorption", "no")
        d["cycles"] = val.integer(ds, "cycles", "0")
        d["sampleCount"] = val.integer(ds, "sampleCount", "100")
        d["grid"] = self._grid(ds.one("dustGrid", "OctTreeDustGrid"))
        if d["grid"]["kind"] == "VoronoiDustGrid":
            raise FatalError("panchromatic dust emission is supported on Cartesian and tree dust grids")
        d["components"] = [dict(geometry=d["geometry"], scale=d["scale"])]
        return d

    def _sph(self, ds, dd, nl):
        """SPHDustDistribution (SPHDustDistribution.cpp:38-100): the particle file's rows x y z h M Z [T]
        (pc, pc, pc, pc, Msun, metallicity, K; TextInFile::readRow with the temperature optional), the dust
        fraction of the metals and the maximum temperature, and its SimpleOligoDustMix."""
        val = self.val
        fdust = val.double(dd, "dustFraction", None, "0.3")
        if not fdust > 0:
            raise FatalError("The dust fraction should be positive")
        tmax = val.double(dd, "maximumTemperature", "temperature", "75000 K")
        fn = val.string(dd, "filename")
        rows = read_particles(os.path.join(self.inputs, fn))
        mix = dd.one("dustMix", "InterstellarDustMix")
        if mix is None:
            raise FatalError("Dust mix was not set")
        _require(mix, "SimpleOligoDustMix")
        d = dict(sph=dict(rows=rows, dustFraction=fdust, maximumTemperature=tmax))
        d["opacities"] = val.doubles(mix, "opacities", "opacity")
        d["albedos"] = val.doubles(mix, "albedos")
        d["asymmetry"] = val.doubles(mix, "asymmetryParameters", None, ",".join(["0"] * nl))
        if not (len(d["opacities"]) == len(d["albedos"]) == len(d["asymmetry"]) == nl):
            raise FatalError("the dust mix needs one opacity, albedo and asymmetry parameter per wavelength")
        d["sampleCount"] = val.integer(ds, "sampleCount", "100")
        d["writeMeanIntensity"] = val.boolean(ds, "writeMeanIntensity", "no")
        d["grid"] = self._grid(ds.one("dustGrid", "OctTreeDustGrid"))
        if d["grid"]["kind"] not in ("OctTreeDustGrid", "BinTreeDustGrid"):
            raise FatalError("an SPH dust distribution is supported on tree dust grids (their cells' exact mass in "
                             "box, TreeDustGrid::density)")
        return d

    def _dustcomp(self, dc, nl):
        """One DustComp (DustComp.cpp): geometry, SimpleOligoDustMix and normalization."""
        val = self.val
        _require(dc, "DustComp")
        d = dict(geometry=_geometry(dc.one("geometry"), val))
        mix = dc.one("mix")
        if mix is None:
            raise FatalError("DustComp requires a mix")
        _require(mix, "SimpleOligoDustMix")
        d["opacities"] = val.doubles(mix, "opacities", "opacity")
        d["albedos"] = val.doubles(mix, "albedos")
        d["asymmetry"] = val.doubles(mix, "asymmetryParameters", None, ",".join(["0"] * nl))
        if not (len(d["opacities"]) == len(d["albedos"]) == len(d["asymmetry"]) == nl):
            raise FatalError("the dust mix needs one opacity, albedo and asymmetry parameter per wavelength")
        norm = dc.one("normalization")
        if norm is None:
            raise FatalError("DustComp requires a normalization")
        if norm.cls == "FaceOnDustCompNormalization":
            lam0 = val.double(norm, "wavelength", "wavelength")
            ells = [ell for ell in range(nl) if abs(lam0 / self.lambdas[ell] - 1.0) < 1e-5]
            if not ells:
                raise FatalError("Optical properties are not defined for this wavelength")
            # FaceOnDustCompNormalization::normalizationFactor (FaceOnDustCompNormalization.cpp:67-74), in pc
            d["scale"] = S.faceon_normalization(d["geometry"], val.double(norm, "opticalDepth"),
                                                d["opacities"][ells[0]])
        elif norm.cls == "DustMassDustCompNormalization":
            # DustMassDustCompNormalization.cpp:47-51: the dust mass (kg) times the geometry's unit-mass
            # density, which is per pc^3 here while the opacity stays in m2/kg
            d["scale"] = val.double(norm, "dustMass", "mass") / (PC * PC)
        else:
            _require(norm, "FaceOnDustCompNormalization", "DustMassDustCompNormalization")
        return d

    def _grid(self, g):
        val = self.val
        spec = dict(kind=g.cls, extent=_box(g, val))
        if g.cls == "CartesianDustGrid":
            bins = []
            for name in ("meshX", "meshY", "meshZ"):
                m = g.one(name, "LinMesh")
                _require(m, "LinMesh")
                bins.append(val.integer(m, "numBins", "100"))
            spec["bins"] = bins
        elif g.cls in ("OctTreeDustGrid", "BinTreeDustGrid"):
            if g.cls == "BinTreeDustGrid" and val.string(g, "directionMethod", "Alternating") != "Alternating":
                raise FatalError("only the Alternating direction method is supported")
            if g.cls == "OctTreeDustGrid" and val.boolean(g, "barycentric", "no"):
                raise FatalError("barycentric octrees are not supported")
            method = val.string(g, "searchMethod", "Neighbor")
            if method not in ("Neighbor", "TopDown", "Bookkeeping"):
                raise FatalError("search method %s is not supported" % method)
            # BinTreeDustGrid::setupSelfBefore (BinTreeDustGrid.cpp:23-24)
            if method == "Bookkeeping" and g.cls == "BinTreeDustGrid":
                raise FatalError("Bookkeeping method is not compatible with binary tree")
            spec.update(minLevel=val.integer(g, "minLevel", "2"), maxLevel=val.integer(g, "maxLevel", "6"),
                        maxMassFraction=val.double(g, "maxMassFraction", None, "1e-6"),
                        maxOpticalDepth=val.double(g, "maxOpticalDepth", None, "0"),
                        maxDensDispFraction=val.double(g, "maxDensDispFraction", None, "0"),
                        sampleCount=val.integer(g, "sampleCount", "100"), search=method)
        elif g.cls == "VoronoiDustGrid":
            spec.update(numParticles=val.integer(g, "numParticles", "500"),
                        distribution=val.string(g, "distribution", "DustDensity"))
            if spec["distribution"] not in ("DustDensity", "Uniform"):
                raise FatalError("Voronoi site distribution %s is not supported" % spec["distribution"])
        else:
            _require(g, "CartesianDustGrid", "OctTreeDustGrid", "BinTreeDustGrid", "VoronoiDustGrid")
        return spec


def read_particles(path):
    """SPH gas particles from a column text file as SPHDustDistribution reads it (TextInFile::readRow with
    seven columns, the last optional: x y z h M Z T): blank lines and lines starting with # are skipped, a
    missing temperature reads as 0, extra columns are ignored. Returns an (n, 7) array in the file's units."""
    rows = []
    try:
        f = open(path)
    except OSError:
        raise FatalError("Could not open the data file " + path)
    with f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            # stringstream reads: a seventh read after trailing blanks fails rather than reaching the end
            if len(t) < 6 or (len(t) == 6 and line.rstrip("\n").rstrip("\r") != line.rstrip()):
                if len(t) < 6:
                    raise FatalError("One or more required value(s) on text line are missing")
                raise FatalError("Input text is not formatted as a floating point number")
            try:
                v = [float(x) for x in t[:7]]
            except ValueError:
                raise FatalError("Input text is not formatted as a floating point number")
            rows.append(v + [0.0] * (7 - len(v)))
    return np.array(rows, dtype=np.float64).reshape(-1, 7)


def _resource(resources, name):
    base = resources or os.environ.get("SKIRT7_AMD_RESOURCES")
    if not base:
        raise FatalError("resource " + name + " needed: pass resources= (SKIRT7's dat directory) or set "
                         "SKIRT7_AMD_RESOURCES")
    path = os.path.join(base, name)
    if not os.path.exists(path):
        raise FatalError("Could not open the data file " + path)
    return path


class Simulation:
    """A ski file's objects on the GPU: grid, dust and stellar systems, instruments and the simulation."""

    def __init__(self, spec, prefix="ski"):
        self.spec, self.prefix = spec, prefix
        sp = spec
        comps = [S.OligoStellarComp(geom, lum) for geom, lum in sp.stellar]
        self.stars = S.StellarSystem(comps, emissionBias=sp.emissionBias)
        self.instr = S.InstrumentSystem(
            [S.FrameInstrument(f["distance"] / PC, f["inclination"], f["azimuth"], f["positionAngle"], f["pixelsX"],
                               f["pixelsY"], f["fovX"] / PC, f["fovY"] / PC, f["centerX"] / PC, f["centerY"] / PC)
             for f in sp.frames], sp.nlambda)
        self.grid = self.dust = self.rho = self.sites = self.sph = self.pan = None
        if sp.pan and sp.dust is not None:
            self._pan(sp)
            return
        if sp.dust is not None and "sph" in sp.dust:
            self._sph_dust(sp.dust, sp.seed)
        elif sp.dust is not None:
            d = sp.dust
            dcs = d["components"]
            geoms, scales = [c["geometry"] for c in dcs], [c["scale"] for c in dcs]
            self.grid = self._grid(d["grid"], geoms, scales, sp.seed)
            # DustSystem::setSampleDensityBody (DustSystem.cpp:145-172): every component's density sampled at
            # the same positions of each cell (its value at the site for a Voronoi cell)
            if d["grid"]["kind"] == "VoronoiDustGrid":
                rhos = [c["scale"] * S.geometry_density(c["geometry"], self.sites) for c in dcs]
            else:
                rhos = [S.sample_density(self.grid, c["geometry"], c["scale"], d["sampleCount"], sp.seed)
                        for c in dcs]
            mixes = [S.SimpleOligoDustMix(c["opacities"], c["albedos"], c["asymmetry"]) for c in dcs]
            self.mix = mixes[0] if len(mixes) == 1 else mixes
            self.rho = rhos[0] if len(rhos) == 1 else rhos
            self.dust = S.DustSystem(self.grid, self.rho, self.mix, writeMeanIntensity=d["writeMeanIntensity"])
        self.sim = S.OligoMonteCarloSimulation(self.dust, self.stars, self.instr, packages=sp.packages,
                                               minWeightReduction=sp.minWeightReduction,
                                               minScattEvents=sp.minScattEvents, scattBias=sp.scattBias,
                                               seed=sp.seed)

    def _pan(self, sp):
        """PanMonteCarloSimulation's objects: the grid and sampled densities as for an oligochromatic run, the
        PanDustSystem (lengths in pc) and the simulation."""
        d = sp.dust
        self.grid = self._grid(d["grid"], [d["geometry"]], [d["scale"]], sp.seed)
        self.rho = S.sample_density(self.grid, d["geometry"], d["scale"], d["sampleCount"], sp.seed)
        self.mix = d["panmix"]
        pdust = P.PanDustSystem(self.grid, self.rho, self.mix, dustEmission=d["dustEmission"],
                                selfAbsorption=d["selfAbsorption"], cycles=d["cycles"],
                                emissionBias=d["emissionBias"], emissionBoost=d["emissionBoost"], length_unit=PC)
        self.dust = pdust.ds
        self.pan = P.PanMonteCarloSimulation(self.stars, pdust, self.instr, packages=sp.packages,
                                             minWeightReduction=sp.minWeightReduction,
                                             minScattEvents=sp.minScattEvents, scattBias=sp.scattBias, seed=sp.seed)
        self.sim = self.pan.sim

    def _sph_dust(self, d, seed):
        """The SPH particles in the transport's units (pc; masses in kg over pc^2, so that kappa rho ds is an
        optical depth with rho per pc^3 and ds in pc, as for DustMassDustCompNormalization), the tree built over
        them on the device and its cells' mass-in-box densities (TreeDustGrid::density)."""
        p = d["sph"]
        rows = np.array(p["rows"], dtype=np.float64, copy=True)
        rows[:, 4] *= MSUN / (PC * PC)
        self.sph = S.SPHDustDistribution(rows, p["dustFraction"], p["maximumTemperature"])
        g = d["grid"]
        cls = S.OctTreeDustGrid if g["kind"] == "OctTreeDustGrid" else S.BinTreeDustGrid
        self.grid, masses = cls.subdivide_sph(g["extent"], self.sph, minlevel=g["minLevel"], maxlevel=g["maxLevel"],
                                              maxmassfraction=g["maxMassFraction"],
                                              maxopticaldepth=g["maxOpticalDepth"],
                                              maxdensdispfraction=g["maxDensDispFraction"], nsamples=g["sampleCount"],
                                              seed=seed, search=_SEARCH[g["search"]], masses=True)
        # with samples (the dispersion criterion) the masses at maxlevel were not evaluated: evaluate the cells
        self.rho = S.sph_density(self.grid, self.sph, None if g["maxDensDispFraction"] > 0 else masses)
        self.mix = S.SimpleOligoDustMix(d["opacities"], d["albedos"], d["asymmetry"])
        self.dust = S.DustSystem(self.grid, self.rho, self.mix, writeMeanIntensity=d["writeMeanIntensity"])

    def _grid(self, g, geoms, scales, seed):
        """The grid of the dust system whose components have geometries geoms and normalisations scales (the
        tree is subdivided on their summed density, Voronoi sites drawn from the components by mass)."""
        ext = g["extent"]
        if g["kind"] == "CartesianDustGrid":
            return S.CartesianDustGrid(ext, *[S.linmesh(ext[a], ext[a + 3], g["bins"][a]) for a in range(3)])
        if g["kind"] in ("OctTreeDustGrid", "BinTreeDustGrid"):
            cls = S.OctTreeDustGrid if g["kind"] == "OctTreeDustGrid" else S.BinTreeDustGrid
            return cls.subdivide(ext, geoms, scales, total_mass=float(sum(scales)), minlevel=g["minLevel"],
                                 maxlevel=g["maxLevel"], maxmassfraction=g["maxMassFraction"],
                                 nsamples=g["sampleCount"], seed=seed, search=_SEARCH[g["search"]],
                                 maxopticaldepth=g["maxOpticalDepth"], maxdensdispfraction=g["maxDensDispFraction"])
        n = g["numParticles"]
        if g["distribution"] == "DustDensity":
            from . import workloads
            # CompDustDistribution::generatePosition: a component chosen with probability proportional to its
            # mass, then a position from its geometry
            if len(geoms) == 1:
                self.sites = workloads.sample_sites(geoms[0], n, extent=ext, seed=seed)
            else:
                counts = np.random.default_rng(seed).multinomial(n, np.asarray(scales) / float(sum(scales)))
                parts = [workloads.sample_sites(gm, int(k), extent=ext, seed=seed + h)
                         for h, (gm, k) in enumerate(zip(geoms, counts)) if k > 0]
                self.sites = np.concatenate(parts)[np.random.default_rng(seed + 1000).permutation(n)]
        else:
            rng = np.random.default_rng(seed)
            self.sites = ext[:3] + rng.random((n, 3)) * (ext[3:] - ext[:3])
        # VoronoiMesh::buildMesh orders the cells by x and drops particles too close to another
        self.sites = S.voronoi_cells(ext, self.sites)
        return S.VoronoiDustGrid(ext, self.sites)

    def run(self, precision=SG_MIXED):
        """The stellar emission phase: every wavelength, Npp packets each; for a panchromatic simulation with
        dust, PanMonteCarloSimulation::runSelf (stellar emission, self-absorption cycles, dust emission)."""
        self.instr.clear()
        if self.pan is not None:
            self.pan.precision = precision
            return self.pan.runSelf()
        return self.sim.runstellaremission(precision)

    def write(self, outdir=".", stamp=None):
        """FrameInstrument::write for every instrument: calibrated <prefix>_<name>_total.fits files."""
        paths = []
        sp = self.spec
        for i, f in enumerate(sp.frames):
            xps, yps = f["fovX"] / f["pixelsX"], f["fovY"] / f["pixelsY"]
            cube = output.calibrate_frame(self.instr.frame(i), sp.lambdas, sp.dlambdas, f["distance"], xps, yps,
                                          sp.units)
            path = os.path.join(outdir, "%s_%s_total.fits" % (self.prefix, f["name"]))
            output.write_frame(path, cube, xps, yps, f["centerX"], f["centerY"], sp.units, stamp)
            paths.append(path)
        return paths


def load(path, resources=None):
    """A ski file from disk; the output prefix is its file name without the extension."""
    with open(path) as f:
        text = f.read()
    return Simulation(Spec(text, resources, os.path.dirname(os.path.abspath(path))),
                      os.path.splitext(os.path.basename(path))[0])


def run(path, outdir=None, resources=None, precision=SG_MIXED):
    """Runs a ski file and writes its frames next to it (or into outdir); returns the FITS paths."""
    s = load(path, resources)
    s.run(precision)
    return s.write(outdir or os.path.dirname(os.path.abspath(path)))
