"""FitSKIRT on the GPU: a genetic algorithm over the parameters of a ski file, each generation's individuals
evaluated together in batched launches.

FitSKIRT (FitSKIRTcore) fits the free parameters of a SKIRT model to observed images. OligoFitScheme::runSelf
(OligoFitScheme.cpp:27-34) runs the genetic algorithm that Optimization sets up (Optimization.cpp:42-90);
every individual is one set of parameter values, written into the ski file's "[label:default]" placeholders
(AdjustableSkirtSimulation::adjustedSkiContent, AdjustableSkirtSimulation.cpp:144-190) and run as a full
simulation (AdjustableSkirtSimulation::performWith, :119-141); its frames are convolved with each reference
image's kernel and the stellar luminosity is fitted (ReferenceImage::chi2, ReferenceImage.cpp:95-114); the
individual's score is the chi2 summed over the reference images (ReferenceImages::chi2,
ReferenceImages.cpp:76-94).

Here the genetic algorithm is the library's (sg_ga_*: GAlib's steady-state algorithm, draw for draw), and a
generation's individuals are evaluated together: each takes a pooled member simulation whose grid, dust
system and instruments were built once. An individual's adjusted ski file is parsed on the host (no device
work), and the member takes the new values in place: the dust density is resampled on the device
(sg_grid_resample_density, Cartesian grids), the stellar components and instrument angles are replaced
(sg_stars_update, sg_instruments_update). Then all members run in one batched transport launch per
wavelength (sg_sim_run_batch) and one fit launch per reference image (sg_fit_frames) scores them all.
A member whose adjusted file changes what cannot be updated in place (grid kind or size, dust mix, instrument
pixels, packet count) is rebuilt in full, as the reference does for every individual.

Supported: one stellar component per fitted frame set (the reference's one-component golden-section
luminosity fit); two or more components, which the reference fits with LumSimplex or GALumfit on per-component
frames, raise FatalError.
"""
import ctypes
import math
import time

import numpy as np

from . import capi, output, ski
from . import skirt as S
from .capi import SG_MIXED, check, dptr, lib

FatalError = capi.SkirtError


# ------------------------------------------------------------------------------------------- parameters


class ParameterRange:
    """ParameterRange (FitSKIRTcore/ParameterRange.hpp:18-38): a label used in the ski file's placeholders,
    the physical quantity ('' for dimensionless, else a Units quantity such as 'length', 'distance', 'mass',
    'posangle') and the range in SI units (radians for angles)."""

    def __init__(self, label, quantityType="length", minimumValue=0.0, maximumValue=1.0):
        if minimumValue > maximumValue:
            raise FatalError("The minimum value of parameter range %s exceeds its maximum" % label)
        self.label = label
        self.quantityType = "" if quantityType in ("", "dimless", None) else quantityType
        self.minimumValue = float(minimumValue)
        self.maximumValue = float(maximumValue)

    def quantityString(self):
        return self.quantityType


def convert_double_to_string(value, quantity, units):
    """DoublePropertyHandler::convertDoubletoString (Discover/DoublePropertyHandler.cpp:203-233): the SI value
    in the units system's unit for the quantity, 10 significant digits, exponent without a plus sign or
    leading zeros, four or more trailing zeros written as an exponent, then the unit."""
    unitspec = ""
    if quantity:
        value = units.out(quantity, value)
        unitspec = " " + units.unit(quantity)
    if math.isnan(value):
        number = "∅"
    elif math.isinf(value):
        number = "-∞" if value < 0 else "∞"
    else:
        number = "%.10g" % value
    number = number.replace("e-0", "e-").replace("e+0", "e").replace("e+", "e")
    zeroes = len(number) - len(number.rstrip("0"))
    if zeroes > 3:
        number = number[:len(number) - zeroes] + "e" + str(zeroes)
    return number + unitspec


def adjusted_ski_content(content, replacements, units):
    """AdjustableSkirtSimulation::adjustedSkiContent (AdjustableSkirtSimulation.cpp:144-190): every
    "[label:default]" becomes the replacement value for label (replacements maps label -> (SI value,
    quantity)) or else its default text."""
    out = []
    index = 0
    while True:
        left = content.find("[", index)
        if left < 0:
            break
        right = content.find("]", left + 1)
        if right < 0:
            raise FatalError("Square brackets not balanced in ski file")
        out.append(content[index:left])
        index = right + 1
        segment = content[left + 1:right]
        if "[" in segment:
            raise FatalError("Square brackets not balanced in ski file")
        colon = segment.find(":")
        if colon < 0:
            raise FatalError("Square brackets don't enclose colon in ski file")
        label = segment[:colon]
        if label in replacements:
            value, qty = replacements[label]
            out.append(convert_double_to_string(value, qty, units))
        else:
            out.append(segment[colon + 1:])
    out.append(content[index:])
    text = "".join(out)
    if "]" in text:
        raise FatalError("Square brackets not balanced in ski file")
    return text


# ------------------------------------------------------------------------------------------- references


class ReferenceImage:
    """ReferenceImage (FitSKIRTcore/ReferenceImage.hpp:22-34): an observed frame (Ny x Nx, in the ski's
    surface-brightness output unit), its convolution kernel and the luminosity bounds of each stellar
    component (only one component is fitted here)."""

    def __init__(self, data, kernel, minLuminosities, maxLuminosities):
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        if self.data.ndim != 2:
            raise FatalError("a reference image is a 2-D frame")
        self.kernel = kernel
        self.minLuminosities = [float(x) for x in np.atleast_1d(minLuminosities)]
        self.maxLuminosities = [float(x) for x in np.atleast_1d(maxLuminosities)]


# ------------------------------------------------------------------------------------------- optimization


class Optimization:
    """Optimization (FitSKIRTcore/Optimization.cpp): the genetic algorithm's settings (defaults popsize 100,
    generations 100, pmut 0.03, pcross 0.65, Optimization.hpp:29-51) and the algorithm itself (the library's
    GAlib-equivalent steady-state GA)."""

    def __init__(self, popsize=100, generations=100, pmut=0.03, pcross=0.65):
        self.popsize, self.generations, self.pmut, self.pcross = int(popsize), int(generations), pmut, pcross
        self._h = None

    def setup(self, ranges, seed):
        lo = np.array([r.minimumValue for r in ranges], dtype=np.float64)
        hi = np.array([r.maximumValue for r in ranges], dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().sg_ga_create(len(ranges), dptr(lo), dptr(hi), self.popsize, self.generations, self.pmut,
                                 self.pcross, seed, ctypes.byref(h)))
        self._h = h
        self.ngenes = len(ranges)

    def initialize(self):
        check(lib().sg_ga_initialize(self._h))

    def step(self):
        check(lib().sg_ga_step(self._h))

    def pending(self):
        n = ctypes.c_int()
        check(lib().sg_ga_pending(self._h, ctypes.byref(n), None))
        genes = np.zeros((n.value, self.ngenes))
        check(lib().sg_ga_pending(self._h, ctypes.byref(n), dptr(genes)))
        return genes

    def set_scores(self, scores):
        s = np.ascontiguousarray(scores, dtype=np.float64)
        check(lib().sg_ga_set_scores(self._h, len(s), dptr(s)))

    def generation(self):
        g, d = ctypes.c_int(), ctypes.c_int()
        check(lib().sg_ga_state(self._h, ctypes.byref(g), ctypes.byref(d)))
        return g.value

    def done(self):
        g, d = ctypes.c_int(), ctypes.c_int()
        check(lib().sg_ga_state(self._h, ctypes.byref(g), ctypes.byref(d)))
        return bool(d.value)

    def best(self):
        genes, score = np.zeros(self.ngenes), ctypes.c_double()
        check(lib().sg_ga_best(self._h, dptr(genes), ctypes.byref(score)))
        return genes, score.value

    def population(self):
        genes, scores = np.zeros((self.popsize, self.ngenes)), np.zeros(self.popsize)
        check(lib().sg_ga_population(self._h, dptr(genes), dptr(scores)))
        return genes, scores

    def __del__(self):
        if getattr(self, "_h", None) is not None and capi._lib is not None:
            capi._lib.sg_ga_destroy(self._h)
            self._h = None


# ------------------------------------------------------------------------------------------- simulation


def _structure(spec):
    """What a pooled member cannot take in place: everything except stellar parameters, instrument angles
    and distances, and a Cartesian grid's dust density."""
    d = spec.dust
    grid = None
    if d is not None:
        g = d["grid"]
        grid = (g["kind"], tuple(np.asarray(g["extent"]).tolist()), tuple(sorted((k, str(v)) for k, v in g.items()
                                                                               if k not in ("kind", "extent"))))
        if g["kind"] != "CartesianDustGrid" or len(d["components"]) > 1:
            # the density decides a tree's or Voronoi grid's structure; several components keep their own
            # densities on the host: either way a new density means a new member
            grid = grid + tuple((repr(c["geometry"].params().tolist()), c["scale"], tuple(c["opacities"]),
                                 tuple(c["albedos"]), tuple(c["asymmetry"])) for c in d["components"])
        grid = grid + (tuple(d["opacities"]), tuple(d["albedos"]), tuple(d["asymmetry"]), d["writeMeanIntensity"],
                       d["sampleCount"], type(d["geometry"]).__name__)
    frames = tuple((f["name"], f["pixelsX"], f["pixelsY"]) for f in spec.frames)
    comps = tuple(type(geom).__name__ for geom, _ in spec.stellar)
    return (grid, frames, comps, spec.nlambda, tuple(spec.lambdas), spec.packages, spec.minWeightReduction,
            spec.minScattEvents, spec.scattBias, spec.seed)


class _Member:
    """A pooled member: a ski.Simulation built once, then updated in place."""

    def __init__(self, spec):
        self.sim = ski.Simulation(spec, prefix="member")
        self.key = _structure(spec)

    def take(self, spec):
        if _structure(spec) != self.key:
            self.__init__(spec)
            return False
        m = self.sim
        m.spec = spec
        comps = [S.OligoStellarComp(geom, lum) for geom, lum in spec.stellar]
        m.stars.update(comps, spec.emissionBias)
        m.instr.update([S.FrameInstrument(f["distance"] / output.PC, f["inclination"], f["azimuth"],
                                          f["positionAngle"], f["pixelsX"], f["pixelsY"], f["fovX"] / output.PC,
                                          f["fovY"] / output.PC, f["centerX"] / output.PC, f["centerY"] / output.PC)
                        for f in spec.frames])
        if spec.dust is not None:
            d = spec.dust
            S.resample_density(m.grid, d["geometry"], d["scale"], d["sampleCount"], spec.seed)
        return True


class AdjustableSkirtSimulation:
    """AdjustableSkirtSimulation (FitSKIRTcore/AdjustableSkirtSimulation.cpp): the ski file with
    "[label:default]" placeholders, the instrument whose frames are fitted, and the pool of member
    simulations that evaluates a generation at once."""

    def __init__(self, skiContent, instrname=None, resources=None, precision=SG_MIXED):
        self.skiContent = skiContent
        self.resources = resources
        self.precision = precision
        base = ski.Spec(adjusted_ski_content(skiContent, {}, _units_of(skiContent)), resources)
        self.units = base.units
        names = [f["name"] for f in base.frames]
        self.instrname = instrname if instrname is not None else names[0]
        if self.instrname not in names:
            raise FatalError("Instrument %s not found in the ski file" % self.instrname)
        self.instrument = names.index(self.instrname)
        if len(base.stellar) != 1:
            raise FatalError("fits of %d stellar components (LumSimplex, GALumfit) are not supported; one "
                             "component is fitted by golden section" % len(base.stellar))
        self.pool = []
        self.rebuilt = 0
        # wall seconds of the evaluations: host preparation (parsing, in-place updates, rebuilds) and the
        # batched transport launches; packets run
        self.timing = dict(prepare=0.0, transport=0.0, packets=0)

    def ncomponents(self):
        return 1

    def nframes(self):
        return self.pool[0].sim.spec.nlambda if self.pool else None

    def specs(self, replacement_dicts):
        return [ski.Spec(adjusted_ski_content(self.skiContent, r, self.units), self.resources)
                for r in replacement_dicts]

    def evaluate(self, specs):
        """Runs every individual's simulation: members updated in place, one batched launch per wavelength.
        Returns the members holding the individuals' frames, in order."""
        t0 = time.perf_counter()
        for i, sp in enumerate(specs):
            if len(sp.stellar) != 1:
                raise FatalError("an individual's ski file has %d stellar components; one is supported"
                                 % len(sp.stellar))
            if i < len(self.pool):
                if not self.pool[i].take(sp):
                    self.rebuilt += 1
            else:
                self.pool.append(_Member(sp))
        members = self.pool[:len(specs)]
        sims = [m.sim for m in members]
        for m in sims:
            m.instr.clear()
        t1 = time.perf_counter()
        nl = specs[0].nlambda
        for ell in range(nl):
            counts = [int(m.sim.setChunkParams(1)) for m in sims]
            S.run_batch([m.sim for m in sims], counts, ells=[ell] * len(sims), precision=self.precision)
            self.timing["packets"] += sum(counts)
        self.timing["prepare"] += t1 - t0
        self.timing["transport"] += time.perf_counter() - t1
        return members


def _units_of(content):
    """The units system a ski file declares (its placeholders' replacements are written in it)."""
    import xml.etree.ElementTree as ET
    root = ET.fromstring(_strip_placeholders(content))
    u = root.find(".//units/*")
    cls = u.tag if u is not None else "ExtragalacticUnits"
    return output.Units(cls, u.get("fluxOutputStyle", "Neutral") if u is not None else "Neutral")


def _strip_placeholders(content):
    out, index = [], 0
    while True:
        left = content.find("[", index)
        if left < 0:
            break
        right = content.find("]", left + 1)
        if right < 0:
            break
        out.append(content[index:left])
        seg = content[left + 1:right]
        out.append(seg[seg.find(":") + 1:])
        index = right + 1
    out.append(content[index:])
    return "".join(out)


# ------------------------------------------------------------------------------------------- fit scheme


class OligoFitScheme:
    """OligoFitScheme (FitSKIRTcore/OligoFitScheme.cpp): the simulation, the parameter ranges, the reference
    images (image k is fitted to the instrument's frame at wavelength k) and the optimization; run() is
    runSelf. Every evaluated individual is recorded as Optimization::writeLine does (generation, parameter
    values in the ski's units, chi2, fitted luminosities, per-image chi2), and each new best one as
    writeBest does."""

    def __init__(self, simulation, parameterRanges, referenceImages, optim, fixedSeed=False):
        self.simulation = simulation
        self.ranges = list(parameterRanges)
        self.images = list(referenceImages)
        self.optim = optim
        self.fixedSeed = fixedSeed
        self.records = []        # (generation, values in output units, chi2, lums, chis)
        self.best_records = []   # (consec, record)
        self._bestChi2 = 1e20

    def _replacements(self, genes):
        return {r.label: (float(v), r.quantityString()) for r, v in zip(self.ranges, genes)}

    def objective(self, genes):
        """OligoFitScheme::objective (OligoFitScheme.cpp:116-152) for a batch of individuals: chi2 summed over
        the reference images, the fitted luminosities and the per-image chi2."""
        if len(genes) == 0:
            return np.zeros(0), np.zeros((0, len(self.images))), np.zeros((0, len(self.images)))
        sim = self.simulation
        t0 = time.perf_counter()
        specs = sim.specs([self._replacements(g) for g in genes])
        sim.timing["prepare"] += time.perf_counter() - t0
        members = sim.evaluate(specs)
        t0 = time.perf_counter()
        nimg = len(self.images)
        lums, chis = np.zeros((len(genes), nimg)), np.zeros((len(genes), nimg))
        for k, ref in enumerate(self.images):
            if k >= specs[0].nlambda:
                raise FatalError("more reference images than wavelengths in the simulation")
            f = specs[0].frames[sim.instrument]
            if ref.data.shape != (f["pixelsY"], f["pixelsX"]):
                raise FatalError("Simulations and Reference Images have different dimensions")
            if len(ref.minLuminosities) != 1 or len(ref.maxLuminosities) != 1:
                raise FatalError("Number of luminosity boundaries differs from 1!")
            # the fit runs on raw detected luminosities; the calibration to the image's unit is one factor
            # per frame, so the luminosity bounds scale by it (the chi2 is unchanged)
            cal = np.array([_calibration(sp, sim.instrument, k) for sp in specs])
            for c in np.unique(cal):
                sel = np.flatnonzero(cal == c)
                lo, hi = ref.minLuminosities[0] * c, ref.maxLuminosities[0] * c
                lum, chi2 = S.fit_frames([members[i].sim.instr for i in sel], ref.data, ref.kernel, lo, hi,
                                         instrument=sim.instrument, ell=k)
                lums[sel, k], chis[sel, k] = lum / c, chi2
        sim.timing["fit"] = sim.timing.get("fit", 0.0) + time.perf_counter() - t0
        return chis.sum(axis=1), lums, chis

    def _evaluate_pending(self):
        genes = self.optim.pending()
        chi2, lums, chis = self.objective(genes)
        gen = self.optim.generation()
        for i, g in enumerate(genes):
            vals = [self.simulation.units.out(r.quantityString(), v) if r.quantityString() else v
                    for r, v in zip(self.ranges, g)]
            rec = (gen, vals, float(chi2[i]), lums[i].tolist(), chis[i].tolist())
            self.records.append(rec)
            if chi2[i] < self._bestChi2:
                self._bestChi2 = chi2[i]
                self.best_records.append((len(self.best_records), rec))
        self.optim.set_scores(chi2)

    def run(self):
        """OligoFitScheme::runSelf: initialize (fixed seed 4357, or one from the clock), then step until
        done, each population evaluation as one batch."""
        self.optim.setup(self.ranges, 4357 if self.fixedSeed else 0)
        self.optim.initialize()
        self._evaluate_pending()
        while not self.optim.done():
            self.optim.step()
            self._evaluate_pending()
        genes, score = self.optim.best()
        return dict(zip([r.label for r in self.ranges], genes.tolist())), score

    def write(self, prefix):
        """<prefix>_allsimulations.dat and <prefix>_BESTsimulations.dat (Optimization.cpp:81-86,276-300)."""
        with open(prefix + "_allsimulations.dat", "w") as f:
            for gen, vals, c2, lm, ch in self.records:
                f.write("%d %s\n" % (gen, " ".join("%.10g" % x for x in list(vals) + [c2] + lm + ch)))
        with open(prefix + "_BESTsimulations.dat", "w") as f:
            for consec, (gen, vals, c2, lm, ch) in self.best_records:
                f.write("%d %s\n" % (consec, " ".join("%.10g" % x for x in list(vals) + [c2] + lm + ch)))


def _calibration(spec, instrument, ell):
    """The factor SingleFrameInstrument::calibrateAndWriteDataCubes applies to instrument's frame ell."""
    f = spec.frames[instrument]
    xps, yps = f["fovX"] / f["pixelsX"], f["fovY"] / f["pixelsY"]
    one = np.ones((spec.nlambda, 1, 1))
    return float(output.calibrate_frame(one, spec.lambdas, spec.dlambdas, f["distance"], xps, yps, spec.units)[ell, 0, 0])
