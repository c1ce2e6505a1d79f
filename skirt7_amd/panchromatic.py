"""PanMonteCarloSimulation on the GPU: stellar emission over a panchromatic wavelength grid with the absorbed
luminosity kept per cell and wavelength, optional dust self-absorption cycles, and the dust emission phase
(SKIRTcore/PanMonteCarloSimulation.cpp:90-340).

The dust's emission spectra are computed on the device from the absorption table (sg_dust_emission_spectra:
mean intensity, each dust population's equilibrium temperature, grey-body emissivity; PanDustSystem,
GreyBodyDustEmissivity, AllCellsDustLib). Dust photon packets are launched from the cells on the device
(sg_sim_run_dust) and transported by the same kernels as stellar packets. Supported: one dust component,
Cartesian and tree grids, the grey-body emissivity with the all-cells library, panchromatic dust mixes given
as populations of absorption and scattering cross sections (Benchmark1DDustMix is provided).
"""
import ctypes
import math

import numpy as np

from . import capi
from . import skirt as S
from .capi import SG_MIXED, check, dptr, lib
from .output import PC

FatalError = capi.SkirtError
KAPPA_V = 2600.0   # Units::kappaV (Units.cpp:29), m2/kg


def log_wavelength_grid(lambdamin, lambdamax, n):
    """LogWavelengthGrid (LogWavelengthGrid.cpp:20-31: NR::loggrid over n points) with PanWavelengthGrid's bin
    widths (PanWavelengthGrid.cpp:24-33, WavelengthGrid.cpp:85-95): borders at the geometric means of
    neighbouring wavelengths, the outer bins ending at the first and last wavelength. Metres."""
    if lambdamin <= 0:
        raise FatalError("the shortest wavelength should be positive")
    if lambdamax <= lambdamin:
        raise FatalError("the longest wavelength should be larger than the shortest")
    if n < 3:
        raise FatalError("There must be at least three bins in a panchromatic wavelength grid")
    logmin = math.log10(lambdamin)
    dlog = math.log10(lambdamax / lambdamin) / (n - 1)
    lam = np.array([10.0 ** (logmin + i * dlog) for i in range(n)])
    lo = np.array([lam[0]] + [math.sqrt(lam[i - 1] * lam[i]) for i in range(1, n)])
    hi = np.array([math.sqrt(lam[i] * lam[i + 1]) for i in range(n - 1)] + [lam[-1]])
    return lam, hi - lo


def wavelength_bins(lam):
    """Each wavelength's bin borders in a PanWavelengthGrid (PanWavelengthGrid.cpp:24-33): the geometric means
    of neighbouring wavelengths, the outer bins ending at the first and last wavelength. (lo, hi)."""
    lam = np.asarray(lam, float)
    n = len(lam)
    lo = np.array([lam[0]] + [math.sqrt(lam[i - 1] * lam[i]) for i in range(1, n)])
    hi = np.array([math.sqrt(lam[i] * lam[i + 1]) for i in range(n - 1)] + [lam[-1]])
    return lo, hi


def nearest_wavelength(lam, x):
    """WavelengthGrid::nearest (WavelengthGrid.cpp:99-106): the index of the grid wavelength nearest to x on a
    logarithmic scale, -1 outside the grid."""
    lam = np.asarray(lam, float)
    if x < lam[0] or x > lam[-1]:
        return -1
    ell = len(lam) - 2 if x == lam[-1] else int(np.searchsorted(lam, x, side="right")) - 1
    return ell if x < math.sqrt(lam[ell] * lam[ell + 1]) else ell + 1


def normalized_sed(Lv):
    """SED::setluminosities (SED.cpp): the luminosities normalised to a unit sum."""
    Lv = np.asarray(Lv, float)
    total = 0.0
    for v in Lv:
        total += v
    if not total > 0:
        raise FatalError("The total luminosity in the SED is zero or negative (%g)" % total)
    return Lv / total


def blackbody_sed(lam, T):
    """BlackBodySED::setupSelfBefore (BlackBodySED.cpp:25-55): each bin's luminosity the trapezoidal average of
    B(lambda) lambda over 101 points log-spaced across the bin, times ln 10 dlog lambda; normalised."""
    if not T > 0:
        raise FatalError("the black body temperature T should be positive")
    lo, hi = wavelength_bins(lam)
    Lv = np.zeros(len(lam))
    N = 100
    for ell in range(len(lam)):
        a, b = math.log10(lo[ell]), math.log10(hi[ell])
        d = (b - a) / N
        acc = 0.0
        for i in range(N + 1):
            w = 0.5 if i in (0, N) else 1.0
            x = 10.0 ** (a + i * d)
            acc += w * float(planck(np.array([x]), T)[0]) * x
        Lv[ell] = acc * math.log(10.0) * d
    return normalized_sed(Lv)


def interpolate_loglog(x, x1, x2, f1, f2):
    """NR::interpolate_loglog (NR.hpp:321-346): linear in log x and, where both values are positive, in log f."""
    x, x1, x2 = math.log10(x), math.log10(x1), math.log10(x2)
    logf = f1 > 0 and f2 > 0
    if logf:
        f1, f2 = math.log10(f1), math.log10(f2)
    fx = f1 + ((x - x1) / (x2 - x1)) * (f2 - f1)
    return 10.0 ** fx if logf else fx


def sun_sed(lam, dlam, data):
    """SunSED::setupSelfBefore (SunSED.cpp:25-52) on the grid: the emissivities of SunSED.dat (lambda in m, j)
    resampled by NR::resample with log-log interpolation (zero beyond the table), times the bin widths;
    normalised (SED::setemissivities)."""
    lv, jv = np.asarray(data[0], float), np.asarray(data[1], float)
    out = np.zeros(len(lam))
    for l, x in enumerate(lam):
        if abs(1.0 - x / lv[0]) < 1e-5:
            out[l] = jv[0]
        elif abs(1.0 - x / lv[-1]) < 1e-5:
            out[l] = jv[-1]
        elif x < lv[0] or x > lv[-1]:
            out[l] = 0.0
        else:
            k = len(lv) - 2 if x == lv[-1] else int(np.searchsorted(lv, x, side="right")) - 1
            out[l] = interpolate_loglog(x, lv[k], lv[k + 1], jv[k], jv[k + 1])
    return normalized_sed(out * np.asarray(dlam, float))


def powgrid(xmin, xmax, n, ratio):
    """NR::powgrid (Fundamentals/NR.hpp:189-204): n bins whose widths grow by `ratio` from first to last."""
    if abs(ratio - 1.0) < 1e-3:
        return np.array([xmin + (xmax - xmin) * i / n for i in range(n + 1)])
    rng = xmax - xmin
    q = ratio ** (1.0 / (n - 1))
    qn = q ** n
    return np.array([xmin + (1.0 - q ** i) / (1.0 - qn) * rng for i in range(n + 1)])


def planck(lam, T):
    """PlanckFunction (PlanckFunction.cpp:21-30) with the constants of Units.cpp:17-19, W/m3/sr."""
    h, c, k = 6.62606957e-34, 2.99792458e8, 1.3806488e-23
    with np.errstate(over="ignore", divide="ignore"):
        x = h * c / (lam * k * T)
        return 2.0 * h * c * c / np.power(lam, 5) / (np.exp(x) - 1.0)


class PanDustMix:
    """A panchromatic DustMix (DustMix.cpp:44-91,236-262) from populations (mu, sigmaabs, sigmasca, g) on the
    simulation's wavelengths: total cross sections, albedo, asymmetry parameter, coefficients per unit mass
    and the Planck-integrated absorption cross sections on the 1000-bin temperature grid up to 5000 K."""

    NT = 1000

    def __init__(self, lambdas, dlambdas, populations):
        self.lambdas, self.dlambdas = np.asarray(lambdas, float), np.asarray(dlambdas, float)
        pops = [p for p in populations if p[0] > 0]
        if not pops:
            raise FatalError("Dust mixture must contain at least one dust population")
        self.muv = np.array([p[0] for p in pops])
        self.sigmaabsvv = np.array([np.asarray(p[1], float) for p in pops])
        self.sigmascavv = np.array([np.asarray(p[2], float) for p in pops])
        self.asymmparvv = np.array([np.asarray(p[3], float) for p in pops])
        nl = len(self.lambdas)
        sumabs, sumsca, sumgsca = np.zeros(nl), np.zeros(nl), np.zeros(nl)
        for c in range(len(pops)):
            sumabs += self.sigmaabsvv[c]
            sumsca += self.sigmascavv[c]
            sumgsca += self.asymmparvv[c] * self.sigmascavv[c]
        sumext = sumabs + sumsca
        self.sigmaabsv = sumabs
        self.albedo = np.where(sumext != 0, sumsca / np.where(sumext != 0, sumext, 1.0), 0.0)
        self.asymmpar = np.where(sumsca != 0, sumgsca / np.where(sumsca != 0, sumsca, 1.0), 0.0)
        mu = 0.0
        for m in self.muv:
            mu += m
        self.mu = mu
        self.kappaabs = sumabs / mu
        self.kappasca = sumsca / mu
        self.kappaext = sumext / mu
        # the temperature grid and the Planck-integrated absorption cross sections (left at 0 for T = 0)
        self.Tv = powgrid(0.0, 5000.0, self.NT, 500.0)
        self.planckabs = np.zeros((len(pops), self.NT + 1))
        for p in range(1, self.NT + 1):
            B = planck(self.lambdas, self.Tv[p])
            for c in range(len(pops)):
                acc = 0.0
                for ell in range(nl):
                    acc += self.sigmaabsvv[c][ell] * B[ell] * self.dlambdas[ell]
                self.planckabs[c, p] = acc

    def Nlambda(self):
        return len(self.lambdas)

    def Npop(self):
        return len(self.muv)


class Benchmark1DDustMix(PanDustMix):
    """Benchmark1DDustMix (Benchmark1DDustMix.cpp:20-57): absorption and scattering cross sections 1 up to one
    micron, then lambda^-1 and lambda^-4; no asymmetry; a dust mass of 2 / kappaV."""

    def __init__(self, lambdas, dlambdas):
        lam = np.asarray(lambdas, float)
        brk = 1e-6
        kabs = np.where(lam <= brk, 1.0, brk / lam)
        ksca = np.where(lam <= brk, 1.0, (brk / lam) ** 4)
        super().__init__(lambdas, dlambdas, [(2.0 / KAPPA_V, kabs, ksca, np.zeros(len(lam)))])


class PanDustSystem:
    """PanDustSystem (PanDustSystem.cpp) with one dust component on a Cartesian or tree grid: the density per cell
    (kappa*rho per grid length unit, as the transport takes it), the panchromatic mix, a grey-body emissivity
    and the all-cells library (the defaults), emission bias 0.5 and boost 1, optional self-absorption with a
    fixed number of cycles (0: until converged)."""

    def __init__(self, grid, rho, mix, dustEmission=True, selfAbsorption=False, cycles=0, emissionBias=0.5,
                 emissionBoost=1.0, length_unit=PC):
        self.grid, self.mix = grid, mix
        self.rho = capi.as_f64(rho)
        self.dustEmission, self.selfAbsorption, self.cycles = dustEmission, selfAbsorption, int(cycles)
        self.emissionBias, self.emissionBoost = float(emissionBias), float(emissionBoost)
        self.length_unit = float(length_unit)
        self.ds = S.DustSystem(grid, self.rho, mix, writeMeanIntensity=dustEmission)

    def emission_spectra(self, labs_stellar, labs_dust=None):
        """calculatedustemission and Labsbolv: (dustlum (ncells, nlambda), Labsbol (ncells))."""
        m = self.mix
        nc, nl = len(self.rho), m.Nlambda()
        out, bol = np.zeros((nc, nl)), np.zeros(nc)
        ls = capi.as_f64(labs_stellar)
        ld = capi.as_f64(labs_dust) if labs_dust is not None else None
        check(lib().sg_dust_emission_spectra(self.ds._h, dptr(self.rho), dptr(ls), dptr(ld) if ld is not None else None,
                                             m.Npop(), dptr(capi.as_f64(m.sigmaabsvv)), m.mu, dptr(capi.as_f64(m.kappaabs)),
                                             dptr(capi.as_f64(m.lambdas)), dptr(capi.as_f64(m.dlambdas)), m.NT,
                                             dptr(capi.as_f64(m.Tv)), dptr(capi.as_f64(m.planckabs)), self.length_unit,
                                             dptr(out), dptr(bol)))
        return out, bol


class PanMonteCarloSimulation:
    """PanMonteCarloSimulation (PanMonteCarloSimulation.cpp:90-340): stellar emission at every wavelength, then,
    with dust emission, the self-absorption cycles (when asked) and the dust emission phase."""

    STAGE_FACTOR = (1.0 / 10.0, 1.0 / 3.0, 1.0)
    STAGE_EPSMAX = (0.010, 0.007, 0.005)

    def __init__(self, stars, dust, instruments, packages=1e6, minWeightReduction=1e4, minScattEvents=0,
                 scattBias=0.5, seed=4357, precision=SG_MIXED):
        self.stars, self.dust, self.instr = stars, dust, instruments
        self.packages = packages
        self.precision = precision
        self.sim = S.OligoMonteCarloSimulation(dust.ds, stars, instruments, packages=packages,
                                               minWeightReduction=minWeightReduction, minScattEvents=minScattEvents,
                                               scattBias=scattBias, seed=seed)
        self.nlambda = dust.mix.Nlambda()
        self.Labs_stellar = None
        self.Labs_dust = None
        self.log = []

    def _npp(self, factor=1.0):
        return int(math.ceil(self.packages * factor))

    def runstellaremission(self):
        """MonteCarloSimulation::runstellaremission with the absorbed luminosity stored (PanDustSystem stores
        absorption whenever it emits)."""
        self.dust.ds.clear()
        npp = self._npp()
        for ell in range(self.nlambda):
            self.sim.run_packets(ell, 0, npp, npp, self.precision)
        self.Labs_stellar = self.dust.ds.Labs() if self.dust.dustEmission else None

    def _dust_phase(self, npp, bias, peeloffs, store, stream):
        dustlum, bol = self.dust.emission_spectra(self.Labs_stellar, self.Labs_dust)
        c = np.zeros(capi.SG_NCOUNTERS, np.uint64)
        for ell in range(self.nlambda):
            Lv = np.ascontiguousarray(bol * dustlum[:, ell])
            check(lib().sg_sim_run_dust(self.sim._h, ell, 0, npp, npp, dptr(Lv), bias, int(peeloffs), int(store),
                                        stream, self.precision, c.ctypes.data_as(capi._U64)))
        return bol

    def rundustselfabsorption(self):
        """The three stages of self-absorption cycles (PanMonteCarloSimulation.cpp:106-157): each cycle re-emits
        the absorbed luminosity with no peel-offs and stores what is absorbed, until the absorbed dust
        luminosity changes by less than the stage's limit (the last stage runs at least two cycles), or for the
        fixed number of cycles."""
        prev = 0.0
        for stage in range(3):
            fixed = self.dust.cycles > 0
            ncmax = self.dust.cycles if fixed else 100
            convergence, cycle = False, 1
            while cycle <= ncmax and (not convergence or fixed):
                npp = self._npp(self.STAGE_FACTOR[stage])
                labs_dust_prev = self.Labs_dust
                self.dust.ds.clear()
                self._dust_phase(npp, 0.0, False, True, 1 + 100 * stage + cycle)
                self.Labs_dust = self.dust.ds.Labs()
                tot = float(np.sum(self.Labs_dust))
                eps = abs((tot - prev) / tot) if tot != 0 else 0.0
                prev = tot
                self.log.append((stage, cycle, tot, eps))
                if (stage < 2 or cycle > 1) and eps < self.STAGE_EPSMAX[stage]:
                    convergence = True
                cycle += 1
                del labs_dust_prev

    def rundustemission(self):
        """PanMonteCarloSimulation::rundustemission (:208-244): packages x emissionBoost dust packets per
        wavelength, biased cell choice, peel-offs, no absorption stored."""
        npp = self._npp(self.dust.emissionBoost)
        self.Labsbol = self._dust_phase(npp, self.dust.emissionBias, True, False, 1000)

    def runSelf(self):
        self.instr.clear()
        self.runstellaremission()
        if self.dust.dustEmission:
            if self.dust.selfAbsorption:
                self.rundustselfabsorption()
            self.rundustemission()
