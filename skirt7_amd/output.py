"""Units systems, frame calibration and FITS output of frame instruments.

Units mirrors SKIRT7's Units, SIUnits and ExtragalacticUnits (SKIRTcore/Units.cpp:15-186,
SIUnits.cpp, ExtragalacticUnits.cpp): program values are SI; `in_`/`out` convert from and to a named
unit, `unit(qty)` names the units system's default unit of a quantity, and the surface brightness
follows the flux output style (Units.cpp:1022-1040). `calibrate_frame` is
SingleFrameInstrument::calibrateAndWriteDataCubes (SingleFrameInstrument.cpp:151-226) for one frame
instrument, and `write_frame` saves its cube with the keywords of Image::saveto / FITSInOut::write
(Image.cpp:131-146,284-303; FITSInOut.cpp:32-95) through the library's sg_fits_write.
"""
import math

import numpy as np

from . import capi
from .capi import SkirtError as FatalError, check, dptr, lib

# physical constants of the units system (Units.cpp:17-26)
C = 2.99792458e8
AU = 1.49597871e11
PC = 3.08567758e16
MSUN = 1.9891e30
LSUN = 3.839e26
K_BOLTZMANN = 1.3806488e-23
ARCSEC = math.pi / (180.0 * 3600.0)

# SI value of one unit, per "quantity unit" (Units.cpp:36-186)
_FACTOR = {}


def _put(qty, table):
    for u, f in table.items():
        _FACTOR[qty + " " + u] = f


for _q in ("length", "distance"):
    _put(_q, {"m": 1.0, "cm": 1e-2, "km": 1e3, "AU": AU, "pc": PC, "kpc": 1e3 * PC, "Mpc": 1e6 * PC})
for _q in ("wavelength", "grainsize"):
    _put(_q, {"m": 1.0, "cm": 1e-2, "mm": 1e-3, "micron": 1e-6, "nm": 1e-9, "A": 1e-10})
_put("section", {"m2": 1.0})
_put("volume", {"m3": 1.0, "AU3": AU ** 3, "pc3": PC ** 3})
_put("velocity", {"m/s": 1.0, "km/s": 1e3})
_put("mass", {"kg": 1.0, "g": 1e-3, "Msun": MSUN})
_put("bulkmass", {"kg": 1.0})
_put("bulkmassdensity", {"kg/m3": 1.0, "g/cm3": 1e3})
_put("masssurfacedensity", {"kg/m2": 1.0, "Msun/AU2": MSUN / AU ** 2, "Msun/pc2": MSUN / PC ** 2})
_put("massvolumedensity", {"kg/m3": 1.0, "g/cm3": 1e3, "Msun/AU3": MSUN / AU ** 3, "Msun/pc3": MSUN / PC ** 3})
_put("opacity", {"m2/kg": 1.0})
_put("energy", {"J": 1.0})
_put("bolluminosity", {"W": 1.0, "Lsun": LSUN})
_put("monluminosity", {"W/m": 1.0, "W/micron": 1e6, "Lsun/micron": LSUN * 1e6})
_put("neutralfluxdensity", {"W/m2": 1.0})
_put("neutralsurfacebrightness", {"W/m2/sr": 1.0, "W/m2/arcsec2": 1.0 / ARCSEC ** 2})
_put("wavelengthfluxdensity", {"W/m3": 1.0, "W/m2/micron": 1e6})
_put("wavelengthsurfacebrightness", {"W/m3/sr": 1.0, "W/m2/micron/sr": 1e6, "W/m2/micron/arcsec2": 1e6 / ARCSEC ** 2})
_put("frequencyfluxdensity", {"W/m2/Hz": 1.0, "Jy": 1e-26, "mJy": 1e-29, "MJy": 1e-20})
_put("frequencysurfacebrightness", {"W/m2/Hz/sr": 1.0, "W/m2/Hz/arcsec2": 1.0 / ARCSEC ** 2, "Jy/sr": 1e-26,
                                    "Jy/arcsec2": 1e-26 / ARCSEC ** 2, "MJy/sr": 1e-20,
                                    "MJy/arcsec2": 1e-20 / ARCSEC ** 2})
_put("temperature", {"K": 1.0})
_put("angle", {"rad": 1.0, "deg": math.pi / 180.0, "arcsec": ARCSEC})
_put("posangle", {"rad": 1.0, "deg": math.pi / 180.0})
_put("solidangle", {"sr": 1.0, "arcsec2": ARCSEC ** 2})
_put("pressure", {"Pa": 1.0, "K/m3": K_BOLTZMANN})

# default output unit per quantity of the two units systems (SIUnits.cpp, ExtragalacticUnits.cpp)
_SI = {"length": "m", "distance": "m", "wavelength": "m", "grainsize": "m", "section": "m2", "volume": "m3",
       "velocity": "m/s", "mass": "kg", "bulkmass": "kg", "bulkmassdensity": "kg/m3", "masssurfacedensity": "kg/m2",
       "massvolumedensity": "kg/m3", "opacity": "m2/kg", "energy": "J", "bolluminosity": "W", "monluminosity": "W/m",
       "neutralfluxdensity": "W/m2", "neutralsurfacebrightness": "W/m2/sr", "wavelengthfluxdensity": "W/m3",
       "wavelengthsurfacebrightness": "W/m3/sr", "frequencyfluxdensity": "W/m2/Hz",
       "frequencysurfacebrightness": "W/m2/Hz/sr", "temperature": "K", "angle": "rad", "posangle": "rad",
       "solidangle": "sr", "pressure": "Pa"}
_EXTRAGALACTIC = dict(_SI, length="pc", distance="Mpc", wavelength="micron", grainsize="micron", volume="pc3",
                      velocity="km/s", mass="Msun", masssurfacedensity="Msun/pc2", massvolumedensity="Msun/pc3",
                      bolluminosity="Lsun", monluminosity="Lsun/micron", wavelengthfluxdensity="W/m2/micron",
                      neutralsurfacebrightness="W/m2/arcsec2", wavelengthsurfacebrightness="W/m2/micron/arcsec2",
                      frequencyfluxdensity="Jy", frequencysurfacebrightness="MJy/sr", angle="arcsec", posangle="deg",
                      solidangle="arcsec2", pressure="K/m3")

FLUX_STYLES = ("Neutral", "Wavelength", "Frequency")


class Units:
    """A units system: 'SIUnits' or 'ExtragalacticUnits', with the flux output style of Units
    (Neutral: lambda F_lambda; Wavelength: F_lambda; Frequency: F_nu)."""

    def __init__(self, system="ExtragalacticUnits", fluxOutputStyle="Neutral"):
        if system == "SIUnits":
            self.defaults = _SI
        elif system == "ExtragalacticUnits":
            self.defaults = _EXTRAGALACTIC
        else:
            raise FatalError("Unknown units system " + system)
        if fluxOutputStyle not in FLUX_STYLES:
            raise FatalError("Unknown flux output style " + fluxOutputStyle)
        self.system, self.style = system, fluxOutputStyle

    @staticmethod
    def factor(qty, unit):
        f = _FACTOR.get(qty + " " + unit)
        if not f:
            raise FatalError("Unknow quantity " + qty + " and/or unit " + unit)
        return f

    def in_(self, qty, value, unit=None):
        """Value in `unit` (default: the system's unit for qty) to SI (Units::in)."""
        return value * self.factor(qty, unit or self.defaults[qty])

    def out(self, qty, value, unit=None):
        """SI value to `unit` (default: the system's unit for qty) (Units::out)."""
        return value / self.factor(qty, unit or self.defaults[qty])

    def unit(self, qty):
        """Units::unit: surface brightness and flux density follow the flux output style."""
        if qty == "surfacebrightness":
            qty = {"Neutral": "neutral", "Wavelength": "wavelength", "Frequency": "frequency"}[self.style] + qty
        elif qty == "fluxdensity":
            qty = {"Neutral": "neutral", "Wavelength": "wavelength", "Frequency": "frequency"}[self.style] + qty
        if qty not in self.defaults:
            raise FatalError("Unknow quantity " + qty)
        return self.defaults[qty]

    def osurfacebrightness(self, lam, flam):
        """Units::osurfacebrightness: f_lambda (W/m3/sr) at wavelength lam (m) to the output style's unit."""
        if self.style == "Wavelength":
            return self.out("wavelengthsurfacebrightness", flam)
        if self.style == "Frequency":
            return self.out("frequencysurfacebrightness", lam * lam * flam / C)
        return self.out("neutralsurfacebrightness", lam * flam)


def calibrate_frame(frame, lambdas, dlambdas, distance, xpsiz, ypsiz, units):
    """SingleFrameInstrument::calibrateAndWriteDataCubes (SingleFrameInstrument.cpp:155-209) for one frame:
    frame (nlambda, Ny, Nx) of detected luminosities in W per wavelength bin; lambdas, dlambdas in m;
    distance, xpsiz, ypsiz in m. Returns the surface brightness in the units system's output unit."""
    f = np.array(frame, dtype=np.float64)
    if f.ndim == 2:
        f = f[None]
    lam = np.asarray(lambdas, dtype=np.float64)
    dl = np.asarray(dlambdas, dtype=np.float64)
    f /= dl[:, None, None]                                     # W -> W/m
    xang = 2.0 * math.atan(xpsiz / (2.0 * distance))
    yang = 2.0 * math.atan(ypsiz / (2.0 * distance))
    f /= xang * yang                                           # -> W/m/sr
    f /= 4.0 * math.pi * distance * distance                   # -> W/m3/sr
    for ell in range(f.shape[0]):
        f[ell] = units.osurfacebrightness(lam[ell], f[ell])
    return f


def fits_write(path, data, incx, incy, xc, yc, dataunits, xyunits, stamp=None):
    """FITSInOut::write through the library: data (nz, ny, nx) or (ny, nx), x fastest in the file."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    check(lib().sg_fits_write(path.encode(), dptr(a), nx, ny, nz, float(incx), float(incy), float(xc), float(yc),
                              dataunits.encode(), xyunits.encode(), stamp.encode() if stamp else None))


def write_frame(path, cube, xpsiz, ypsiz, xpc, ypc, units, stamp=None):
    """Image(item, Nx, Ny, Nlambda, xpsiz, ypsiz, xpc, ypc, "surfacebrightness") + Image::saveto
    (Image.cpp:131-146,284-303): increments and centre in the system's length unit."""
    fits_write(path, cube, units.out("length", xpsiz), units.out("length", ypsiz), units.out("length", xpc),
               units.out("length", ypc), units.unit("surfacebrightness"), units.unit("length"), stamp)
