"""ctypes access to the CPU oracle (oracle/_build/liboracle.so) for tests, smoke() and bench.py's
CPU baseline. The oracle is test infrastructure; it is never part of the product path."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
# the same restatement built with -O3 -march=x86-64-v3 (the CPU baseline's optimised build)
ORACLE_V3 = os.path.join(ROOT, "oracle", "_build", "liboracle_v3.so")
REFSHIM = os.path.join(ROOT, "oracle", "_ref", "libnrref.so")

_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_V = ctypes.c_void_p

_SIGS = {
    "or_scene_create": (_V, []),
    "or_scene_destroy": (None, [_V]),
    "or_scene_set_cartesian": (ctypes.c_int, [_V, _D, _D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int]),
    "or_scene_set_tree": (ctypes.c_int, [_V, ctypes.c_int, _D, _U8, ctypes.c_int64, ctypes.c_int]),
    "or_scene_set_voronoi": (ctypes.c_int, [_V, _D, _D, ctypes.c_int64, _I64, _I32]),
    "or_scene_num_cells": (ctypes.c_int64, [_V]),
    "or_scene_set_dust": (ctypes.c_int, [_V, _D, ctypes.c_int64, ctypes.c_int, _D, _D, _D, ctypes.c_int]),
    "or_scene_set_dust_multi": (ctypes.c_int, [_V, ctypes.c_int, _D, ctypes.c_int64, ctypes.c_int, _D, _D, _D, _D,
                                               ctypes.c_int]),
    "or_scene_add_star": (ctypes.c_int, [_V, _D, _D, _D, _D, ctypes.c_int]),
    "or_scene_add_instrument": (ctypes.c_int, [_V, _D]),
    "or_scene_path": (ctypes.c_int, [_V, _D, ctypes.c_int, _I32, _D]),
    "or_scene_whichcell": (ctypes.c_int, [_V, _D]),
    "or_scatter_phase": (ctypes.c_double, [_V, ctypes.c_int, _D, _D, _D]),
    "or_dust_spectra": (ctypes.c_int, [_V, _D, _D, ctypes.c_int, _D, ctypes.c_double, _D, _D, _D, ctypes.c_int, _D,
                                       _D, ctypes.c_double, _D, _D]),
    "or_run_dust_packets": (ctypes.c_int, [_V, _D, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                           _D, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                           ctypes.c_int, _D, _D, _U64]),
    "or_random_mix": (ctypes.c_int, [_V, ctypes.c_int, _D, ctypes.c_double]),
    "or_scene_opticaldepth": (ctypes.c_double, [_V, _D, ctypes.c_int]),
    "or_scene_pathlength": (None, [_V, _D, ctypes.c_int, ctypes.c_int, _D, _D]),
    "or_run_reference": (ctypes.c_int, [_V, _D, ctypes.c_int, _D, _D, _U64]),
    "or_run_packets": (ctypes.c_int, [_V, _D, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.c_int, _D, _D, _U64]),
    "or_mt_first_uniforms": (ctypes.c_double, [ctypes.c_uint32, ctypes.c_int, _D]),
    "or_mt_first_words": (None, [ctypes.c_uint32, ctypes.c_int, _U32]),
    "or_philox4x32": (None, [_U32, _U32, _U32]),
    "or_lambertw1": (ctypes.c_double, [ctypes.c_double]),
    "or_sersic_table": (None, [ctypes.c_double, _D, _D]),
    "or_hg_sample": (None, [ctypes.c_double, _D, ctypes.c_double, ctypes.c_double, _D]),
    "or_hg_value": (ctypes.c_double, [ctypes.c_double, _D, _D]),
    "or_gaussian_kernel": (None, [ctypes.c_double, ctypes.c_int, _D]),
    "or_convolve_nested": (None, [_D, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _D]),
    "or_golden_section": (None, [_D, _D, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D]),
    "or_last_error": (ctypes.c_char_p, []),
    "or_nr_locate_clip": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_double]),
    "or_nr_locate_fail": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_double]),
    "or_nr_locate_tau": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_double]),
    "or_nr_interp_loglog": (ctypes.c_double, [ctypes.c_double] * 5),
    "or_star_positions": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, _D]),
    "or_launch_sample": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.c_uint64, _I32, _D]),
    "or_exponcutoff_sample": (None, [ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, _D]),
    "or_pixelondetector": (None, [_D, ctypes.c_int64, _D, _I32, _D]),
    "or_sph_last_error": (ctypes.c_char_p, []),
    "or_sph_create": (_V, [_D, ctypes.c_int64, ctypes.c_double, ctypes.c_double]),
    "or_sph_destroy": (None, [_V]),
    "or_sph_count": (ctypes.c_int64, [_V]),
    "or_sph_mass": (ctypes.c_double, [_V]),
    "or_sph_myerf": (ctypes.c_double, [ctypes.c_double]),
    "or_sph_grid": (None, [_V, _D]),
    "or_sph_cell_list": (ctypes.c_int64, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I32, ctypes.c_int64]),
    "or_sph_mass_in_box": (None, [_V, _D, ctypes.c_int64, _D]),
    "or_sph_density": (None, [_V, _D, ctypes.c_int64, _D]),
    "or_sph_subdivide": (ctypes.c_int, [_V, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, _U8, _D, ctypes.c_int64, _I64]),
}

_lib = None
_libs = {}


def lib(variant=None):
    """The oracle library: the -O2 build (default), or variant "v3" (-O3 -march=x86-64-v3)."""
    global _lib
    path = ORACLE_V3 if variant == "v3" else ORACLE
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError("oracle not built (make -C oracle)")
        L = ctypes.CDLL(path)
        for n, (r, a) in _SIGS.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _libs[path] = L
    if path == ORACLE:
        _lib = _libs[path]
    return _libs[path]


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dp(a):
    return a.ctypes.data_as(_D)


class Scene:
    """An oracle scene: grid, dust, stars and instruments (all copies of the caller's arrays)."""

    def __init__(self, variant=None):
        self.L = lib(variant)
        self.h = self.L.or_scene_create()
        self._keep = []
        self.ninstr = []
        self.nlambda = 1

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "L", None) is not None:
            self.L.or_scene_destroy(self.h)
            self.h = None

    def keep(self, *a):
        self._keep.extend(a)
        return a[0] if len(a) == 1 else a

    def cartesian(self, extent, xv, yv, zv):
        e, x, y, z = self.keep(f64(extent), f64(xv), f64(yv), f64(zv))
        assert self.L.or_scene_set_cartesian(self.h, dp(e), dp(x), len(x), dp(y), len(y), dp(z), len(z)) == 0
        return self

    def tree(self, arity, extent, flags, search):
        e = self.keep(f64(extent))
        fl = self.keep(np.ascontiguousarray(flags, dtype=np.uint8))
        rc = self.L.or_scene_set_tree(self.h, arity, dp(e), fl.ctypes.data_as(_U8), len(fl), search)
        assert rc == 0, self.L.or_last_error()
        return self

    def voronoi(self, extent, sites, off, ids):
        e = self.keep(f64(extent))
        s = self.keep(f64(sites).reshape(-1, 3))
        o = self.keep(np.ascontiguousarray(off, dtype=np.int64))
        i = self.keep(np.ascontiguousarray(ids, dtype=np.int32))
        assert self.L.or_scene_set_voronoi(self.h, dp(e), dp(s), s.shape[0], o.ctypes.data_as(_I64),
                                          i.ctypes.data_as(_I32)) == 0
        return self

    def dust(self, rho, kext, albedo, g, store_abs=False):
        rho, k, a, gg = self.keep(f64(rho), f64(kext), f64(albedo), f64(g))
        self.nlambda = len(k)
        self.ncells = len(rho)
        assert self.L.or_scene_set_dust(self.h, dp(rho), len(rho), len(k), dp(k), dp(a), dp(gg), int(store_abs)) == 0
        return self

    def dust_multi(self, rhos, mixes, store_abs=False):
        """Several dust components: one density array and one SimpleOligoDustMix-like object per component."""
        r = self.keep(f64(np.concatenate([f64(x) for x in rhos])))
        cat = lambda name: self.keep(f64(np.concatenate([getattr(m, name) for m in mixes])))
        k, ks, a, gg = cat("kappaext"), cat("kappasca"), cat("albedo"), cat("asymmpar")
        n = len(rhos[0])
        self.nlambda = len(mixes[0].kappaext)
        self.ncells = n
        assert self.L.or_scene_set_dust_multi(self.h, len(rhos), dp(r), n, len(mixes[0].kappaext), dp(k), dp(ks), dp(a),
                                              dp(gg), int(store_abs)) == 0
        return self

    def star(self, params, lum, ts=None, tM=None):
        p, l = self.keep(f64(params), f64(lum))
        if ts is not None:
            s, M = self.keep(f64(ts), f64(tM))
            assert self.L.or_scene_add_star(self.h, dp(p), dp(l), dp(s), dp(M), len(s)) == 0
        else:
            assert self.L.or_scene_add_star(self.h, dp(p), dp(l), None, None, 0) == 0
        return self

    def instrument(self, params):
        p = self.keep(f64(params))
        assert self.L.or_scene_add_instrument(self.h, dp(p)) == 0
        self.ninstr.append((int(p[4]), int(p[5])))
        return self

    def star_positions(self, comp, n, seed=4357, first=0):
        """GeometricStellarComp::launch positions of component comp on packet streams [first, first+n)."""
        out = np.zeros((n, 3))
        rc = self.L.or_star_positions(self.h, comp, seed, first, n, dp(out))
        assert rc == 0, self.L.or_last_error()
        return out

    def launch_sample(self, n, ell=0, bias=0.5, seed=4357, first=0):
        """StellarSystem::launch with L = 1: (component index, luminosity weight) per packet."""
        comp = np.zeros(n, np.int32)
        w = np.zeros(n)
        rc = self.L.or_launch_sample(self.h, ell, bias, seed, first, n, comp.ctypes.data_as(_I32), dp(w))
        assert rc == 0, self.L.or_last_error()
        return comp, w

    def path(self, r, k, maxseg=100000):
        rk = f64(np.concatenate([r, k]))
        m = np.zeros(maxseg, np.int32)
        ds = np.zeros(maxseg)
        n = self.L.or_scene_path(self.h, dp(rk), maxseg, m.ctypes.data_as(_I32), dp(ds))
        assert n >= 0
        return m[:n].copy(), ds[:n].copy()

    def whichcell(self, r):
        return self.L.or_scene_whichcell(self.h, dp(f64(r)))

    def opticaldepth(self, r, k, ell=0):
        return self.L.or_scene_opticaldepth(self.h, dp(f64(np.concatenate([r, k]))), ell)

    def pathlength(self, r, k, taus, ell=0):
        t = f64(taus)
        out = np.zeros(len(t))
        self.L.or_scene_pathlength(self.h, dp(f64(np.concatenate([r, k]))), ell, len(t), dp(t), dp(out))
        return out

    def _buffers(self):
        nf = sum(self.nlambda * nx * ny for nx, ny in self.ninstr)
        frames = np.zeros(max(nf, 1))
        labs = np.zeros(max(getattr(self, "ncells", 1) * self.nlambda, 1))
        return frames, labs

    def run_packets(self, w, ell, first, count, npp, threads=8):
        frames, labs = self._buffers()
        c = np.zeros(8, np.uint64)
        mc = f64(w["sim"].mc)
        rc = self.L.or_run_packets(self.h, dp(mc), ell, first, count, npp, threads, dp(frames), dp(labs),
                                  c.ctypes.data_as(_U64))
        assert rc == 0, self.L.or_last_error()
        self.labs = labs
        return frames, c

    def dust_spectra(self, mix, labs_s, labs_d=None, len2=None):
        """or_dust_spectra for a panchromatic mix (skirt7_amd.panchromatic.PanDustMix): (dustlum, labsbol)."""
        from skirt7_amd.output import PC
        nc, nl = self.ncells, self.nlambda
        out, bol = np.zeros((nc, nl)), np.zeros(nc)
        ls = f64(labs_s)
        ld = f64(labs_d) if labs_d is not None else None
        rc = self.L.or_dust_spectra(self.h, dp(ls), dp(ld) if ld is not None else None, mix.Npop(),
                                    dp(f64(mix.sigmaabsvv)), mix.mu, dp(f64(mix.kappaabs)), dp(f64(mix.lambdas)),
                                    dp(f64(mix.dlambdas)), mix.NT, dp(f64(mix.Tv)), dp(f64(mix.planckabs)),
                                    PC * PC if len2 is None else len2, dp(out), dp(bol))
        assert rc == 0, self.L.or_last_error()
        return out, bol

    def run_dust_packets(self, mc, ell, first, count, npp, Lv, xi, peeloffs, store_abs, stream, threads=16):
        frames, labs = self._buffers()
        c = np.zeros(8, np.uint64)
        rc = self.L.or_run_dust_packets(self.h, dp(f64(mc)), ell, first, count, npp, dp(f64(Lv)), xi, int(peeloffs),
                                        int(store_abs), stream, threads, dp(frames), dp(labs), c.ctypes.data_as(_U64))
        assert rc == 0, self.L.or_last_error()
        self.labs = labs
        return frames, c

    def run_reference(self, mc, threads=8):
        frames, labs = self._buffers()
        c = np.zeros(8, np.uint64)
        m = f64(mc)
        rc = self.L.or_run_reference(self.h, dp(m), threads, dp(frames), dp(labs), c.ctypes.data_as(_U64))
        assert rc == 0, self.L.or_last_error()
        self.labs = labs
        return frames, c


def gaussian_kernel(fwhm, dimension):
    out = np.zeros((dimension, dimension))
    lib().or_gaussian_kernel(fwhm, dimension, dp(out))
    return out


def convolve(image, kernel):
    """Convolution::nested_loop of image (Ny, Nx) with kernel (ky, kx)."""
    image, kernel = f64(image), f64(kernel)
    out = np.zeros_like(image)
    Ny, Nx = image.shape
    ky, kx = kernel.shape
    lib().or_convolve_nested(dp(image), Nx, Ny, dp(kernel), kx, ky, dp(out))
    return out


def golden_section(ref, frame, minlum, maxlum):
    """GoldenSection::optimize: (lum, chi2, frame as returned)."""
    ref = f64(ref)
    fr = f64(frame).copy()
    lum = np.zeros(1)
    chi2 = np.zeros(1)
    lib().or_golden_section(dp(ref), dp(fr), ref.size, minlum, maxlum, dp(lum), dp(chi2))
    return float(lum[0]), float(chi2[0]), fr


def scene_from_workload(w, variant=None):
    """An oracle scene with exactly the inputs of a skirt7_amd.workloads setup."""
    sc = Scene(variant)
    g = w["grid"]
    if w["kind"] == "cartesian":
        sc.cartesian(g.extent, g.xv, g.yv, g.zv)
    elif w["kind"] in ("octree", "bintree"):
        sc.tree(g.arity, g.extent, g.subdivided, g.search)
    else:
        sc.voronoi(g.extent, g.sites, g.nb_offsets, g.nb_ids)
    mix = w["mix"]
    if isinstance(mix, (list, tuple)):
        sc.dust_multi(w["rho"], mix, w.get("store_abs", False))
    else:
        sc.dust(w["rho"], mix.kappaext, mix.albedo, mix.asymmpar, w.get("store_abs", False))
    for c in w["comps"]:
        geo = c.geometry
        ts = getattr(geo, "s", None)
        if ts is not None:
            sc.star(geo.params(), c.luminosities, geo.s, geo.M)
        else:
            sc.star(geo.params(), c.luminosities)
    for ins in w["instr"].instruments:
        sc.instrument(ins.p)
    return sc


def exponcutoff_sample(xmax, n, seed=4357, first=0):
    """Random::exponcutoff(xmax) on packet streams [first, first+n)."""
    out = np.zeros(n)
    lib().or_exponcutoff_sample(xmax, seed, first, n, dp(out))
    return out


def pixelondetector(params10, xyz):
    """SingleFrameInstrument::pixelondetector of positions xyz[n, 3]: (pixel indices, bfkobs)."""
    r = f64(xyz).reshape(-1, 3)
    pix = np.zeros(r.shape[0], np.int32)
    kobs = np.zeros(3)
    lib().or_pixelondetector(dp(f64(params10)), r.shape[0], dp(r), pix.ctypes.data_as(_I32), dp(kobs))
    return pix, kobs


class SphMedium:
    """The oracle's SPHDustDistribution (oracle/sph.cpp) over rows of x y z h M Z T."""

    def __init__(self, rows, fdust=1.0, tmax=0.0):
        self.L = lib()
        self.rows = f64(rows).reshape(-1, 7)
        self.h = self.L.or_sph_create(dp(self.rows), self.rows.shape[0], fdust, tmax)

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "L", None) is not None:
            self.L.or_sph_destroy(self.h)
            self.h = None

    def count(self):
        return self.L.or_sph_count(self.h)

    def mass(self):
        return self.L.or_sph_mass(self.h)

    def grid(self):
        out = np.zeros((3, 21))
        self.L.or_sph_grid(self.h, dp(out))
        return out

    def cell_list(self, i, j, k):
        n = self.L.or_sph_cell_list(self.h, i, j, k, None, 0)
        out = np.zeros(n, np.int32)
        self.L.or_sph_cell_list(self.h, i, j, k, out.ctypes.data_as(_I32), n)
        return out

    def mass_in_box(self, boxes):
        b = f64(boxes).reshape(-1, 6)
        out = np.zeros(b.shape[0])
        self.L.or_sph_mass_in_box(self.h, dp(b), b.shape[0], dp(out))
        return out

    def density(self, pos):
        p = f64(pos).reshape(-1, 3)
        out = np.zeros(p.shape[0])
        self.L.or_sph_density(self.h, dp(p), p.shape[0], dp(out))
        return out

    def subdivide(self, arity, extent, minlevel, maxlevel, maxmassfraction, maxopticaldepth=0.0, cap=1 << 22):
        """TreeDustGrid's subdivision with the box calculator: (flags, masses) in creation order."""
        e = f64(extent)
        flags = np.zeros(cap, np.uint8)
        masses = np.zeros(cap)
        n = ctypes.c_int64()
        rc = self.L.or_sph_subdivide(self.h, arity, dp(e), minlevel, maxlevel, maxmassfraction, maxopticaldepth,
                                     flags.ctypes.data_as(_U8), dp(masses), cap, ctypes.byref(n))
        assert rc == 0, self.L.or_sph_last_error()
        return flags[:n.value].copy(), masses[:n.value].copy()


def myerf(x):
    return lib().or_sph_myerf(float(x))
