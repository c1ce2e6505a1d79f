"""Host-side mirror of SKIRT7's simulation items on the photon-packet transport path.

The classes keep the reference's names and argument meanings (file:line in each docstring) and
forward all work to the HIP library through the C-ABI (skirt7_amd.capi). Errors from the library
raise FatalError, as the reference's FATALERROR does.
"""
import ctypes
import math

import numpy as np

from . import capi
from .capi import SG_FP64, SG_MIXED, SG_NEIGHBOR, SG_TOPDOWN, as_f64, as_i32, as_i64, as_u8, check, dptr, lib

FatalError = capi.SkirtError


def _handle():
    return ctypes.c_void_p()


# ----------------------------------------------------------------------------------- grids


class DustGridPath:
    """Segments of one path (DustGridPath.hpp:27-168): cell numbers m, lengths ds, cumulative s."""

    def __init__(self, m, ds):
        self.m = np.asarray(m, dtype=np.int32)
        self.ds = np.asarray(ds, dtype=np.float64)
        self.s = np.cumsum(self.ds) if len(self.ds) else np.zeros(0)

    def size(self):
        return len(self.m)


class DustGrid:
    """Common interface: path (DustGrid.hpp:106) and whichcell (DustGrid.hpp:89)."""

    _h = None

    def numCells(self):
        n = ctypes.c_int64()
        check(lib().sg_grid_num_cells(self._h, ctypes.byref(n)))
        return n.value

    def paths(self, rk, max_seg=4096, precision=SG_FP64):
        """Segment lists for n rays rk[n,6] = (rx,ry,rz,kx,ky,kz): (counts, m[n,max_seg], ds[n,max_seg])."""
        rk = as_f64(rk).reshape(-1, 6)
        n = rk.shape[0]
        counts = np.zeros(n, np.int32)
        m = np.zeros((n, max_seg), np.int32)
        ds = np.zeros((n, max_seg), np.float64)
        check(lib().sg_grid_paths(self._h, n, dptr(rk), max_seg, precision, counts.ctypes.data_as(capi._I32),
                                  m.ctypes.data_as(capi._I32), dptr(ds)))
        return counts, m, ds

    def path(self, r, k, precision=SG_FP64, max_seg=65536):
        counts, m, ds = self.paths(np.concatenate([r, k])[None, :], max_seg, precision)
        if counts[0] < 0:
            raise FatalError("path longer than %d segments" % max_seg)
        return DustGridPath(m[0, :counts[0]], ds[0, :counts[0]])

    def whichcell(self, r):
        r = as_f64(r).reshape(-1, 3)
        out = np.zeros(r.shape[0], np.int32)
        check(lib().sg_grid_whichcell(self._h, r.shape[0], dptr(r), out.ctypes.data_as(capi._I32)))
        return out

    def __del__(self):
        if self._h is not None and capi._lib is not None:
            capi._lib.sg_grid_destroy(self._h)
            self._h = None


def linmesh(vmin, vmax, n):
    """Cell borders of a LinMesh scaled to [vmin, vmax] (LinMesh.cpp:19-24, NR.hpp lingrid,
    CartesianDustGrid.cpp:33-35): (0 + i*(1/n)) * (vmax - vmin) + vmin."""
    dx = (1.0 - 0.0) / n
    t = np.array([0.0 + i * dx for i in range(n + 1)])
    return t * (vmax - vmin) + vmin


class CartesianDustGrid(DustGrid):
    """CartesianDustGrid (CartesianDustGrid.cpp:28-43,136-283)."""

    def __init__(self, extent, xv, yv, zv):
        self.extent = as_f64(extent)
        self.xv, self.yv, self.zv = as_f64(xv), as_f64(yv), as_f64(zv)
        h = _handle()
        check(lib().sg_grid_create_cartesian(dptr(self.extent), dptr(self.xv), len(self.xv), dptr(self.yv),
                                             len(self.yv), dptr(self.zv), len(self.zv), ctypes.byref(h)))
        self._h = h

    @classmethod
    def uniform(cls, extent, nx, ny, nz):
        e = as_f64(extent)
        return cls(e, linmesh(e[0], e[3], nx), linmesh(e[1], e[4], ny), linmesh(e[2], e[5], nz))

    def cell_boxes(self):
        Nx, Ny, Nz = len(self.xv) - 1, len(self.yv) - 1, len(self.zv) - 1
        i, j, k = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
        return np.stack([self.xv[i], self.yv[j], self.zv[k], self.xv[i + 1], self.yv[j + 1], self.zv[k + 1]],
                        -1).reshape(-1, 6)


class TreeDustGrid(DustGrid):
    """TreeDustGrid (TreeDustGrid.cpp:49-162,371-643) from a frozen subdivision: one flag per node
    in the reference's breadth-first node numbering."""

    arity = 8

    def __init__(self, extent, subdivided, search=SG_NEIGHBOR):
        self.extent = as_f64(extent)
        self.subdivided = as_u8(subdivided)
        self.search = search
        h = _handle()
        check(lib().sg_grid_create_tree(self.arity, dptr(self.extent), self.subdivided.ctypes.data_as(capi._U8),
                                        len(self.subdivided), search, ctypes.byref(h)))
        self._h = h

    @classmethod
    def subdivide(cls, extent, geometry, scale, total_mass, minlevel=2, maxlevel=6, maxmassfraction=1e-6,
                  nsamples=100, seed=4357, search=SG_NEIGHBOR, maxopticaldepth=0.0, maxdensdispfraction=0.0):
        """Builds the subdivision from a dust geometry density (TreeDustGrid.cpp:166-230) with the reference's
        stopping criteria (0 switches one off; all three off subdivides to maxlevel); geometry and scale may be
        lists, one per dust component, for a CompDustDistribution's summed density."""
        e = as_f64(extent)
        cap = 1 << 27
        flags = np.zeros(cap, np.uint8)
        n = ctypes.c_int64()
        geoms = list(geometry) if isinstance(geometry, (list, tuple)) else [geometry]
        scales = as_f64(np.atleast_1d(scale))
        if len(scales) != len(geoms):
            raise FatalError("one scale per dust component is needed")
        g = as_f64(np.concatenate([x.params() for x in geoms]))
        check(lib().sg_tree_subdivide_multi(cls.arity, dptr(e), len(geoms), dptr(g), dptr(scales), total_mass,
                                            minlevel, maxlevel, maxmassfraction, maxopticaldepth,
                                            maxdensdispfraction, nsamples, seed, flags.ctypes.data_as(capi._U8),
                                            cap, ctypes.byref(n)))
        return cls(e, flags[:n.value].copy(), search)


    @classmethod
    def subdivide_device(cls, extent, geometry, scale, total_mass, minlevel=2, maxlevel=6, maxmassfraction=1e-6,
                         nsamples=100, seed=4357, search=SG_NEIGHBOR, maxopticaldepth=0.0, maxdensdispfraction=0.0):
        """subdivide built on the device, level by level (exponential-disc geometries): the same sample
        positions and stopping criteria, so the same subdivision up to the rounding of the device's exp, log
        and pow at a node whose criterion value lies within it of the threshold."""
        e = as_f64(extent)
        geoms = list(geometry) if isinstance(geometry, (list, tuple)) else [geometry]
        scales = as_f64(np.atleast_1d(scale))
        if len(scales) != len(geoms):
            raise FatalError("one scale per dust component is needed")
        g = as_f64(np.concatenate([x.params() for x in geoms]))
        cap = max(1 << 22, int(4.0 / maxmassfraction) if maxmassfraction > 0 else 0)
        while True:
            flags = np.zeros(cap, np.uint8)
            n = ctypes.c_int64()
            st = lib().sg_tree_subdivide_device(cls.arity, dptr(e), len(geoms), dptr(g), dptr(scales), total_mass,
                                                minlevel, maxlevel, maxmassfraction, maxopticaldepth,
                                                maxdensdispfraction, nsamples, seed, flags.ctypes.data_as(capi._U8),
                                                cap, ctypes.byref(n))
            if st == -5:
                cap = n.value
                continue
            check(st)
            return cls(e, flags[:n.value].copy(), search)

    @classmethod
    def subdivide_sph(cls, extent, sph, minlevel=2, maxlevel=6, maxmassfraction=1e-6, maxopticaldepth=0.0,
                      maxdensdispfraction=0.0, nsamples=100, seed=4357, search=SG_NEIGHBOR, masses=False):
        """TreeDustGrid over an SPHDustDistribution (TreeDustGrid.cpp:49-230), built on the device: node masses
        from the particles' mass in the node's box (TreeNodeBoxDensityCalculator), or from density samples
        when the dispersion criterion is set (TreeDustGrid.cpp:70). With masses=True also returns each node's
        mass (with masses in box: every node beyond minlevel; with samples: the nodes the criteria evaluated;
        -1 elsewhere)."""
        e = as_f64(extent)
        # a build that outgrows cap reports its node count: the second try fits
        cap = max(1 << 22, int(4.0 / maxmassfraction) if maxmassfraction > 0 else 0)
        while True:
            flags = np.zeros(cap, np.uint8)
            mass = np.zeros(cap) if masses else None
            n = ctypes.c_int64()
            st = lib().sg_tree_subdivide_sph(cls.arity, dptr(e), sph._h, minlevel, maxlevel, maxmassfraction,
                                             maxopticaldepth, maxdensdispfraction, nsamples, seed,
                                             flags.ctypes.data_as(capi._U8), dptr(mass) if masses else None, cap,
                                             ctypes.byref(n))
            if st == -5:
                cap = n.value
                continue
            check(st)
            grid = cls(e, flags[:n.value].copy(), search)
            return (grid, mass[:n.value].copy()) if masses else grid

    @classmethod
    def subdivide_field(cls, extent, field, minlevel=2, maxlevel=30, maxmassfraction=1e-6, search=SG_NEIGHBOR,
                        maxopticaldepth=0.0):
        """The subdivision with exact node masses from a voxel density field field[i, j, k] over the box (mass
        fraction and optical-depth criteria)."""
        e = as_f64(extent)
        f = as_f64(field)
        nx, ny, nz = f.shape
        cap = 1 << 22
        while True:
            flags = np.zeros(cap, np.uint8)
            n = ctypes.c_int64()
            st = lib().sg_tree_subdivide_field_od(cls.arity, dptr(e), dptr(f), nx, ny, nz, minlevel, maxlevel,
                                                  maxmassfraction, maxopticaldepth, flags.ctypes.data_as(capi._U8),
                                                  cap, ctypes.byref(n))
            if st == -5:
                cap = n.value
                continue
            check(st)
            return cls(e, flags[:n.value].copy(), search)


class OctTreeDustGrid(TreeDustGrid):
    """OctTreeDustGrid (OctTreeDustGrid.cpp:32-43, OctTreeNode.cpp)."""
    arity = 8


class BinTreeDustGrid(TreeDustGrid):
    """BinTreeDustGrid with the Alternating direction method (BinTreeDustGrid.cpp:19-53)."""
    arity = 2


def voronoi_neighbors(extent, sites):
    """Neighbour lists of the Voronoi cells of `sites` in the box (VoronoiMesh.cpp:297-405): (offsets, ids)."""
    e = as_f64(extent)
    s = as_f64(sites).reshape(-1, 3)
    n = s.shape[0]
    off = np.zeros(n + 1, np.int64)
    cap = 24 * n + 64
    nids = ctypes.c_int64()
    while True:
        ids = np.zeros(cap, np.int32)
        st = lib().sg_voronoi_neighbors(dptr(e), dptr(s), n, off.ctypes.data_as(capi._I64),
                                        ids.ctypes.data_as(capi._I32), cap, ctypes.byref(nids))
        if st == -5:
            cap = nids.value
            continue
        check(st)
        return off, ids[:nids.value].copy()


def voronoi_cells(extent, particles):
    """VoronoiMesh::buildMesh with nearby particles removed (VoronoiMesh.cpp:297-350, as VoronoiDustGrid builds
    its mesh from generated or imported particles): the particles sorted on x; one that lies within
    eps = 1e-12 |box widths| of a later particle in that order is dropped; the survivors, in sorted order, are
    the cells (cell m is the m-th survivor). Returns their positions."""
    e = as_f64(extent)
    p = as_f64(particles).reshape(-1, 3)
    n = p.shape[0]
    order = np.argsort(p[:, 0], kind="stable")
    x = p[order, 0]
    eps = 1e-12 * float(np.linalg.norm(e[3:] - e[:3]))
    keep = np.ones(n, bool)
    for i in np.flatnonzero(np.diff(x) < eps):
        j = i + 1
        while j < n and x[j] - x[i] < eps:
            d = p[order[j]] - p[order[i]]
            if float(d @ d) < eps * eps:
                keep[i] = False
                break
            j += 1
    return p[order[keep]]


class VoronoiDustGrid(DustGrid):
    """VoronoiDustGrid / VoronoiMesh (VoronoiMesh.cpp:297-405,541-570,778-873) from generator sites and
    neighbour lists (walls -1..-6); the lists are built from the sites when not given. The sites are the cells
    in the given order; voronoi_cells turns a particle set into them as the reference's mesh does."""

    def __init__(self, extent, sites, nb_offsets=None, nb_ids=None):
        self.extent = as_f64(extent)
        self.sites = as_f64(sites).reshape(-1, 3)
        if nb_offsets is None:
            nb_offsets, nb_ids = voronoi_neighbors(self.extent, self.sites)
        self.nb_offsets = as_i64(nb_offsets)
        self.nb_ids = as_i32(nb_ids)
        h = _handle()
        check(lib().sg_grid_create_voronoi(dptr(self.extent), dptr(self.sites), self.sites.shape[0],
                                           self.nb_offsets.ctypes.data_as(capi._I64),
                                           self.nb_ids.ctypes.data_as(capi._I32), ctypes.byref(h)))
        self._h = h


# ----------------------------------------------------------------------------------- geometry


class ExpDiskGeometry:
    """ExpDiskGeometry (ExpDiskGeometry.cpp:23-43,117-161): scale length hR, scale height hz, optional
    truncations."""

    def __init__(self, hR, hz, Rmax=0.0, zmax=0.0, Rmin=0.0):
        self.hR, self.hz, self.Rmax, self.zmax, self.Rmin = hR, hz, Rmax, zmax, Rmin

    def params(self):
        p = np.zeros(capi.SG_STAR_NPARAM)
        p[0] = 1
        p[1:6] = [self.hR, self.hz, self.Rmax, self.zmax, self.Rmin]
        return p

    def SigmaZ(self):
        """Face-on surface density of the normalised geometry (ExpDiskGeometry.cpp:175-187)."""
        intphi = 2.0 * math.pi
        intz = -2.0 * self.hz * math.expm1(-self.zmax / self.hz) if self.zmax > 0 else 2.0 * self.hz
        tmin = math.exp(-self.Rmin / self.hR) * (1.0 + self.Rmin / self.hR) if self.Rmin > 0 else 1.0
        tmax = math.exp(-self.Rmax / self.hR) * (1.0 + self.Rmax / self.hR) if self.Rmax > 0 else 0.0
        rho0 = 1.0 / (self.hR * self.hR * (tmin - tmax) * intphi * intz)
        if self.Rmin > 0:
            return 0.0
        if self.zmax > 0:
            return -2.0 * rho0 * self.hz * math.expm1(-self.zmax / self.hz)
        return 2.0 * rho0 * self.hz


class SersicGeometry:
    """SersicGeometry (SersicGeometry.cpp:30-42,85-90) with its SersicFunction mass table."""

    def __init__(self, n, reff):
        self.n, self.reff = n, reff
        self.s = np.zeros(101)
        self.M = np.zeros(101)
        check(lib().sg_sersic_table(n, dptr(self.s), dptr(self.M)))

    def params(self):
        p = np.zeros(capi.SG_STAR_NPARAM)
        p[0] = 2
        p[1:3] = [self.n, self.reff]
        return p


class SpiralStructureGeometryDecorator:
    """SpiralStructureGeometryDecorator (SpiralStructureGeometryDecorator.cpp:23-40,165-229)."""

    def __init__(self, geometry, arms=1, pitch=0.1745329252, radius=1.0, phase=0.0, perturbWeight=0.3, index=1):
        self.geometry = geometry
        self.arms, self.pitch, self.radius, self.phase, self.perturbWeight, self.index = (
            arms, pitch, radius, phase, perturbWeight, index)

    def params(self):
        p = self.geometry.params()
        p[6] = 1
        p[7:13] = [self.arms, self.pitch, self.radius, self.phase, self.perturbWeight, self.index]
        return p

    def SigmaZ(self):
        return self.geometry.SigmaZ()

    @property
    def s(self):
        return getattr(self.geometry, "s", None)

    @property
    def M(self):
        return getattr(self.geometry, "M", None)


def geometry_density(geometry, xyz):
    xyz = as_f64(xyz).reshape(-1, 3)
    out = np.zeros(xyz.shape[0])
    check(lib().sg_geometry_density(dptr(as_f64(geometry.params())), xyz.shape[0], dptr(xyz), dptr(out)))
    return out


# ----------------------------------------------------------------------------------- dust


class SimpleOligoDustMix:
    """SimpleOligoDustMix (SimpleOligoDustMix.cpp:22-51) with the derived per-wavelength
    properties of DustMix::setupSelfAfter (DustMix.cpp:44-91)."""

    def __init__(self, opacities, albedos, asymmetryParameters):
        kext = np.atleast_1d(as_f64(opacities))
        alb = np.atleast_1d(as_f64(albedos))
        g = np.atleast_1d(as_f64(asymmetryParameters))
        if not (len(kext) == len(alb) == len(g)):
            raise FatalError("The number of opacities, albedos and asymmetry parameters differ")
        kabs = kext * (1.0 - alb)
        ksca = kext * alb
        sumabs = 0.0 + kabs
        sumsca = 0.0 + ksca
        sumgsca = 0.0 + g * ksca
        sumext = sumabs + sumsca
        self.albedo = np.where(sumext != 0, sumsca / np.where(sumext != 0, sumext, 1.0), 0.0)
        self.asymmpar = np.where(sumsca != 0, sumgsca / np.where(sumsca != 0, sumsca, 1.0), 0.0)
        mu = 0.0 + 1.0
        self.kappaext = sumext / mu
        self.kappaabs = sumabs / mu
        self.kappasca = sumsca / mu

    def Nlambda(self):
        return len(self.kappaext)


class DustSystem:
    """OligoDustSystem (DustSystem.cpp:935-976, OligoDustSystem.cpp:53-87): the cell densities rho[m] of a
    dust component on a grid and its dust mix; or, for several components (a CompDustDistribution), a list
    of density arrays and a list of mixes, one per component."""

    def __init__(self, grid, rho, mix, writeMeanIntensity=False):
        self.grid = grid
        multi = isinstance(mix, (list, tuple))
        self.mixes = list(mix) if multi else [mix]
        self.mix = self.mixes[0]
        rhos = [as_f64(r) for r in rho] if multi else [as_f64(rho)]
        if len(rhos) != len(self.mixes):
            raise FatalError("one density array per dust component is needed")
        self.rhos = rhos
        self.rho = rhos[0] if len(rhos) == 1 else sum(rhos[1:], rhos[0].copy())
        self.ncomp = len(rhos)
        self.nlambda = self.mix.Nlambda()
        h = _handle()
        if self.ncomp == 1:
            check(lib().sg_dust_create(grid._h, dptr(rhos[0]), len(rhos[0]), self.nlambda, dptr(mix.kappaext),
                                       dptr(mix.albedo), dptr(mix.asymmpar), int(bool(writeMeanIntensity)),
                                       ctypes.byref(h)))
        else:
            if any(m.Nlambda() != self.nlambda for m in self.mixes):
                raise FatalError("the dust mixes disagree on the number of wavelengths")
            allrho = as_f64(np.concatenate(rhos))
            stack = lambda name: as_f64(np.concatenate([getattr(m, name) for m in self.mixes]))
            self._keep = (allrho, stack("kappaext"), stack("kappasca"), stack("albedo"), stack("asymmpar"))
            check(lib().sg_dust_create_multi(grid._h, self.ncomp, dptr(allrho), len(rhos[0]), self.nlambda,
                                             dptr(self._keep[1]), dptr(self._keep[2]), dptr(self._keep[3]),
                                             dptr(self._keep[4]), int(bool(writeMeanIntensity)), ctypes.byref(h)))
        self._h = h
        self.writeMeanIntensity = bool(writeMeanIntensity)

    def storeabsorptionrates(self):
        return self.writeMeanIntensity

    def opticaldepths(self, rk, ell=0, precision=SG_FP64):
        """DustSystem::opticaldepth (DustSystem.cpp:960-976) for rays rk[n,6]."""
        rk = as_f64(rk).reshape(-1, 6)
        out = np.zeros(rk.shape[0])
        check(lib().sg_dust_optical_depths(self._h, ell, rk.shape[0], dptr(rk), precision, dptr(out)))
        return out

    def pathlengths(self, rk, taus, ell=0, precision=SG_FP64):
        """DustGridPath::pathlength after fillOpticalDepth (DustGridPath.cpp:162-173)."""
        rk = as_f64(rk).reshape(-1, 6)
        taus = as_f64(taus)
        out = np.zeros(rk.shape[0])
        check(lib().sg_dust_path_lengths(self._h, ell, rk.shape[0], dptr(rk), dptr(taus), precision, dptr(out)))
        return out

    def Labs(self):
        """OligoDustSystem::Labs table, (ncells, nlambda) (OligoDustSystem.cpp:82-85)."""
        out = np.zeros((len(self.rho), self.nlambda))
        check(lib().sg_dust_read_absorption(self._h, dptr(out)))
        return out

    def clear(self):
        check(lib().sg_dust_clear_absorption(self._h))

    def __del__(self):
        if getattr(self, "_h", None) is not None and capi._lib is not None:
            capi._lib.sg_dust_destroy(self._h)
            self._h = None


def sample_density(grid, geometry, scale, nsamples=100, seed=4357, device=False):
    """Cell densities by DustSystem::setSampleDensityBody (DustSystem.cpp:145-172), on the host or (device=True,
    exponential discs) on the device at the same positions."""
    out = np.zeros(grid.numCells())
    fn = lib().sg_grid_sample_density_device if device else lib().sg_grid_sample_density
    check(fn(grid._h, dptr(as_f64(geometry.params())), scale, nsamples, seed, dptr(out)))
    return out


def resample_density(grid, geometry, scale, nsamples=100, seed=4357):
    """The same sampling on the device, stored as a Cartesian grid's densities (no host round trip)."""
    check(lib().sg_grid_resample_density(grid._h, dptr(as_f64(geometry.params())), scale, nsamples, seed))


class SPHDustDistribution:
    """SPHDustDistribution (SPHDustDistribution.cpp:38-256) from a table of SPH gas particles, rows of
    x y z h M Z T (lengths in the grid's unit, masses in any mass unit): particles hotter than
    max_temperature are ignored (when both are positive), the dust is dust_fraction of the metal mass, a
    negative total gas or metal mass suppresses all dust. Densities and masses in a box are evaluated on the
    device over the reference's 20^3 particle grid (SPHGasParticleGrid.cpp)."""

    def __init__(self, particles, dust_fraction, max_temperature=0.0):
        rows = as_f64(particles).reshape(-1, 7)
        h = _handle()
        check(lib().sg_sph_create(dptr(rows), rows.shape[0], dust_fraction, max_temperature, ctypes.byref(h)))
        self._h = h
        cnt, mass, ent = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
        self.grid_borders = np.zeros((3, 21))
        check(lib().sg_sph_info(self._h, ctypes.byref(cnt), ctypes.byref(mass), dptr(self.grid_borders),
                                ctypes.byref(ent)))
        self.count = cnt.value
        self.list_entries = ent.value
        self._mass = mass.value

    def __del__(self):
        if getattr(self, "_h", None) is not None and capi._lib is not None:
            capi._lib.sg_sph_destroy(self._h)
            self._h = None

    def mass(self):
        """SPHDustDistribution::mass (SPHDustDistribution.cpp:248-256)."""
        return self._mass

    def massInBox(self, boxes):
        """SPHDustDistribution::massInBox (SPHDustDistribution.cpp:226-236) of boxes[n, 6] (min then max)."""
        b = as_f64(boxes).reshape(-1, 6)
        out = np.zeros(b.shape[0])
        check(lib().sg_sph_mass_in_box(self._h, dptr(b), b.shape[0], dptr(out)))
        return out

    def density(self, positions):
        """SPHDustDistribution::density (SPHDustDistribution.cpp:192-202) at positions[n, 3]."""
        p = as_f64(positions).reshape(-1, 3)
        out = np.zeros(p.shape[0])
        check(lib().sg_sph_density(self._h, dptr(p), p.shape[0], dptr(out)))
        return out

    def cell_list(self, i, j, k):
        """The particle grid's list of cell (i, j, k) (SPHGasParticleGrid.cpp:94-132)."""
        n = ctypes.c_int64()
        check(lib().sg_sph_cell_list(self._h, i, j, k, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, np.int32)
        check(lib().sg_sph_cell_list(self._h, i, j, k, out.ctypes.data_as(capi._I32), n.value, ctypes.byref(n)))
        return out


def sph_density(grid, sph, node_masses=None):
    """TreeDustGrid::density (TreeDustGrid.cpp:655-660) over an SPHDustDistribution: each cell's mass in its
    box over its volume; node_masses, the masses subdivide_sph(masses=True) returned for this tree, serve as
    the cells' masses, else they are evaluated on the device."""
    out = np.zeros(grid.numCells())
    if node_masses is None:
        check(lib().sg_grid_sph_density(grid._h, sph._h, None, 0, dptr(out)))
    else:
        m = as_f64(node_masses)
        check(lib().sg_grid_sph_density(grid._h, sph._h, dptr(m), len(m), dptr(out)))
    return out


def knn_distance(positions, k=32):
    """Distance from each point to its k-th nearest other point (on the device): smoothing lengths for a
    generated particle set."""
    p = as_f64(positions).reshape(-1, 3)
    out = np.zeros(p.shape[0])
    check(lib().sg_knn_distance(dptr(p), p.shape[0], k, dptr(out)))
    return out


def field_density(grid, field):
    """Cell densities from a voxel density field: mass in the cell's box over its volume."""
    f = as_f64(field)
    out = np.zeros(grid.numCells())
    check(lib().sg_grid_field_density(grid._h, dptr(f), f.shape[0], f.shape[1], f.shape[2], dptr(out)))
    return out


def faceon_normalization(geometry, tau, kappaext):
    """FaceOnDustCompNormalization::normalizationFactor (FaceOnDustCompNormalization.cpp:67-74): only an
    axisymmetric geometry (an AxGeometry: here the exponential disc) has a face-on surface density; a Sersic
    sphere or a spiral decorator (a GenGeometry) raises the reference's fatal error."""
    if not isinstance(geometry, ExpDiskGeometry):
        raise FatalError("Geometry is not axisymmetric")
    return tau / (geometry.SigmaZ() * kappaext)


# ----------------------------------------------------------------------------------- stars


class OligoStellarComp:
    """OligoStellarComp (OligoStellarComp.cpp:26-76): a geometry with one luminosity per wavelength
    (here directly in the caller's power unit)."""

    def __init__(self, geometry, luminosities):
        self.geometry = geometry
        self.luminosities = np.atleast_1d(as_f64(luminosities))


class StellarSystem:
    """StellarSystem (StellarSystem.cpp:116-158) with its emission bias (default 0.5)."""

    def __init__(self, components, emissionBias=0.5):
        self.components = list(components)
        self.emissionBias = emissionBias
        nl = len(self.components[0].luminosities)
        params, lum, s_arr, M_arr = self._records(self.components)
        h = _handle()
        check(lib().sg_stars_create(len(self.components), dptr(params), nl, dptr(lum), dptr(s_arr), dptr(M_arr),
                                    101, emissionBias, ctypes.byref(h)))
        self._h = h
        self.nlambda = nl

    @staticmethod
    def _records(components):
        params = as_f64(np.concatenate([c.geometry.params() for c in components]))
        lum = as_f64(np.concatenate([c.luminosities for c in components]))
        ts = [c.geometry.s for c in components if getattr(c.geometry, "s", None) is not None]
        tM = [c.geometry.M for c in components if getattr(c.geometry, "M", None) is not None]
        s_arr = as_f64(np.concatenate(ts)) if ts else np.zeros(1)
        M_arr = as_f64(np.concatenate(tM)) if tM else np.zeros(1)
        return params, lum, s_arr, M_arr

    def update(self, components, emissionBias=None):
        """New components (same count and wavelengths) in place of the current ones."""
        components = list(components)
        if emissionBias is not None:
            self.emissionBias = emissionBias
        params, lum, s_arr, M_arr = self._records(components)
        check(lib().sg_stars_update(self._h, len(components), dptr(params), dptr(lum), dptr(s_arr), dptr(M_arr),
                                    101, self.emissionBias))
        self.components = components

    def luminosity(self, ell):
        return sum(c.luminosities[ell] for c in self.components)

    def __del__(self):
        if getattr(self, "_h", None) is not None and capi._lib is not None:
            capi._lib.sg_stars_destroy(self._h)
            self._h = None


# ----------------------------------------------------------------------------------- instruments


class FrameInstrument:
    """FrameInstrument (FrameInstrument.cpp:33-47; DistantInstrument.cpp:27-50;
    SingleFrameInstrument.cpp:26-42,130-147). Angles in radians."""

    def __init__(self, distance, inclination, azimuth, positionangle, pixelsX, pixelsY, fieldOfViewX,
                 fieldOfViewY, centerX=0.0, centerY=0.0):
        self.p = np.array([distance, inclination, azimuth, positionangle, pixelsX, pixelsY, fieldOfViewX,
                           fieldOfViewY, centerX, centerY], dtype=np.float64)
        self.Nx, self.Ny = int(pixelsX), int(pixelsY)


class InstrumentSystem:
    def __init__(self, instruments, nlambda=1):
        self.instruments = list(instruments)
        self.nlambda = nlambda
        params = as_f64(np.concatenate([i.p for i in self.instruments])) if self.instruments else np.zeros(1)
        h = _handle()
        check(lib().sg_instruments_create(len(self.instruments), dptr(params), nlambda, ctypes.byref(h)))
        self._h = h

    def frame(self, i):
        """Raw detected luminosities of instrument i, shape (nlambda, Ny, Nx)."""
        ins = self.instruments[i]
        out = np.zeros((self.nlambda, ins.Ny, ins.Nx))
        check(lib().sg_instruments_read(self._h, i, dptr(out)))
        return out

    def clear(self):
        check(lib().sg_instruments_clear(self._h))

    def update(self, instruments):
        """New viewing parameters (same pixel counts) in place of the current ones; frames keep their tallies."""
        instruments = list(instruments)
        if len(instruments) != len(self.instruments):
            raise FatalError("an instrument update keeps the number of instruments")
        params = as_f64(np.concatenate([i.p for i in instruments])) if instruments else np.zeros(1)
        check(lib().sg_instruments_update(self._h, dptr(params)))
        self.instruments = instruments

    def __del__(self):
        if getattr(self, "_h", None) is not None and capi._lib is not None:
            capi._lib.sg_instruments_destroy(self._h)
            self._h = None


# ----------------------------------------------------------------------------------- simulation


class OligoMonteCarloSimulation:
    """OligoMonteCarloSimulation / MonteCarloSimulation (MonteCarloSimulation.cpp:68-114,241-297) on one
    GPU. Each packet draws from its own Philox stream keyed by (seed, member)."""

    def __init__(self, dustSystem, stellarSystem, instrumentSystem, packages=1e6, minWeightReduction=1e4,
                 minScattEvents=0, scattBias=0.5, seed=4357, member=0, tauCut=0.0):
        self.ds, self.ss, self.instr = dustSystem, stellarSystem, instrumentSystem
        self.packages = packages
        self.mc = np.array([packages, minWeightReduction, minScattEvents, scattBias, seed,
                            stellarSystem.emissionBias, member, tauCut], dtype=np.float64)
        h = _handle()
        check(lib().sg_sim_create(dustSystem._h if dustSystem is not None else None, stellarSystem._h,
                                  instrumentSystem._h, dptr(self.mc), ctypes.byref(h)))
        self._h = h
        self.counters = None

    def setChunkParams(self, threads=1):
        """Npp of MonteCarloSimulation::setChunkParams (MonteCarloSimulation.cpp:68-114) for one process."""
        nl = self.ss.nlambda
        total = 1 if threads == 1 else math.ceil(max(10.0 * threads / nl, self.packages / 1e7))
        chunksize = math.ceil(self.packages / total)
        return total * chunksize

    def run_packets(self, ell, first, count, npp, precision=SG_FP64):
        """Packets [first, first+count) of wavelength ell with normalisation L(ell)/npp."""
        c = np.zeros(capi.SG_NCOUNTERS, np.uint64)
        check(lib().sg_sim_run(self._h, ell, int(first), int(count), int(npp), precision,
                               c.ctypes.data_as(capi._U64)))
        self.counters = dict(zip(capi.COUNTER_NAMES, (int(x) for x in c)))
        return self.counters

    def run_packets_static(self, ell, first, count, npp, precision=SG_FP64):
        """run_packets without any device allocation (work memory in the library's static device array);
        at most 2^19 packets per call."""
        c = np.zeros(capi.SG_NCOUNTERS, np.uint64)
        check(lib().sg_sim_run_static(self._h, ell, int(first), int(count), int(npp), precision,
                                      c.ctypes.data_as(capi._U64)))
        self.counters = dict(zip(capi.COUNTER_NAMES, (int(x) for x in c)))
        return self.counters

    def runstellaremission(self, precision=SG_FP64):
        """The stellar emission phase (MonteCarloSimulation.cpp:241-255): all wavelengths, Npp packets each."""
        npp = self.setChunkParams(1)
        total = None
        for ell in range(self.ss.nlambda):
            c = self.run_packets(ell, 0, npp, npp, precision)
            total = c if total is None else {k: total[k] + c[k] for k in c}
        self.counters = total
        return total

    def __del__(self):
        if getattr(self, "_h", None) is not None and capi._lib is not None:
            capi._lib.sg_sim_destroy(self._h)
            self._h = None


def run_batch(simulations, counts, ells=None, precision=SG_FP64):
    """Batched population: simulation i runs packets [0, counts[i]) at wavelength ells[i] (default 0)
    with normalisation L/counts[i], all members in one persistent launch per round. Each member's result
    equals its own run_packets(ell, 0, count, count); the returned counters are summed over members."""
    sims = list(simulations)
    n = len(sims)
    if n == 0:
        raise ValueError("a batched run needs at least one member")
    counts = np.ascontiguousarray(counts, dtype=np.uint64)
    ells = np.zeros(n, np.int32) if ells is None else np.ascontiguousarray(ells, dtype=np.int32)
    if counts.shape != (n,) or ells.shape != (n,):
        raise ValueError("counts and ells need one entry per member")
    handles = (ctypes.c_void_p * n)(*[s._h.value for s in sims])
    c = np.zeros(capi.SG_NCOUNTERS, np.uint64)
    check(lib().sg_sim_run_batch(n, handles, ells.ctypes.data_as(capi._I32), counts.ctypes.data_as(capi._U64),
                                 precision, c.ctypes.data_as(capi._U64)))
    return dict(zip(capi.COUNTER_NAMES, (int(x) for x in c)))


# ----------------------------------------------------------------------- FitSKIRT evaluation


class GaussianKernel:
    """GaussianKernel (FitSKIRTcore/GaussianKernel.cpp:19-46), normalised to unit sum as
    ConvolutionKernel::setupSelfAfter does (ConvolutionKernel.cpp:23-28). data[y, x]."""

    def __init__(self, fwhm=3.0, dimension=6):
        self.fwhm, self.dimension = float(fwhm), int(dimension)
        self.data = np.zeros((self.dimension, self.dimension))
        check(lib().sg_gaussian_kernel(self.fwhm, self.dimension, dptr(self.data)))


def fit_frames(instrument_systems, reference, kernel, minlum, maxlum, instrument=0, ell=0, return_frames=False):
    """ReferenceImage::chi2 for one stellar component (ReferenceImage.cpp:95-114) over a population: each
    member's frame is convolved with the kernel and its luminosity scale is fitted to the reference image by
    golden section. Returns (lum, chi2) arrays, plus the convolved, masked frames when asked."""
    systems = list(instrument_systems)
    n = len(systems)
    if n == 0:
        raise ValueError("a fit needs at least one member")
    ref = as_f64(reference)
    k = as_f64(kernel.data if isinstance(kernel, GaussianKernel) else kernel)
    if k.ndim != 2 or ref.ndim != 2:
        raise ValueError("kernel and reference must be 2-D (y, x) arrays")
    handles = (ctypes.c_void_p * n)(*[s._h.value for s in systems])
    lum, chi2 = np.zeros(n), np.zeros(n)
    frames = np.zeros((n,) + ref.shape) if return_frames else None
    check(lib().sg_fit_frames(n, handles, instrument, ell, dptr(k), k.shape[1], k.shape[0], dptr(ref), float(minlum),
                              float(maxlum), dptr(lum), dptr(chi2), dptr(frames) if return_frames else None))
    return (lum, chi2, frames) if return_frames else (lum, chi2)


_pipeline_mode = [1]


def set_pipeline(mode):
    """1: event and walk kernels over device queues (default); 0: one persistent lane-pipeline kernel."""
    check(lib().sg_set_pipeline(int(mode)))
    _pipeline_mode[0] = int(mode)


def set_batch_alone(min_packets):
    """Members of run_batch with at least this many packets run one at a time through the event and walk
    kernels; the rest share the persistent batched launches (0: every member shares them)."""
    check(lib().sg_set_batch_alone(int(min_packets)))


def set_mixed_topdown(descend):
    """Mixed-precision walks of TopDown trees: False (default) follow the ropes between leaves (the same
    cells); True descend from the root at every step."""
    check(lib().sg_set_mixed_topdown(1 if descend else 0))


def set_walk_order(order=0):
    """Event/walk pipeline: 1 (or True) takes each pass's walks sorted by a coherence key (peel-offs by
    instrument and pixel, other walks by start point), 2 sorted and claimed by XCD segment (each workgroup
    first takes the walks of its XCD's eighth of the order), 0 (default, or False) as the event kernel
    appended them."""
    check(lib().sg_set_walk_order(int(order)))


def set_emit_batch(packets=0):
    """Packets emitted at once by run_packets (0: the default); longer ranges are transported batch by batch."""
    check(lib().sg_set_emit_batch(int(packets)))


def set_slot_groups(groups=2):
    """Event/walk pipeline: 2 (default) slot groups alternating on two streams; 1 runs every pass on one stream,
    one kernel at a time."""
    check(lib().sg_set_slot_groups(int(groups)))


def set_tuning(block=0, waves_per_cu=0, compact_lanes=0):
    """Launch parameters of sg_sim_run: 0 keeps the current value, a negative value restores the default."""
    check(lib().sg_set_tuning(block, waves_per_cu, compact_lanes))


def get_pipeline():
    return _pipeline_mode[0]


def set_slots(slots=0):
    """In-flight packet slots of the event/walk pipeline (0: the default)."""
    check(lib().sg_set_slots(int(slots)))


def autotune(sim, ell=0, first=0, sample=2_000_000, precision=1, choose_pipeline=False):
    """Times the simulation on packets [first, first+sample) over the launch parameters (the transport
    organisation when choose_pipeline, then compaction threshold, in-flight slots or waves per CU, block
    size), keeps the fastest set for later runs and returns it with the sample's rate; the tallies are left
    as they were."""
    best = np.zeros(5, np.int64)
    rate = ctypes.c_double(0.0)
    check(lib().sg_autotune(sim._h, int(ell), int(first), int(sample), int(precision), int(bool(choose_pipeline)),
                            best.ctypes.data_as(capi._I64), ctypes.byref(rate)))
    _pipeline_mode[0] = int(best[3])
    key = "slots" if best[3] == 1 else "waves_per_cu"
    out = {"pipeline": int(best[3]), "block": int(best[0]), key: int(best[1]), "compact_lanes": int(best[2]),
           "sample_packets_per_s": rate.value}
    if best[3] == 1:
        out["walk_order"] = int(best[4])
    return out


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().sg_device_name(buf, 256))
    return buf.value.decode()
