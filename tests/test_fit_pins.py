"""Pins of the oracle's FitSKIRT evaluation (PSF kernel, nested-loop convolution, golden-section luminosity
fit) against what the mathematics fixes, the product's host-side kernel builder against the oracle, and the
product's host-side voxel-field tree subdivision against brute-force masses. CPU only."""
import math

import numpy as np
import pytest
from scipy import optimize, signal

from tests import oracle_lib as O


def test_gaussian_kernel_shape_and_centre():
    # GaussianKernel.cpp:35-43: the centre is pixel (dimension+1)/2 in 1-based counting, so a 6-pixel
    # kernel peaks at 0-based index 2 and is symmetric about it (indices 0..4), with index 5 unpaired.
    k = O.gaussian_kernel(3.0, 6)
    assert k.shape == (6, 6)
    assert abs(k.sum() - 1.0) < 1e-15
    assert np.unravel_index(np.argmax(k), k.shape) == (2, 2)
    for a, b in ((0, 4), (1, 3)):
        assert np.array_equal(k[:, a], k[:, b]) and np.array_equal(k[a, :], k[b, :])
    assert np.array_equal(k, k.T)
    # odd dimension: centred kernel, symmetric both ways
    k7 = O.gaussian_kernel(3.0, 7)
    assert np.unravel_index(np.argmax(k7), k7.shape) == (3, 3)
    assert np.allclose(k7, k7[::-1, ::-1], rtol=0, atol=0)


def test_gaussian_kernel_half_maximum_at_half_fwhm():
    # The profile falls to one half of its peak at a distance FWHM/2 from the centre: this fixes the
    # sigma = FWHM / 2.3548 conversion (2 sqrt(2 ln 2) = 2.35482...).
    fwhm = 10.0
    k = O.gaussian_kernel(fwhm, 41)
    c = 20
    assert abs(k[c, c + 5] / k[c, c] - 0.5) < 1e-4
    assert abs(k[c, c - 5] / k[c, c] - 0.5) < 1e-4


def test_product_kernel_matches_oracle():
    from skirt7_amd import skirt as S
    for fwhm, dim in ((3.0, 6), (3.0, 7), (1.5, 5), (8.0, 21)):
        assert np.array_equal(S.GaussianKernel(fwhm, dim).data, O.gaussian_kernel(fwhm, dim))


def test_convolution_point_source_places_the_kernel():
    # Convolution.cpp:46-58 scatters input pixel (xi, yi) to (xi - cx + xk, yi - cy + yk): a point source
    # reproduces the kernel itself, unflipped, anchored at (xi - cx, yi - cy). An asymmetric kernel catches a
    # transposed or flipped kernel index.
    rng = np.random.default_rng(3)
    ker = rng.random((4, 5))
    img = np.zeros((20, 30))
    yi, xi = 9, 13
    img[yi, xi] = 2.5
    out = O.convolve(img, ker)
    cy, cx = (4 - 1) // 2, (5 - 1) // 2
    expect = np.zeros_like(img)
    expect[yi - cy:yi - cy + 4, xi - cx:xi - cx + 5] = 2.5 * ker
    assert np.array_equal(out, expect)


def test_convolution_matches_library_full_convolution():
    # out[yo, xo] = sum over in-frame inputs of img[yo + cy - yk, xo + cx - xk] ker[yk, xk]: the 'full'
    # discrete convolution cropped at offset (cy, cx).
    rng = np.random.default_rng(4)
    img = rng.random((17, 23))
    for shape in ((6, 6), (3, 5), (1, 1), (7, 4)):
        ker = rng.random(shape)
        full = signal.convolve2d(img, ker, mode="full")
        cy, cx = (shape[0] - 1) // 2, (shape[1] - 1) // 2
        ref = full[cy:cy + img.shape[0], cx:cx + img.shape[1]]
        assert np.allclose(O.convolve(img, ker), ref, rtol=1e-13, atol=1e-15)


def test_convolution_conserves_flux_away_from_edges():
    rng = np.random.default_rng(5)
    img = np.zeros((40, 40))
    img[10:30, 10:30] = rng.random((20, 20))
    ker = O.gaussian_kernel(3.0, 6)
    assert abs(O.convolve(img, ker).sum() - img.sum()) < 1e-12 * img.sum()


def _chi2(ref, frame, x):
    m = ref != 0
    t = x * frame[m]
    return float(np.sum((ref[m] - t) ** 2 / (np.abs(ref[m]) + t)))


def test_golden_section_finds_the_chi2_minimum():
    # chi2(x) = sum (r - x s)^2 / (|r| + x s) is convex in x for s >= 0, so the golden-section bracket
    # always holds the minimiser that a library bounded minimisation finds. GoldenSection.cpp:116-118 ends
    # the search when the same trial point wins twice in a row after iteration 20, and the bracket has
    # then shrunk at least 21 times by the golden ratio: that bounds the distance to the minimiser.
    rng = np.random.default_rng(6)
    frame = rng.random((32, 48)) * 3.0
    truth = 170.0
    ref = rng.poisson(frame * truth).astype(np.float64)
    ref[0:3, :] = 0.0   # masked rows
    lum, chi2, fr = O.golden_section(ref, frame, 1.0, 1000.0)
    res = optimize.minimize_scalar(lambda x: _chi2(ref, frame, x), bounds=(1.0, 1000.0), method="bounded",
                                   options=dict(xatol=1e-10))
    assert abs(lum - res.x) <= 999.0 * 0.6180339887 ** 21
    assert _chi2(ref, frame, lum) >= res.fun * (1 - 1e-12)
    assert abs(chi2 - _chi2(ref, frame, lum)) < 1e-9 * chi2
    assert abs(lum - truth) < 0.02 * truth
    # GoldenSection::function zeroes the masked pixels of the frame it returns
    assert np.all(fr[0:3, :] == 0.0) and np.array_equal(fr[3:, :], frame[3:, :])


def test_golden_section_at_the_bounds():
    # a minimum outside the interval drives the bracket to the nearer end
    rng = np.random.default_rng(7)
    frame = rng.random(500) + 0.5
    ref = frame * 50.0
    lum, _, _ = O.golden_section(ref, frame, 100.0, 200.0)
    assert abs(lum - 100.0) < 1e-5
    lum, _, _ = O.golden_section(ref, frame, 1.0, 10.0)
    assert abs(lum - 10.0) < 1e-6


def test_voxel_field_tree_masses_exact():
    """sg_tree_subdivide_field / sg_grid_field_density (setup helpers): the mass in a box of a voxel field
    equals the brute-force sum over voxels of density x overlap volume; checked through the subdivision of
    a tiny field, whose node masses decide every split."""
    import ctypes
    from skirt7_amd import capi
    rng = np.random.default_rng(8)
    n = 6
    field = rng.random((n, n, n)) ** 4
    ext = np.array([-1.0, -2.0, 0.0, 2.0, 1.0, 3.0])
    w = (ext[3:] - ext[:3]) / n

    def brute(q):
        m = 0.0
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    lo = ext[:3] + w * np.array([i, j, k])
                    ov = np.clip(np.minimum(q[3:], lo + w) - np.maximum(q[:3], lo), 0, None)
                    m += field[i, j, k] * np.prod(ov)
        return m

    total = brute(ext)
    for maxmf in (0.2, 0.05, 0.01):
        cap = 4096
        flags = np.zeros(cap, np.uint8)
        nn = ctypes.c_int64()
        assert capi.lib().sg_tree_subdivide_field(2, capi.dptr(ext), capi.dptr(np.ascontiguousarray(field)), n, n, n,
                                                  1, 12, maxmf, flags.ctypes.data_as(capi._U8), cap,
                                                  ctypes.byref(nn)) == 0
        # rebuild the node boxes breadth-first and check every decision with the brute-force mass
        boxes = [ext.copy()]
        levels = [0]
        for l in range(nn.value):
            q, lev = boxes[l], levels[l]
            should = lev <= 1 or (lev < 12 and brute(q) / total >= maxmf)
            assert bool(flags[l]) == should, (l, lev)
            if flags[l]:
                a = lev % 3
                c = 0.5 * (q[a] + q[a + 3])
                lo, hi = q.copy(), q.copy()
                lo[a + 3] = c
                hi[a] = c
                boxes += [lo, hi]
                levels += [lev + 1, lev + 1]


def test_voronoi_cells_order_and_nearby_removal():
    """VoronoiMesh::buildMesh with removeNearby (VoronoiMesh.cpp:297-350): cells are the particles sorted on x;
    a particle within eps = 1e-12 |widths| of one later in that order is dropped (the later one stays), so a
    set with planted near-duplicates loses exactly those, and particles at least eps apart all stay."""
    from skirt7_amd import skirt as S
    ext = np.array([-1.0, -2.0, -3.0, 1.0, 2.0, 3.0])
    eps = 1e-12 * np.linalg.norm([2.0, 4.0, 6.0])
    rng = np.random.default_rng(3)
    p = rng.uniform(-1, 1, (500, 3)) * [1, 2, 3]
    dup = p[[5, 17, 300]] + np.array([[0.3, 0.0, 0.0], [0.0, 0.5, 0.0], [0.2, 0.2, 0.2]]) * eps
    far = p[[7]] + np.array([[0.0, 0.0, 1.5]]) * eps            # same x, farther than eps: stays
    allp = np.concatenate([p, dup, far])
    cells = S.voronoi_cells(ext, allp)
    assert cells.shape[0] == allp.shape[0] - 3
    assert np.all(np.diff(cells[:, 0]) >= 0)
    # every original particle position or its near twin is present once
    for q in p:
        d = np.linalg.norm(cells - q, axis=1)
        assert np.sum(d < eps) == 1
    assert np.any(np.all(cells == far[0], axis=1))
    # which of a pair stays: the later one in x order
    a = np.array([[0.1, 0.2, 0.3], [0.1 + 0.5 * eps, 0.2, 0.3]])
    kept = S.voronoi_cells(ext, a[::-1])
    assert kept.shape == (1, 3) and np.array_equal(kept[0], a[1])


def test_tree_criteria_optical_depth_and_combinations():
    """TreeDustGrid::subdivide's stopping criteria (TreeDustGrid.cpp:194-220) on a voxel field's exact node
    masses: a node below maxlevel is split when no criterion is set, or when its mass fraction reaches
    maxMassFraction or its optical depth kappaV M / V^(2/3) (kappaV = 2600) reaches maxOpticalDepth. Every
    decision is checked against brute-force masses, for the optical depth alone, both criteria together and
    none (every node split to maxlevel)."""
    import ctypes
    from skirt7_amd import capi
    rng = np.random.default_rng(9)
    n = 5
    field = 2e-3 * rng.random((n, n, n)) ** 3
    ext = np.array([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0])
    w = (ext[3:] - ext[:3]) / n
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    vlo = ext[:3] + w * idx
    fv = field.reshape(-1)

    def brute(q):
        ov = np.clip(np.minimum(q[3:], vlo + w) - np.maximum(q[:3], vlo), 0, None)
        return float(np.sum(fv * np.prod(ov, axis=1)))

    total = brute(ext)
    for maxmf, maxod, maxlevel in ((0.0, 0.05, 5), (0.02, 0.2, 5), (0.0, 0.0, 3)):
        cap = 1 << 20
        flags = np.zeros(cap, np.uint8)
        nn = ctypes.c_int64()
        assert capi.lib().sg_tree_subdivide_field_od(8, capi.dptr(ext), capi.dptr(np.ascontiguousarray(field)), n, n,
                                                     n, 1, maxlevel, maxmf, maxod, flags.ctypes.data_as(capi._U8),
                                                     cap, ctypes.byref(nn)) == 0
        boxes, levels = [ext.copy()], [0]
        nsplit_by_od = nstop = 0
        for l in range(nn.value):
            q, lev = boxes[l], levels[l]
            m = brute(q)
            vol = np.prod(q[3:] - q[:3])
            by_mf = maxmf > 0 and m / total >= maxmf
            by_od = maxod > 0 and 2600.0 * m / vol ** (2.0 / 3.0) >= maxod
            should = lev <= 1 or (lev < maxlevel and ((maxmf == 0 and maxod == 0) or by_mf or by_od))
            assert bool(flags[l]) == should, (l, lev, maxmf, maxod)
            nsplit_by_od += bool(flags[l]) and lev > 1 and by_od and not by_mf
            nstop += (not flags[l]) and 1 < lev < maxlevel
            if flags[l]:
                c = 0.5 * (q[:3] + q[3:])
                for k in range(8):
                    b = np.array([c[a] if (k >> a) & 1 else q[a] for a in range(3)] +
                                 [q[a + 3] if (k >> a) & 1 else c[a] for a in range(3)])
                    boxes.append(b)
                    levels.append(lev + 1)
        if maxod > 0:
            assert nsplit_by_od > 0 and nstop > 0
        if maxmf == 0 and maxod == 0:
            assert nn.value == sum(8 ** k for k in range(maxlevel + 1))


def test_tree_density_dispersion_criterion():
    """The density-dispersion criterion (TreeNodeSampleDensityCalculator::densityDispersion, (max - min) / max
    of the node's density samples) on an exponential disc: a node split by it alone has a true dispersion of
    at least the threshold (samples lie inside the node, so their spread cannot exceed the node's; the disc's
    extremes in a box are at the points nearest to and farthest from the centre in R and |z|), and with many
    samples most unsplit nodes below maxlevel have a true dispersion under it."""
    from skirt7_amd import skirt as S
    hR, hz = 5.0, 1.0
    disc = S.ExpDiskGeometry(hR, hz)
    ext = np.array([-10.0, -10.0, -2.0, 10.0, 10.0, 2.0])
    f = 0.6

    def true_disp(q):
        x = np.clip(0.0, q[0], q[3]); y = np.clip(0.0, q[1], q[4]); z = np.clip(0.0, q[2], q[5])
        rmin, zmin = np.hypot(x, y), abs(z)
        rmax = np.hypot(max(abs(q[0]), abs(q[3])), max(abs(q[1]), abs(q[4])))
        zmax = max(abs(q[2]), abs(q[5]))
        hi = np.exp(-rmin / hR - zmin / hz)
        lo = np.exp(-rmax / hR - zmax / hz)
        return (hi - lo) / hi

    import ctypes
    from skirt7_amd import capi
    cap = 1 << 20
    buf = np.zeros(cap, np.uint8)
    nn = ctypes.c_int64()
    assert capi.lib().sg_tree_subdivide_multi(8, capi.dptr(ext), 1, capi.dptr(capi.as_f64(disc.params())),
                                              capi.dptr(np.ones(1)), 1.0, 1, 5, 0.0, 0.0, f, 400, 11,
                                              buf.ctypes.data_as(capi._U8), cap, ctypes.byref(nn)) == 0
    flags = buf[:nn.value]
    boxes, levels = [ext.copy()], [0]
    unsplit_ok = unsplit = 0
    for l in range(len(flags)):
        q, lev = boxes[l], levels[l]
        if 1 < lev < 5:
            if flags[l]:
                assert true_disp(q) >= f - 1e-12, (l, lev, true_disp(q))
            else:
                unsplit += 1
                unsplit_ok += true_disp(q) < f + 0.05
        if flags[l]:
            c = 0.5 * (q[:3] + q[3:])
            for k in range(8):
                boxes.append(np.array([c[a] if (k >> a) & 1 else q[a] for a in range(3)] +
                                      [q[a + 3] if (k >> a) & 1 else c[a] for a in range(3)]))
                levels.append(lev + 1)
    assert int(np.sum(flags)) > 10 and unsplit > 10
    assert unsplit_ok >= 0.9 * unsplit
