"""The oracle's SPH dust medium and its tree subdivision (oracle/sph.cpp) pinned on the CPU.

- SPHGasParticle's sampled erf (SPHGasParticle.cpp:19-44): the table holds math.erf at its 5001 points, a
  query takes the point below it (truncation), the function is odd and saturates at +-1 beyond 2.5.
- metalMassInBox (the Gaussian approximation, s = 2.42 / h): a box reaching past every particle's support
  holds the total dust mass; a half space through a particle's centre holds half of its mass; a slab at
  distance a from the centre holds (1 - erf(2.42 a / h)) / 2 of it, to the table's step.
- metalDensity (the cubic spline): its integral over space is the metal mass, its peak 8/pi M Z / h^3, and
  it is continuous at u = 1/2.
- SPHGasParticleGrid: the borders are infinite at both ends, increase, and split the particle centres into
  slabs of about n / 20 each; every cell's list is, in particle order, the set of particles whose sphere
  reaches into the cell (nearest point of the cell's box closer than h, computed here by clamping).
- TreeDustGrid's subdivision with the box calculator: every node rebuilt from the flags alone (children
  as OctTreeNode / BinTreeNode create them), each evaluated node's mass equal to massInBox of its box,
  and each decision equal to the criteria applied to that mass.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from tests import oracle_lib as O


def particles(n, seed=7, k=32, clustered=True):
    rng = np.random.default_rng(seed)
    if clustered:
        centres = rng.uniform(-8, 8, (12, 3))
        pos = centres[rng.integers(0, 12, n)] + rng.normal(0, 1.5, (n, 3))
        pos = np.clip(pos, -9.9, 9.9)
    else:
        pos = rng.uniform(-10, 10, (n, 3))
    h = cKDTree(pos).query(pos, k + 1)[0][:, -1]
    M = rng.uniform(0.5, 1.5, n)
    Z = rng.uniform(0.01, 0.03, n)
    return np.column_stack([pos, h, M, Z, np.zeros(n)])


def test_erf_table():
    for i in (0, 1, 17, 2000, 4999, 5000):
        x = i * 2.5 / 5000
        assert O.myerf(x) == (1.0 if x >= 2.5 else math.erf(x))
    rng = np.random.default_rng(1)
    for x in rng.uniform(0, 2.5, 200):
        below = math.floor(x * 2000) / 2000
        assert abs(O.myerf(x) - math.erf(below)) < 1e-15
        assert O.myerf(-x) == -O.myerf(x)
    assert O.myerf(2.5) == 1.0 and O.myerf(7.0) == 1.0 and O.myerf(-2.5) == -1.0 and O.myerf(-9.0) == -1.0


def test_total_and_half_masses():
    rows = particles(3000)
    md = O.SphMedium(rows, fdust=0.3)
    total = 0.3 * float(np.sum(rows[:, 4] * rows[:, 5]))
    assert abs(md.mass() / total - 1) < 1e-13
    reach = 1.04 * rows[:, 3].max()
    lo, hi = rows[:, :3].min(axis=0) - reach, rows[:, :3].max(axis=0) + reach
    assert abs(md.mass_in_box(np.concatenate([lo, hi]))[0] / total - 1) < 1e-13
    # a single particle: half spaces through its centre, and slabs at distance a
    one = np.array([[0.3, -0.2, 0.1, 0.8, 2.0, 0.02, 0.0]])
    m1 = O.SphMedium(one, fdust=1.0)
    mz = 2.0 * 0.02
    big = 100.0
    for ax in range(3):
        b = np.array([-big, -big, -big, big, big, big])
        b[ax] = one[0, ax]
        assert m1.mass_in_box(b)[0] == mz / 2
        b = np.array([-big, -big, -big, big, big, big])
        b[ax + 3] = one[0, ax]
        assert m1.mass_in_box(b)[0] == mz / 2
    for a in (0.05, 0.2, 0.4, 0.7):
        b = np.array([-big, -big, -big, big, big, big])
        b[0] = one[0, 0] + a
        frac = m1.mass_in_box(b)[0] / mz
        assert abs(frac - 0.5 * (1 - math.erf(2.42 * a / 0.8))) < 3e-4, a


def test_spline_density():
    h, M, Z = 0.7, 3.0, 0.02
    one = np.array([[0.0, 0.0, 0.0, h, M, Z, 0.0]])
    md = O.SphMedium(one, fdust=0.5)
    peak = md.density(np.zeros((1, 3)))[0]
    assert abs(peak / (0.5 * 8 / math.pi * M * Z / h ** 3) - 1) < 1e-14
    # continuity at u = 1/2: both branches give rho0 / 4
    r = np.array([[0.5 * h, 0, 0], [0.5 * h * (1 - 1e-12), 0, 0]])
    d = md.density(r)
    assert abs(d[0] / (peak / 4) - 1) < 1e-10 and abs(d[1] / (peak / 4) - 1) < 1e-10
    assert md.density(np.array([[h * 1.0001, 0, 0]]))[0] == 0.0
    # midpoint rule over the sphere's cube
    n = 120
    c = (np.arange(n) + 0.5) / n * 2 * h - h
    X, Y, Zc = np.meshgrid(c, c, c, indexing="ij")
    vals = md.density(np.stack([X.ravel(), Y.ravel(), Zc.ravel()], 1))
    integral = vals.sum() * (2 * h / n) ** 3
    assert abs(integral / (0.5 * M * Z) - 1) < 2e-3


def test_particle_grid_and_lists():
    rows = particles(8000, seed=3)
    md = O.SphMedium(rows)
    g = md.grid()
    n = rows.shape[0]
    for a in range(3):
        assert g[a, 0] == -np.inf and g[a, 20] == np.inf
        assert np.all(np.diff(g[a, 1:20]) > 0)
        cnt = np.histogram(rows[:, a], bins=g[a])[0]
        # each slab holds about n / 20 centres: the border lies at the first bin (100 per cell) past it
        binw = (np.max(rows[:, a] + rows[:, 3]) - np.min(rows[:, a] - rows[:, 3])) / 2000
        width_counts = np.histogram(rows[:, a], bins=2000)[0].max()
        assert np.all(np.abs(cnt[:-1] - n // 20) <= 2 * width_counts + 2), (a, cnt)
        assert binw > 0
    pos, h = rows[:, :3], rows[:, 3]
    rng = np.random.default_rng(5)
    for (i, j, k) in [(0, 0, 0), (19, 19, 19), (9, 10, 11)] + [tuple(rng.integers(0, 20, 3)) for _ in range(12)]:
        lo = np.array([g[0, i], g[1, j], g[2, k]])
        hi = np.array([g[0, i + 1], g[1, j + 1], g[2, k + 1]])
        near = np.clip(pos, lo, hi)
        d2 = np.sum((pos - near) ** 2, axis=1)
        expect = np.flatnonzero(d2 < h * h)
        got = md.cell_list(i, j, k)
        assert np.array_equal(got, expect), (i, j, k)


def _tree_nodes(flags, arity, extent):
    """Boxes and levels of every node, rebuilt from the flags in creation order."""
    boxes = [np.array(extent, float)]
    levels = [0]
    for l, f in enumerate(flags):
        if not f:
            continue
        q, lev = boxes[l], levels[l]
        if arity == 8:
            c = 0.5 * (q[:3] + q[3:])
            for k in range(8):
                up = np.array([(k >> a) & 1 for a in range(3)], bool)
                boxes.append(np.concatenate([np.where(up, c, q[:3]), np.where(up, q[3:], c)]))
                levels.append(lev + 1)
        else:
            a = lev % 3
            c = 0.5 * (q[a] + q[a + 3])
            lo, hi = q.copy(), q.copy()
            lo[a + 3] = c
            hi[a] = c
            boxes += [lo, hi]
            levels += [lev + 1, lev + 1]
    assert len(boxes) == len(flags)
    return np.array(boxes), np.array(levels)


@pytest.mark.parametrize("arity,mmf,mod", [(8, 2e-3, 0.0), (2, 1e-3, 0.0), (8, 0.0, 60.0), (2, 2e-3, 60.0)])
def test_subdivision_decisions(arity, mmf, mod):
    rows = particles(4000, seed=11)
    md = O.SphMedium(rows)
    ext = np.array([-10.0, -10, -10, 10, 10, 10])
    minlevel, maxlevel = (1, 6) if arity == 8 else (3, 16)
    flags, masses = md.subdivide(arity, ext, minlevel, maxlevel, mmf, mod)
    boxes, levels = _tree_nodes(flags, arity, ext)
    evaluated = (levels > minlevel) & (levels < maxlevel)
    assert np.all(masses[~evaluated] == -1)
    ref = md.mass_in_box(boxes[evaluated])
    assert np.array_equal(masses[evaluated], ref)
    vol = np.prod(boxes[:, 3:] - boxes[:, :3], axis=1)
    total = md.mass()
    need = np.zeros(len(flags), bool)
    if mmf > 0:
        need |= masses / total >= mmf
    if mod > 0:
        need |= 2600.0 * masses / vol ** (2.0 / 3.0) >= mod
    expect = (levels <= minlevel) | (evaluated & need)
    assert np.array_equal(flags.astype(bool), expect)
    assert flags.sum() > 20 and (levels == maxlevel).sum() + (~flags.astype(bool) & evaluated).sum() > 50
