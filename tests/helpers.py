"""Test inputs shared by the CPU and GPU tests: random trees, rays and Voronoi neighbour sets."""
import numpy as np

EXT = np.array([-20.0, -20.0, -20.0, 20.0, 20.0, 20.0])


def random_tree_flags(arity, maxlevel, p, seed, minlevel=1):
    """Subdivision flags in breadth-first node order for a random tree."""
    rng = np.random.default_rng(seed)
    flags, levels, i = [], [0], 0
    while i < len(levels):
        lev = levels[i]
        split = lev < minlevel or (lev < maxlevel and rng.random() < p)
        flags.append(1 if split else 0)
        if split:
            levels.extend([lev + 1] * arity)
        i += 1
    return np.array(flags, np.uint8)


def random_rays(n, seed, ext=EXT, outside=0.3, special=True):
    """Rays (r, k) with starts inside and outside the box; includes axis-aligned, diagonal and
    grid-plane-aligned cases when special is set."""
    rng = np.random.default_rng(seed)
    lo, hi = ext[:3], ext[3:]
    w = hi - lo
    r = lo + rng.random((n, 3)) * w
    nout = int(n * outside)
    r[:nout] = lo - 0.5 * w + rng.random((nout, 3)) * 2.0 * w
    k = rng.normal(size=(n, 3))
    k /= np.linalg.norm(k, axis=1)[:, None]
    if special:
        axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
        diag = np.array([[1, 1, 1], [-1, 1, -1], [1, -1, 0], [0, 1, 1]], float)
        diag /= np.linalg.norm(diag, axis=1)[:, None]
        spec = np.concatenate([axes, diag])
        m = min(len(spec) * 4, n // 4)
        k[nout:nout + m] = spec[np.arange(m) % len(spec)]
        # starts exactly on grid planes (multiples of 2.5 kpc) and at the centre
        r[nout + m:nout + 2 * m] = np.round(r[nout + m:nout + 2 * m] / 2.5) * 2.5
        r[nout + 2 * m] = 0.0
        # direction components too small for the 1e-15 guard
        k[nout + 2 * m + 1] = np.array([1e-16, 0.6, 0.8])
    return np.ascontiguousarray(np.concatenate([r, k], axis=1))


def voronoi_superset(sites):
    """Neighbour lists naming every other site and all six walls. The walk takes the nearest
    bisector plane ahead, which is always a true Voronoi neighbour, so a superset of the true lists
    gives the same path."""
    n = len(sites)
    ids = []
    off = [0]
    for i in range(n):
        lst = [j for j in range(n) if j != i] + [-1, -2, -3, -4, -5, -6]
        ids.extend(lst)
        off.append(len(ids))
    return np.array(off, np.int64), np.array(ids, np.int32)
