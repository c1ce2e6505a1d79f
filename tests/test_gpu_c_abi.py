"""The C-ABI from a plain C host (tests/c_abi/c_abi_smoke.c, built by __graft_entry__.build() with gcc
against include/skirt7_amd.h alone): its paths, cell lookups and error report equal the Python binding's
on the same grids and rays, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from skirt7_amd import skirt as S
from skirt7_amd.capi import SG_FP64, SG_NEIGHBOR

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "c_abi", "_build", "c_abi_smoke")

EXT = np.array([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0])
RK = np.array([[-0.9, -0.8, -0.7, 0.48, 0.6, 0.64],
               [-3.0, 0.1, 0.2, 1.0, 0.0, 0.0],
               [0.3, -0.2, 0.9, -0.6, 0.0, -0.8],
               [2.0, 2.0, 2.0, 1.0, 0.0, 0.0]])


def c_output():
    assert os.path.exists(BIN), "the C caller is not built: run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.splitlines()


def parse_paths(lines, name):
    out = {}
    for ln in lines:
        f = ln.split()
        if f[:2] == ["path", name]:
            segs = [s.split(":") for s in f[4:]]
            out[int(f[2])] = (int(f[3]), np.array([int(a) for a, _ in segs]), np.array([float(b) for _, b in segs]))
    return out


def test_c_host_matches_python_binding():
    lines = c_output()
    print("\n".join(lines[:3]), flush=True)
    cart = S.CartesianDustGrid(EXT, [-1.0 + 0.5 * i for i in range(5)], [-1.0 + 2.0 * i / 3.0 for i in range(4)],
                               [-1.0 + 0.4 * i for i in range(6)])
    tree = S.BinTreeDustGrid(EXT, np.array([1, 1, 0, 0, 0], np.uint8), SG_NEIGHBOR)
    assert "cells cartesian %d" % cart.numCells() in lines
    assert "cells tree %d" % tree.numCells() in lines
    for name, g in (("cartesian", cart), ("tree", tree)):
        got = parse_paths(lines, name)
        counts, m, ds = g.paths(RK, 256, SG_FP64)
        assert sorted(got) == [0, 1, 2, 3]
        for i in range(len(RK)):
            n, cm, cds = got[i]
            assert n == counts[i]
            assert np.array_equal(cm, m[i, :n]) and np.array_equal(cds, ds[i, :n]), (name, i)
    assert counts[1] > 0 and counts[3] == 0
    wc = cart.whichcell(np.array([[0.1, 0.2, 0.3], [-0.99, 0.99, -0.5], [5.0, 0.0, 0.0]]))
    assert "whichcell cartesian %d %d %d" % tuple(wc) in lines
    bad = [ln for ln in lines if ln.startswith("invalid status")]
    assert len(bad) == 1 and int(bad[0].split()[2]) < 0 and "bin" in bad[0]
    assert lines[-1] == "done"
