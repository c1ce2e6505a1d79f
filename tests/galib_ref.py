"""The reference's GAlib as a checker: oracle/_ref/ga_ref (built by oracle/ref/Makefile from the reference's
GAlib sources and oracle/ref/ga_driver.cpp) runs one FitSKIRT-configured steady-state GA per process and asks
for the scores of each population evaluation over a pipe. Values travel as C99 hexadecimal floats, so both
sides see the same doubles."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GA_REF = os.path.join(ROOT, "oracle", "_ref", "ga_ref")


def available():
    return os.path.exists(GA_REF)


class GaRef:
    def __init__(self, lower, upper, popsize, generations, pmut, pcross, seed):
        self.L = len(lower)
        self.popsize = popsize
        self.p = subprocess.Popen([GA_REF], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
        cfg = "%d %d %d %d\n%s %s\n" % (self.L, popsize, generations, seed, float(pmut).hex(), float(pcross).hex())
        cfg += "".join("%s %s\n" % (float(a).hex(), float(b).hex()) for a, b in zip(lower, upper))
        self.p.stdin.write(cfg)
        self.p.stdin.flush()

    def _line(self):
        line = self.p.stdout.readline()
        if not line:
            raise RuntimeError("ga_ref ended early (exit %s)" % self.p.poll())
        return line.split()

    def request(self):
        """('EVAL', generation, genes (n x L)) or ('DONE', generation, (genes, scores), (best genes, best score))."""
        head = self._line()
        if head[0] == "EVAL":
            n, gen = int(head[1]), int(head[2])
            genes = np.array([[float.fromhex(x) for x in self._line()] for _ in range(n)]).reshape(n, self.L)
            return "EVAL", gen, genes
        assert head[0] == "DONE", head
        rows = np.array([[float.fromhex(x) for x in self._line()] for _ in range(self.popsize)])
        best = [float.fromhex(x) for x in self._line()[1:]]
        self.p.wait(timeout=30)
        return "DONE", int(head[1]), (rows[:, :self.L], rows[:, self.L]), (np.array(best[:self.L]), best[self.L])

    def scores(self, values):
        self.p.stdin.write("".join(float(v).hex() + "\n" for v in values))
        self.p.stdin.flush()

    def close(self):
        if self.p.poll() is None:
            self.p.kill()
            self.p.wait()
