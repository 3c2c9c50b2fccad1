"""The reference's own build-to-build spread, next to the 1e-11 parity bar.

The reference's Makefile builds with `-O3 -march=native` and lets the compiler fuse
a*b + c into FMA instructions (Makefile:21; g++ does so in -std=c++20 mode), while the
oracle and the GPU path are built with `-ffp-contract=off`, the reference's arithmetic
operation for operation.  This test compiles the oracle's source a second time with the
Makefile's flags and FMA contraction on, runs the same cases with both builds and
reports how far a legitimate rebuild of the same code moves the results: the 2-D bump
129 x 65 (JST, CFL 2) for 2000 iterations and the 25 x 9 x 17 rotor (Hall, JST) for 200.
The two builds differ (the flags change the arithmetic) but stay far inside the 1e-11
bar that the GPU path is held to against the oracle; on MI355X the GPU path's distance
from the oracle on the bump workload (2.0e-14 history, 2.4e-14 fields after 2000
iterations) is of the same size as this spread between two CPU builds."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from cturbobfm_amd import cases

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

CASES = {
    "bump2d_129x65": (lambda: cases.bump2d(129, 65, n_iterations=2000), 2000),
    "rotor3d_25x9x17": (lambda: cases.rotor3d(25, 9, 17, n_iterations=200), 200),
}

RUNNER = r"""
import sys, numpy as np
sys.path.insert(0, %(root)r); sys.path.insert(0, %(oracle)r)
import oracle as O
from cturbobfm_amd import cases
mk = {"bump2d_129x65": lambda: cases.bump2d(129, 65, n_iterations=2000),
      "rotor3d_25x9x17": lambda: cases.rotor3d(25, 9, 17, n_iterations=200)}
name, n, out = sys.argv[1], int(sys.argv[2]), sys.argv[3]
oc = O.OracleCase(mk[name]())
lr, conv, _ = oc.run(n, threads=4)
np.savez(out, lr=lr, U=oc.state())
"""


@pytest.fixture(scope="module")
def fma_build(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc absent")
    d = tmp_path_factory.mktemp("fma")
    so = str(d / "liboracle_fma.so")
    cmd = [gcc, "-std=c11", "-O3", "-march=native", "-ffp-contract=fast", "-fopenmp", "-fPIC", "-shared",
           "-o", so, os.path.join(ROOT, "oracle", "tbfm_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True)
    return so, str(d)


@pytest.mark.parametrize("name", list(CASES))
def test_fma_build_spread(name, fma_build):
    so, d = fma_build
    mk, n = CASES[name]
    script = os.path.join(d, "run.py")
    with open(script, "w") as f:
        f.write(RUNNER % {"root": ROOT, "oracle": os.path.join(ROOT, "oracle")})
    out = os.path.join(d, name + ".npz")
    env = dict(os.environ, TBFM_ORACLE_LIB=so)
    subprocess.run([sys.executable, script, name, str(n), out], check=True, env=env)
    g = np.load(out, allow_pickle=False)
    oc = O.OracleCase(mk())
    lr, conv, _ = oc.run(n, threads=4)
    U = oc.state()
    hist = float(np.max(np.abs(g["lr"] - lr)))
    field = max(float(np.max(np.abs(g["U"][v] - U[v])) / np.max(np.abs(U[v]))) for v in range(5))
    print("%s: FMA build against the -ffp-contract=off build after %d iterations: log-residual history "
          "%.3e (absolute), fields %.3e (relative)" % (name, n, hist, field))
    # the flags change the arithmetic; the spread stays inside the parity bar
    assert hist > 0.0 or field > 0.0
    assert hist <= 1e-11 and field <= 1e-11
