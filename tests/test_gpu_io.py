"""The reference's files on the GPU path.

- `turbobfm input.ini` (cturbobfm_amd/turbobfm) run in a case folder writes
  Volume_CSV/results.csv and residuals.csv like the reference; the reference's own
  regression check (test/regression_tests/test_cases.py:45-88: read the structured CSV,
  relative L2 of Density, Pressure and Temperature against reference.csv below 3e-2)
  passes on every regression case this path covers.
- tbfm_gpu_write_solution of a device case equals the host writer (pinned byte-exact
  to the reference's Output in test_io.py) on the same state and side outputs.
- A case created with RESTART_SOLUTION = yes starts from exactly the restart state and
  body forces of the host reader, and continues like the CPU restatement from it.
"""
import os
import subprocess

import numpy as np
import pytest

import cturbobfm_amd as T
from cturbobfm_amd import Solver, cases

import oracle as O
import regression_cases as RC

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cturbobfm_amd", "turbobfm")
GOLD = os.path.join(ROOT, "tests", "golden")


def read_structured_csv(path):
    """The reference harness's reader: NI/NJ/NK lines, then a CSV with a header."""
    with open(path) as f:
        dims = [int(f.readline().strip().split("=")[1]) for _ in range(3)]
        header = f.readline().strip().split(",")
    data = np.loadtxt(path, delimiter=",", skiprows=4, ndmin=2)
    assert data.shape[0] == dims[0] * dims[1] * dims[2]
    return {h: data[:, c].reshape(dims) for c, h in enumerate(header)}


@pytest.mark.parametrize("name", list(RC.CONFIGS))
def test_cli_regression_case(name, tmp_path):
    print("case", name, flush=True)
    folder = str(tmp_path / name)
    RC.case_files(name, folder)
    r = subprocess.run([CLI, "input.ini"], cwd=folder, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert ("Convergence reached at iteration" in r.stdout) == RC.converges(name)
    res = read_structured_csv(os.path.join(folder, "Volume_CSV", "results.csv"))
    ref = RC.load_reference_solution(name)
    for key in ("Density", "Pressure", "Temperature"):
        err = RC.rel_l2(ref[key], res[key])
        print("%s %s %.3e" % (name, key, err), flush=True)
        assert err < 3e-2
    with open(os.path.join(folder, "residuals.csv")) as f:
        lines = f.read().splitlines()
    assert lines[0] == "Rho,RhoU,RhoV,RhoW,RhoE"
    if RC.converges(name):
        conv = int(r.stdout.split("Convergence reached at iteration")[1].split()[0])
    else:
        conv = RC.max_iterations(name)
    assert len(lines) == conv + 1
    if RC.CONFIGS[name].get("TURBO_OUTPUT") == "yes":
        with open(os.path.join(folder, "turbo.csv"), encoding="utf-8") as f:
            tl = f.read().splitlines()
        assert tl[0] == "Time[\u03bcs],Massflow[kg/s],PRtt,TRtt,ETAtt"
        assert len(tl) == conv + 1
        last = [float(v) for v in tl[-1].split(",")]
        print("%s turbo last row %s" % (name, last), flush=True)
        assert last[1] > 0 and 0.5 < last[2] < 3.0


def test_write_solution_matches_host_writer(tmp_path):
    case = cases.rotor3d(17, 6, 9, bfm_model="Hall-Thollet")
    case["config"].update({"OUTPUT_FIELDS_TYPE": "turbo_bfm", "SOLUTION_NAME": "results", "SAVE_UNSTEADY": "no"})
    s = Solver.from_case(case)
    s.advance(30)
    a = str(tmp_path / "device.csv")
    s.write_solution(a)
    fi, fv, dv, _ = s.side_outputs()
    ini = cases.write_case(case, str(tmp_path / "case"))
    b = str(tmp_path / "host.csv")
    T.write_solution_from_ini(ini, s.state(), b, fi, fv, dv)
    with open(a, "rb") as fa, open(b, "rb") as fb:
        assert fa.read() == fb.read()


def test_restart_state_and_continuation(tmp_path):
    g = np.load(os.path.join(GOLD, "volume_rotor3d_turbo_bfm.npz"), allow_pickle=False)
    path = str(tmp_path / "restart.csv")
    with open(path, "wb") as f:
        f.write(g["text"].tobytes())
    case = cases.rotor3d(9, 4, 5, bfm_model="Hall-Thollet")
    case["config"].update({"RESTART_SOLUTION": "yes", "RESTART_SOLUTION_FILEPATH": path})
    ini = cases.write_case(case, str(tmp_path / "case"))
    U0, fi0, fv0 = T.initial_state_from_ini(ini)
    s = Solver(ini)
    assert np.array_equal(s.state(), U0)
    fi, fv, _, _ = s.side_outputs()
    assert np.array_equal(fi, fi0) and np.array_equal(fv, fv0)
    oc = O.OracleCase(case)
    oc.set_state(U0)
    lo, _, _ = oc.run(20, threads=4)
    lg = s.advance(20)
    print("history err %.3e" % np.max(np.abs(lg - lo)), flush=True)
    assert np.max(np.abs(lg - lo)) <= 1e-11
    Ug, Uo = s.state(), oc.state()
    for v in range(5):
        assert np.max(np.abs(Ug[v] - Uo[v])) <= 1e-11 * max(np.max(np.abs(Uo[v])), 1e-300)


def test_turbo_history_matches_host_turbo():
    """Per-iteration device rows against the host evaluation of the same states (face-order
    sums) and the time against the running sum of the minimum time step."""
    case = cases.rotor3d(17, 6, 9, n_iterations=50, turbo_output=True)
    s = Solver.from_case(case)
    t = 0.0
    for it in range(1, 9):
        host = s.turbo()
        s.advance(1)
        dt = s.side_outputs()[3]
        t = t + dt.min()
        row = s.turbo_history(it, 1)[0]
        print(it, row, host[6:], flush=True)
        assert row[0] == t
        assert np.max(np.abs(row[1:4] - host[6:9]) / np.abs(host[6:9])) < 1e-12
        # the efficiency divides by TR - 1, which amplifies last-bit differences
        assert abs(row[4] - host[9]) <= 1e-8 * max(1.0, abs(host[9]))


def test_turbo_history_alone_equals_batch():
    from cturbobfm_amd import Batch
    case = cases.rotor3d(17, 6, 9, n_iterations=50, turbo_output=True)
    a = Solver.from_case(case)
    a.advance(20)
    b = Batch()
    m = Solver.from_case(case)
    other = Solver.from_case(cases.bump2d(33, 17))
    b.admit(m)
    b.admit(other)
    b.advance(20)
    m.status()
    assert np.array_equal(a.turbo_history(1, 20), m.turbo_history(1, 20))
    b.close()
