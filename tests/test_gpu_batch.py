"""Batched advance (tbfm_gpu_batch_*): every kernel of an iteration is one launch over
all members and the iteration is replayed as a captured HIP graph.  Each member keeps
its own buffers, counters, reductions and convergence test, so the bar here is bitwise
equality: a case advanced in any mix equals the same case advanced alone (states and
residual histories), and the solo graph path equals the solo eager path.
"""
import numpy as np
import pytest

from cturbobfm_amd import Batch, Solver, SolverError, cases

pytestmark = pytest.mark.gpu


def rotor(scheme="JST", **kw):
    return cases.rotor3d(33, 9, 17, n_iterations=400, scheme=scheme, **kw)


def bump(**kw):
    return cases.bump2d(49, 25, n_iterations=400, **kw)


MIX = {
    "rotor_jst": lambda: rotor(),
    "rotor_roe_muscl_mixed": lambda: rotor("ROE", muscl=True, limiter="van albada", cfl=1.2),
    "bump_jst": lambda: bump(),
    "bump_roe": lambda: bump(scheme="ROE", cfl=1.0),
}
MIX_OVERRIDES = {"rotor_roe_muscl_mixed": {"GPU_PRECISION": "mixed"}}


def make(name):
    return Solver.from_case(MIX[name](), MIX_OVERRIDES.get(name))


def test_alone_equals_mix():
    n = 60
    alone = {}
    for name in MIX:
        print("alone", name, flush=True)
        s = make(name)
        s.advance(n)
        alone[name] = (s.state(), s.history.copy())
        s.close()
    b = Batch()
    members = {name: make(name) for name in MIX}
    for s in members.values():
        b.admit(s)
    assert len(b) == len(MIX)
    print("batch of", len(b), flush=True)
    b.advance(n // 2)
    b.advance(n - n // 2)
    for name, s in members.items():
        it, conv, stopped, lr = s.status()
        assert it == n and conv == 0 and not stopped
        U = s.state()
        assert np.array_equal(U, alone[name][0]), "state of %s differs in the mix" % name
        assert np.array_equal(lr, alone[name][1][-1]), "last residual of %s differs" % name
    b.close()


def test_retire_and_readmit():
    n = 30
    ref = make("rotor_jst")
    ref.advance(2 * n)
    U_ref = ref.state()
    a, c = make("rotor_jst"), make("bump_jst")
    b = Batch()
    b.admit(a)
    b.admit(c)
    b.advance(n)
    b.retire(c)
    b.advance(n // 2)
    b.admit(c)
    b.advance(n - n // 2)
    assert a.status()[0] == 2 * n
    assert c.status()[0] == n + n - n // 2
    assert np.array_equal(a.state(), U_ref)
    # a retired member advances alone again, continuing its own counters
    b.retire(a)
    a.advance(5)
    ref.advance(5)
    assert np.array_equal(a.state(), ref.state())
    b.close()


def test_member_convergence_stops_only_that_member():
    drop = 0.5
    alone = make("bump_jst")
    alone.set_convergence(True, drop)     # half an order of density drop
    alone.solve(4000, chunk=500)
    conv = alone.converged_at
    assert conv > 0, "the bump did not drop %g orders in 4000 iterations" % drop
    fast = make("bump_jst")
    fast.set_convergence(True, drop)
    slow = make("rotor_jst")
    b = Batch()
    b.admit(fast)
    b.admit(slow)
    steps = 0
    while steps < conv + 50:
        b.advance(50)
        steps += 50
    it_f, conv_f, stopped_f, _ = fast.status()
    it_s, conv_s, stopped_s, _ = slow.status()
    assert stopped_f and conv_f == conv and it_f == conv
    assert not stopped_s and it_s == steps
    assert np.array_equal(fast.state(), alone.state())
    b.close()


def test_graph_equals_eager():
    g = make("rotor_jst")
    e = make("rotor_jst")
    e.profile(True, True)          # profiling launches eagerly with events
    g.advance(25)
    e.advance(25)
    assert np.array_equal(g.state(), e.state())
    assert np.array_equal(g.history, e.history)


def test_rk_scheme_mismatch_rejected():
    a = make("bump_jst")
    c = Solver.from_case(bump(time_integration=1))
    b = Batch()
    b.admit(a)
    with pytest.raises(SolverError, match="same RK scheme"):
        b.admit(c)
    b.close()
