"""Serving loop (cturbobfm_amd/serve.py): cases that arrive while others are resident
are admitted to the running batch, advance together with them and retire on their own
density-residual drop.  The bar: every served result equals the same case solved alone
on the plain path, bit for bit (final state, iteration of retirement), and admission
respects the device-memory budget.
"""
import numpy as np
import pytest

from cturbobfm_amd import serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lut_path(tmp_path_factory):
    return serve.stream_table(str(tmp_path_factory.mktemp("lut")))


def small_stream(lut_path=None):
    specs = [serve.bump_spec(33, 17, 0.5, "fp64", max_iter=1200),
             serve.rotor_spec(25, 9, 17, 95000.0, "fp64", max_iter=1200),
             serve.bump_spec(33, 17, 0.45, "mixed", max_iter=1200),
             serve.rotor_spec(33, 9, 17, 93000.0, "mixed", max_iter=1200),
             serve.bump_spec(49, 25, 0.55, "fp64", max_iter=1200),
             serve.rotor_spec(25, 9, 17, 97000.0, "mixed", max_iter=1200)]
    if lut_path:
        # look-up-table fluid requests in the same stream
        specs.insert(2, serve.rotor_spec(25, 9, 17, 95000.0, "fp64", max_iter=1200, lut_path=lut_path))
        specs.append(serve.bump_spec(33, 17, 0.5, "fp64", max_iter=1200, lut_path=lut_path))
    return specs


def test_served_equals_alone(lut_path):
    specs = small_stream(lut_path)
    arrivals = serve.poisson_arrivals(len(specs), rate=50.0, seed=3)
    srv = serve.Server(drop=1.5, max_iter=1200, chunk=100, keep_states=True)
    served = srv.run(specs, arrivals)
    alone = serve.sequential(specs, drop=1.5, max_iter=1200, chunk=100, keep_states=True)
    for a, b in zip(served, alone):
        print(a.name, a.iterations, a.converged, b.iterations, b.converged, flush=True)
        assert not a.diverged and not b.diverged
        assert a.iterations == b.iterations and a.converged == b.converged
        assert np.array_equal(a.state, b.state), "served result of %s differs from the case alone" % a.name
        assert a.admitted >= a.arrival - 1e-9 and a.finished > a.admitted
    assert any(r.converged for r in served)


def test_admission_respects_budget():
    specs = small_stream()[:4]
    arrivals = np.zeros(len(specs))          # a burst: everything queued at once
    # a budget that holds only the first case at a time
    srv = serve.Server(drop=1.5, max_iter=300, chunk=100, budget=1)
    res = srv.run(specs, arrivals)
    order = sorted(res, key=lambda r: r.admitted)
    for prev, cur in zip(order, order[1:]):
        assert cur.admitted >= prev.finished - 1e-9, "a case was admitted before the previous one left"
