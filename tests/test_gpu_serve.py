"""The continuous-batching service: requests admitted while others run give the same results as
each served alone (bit for bit: regions never interact), and the stop criterion retires them.  The
requests span the serving workload's four depths (50 / 100 / 200 / 400 mm), with notches, and their
inputs are generated after they arrive."""
import numpy as np
import pytest

from bb_pd_amd import _lib
from bb_pd_amd.serve import Request, serve, slot_bytes

pytestmark = pytest.mark.gpu

HIST = ("reaction", "reaction_static", "deflection", "cmod", "disp", "cn", "inc_end")


def workload_requests():
    """One request per depth with the three notch ratios among them, staggered arrivals."""
    ps = [dict(depth=400, ratio=0.2, E=37e9, s0=1.0e-4), dict(depth=50, ratio=0.0, E=31e9, s0=0.9e-4),
          dict(depth=200, ratio=0.4, E=44e9, s0=1.15e-4), dict(depth=100, ratio=0.2, E=35e9, s0=0.85e-4),
          dict(depth=50, ratio=0.4, E=40e9, s0=1.1e-4)]
    return [Request(i, params=p, arrival=0.05 * i) for i, p in enumerate(ps)]


@pytest.mark.parametrize("precision", [_lib.MIXED, _lib.FP64])
def test_served_results_equal_alone(gpu, precision):
    reqs = workload_requests()
    print("case: serve", [r.params["depth"] for r in reqs], flush=True)
    # a budget that holds the 400 mm beam and one or two others at a time: later requests wait
    budget = 1.3 * (320000 + 80000) * slot_bytes(precision)
    kw = dict(chunk=50, ipi=50, increment_um=25.0, keep_history=True, keep_cases=True)
    mix = serve(reqs, precision, budget, **kw)
    assert mix["requests"] == len(reqs) and mix["rejected"] == 0
    assert {r.params["depth"] for r in mix["done"]} == {50, 100, 200, 400}
    waited = [r for r in mix["done"] if r.admitted - r.arrival > 0.01]
    assert waited, "no request waited for slots: the mix did not overlap"
    for r in mix["done"]:
        assert r.built >= r.arrival                      # inputs generated after arrival
        alone = serve([Request(0, case=r.case)], precision, budget, **kw)["done"][0]
        n = min(len(r.history["reaction_static"]), len(alone.history["reaction_static"]))
        assert n >= 50
        for k in HIST:
            a, b = r.history[k][:n], alone.history[k][:n]
            assert np.array_equal(np.asarray(a, np.float64).view(np.int64),
                                  np.asarray(b, np.float64).view(np.int64)), (r.name, k)
        assert r.peak == alone.peak and r.steps == alone.steps
        # the criterion fired: 0.5 mm of travel (20 increments of 25 um) or the load fell below 20% of its peak
        assert r.steps <= 21 * 50 + 50
        print(f"  {r.name}: {r.steps} steps, same alone and in the mix", flush=True)


def test_request_larger_than_batch_is_rejected(gpu):
    """A request that needs more node slots than the whole batch holds is rejected, not retried."""
    reqs = [Request(0, params=dict(depth=400, ratio=0.0, E=37e9, s0=1e-4)),
            Request(1, params=dict(depth=50, ratio=0.0, E=37e9, s0=1e-4), arrival=0.01)]
    budget = 100000 * slot_bytes(_lib.MIXED)
    res = serve(reqs, _lib.MIXED, budget, chunk=50, ipi=50, increment_um=25.0)
    assert res["rejected"] == 1 and res["requests"] == 1
    assert res["done"][0].rid == 1
