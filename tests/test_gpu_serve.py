"""The continuous-batching service: requests admitted while others run give the same results as
each served alone (bit for bit: regions never interact), and the stop criterion retires them."""
import numpy as np
import pytest

from bb_pd_amd import _lib
from bb_pd_amd.serve import Request, request_case, serve

pytestmark = pytest.mark.gpu


def small_requests(n, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        c = request_case(rng, ipi=50, increment_um=25.0)
        if c.nNodes <= 20000:  # depth 50 and 100 mm beams keep the test short
            out.append(c)
    return out


@pytest.mark.parametrize("precision", [_lib.MIXED, _lib.FP64])
def test_served_results_equal_alone(gpu, precision):
    cases = small_requests(4)
    print("case: serve 4 requests", [c.name for c in cases], flush=True)
    # in the mix: staggered arrivals, a budget that holds about two requests at a time
    reqs = [Request(i, c, arrival=0.02 * i) for i, c in enumerate(cases)]
    budget = 2.2 * 20480 * 1200
    mix = serve(reqs, precision, budget, chunk=50, keep_history=True)
    assert mix["requests"] == 4
    for r in mix["done"]:
        alone = serve([Request(0, r.case)], precision, budget, chunk=50, keep_history=True)["done"][0]
        n = min(len(r.history["reaction_static"]), len(alone.history["reaction_static"]))
        assert n >= 50
        for k in ("reaction_static", "deflection", "cmod"):
            a, b = r.history[k][:n], alone.history[k][:n]
            assert np.array_equal(a.view(np.int64), b.view(np.int64)), (r.case.name, k)
        assert r.peak == alone.peak
        # the criterion fired: 0.5 mm of travel (20 increments of 25 um) or the load fell below 20% of its peak
        assert r.steps <= 21 * 50 + 50
