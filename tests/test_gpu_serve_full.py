"""The serving loop at its benchmark scale: a real-time Poisson trace (10 requests
per second) of 24 seeded requests drawn as the serving workload draws them (N
log-uniform in [1e4, 1e6] cells as 6 n^3, M in {200, 400, 800}, Thomson or HG g
in {0.3, 0.6, 0.9}), FP64 to 1e-8 and bf16x3 to 1e-6, continuous batching under
a 200 GB budget.

Checked: every request converges below its tolerance; three requests (the
largest, by N x M, and two others) run again alone in a fresh server reach the
same state bit for bit and in the same number of iterations (the mix changes
neither the kernels nor the summation order). The FP64 iteration itself is
pinned against the oracle in test_gpu_iter.py and test_gpu_serve.py.
"""
import numpy as np
import pytest

from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs

pytestmark = pytest.mark.gpu

GB = 1 << 30


def trace_problems(seed=0x5EED + 7, count=24, n_lo=1e4, n_hi=1e6):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n_target = float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        edge = max(1, int(round((n_target / 6.0) ** (1.0 / 3.0))))
        m = int(rng.choice([200, 400, 800]))
        ph = int(rng.integers(0, 4))
        phase, g = ("thomson", 0.0) if ph == 0 else ("hg", (0.3, 0.6, 0.9)[ph - 1])
        out.append(dict(n=6 * edge ** 3, m=m, phase=phase, g=g, seed=int(rng.integers(0, 1 << 30))))
    return out


def state(srv, rid, m, n):
    out = D.DeviceArray((m, n))
    srv.result(rid, out)
    h = out.to_host()
    out.free()
    return h


@pytest.mark.parametrize("form,tol", [("fp64", 1e-8), ("bf16x3", 1e-6)])
def test_full_scale_trace_converges_and_alone_equals_mix(gpu, form, tol):
    probs = trace_problems()
    assert max(p["n"] for p in probs) > 500_000
    rng = np.random.default_rng(0x5EED + 77)
    arrivals = np.cumsum(rng.exponential(0.1, len(probs)))
    descs = [rs.request(p["n"], p["m"], p["phase"], p["g"], tolerance=tol, form=form, seed=p["seed"]) for p in probs]
    big = int(np.argmax([p["n"] * p["m"] for p in probs]))
    picks = sorted({big, 3, 17})
    with rs.Server(200 * GB, 64, iters_per_step=2, keep_results=True) as srv:
        lat, wall = srv.run_trace(descs, arrivals)
        infos = [srv.poll(i) for i in range(len(probs))]
        for i, info in enumerate(infos):
            assert info["status"] == "converged" and info["residual"] < tol, (i, probs[i], info)
        mixed = {i: state(srv, i, probs[i]["m"], probs[i]["n"]) for i in picks}
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)
    for i in picks:
        with rs.Server(32 * GB, 4, iters_per_step=2, keep_results=True) as solo:  # the largest request needs 19.2 GB
            rid = solo.submit(descs[i])
            for _ in range(1000):
                running, _ = solo.step()
                if running == 0 and solo.stats()["pending"] == 0:
                    break
            info = solo.poll(rid)
            assert info["status"] == "converged" and info["iterations"] == infos[i]["iterations"], (i, info, infos[i])
            assert np.array_equal(state(solo, rid, probs[i]["m"], probs[i]["n"]), mixed[i]), (i, probs[i])
