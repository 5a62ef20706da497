"""GPU serving loop (rsh_server_*): continuous batching of source-iteration requests.

* a request's result is the same alone and in a mix (bit-identical state and
  iteration count, FP64 and split forms: batching changes neither the kernels
  nor the summation order);
* every request converges below its FP64 residual tolerance and matches the
  oracle's iteration of the same seeded problem (state within 1e-12 for FP64);
* admission respects the HBM budget; cancellation frees the slot;
* a real-time Poisson trace completes with finite latencies.
"""
import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs

pytestmark = pytest.mark.gpu

GB = 1 << 30


def result(srv, rid, m, n):
    out = D.DeviceArray((m, n))
    srv.result(rid, out)
    h = out.to_host()
    out.free()
    return h


def drain(srv, limit=500):
    for _ in range(limit):
        running, _ = srv.step()
        st = srv.stats()
        if running == 0 and st["pending"] == 0:
            return
    raise AssertionError("server did not drain")


@pytest.mark.parametrize("form,tol,n_target,scale", [("fp64", 1e-8, 3000, 1), ("bf16x3", 1e-6, 3000, 1),
                                                     ("fp16x2", 1e-6, 250000, 40)])
def test_alone_equals_in_mix(gpu, form, tol, n_target, scale):
    """A request's state and iteration count are the same alone and in a mix that
    arrives around it (FP64 and split forms; the last case at 250,000 cells with
    a mix of 40,000-220,000-cell requests)."""
    target = rs.request(n_target, 400, "hg", 0.6, tolerance=tol, form=form, seed=5)
    others = [rs.request((1000 + 700 * i) * scale, m, ph, g, tolerance=tol, form=form, seed=100 + i)
              for i, (m, ph, g) in enumerate([(200, "thomson", 0.0), (800, "hg", 0.9), (400, "hg", 0.3),
                                              (200, "hg", 0.6)])]
    with rs.Server(64 * GB, 8, iters_per_step=2, keep_results=True) as srv:
        rid = srv.submit(target)
        drain(srv)
        alone, info_alone = result(srv, rid, 400, n_target), srv.poll(rid)
    with rs.Server(64 * GB, 8, iters_per_step=2, keep_results=True) as srv:
        ids = [srv.submit(others[0]), srv.submit(others[1])]
        srv.step()                      # the target arrives while others are on the GPU
        rid = srv.submit(target)
        ids += [srv.submit(o) for o in others[2:]]
        drain(srv)
        mixed, info_mix = result(srv, rid, 400, n_target), srv.poll(rid)
        for i in ids:
            assert srv.poll(i)["status"] == "converged"
    assert info_alone["status"] == info_mix["status"] == "converged"
    assert info_alone["iterations"] == info_mix["iterations"]
    assert info_alone["residual"] < tol
    assert np.array_equal(alone, mixed)


def test_converged_result_matches_oracle(gpu):
    n, m = 1200, 200
    desc = rs.request(n, m, "hg", 0.8, tolerance=1e-8, seed=9)
    with rs.Server(16 * GB, 4, iters_per_step=3, keep_results=True) as srv:
        rid = srv.submit(desc)
        drain(srv)
        info = srv.poll(rid)
        got = result(srv, rid, m, n)
    assert info["status"] == "converged" and info["residual"] < 1e-8
    # regenerate the request's seeded inputs on the device and iterate with the oracle
    I0, B = D.DeviceArray((m, n)), D.DeviceArray((m, n))
    rs.fill_lognormal_dev(I0, 0.0, 1.0, seed=9 * 2 + 1)
    rs.fill_uniform_dev(B, 0.5, 2.0, seed=9 * 2 + 2)
    D.synchronize()
    xyz, w, F = rs.fibonacci_directions(m)
    P = orc.phase_matrix(xyz, w, F, 1, 0.8, True)
    ref, resid = orc.source_iteration(I0.to_host(), B.to_host(), P, 0.5, info["iterations"], threads=16)
    assert float(np.max(np.abs(got - ref) / ref)) < 1e-12
    assert abs(resid[-1] - info["residual"]) <= 1e-6 * resid[-1]
    I0.free()
    B.free()


def test_admission_respects_budget_and_cancel(gpu):
    desc = rs.request(20000, 400, tolerance=1e-12, max_iterations=6, seed=3)
    fp = rs.Batch.footprint("fp64", 20000, 400)
    with rs.Server(int(2.5 * fp), 8, iters_per_step=2) as srv:
        ids = [srv.submit(desc) for _ in range(5)]
        srv.step()
        st = srv.stats()
        assert st["running"] == 2 and st["pending"] == 3 and st["used_bytes"] == 2 * fp
        srv.cancel(ids[4])              # pending request
        srv.cancel(ids[0])              # running request
        assert srv.stats()["used_bytes"] == fp
        drain(srv)
        assert [srv.poll(i)["status"] for i in ids] == ["cancelled", "maxed", "maxed", "maxed", "cancelled"]
        assert all(srv.poll(i)["iterations"] == 6 for i in ids[1:4])


@pytest.mark.parametrize("form,tol", [("fp64", 1e-8), ("fp16x2", 1e-6)])
def test_poisson_trace(gpu, form, tol):
    rng = np.random.default_rng(17)
    k = 12
    descs = [rs.request(int(rng.integers(2000, 20000)), int(rng.choice([200, 400, 800])), "hg",
                        float(rng.choice([0.3, 0.6, 0.9])), tolerance=tol, form=form, seed=int(s))
             for s in rng.integers(0, 1 << 30, k)]
    arrivals = np.cumsum(rng.exponential(0.01, k))
    with rs.Server(32 * GB, 16, iters_per_step=2) as srv:
        lat, wall = srv.run_trace(descs, arrivals)
        for i in range(k):
            info = srv.poll(i)
            assert info["status"] == "converged" and info["residual"] < tol
    assert np.all(lat > 0) and np.all(np.isfinite(lat)) and wall > 0


@pytest.mark.parametrize("form", ["fp64", "bf16x3"])
def test_request_filling_the_whole_budget_is_admitted(gpu, form):
    """The footprint counts the arena's 256-byte rounding of every slot buffer:
    a request with an odd cell count whose footprint equals the budget is
    admitted into the empty arena and converges (it used to wait for ever)."""
    n, m = 2001, 200
    fp = rs.Batch.footprint(form, n, m)
    desc = rs.request(n, m, "hg", 0.6, tolerance=1e-6, form=form, seed=12)
    with rs.Server(fp, 4, iters_per_step=2) as srv:
        rid = srv.submit(desc)
        drain(srv, limit=100)
        assert srv.poll(rid)["status"] == "converged"


def test_unservable_requests_are_refused_at_submission(gpu):
    """A request the batch cannot hold (FP64 with 2^24 cells or more) or larger
    than the whole budget is refused by submit, so nothing can block the queue."""
    from relativistic_solver_hllc_amd._native import RshError

    with rs.Server(1 * GB, 4, iters_per_step=2) as srv:
        with pytest.raises(RshError) as e:
            srv.submit(rs.request(1 << 24, 200, tolerance=1e-8))
        assert e.value.status == 6
        with pytest.raises(RshError) as e:
            srv.submit(rs.request(1 << 20, 800, tolerance=1e-8))
        assert e.value.status == 3
        assert srv.stats()["pending"] == 0
