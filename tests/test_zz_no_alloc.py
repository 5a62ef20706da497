"""rsh_scatter_f64_dev_noalloc: the allocation-free FP64 entry point.

It builds the phase panel into the library's fixed per-device scratch instead
of the direction set's (allocating) panel cache and runs the same MFMA kernel,
so its results must equal rsh_scatter_f64_dev's bit for bit. Checked on two
calls in a row, the second with a smaller input (fewer cells, one face per
cell, an inner output range), with no change of the device's free memory
across the allocation-free calls (after the first call on the device, which
finds the scratch array), and with panels beyond the scratch refused.
"""
import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import _native as N
from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs

pytestmark = pytest.mark.gpu


def mem_free():
    import ctypes as C

    f, t = C.c_int64(), C.c_int64()
    N.call("rsh_dev_mem_info", 0, C.byref(f), C.byref(t))
    return f.value


def test_noalloc_equals_allocating_entry_point(gpu):
    m = 400
    xyz, w, F = rs.fibonacci_directions(m)
    rng = np.random.default_rng(71)
    calls = [(3000, 4, 0, m), (1000, 1, 37, 251)]   # the second input is smaller
    inputs = [rng.lognormal(0.0, 1.5, size=(m, n, faces)) for n, faces, _, _ in calls]
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.6, normalize=True) as ds:
        d_in = [D.DeviceArray.from_host(a) for a in inputs]
        d_out = [D.DeviceArray((k1 - k0, n)) for n, _, k0, k1 in calls]
        # The first call on the device looks up the scratch array's address; the
        # runtime places the library's device globals (the array included) on
        # the device when its code is first used there. From then on the calls
        # allocate nothing.
        ds.scatter_f64_noalloc(d_in[1], 1, 16, 0, 1, d_out[1])
        D.synchronize()
        before = mem_free()
        for (n, faces, k0, k1), a, o in zip(calls, d_in, d_out):
            ds.scatter_f64_noalloc(a, faces, n, k0, k1, o)
            D.synchronize()
        assert mem_free() == before          # no device allocation in either call
        got = [o.to_host() for o in d_out]
        for (n, faces, k0, k1), a, o, g in zip(calls, d_in, d_out, got):
            o.fill(np.nan)
            ds.scatter_f64(a, faces, n, k0, k1, o)
            D.synchronize()
            assert np.array_equal(g, o.to_host())
        P = orc.phase_matrix(xyz, w, F, 1, 0.6, True)
        ref = orc.scatter_matrix(inputs[0], P, threads=16)
        assert float(np.max(np.abs(got[0] - ref) / ref)) < 1e-12
        for x in d_in + d_out:
            x.free()


def test_noalloc_refuses_panels_beyond_its_scratch(gpu):
    assert N.lib.rsh_noalloc_scratch_bytes() == 8 << 20
    m = 1200                                  # FP64 panel of M = 1200 rows: 1200 / 4 j-steps x 75 row blocks x 512 B
    xyz, w, F = rs.fibonacci_directions(m)
    with rs.DirectionSet(xyz, w, F) as ds:
        a, o = D.DeviceArray((m, 16, 1)), D.DeviceArray((m, 16))
        with pytest.raises(N.RshError) as e:
            ds.scatter_f64_noalloc(a, 1, 16, 0, m, o)
        assert e.value.status == 6 and b"scratch" in N.lib.rsh_last_error()
        ds.scatter_f64_noalloc(a, 1, 16, 0, 400, o)   # a smaller output range fits
        a.free()
        o.free()
