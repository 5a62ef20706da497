"""The split-precision workload at full size, every result checked.

1,572,864 cells x 4 faces (log-normal(0, 1.5), seeded on the device), M = 400
Fibonacci directions, Henyey-Greenstein g = 0.8 with FP64 row normalisation.
The FP64 MFMA result is pinned to the oracle on sampled cells (<= 1e-12
relative); each split form's 629,145,600 results are then compared with it on
the device (rsh_dev_rel_error). Tolerances, per result:
  fp16x2, bf16x3 : max relative error <= 1e-6
  bf16x2         : max <= 3e-5, mean <= 3e-6 (16-bit operands; not a 1e-6 form)
"""
import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs

pytestmark = pytest.mark.gpu

N, M = 6 * 64 ** 3, 400
TOL = {"fp16x2": (1e-6, 1e-6), "bf16x3": (1e-6, 1e-6), "bf16x2": (3e-5, 3e-6)}


@pytest.fixture(scope="module")
def workload(gpu):
    xyz, w, F = rs.fibonacci_directions(M)
    illum = D.DeviceArray((M, N, 4))
    rs.fill_lognormal_dev(illum, 0.0, 1.5, seed=0x5EED + 3)
    ds = rs.DirectionSet(xyz, w, F, phase="hg", g=0.8, normalize=True)
    ref = D.DeviceArray((M, N))
    ds.scatter_f64(illum, 4, N, 0, M, ref)
    D.synchronize()
    cells = np.unique(np.concatenate([[0, 15, 16, N - 1], np.random.default_rng(5).integers(0, N, 12)]))
    want = orc.scatter_matrix(illum.gather_cells(cells), orc.phase_matrix(xyz, w, F, 1, 0.8, True), threads=16)
    got = ref.gather_columns(cells)
    assert float(np.max(np.abs(got - want) / want)) < 1e-12
    yield illum, ds, ref
    for a in (illum, ref):
        a.free()
    ds.close()


@pytest.mark.parametrize("form", ["fp16x2", "bf16x3", "bf16x2"])
@pytest.mark.parametrize("out_f32", [False, True])
def test_every_result_of_the_full_workload(workload, form, out_f32):
    illum, ds, ref = workload
    sb, cb = rs.split_sizes(form, 1, N, M)
    state = D.DeviceArray((sb,), np.uint8)
    scale = D.DeviceArray((cb // 4,), np.float32)
    out = D.DeviceArray((M, N), np.float32 if out_f32 else np.float64)
    rs.pack_split_avg_dev(form, illum, N, M, state, scale)
    ds.scatter_split(form, state, scale, 1, N, 0, M, out, out_f32=out_f32)
    e = rs.rel_error_dev(out, ref, M, N, threshold=TOL[form][0])
    for a in (state, scale, out):
        a.free()
    print(f"{form} {'fp32' if out_f32 else 'fp64'} out: max {e['max_rel_err']:.3e} at (row, cell) "
          f"{e['worst_row_cell']}, mean {e['mean_rel_err']:.3e}, above {TOL[form][0]:g}: {e['count_above']} of "
          f"{e['results']}")
    assert e["results"] == M * N
    assert e["max_rel_err"] <= TOL[form][0] and e["mean_rel_err"] <= TOL[form][1]


def test_rel_error_reduction_counts_planted_errors(gpu):
    """The reduction itself: planted relative errors of known size at known places
    (an FP64 array against a scaled copy), max, mean and count above exact."""
    m, n = 37, 5003
    rng = np.random.default_rng(2)
    ref_h = rng.lognormal(size=(m, n))
    got_h = ref_h.copy()
    got_h[3, 17] *= 1 + 4e-6
    got_h[36, n - 1] *= 1 - 2e-6
    got_h[20, 2500] *= 1 + 5e-7
    ref, got = D.DeviceArray.from_host(ref_h), D.DeviceArray.from_host(got_h)
    e = rs.rel_error_dev(got, ref, m, n, threshold=1e-6)
    want = np.abs(got_h - ref_h) / ref_h
    assert e["max_rel_err"] == want.max() and e["count_above"] == 2 and e["worst_row_cell"] == (3, 17)
    assert e["mean_rel_err"] == pytest.approx(want.mean(), rel=1e-12)
    # a NaN counts as an infinite error
    got_h[0, 0] = np.nan
    got2 = D.DeviceArray.from_host(got_h)
    assert rs.rel_error_dev(got2, ref, m, n)["max_rel_err"] == np.inf
    for a in (ref, got, got2):
        a.free()
