"""GPU parity of the FP64 scattering integral (through the C ABI).

Tolerances:
* direct kernel (reference order, no FMA) vs oracle parity build: bit-exact;
* MFMA GEMM vs oracle: max relative error <= 1e-12 (all terms positive, so any
  summation order stays within ~M u of the sequential sum; 1e-12 is the
  required bound);
* golden fixtures (exact arithmetic): <= 1e-14 for the direct kernel.
"""
import os

import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs

pytestmark = pytest.mark.gpu

TILES = [(8, 2), (4, 2), (4, 4)]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0


def run(ds, illum, faces, k0, k1, kernel="auto", tile=None, ld=None):
    M, N = illum.shape[0], illum.shape[1]
    d_illum = D.DeviceArray.from_host(np.ascontiguousarray(illum, dtype=np.float64))
    ld = N if ld is None else ld
    out = D.DeviceArray((max(k1 - k0, 1), ld))
    out.fill(-7.0)
    ds.scatter_f64(d_illum, faces, N, k0, k1, out, ld=ld, kernel=kernel, tile=tile)
    D.synchronize()
    res = out.to_host()[: k1 - k0, :N]
    d_illum.free()
    out.free()
    return res


def test_mfma_f64_lane_layout(gpu):
    assert rs.selftest_mfma_f64(0) == 0


@pytest.mark.parametrize("name", ["tiny_thomson_N64_M26.npz", "tiny_thomson_random_N37_M11.npz"])
def test_golden_thomson(gpu, golden_dir, name):
    d = np.load(os.path.join(golden_dir, name))
    M = len(d["w"])
    with rs.DirectionSet(d["xyz"], d["w"], float(d["F"])) as ds:
        Sd = run(ds, d["illum"], 4, 0, M, kernel="direct")
        assert rel(Sd, d["S"]) < 1e-14
        Sm = run(ds, d["illum"], 4, 0, M)
        assert rel(Sm, d["S"]) < 1e-12


def test_golden_hg_normalised(gpu, golden_dir):
    d = np.load(os.path.join(golden_dir, "tiny_hg08_norm_N48_M26.npz"))
    with rs.DirectionSet(d["xyz"], d["w"], float(d["F"]), phase="hg", g=float(d["g"]), normalize=True) as ds:
        assert rel(ds.phase_matrix(), d["P"]) < 1e-13
        Sm = run(ds, d["illum"], 4, 0, 26)
        assert rel(Sm, d["S"]) < 1e-12
        Sd = run(ds, d["illum"], 4, 0, 26, kernel="direct")
        assert rel(Sd, d["S"]) < 1e-13


def test_phase_matrix_matches_oracle(gpu):
    xyz, w, F = rs.fibonacci_directions(400)
    with rs.DirectionSet(xyz, w, F) as ds:
        P = ds.phase_matrix()
        assert np.array_equal(P, orc.phase_matrix(xyz, w, F))  # same ops, no FMA: bit-exact
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.8, normalize=True) as ds:
        P = ds.phase_matrix(100, 300)
        assert rel(P, orc.phase_matrix(xyz, w, F, 1, 0.8, True, 100, 300)) < 1e-13
        assert np.max(np.abs(P.sum(axis=1) - 1)) < 1e-14


def test_24576x200_parity(gpu):
    """16^3 x 6 = 24,576 cells, M = 200 Fibonacci, Thomson, log-normal(0,1)."""
    n, m = 6 * 16 ** 3, 200
    rng = np.random.default_rng(0x5EED + 1)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    ref = orc.scatter_reference(illum, xyz, w, F, threads=16)
    with rs.DirectionSet(xyz, w, F) as ds:
        Sd = run(ds, illum, 4, 0, m, kernel="direct")
        assert np.array_equal(Sd, ref), f"direct kernel not bit-exact: max rel {rel(Sd, ref):.3e}"
        Sm = run(ds, illum, 4, 0, m)
        err = rel(Sm, ref)
        print(f"24,576 x 200 MFMA max rel err {err:.3e}")
        assert err < 1e-12


CASES = [  # (n_cells, M, k0, k1, faces)
    (1, 1, 0, 1, 4),
    (15, 3, 0, 3, 4),
    (17, 26, 5, 21, 4),
    (100, 5, 2, 3, 1),
    (1007, 130, 17, 129, 1),
    (4097, 401, 0, 401, 4),
    (300, 800, 0, 800, 4),
    (129, 800, 390, 800, 1),
    (2000, 64, 0, 64, 4),
]


@pytest.mark.parametrize("n,m,k0,k1,faces", CASES)
def test_ragged_shapes(gpu, n, m, k0, k1, faces):
    rng = np.random.default_rng(n * 1000 + m)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, faces))
    xyz, w, F = rs.fibonacci_directions(m)
    ref = orc.scatter_reference(illum, xyz, w, F, k_disp=k0, k_count=k1 - k0, threads=16)
    with rs.DirectionSet(xyz, w, F) as ds:
        Sd = run(ds, illum, faces, k0, k1, kernel="direct")
        assert np.array_equal(Sd, ref)
        for tile in TILES:
            Sm = run(ds, illum, faces, k0, k1, kernel="mfma", tile=tile, ld=n + 3)
            assert rel(Sm, ref) < 1e-12, f"tile {tile}"


def test_tiles_bit_identical(gpu):
    n, m = 5000, 400
    rng = np.random.default_rng(7)
    illum = rng.lognormal(0.0, 1.5, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    with rs.DirectionSet(xyz, w, F) as ds:
        outs = [run(ds, illum, 4, 0, m, kernel="mfma", tile=t) for t in TILES]
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_hg_normalised_vs_oracle(gpu):
    n, m = 3000, 400
    rng = np.random.default_rng(11)
    illum = rng.lognormal(0.0, 1.5, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    P = orc.phase_matrix(xyz, w, F, 1, 0.8, True)
    ref = orc.scatter_matrix(illum, P, threads=16)
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.8, normalize=True) as ds:
        Sm = run(ds, illum, 4, 0, m)
    assert rel(Sm, ref) < 1e-12


def test_empty_inputs_are_noops(gpu):
    xyz, w, F = rs.fibonacci_directions(8)
    with rs.DirectionSet(xyz, w, F) as ds:
        out = D.DeviceArray((1, 4))
        out.fill(5.0)
        illum = D.DeviceArray((8, 4, 4))
        illum.fill(1.0)
        ds.scatter_f64(illum, 4, 0, 0, 8, out)
        ds.scatter_f64(illum, 4, 4, 3, 3, out)
        D.synchronize()
        assert np.all(out.to_host() == 5.0)


def test_problem_api_streams_and_offsets(gpu):
    """InitDevice / CudaSendIllumAsync (both overloads) / CalculateIntScatteringAsyncMPIStream
    on two streams / CudaSyncStream / event timing / ClearDevice."""
    n, m = 3001, 90
    kb, ke = 10, 70  # this rank's local output directions
    rng = np.random.default_rng(5)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    host = D.PinnedArray((illum.size,))
    host.a[:] = illum.reshape(-1)
    host_S = D.PinnedArray(((ke - kb) * n,))
    host_S.a[:] = -1.0
    p = rs.Problem(n, xyz, w, F, kb, ke)
    try:
        per_dir = 4 * n
        # first 30 directions in one block on stream 0 (overload 1: same offset)
        p.send_illum_async(host, 30 * per_dir, 0, stream=0)
        # the rest direction by direction on stream 1 (overload 2: separate offsets)
        for j in range(30, m):
            p.send_illum_async(host, per_dir, dst_off=j * per_dir, src_off=j * per_dir, stream=1)
        p.calculate_int_scattering(kb, 40, stream=0, host_out=host_S)
        p.calculate_int_scattering(40, ke, stream=1, host_out=host_S)
        p.sync_stream(0)
        p.sync_stream(1)
        times = p.event_times()
        assert times[0] > 0 and times[1] > 0 and times[2] == -1
        got = host_S.a.reshape(ke - kb, n).copy()
    finally:
        p.close()
        host.free()
        host_S.free()
    ref = orc.scatter_reference(illum, xyz, w, F, k_disp=kb, k_count=ke - kb, threads=16)
    assert rel(got, ref) < 1e-12


def test_problem_sync_copy_and_full_range(gpu):
    n, m = 700, 40
    rng = np.random.default_rng(9)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    out = np.zeros((m, n))
    with rs.Problem(n, xyz, w, F) as p:
        p.copy_illum_on_device(illum.reshape(-1), illum.size)
        p.calculate_int_scattering(0, m, stream=2, host_out=out)
        p.wait()
    assert rel(out, orc.scatter_reference(illum, xyz, w, F)) < 1e-12


@pytest.mark.parametrize("chunks", [1, 3, 8])
def test_scatter_from_host_pipelined(gpu, chunks):
    """CalculateIntScattering's upload + compute + copy back pipelined by cell range
    (rsh_scatter_from_host): S for a local direction range equals the unpipelined
    sequence bit for bit (same kernel, same per-cell sums) and the oracle to 1e-12;
    the device illum holds the whole upload afterwards (a later scatter of the
    full range from it matches too). Ragged cell ranges (n not a multiple of chunks)."""
    n, m = 2999, 60
    kb, ke = 5, 50
    rng = np.random.default_rng(11)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    host = D.PinnedArray((illum.size,))
    host.a[:] = illum.reshape(-1)
    a = D.PinnedArray(((ke - kb) * n,))
    b = D.PinnedArray(((ke - kb) * n,))
    a.a[:] = -1.0
    b.a[:] = -2.0
    p = rs.Problem(n, xyz, w, F, kb, ke)
    try:
        p.calculate_int_scattering_from_host(host, kb + 3, ke, host_out=a, chunks=chunks)
        p.copy_illum_on_device(host, illum.size)
        p.calculate_int_scattering(kb + 3, ke, stream=0, host_out=b)
        p.sync_stream(0)
        got_a = a.a.reshape(ke - kb, n).copy()
        got_b = b.a.reshape(ke - kb, n).copy()
        # rows before kb + 3 are not written by either call
        assert np.all(got_a[:3] == -1.0) and np.all(got_b[:3] == -2.0)
        assert np.array_equal(got_a[3:], got_b[3:])
        # the device illum is complete after the pipelined call: recompute from it
        p.calculate_int_scattering_from_host(host, kb, kb, host_out=None, chunks=chunks)
        p.calculate_int_scattering(kb, ke, stream=1, host_out=b)
        p.sync_stream(1)
        full = b.a.reshape(ke - kb, n).copy()
    finally:
        p.close()
        host.free()
        a.free()
        b.free()
    ref = orc.scatter_reference(illum, xyz, w, F, k_disp=kb, k_count=ke - kb, threads=16)
    assert rel(got_a[3:], ref[3:]) < 1e-12
    assert rel(full, ref) < 1e-12


@pytest.mark.big
def test_1572864x400_full_size(gpu):
    """64^3 x 6 = 1,572,864 cells, M = 400, Thomson, log-normal(0, 1.5) (20.1 GB of faces)."""
    n, m = 6 * 64 ** 3, 400
    xyz, w, F = rs.fibonacci_directions(m)
    illum = D.DeviceArray((m, n, 4))
    rs.fill_lognormal_dev(illum, 0.0, 1.5, seed=0x5EED + 2)
    out = D.DeviceArray((m, n))
    out2 = D.DeviceArray((m, n))
    try:
        with rs.DirectionSet(xyz, w, F) as ds:
            ds.scatter_f64(illum, 4, n, 0, m, out)
            D.synchronize()
            # sampled cells vs the oracle (edges of the cell range + random interior cells)
            rng = np.random.default_rng(12)
            cells = np.unique(np.concatenate([[0, 1, 15, 16, 63, 64, n - 65, n - 2, n - 1],
                                              rng.integers(0, n, 55)]))
            sub = illum.gather_cells(cells)
            ref = orc.scatter_reference(sub, xyz, w, F, threads=16)
            got = out.gather_columns(cells)
            err = rel(got, ref)
            print(f"1,572,864 x 400 sampled max rel err {err:.3e}")
            assert err < 1e-12
            # linearity: doubling the input doubles the output exactly
            illum.scale(2.0)
            ds.scatter_f64(illum, 4, n, 0, m, out2)
            D.synchronize()
            a, b = out.gather_columns(cells), out2.gather_columns(cells)
            assert np.array_equal(b, 2.0 * a)
            # isotropic input: S[k][c] = sum_j P[k][j] for every cell
            illum.fill(1.0)
            ds.scatter_f64(illum, 4, n, 0, m, out2)
            D.synchronize()
            rows = ds.phase_matrix().sum(axis=1)
            iso = out2.gather_columns(np.arange(0, n, 4099))
            assert rel(iso, np.repeat(rows[:, None], iso.shape[1], axis=1)) < 1e-13
    finally:
        illum.free()
        out.free()
        out2.free()


def test_scatter_from_host_cell_level_state(gpu):
    """The pipelined host path on a cell-level (faces = 1) problem, HG phase with
    row normalisation, with more cell ranges than a few cells each (chunks = 64
    on 100 cells) and a host array that is not pinned: equal to the unpipelined
    sequence bit for bit and to the oracle to 1e-12."""
    n, m = 100, 33
    rng = np.random.default_rng(12)
    cells = rng.lognormal(0.0, 1.0, size=(m, n))
    xyz, w, F = rs.fibonacci_directions(m)
    a = np.full((m, n), -1.0)
    b = np.full((m, n), -2.0)
    with rs.Problem(n, xyz, w, F, faces=1, phase="hg", g=0.6, normalize=True) as p:
        p.calculate_int_scattering_from_host(cells, 0, m, host_out=a, chunks=64)
        p.copy_illum_on_device(cells, cells.size)
        p.calculate_int_scattering(0, m, stream=2, host_out=b)
        p.wait()
    assert np.array_equal(a, b)
    P = orc.phase_matrix(xyz, w, F, 1, 0.6, True)
    ref = orc.scatter_matrix(cells[:, :, None], P)
    assert rel(a, ref) < 1e-12
