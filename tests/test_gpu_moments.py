"""GPU parity of the radiation moments (through the C ABI).

The moments kernel sums each face's directional integrals in its own order
(psi[j][q] = phi_q(omega_j) w_j / F, fused multiply-add, faces combined last),
so it is not bit-identical to the reference's sequence. Tolerance, per cell:
|got - ref| <= 1e-12 * e for energy, stream and impuls, and <= 1e-12 * e *
sum_f area_f / volume for divstream and divimpuls, where e = max_f sum_j w_j
|I_jf| / F bounds every term of those sums (flux and pressure components and
the divergences cancel, so their own magnitude is no scale).
"""
import os

import numpy as np
import pytest

from mesh_util import linear_flux_illum, moment_scales, random_tetra_cells, six_axis_directions
from oracle import orc
from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs

pytestmark = pytest.mark.gpu

TOL = 1e-12
WIDTH = {"energy": 1, "stream": 3, "impuls": 9, "divstream": 1, "divimpuls": 3}


def check(got, ref, illum, w, F, areas, volume, tol=TOL):
    e, d = moment_scales(illum, w, F, areas, volume)
    for k, v in ref.items():
        if k not in got:
            continue
        scale = d if k.startswith("div") else e
        err = np.abs(np.asarray(got[k]).reshape(v.shape) - v)
        err = err if err.ndim == 1 else err.max(axis=1)
        worst = float(np.max(err / scale)) if err.size else 0.0
        assert worst <= tol, (k, worst)


def run(ds, illum, normals=None, areas=None, volume=None, fpt=0, which=tuple(WIDTH)):
    M, N, faces = illum.shape
    d_illum = D.DeviceArray.from_host(np.ascontiguousarray(illum, dtype=np.float64))
    geo = [None if a is None else D.DeviceArray.from_host(np.ascontiguousarray(a, dtype=np.float64))
           for a in (normals, areas, volume)]
    outs = {k: D.DeviceArray((max(N, 1), WIDTH[k])) for k in which}
    for o in outs.values():
        o.fill(-7.0)
    ds.moments(d_illum, faces, N, energy=outs.get("energy"), stream_out=outs.get("stream"),
               impuls=outs.get("impuls"), divstream=outs.get("divstream"), divimpuls=outs.get("divimpuls"),
               normals=geo[0], areas=geo[1], volume=geo[2], faces_per_thread=fpt)
    D.synchronize()
    res = {}
    for k, o in outs.items():
        h = o.to_host()[:N]
        res[k] = h[:, 0] if WIDTH[k] == 1 else h
        o.free()
    d_illum.free()
    for g in geo:
        if g is not None:
            g.free()
    return res


@pytest.mark.parametrize("fpt", [0, 1, 2, 4])
def test_moments_golden_exact(gpu, golden_dir, fpt):
    z = np.load(os.path.join(golden_dir, "moments_N7_M14.npz"))
    with rs.DirectionSet(z["xyz"], z["w"], float(z["F"])) as ds:
        got = run(ds, z["illum"], z["normals"], z["areas"], z["volume"], fpt=fpt)
    check(got, {k: z[k] for k in WIDTH}, z["illum"], z["w"], float(z["F"]), z["areas"], z["volume"])


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 4099])
@pytest.mark.parametrize("fpt", [1, 2, 4])
def test_moments_match_oracle_ragged(gpu, n, fpt):
    rng = np.random.default_rng(1000 + n)
    xyz, w, F = rs.fibonacci_directions(26)
    normals, areas, volume, _ = random_tetra_cells(n, rng)
    illum = rng.lognormal(0.0, 1.0, size=(26, n, 4))
    ref = orc.moments(illum, xyz, w, F, normals, areas, volume, threads=8)
    with rs.DirectionSet(xyz, w, F) as ds:
        got = run(ds, illum, normals, areas, volume, fpt=fpt)
    check(got, ref, illum, w, F, areas, volume)


def test_moments_24576x200_sizes(gpu):
    """24,576 cells (16^3 x 6), M = 200 Fibonacci directions, log-normal(0, 1) faces."""
    n, m = 6 * 16 ** 3, 200
    rng = np.random.default_rng(77)
    xyz, w, F = rs.fibonacci_directions(m)
    normals, areas, volume, _ = random_tetra_cells(n, rng)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    ref = orc.moments(illum, xyz, w, F, normals, areas, volume, threads=16)
    with rs.DirectionSet(xyz, w, F) as ds:
        got = run(ds, illum, normals, areas, volume)
    check(got, ref, illum, w, F, areas, volume)


def test_moments_cell_level_and_partial_outputs(gpu):
    """faces = 1 input; only some outputs requested; no geometry needed without divergences."""
    rng = np.random.default_rng(4)
    xyz, w, F = rs.fibonacci_directions(40)
    illum = rng.lognormal(size=(40, 777, 1))
    ref = orc.moments(illum, xyz, w, F, which=("energy", "stream", "impuls"))
    with rs.DirectionSet(xyz, w, F) as ds:
        got = run(ds, illum, which=("energy", "impuls"))
    check(got, ref, illum, w, F, None, None)


def test_moments_closed_forms(gpu):
    """Six-axis rule: isotropic input gives energy = I, flux = 0, pressure = I/3 Id,
    zero divergences; a linear face flux (3x, 0, 0) gives divstream = 3 and
    divimpuls = (3, 0, 0) (the reference's TestDivStream field)."""
    xyz, w, F = six_axis_directions()
    rng = np.random.default_rng(9)
    n = 300
    normals, areas, volume, cents = random_tetra_cells(n, rng)
    c = rng.uniform(0.5, 3.0, size=n)
    iso = np.broadcast_to(c[None, :, None], (6, n, 4)).copy()
    lin = linear_flux_illum(cents, base=2.0, slope=3.0)
    with rs.DirectionSet(xyz, w, F) as ds:
        got = run(ds, iso, normals, areas, volume)
        e, d = moment_scales(iso, w, F, areas, volume)
        np.testing.assert_allclose(got["energy"], c, rtol=1e-14)
        assert np.max(np.abs(got["stream"]).max(axis=1) / e) < 1e-15
        np.testing.assert_allclose(got["impuls"], (c[:, None] / 3.0) * np.eye(3).reshape(1, 9), rtol=1e-14,
                                   atol=1e-15)
        assert np.max(np.abs(got["divstream"]) / d) < 1e-14
        assert np.max(np.abs(got["divimpuls"]).max(axis=1) / d) < 1e-14
        got = run(ds, lin, normals, areas, volume)
        _, d = moment_scales(lin, w, F, areas, volume)
        assert np.max(np.abs(got["divstream"] - 3.0) / d) < 1e-13
        assert np.max(np.abs(got["divimpuls"] - [3.0, 0.0, 0.0]).max(axis=1) / d) < 1e-13


def test_moments_argument_errors(gpu):
    xyz, w, F = rs.fibonacci_directions(8)
    d_illum = D.DeviceArray((8, 16, 4))
    out = D.DeviceArray((16,))
    with rs.DirectionSet(xyz, w, F) as ds:
        with pytest.raises(RuntimeError, match="divergences need"):
            ds.moments(d_illum, 4, 16, divstream=out)
        with pytest.raises(RuntimeError, match="faces_per_thread"):
            ds.moments(d_illum, 4, 16, energy=out, faces_per_thread=3)
        ds.moments(d_illum, 4, 0, energy=out)  # empty mesh: nothing to do
    with rs.DirectionSet(xyz, w, 0.0, normalize=True) as ds:
        with pytest.raises(RuntimeError, match="full_area"):
            ds.moments(d_illum, 4, 16, energy=out)
    d_illum.free()
    out.free()


def test_problem_all_param_stream_ordering(gpu):
    """Problem API: geometry once, then twice an upload on stream 0 followed by
    CalculateAllParamStream on the params stream with no host wait in between
    (the upload is ordered before the moments) into pinned host outputs; the
    second pass must see the second upload."""
    rng = np.random.default_rng(31)
    n, m = 3000, 50
    xyz, w, F = rs.fibonacci_directions(m)
    normals, areas, volume, _ = random_tetra_cells(n, rng)
    illum = rng.lognormal(size=(m, n, 4))
    illum2 = rng.lognormal(size=(m, n, 4))
    host = D.PinnedArray((m, n, 4))
    outs = {k: D.PinnedArray((n, WIDTH[k])) for k in WIDTH}
    with rs.Problem(n, xyz, w, F, 0, m) as p:
        p.set_geometry(normals, areas, volume)
        for data in (illum, illum2):
            host.a[...] = data
            p.send_illum_async(host, m * n * 4, stream=rs.STREAM_SCATTERING_1)
            p.calculate_all_param(rs.STREAM_PARAMS, energy=outs["energy"], stream_out=outs["stream"],
                                  impuls=outs["impuls"], divstream=outs["divstream"], divimpuls=outs["divimpuls"],
                                  sync=False)
            p.sync_stream(rs.STREAM_PARAMS)
            got = {k: (o.a[:, 0] if WIDTH[k] == 1 else o.a).copy() for k, o in outs.items()}
            ref = orc.moments(data, xyz, w, F, normals, areas, volume, threads=16)
            check(got, ref, data, w, F, areas, volume)
        assert all(b for b in p.moment_buffers())
    host.free()
    for o in outs.values():
        o.free()


@pytest.mark.big
def test_moments_1572864x400_full_size(gpu):
    """1,572,864 cells, M = 400, log-normal(0, 1.5) faces (20.1 GB): sampled cells vs
    the oracle, and linearity (doubling the input doubles every moment exactly)."""
    n, m = 6 * 64 ** 3, 400
    xyz, w, F = rs.fibonacci_directions(m)
    rng = np.random.default_rng(12)
    normals, areas, volume, _ = random_tetra_cells(n, rng)
    illum = D.DeviceArray((m, n, 4))
    rs.fill_lognormal_dev(illum, 0.0, 1.5, seed=0x5EED + 2)
    geo = [D.DeviceArray.from_host(a) for a in (normals, areas, volume)]
    outs = {k: D.DeviceArray((n, WIDTH[k])) for k in WIDTH}
    outs2 = {k: D.DeviceArray((n, WIDTH[k])) for k in WIDTH}
    try:
        with rs.DirectionSet(xyz, w, F) as ds:
            def launch(o):
                ds.moments(illum, 4, n, energy=o["energy"], stream_out=o["stream"], impuls=o["impuls"],
                           divstream=o["divstream"], divimpuls=o["divimpuls"], normals=geo[0], areas=geo[1],
                           volume=geo[2])
                D.synchronize()
            launch(outs)
            cells = np.unique(np.concatenate([[0, 1, 63, 64, n - 65, n - 1], rng.integers(0, n, 58)]))
            sub = illum.gather_cells(cells)
            ref = orc.moments(sub, xyz, w, F, normals[cells], areas[cells], volume[cells], threads=16)
            host = {k: o.to_host() for k, o in outs.items()}
            got = {k: (h[cells, 0] if WIDTH[k] == 1 else h[cells]) for k, h in host.items()}
            check(got, ref, sub, w, F, areas[cells], volume[cells])
            illum.scale(2.0)
            launch(outs2)
            for k in WIDTH:
                assert np.array_equal(outs2[k].to_host(), 2.0 * host[k]), k
    finally:
        for a in [illum, *geo, *outs.values(), *outs2.values()]:
            a.free()


def test_write_solution_files(gpu, tmp_path):
    """Solution dump after CalculateAllParam: Illum.bin is the first direction's
    face average (bit-exact: faces summed in order, then / 4), the moment files
    hold N elements of 1, 3 or 9 doubles and match the oracle; without geometry
    the divergence files are not written; CompareFiles of a file with itself is 0."""
    rng = np.random.default_rng(44)
    n, m = 2500, 30
    xyz, w, F = rs.fibonacci_directions(m)
    normals, areas, volume, _ = random_tetra_cells(n, rng)
    illum = rng.lognormal(size=(m, n, 4))
    ref = orc.moments(illum, xyz, w, F, normals, areas, volume, threads=16)
    expect_illum = (((illum[0, :, 0] + illum[0, :, 1]) + illum[0, :, 2]) + illum[0, :, 3]) / 4
    for with_geo in (True, False):
        d = tmp_path / ("geo" if with_geo else "nogeo")
        d.mkdir()
        with rs.Problem(n, xyz, w, F, 0, m) as p:
            if with_geo:
                p.set_geometry(normals, areas, volume)
            p.copy_illum_on_device(np.ascontiguousarray(illum), m * n * 4)
            p.write_solution(str(d) + "/")  # before any moments: Illum.bin only
            assert sorted(os.listdir(d)) == ["Illum.bin"]
            p.calculate_all_param(rs.STREAM_PARAMS)
            p.write_solution(str(d) + "/")
        names = sorted(os.listdir(d))
        want = sorted(k for k in rs.SOLUTION_FILES if with_geo or not k.startswith("div"))
        assert names == want
        got_illum = rs.read_simple_bin(str(d / "Illum.bin"))
        assert got_illum.shape == (n,) and np.array_equal(got_illum, expect_illum)
        got = {}
        for k in WIDTH:
            f = d / f"{k}.bin"
            if f.exists():
                a = rs.read_simple_bin(str(f), np.float64, rs.SOLUTION_FILES[f"{k}.bin"])
                assert a.shape[0] == n
                got[k] = a if WIDTH[k] > 1 else a.reshape(n)
        check(got, ref, illum, w, F, areas, volume)
        idx, worst, total = rs.compare_files(str(d / "stream.bin"), str(d / "stream.bin"), "Vector3")
        assert (worst, total) == (0.0, 0.0) and idx == 0
