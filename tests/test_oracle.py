"""Pins the CPU oracle (oracle/scatter_oracle.cpp) before it is trusted.

Each oracle function is checked against something that fixes it
independently of its own code, chosen so that a dropped term, a wrong index
or sign, or a transposed operand fails:

* golden vectors from exact rational / 60-digit decimal brute force
  (tests/golden/make_golden.py) -- Thomson reference path, HG normalised path,
  5-step source iteration with residual trace, direction partition, direction file;
* closed forms: Gamma(0) = 3/4, Gamma(+-1) = 3/2, mean of the phase kernel over the
  sphere = 1 by Gauss-Legendre x uniform-phi quadrature, HG g = 0 is isotropic;
* invariants: M = 1 closed form, isotropic input, linearity, cell-permutation
  equivariance, joint permutation of input directions with their weights,
  reciprocity w_k P[k][j] = w_j P[j][k] of the symmetric Thomson kernel,
  GetDisp = exclusive prefix sum of GetSend;
* radiation moments: exact rational golden vector (energy, flux, pressure
  tensor, both divergences on random tetrahedra), the isotropic closed form on
  the six-axis rule (energy = I, flux = 0, pressure = I/3 Id), the divergence
  theorem for a linear face flux (the reference's TestDivStream check: div = 3),
  a zero divergence for face-uniform fields (sum_f n_f A_f = 0), face and cell
  permutation, and the simple-bin geometry file format against numpy.

The reference itself cannot be built here (its CPU path needs Eigen and MPI
headers absent from the image, its GPU path needs nvcc), so there is no
oracle/_ref comparison; the reference ships no golden data of its own.
"""
import math
import os

import numpy as np
import pytest

from oracle import orc
from mesh_util import linear_flux_illum, moment_scales, random_tetra_cells, six_axis_directions


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_gamma_closed_form():
    assert orc.gamma(0.0) == 0.75
    assert orc.gamma(1.0) == 1.5
    assert orc.gamma(-1.0) == 1.5
    for mu in np.linspace(-1, 1, 17):
        assert orc.gamma(mu) == orc.gamma(-mu)
        assert orc.phase(0, 0.3, mu) == orc.gamma(mu)


@pytest.mark.parametrize("kind,g", [(0, 0.0), (1, 0.0), (1, 0.3), (1, 0.8), (1, -0.6)])
def test_phase_kernel_mean_is_one(kind, g):
    # (1/4pi) int p dOmega = (1/2) int_{-1}^{1} p(mu) dmu (azimuth integrates exactly).
    # Composite 40-point Gauss-Legendre on panels graded toward the forward peak mu = 1.
    x, wq = np.polynomial.legendre.leggauss(40)
    edges = [-1.0, 0.0, 0.5, 0.8, 0.9, 0.95, 0.98, 0.99, 0.995, 0.998, 0.999, 1.0]
    terms = []
    for a, b in zip(edges[:-1], edges[1:]):
        for xi, wi in zip(x, wq):
            terms.append(0.5 * (b - a) * wi * orc.phase(kind, g, 0.5 * (a + b) + 0.5 * (b - a) * xi))
    val = 0.5 * math.fsum(terms)
    assert abs(val - 1.0) < 1e-13


def test_hg_isotropic_and_peak():
    for mu in (-1.0, -0.2, 0.5, 1.0):
        assert orc.phase(1, 0.0, mu) == 1.0
    g = 0.8
    assert abs(orc.phase(1, g, 1.0) - (1 + g) / (1 - g) ** 2) < 1e-12
    assert abs(orc.phase(1, g, -1.0) - (1 - g) / (1 + g) ** 2) < 1e-12


def test_reference_path_golden_thomson(golden_dir):
    d = np.load(os.path.join(golden_dir, "tiny_thomson_N64_M26.npz"))
    S = orc.scatter_reference(d["illum"], d["xyz"], d["w"], float(d["F"]))
    # sequential FP64 sum of M = 26 positive terms vs exact: |err| <= ~(M + 6) u
    assert rel(S, d["S"]) < 1e-14


def test_reference_path_golden_random_dirs(golden_dir):
    d = np.load(os.path.join(golden_dir, "tiny_thomson_random_N37_M11.npz"))
    S = orc.scatter_reference(d["illum"], d["xyz"], d["w"], float(d["F"]))
    assert rel(S, d["S"]) < 1e-14


def test_reference_path_output_range_and_cells(golden_dir):
    d = np.load(os.path.join(golden_dir, "tiny_thomson_N64_M26.npz"))
    S = orc.scatter_reference(d["illum"], d["xyz"], d["w"], float(d["F"]), k_disp=5, k_count=9, cell_begin=3,
                              cell_end=40)
    assert rel(S[:, 3:40], d["S"][5:14, 3:40]) < 1e-14
    assert np.all(S[:, :3] == 0) and np.all(S[:, 40:] == 0)


def test_phase_matrix_and_scatter_golden_hg(golden_dir):
    d = np.load(os.path.join(golden_dir, "tiny_hg08_norm_N48_M26.npz"))
    P = orc.phase_matrix(d["xyz"], d["w"], float(d["F"]), kind=1, g=float(d["g"]), normalize=True)
    assert rel(P, d["P"]) < 1e-13
    assert np.max(np.abs(P.sum(axis=1) - 1.0)) < 1e-14
    S = orc.scatter_matrix(d["illum"], P)
    assert rel(S, d["S"]) < 1e-13


def test_source_iteration_golden(golden_dir):
    d = np.load(os.path.join(golden_dir, "iter5_thomson_norm_N32_M14.npz"))
    P = orc.phase_matrix(d["xyz"], d["w"], float(d["F"]), kind=0, normalize=True)
    I, resid = orc.source_iteration(d["I0"], d["B"], P, float(d["albedo"]), 5)
    assert rel(I, d["trace"][-1]) < 1e-13
    assert rel(resid, d["resid"]) < 1e-12


def test_recalc_norm_small():
    old = np.array([1.0, 2.0, 4.0, 1.0])
    new = np.array([1.0, 1.0, 5.0, 2.0])
    r, upd = orc.recalc_norm(old, new)
    assert r == 1.0  # |2 - 1| / 1
    assert np.array_equal(upd, new)
    r, _ = orc.recalc_norm(new, new)
    assert r == 0.0


def test_partition_golden(golden_dir):
    d = np.load(os.path.join(golden_dir, "partition.npz"))
    for i in range(len(d["np"])):
        p, n = int(d["np"][i]), int(d["n"][i])
        assert np.array_equal(orc.get_send(p, n), d["send"][i][:p])
        assert np.array_equal(orc.get_disp(p, n), d["disp"][i][:p])


def test_partition_prefix_sum_property():
    for p in range(1, 13):
        for n in range(0, 60):
            s, dsp = orc.get_send(p, n), orc.get_disp(p, n)
            assert s.sum() == n and s.max() - s.min() <= 1
            assert np.array_equal(dsp, np.concatenate([[0], np.cumsum(s)[:-1]]))


def test_m1_closed_form():
    rng = np.random.default_rng(1)
    illum = rng.lognormal(size=(1, 10, 4))
    xyz = np.array([[0.0, 0.6, 0.8]])
    w = np.array([0.7])
    S = orc.scatter_reference(illum, xyz, w, 2.5)
    expect = 1.5 * illum.mean(axis=2)[0] * 0.7 / 2.5
    assert rel(S[0], expect) < 1e-15


def _setup(m=20, n=30, seed=2):
    rng = np.random.default_rng(seed)
    xyz, w, F = orc.fibonacci(m)
    illum = rng.lognormal(size=(m, n, 4))
    return rng, xyz, w, F, illum


def test_isotropic_input():
    rng, xyz, w, F, _ = _setup()
    illum = np.full((20, 30, 4), 3.0)
    S = orc.scatter_reference(illum, xyz, w, F)
    P = orc.phase_matrix(xyz, w, F)
    assert rel(S, 3.0 * P.sum(axis=1)[:, None] * np.ones((1, 30))) < 1e-14
    Pn = orc.phase_matrix(xyz, w, F, normalize=True)
    assert rel(orc.scatter_matrix(illum, Pn), np.full((20, 30), 3.0)) < 1e-14


def test_linearity_and_scaling():
    rng, xyz, w, F, illum = _setup()
    other = rng.lognormal(size=illum.shape)
    a = orc.scatter_reference(illum, xyz, w, F)
    b = orc.scatter_reference(other, xyz, w, F)
    ab = orc.scatter_reference(illum + other, xyz, w, F)
    assert rel(ab, a + b) < 1e-13
    # scaling by 2 is exact in binary floating point
    assert np.array_equal(orc.scatter_reference(2 * illum, xyz, w, F), 2 * a)


def test_cell_permutation_equivariance():
    rng, xyz, w, F, illum = _setup()
    perm = rng.permutation(illum.shape[1])
    a = orc.scatter_reference(illum, xyz, w, F)
    b = orc.scatter_reference(illum[:, perm, :], xyz, w, F)
    assert np.array_equal(b, a[:, perm])


def test_face_permutation():
    rng, xyz, w, F, illum = _setup()
    a = orc.scatter_reference(illum, xyz, w, F)
    b = orc.scatter_reference(illum[:, :, ::-1], xyz, w, F)
    assert rel(a, b) < 1e-14


def test_joint_direction_permutation():
    rng, xyz, w, F, illum = _setup()
    w = rng.uniform(0.5, 1.5, size=len(w))
    perm = rng.permutation(len(w))
    a = orc.scatter_reference(illum, xyz, w, F)
    b = orc.scatter_reference(illum[perm], xyz[perm], w[perm], F)
    # output direction k moves to position perm^-1(k); summation order changes -> rounding only
    assert rel(b, a[perm]) < 1e-13


def test_reciprocity_thomson():
    rng = np.random.default_rng(3)
    v = rng.normal(size=(15, 3))
    xyz = v / np.linalg.norm(v, axis=1, keepdims=True)
    w = rng.uniform(0.2, 2.0, size=15)
    P = orc.phase_matrix(xyz, w, 7.0)
    lhs = w[:, None] * P
    assert rel(lhs, lhs.T) < 1e-14


def test_face_average_matches_reference_helper():
    rng, xyz, w, F, illum = _setup()
    import ctypes as C

    out = np.empty(illum.shape[1])
    orc.load().orc_face_average(illum.ctypes.data, illum.shape[1], 4, 7, out.ctypes.data)
    assert rel(out, illum[7].mean(axis=1)) < 1e-15
    const = np.full_like(illum, 0.3)
    orc.load().orc_face_average(const.ctypes.data, illum.shape[1], 4, 2, out.ctypes.data)
    assert np.all(out == 0.3)


def test_fibonacci_properties():
    xyz, w, F = orc.fibonacci(400)
    assert np.max(np.abs(np.linalg.norm(xyz, axis=1) - 1)) < 1e-15
    assert abs(F - 4 * math.pi) < 1e-12
    assert np.max(np.abs(xyz.mean(axis=0))) < 1e-2
    # the quadrature integrates the Thomson kernel to ~1 (error of the Fibonacci set)
    P = orc.phase_matrix(xyz, w, F)
    assert np.max(np.abs(P.sum(axis=1) - 1.0)) < 5e-3


def test_direction_file_golden(golden_dir):
    xyz, w, F = orc.read_directions(os.path.join(golden_dir, "dirs_M26.txt"))
    d = np.load(os.path.join(golden_dir, "tiny_thomson_N64_M26.npz"))
    assert np.array_equal(xyz, d["xyz"]) and np.array_equal(w, d["w"]) and F == float(d["F"])


# ---------------------------------------------------------------- radiation moments

def _check_moments(got, ref, illum, w, F, areas, volume, tol):
    e, d = moment_scales(illum, w, F, areas, volume)
    for k, v in ref.items():
        if k not in got:
            continue
        scale = d if k.startswith("div") else e
        err = np.abs(np.asarray(got[k]) - v)
        err = err if err.ndim == 1 else err.max(axis=1)
        assert float(np.max(err / scale)) <= tol, k


def test_moments_golden_exact(golden_dir):
    z = np.load(os.path.join(golden_dir, "moments_N7_M14.npz"))
    got = orc.moments(z["illum"], z["xyz"], z["w"], float(z["F"]), z["normals"], z["areas"], z["volume"])
    ref = {k: z[k] for k in orc.MOMENTS}
    _check_moments(got, ref, z["illum"], z["w"], float(z["F"]), z["areas"], z["volume"], 1e-14)


def test_moments_isotropic_six_axis():
    xyz, w, F = six_axis_directions()
    rng = np.random.default_rng(5)
    n = 9
    normals, areas, volume, _ = random_tetra_cells(n, rng)
    c = rng.uniform(0.5, 3.0, size=n)
    illum = np.broadcast_to(c[None, :, None], (6, n, 4)).copy()
    got = orc.moments(illum, xyz, w, F, normals, areas, volume)
    np.testing.assert_allclose(got["energy"], c, rtol=1e-15)
    assert np.max(np.abs(got["stream"])) == 0.0
    np.testing.assert_allclose(got["impuls"], (c[:, None] / 3.0) * np.eye(3).reshape(1, 9), rtol=1e-15, atol=0)
    # face-uniform fields have zero discrete divergence: sum_f n_f A_f = 0 for a closed cell
    _, d = moment_scales(illum, w, F, areas, volume)
    assert np.max(np.abs(got["divstream"]) / d) < 1e-15
    assert np.max(np.abs(got["divimpuls"]).max(axis=1) / d) < 1e-14


def test_moments_divergence_theorem():
    """Linear face flux F(x) = (3x, 0, 0) at face centroids: discrete divergence 3
    (solve_module/illum/Illum_async.cpp:1351-1387 sets the same field)."""
    xyz, w, F = six_axis_directions()
    rng = np.random.default_rng(11)
    normals, areas, volume, cents = random_tetra_cells(50, rng)
    illum = linear_flux_illum(cents, base=2.0, slope=3.0)
    got = orc.moments(illum, xyz, w, F, normals, areas, volume)
    _, d = moment_scales(illum, w, F, areas, volume)
    assert np.max(np.abs(got["divstream"] - 3.0) / d) < 1e-14
    assert np.max(np.abs(got["divimpuls"] - [3.0, 0.0, 0.0]).max(axis=1) / d) < 1e-14
    # the face-averaged flux is the field at the mean face centroid
    np.testing.assert_allclose(got["stream"][:, 0], 3.0 * cents[:, :, 0].mean(axis=1), rtol=1e-13)


def test_moments_permutations(golden_dir):
    z = np.load(os.path.join(golden_dir, "moments_N7_M14.npz"))
    illum, xyz, w, F = z["illum"], z["xyz"], z["w"], float(z["F"])
    normals, areas, volume = z["normals"], z["areas"], z["volume"]
    base = orc.moments(illum, xyz, w, F, normals, areas, volume)
    rng = np.random.default_rng(3)
    pc = rng.permutation(illum.shape[1])
    got = orc.moments(illum[:, pc], xyz, w, F, normals[pc], areas[pc], volume[pc])
    for k in orc.MOMENTS:
        np.testing.assert_array_equal(got[k], base[k][pc])
    pf = rng.permutation(4)
    got = orc.moments(illum[:, :, pf], xyz, w, F, normals[:, pf], areas[:, pf], volume)
    _check_moments(got, base, illum, w, F, areas, volume, 1e-14)
    pj = rng.permutation(illum.shape[0])
    got = orc.moments(illum[pj], xyz[pj], w[pj], F, normals, areas, volume)
    _check_moments(got, base, illum, w, F, areas, volume, 1e-14)


def test_moments_cell_level_input():
    """faces = 1: the face average is the value itself; matches faces = 4 with equal faces."""
    rng = np.random.default_rng(8)
    xyz, w, F = orc.fibonacci(26)
    cell = rng.lognormal(size=(26, 5, 1))
    got1 = orc.moments(cell, xyz, w, F, which=("energy", "stream", "impuls"))
    got4 = orc.moments(np.repeat(cell, 4, axis=2), xyz, w, F, which=("energy", "stream", "impuls"))
    for k in got1:
        np.testing.assert_allclose(got1[k], got4[k], rtol=1e-14, atol=1e-16)


def test_simple_bin_format(tmp_path):
    """int32 count + raw elements (file_module/reader_bin.h:13-31, reader_bin.cpp:49-69)."""
    rng = np.random.default_rng(2)
    vol = rng.uniform(size=11)
    p = str(tmp_path / "volume.bin")
    with open(p, "wb") as f:
        f.write(np.int32(11).tobytes() + vol.tobytes())
    np.testing.assert_array_equal(orc.read_simple_bin(p), vol)
