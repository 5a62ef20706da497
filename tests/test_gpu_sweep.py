"""GPU parity of the transport sweep (rsh_sweep_* through the C ABI) against the
sequential oracle (oracle/sweep_oracle.cpp, Illum_async.cpp:725-784 +
ReCalcIllum :472-510).

The GPU walks the cells of each direction level by level; the oracle walks the
reference's sorted_id_cell order. The face value arithmetic is the same
expression sequence, so the results differ only by the rounding of exp (device
math library against glibc) and FMA contraction, propagated along the sweep:
tolerance |gpu - oracle| <= 1e-12 * max(|oracle|, 1e-300) per face, and the
per-direction residuals within 1e-10 relative (a ratio of such values).
"""
import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs
from relativistic_solver_hllc_amd import sweep as sw

pytestmark = pytest.mark.gpu

TOL = 1e-12


def six_axis():
    return np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)


def gpu_sweep(v, pairs, per, S, old, cls, nodes=None):
    M, n = S.shape
    with sw.Sweep(n, M, pairs, nodes) as s:
        s.set_all(per)
        d_S, d_old = D.DeviceArray.from_host(S), D.DeviceArray.from_host(old)
        d_new, d_norm = D.DeviceArray((M, n, 4)), D.DeviceArray((M,))
        d_new.fill(-99.0)
        s.run(cls, d_S, d_new, d_old, d_norm)
        D.synchronize()
        out, norm = d_new.to_host(), d_norm.to_host()
        levels = [s.levels(k) for k in range(M)]
        for a in (d_S, d_old, d_new, d_norm):
            a.free()
    return out, norm, levels


def check(got, ref):
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    assert np.isfinite(got).all()
    assert float(err.max()) <= TOL, float(err.max())


@pytest.mark.parametrize("ne,m,cls", [(2, 14, 1), (3, 26, 1), (4, 26, 10), (4, 26, 2), (4, 26, 3), (5, 50, 0),
                                      (6, 26, 1)])
def test_sweep_matches_oracle(gpu, ne, m, cls):
    v, pairs = sw.mesh_cube_tetra(ne)
    n = v.shape[0]
    xyz, _, _ = rs.fibonacci_directions(m)
    dirs = np.vstack([xyz, six_axis()])        # axis directions: faces out on both sides
    M = len(dirs)
    per = sw.prepare(v, pairs, dirs)
    rng = np.random.default_rng(ne * 100 + m)
    S = rng.lognormal(size=(M, n))
    old = rng.lognormal(size=(M, n, 4))
    nodes = sw.tetra_geometry(v)[3] if cls == 0 else None
    ref, ref_norm = orc.sweep(per, pairs, S, old, cls=cls, nodes=nodes)
    assert ref_norm != -2.0
    got, norm, levels = gpu_sweep(v, pairs, per, S, old, cls, nodes)
    check(got, ref)
    # per-direction residuals: the oracle's is the max over directions
    assert max(levels) > 1
    assert norm.max() == pytest.approx(ref_norm, rel=1e-10) if np.isfinite(ref_norm) else np.isinf(norm.max())


@pytest.mark.parametrize("cls", [1, 0])
def test_sweep_on_reference_placement_matches_oracle(gpu, cls):
    """The sweep fed the records of the reference's own MAKE step (its VTK face numbering, node placement and
    order: rsh_mesh_vtk_arrays, rsh_make_nodes, rsh_make_direction) against the oracle on the same records,
    and, renumbered, against the sweep on this library's own records."""
    v, pairs = sw.mesh_cube_tetra(5)
    n = v.shape[0]
    # Seeded random directions: a Fibonacci set holds directions with a zero component (y = 0 at k = 0,
    # z = 0 at k = (M - 1) / 2 for odd M), along which the reference's in-face search stops.
    dirs = np.random.default_rng(2024).normal(size=(48, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    M = len(dirs)
    own = sw.prepare(v, pairs, dirs)
    vm, faces, normals, pv = sw.mesh_vtk_arrays(v, pairs)
    ref_per = []
    for k in range(M):
        st, s, f, _ = sw.make_direction(vm, faces, normals, pv, own[k][0], dirs[k])
        ref_per.append((own[k][0], st, s, f))
    rng = np.random.default_rng(77 + cls)
    S = rng.lognormal(size=(M, n))
    old = rng.lognormal(size=(M, n, 4))
    nodes_ref = sw.make_nodes(vm) if cls == 0 else None
    ref, ref_norm = orc.sweep(ref_per, pv, S, old, cls=cls, nodes=nodes_ref)
    assert ref_norm != -2.0
    got, norm, levels = gpu_sweep(v, pv, ref_per, S, old, cls, nodes_ref)
    check(got, ref)
    assert norm.max() == pytest.approx(ref_norm, rel=1e-10) if np.isfinite(ref_norm) else np.isinf(norm.max())
    mine, _, _ = gpu_sweep(v, pairs, own, S, old, cls, sw.tetra_geometry(v)[3] if cls == 0 else None)
    check(got[:, :, [1, 2, 0, 3]], mine)  # face g of this library's numbering is VTK face [1, 2, 0, 3][g]


def test_sweep_order_and_files(gpu, tmp_path):
    """The same directions loaded from reference-format files (graph, state_face,
    LocX0) and from a reversed-within-level order give identical GPU results."""
    v, pairs = sw.mesh_cube_tetra(4)
    n = v.shape[0]
    xyz, _, _ = rs.fibonacci_directions(20)
    per = sw.prepare(v, pairs, xyz)
    for k, d in enumerate(per):
        sw.write_direction_files(str(tmp_path), k, *d)
    per_f = [sw.read_direction_files(str(tmp_path), k) for k in range(len(per))]
    rng = np.random.default_rng(2)
    S = rng.lognormal(size=(20, n))
    old = rng.lognormal(size=(20, n, 4))
    a, na, _ = gpu_sweep(v, pairs, per, S, old, 1)
    b, nb, _ = gpu_sweep(v, pairs, per_f, S, old, 1)
    assert np.array_equal(a, b) and np.array_equal(na, nb)


def test_sweep_rejects_bad_input(gpu):
    v, pairs = sw.mesh_cube_tetra(2)
    n = v.shape[0]
    xyz, _, _ = rs.fibonacci_directions(4)
    per = sw.prepare(v, pairs, xyz)
    order, state, s, face = per[0]
    with sw.Sweep(n, 4, pairs) as sp:
        with pytest.raises(rs.N.RshError):
            sp.set_direction(0, order[::-1].copy(), state, s, face)          # not upwind
        bad_face = face.copy()
        bad_face[0] = [g for g in range(4) if not (state[order[0]] >> g) & 1][0]
        with pytest.raises(rs.N.RshError):
            sp.set_direction(0, order, state, s, bad_face)                   # record names an out face
        with pytest.raises(rs.N.RshError):
            sp.set_direction(0, order, state, s[:-3], face[:-3])             # record count
        sp.set_all(per[:3])
        d_S, d_new = D.DeviceArray((4, n)), D.DeviceArray((4, n, 4))
        with pytest.raises(rs.N.RshError):
            sp.run(1, d_S, d_new)                                            # direction 3 not set
        sp.set_direction(3, *per[3])
        with pytest.raises(rs.N.RshError):
            sp.run(0, d_S, d_new)                                            # class 0 without nodes
        with pytest.raises(rs.N.RshError):
            sp.run(11, d_S, d_new)
        # illum not 16-byte aligned (rejected on the host; the array has room
        # for the shifted extent all the same)
        d_pad = D.DeviceArray((4 * n * 4 + 2,))
        with pytest.raises(rs.N.RshError):
            sp.run(1, d_S, d_pad.ptr + 8)
        d_pad.free()
        d_S.fill(1.0)
        sp.run(1, d_S, d_new)
        D.synchronize()
        d_S.free()
        d_new.free()
    bad = pairs.copy()
    bad[0, 0] = -7
    with pytest.raises(rs.N.RshError):
        sw.Sweep(n, 4, bad)


def test_transport_iteration_matches_oracle(gpu):
    """The reference's illumination iteration on the GPU: sweep (S -> I) then the
    FP64 scattering integral (I -> S), three times, against the oracle pipeline
    (orc.sweep + the GetS restatement); Thomson phase, F = 4 pi."""
    v, pairs = sw.mesh_cube_tetra(3)
    n = v.shape[0]
    m = 26
    xyz, w, F = rs.fibonacci_directions(m)
    per = sw.prepare(v, pairs, xyz)
    S = np.zeros((m, n))
    I_ref = np.ones((m, n, 4))
    with sw.Sweep(n, m, pairs) as sp, rs.DirectionSet(xyz, w, F) as ds:
        sp.set_all(per)
        d_S = D.DeviceArray.from_host(S)
        d_a, d_b, d_norm = D.DeviceArray.from_host(I_ref), D.DeviceArray((m, n, 4)), D.DeviceArray((m,))
        for it in range(3):
            I_new, norm_ref = orc.sweep(per, pairs, S, I_ref, cls=1)
            sp.run(1, d_S, d_b, d_a, d_norm)
            ds.scatter_f64(d_b, 4, n, 0, m, d_S)
            D.synchronize()
            got = d_b.to_host()
            check_tol = np.abs(got - I_new) / np.maximum(np.abs(I_new), 1e-300)
            assert check_tol.max() <= 1e-11 * (it + 1), (it, check_tol.max())
            S = orc.scatter_reference(I_new, xyz, w, F)
            got_S = d_S.to_host()
            assert (np.abs(got_S - S) / np.maximum(np.abs(S), 1e-300)).max() <= 1e-11 * (it + 1)
            assert d_norm.to_host().max() == pytest.approx(norm_ref, rel=1e-9)
            I_ref = I_new
            d_a, d_b = d_b, d_a
        for a in (d_S, d_a, d_b, d_norm):
            a.free()


@pytest.mark.parametrize("cls", [1, 0])
def test_sweep_multipass_levels_match_oracle(gpu, cls):
    """A mesh whose levels hold more cells than one pass (512), so that passes of
    one level read upstream values from the LDS ring of earlier passes and from
    global memory within the level, and levels end on their last pass's ring.
    Class 0 (the sphere-source test task) also carries its per-node source flags
    in the out faces' ring-code fields: on this mesh many nodes lie in the source
    and share ring words with in-face ring codes across several passes."""
    v, pairs = sw.mesh_cube_tetra(24)
    n = v.shape[0]
    dirs = np.random.default_rng(31).normal(size=(6, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    per = sw.prepare(v, pairs, dirs)
    rng = np.random.default_rng(32)
    S = rng.lognormal(size=(6, n))
    old = rng.lognormal(size=(6, n, 4))
    nodes = sw.tetra_geometry(v)[3] if cls == 0 else None
    if cls == 0:
        centre = np.array([0.5, 0.0, 0.0])
        inside = np.linalg.norm(nodes.reshape(-1, 3) - centre, axis=1) < 0.05
        assert inside.sum() > 100     # many source nodes
    ref, ref_norm = orc.sweep(per, pairs, S, old, cls=cls, nodes=nodes)
    got, norm, levels = gpu_sweep(v, pairs, per, S, old, cls, nodes)
    assert n / min(levels) > 512          # some level of every direction spans several passes
    check(got, ref)
    assert norm.max() == pytest.approx(ref_norm, rel=1e-10)


def test_sweep_all_directions_at_196608_cells(gpu):
    """The bench's sweep workload, every direction: 6 x 32^3 = 196,608 tetrahedra,
    M = 400 Fibonacci directions, class 1, against the sequential oracle. Every
    face value of every direction is compared, so the LDS-ring passes, the
    reads of older upstream values from global memory and the store ordering of
    faces out on both sides are all exercised at the size that is timed."""
    v, pairs = sw.mesh_cube_tetra(32)
    n, m = v.shape[0], 400
    xyz, _, _ = rs.fibonacci_directions(m)
    per = sw.prepare(v, pairs, xyz, threads=16)
    rng = np.random.default_rng(0x5EED + 7)
    S = rng.lognormal(size=(m, n))
    old = np.ones((m, n, 4))
    ref, ref_norm = orc.sweep(per, pairs, S, old, cls=1, threads=16)
    got, norm, levels = gpu_sweep(v, pairs, per, S, old, 1)
    assert n == 196608 and len(levels) == m and max(levels) > 100
    check(got, ref)
    assert norm.max() == pytest.approx(ref_norm, rel=1e-10)
