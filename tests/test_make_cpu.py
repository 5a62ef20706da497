"""CPU tests of the reference's sweep preparation (its MAKE step) in the
reference's own numbering: rsh_mesh_vtk_arrays, rsh_make_nodes,
rsh_make_direction and the file driver run_make_module, against the oracle
(oracle/make_oracle.cpp, a restatement of make_illum_struct/
struct_short_characteristics_main.cpp:128-165, MakeArrayX, GetNodes,
CalculateNodeValue and CalculateIllumeOnInnerFace).

What pins the oracle (the reference ships no mesh or LocX0 file):
* closed forms for the node placement: node j of VTK face f is at weight 2/3
  on one point of the face and 1/6 on the other two, in the order written
  here as a table of barycentric weights (a transposed plane rotation or a
  swapped face case moves points off these positions);
* a unit right tetrahedron (points e0, e1, e2, origin: local coordinates equal
  global ones) with directions whose characteristics are traced by hand:
  distance to the in-face plane, the in face that is hit, and the local point
  on it, including the inclined face's plane coordinates ((l1 - l0)/sqrt 2,
  l2 sqrt(3/2)) and (0, 0) behind a boundary face;
* invariants on random tetrahedra and directions: x - s omega lies on the
  named in face's plane strictly inside its triangle, no lower-numbered in face
  holds the point, and the local point is the in face's pair of barycentric
  coordinates, computed here by numpy's solver;
* the sweep itself: records of the reference's placement and of this
  library's own placement give the same face intensities through the sweep
  oracle (the two place a face's nodes at the same points in another order).
"""
import os

import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import scatter as rs
from relativistic_solver_hllc_amd import sweep as sw
from relativistic_solver_hllc_amd._native import RshError

VTK_FACE = [(0, 1, 3), (1, 2, 3), (2, 0, 3), (0, 2, 1)]
VTK_OPP = [2, 0, 1, 3]  # VTK face f is the face opposite point VTK_OPP[f] (this library's face VTK_OPP[f])
# Barycentric weights on points (0, 1, 2, 3) of node j of VTK face f.
S6, T3 = 1.0 / 6.0, 2.0 / 3.0
NODE_WEIGHTS = np.array([
    [[S6, S6, 0, T3], [T3, S6, 0, S6], [S6, T3, 0, S6]],
    [[0, S6, S6, T3], [0, T3, S6, S6], [0, S6, T3, S6]],
    [[S6, 0, S6, T3], [T3, 0, S6, S6], [S6, 0, T3, S6]],
    [[S6, S6, T3, 0], [S6, T3, S6, 0], [T3, S6, S6, 0]],
])


def random_tetra(rng, count):
    """Random non-degenerate tetrahedra [count, 4, 3]."""
    v = rng.normal(size=(count, 4, 3))
    vol = np.einsum("ci,ci->c", v[:, 1] - v[:, 0], np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))
    return v[np.abs(vol) > 0.1]


def unit_right_tetra():
    return np.array([[[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [0, 0, 0]]])


def test_vtk_arrays_layout():
    v, pairs = sw.mesh_cube_tetra(2)
    vm, faces, normals, pv = sw.mesh_vtk_arrays(v, pairs)
    n = len(v)
    np.testing.assert_array_equal(vm[:, :, :3], v)
    np.testing.assert_array_equal(vm[:, :, 3], np.ones((n, 4)))
    for f in range(4):
        np.testing.assert_array_equal(faces[:, f], v[:, list(VTK_FACE[f])])
        opp = v[:, VTK_OPP[f]]
        # outward: the opposite point is behind the face
        assert np.all(np.einsum("ci,ci->c", normals[:, f], opp - faces[:, f, 0]) < 0)
        np.testing.assert_allclose(np.linalg.norm(normals[:, f], axis=1), 1.0, rtol=1e-15)
    # renumbered pairs are symmetric and join faces with the same three points
    flat = pv.reshape(-1)
    inner = np.nonzero(flat >= 0)[0]
    np.testing.assert_array_equal(flat[flat[inner]], inner)
    fa = faces.reshape(-1, 3, 3)
    for fid in inner[:200]:
        a = sorted(map(tuple, fa[fid]))
        b = sorted(map(tuple, fa[flat[fid]]))
        assert a == b
    assert np.array_equal(np.sort(flat[flat < 0]), np.sort(pairs.reshape(-1)[pairs.reshape(-1) < 0]))


def test_nodes_closed_form_oracle_and_product():
    rng = np.random.default_rng(11)
    v = random_tetra(rng, 300)
    vm, _, _, _ = sw.mesh_vtk_arrays(v, np.full((len(v), 4), -1, np.int32))
    want = np.einsum("fjp,cpi->cfji", NODE_WEIGHTS, v)
    scale = np.abs(v).max()
    assert np.abs(orc.make_nodes(vm) - want).max() < 4e-16 * scale * 8
    assert np.abs(sw.make_nodes(vm) - want).max() < 4e-16 * scale * 8


def test_oracle_single_tetra_inclined_in_face():
    """omega = -(1,1,1)/sqrt 3: the inclined face 3 (x + y + z = 1) is the only in face; every node of faces
    0, 1, 2 is traced to it at s = (1 - sum x)/sqrt 3, x0 = x + (1 - sum x)/3 (1,1,1)."""
    v = unit_right_tetra()
    w = -np.ones(3) / np.sqrt(3.0)
    for neighbour in (7, sw.BOUND_FREE):  # an interior neighbour (any id >= 0) or a boundary face
        pairs = np.array([[sw.BOUND_FREE, sw.BOUND_FREE, sw.BOUND_FREE, neighbour]], np.int32)
        vm, faces, normals, _ = sw.mesh_vtk_arrays(v, np.full((1, 4), -1, np.int32))
        for impl in ("oracle", "product"):
            if impl == "oracle":
                rc, st, s, f, x0 = orc.make_direction(vm, faces, normals, pairs, [0], w)
                assert rc == 0
            else:
                st, s, f, x0 = sw.make_direction(vm, faces, normals, pairs, [0], w)
            assert st[0] == 0b1000 and len(s) == 9 and np.all(f == 3)
            r = 0
            for face in range(3):
                for j in range(3):
                    x = NODE_WEIGHTS[face, j, :3]  # local = global coordinates on this tetrahedron
                    sx = x.sum()
                    assert abs(s[r] - (1 - sx) / np.sqrt(3.0)) < 1e-15
                    p = x + (1 - sx) / 3.0
                    if neighbour >= 0:
                        want = ((p[1] - p[0]) / np.sqrt(2.0), p[2] * np.sqrt(1.5))
                    else:
                        want = (0.0, 0.0)
                    assert np.abs(x0[r] - want).max() < 1e-15, (impl, face, j, x0[r], want)
                    r += 1


def test_oracle_single_tetra_axis_in_faces():
    """A generic direction into the positive octant: faces 0, 1, 2 (z = 0, x = 0, y = 0) are in faces and each
    node of the inclined face is traced back to the coordinate plane x_i = 0 with the smallest x_i / w_i."""
    v = unit_right_tetra()
    vm, faces, normals, _ = sw.mesh_vtk_arrays(v, np.full((1, 4), -1, np.int32))
    pairs = np.array([[5, 6, 9, sw.BOUND_FREE]], np.int32)
    plane_face = {2: 0, 0: 1, 1: 2}  # coordinate i = 0 is VTK face plane_face[i]
    rng = np.random.default_rng(3)
    for _ in range(20):
        w = rng.uniform(0.2, 1.0, 3)
        w /= np.linalg.norm(w)
        for impl in ("oracle", "product"):
            if impl == "oracle":
                rc, st, s, f, x0 = orc.make_direction(vm, faces, normals, pairs, [0], w)
                assert rc == 0
            else:
                st, s, f, x0 = sw.make_direction(vm, faces, normals, pairs, [0], w)
            assert st[0] == 0b0111 and len(s) == 3
            for j in range(3):
                x = NODE_WEIGHTS[3, j, :3]
                i = int(np.argmin(x / w))
                t = x[i] / w[i]
                p = x - t * w
                assert f[j] == plane_face[i]
                assert abs(s[j] - t) < 1e-15
                want = {0: (p[0], p[1]), 1: (p[1], p[2]), 2: (p[0], p[2])}[plane_face[i]]
                assert np.abs(x0[j] - want).max() < 1e-15


def _check_records(v, order, w, st, s, f, x0, pairs):
    """Invariants of a record set in the reference's numbering, by numpy geometry."""
    nodes = np.einsum("fjp,cpi->cfji", NODE_WEIGHTS, v)
    r = 0
    for c in order:
        for face in range(4):
            if (st[c] >> face) & 1:
                continue
            for j in range(3):
                x = nodes[c, face, j]
                p = x - s[r] * w
                g = int(f[r])
                assert (st[c] >> g) & 1
                tri = v[c, list(VTK_FACE[g])]
                # barycentric coordinates of p in the in face's triangle
                mat = np.vstack([tri.T, np.ones(3)])
                lam = np.linalg.lstsq(mat, np.append(p, 1.0), rcond=None)[0]
                assert np.abs(mat @ lam - np.append(p, 1.0)).max() < 1e-12  # on the plane
                assert lam.min() > 0
                for h in range(g):  # no lower-numbered in face holds the line's crossing
                    if not (st[c] >> h) & 1:
                        continue
                    th = v[c, list(VTK_FACE[h])]
                    nh = np.cross(th[1] - th[0], th[2] - th[0])
                    q = x + (nh @ (th[0] - x)) / (nh @ w) * w
                    mh = np.vstack([th.T, np.ones(3)])
                    lh = np.linalg.lstsq(mh, np.append(q, 1.0), rcond=None)[0]
                    assert lh.min() < 1e-9
                if pairs[c, g] in (-1, -2, -3, -4):
                    assert np.all(x0[r] == 0)
                else:
                    e = v[c, :3] - v[c, 3]
                    l = np.linalg.solve(e.T, p - v[c, 3])
                    want = {0: (l[0], l[1]), 1: (l[1], l[2]), 2: (l[0], l[2]),
                            3: ((l[1] - l[0]) / np.sqrt(2.0), (1 - l[0] - l[1] + 2 * l[2]) / np.sqrt(6.0))}[g]
                    assert np.abs(x0[r] - want).max() < 1e-9
                r += 1
    assert r == len(s)


def test_random_tetra_oracle_invariants_and_product_parity():
    rng = np.random.default_rng(7)
    v = random_tetra(rng, 400)
    n = len(v)
    pairs = rng.choice([-1, -3, 0, 5], size=(n, 4)).astype(np.int32)
    vm, faces, normals, _ = sw.mesh_vtk_arrays(v, np.full((n, 4), -1, np.int32))
    order = rng.permutation(n).astype(np.int32)
    for _ in range(6):
        w = rng.normal(size=3)
        w /= np.linalg.norm(w)
        rc, st, s, f, x0 = orc.make_direction(vm, faces, normals, pairs, order, w)
        assert rc == 0
        _check_records(v, order, w, st, s, f, x0, pairs)
        pst, ps, pf, px0 = sw.make_direction(vm, faces, normals, pairs, order, w)
        np.testing.assert_array_equal(pst, st)
        np.testing.assert_array_equal(pf, f)
        assert np.abs(ps - s).max() < 1e-12 * max(1.0, s.max())
        assert np.abs(px0 - x0).max() < 1e-11


@pytest.fixture(scope="module")
def cube3():
    v, pairs = sw.mesh_cube_tetra(3)
    xyz, _, _ = rs.fibonacci_directions(26)
    vm, faces, normals, pv = sw.mesh_vtk_arrays(v, pairs)
    own = sw.prepare(v, pairs, xyz[1:])  # direction 0 has y = 0 exactly (below)
    return v, pairs, xyz, vm, faces, normals, pv, own


def test_cube_mesh_product_matches_oracle(cube3):
    v, pairs, xyz, vm, faces, normals, pv, own = cube3
    for k, d in enumerate(own, start=1):
        order = d[0]
        rc, st, s, f, x0 = orc.make_direction(vm, faces, normals, pv, order, xyz[k])
        assert rc == 0
        pst, ps, pf, px0 = sw.make_direction(vm, faces, normals, pv, order, xyz[k])
        np.testing.assert_array_equal(pst, st)
        np.testing.assert_array_equal(pf, f)
        assert np.abs(ps - s).max() < 1e-14 and np.abs(px0 - x0).max() < 1e-13
        # the same face states as this library's own preparation, renumbered
        mine = np.zeros_like(st)
        for face in range(4):
            mine |= ((st >> face) & 1) << VTK_OPP[face]
        np.testing.assert_array_equal(mine, d[1])
        # per cell, the same distances as this library's own placement (same points, other order)
        pos = posm = 0
        for c in order:
            cnt = 3 * (4 - bin(int(st[c])).count("1"))
            a, b = np.sort(ps[pos:pos + cnt]), np.sort(d[2][posm:posm + cnt])
            assert cnt == 0 or np.abs(a - b).max() < 1e-14
            pos += cnt
            posm += cnt


def test_grazing_direction_stops_like_the_reference(cube3):
    """Fibonacci direction 0 has y = 0: faces parallel to it are out faces whose nodes trace back onto an
    edge of two in faces, so the strict in-triangle test holds for neither. The reference stops there
    ("Not found inner face"); the oracle flags it and the library refuses the direction."""
    v, pairs, xyz, vm, faces, normals, pv, own = cube3
    assert xyz[0, 1] == 0.0
    order = sw.prepare_direction(v, pairs, xyz[0])[0]
    rc = orc.make_direction(vm, faces, normals, pv, order, xyz[0])[0]
    assert rc == 1
    with pytest.raises(RshError, match="Not found inner face"):
        sw.make_direction(vm, faces, normals, pv, order, xyz[0])
    with pytest.raises(RshError, match="permutation"):
        sw.make_direction(vm, faces, normals, pv, np.zeros(len(v), np.int32), xyz[1])


@pytest.mark.parametrize("cls", [1, 0])
def test_reference_placement_sweeps_like_own_placement(cube3, cls):
    """Sweep oracle on the reference's records (VTK numbering, its nodes for class 0) and on this library's
    records: the same face intensities, face f of the first being face VTK_OPP[f] of the second."""
    v, pairs, xyz, vm, faces, normals, pv, own = cube3
    n, m = len(v), len(own)
    rng = np.random.default_rng(5)
    S = rng.uniform(0.5, 2.0, (m, n))
    old = np.ones((m, n, 4))
    _, _, _, my_nodes = sw.tetra_geometry(v)
    mine, _ = orc.sweep(own, pairs, S, old, cls=cls, nodes=my_nodes)
    ref_dirs = []
    for k, d in enumerate(own, start=1):
        st, s, f, _ = sw.make_direction(vm, faces, normals, pv, d[0], xyz[k])
        ref_dirs.append((d[0], st, s, f))
    ref, _ = orc.sweep(ref_dirs, pv, S, old, cls=cls, nodes=sw.make_nodes(vm))
    assert np.isfinite(ref).all()
    back = ref[:, :, [1, 2, 0, 3]]  # this library's face g is VTK face [1, 2, 0, 3][g]
    assert np.abs(back - mine).max() <= 1e-13 * np.abs(mine).max()


def test_run_make_module_files(cube3, tmp_path):
    v, pairs, xyz, vm, faces, normals, pv, own = cube3
    base, geo = str(tmp_path) + "/", str(tmp_path / "illum_geo_")
    sw.write_make_inputs(base, vm, faces, normals, pv)
    for k, d in enumerate(own, start=1):
        rs.write_simple_bin(f"{base}graph{k}.bin", d[0])
    dirs = xyz.copy()
    assert sw.run_make_module(base, geo, dirs, base + "graph", a=1, b=len(xyz)) == len(xyz) - 1
    x = rs.read_simple_bin(geo + "X.bin", np.float64, 36).reshape(-1, 4, 3, 3)
    np.testing.assert_array_equal(x, sw.make_nodes(vm))
    assert rs.read_simple_bin(geo + "ResBound.bin").size == 0
    for k in (1, 13, len(xyz) - 1):
        st, s, f, x0 = sw.make_direction(vm, faces, normals, pv, own[k - 1][0], xyz[k])
        np.testing.assert_array_equal(rs.read_simple_bin(f"{geo}state_face{k}.bin", np.int32), st)
        fs, ff, fx0 = sw.read_locx0(f"{geo}LocX0{k}.bin")
        np.testing.assert_array_equal(fs, s)
        np.testing.assert_array_equal(ff, f)
        np.testing.assert_array_equal(fx0, x0)
    assert os.path.getsize(f"{geo}LocX0{1}.bin") == 4 + 48 * len(own[0][2])
