"""CPU tests of the transport sweep: the host preparation (mesh, face states,
upwind order, vec_x0 records) and the pinning of the sweep oracle
(oracle/sweep_oracle.cpp, a restatement of Illum_async.cpp:725-784 +
ReCalcIllum :472-510 + GetCurIllum :259-430).

What pins the oracle (no reference binary or fixture exists for the sweep):
* closed forms: the node update of GetCurIllum is the exact solution of
  dI/ds = -k I + (Q + beta S) over a segment of length s, so on a single cell
  every out-face node must equal I0 e^{-ks} + E (1 - e^{-ks}), E = (Q + beta
  S) / k, written here in that independent form; with S = 1 and emitting
  boundaries the test-task class is at equilibrium (E = 1) everywhere;
  pure transport (class 3) carries the boundary value unchanged; class 2's
  absorber saturates at Q = 1 for any s > 0;
* order independence: any other upwind order (a random topological sort,
  records permuted with their cells) gives bit-identical face values, which a
  wrong neighbour index, a wrong in/out bit or a missing dependency breaks;
* NaN start: the oracle's coefficient array starts as NaN for every
  direction, so a face read before it is written shows up as NaN;
* the residual: a second sweep with the same scattering returns norm 0; the
  first equals max |old - new| / new computed here with NaN ratios ignored.
"""
import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import scatter as rs
from relativistic_solver_hllc_amd import sweep as sw


def six_axis():
    return np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)


@pytest.fixture(scope="module")
def mesh4():
    v, pairs = sw.mesh_cube_tetra(4)
    xyz, w, F = rs.fibonacci_directions(26)
    dirs = np.vstack([xyz, six_axis(), [[1, 1, 0]] / np.sqrt(2)])
    return v, pairs, dirs, sw.prepare(v, pairs, dirs)


def test_mesh_is_a_closed_conforming_cube():
    for ne in (1, 3, 5):
        v, pairs = sw.mesh_cube_tetra(ne)
        n = 6 * ne ** 3
        assert v.shape == (n, 4, 3) and pairs.shape == (n, 4)
        normals, areas, vol, nodes = sw.tetra_geometry(v)
        assert abs(vol.sum() - 1.0) < 1e-12 and np.all(vol > 0)
        assert np.abs((normals * areas[..., None]).sum(axis=1)).max() < 1e-14   # closed cells
        inner = pairs >= 0
        flat = pairs.reshape(-1)
        ids = np.arange(4 * n)
        assert np.array_equal(flat[flat[inner.reshape(-1)]], ids[inner.reshape(-1)])  # symmetric
        # neighbouring faces: opposite normals, equal areas
        cf = ids[inner.reshape(-1)]
        nb = flat[cf]
        np.testing.assert_allclose(normals.reshape(-1, 3)[cf], -normals.reshape(-1, 3)[nb], atol=1e-15)
        assert (pairs < 0).sum() == 12 * ne * ne and (pairs == sw.BOUND_OUT_SOURCE).sum() == 2 * ne * ne
        # face nodes lie on their face plane
        a = v[:, [1, 0, 0, 0]]
        d = np.einsum("cfk,cfjk->cfj", normals, nodes - a[:, :, None, :])
        assert np.abs(d).max() < 1e-15


def test_prepare_direction_invariants(mesh4):
    v, pairs, dirs, per = mesh4
    normals, _, _, nodes = sw.tetra_geometry(v)
    n = v.shape[0]
    for w, (order, state, s, face) in zip(dirs, per):
        assert sorted(order.tolist()) == list(range(n))
        pos = np.empty(n, np.int64)
        pos[order] = np.arange(n)
        dots = normals @ w
        for c in range(n):
            for f in range(4):
                p = pairs[c, f]
                in_ = (state[c] >> f) & 1
                if dots[c, f] < -1e-9:
                    assert in_
                elif dots[c, f] > 1e-9:
                    assert not in_
                if in_ and p >= 0:
                    assert pos[p // 4] < pos[c]        # upwind
        # records: traced point lies on the named in face and inside the cell
        r = 0
        for c in order:
            for f in range(4):
                if (state[c] >> f) & 1:
                    continue
                for j in range(3):
                    g = face[r]
                    assert (state[c] >> g) & 1
                    x0 = nodes[c, f, j] - s[r] * w / np.linalg.norm(w)
                    assert abs(normals[c, g] @ (x0 - v[c, [1, 0, 0, 0][g]])) < 1e-12
                    T = np.vstack([(v[c, 1:] - v[c, 0]).T])
                    lam = np.linalg.solve(T, x0 - v[c, 0])
                    assert lam.min() > -1e-12 and lam.sum() < 1 + 1e-12 and s[r] >= 0
                    r += 1
        assert r == len(s)


def test_equilibrium_and_pure_transport(mesh4):
    v, pairs, dirs, _ = mesh4
    pb = pairs.copy()
    pb[pb < 0] = sw.BOUND_OUT_SOURCE
    per = sw.prepare(v, pb, dirs)
    M, n = len(dirs), v.shape[0]
    new, norm = orc.sweep(per, pb, np.ones((M, n)), np.full((M, n, 4), 0.5), cls=1)
    assert np.abs(new - 1).max() < 1e-14                 # E = (1 + 0.5 * 1) / 1.5 = 1 = boundary value
    assert norm == pytest.approx(0.5, rel=1e-12)
    new3, _ = orc.sweep(per, pb, np.full((M, n), 7.0), np.ones((M, n, 4)), cls=3)
    assert np.array_equal(new3, np.ones_like(new3))      # k = 0: the boundary value is carried unchanged
    pf = pairs.copy()
    pf[pf < 0] = sw.BOUND_FREE
    perf = sw.prepare(v, pf, dirs)
    new0, _ = orc.sweep(perf, pf, np.full((M, n), 7.0), np.ones((M, n, 4)), cls=3)
    assert np.array_equal(new0, np.zeros_like(new0))
    new2, _ = orc.sweep(per, pb, np.ones((M, n)), np.ones((M, n, 4)), cls=2)
    assert np.abs(new2 - 1).max() < 1e-15                # absorber saturates at Q / alpha = 1


def test_single_cell_closed_form():
    """One tetrahedron with emitting boundaries: every out face equals the mean
    over its nodes of I0 e^{-ks} + E (1 - e^{-ks}) (test task: Q = 1, alpha = 1,
    beta = 1/2); the in faces hold the boundary value 1."""
    v = np.array([[[0.1, 0.0, 0.0], [1.0, 0.2, 0.0], [0.3, 1.1, 0.1], [0.2, 0.3, 0.9]]])
    pairs = np.full((1, 4), sw.BOUND_OUT_SOURCE, np.int32)
    dirs = np.random.default_rng(3).normal(size=(40, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    per = sw.prepare(v, pairs, dirs)
    S = np.random.default_rng(4).uniform(0.1, 3.0, size=(40, 1))
    new, _ = orc.sweep(per, pairs, S, np.ones((40, 1, 4)), cls=1)
    k, Q, beta = 1.5, 1.0, 0.5
    for d, (order, state, s, face) in enumerate(per):
        E = (Q + beta * S[d, 0]) / k
        r = 0
        for f in range(4):
            if (state[0] >> f) & 1:
                assert new[d, 0, f] == 1.0
                continue
            nodes = [np.exp(-k * s[r + j]) * 1.0 + E * (1 - np.exp(-k * s[r + j])) for j in range(3)]
            r += 3
            assert new[d, 0, f] == pytest.approx(sum(nodes) / 3, rel=2e-15)


def _random_upwind_order(n, pairs, order, state, rng):
    """A random topological order of the same dependencies: in faces, and faces out
    on both sides (written by both cells), which keep their relative order."""
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    deps = [[] for _ in range(n)]
    for c in range(n):
        for f in range(4):
            p = pairs[c, f]
            if p < 0:
                continue
            u, g = p // 4, p % 4
            if (state[c] >> f) & 1 or (not (state[u] >> g) & 1 and pos[u] < pos[c]):
                deps[c].append(u)
    indeg = np.array([len(d) for d in deps])
    users = [[] for _ in range(n)]
    for c in range(n):
        for u in deps[c]:
            users[u].append(c)
    ready = [c for c in range(n) if indeg[c] == 0]
    out = []
    while ready:
        c = ready.pop(rng.integers(len(ready)))
        out.append(c)
        for x in users[c]:
            indeg[x] -= 1
            if indeg[x] == 0:
                ready.append(x)
    return np.array(out, np.int32)


def _reorder(order, state, s, face, new_order):
    """Records of the same cells re-listed in new_order's traversal order."""
    counts = np.array([3 * (4 - bin(int(st)).count("1")) for st in state])
    start = np.zeros(len(state), np.int64)
    start[order] = np.concatenate([[0], np.cumsum(counts[order])[:-1]])
    idx = np.concatenate([np.arange(start[c], start[c] + counts[c]) for c in new_order])
    return new_order, state, s[idx], face[idx]


def test_order_independence_and_residual(mesh4):
    v, pairs, dirs, per = mesh4
    # Fibonacci and axis directions: the latter have faces out on both sides
    M, n = len(dirs), v.shape[0]
    rng = np.random.default_rng(9)
    S = rng.lognormal(size=(M, n))
    old = rng.lognormal(size=(M, n, 4))
    for cls in (1, 10):
        a, norm_a = orc.sweep(per, pairs, S, old, cls=cls)
        assert not np.isnan(a).any()
        per2 = [_reorder(*d, _random_upwind_order(n, pairs, d[0], d[1], rng)) for d in per]
        assert any(not np.array_equal(p[0], q[0]) for p, q in zip(per, per2))
        b, norm_b = orc.sweep(per2, pairs, S, old, cls=cls)
        assert np.array_equal(a, b) and norm_a == norm_b
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = np.abs(old - a) / a
        assert norm_a == np.nanmax(ratio)
        again, norm2 = orc.sweep(per, pairs, S, a, cls=cls)
        assert np.array_equal(again, a) and norm2 == 0.0


def test_oracle_flags_malformed_input(mesh4):
    v, pairs, dirs, per = mesh4
    M, n = 2, v.shape[0]
    order, state, s, face = per[0]
    bad = [(order, state, s[:-3], face[:-3]), per[1]]
    _, norm = orc.sweep(bad, pairs, np.ones((M, n)), np.ones((M, n, 4)), cls=1)
    assert norm == -2.0


def test_direction_files_roundtrip(tmp_path, mesh4):
    v, pairs, dirs, per = mesh4
    for k in (0, 5):
        sw.write_direction_files(str(tmp_path), k, *per[k])
        back = sw.read_direction_files(str(tmp_path), k)
        for a, b in zip(per[k], back):
            assert np.array_equal(a, b)
    raw = np.fromfile(tmp_path / "LocX05.bin", np.uint8)
    assert raw[:4].view(np.int32)[0] == len(per[5][2]) and raw.size == 4 + 48 * len(per[5][2])
