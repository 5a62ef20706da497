"""Transport sweep (short characteristics): Python mirror of the sweep part of
the C ABI (include/rsh_scatter.h), the reference's sweep input formats, and
the synthetic tetrahedral mesh.

Reference: the direction loop of MPI_CalculateIllumAsync
(solve_module/illum/Illum_async.cpp:725-784), ReCalcIllum (:472-510), the
inputs of ReadIllumGeometry (file_module/reader_bin.cpp:164-203) and their
producer make_illum_struct (struct_short_characteristics_main.cpp:128-165).
"""
from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor
from typing import List, Optional, Tuple

import numpy as np

from . import _native as N

BOUND_FREE, BOUND_LOCK, BOUND_OUT_SOURCE, BOUND_INNER_SOURCE = -1, -2, -3, -4
CLASSES = (0, 1, 2, 3, 10)

# cell_local (global_def.h:101-108) as written by WriteSimpleFileBin on x86-64
# with Eigen's default 16-byte alignment of Vector2d: s at 0, x0 at 16,
# in_face_id at 32, 48 bytes per record.
CELL_LOCAL = np.dtype({"names": ["s", "x0", "in_face"], "formats": ["<f8", ("<f8", 2), "u1"],
                       "offsets": [0, 16, 32], "itemsize": 48})


def _p(a):
    return None if a is None else a.ctypes.data


def mesh_cube_tetra(n_edge: int) -> Tuple[np.ndarray, np.ndarray]:
    """Unit cube in n_edge^3 cubes of 6 tetrahedra: vertices [N,4,3], pairs [N,4] int32."""
    n = 6 * n_edge ** 3
    v = np.empty((n, 4, 3))
    pairs = np.empty((n, 4), np.int32)
    N.call("rsh_mesh_cube_tetra", n_edge, v.ctypes.data, pairs.ctypes.data)
    return v, pairs


def tetra_geometry(vertices: np.ndarray):
    """(normals [N,4,3], areas [N,4], volume [N], nodes [N,4,3,3]) of tetrahedra vertices [N,4,3]."""
    v = np.ascontiguousarray(vertices, np.float64)
    n = v.shape[0]
    normals, areas, volume, nodes = np.empty((n, 4, 3)), np.empty((n, 4)), np.empty(n), np.empty((n, 4, 3, 3))
    N.call("rsh_tetra_geometry", n, v.ctypes.data, normals.ctypes.data, areas.ctypes.data, volume.ctypes.data,
           nodes.ctypes.data)
    return normals, areas, volume, nodes


def prepare_direction(vertices: np.ndarray, pairs: np.ndarray, omega) -> Tuple[np.ndarray, ...]:
    """(order [N] int32, face_state [N] int32, s [R], in_face [R] uint8) for one unit direction."""
    v = np.ascontiguousarray(vertices, np.float64)
    pr = np.ascontiguousarray(pairs, np.int32)
    w = np.ascontiguousarray(omega, np.float64)
    n = v.shape[0]
    order, state = np.empty(n, np.int32), np.empty(n, np.int32)
    cap = 9 * n
    s, face = np.empty(cap), np.empty(cap, np.uint8)
    count = C.c_int64()
    N.call("rsh_sweep_prepare_direction", n, v.ctypes.data, pr.ctypes.data, w.ctypes.data, order.ctypes.data,
           state.ctypes.data, s.ctypes.data, face.ctypes.data, cap, C.byref(count))
    r = count.value
    return order, state, s[:r].copy(), face[:r].copy()


def prepare(vertices, pairs, dirs, threads: int = 16) -> List[Tuple[np.ndarray, ...]]:
    """prepare_direction for every row of dirs [M,3], on a thread pool (the C call drops the GIL)."""
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda w: prepare_direction(vertices, pairs, w), dirs))


# ---------------------------------------------------------------- the reference's MAKE step

def mesh_vtk_arrays(vertices, pairs):
    """The reference's input arrays for a mesh given in this library's numbering (face f opposite vertex f),
    in its VTK face numbering (face f = points (0,1,3), (1,2,3), (2,0,3), (0,2,1)): (vertex_matrix [N,4,4]
    (vertex.bin; row j = point j as (x, y, z, 1), Eigen's column-major Matrix4d), faces [N,4,3,3] (grid.bin),
    outward unit normals [N,4,3] (normals.bin), pairs [N,4] int32 (pairs.bin))."""
    v = np.ascontiguousarray(vertices, np.float64)
    pr = np.ascontiguousarray(pairs, np.int32)
    n = v.shape[0]
    vm, faces, normals, pv = np.empty((n, 4, 4)), np.empty((n, 4, 3, 3)), np.empty((n, 4, 3)), np.empty((n, 4), np.int32)
    N.call("rsh_mesh_vtk_arrays", n, v.ctypes.data, pr.ctypes.data, vm.ctypes.data, faces.ctypes.data,
           normals.ctypes.data, pv.ctypes.data)
    return vm, faces, normals, pv


def make_nodes(vertex_matrix) -> np.ndarray:
    """Face nodes [N,4,3,3] in the reference's placement and order (X.bin; MakeArrayX)."""
    vm = np.ascontiguousarray(vertex_matrix, np.float64)
    nodes = np.empty((vm.shape[0], 4, 3, 3))
    N.call("rsh_make_nodes", vm.shape[0], vm.ctypes.data, nodes.ctypes.data)
    return nodes


def make_direction(vertex_matrix, faces, normals, pairs, order, omega) -> Tuple[np.ndarray, ...]:
    """One direction of the reference's MAKE loop: (face_state [N] int32, s [R], in_face [R] uint8, x0 [R,2])
    for the cells in `order` (graph<k>.bin), in the reference's numbering."""
    vm = np.ascontiguousarray(vertex_matrix, np.float64)
    fc, nr = np.ascontiguousarray(faces, np.float64), np.ascontiguousarray(normals, np.float64)
    pr, od = np.ascontiguousarray(pairs, np.int32), np.ascontiguousarray(order, np.int32)
    w = np.ascontiguousarray(omega, np.float64)
    n = vm.shape[0]
    state = np.empty(n, np.int32)
    cap = 9 * n
    s, x0, face = np.empty(cap), np.empty((cap, 2)), np.empty(cap, np.uint8)
    count = C.c_int64()
    N.call("rsh_make_direction", n, vm.ctypes.data, fc.ctypes.data, nr.ctypes.data, pr.ctypes.data, od.ctypes.data,
           w.ctypes.data, state.ctypes.data, s.ctypes.data, x0.ctypes.data, face.ctypes.data, cap, C.byref(count))
    r = count.value
    return state, s[:r].copy(), face[:r].copy(), x0[:r].copy()


def make(vertex_matrix, faces, normals, pairs, orders, dirs, threads: int = 16):
    """make_direction for every direction (row of dirs [M,3], with its order) on a thread pool:
    a list of (order, face_state, s, in_face, x0)."""
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)

    def one(k):
        st, s, f, x0 = make_direction(vertex_matrix, faces, normals, pairs, orders[k], dirs[k])
        return np.asarray(orders[k], np.int32), st, s, f, x0
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(one, range(len(dirs))))


def write_make_inputs(base_dir: str, vertex_matrix, faces, normals, pairs) -> None:
    """vertex.bin, grid.bin, normals.bin and pairs.bin as the reference's MAKE step reads them
    (ReadSimpleFileBin, ReadCellFaces, ReadNormalFile: an int32 cell count then the records)."""
    from .scatter import write_simple_bin
    n = len(vertex_matrix)
    write_simple_bin(os.path.join(base_dir, "vertex.bin"), np.ascontiguousarray(vertex_matrix, np.float64), 16)
    write_simple_bin(os.path.join(base_dir, "grid.bin"), np.ascontiguousarray(faces, np.float64).reshape(n, 36), 36)
    write_simple_bin(os.path.join(base_dir, "normals.bin"), np.ascontiguousarray(normals, np.float64).reshape(n, 12),
                     12)
    write_simple_bin(os.path.join(base_dir, "pairs.bin"), np.ascontiguousarray(pairs, np.int32).reshape(-1))


def run_make_module(base_dir: str, illum_geo: str, dirs, graph: str, a: int = 0, b: int = 0,
                    threads: int = 16) -> int:
    """The reference's RunMakeModule (make_illum_struct/struct_short_characteristics_main.cpp:14-177) on its
    files: reads pairs.bin, grid.bin, vertex.bin, normals.bin from base_dir and graph<k>.bin (prefix `graph`),
    writes state_face<k>.bin and LocX0<k>.bin for directions [a, b) (b = 0: all), X.bin and ResBound.bin
    (empty: no inner-source faces are traced) under the prefix illum_geo. dirs: [M,3] unit vectors.
    Returns the number of directions written."""
    from .scatter import read_simple_bin, write_simple_bin
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    pairs = read_simple_bin(os.path.join(base_dir, "pairs.bin"), np.int32)
    faces = read_simple_bin(os.path.join(base_dir, "grid.bin"), np.float64, 36)
    vm = read_simple_bin(os.path.join(base_dir, "vertex.bin"), np.float64, 16)
    normals = read_simple_bin(os.path.join(base_dir, "normals.bin"), np.float64, 12)
    n = vm.shape[0]
    if faces.shape[0] != n or normals.shape[0] != n or pairs.size != 4 * n:
        raise ValueError("vertex.bin, grid.bin, normals.bin and pairs.bin disagree on the cell count")
    b = len(dirs) if b == 0 else b
    ks = list(range(a, b))
    orders = {k: read_simple_bin(f"{graph}{k}.bin", np.int32) for k in ks}

    def one(k):
        st, s, f, x0 = make_direction(vm, faces, normals, pairs, orders[k], dirs[k])
        write_simple_bin(f"{illum_geo}state_face{k}.bin", st)
        _write_locx0(f"{illum_geo}LocX0{k}.bin", s, f, x0)
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(one, ks))
    write_simple_bin(f"{illum_geo}X.bin", make_nodes(vm).reshape(n, 36), 36)
    write_simple_bin(f"{illum_geo}ResBound.bin", np.empty(0))
    return len(ks)


# ---------------------------------------------------------------- files

def _write_locx0(path: str, s, in_face, x0=None) -> None:
    from .scatter import write_simple_bin
    rec = np.zeros(len(s), CELL_LOCAL)
    rec["s"], rec["in_face"] = s, in_face
    if x0 is not None:
        rec["x0"] = x0
    write_simple_bin(path, rec.view(np.uint8), CELL_LOCAL.itemsize)


def read_locx0(path: str):
    """LocX0<k>.bin: (s [R], in_face [R] uint8, x0 [R,2])."""
    from .scatter import read_simple_bin
    raw = read_simple_bin(path, np.uint8, CELL_LOCAL.itemsize)
    rec = np.ascontiguousarray(raw).view(CELL_LOCAL).reshape(-1)
    return np.ascontiguousarray(rec["s"]), np.ascontiguousarray(rec["in_face"]), np.ascontiguousarray(rec["x0"])


def write_direction_files(main_dir: str, k: int, order, face_state, s, in_face, x0=None) -> None:
    """graph<k>.bin, state_face<k>.bin (int32) and LocX0<k>.bin (48-byte cell_local records; x0, the local
    point on the in face, zero when not given: the sweep does not read it) in simple-bin format."""
    from .scatter import write_simple_bin
    write_simple_bin(os.path.join(main_dir, f"graph{k}.bin"), np.ascontiguousarray(order, np.int32))
    write_simple_bin(os.path.join(main_dir, f"state_face{k}.bin"), np.ascontiguousarray(face_state, np.int32))
    _write_locx0(os.path.join(main_dir, f"LocX0{k}.bin"), s, in_face, x0)


def read_direction_files(main_dir: str, k: int):
    """Inverse of write_direction_files: (order, face_state, s, in_face)."""
    from .scatter import read_simple_bin
    order = read_simple_bin(os.path.join(main_dir, f"graph{k}.bin"), np.int32)
    state = read_simple_bin(os.path.join(main_dir, f"state_face{k}.bin"), np.int32)
    s, in_face, _ = read_locx0(os.path.join(main_dir, f"LocX0{k}.bin"))
    return order, state, s, in_face


# ---------------------------------------------------------------- GPU sweep

class Sweep:
    """Device sweep structures for all directions (rsh_sweep_*)."""

    def __init__(self, n_cells: int, n_dirs: int, pairs, nodes=None, device: int = 0):
        self.n, self.m = n_cells, n_dirs
        self._pairs = np.ascontiguousarray(pairs, np.int32)
        nd = None if nodes is None else np.ascontiguousarray(nodes, np.float64)
        h = C.c_void_p()
        N.call("rsh_sweep_create", device, n_cells, n_dirs, self._pairs.ctypes.data, _p(nd), C.byref(h))
        self.handle = h

    def set_direction(self, k: int, order, face_state, s, in_face) -> None:
        o, st = np.ascontiguousarray(order, np.int32), np.ascontiguousarray(face_state, np.int32)
        ss, f = np.ascontiguousarray(s, np.float64), np.ascontiguousarray(in_face, np.uint8)
        N.call("rsh_sweep_set_direction", self.handle, k, o.ctypes.data, st.ctypes.data, _p(ss) if ss.size else None,
               _p(f) if f.size else None, ss.size)

    def set_all(self, per_dir) -> None:
        for k, d in enumerate(per_dir):
            self.set_direction(k, *d)

    def levels(self, k: int) -> int:
        v = C.c_int32()
        N.call("rsh_sweep_levels", self.handle, k, C.byref(v))
        return v.value

    def run(self, cls: int, scattering, illum_new, illum_old=None, norm=None, ld: Optional[int] = None,
            stream=None) -> None:
        """One sweep of every direction on device arrays: scattering [M][ld], illum [M][N][4], norm [M]."""
        from .scatter import _ptr, _stream
        N.call("rsh_sweep_run", self.handle, cls, _ptr(scattering), self.n if ld is None else ld,
               None if illum_old is None else _ptr(illum_old), _ptr(illum_new),
               None if norm is None else _ptr(norm), _stream(stream))

    def close(self) -> None:
        if self.handle:
            N.call("rsh_sweep_destroy", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
