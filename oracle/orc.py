"""ctypes wrapper of the CPU oracle (oracle/scatter_oracle.cpp, sweep_oracle.cpp, make_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline, never
as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS = {"parity": "liborc.so", "o3": "liborc_o3.so", "ofast": "liborc_ofast.so"}

i32, i64, f64, vp = C.c_int32, C.c_int64, C.c_double, C.c_void_p
_PROTOS = {
    "orc_thomson_gamma": (f64, [f64]),
    "orc_hg_gamma": (f64, [f64, f64]),
    "orc_phase": (f64, [C.c_int, f64, f64]),
    "orc_scatter_reference": (None, [vp, i64, i32, i32, vp, vp, f64, i32, i32, i64, i64, vp, i32]),
    "orc_face_average": (None, [vp, i64, i32, i32, vp]),
    "orc_phase_matrix": (None, [i32, vp, vp, f64, i32, f64, i32, i32, i32, vp]),
    "orc_scatter_matrix": (None, [vp, i64, i32, i32, vp, i32, i64, i64, vp, i32]),
    "orc_get_send": (None, [i32, i32, vp]),
    "orc_get_disp": (None, [i32, i32, vp]),
    "orc_recalc_norm": (f64, [vp, vp, i64]),
    "orc_source_iteration": (None, [vp, vp, i64, i32, vp, f64, i32, vp, i32]),
    "orc_fibonacci": (None, [i32, vp, vp, C.POINTER(f64)]),
    "orc_read_directions": (C.c_int, [C.c_char_p, i32, C.POINTER(i32), vp, vp, C.POINTER(f64)]),
    "orc_max_threads": (C.c_int, []),
    "orc_moments": (None, [vp, i64, i32, i32, vp, vp, f64, vp, vp, vp, vp, vp, vp, vp, vp, i32]),
    "orc_read_simple_bin": (i64, [C.c_char_p, i32, vp, i64]),
    "orc_sweep": (f64, [i64, i32, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp, i32]),
    "orc_make_nodes": (None, [i64, vp, vp]),
    "orc_make_direction": (C.c_int, [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64, C.POINTER(i64)]),
}
_cache = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(variant: str = "parity"):
    if variant in _cache:
        return _cache[variant]
    path = os.path.join(HERE, VARIANTS[variant])
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    _cache[variant] = lib
    return lib


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def scatter_reference(illum, xyz, w, F, k_disp=0, k_count=None, cell_begin=0, cell_end=None, threads=0,
                      variant="parity", out=None):
    """GetS + CalculateIntCPU restatement. illum [M][N][faces] -> S [k_count][N]."""
    lib = load(variant)
    illum = _c(illum)
    M, N, faces = illum.shape
    k_count = M - k_disp if k_count is None else k_count
    cell_end = N if cell_end is None else cell_end
    S = np.zeros((k_count, N)) if out is None else out
    xyz, w = _c(xyz), _c(w)
    lib.orc_scatter_reference(illum.ctypes.data, N, M, faces, xyz.ctypes.data, w.ctypes.data, float(F), k_disp,
                              k_count, cell_begin, cell_end, S.ctypes.data, threads)
    return S


def phase_matrix(xyz, w, F, kind=0, g=0.0, normalize=False, k0=0, k1=None):
    lib = load()
    xyz, w = _c(xyz), _c(w)
    M = len(w)
    k1 = M if k1 is None else k1
    P = np.empty((k1 - k0, M))
    lib.orc_phase_matrix(M, xyz.ctypes.data, w.ctypes.data, float(F), kind, float(g), int(normalize), k0, k1,
                         P.ctypes.data)
    return P


def scatter_matrix(illum, P, cell_begin=0, cell_end=None, threads=0):
    """S[k][c] = sum_j P[k][j] * mean_f illum[j][c][f]; illum [M][N][faces]."""
    lib = load()
    illum, P = _c(illum), _c(P)
    M, N, faces = illum.shape
    cell_end = N if cell_end is None else cell_end
    S = np.zeros((P.shape[0], N))
    lib.orc_scatter_matrix(illum.ctypes.data, N, M, faces, P.ctypes.data, P.shape[0], cell_begin, cell_end,
                           S.ctypes.data, threads)
    return S


def get_send(np_, n):
    out = np.empty(np_, dtype=np.int32)
    load().orc_get_send(np_, n, out.ctypes.data)
    return out


def get_disp(np_, n):
    out = np.empty(np_, dtype=np.int32)
    load().orc_get_disp(np_, n, out.ctypes.data)
    return out


def recalc_norm(old, new):
    old = _c(old).copy()
    new = _c(new)
    r = load().orc_recalc_norm(old.ctypes.data, new.ctypes.data, old.size)
    return r, old


def source_iteration(I0, B, P, albedo, n_iter, threads=0):
    I = _c(I0).copy()
    B, P = _c(B), _c(P)
    M, N = I.shape
    resid = np.empty(n_iter)
    load().orc_source_iteration(I.ctypes.data, B.ctypes.data, N, M, P.ctypes.data, float(albedo), n_iter,
                                resid.ctypes.data, threads)
    return I, resid


def fibonacci(m):
    xyz = np.empty((m, 3))
    w = np.empty(m)
    F = C.c_double()
    load().orc_fibonacci(m, xyz.ctypes.data, w.ctypes.data, C.byref(F))
    return xyz, w, F.value


def read_directions(path, cap=1 << 16):
    xyz = np.empty((cap, 3))
    w = np.empty(cap)
    m = C.c_int32()
    F = C.c_double()
    rc = load().orc_read_directions(path.encode(), cap, C.byref(m), xyz.ctypes.data, w.ctypes.data, C.byref(F))
    if rc:
        raise IOError(f"orc_read_directions({path}) -> {rc}")
    return xyz[: m.value].copy(), w[: m.value].copy(), F.value


def gamma(mu):
    return load().orc_thomson_gamma(float(mu))


def phase(kind, g, mu):
    return load().orc_phase(int(kind), float(g), float(mu))


MOMENTS = ("energy", "stream", "impuls", "divstream", "divimpuls")
MOMENT_WIDTH = {"energy": 1, "stream": 3, "impuls": 9, "divstream": 1, "divimpuls": 3}


def moments(illum, xyz, w, F, normals=None, areas=None, volume=None, which=MOMENTS, threads=0):
    """Radiation moments of illum[j][c][f] (orc_moments); returns a dict of the
    requested outputs, shaped [n] or [n][width]."""
    illum = _c(illum)
    m, n, faces = illum.shape
    geo = [None if a is None else _c(a) for a in (normals, areas, volume)]
    if any(k.startswith("div") for k in which) and any(a is None for a in geo):
        raise ValueError("divergences need normals, areas and volume")
    outs = {k: np.empty((n, MOMENT_WIDTH[k])) for k in which}
    ptr = lambda a: None if a is None else a.ctypes.data
    load().orc_moments(illum.ctypes.data, n, m, faces, _c(xyz).ctypes.data, _c(w).ctypes.data, float(F),
                       ptr(geo[0]), ptr(geo[1]), ptr(geo[2]),
                       *[ptr(outs.get(k)) for k in MOMENTS], int(threads))
    return {k: (v[:, 0] if MOMENT_WIDTH[k] == 1 else v) for k, v in outs.items()}


def read_simple_bin(path, dtype=np.float64):
    """Element array of a simple-bin file (int32 count + raw elements)."""
    dt = np.dtype(dtype)
    lib = load()
    n = lib.orc_read_simple_bin(path.encode(), dt.itemsize, None, 0)
    if n < 0:
        raise IOError(f"orc_read_simple_bin({path}) failed")
    out = np.empty(n, dt)
    if lib.orc_read_simple_bin(path.encode(), dt.itemsize, out.ctypes.data, n) != n:
        raise IOError(f"orc_read_simple_bin({path}) failed")
    return out


def sweep(per_dir, pairs, scattering, illum_old, cls=1, nodes=None, threads=0, variant="parity"):
    """Sequential transport sweep (oracle/sweep_oracle.cpp, Illum_async.cpp:725-784 + ReCalcIllum).
    per_dir: list over directions of (order, face_state, s, in_face); scattering [M][N];
    illum_old [M][N][4]. Returns (illum_new, norm); norm = -2 flags malformed input."""
    lib = load(variant)
    M = len(per_dir)
    pairs = _c(pairs, np.int32)
    n = pairs.shape[0]
    order = _c(np.stack([d[0] for d in per_dir]), np.int32)
    state = _c(np.stack([d[1] for d in per_dir]), np.int32)
    counts = [len(d[2]) for d in per_dir]
    begin = _c(np.concatenate([[0], np.cumsum(counts)]), np.int64)
    s = _c(np.concatenate([d[2] for d in per_dir]) if sum(counts) else np.zeros(1))
    f = _c(np.concatenate([d[3] for d in per_dir]) if sum(counts) else np.zeros(1, np.uint8), np.uint8)
    S = _c(scattering)
    illum = _c(illum_old).copy()
    nd = None if nodes is None else _c(nodes)
    norm = lib.orc_sweep(n, M, order.ctypes.data, state.ctypes.data, s.ctypes.data, f.ctypes.data, begin.ctypes.data,
                         pairs.ctypes.data, None if nd is None else nd.ctypes.data, S.ctypes.data, cls,
                         illum.ctypes.data, threads)
    return illum, norm


def make_nodes(vertex_matrix):
    """MakeArrayX (oracle/make_oracle.cpp): face nodes [N,4,3,3] from vertex matrices [N,4,4]."""
    vm = _c(vertex_matrix)
    out = np.empty((vm.shape[0], 4, 3, 3))
    load().orc_make_nodes(vm.shape[0], vm.ctypes.data, out.ctypes.data)
    return out


def make_direction(vertex_matrix, faces, normals, pairs, order, omega):
    """The reference's MAKE loop for one direction: (status, face_state, s, in_face, x0 [R,2]);
    status 1 = a node found no in face strictly inside its triangle."""
    vm, fc, nr = _c(vertex_matrix), _c(faces), _c(normals)
    pr, od, w = _c(pairs, np.int32), _c(order, np.int32), _c(omega)
    n = vm.shape[0]
    cap = 9 * n
    st, s, x0, f = np.empty(n, np.int32), np.empty(cap), np.empty((cap, 2)), np.empty(cap, np.uint8)
    cnt = C.c_int64()
    rc = load().orc_make_direction(n, vm.ctypes.data, fc.ctypes.data, nr.ctypes.data, pr.ctypes.data, od.ctypes.data,
                                   w.ctypes.data, st.ctypes.data, s.ctypes.data, x0.ctypes.data, f.ctypes.data, cap,
                                   C.byref(cnt))
    r = min(cnt.value, cap)
    return rc, st, s[:r].copy(), f[:r].copy(), x0[:r].copy()
