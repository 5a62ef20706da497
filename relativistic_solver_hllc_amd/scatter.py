"""Python host mirror of the scattering-integral boundary.

Thin wrappers over the C ABI (include/rsh_scatter.h). Naming follows the
reference's host API (solver_hllc_illum_cuda/cuda/cuda_solve.h:5-37):

    reference                                   here
    ReadSphereDirectionDecartToSpherical        read_directions
    GetSend / GetDisp                           partition
    grid_directions_init_copy_host_to_device    DirectionSet
    InitDevice / ClearDevice                    Problem(...) / Problem.close()
    CudaSendIllumAsync / CopyIllumOnDevice      Problem.send_illum_async / copy_illum_on_device
    CalculateIntScattering*                     Problem.calculate_int_scattering
    CudaSyncStream / CudaWait                   Problem.sync_stream / wait
    CudaEventPrint                              Problem.event_times

Device memory handed to the *_dev functions is any object with a device
address: relativistic_solver_hllc_amd.device.DeviceArray (the library's own
allocations), or a raw integer address.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _native as N

PHASES = {"thomson": 0, "hg": 1}
FORMS = {"fp64": 0, "bf16x2": 1, "fp16x2": 2, "bf16x3": 3}
F64_KERNELS = {"auto": 0, "direct": 1, "mfma": 2}
STREAM_SCATTERING_1, STREAM_SCATTERING_2, STREAM_PARAMS = 0, 1, 2


def _ptr(a) -> int:
    """Device or host address of a torch tensor or numpy array."""
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a)


def _f64c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _stream(stream) -> Optional[int]:
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return int(stream)


# ------------------------------------------------------------------ host helpers

def fibonacci_directions(m: int) -> Tuple[np.ndarray, np.ndarray, float]:
    xyz = np.empty((m, 3))
    w = np.empty(m)
    F = C.c_double()
    N.call("rsh_fibonacci_directions", m, xyz.ctypes.data, w.ctypes.data, C.byref(F))
    return xyz, w, F.value


def read_directions(path: str, cap: int = 1 << 16) -> Tuple[np.ndarray, np.ndarray, float]:
    xyz = np.empty((cap, 3))
    w = np.empty(cap)
    m = C.c_int32()
    F = C.c_double()
    N.call("rsh_read_directions", path.encode(), cap, C.byref(m), xyz.ctypes.data, w.ctypes.data, C.byref(F))
    return xyz[: m.value].copy(), w[: m.value].copy(), F.value


def write_directions(path: str, xyz, w, full_area: float) -> None:
    xyz, w = _f64c(xyz), _f64c(w)
    N.call("rsh_write_directions", path.encode(), len(w), xyz.ctypes.data, w.ctypes.data, float(full_area))


def read_simple_bin(path: str, dtype=np.float64, width: int = 1) -> np.ndarray:
    """Simple-bin file (int32 count + raw elements of `width` values of dtype):
    squares.bin / volume.bin (width 1), normals.bin (width = faces * 3)."""
    dt = np.dtype(dtype)
    esize = dt.itemsize * width
    count = C.c_int64()
    N.call("rsh_read_simple_bin", path.encode(), esize, None, 0, C.byref(count))
    out = np.empty((count.value, width) if width > 1 else count.value, dt)
    N.call("rsh_read_simple_bin", path.encode(), esize, out.ctypes.data if out.size else None, count.value,
           C.byref(count))
    return out


def write_simple_bin(path: str, data, width: int = 1) -> None:
    """Writes a simple-bin file; the count is len(data) // width elements."""
    a = np.ascontiguousarray(data)
    esize = a.dtype.itemsize * width
    N.call("rsh_write_simple_bin", path.encode(), esize, a.ctypes.data if a.size else None, a.size // width)


# Element widths (doubles) of the solution files, as CompareFiles reads them
# (double, Vector3, Matrix3; utils/rebuild_solve.cpp:601-669).
SOLUTION_FILES = {"Illum.bin": 1, "energy.bin": 1, "stream.bin": 3, "impuls.bin": 9, "divstream.bin": 1,
                  "divimpuls.bin": 3}
COMPARE_WIDTH = {"double": 1, "Vector3": 3, "Matrix3": 9}


def compare_files(file1: str, file2: str, type_: str) -> Tuple[int, float, float]:
    """CompareFiles (utils/rebuild_solve.cpp:579-669) for double / Vector3 / Matrix3 files:
    (index of the largest difference, that difference, sum of differences); differences are
    |a - b| for doubles and the Euclidean (Frobenius) norm of a - b for vectors and matrices.
    The first file's count drives the loop, as in the reference."""
    if type_ not in COMPARE_WIDTH:
        raise ValueError(f"unknown type {type_!r}: double, Vector3 or Matrix3")
    wd = COMPARE_WIDTH[type_]
    a = read_simple_bin(file1, np.float64, wd).reshape(-1, wd)
    b = read_simple_bin(file2, np.float64, wd).reshape(-1, wd)
    if len(b) < len(a):
        raise ValueError(f"{file2} holds {len(b)} elements, {file1} {len(a)}")
    d = np.abs(a[:, 0] - b[: len(a), 0]) if wd == 1 else np.sqrt(((a - b[: len(a)]) ** 2).sum(axis=1))
    best, idx, total = -1.0, -1, 0.0
    for i, x in enumerate(d.tolist()):  # sequential sum and first maximum, as the reference loop
        total += x
        if best < x:
            best, idx = x, i
    return idx, best, total


def partition(np_: int, n: int) -> Tuple[np.ndarray, np.ndarray]:
    send = np.empty(np_, dtype=np.int32)
    disp = np.empty(np_, dtype=np.int32)
    N.call("rsh_partition", np_, n, send.ctypes.data, disp.ctypes.data)
    return send, disp


def device_count() -> int:
    n = C.c_int32()
    N.call("rsh_device_count", C.byref(n))
    return n.value


def selftest_mfma_f64(device: int = 0) -> int:
    bad = C.c_int32()
    N.call("rsh_selftest_mfma_f64", device, C.byref(bad))
    return bad.value


def fill_lognormal_dev(out, mu: float, sigma: float, seed: int, stream=None) -> None:
    N.call("rsh_fill_lognormal_dev", _ptr(out), out.numel(), float(mu), float(sigma), int(seed) & (2**64 - 1),
           _stream(stream))


def split_sizes(form: str, faces: int, n_cells: int, m: int) -> Tuple[int, int]:
    """Bytes of the packed split state and of its per-cell scale array."""
    a, b = C.c_int64(), C.c_int64()
    N.call("rsh_split_sizes", FORMS[form], faces, n_cells, m, C.byref(a), C.byref(b))
    return a.value, b.value


def pack_split_dev(form: str, illum, faces: int, n_cells: int, m: int, state, cell_scale, stream=None) -> None:
    """FP64 illum [M][n_cells][faces] -> resident split state + per-cell inverse scales."""
    N.call("rsh_pack_split_dev", FORMS[form], _ptr(illum), faces, n_cells, m, _ptr(state), _ptr(cell_scale),
           _stream(stream))


# ------------------------------------------------------------------ direction set

def pack_split_avg_dev(form: str, illum, n_cells: int, m: int, state, cell_scale, stream=None) -> None:
    """Cell-level split state (faces = 1) of the face average of 4-face FP64 illum."""
    N.call("rsh_pack_split_avg_dev", FORMS[form], _ptr(illum), n_cells, m, _ptr(state), _ptr(cell_scale),
           _stream(stream))


class DirectionSet:
    """Directions, weights and the device FP64 phase matrix of one direction set."""

    def __init__(self, xyz, w, full_area: float, phase: str = "thomson", g: float = 0.0,
                 normalize: bool = False, device: int = 0):
        self.xyz, self.w = _f64c(xyz).reshape(-1, 3), _f64c(w)
        self.m = len(self.w)
        self.full_area = float(full_area)
        self.phase, self.g, self.normalize, self.device = phase, float(g), bool(normalize), device
        h = C.c_void_p()
        N.call("rsh_dirset_create", device, self.m, self.xyz.ctypes.data, self.w.ctypes.data, self.full_area,
               PHASES[phase], self.g, int(self.normalize), C.byref(h))
        self.handle = h

    def phase_matrix(self, k0: int = 0, k1: Optional[int] = None) -> np.ndarray:
        k1 = self.m if k1 is None else k1
        out = np.empty((k1 - k0, self.m))
        N.call("rsh_dirset_phase_matrix", self.handle, k0, k1, out.ctypes.data)
        return out

    def scatter_f64(self, illum, faces: int, n_cells: int, k0: int, k1: int, out, ld: Optional[int] = None,
                    stream=None, kernel: str = "auto", tile: Optional[Tuple[int, int]] = None) -> None:
        """S[k - k0][c] = sum_j P[k][j] * mean_f illum[j][c][f] on device memory."""
        t = C.byref(N.Tile(*tile)) if tile is not None else None
        N.call("rsh_scatter_f64_dev", self.handle, _ptr(illum), faces, n_cells, k0, k1, _ptr(out),
               n_cells if ld is None else ld, _stream(stream), F64_KERNELS[kernel], t)

    def scatter_f64_noalloc(self, illum, faces: int, n_cells: int, k0: int, k1: int, out,
                            ld: Optional[int] = None, stream=None) -> None:
        """rsh_scatter_f64_dev_noalloc: the MFMA path with the phase panel built in the library's fixed
        per-device scratch (no device allocation; overlapping calls on one device race)."""
        N.call("rsh_scatter_f64_dev_noalloc", self.handle, _ptr(illum), faces, n_cells, k0, k1, _ptr(out),
               ld or n_cells, _stream(stream))

    def scatter_split(self, form: str, state, cell_scale, faces: int, n_cells: int, k0: int, k1: int, out,
                      ld: Optional[int] = None, out_f32: bool = False, stream=None, waves: int = 0) -> None:
        """Split-precision S rows [k0, k1) from a packed state (see pack_split_dev)."""
        N.call("rsh_scatter_split_dev", self.handle, FORMS[form], _ptr(state), _ptr(cell_scale), faces, n_cells, k0,
               k1, _ptr(out), n_cells if ld is None else ld, int(out_f32), _stream(stream), waves)

    def moments(self, illum, faces: int, n_cells: int, energy=None, stream_out=None, impuls=None, divstream=None,
                divimpuls=None, normals=None, areas=None, volume=None, faces_per_thread: int = 0,
                stream=None) -> None:
        """Radiation moments over all directions on device memory (rsh_moments_f64_dev).
        Outputs: energy [N], stream [N][3], impuls [N][9], divstream [N], divimpuls [N][3];
        geometry: normals [N][faces][3], areas [N][faces], volume [N]. None = not computed / absent."""
        opt = lambda a: None if a is None else _ptr(a)
        N.call("rsh_moments_f64_dev", self.handle, _ptr(illum), faces, n_cells, opt(normals), opt(areas),
               opt(volume), opt(energy), opt(stream_out), opt(impuls), opt(divstream), opt(divimpuls),
               faces_per_thread, _stream(stream))

    def close(self) -> None:
        if self.handle:
            N.call("rsh_dirset_destroy", self.handle)
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


# ------------------------------------------------------------------ autotuner and batches

def autotune_f64(m: int, faces: int = 1, n_cells: int = 65536, device: int = 0) -> Tuple[Tuple[int, int], float]:
    """Times every FP64 tile for the row-block class of m and records the fastest."""
    t = N.Tile()
    ms = C.c_float()
    N.call("rsh_autotune_f64", device, m, faces, n_cells, C.byref(t), C.byref(ms))
    return (t.waves, t.chunk_steps), ms.value


def tuned_tile_f64(m: int, device: int = 0) -> Tuple[int, int]:
    t = N.Tile()
    N.call("rsh_tuned_tile_f64", device, m, C.byref(t))
    return t.waves, t.chunk_steps


def batch_autotune(form: str, m: int, n_cells: int = 100000, iterations: int = 6, device: int = 0):
    """Times the grouped-launch candidates of the batch class of (form, m) and records the fastest
    for the device: FP64 -> ("tile", (waves, j-steps)); split forms -> ("kernel", "resident" | "resident4" |
    "streaming").
    Returns (choice, ms per iteration)."""
    t, ms = N.Tile(), C.c_float()
    N.call("rsh_batch_autotune", device, FORMS[form], m, int(n_cells), iterations, C.byref(t), C.byref(ms))
    return _grouped_choice(form, t), ms.value


def batch_tuned(form: str, m: int, device: int = 0):
    t = N.Tile()
    N.call("rsh_batch_tuned", device, FORMS[form], m, C.byref(t))
    return _grouped_choice(form, t)


def _grouped_choice(form: str, t):
    if form == "fp64":
        return ("tile", (t.waves, t.chunk_steps))
    return ("kernel", {0: "auto", 1: "resident", 2: "streaming", 3: "resident4"}[t.waves])


class Batch:
    """Uneven source-iteration problems sharing grouped launches (rsh_batch_*)."""

    def __init__(self, capacity: int, device: int = 0, use_graph: bool = True):
        h = C.c_void_p()
        N.call("rsh_batch_create", device, capacity, int(use_graph), C.byref(h))
        self.handle, self.capacity = h, capacity

    @staticmethod
    def footprint(form: str, n_cells: int, m: int) -> int:
        b = C.c_int64()
        N.call("rsh_batch_footprint", FORMS[form], n_cells, m, C.byref(b))
        return b.value

    def add(self, ds: DirectionSet, n_cells: int, I0, B, albedo: float, form: str = "fp64") -> int:
        slot = C.c_int32()
        N.call("rsh_batch_add", self.handle, ds.handle, n_cells, _ptr(I0), _ptr(B), float(albedo), FORMS[form],
               C.byref(slot))
        return slot.value

    def add_seeded(self, ds: DirectionSet, n_cells: int, seed: int, albedo: float, form: str = "fp64") -> int:
        """Slot with I0 ~ log-normal(0,1) and B ~ U(0.5,2) generated on the device from seed."""
        slot = C.c_int32()
        N.call("rsh_batch_add_seeded", self.handle, ds.handle, n_cells, int(seed) & (2**64 - 1), float(albedo),
               FORMS[form], C.byref(slot))
        return slot.value

    def set_split_kernel(self, kind: str) -> None:
        """Kernel for split-form slots added afterwards: "auto", "resident" (resident panel, 8 waves of 2 cell
        tiles), "resident4" (resident panel, 4 waves of 2 cell tiles, two workgroups per CU where the LDS allows)
        or "streaming"."""
        N.call("rsh_batch_set_split_kernel", self.handle,
               {"auto": 0, "resident": 1, "streaming": 2, "resident4": 3}[kind])

    def set_f64_tile(self, tile: Optional[Tuple[int, int]]) -> None:
        """FP64 grouped tile (waves, j-steps) of the classes created from now on; None = tuned or default."""
        t = N.Tile(*(tile or (0, 0)))
        N.call("rsh_batch_set_f64_tile", self.handle, C.byref(t))

    def slot_config(self, slot: int, form: str):
        """What a slot runs with: its FP64 class's tile, or the split kernel."""
        t = N.Tile()
        N.call("rsh_batch_slot_config", self.handle, slot, C.byref(t))
        return _grouped_choice(form, t)

    def reserve(self, nbytes: int) -> None:
        """One device arena for the slot buffers (batch must be empty)."""
        N.call("rsh_batch_reserve", self.handle, int(nbytes))

    def remove(self, slot: int) -> None:
        N.call("rsh_batch_remove", self.handle, slot)

    def run(self, iterations: int) -> None:
        N.call("rsh_batch_run", self.handle, iterations)

    def sync(self) -> None:
        N.call("rsh_batch_sync", self.handle)

    def residuals(self) -> Tuple[np.ndarray, np.ndarray]:
        r = np.empty(self.capacity)
        it = np.empty(self.capacity, dtype=np.int32)
        N.call("rsh_batch_residuals", self.handle, r.ctypes.data, it.ctypes.data)
        return r, it

    def state(self, slot: int, out) -> None:
        N.call("rsh_batch_get_state", self.handle, slot, _ptr(out))

    def close(self) -> None:
        if self.handle:
            N.call("rsh_batch_destroy", self.handle)
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


REQ_STATUS = {0: "pending", 1: "running", 2: "converged", 3: "maxed", 4: "cancelled", 5: "failed"}


def request(n_cells: int, m: int, phase: str = "thomson", g: float = 0.0, albedo: float = 0.5,
            tolerance: float = 1e-8, max_iterations: int = 200, form: str = "fp64", seed: int = 0) -> N.RequestDesc:
    return N.RequestDesc(n_cells, m, PHASES[phase], float(g), float(albedo), float(tolerance), max_iterations,
                         FORMS[form], int(seed) & (2**64 - 1))


class Server:
    """Continuous-batching server of source-iteration requests (rsh_server_*)."""

    def __init__(self, hbm_budget_bytes: int, max_slots: int = 64, iters_per_step: int = 2,
                 keep_results: bool = False, device: int = 0):
        h = C.c_void_p()
        N.call("rsh_server_create", device, int(hbm_budget_bytes), max_slots, iters_per_step, int(keep_results),
               C.byref(h))
        self.handle = h

    def submit(self, desc: N.RequestDesc) -> int:
        rid = C.c_int32()
        N.call("rsh_server_submit", self.handle, C.byref(desc), C.byref(rid))
        return rid.value

    def cancel(self, rid: int) -> None:
        N.call("rsh_server_cancel", self.handle, rid)

    def step(self) -> Tuple[int, int]:
        running, done = C.c_int32(), C.c_int32()
        N.call("rsh_server_step", self.handle, C.byref(running), C.byref(done))
        return running.value, done.value

    def poll(self, rid: int) -> dict:
        info = N.RequestInfo()
        N.call("rsh_server_poll", self.handle, rid, C.byref(info))
        return dict(status=REQ_STATUS[info.status], iterations=info.iterations, residual=info.residual,
                    t_submit=info.t_submit, t_admit=info.t_admit, t_done=info.t_done,
                    footprint_bytes=info.footprint_bytes)

    def result(self, rid: int, out) -> None:
        N.call("rsh_server_result", self.handle, rid, _ptr(out))

    def stats(self) -> dict:
        used, pend, run, comp = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int64()
        N.call("rsh_server_stats", self.handle, C.byref(used), C.byref(pend), C.byref(run), C.byref(comp))
        return dict(used_bytes=used.value, pending=pend.value, running=run.value, completed=comp.value)

    def run_trace(self, descs, arrivals) -> Tuple[np.ndarray, float]:
        """Replays arrival times (seconds, ascending) in real time; returns per-request latency and wall time."""
        n = len(descs)
        arr = (N.RequestDesc * n)(*descs)
        t = np.ascontiguousarray(arrivals, dtype=np.float64)
        lat = np.empty(n)
        wall = C.c_double()
        N.call("rsh_server_run_trace", self.handle, n, C.cast(arr, C.c_void_p), t.ctypes.data, lat.ctypes.data,
               C.byref(wall))
        return lat, wall.value

    def close(self) -> None:
        if self.handle:
            N.call("rsh_server_destroy", self.handle)
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


def fill_uniform_dev(out, lo: float, hi: float, seed: int, stream=None) -> None:
    N.call("rsh_fill_uniform_dev", _ptr(out), out.numel(), float(lo), float(hi), int(seed) & (2**64 - 1),
           _stream(stream))


def mfma_peak(kind: str = "fp64", device: int = 0) -> float:
    """Measured register-only MFMA rate in TFLOP/s ("fp64": 16x16x4 FP64, "bf16": 16x16x32 BF16)."""
    t = C.c_double()
    N.call("rsh_mfma_peak", device, {"fp64": 0, "bf16": 1}[kind], C.byref(t))
    return t.value


def rel_error_dev(got, ref, rows: int, n_cells: int, ld_got: Optional[int] = None, ld_ref: Optional[int] = None,
                  threshold: float = 1e-6, stream=None) -> dict:
    """Relative error |got - ref| / |ref| of every one of rows x n_cells results (got FP64 or FP32,
    ref FP64, both device arrays) reduced on the device: max, mean and the count above threshold."""
    got_f32 = 1 if np.dtype(got.dtype) == np.float32 else 0
    mx, mean, above, at = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
    N.call("rsh_dev_rel_error", _ptr(got), got_f32, int(ld_got or n_cells), _ptr(ref), int(ld_ref or n_cells),
           int(rows), int(n_cells), float(threshold), C.byref(mx), C.byref(mean), C.byref(above), C.byref(at),
           _stream(stream))
    worst = (at.value // int(n_cells), at.value % int(n_cells)) if at.value >= 0 else None
    return {"max_rel_err": mx.value, "mean_rel_err": mean.value, "count_above": above.value,
            "threshold": threshold, "results": int(rows) * int(n_cells), "worst_row_cell": worst}


# ------------------------------------------------------------------ problem (InitDevice ...)

class Problem:
    """Reference-shaped problem: device illum/scattering buffers + 3 streams."""

    def __init__(self, n_cells: int, xyz, w, full_area: float, k_begin: int = 0, k_end: Optional[int] = None,
                 faces: int = 4, phase: str = "thomson", g: float = 0.0, normalize: bool = False, device: int = 0):
        xyz, w = _f64c(xyz).reshape(-1, 3), _f64c(w)
        m = len(w)
        k_end = m if k_end is None else k_end
        self.n_cells, self.m, self.faces, self.k_begin, self.k_end = n_cells, m, faces, k_begin, k_end
        d = N.ProblemDesc(device, n_cells, faces, m, xyz.ctypes.data, w.ctypes.data, float(full_area), k_begin,
                          k_end, PHASES[phase], float(g), int(normalize))
        h = C.c_void_p()
        N.call("rsh_create", C.byref(d), C.byref(h))
        self.handle = h

    def send_illum_async(self, host, count: int, dst_off: int = 0, src_off: Optional[int] = None,
                         stream: int = STREAM_SCATTERING_1) -> None:
        """CudaSendIllumAsync(size, shift[, shift_host], Illum_host)."""
        src_off = dst_off if src_off is None else src_off
        N.call("rsh_upload_illum", self.handle, _ptr(host), count, dst_off, src_off, stream)

    def copy_illum_on_device(self, host, count: int) -> None:
        N.call("rsh_copy_illum_sync", self.handle, _ptr(host), count)

    def calculate_int_scattering(self, k0: int, k1: int, stream: int = STREAM_SCATTERING_1, host_out=None) -> None:
        N.call("rsh_scatter", self.handle, k0, k1, stream, None if host_out is None else _ptr(host_out))

    def calculate_int_scattering_from_host(self, host_illum, k0: int, k1: int, host_out=None, chunks: int = 8) -> None:
        """CalculateIntScattering's upload + compute + copy back, pipelined over `chunks` cell ranges
        (rsh_scatter_from_host); returns when done."""
        N.call("rsh_scatter_from_host", self.handle, _ptr(host_illum), k0, k1,
               None if host_out is None else _ptr(host_out), chunks)

    def sync_stream(self, stream: int) -> None:
        N.call("rsh_sync", self.handle, stream)

    def wait(self) -> None:
        N.call("rsh_sync", self.handle, -1)

    def buffers(self) -> Tuple[int, int]:
        a, b = C.c_void_p(), C.c_void_p()
        N.call("rsh_problem_buffers", self.handle, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_geometry(self, normals, areas, volume) -> None:
        """normals [N][faces][3], areas [N][faces], volume [N] (InitDevice's normals.bin,
        squares.bin, volume.bin)."""
        self._geo = [_f64c(normals), _f64c(areas), _f64c(volume)]
        N.call("rsh_problem_set_geometry", self.handle, *[a.ctypes.data for a in self._geo])

    def calculate_all_param(self, stream: int = STREAM_PARAMS, energy=None, stream_out=None, impuls=None,
                            divstream=None, divimpuls=None, sync: bool = True) -> None:
        """CalculateAllParam(Stream): moments on a stream id, copied into the given host arrays."""
        opt = lambda a: None if a is None else _ptr(a)
        N.call("rsh_moments", self.handle, stream, opt(energy), opt(stream_out), opt(impuls), opt(divstream),
               opt(divimpuls))
        if sync:
            N.call("rsh_sync", self.handle, stream)

    def write_solution(self, main_dir: str) -> None:
        """Solution dump (WriteFileSolution, file_module/writer_bin.cpp:15-44): Illum.bin and, after
        calculate_all_param, energy / stream / impuls (/ divstream / divimpuls with geometry) .bin files
        named main_dir + name."""
        N.call("rsh_write_solution", self.handle, main_dir.encode())

    def moment_buffers(self) -> Tuple[int, int, int, int, int]:
        ptrs = [C.c_void_p() for _ in range(5)]
        N.call("rsh_problem_moment_buffers", self.handle, *[C.byref(p) for p in ptrs])
        return tuple(p.value for p in ptrs)

    def event_times(self) -> np.ndarray:
        ms = (C.c_float * 3)()
        N.call("rsh_event_times", self.handle, ms)
        return np.array(list(ms))

    def close(self) -> None:
        if self.handle:
            N.call("rsh_destroy", self.handle)
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
