"""The C++ drop-in of cuda/cuda_solve.h (include/rsh_cuda_solve.hpp).

CPU: every function the header declares is defined by the shim (checked on the
built test program's symbol table) and the signatures match the reference's.
GPU: the test program replays the reference's illumination-iteration call sequence
(Illum_async.cpp:663-1072) through the shim; both the async two-stream result
and CalculateIntScattering match the oracle, and on the rank with start == 0
CalculateAllParamStream (with geometry read from normals.bin / squares.bin /
volume.bin) and CalculateEnergy match the oracle's moments.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from mesh_util import moment_scales, random_tetra_cells
from oracle import orc
from relativistic_solver_hllc_amd import build as B

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Reference declarations (solver_hllc_illum_cuda/cuda/cuda_solve.h:5-37), as demangled by nm -C.
REFERENCE_SIGNATURES = [
    "CudaEventPrint()",
    "CudaSyncStream(e_cuda_stream_id_t)",
    "CalculateIntScatteringAsyncMPIStream(grid_directions_t const&, grid_t&, int, int, e_cuda_stream_id_t)",
    "CudaSendIllumAsync(int, int, double const*)",
    "CudaSendIllumAsync(int, int, int, double const*)",
    "CalculateIntScatteringAsyncMPI(grid_directions_t const&, grid_t&, int)",
    "CopyIllumOnDevice(int, double const*)",
    "CalculateIntScattering(grid_directions_t const&, grid_t&)",
    "CalculateIntScatteringAsync(grid_directions_t const&, grid_t&)",
    "CalculateEnergy(grid_directions_t const&, grid_t&)",
    "CalculateStream(grid_directions_t const&, grid_t&)",
    "CalculateImpuls(grid_directions_t const&, grid_t&)",
    "CalculateDivImpuls(grid_directions_t const&, grid_t&)",
    "CalculateDivStream(grid_directions_t const&, grid_t&)",
    "CalculateAllParam(grid_directions_t const&, grid_t&)",
    "CalculateAllParamStream(grid_directions_t const&, grid_t&, e_cuda_stream_id_t)",
    "SetDevice(int)",
    "InitDevice(grid_directions_t const&, grid_t&, int, int)",
    "ClearDevice()",
    "ClearHost(grid_t&)",
    "CudaWait()",
]


def _driver():
    if not os.path.exists(B.SHIM_DRIVER):
        pytest.fail(f"{B.SHIM_DRIVER} not built; run python -m relativistic_solver_hllc_amd.build")
    return B.SHIM_DRIVER


def test_shim_defines_every_reference_entry_point():
    out = subprocess.run(["nm", "-C", "--defined-only", _driver()], capture_output=True, text=True,
                         check=True).stdout
    defined = {re.sub(r"^[0-9a-f]+ [TtWw] ", "", l).strip() for l in out.splitlines()}
    missing = [s for s in REFERENCE_SIGNATURES if s not in defined]
    assert not missing, missing
    # 64-bit overloads for offsets beyond 2^31
    assert "CudaSendIllumAsync(long, long, long, double const*)" in defined


def write_geometry(dirpath, normals, areas, volume):
    """normals.bin (int cells + 4 Vector3 per cell), squares.bin, volume.bin
    (file_module/reader_bin.cpp:49-69, reader_bin.h:13-31)."""
    n = len(volume)
    for name, arr, count in (("normals.bin", normals, n), ("squares.bin", areas, areas.size),
                             ("volume.bin", volume, n)):
        with open(os.path.join(dirpath, name), "wb") as f:
            f.write(np.int32(count).tobytes() + np.ascontiguousarray(arr, dtype=np.float64).tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("start,end", [(12, 48), (0, 20)])
def test_shim_replays_reference_call_sequence(gpu, tmp_path, start, end):
    n, m = 2500, 60
    rng = np.random.default_rng(21)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    xyz, w, F = orc.fibonacci(m)
    normals, areas, volume, _ = random_tetra_cells(n, rng)
    write_geometry(tmp_path, normals, areas, volume)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.int64(n).tobytes() + np.int32(m).tobytes() + np.int32(start).tobytes() +
                np.int32(end).tobytes())
        f.write(xyz.astype(np.float64).tobytes() + w.astype(np.float64).tobytes() + np.float64(F).tobytes())
        f.write(illum.astype(np.float64).tobytes())
    r = subprocess.run([_driver(), str(inp), str(outp), str(tmp_path) + os.sep], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "time kernel1" in r.stdout and "shim_driver ok" in r.stdout
    raw = np.fromfile(outp, dtype=np.float64)
    ns = 2 * (end - start) * n
    res = raw[:ns].reshape(2, end - start, n)
    ref = orc.scatter_reference(illum, xyz, w, F, k_disp=start, k_count=end - start, threads=16)
    for got in res:
        assert float(np.max(np.abs(got - ref) / ref)) < 1e-12
    if start != 0:
        assert raw.size == ns  # moments only on the rank with start == 0
        return
    assert raw.size == ns + 18 * n
    mref = orc.moments(illum, xyz, w, F, normals, areas, volume, threads=16)
    e, d = moment_scales(illum, w, F, areas, volume)
    off = ns
    for k, width in (("energy", 1), ("stream", 3), ("impuls", 9), ("divstream", 1), ("divimpuls", 3)):
        got = raw[off:off + width * n].reshape(n, width) if width > 1 else raw[off:off + n]
        off += width * n
        err = np.abs(got - mref[k])
        err = err if err.ndim == 1 else err.max(axis=1)
        assert float(np.max(err / (d if k.startswith("div") else e))) < 1e-12, k
    energy2 = raw[off:]
    assert np.array_equal(energy2, raw[ns:ns + n])  # CalculateEnergy = energy of CalculateAllParamStream
