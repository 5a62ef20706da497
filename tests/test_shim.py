"""The C++ drop-in of cuda/cuda_solve.h (include/rsh_cuda_solve.hpp).

CPU: every function the header declares is defined by the shim (checked on the
built driver's symbol table) and the signatures match the reference's.
GPU: the driver replays the reference's illumination-iteration call sequence
(Illum_async.cpp:663-1072) through the shim; both the async two-stream result
and CalculateIntScattering match the oracle.
"""
import os
import re
import subprocess

import numpy as np
import pytest

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


@pytest.mark.gpu
def test_shim_replays_reference_call_sequence(gpu, tmp_path):
    n, m, start, end = 2500, 60, 12, 48
    rng = np.random.default_rng(21)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    xyz, w, F = orc.fibonacci(m)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.int64(n).tobytes() + np.int32(m).tobytes() + np.int32(start).tobytes() +
                np.int32(end).tobytes())
        f.write(xyz.astype(np.float64).tobytes() + w.astype(np.float64).tobytes() + np.float64(F).tobytes())
        f.write(illum.astype(np.float64).tobytes())
    r = subprocess.run([_driver(), str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "time kernel1" in r.stdout and "shim_driver ok" in r.stdout
    res = np.fromfile(outp, dtype=np.float64).reshape(2, end - start, n)
    ref = orc.scatter_reference(illum, xyz, w, F, k_disp=start, k_count=end - start, threads=16)
    for got in res:
        assert float(np.max(np.abs(got - ref) / ref)) < 1e-12
