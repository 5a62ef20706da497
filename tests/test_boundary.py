"""CPU tests of the C-ABI boundary (no GPU compute calls).

* librsh_scatter.so loads and exports every function include/*.h declares;
* host-only entry points (direction generator, direction file I/O, partition)
  agree with the oracle's restatement of the reference;
* invalid arguments come back as status codes with a message, before any
  device work is attempted.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import _native as N
from relativistic_solver_hllc_amd import scatter as rs

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(rsh_[A-Za-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "rsh_scatter" in names and "rsh_create" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    missing = [n for n in declared_functions() if not hasattr(N.lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    # and the Python mirror binds every one of them
    assert set(declared_functions()) <= set(N.PROTOTYPES)


def test_abi_version():
    assert N.lib.rsh_abi_version() == 1


@pytest.mark.parametrize("m", [1, 2, 26, 200, 400, 800])
def test_fibonacci_matches_oracle(m):
    xyz, w, F = rs.fibonacci_directions(m)
    oxyz, ow, oF = orc.fibonacci(m)
    assert np.array_equal(xyz, oxyz) and np.array_equal(w, ow) and F == oF


def test_direction_file_roundtrip(tmp_path, golden_dir):
    xyz, w, F = rs.read_directions(os.path.join(golden_dir, "dirs_M26.txt"))
    oxyz, ow, oF = orc.read_directions(os.path.join(golden_dir, "dirs_M26.txt"))
    assert np.array_equal(xyz, oxyz) and np.array_equal(w, ow) and F == oF
    p = str(tmp_path / "d.txt")
    rs.write_directions(p, xyz, w, F)
    x2, w2, F2 = rs.read_directions(p)
    assert np.array_equal(x2, xyz) and np.array_equal(w2, w) and F2 == F
    # the oracle reads what the library writes
    x3, w3, F3 = orc.read_directions(p)
    assert np.array_equal(x3, xyz) and F3 == F


def test_direction_file_errors(tmp_path):
    with pytest.raises(N.RshError) as e:
        rs.read_directions(str(tmp_path / "missing.txt"))
    assert e.value.status == 5
    bad = tmp_path / "bad.txt"
    bad.write_text("3\n1 0 0 1\n1 0 1\n")
    with pytest.raises(N.RshError) as e:
        rs.read_directions(str(bad))
    assert e.value.status == 5
    big = tmp_path / "big.txt"
    big.write_text("5\n" + "1 0 0 1\n" * 5 + "5\n")
    with pytest.raises(N.RshError) as e:
        rs.read_directions(str(big), cap=4)
    assert e.value.status == 1


def test_partition_matches_oracle(golden_dir):
    for p in range(1, 17):
        for n in (0, 1, 5, 126, 364, 400, 401, 800):
            send, disp = rs.partition(p, n)
            assert np.array_equal(send, orc.get_send(p, n))
            assert np.array_equal(disp, orc.get_disp(p, n))
    with pytest.raises(N.RshError):
        rs.partition(0, 5)


def test_argument_errors_without_device():
    null = C.c_void_p()
    # null direction set
    assert N.lib.rsh_scatter_f64_dev(null, null, 4, 10, 0, 1, null, 10, null, 0, None) == 1
    assert b"direction set" in N.lib.rsh_last_error()
    # bad problem descriptor (faces = 3) is rejected before any device call
    d = N.ProblemDesc(0, 10, 3, 4, None, None, 1.0, 0, 4, 0, 0.0, 0)
    h = C.c_void_p()
    assert N.lib.rsh_create(C.byref(d), C.byref(h)) == 1
    d = N.ProblemDesc(0, 10, 4, 4, None, None, 1.0, 2, 1, 0, 0.0, 0)
    assert N.lib.rsh_create(C.byref(d), C.byref(h)) == 1
    assert N.lib.rsh_scatter(null, 0, 1, 0, None) == 1
    assert N.lib.rsh_sync(null, 0) == 1
    assert N.lib.rsh_upload_illum(null, None, 1, 0, 0, 0) == 1
    xyz, w, F = rs.fibonacci_directions(4)
    # HG with |g| >= 1 and unknown phase ids are argument errors
    assert N.lib.rsh_dirset_create(0, 4, xyz.ctypes.data, w.ctypes.data, F, 1, 1.0, 1, C.byref(h)) == 1
    assert N.lib.rsh_dirset_create(0, 4, xyz.ctypes.data, w.ctypes.data, F, 7, 0.0, 0, C.byref(h)) == 1


def test_device_count_is_callable():
    assert rs.device_count() >= 0


def test_compare_files_matches_definition(tmp_path):
    """CompareFiles (utils/rebuild_solve.cpp:579-669) on simple-bin files written
    through the library: first index of the largest difference and the
    sequential sum, |a-b| for doubles and the Euclidean norm for Vector3 /
    Matrix3 elements; the first file's count drives the loop."""
    rng = np.random.default_rng(5)
    for type_, wd in (("double", 1), ("Vector3", 3), ("Matrix3", 9)):
        a = rng.integers(-8, 8, size=(40, wd)).astype(np.float64)  # exact differences
        b = a.copy()
        b[7] += 0.5           # largest difference, planted twice: the first index wins
        b[31] += 0.5
        b[12] -= 0.25
        f1, f2 = str(tmp_path / f"a{wd}.bin"), str(tmp_path / f"b{wd}.bin")
        rs.write_simple_bin(f1, a, wd)
        rs.write_simple_bin(f2, np.vstack([b, rng.normal(size=(3, wd))]), wd)  # longer second file
        idx, worst, total = rs.compare_files(f1, f2, type_)
        d = [abs(x[0] - y[0]) if wd == 1 else float(np.sqrt(sum((xi - yi) ** 2 for xi, yi in zip(x, y))))
             for x, y in zip(a.tolist(), b.tolist())]
        assert idx == 7 and worst == max(d)
        assert total == pytest.approx(sum(d), rel=1e-15)
    with pytest.raises(ValueError):
        rs.compare_files(f1, f2, "int64")



@pytest.mark.parametrize("m,rows", [(1, 16), (16, 16), (26, 32), (200, 208), (208, 208), (209, 224), (400, 416),
                                    (800, 832)])
def test_batch_footprint_counts_padded_rows(m, rows):
    """rsh_batch_footprint (host arithmetic, no device) counts whole row blocks.
    FP64 slots hold I twice and B with the iteration plan's rows: passes of at
    most 13 row blocks of 16 (M = 209 -> 2 passes of 7 row blocks = 224 rows)."""
    n = 1000
    assert rs.Batch.footprint("fp64", n, m) == 3 * rows * n * 8


@pytest.mark.parametrize("form", ["bf16x2", "fp16x2", "bf16x3"])
def test_split_footprint_pads_source_rows(form):
    """Split slots hold B in FP32 in whole 16-row blocks and 16-cell tiles: M = 17
    and M = 32 need the same memory (one 32-direction K-step of state, 32 rows of
    B), M = 16 one row block of B less (16 rows x 1008 cells for N = 1000)."""
    n = 1000
    assert rs.Batch.footprint(form, n, 17) == rs.Batch.footprint(form, n, 32)
    assert rs.Batch.footprint(form, n, 17) - rs.Batch.footprint(form, n, 16) == 16 * 1008 * 4


@pytest.mark.parametrize("form", ["fp64", "bf16x2", "fp16x2", "bf16x3"])
def test_footprint_is_what_the_arena_carves(form):
    """The footprint counts every slot buffer rounded up to the arena's 256-byte
    alignment, as rsh_batch_reserve's arena carves it: with an odd cell count a
    request whose footprint equals the budget still fits an empty arena."""
    def r256(b):
        return -(-b // 256) * 256

    n, m = 1001, 200
    if form == "fp64":
        want = 3 * r256(208 * n * 8)                  # I ping-pong + B, 13-row-block passes
    else:
        nt = 3 if form == "bf16x3" else 2
        steps = (m + 31) // 32
        state = -(-n // 16) * steps * nt * 64 * 16    # packed state units of 16 bytes
        want = 2 * r256(state) + r256(208 * (-(-n // 16) * 16) * 4) + r256(4 * n)   # FP32 B in 16-cell tiles
    assert rs.Batch.footprint(form, n, m) == want


def test_footprint_refuses_what_a_batch_cannot_hold():
    """FP64 slots address one row block with 32-bit offsets: 2^24 cells or more
    are refused by the footprint itself (so a server rejects such a request at
    submission instead of queueing it for ever); split forms accept them."""
    from relativistic_solver_hllc_amd._native import RshError

    with pytest.raises(RshError):
        rs.Batch.footprint("fp64", 1 << 24, 200)
    assert rs.Batch.footprint("fp64", (1 << 24) - 1, 16) > 0
    assert rs.Batch.footprint("bf16x3", 1 << 24, 200) > 0


def test_import_fails_loudly_without_the_library(tmp_path):
    """A copy of the package without librsh_scatter.so refuses to load (no Python
    or CPU fallback): importing its scatter module raises ImportError naming the
    build command."""
    import shutil
    import subprocess
    import sys
    src = os.path.join(REPO, "relativistic_solver_hllc_amd")
    dst = tmp_path / "relativistic_solver_hllc_amd"
    dst.mkdir()
    for f in os.listdir(src):
        if f.endswith(".py"):
            shutil.copy(os.path.join(src, f), dst / f)
    proc = subprocess.run([sys.executable, "-c", "import relativistic_solver_hllc_amd.scatter"], cwd=tmp_path,
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode != 0
    assert "ImportError" in proc.stderr and "is not built" in proc.stderr
