"""GPU parity of the split-precision forms against the FP64 oracle.

The forms differ in how many significand bits their operands keep:
  bf16x2  hi+lo BF16   ~16 bits  -> per-element representation error <= 2^-16
  fp16x2  hi+lo FP16   ~22 bits  (power-of-two scaled per cell and per row)
  bf16x3  hi+mid+lo BF16 ~24 bits
all with FP32 MFMA accumulation (main and cross-term accumulators, combined in
FP64). Tolerances written here, max / mean relative error vs the FP64 oracle:
  fp16x2, bf16x3 : max <= 1e-6
  bf16x2         : max <= 3e-5, mean <= 3e-6  (16-bit operands cannot reach 1e-6 per result;
                   the measured errors are printed and recorded by bench.py)
"""
import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs

pytestmark = pytest.mark.gpu

TOL = {"fp16x2": (1e-6, 1e-6), "bf16x3": (1e-6, 1e-6), "bf16x2": (3e-5, 3e-6)}


def errs(got, ref):
    r = np.abs(got - ref) / np.abs(ref)
    return float(r.max()), float(r.mean())


def run_split(ds, form, illum, faces, k0, k1, out_f32=False, ld=None):
    m, n = illum.shape[0], illum.shape[1]
    d_illum = D.DeviceArray.from_host(np.ascontiguousarray(illum, dtype=np.float64))
    sb, cb = rs.split_sizes(form, faces, n, m)
    state = D.DeviceArray((sb,), np.uint8)
    scale = D.DeviceArray((max(cb // 4, 1),), np.float32)
    rs.pack_split_dev(form, d_illum, faces, n, m, state, scale)
    ld = n if ld is None else ld
    out = D.DeviceArray((max(k1 - k0, 1), ld), np.float32 if out_f32 else np.float64)
    ds.scatter_split(form, state, scale, faces, n, k0, k1, out, ld=ld, out_f32=out_f32)
    D.synchronize()
    res = out.to_host()[: k1 - k0, :n].astype(np.float64)
    for a in (d_illum, state, scale, out):
        a.free()
    return res


@pytest.mark.parametrize("form", ["bf16x2", "fp16x2", "bf16x3"])
def test_hg08_reduced_n(gpu, form):
    """The HG workload's shape at reduced N: HG g = 0.8, FP64 row normalisation, log-normal(0, 1.5) faces."""
    n, m = 4096, 400
    rng = np.random.default_rng(0x5EED + 3)
    illum = rng.lognormal(0.0, 1.5, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    ref = orc.scatter_matrix(illum, orc.phase_matrix(xyz, w, F, 1, 0.8, True), threads=16)
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.8, normalize=True) as ds:
        got = run_split(ds, form, illum, 4, 0, m)
    mx, mean = errs(got, ref)
    print(f"{form}: max rel err {mx:.3e}, mean {mean:.3e}")
    assert mx <= TOL[form][0] and mean <= TOL[form][1]


@pytest.mark.parametrize("form", ["bf16x2", "fp16x2", "bf16x3"])
@pytest.mark.parametrize("n,m,k0,k1,faces", [(1, 1, 0, 1, 4), (37, 26, 3, 21, 4), (1000, 130, 17, 129, 1),
                                             (513, 800, 0, 800, 4), (300, 33, 0, 33, 1)])
def test_split_ragged_thomson(gpu, form, n, m, k0, k1, faces):
    rng = np.random.default_rng(n + 7 * m)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, faces))
    xyz, w, F = rs.fibonacci_directions(m)
    ref = orc.scatter_reference(illum, xyz, w, F, k_disp=k0, k_count=k1 - k0, threads=16)
    with rs.DirectionSet(xyz, w, F) as ds:
        got = run_split(ds, form, illum, faces, k0, k1, ld=n + 5)
    mx, mean = errs(got, ref)
    assert mx <= TOL[form][0] and mean <= TOL[form][1], (mx, mean)


def test_split_fp32_output(gpu):
    n, m = 2000, 200
    rng = np.random.default_rng(3)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    ref = orc.scatter_reference(illum, xyz, w, F, threads=16)
    with rs.DirectionSet(xyz, w, F) as ds:
        got = run_split(ds, "fp16x2", illum, 4, 0, m, out_f32=True)
    mx, _ = errs(got, ref)
    assert mx <= 1e-6


def test_split_isotropic_normalised(gpu):
    """I = c everywhere with row-normalised P: S = c up to the split error."""
    n, m = 1000, 400
    illum = np.full((m, n, 4), 3.25)
    xyz, w, F = rs.fibonacci_directions(m)
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.6, normalize=True) as ds:
        for form in ("bf16x2", "fp16x2", "bf16x3"):
            got = run_split(ds, form, illum, 4, 0, m)
            assert np.max(np.abs(got - 3.25) / 3.25) <= TOL[form][0]


@pytest.mark.parametrize("form", ["bf16x2", "fp16x2", "bf16x3"])
def test_split_tiles_bit_identical(gpu, form):
    """One wave per tile (waves=4) or two waves splitting the rows (waves=8): same sums, same bits."""
    n, m = 3000, 400
    rng = np.random.default_rng(31)
    illum = rng.lognormal(0.0, 1.5, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.8, normalize=True) as ds:
        d_illum = D.DeviceArray.from_host(illum)
        sb, cb = rs.split_sizes(form, 4, n, m)
        state, scale = D.DeviceArray((sb,), np.uint8), D.DeviceArray((cb // 4,), np.float32)
        rs.pack_split_dev(form, d_illum, 4, n, m, state, scale)
        outs = []
        for waves in (4, 8):
            out = D.DeviceArray((m, n))
            ds.scatter_split(form, state, scale, 4, n, 0, m, out, waves=waves)
            D.synchronize()
            outs.append(out.to_host())
            out.free()
        for a in (d_illum, state, scale):
            a.free()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("form", ["bf16x2", "fp16x2", "bf16x3"])
@pytest.mark.parametrize("n,m,k0,k1", [(70001, 400, 0, 400), (5000, 200, 0, 200), (3000, 400, 37, 301),
                                       (17, 26, 0, 26), (9000, 800, 0, 800)])
def test_resident_panel_kernel_bit_identical(gpu, form, n, m, k0, k1):
    """Cell-level state: the resident-panel kernel (waves=1, the default for faces = 1)
    and both streaming kernels sum the same products in the same order."""
    rng = np.random.default_rng(n + m + k0)
    cell = rng.lognormal(0.0, 1.5, size=(m, n, 1))
    xyz, w, F = rs.fibonacci_directions(m)
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.8, normalize=True) as ds:
        d_cell = D.DeviceArray.from_host(cell)
        sb, cb = rs.split_sizes(form, 1, n, m)
        state, scale = D.DeviceArray((sb,), np.uint8), D.DeviceArray((cb // 4,), np.float32)
        rs.pack_split_dev(form, d_cell, 1, n, m, state, scale)
        outs = {}
        for waves in (1, 0, 4, 8):
            out = D.DeviceArray((k1 - k0, n + 3))
            out.fill(-1.0)
            ds.scatter_split(form, state, scale, 1, n, k0, k1, out, ld=n + 3, waves=waves)
            D.synchronize()
            outs[waves] = out.to_host()
            out.free()
        for a in (d_cell, state, scale):
            a.free()
    ref = outs[4]
    assert np.all(ref[:, n:] == -1.0)  # padding columns untouched
    for wv, o in outs.items():
        assert np.array_equal(o, ref), wv
    if m <= 400 and n <= 5000:
        P = orc.phase_matrix(xyz, w, F, 1, 0.8, True, k0, k1)
        exact = orc.scatter_matrix(cell, P, threads=16)
        mx, mean = errs(ref[:, :n], exact)
        assert mx <= TOL[form][0] and mean <= TOL[form][1], (mx, mean)


@pytest.mark.parametrize("form", ["bf16x2", "fp16x2", "bf16x3"])
def test_face_averaged_pack_vs_oracle(gpu, form):
    """rsh_pack_split_avg_dev: 4-face input reduced to a cell-level state in the pack;
    the resident-panel kernel on it matches the FP64 oracle on the 4-face input."""
    n, m = 4096, 400
    rng = np.random.default_rng(0x5EED + 33)
    illum = rng.lognormal(0.0, 1.5, size=(m, n, 4))
    xyz, w, F = rs.fibonacci_directions(m)
    ref = orc.scatter_matrix(illum, orc.phase_matrix(xyz, w, F, 1, 0.8, True), threads=16)
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.8, normalize=True) as ds:
        d_illum = D.DeviceArray.from_host(illum)
        sb, cb = rs.split_sizes(form, 1, n, m)
        state, scale = D.DeviceArray((sb,), np.uint8), D.DeviceArray((cb // 4,), np.float32)
        rs.pack_split_avg_dev(form, d_illum, n, m, state, scale)
        out = D.DeviceArray((m, n))
        ds.scatter_split(form, state, scale, 1, n, 0, m, out)
        D.synchronize()
        got = out.to_host()
        for a in (d_illum, state, scale, out):
            a.free()
    mx, mean = errs(got, ref)
    print(f"{form} face-averaged pack: max rel err {mx:.3e}, mean {mean:.3e}")
    assert mx <= TOL[form][0] and mean <= TOL[form][1]


def test_resident_kernel_rejects_face_state(gpu):
    xyz, w, F = rs.fibonacci_directions(26)
    with rs.DirectionSet(xyz, w, F) as ds:
        sb, cb = rs.split_sizes("bf16x2", 4, 64, 26)
        state, scale = D.DeviceArray((sb,), np.uint8), D.DeviceArray((cb // 4,), np.float32)
        out = D.DeviceArray((26, 64))
        with pytest.raises(RuntimeError, match="resident-panel"):
            ds.scatter_split("bf16x2", state, scale, 4, 64, 0, 26, out, waves=1)
        for a in (state, scale, out):
            a.free()


def _bf16_rne(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32 values."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _part_values(bits, form):
    """16-bit part patterns -> their float32 values."""
    if form == "fp16x2":
        return bits.view(np.float16).astype(np.float32)
    return (bits.astype(np.uint32) << 16).view(np.float32)


def _nearest(r, form):
    """The value of the part format nearest to r (ties to even)."""
    return r.astype(np.float16).astype(np.float32) if form == "fp16x2" else _bf16_rne(r)


@pytest.mark.parametrize("form", ["bf16x2", "fp16x2", "bf16x3"])
@pytest.mark.parametrize("faces", [4, 1])
def test_cell_pack_matches_layout_definition(gpu, form, faces):
    """The one-pass cell-level pack against a numpy restatement of its definition:
    cell scale 2^-e, e = 11 - ilogb(max_{j,f} |I[j][c][f]|); value (((f0 + f1) + f2) + f3) / 4
    (faces = 4) or I[j][c] (faces = 1), times 2^e, rounded to FP32, split into parts that
    each round what is left to the 16-bit format; stored MFMA-native: the 16-byte vector
    ((tile * steps + s) * parts + part) * 64 + lane holds cells 16 tile + lane % 16,
    directions 32 s + 8 (lane / 16) + e. Ragged cells and directions pad with zeros.
    The device's float -> 16-bit conversions may break an exact tie either way
    (the split stays exact: each remainder is computed in FP32), so a part must
    equal the round-to-nearest-even value or lie at the same distance from the
    remainder."""
    n, m = 1000, 200            # 62.5 cell tiles, 6.25 K-steps
    rng = np.random.default_rng(31 + faces)
    illum = rng.lognormal(0.0, 1.5, size=(m, n, faces))
    illum[:, 17, :] *= 2.0 ** -40      # a cell far from the others' magnitude
    d_illum = D.DeviceArray.from_host(illum)
    sb, cb = rs.split_sizes(form, 1, n, m)
    state = D.DeviceArray((sb,), np.uint8)
    scale = D.DeviceArray((cb // 4,), np.float32)
    if faces == 4:
        rs.pack_split_avg_dev(form, d_illum, n, m, state, scale)
    else:
        rs.pack_split_dev(form, d_illum, 1, n, m, state, scale)
    D.synchronize()
    got_state, got_scale = state.to_host(), scale.to_host()
    for a in (d_illum, state, scale):
        a.free()
    mx = np.max(np.abs(illum), axis=(0, 2))
    e = 11 - (np.frexp(mx)[1] - 1)
    assert np.array_equal(got_scale, np.ldexp(np.float32(1.0), -e).astype(np.float32))
    v = (((illum[..., 0] + illum[..., 1]) + illum[..., 2]) + illum[..., 3]) * 0.25 if faces == 4 else illum[..., 0]
    x = (v * np.ldexp(1.0, e)[None, :]).astype(np.float32)          # [m][n]
    steps, tiles = (m + 31) // 32, (n + 15) // 16
    xp = np.zeros((steps * 32, tiles * 16), np.float32)
    xp[:m, :n] = x
    nt = 3 if form == "bf16x3" else 2
    got = got_state.view(np.uint16).reshape(tiles, steps, nt, 64, 8)
    # [tile][s][part][lane = 16 g + c][e] -> [part][direction 32 s + 8 g + e][cell 16 tile + c]
    vals = got.reshape(tiles, steps, nt, 4, 16, 8).transpose(2, 1, 3, 5, 0, 4).reshape(nt, steps * 32, tiles * 16)
    # Each part is a nearest value of its format to what the parts before it left
    # (the device's conversion may break exact ties either way); what is left after
    # each part is computed exactly in FP32, as the kernel does.
    r = xp.copy()
    for t in range(nt):
        p = _part_values(vals[t], form)
        q = _nearest(r, form)
        tie = np.abs(r.astype(np.float64) - p) == np.abs(r.astype(np.float64) - q)
        bad = np.argwhere(~((p == q) | tie))
        if len(bad):
            j, c = bad[0]
            raise AssertionError(f"part {t}: {len(bad)} of {p.size} parts are not a nearest value, first at "
                                 f"(dir {j}, cell {c}): remainder {r[j, c]!r}, got {p[j, c]!r}, nearest {q[j, c]!r}")
        r = (r - p).astype(np.float32)
    assert np.array_equal(vals[:, m:, :], np.zeros_like(vals[:, m:, :]))    # padding directions
    assert np.array_equal(vals[:, :, n:], np.zeros_like(vals[:, :, n:]))    # padding cells
