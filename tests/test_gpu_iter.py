"""GPU source iteration I <- (1 - a) B + a S(I) in batches (rsh_batch_*), FP64.

Checked against the oracle's source iteration (orc_source_iteration, residual of
ReCalcIllum, Illum_async.cpp:472-510) and the 60-digit golden trace:
state max relative error <= 1e-12, residuals within 1e-9 relative.
Batching must not change results: a problem alone and in a mix, eager or
graph-replayed, gives bit-identical states.
"""
import os

import numpy as np
import pytest

from oracle import orc
from relativistic_solver_hllc_amd import device as D
from relativistic_solver_hllc_amd import scatter as rs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b))))


def make(m, n, phase, g, seed):
    xyz, w, F = rs.fibonacci_directions(m)
    rng = np.random.default_rng(seed)
    I0 = rng.lognormal(0.0, 1.0, size=(m, n))
    B = rng.uniform(0.5, 2.0, size=(m, n))
    kind = 0 if phase == "thomson" else 1
    P = orc.phase_matrix(xyz, w, F, kind, g, True)
    ds = rs.DirectionSet(xyz, w, F, phase=phase, g=g, normalize=True)
    return dict(m=m, n=n, xyz=xyz, w=w, F=F, I0=I0, B=B, P=P, ds=ds)


def add(batch, p, albedo=0.5):
    dI, dB = D.DeviceArray.from_host(p["I0"]), D.DeviceArray.from_host(p["B"])
    slot = batch.add(p["ds"], p["n"], dI, dB, albedo)
    dI.free()
    dB.free()
    return slot


def state(batch, slot, p):
    out = D.DeviceArray((p["m"], p["n"]))
    batch.state(slot, out)
    h = out.to_host()
    out.free()
    return h


def test_golden_trace(gpu, golden_dir):
    d = np.load(os.path.join(golden_dir, "iter5_thomson_norm_N32_M14.npz"))
    m, n = d["I0"].shape
    with rs.DirectionSet(d["xyz"], d["w"], float(d["F"]), normalize=True) as ds, rs.Batch(4) as b:
        dI, dB = D.DeviceArray.from_host(d["I0"]), D.DeviceArray.from_host(d["B"])
        slot = b.add(ds, n, dI, dB, float(d["albedo"]))
        for t in range(5):
            b.run(1)
            r, it = b.residuals()
            assert it[slot] == t + 1
            assert abs(r[slot] - d["resid"][t]) <= 1e-9 * d["resid"][t]
            out = D.DeviceArray((m, n))
            b.state(slot, out)
            assert rel(out.to_host(), d["trace"][t]) < 1e-12
            out.free()


MIX = [(200, 1000, "thomson", 0.0), (400, 3001, "hg", 0.6), (800, 517, "hg", 0.9), (26, 2048, "hg", 0.3),
       (130, 700, "thomson", 0.0), (400, 40, "hg", 0.3)]


def test_uneven_mix_vs_oracle(gpu):
    probs = [make(m, n, ph, g, seed=i) for i, (m, n, ph, g) in enumerate(MIX)]
    iters = 7
    with rs.Batch(8) as b:
        slots = [add(b, p) for p in probs]
        b.run(iters)
        r, it = b.residuals()
        for p, s in zip(probs, slots):
            ref, rres = orc.source_iteration(p["I0"], p["B"], p["P"], 0.5, iters, threads=16)
            assert it[s] == iters
            assert rel(state(b, s, p), ref) < 1e-12, (p["m"], p["n"])
            assert abs(r[s] - rres[-1]) <= 1e-9 * rres[-1]
    for p in probs:
        p["ds"].close()


def test_alone_equals_in_mix_and_graph_equals_eager(gpu):
    probs = [make(m, n, ph, g, seed=10 + i) for i, (m, n, ph, g) in enumerate(MIX)]
    target = 1
    results = {}
    for label, members, graph in (("alone", [target], True), ("mix", list(range(len(probs))), True),
                                  ("eager", list(range(len(probs))), False)):
        with rs.Batch(8, use_graph=graph) as b:
            slots = {i: add(b, probs[i]) for i in members}
            b.run(6)
            results[label] = state(b, slots[target], probs[target])
    assert np.array_equal(results["alone"], results["mix"])
    assert np.array_equal(results["mix"], results["eager"])
    for p in probs:
        p["ds"].close()


def test_join_and_leave(gpu):
    A, Bp, Cp = make(400, 1500, "hg", 0.6, 21), make(200, 900, "thomson", 0.0, 22), make(800, 300, "hg", 0.9, 23)
    with rs.Batch(4) as b:
        sa, sb = add(b, A), add(b, Bp)
        b.run(3)
        sc = add(b, Cp)          # joins after 3 iterations of the others
        b.run(4)
        b.remove(sa)             # leaves after 7 iterations
        b.run(2)
        r, it = b.residuals()
        assert it[sa] == -1 and it[sb] == 9 and it[sc] == 6
        for p, s, k in ((Bp, sb, 9), (Cp, sc, 6)):
            ref, rres = orc.source_iteration(p["I0"], p["B"], p["P"], 0.5, k, threads=16)
            assert rel(state(b, s, p), ref) < 1e-12
            assert abs(r[s] - rres[-1]) <= 1e-9 * rres[-1]
    for p in (A, Bp, Cp):
        p["ds"].close()


def test_autotune_records_a_supported_tile(gpu):
    tile, ms = rs.autotune_f64(400, faces=1, n_cells=32768)
    assert tile in [(8, 2), (4, 2), (4, 4)] and ms > 0
    assert rs.tuned_tile_f64(400) == tile


# ---------------------------------------------------------------- split-precision state
# State error bound after the iterations below, vs the FP64 oracle iteration (relative):
SPLIT_TOL = {"fp16x2": 1e-6, "bf16x3": 1e-6, "bf16x2": 6e-5}


def add_form(batch, p, form, albedo=0.5):
    dI, dB = D.DeviceArray.from_host(p["I0"]), D.DeviceArray.from_host(p["B"])
    slot = batch.add(p["ds"], p["n"], dI, dB, albedo, form=form)
    dI.free()
    dB.free()
    return slot


@pytest.mark.parametrize("form", ["fp16x2", "bf16x3", "bf16x2"])
def test_split_iteration_vs_oracle_side_by_side(gpu, form):
    cases = [(400, 2000, "hg", 0.6), (200, 900, "thomson", 0.0), (800, 333, "hg", 0.9), (161, 257, "hg", 0.3)]
    probs = [make(m, n, ph, g, seed=40 + i) for i, (m, n, ph, g) in enumerate(cases)]
    iters = 8
    with rs.Batch(16) as b:
        split_slots = [add_form(b, p, form) for p in probs]
        f64_slots = [add_form(b, p, "fp64") for p in probs]  # same problems in FP64, same launches
        b.run(iters)
        r, it = b.residuals()
        for p, ss, fs in zip(probs, split_slots, f64_slots):
            ref, rres = orc.source_iteration(p["I0"], p["B"], p["P"], 0.5, iters, threads=16)
            assert it[ss] == iters and it[fs] == iters
            err = rel(state(b, ss, p), ref)
            print(f"{form} M={p['m']} N={p['n']}: state max rel err {err:.2e}")
            assert err <= SPLIT_TOL[form]
            assert rel(state(b, fs, p), ref) < 1e-12
            assert abs(r[ss] - rres[-1]) <= 1e-3 * rres[-1]
    for p in probs:
        p["ds"].close()


def test_split_alone_equals_in_mix(gpu):
    probs = [make(m, n, ph, g, seed=60 + i) for i, (m, n, ph, g) in enumerate(MIX)]
    out = {}
    for label, members in (("alone", [2]), ("mix", list(range(len(probs))))):
        with rs.Batch(8) as b:
            slots = {i: add_form(b, probs[i], "fp16x2") for i in members}
            b.run(5)
            out[label] = state(b, slots[2], probs[2])
    assert np.array_equal(out["alone"], out["mix"])
    for p in probs:
        p["ds"].close()


@pytest.mark.parametrize("form", ["fp16x2", "bf16x3"])
def test_split_converges_below_1e_6(gpu, form):
    p = make(400, 1500, "hg", 0.8, seed=77)
    with rs.Batch(2) as b:
        s = add_form(b, p, form)
        for _ in range(60):
            b.run(1)
            r, _ = b.residuals()
            if r[s] < 1e-6:
                break
        assert r[s] < 1e-6
        ref, _ = orc.source_iteration(p["I0"], p["B"], p["P"], 0.5, int(b.residuals()[1][s]), threads=16)
        assert rel(state(b, s, p), ref) <= 1e-6
    p["ds"].close()


@pytest.mark.parametrize("form", ["fp64", "fp16x2"])
def test_add_is_ordered_after_null_stream_work(gpu, form):
    """Inputs written by kernels on the null stream and handed to rsh_batch_add
    with no host wait in between: the slot must see the finished arrays.
    Columns (cells) iterate independently, so sampled columns are checked
    against the oracle's iteration of the same columns."""
    m, n, iters = 400, 6 * 40 ** 3, 3
    xyz, w, F = rs.fibonacci_directions(m)
    P = orc.phase_matrix(xyz, w, F, 1, 0.6, True)
    cells = np.array([0, 1, 777, n // 2, n - 2, n - 1])
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.6, normalize=True) as ds, rs.Batch(2) as b:
        I0, B = D.DeviceArray((m, n)), D.DeviceArray((m, n))
        rs.fill_lognormal_dev(I0, 0.0, 1.0, seed=501)
        rs.fill_uniform_dev(B, 0.5, 2.0, seed=502)
        s = b.add(ds, n, I0, B, 0.5, form=form)
        b.run(iters)
        out = D.DeviceArray((m, n))
        b.state(s, out)
        got = out.gather_columns(cells)
        ref, _ = orc.source_iteration(I0.gather_columns(cells), B.gather_columns(cells), P, 0.5, iters)
        assert rel(got, ref) < (1e-12 if form == "fp64" else 1e-6)
        for a in (I0, B, out):
            a.free()


@pytest.mark.parametrize("form", ["fp64", "bf16x3"])
@pytest.mark.parametrize("arena", [False, True])
def test_seeded_slot_equals_copied_slot(gpu, form, arena):
    """rsh_batch_add_seeded generates in the slot exactly what rsh_fill_*_dev with
    the same seeds produce; arena-carved slots compute the same bits."""
    m, n, seed = 200, 3000, 1234
    xyz, w, F = rs.fibonacci_directions(m)
    with rs.DirectionSet(xyz, w, F, phase="hg", g=0.6, normalize=True) as ds, rs.Batch(4) as b:
        if arena:
            b.reserve(64 << 20)
        I0, B = D.DeviceArray((m, n)), D.DeviceArray((m, n))
        rs.fill_lognormal_dev(I0, 0.0, 1.0, seed=seed * 2 + 1)
        rs.fill_uniform_dev(B, 0.5, 2.0, seed=seed * 2 + 2)
        s1 = b.add(ds, n, I0, B, 0.5, form=form)
        s2 = b.add_seeded(ds, n, seed, 0.5, form=form)
        b.run(5)
        r, it = b.residuals()
        out1, out2 = D.DeviceArray((m, n)), D.DeviceArray((m, n))
        b.state(s1, out1)
        b.state(s2, out2)
        assert np.array_equal(out1.to_host(), out2.to_host())
        assert r[s1] == r[s2] and it[s1] == it[s2] == 5
        for a in (I0, B, out1, out2):
            a.free()


def test_arena_reuse_and_exhaustion(gpu):
    """Slots carved from a reserved arena are freed and coalesced on removal;
    a problem larger than any free block is refused with RSH_E_NOMEM and fits
    again once its neighbours are gone."""
    m = 200
    xyz, w, F = rs.fibonacci_directions(m)
    per = rs.Batch.footprint("fp64", 1000, m)
    with rs.DirectionSet(xyz, w, F, normalize=True) as ds, rs.Batch(8) as b:
        b.reserve(3 * per + 3 * 256 * 3)
        a1 = b.add_seeded(ds, 1000, 1, 0.5)
        a2 = b.add_seeded(ds, 1000, 2, 0.5)
        a3 = b.add_seeded(ds, 1000, 3, 0.5)
        with pytest.raises(RuntimeError, match="RSH_E_NOMEM"):
            b.add_seeded(ds, 1000, 4, 0.5)
        b.run(2)
        b.remove(a1)
        b.remove(a2)
        with pytest.raises(RuntimeError, match="RSH_E_NOMEM"):
            b.add_seeded(ds, 2500, 5, 0.5)  # more than the two freed blocks together
        big = b.add_seeded(ds, 2000, 5, 0.5)  # fits the coalesced block of a1 + a2
        b.run(3)
        r, it = b.residuals()
        assert it[big] == 3 and it[a3] == 5
        # the problem in the reused memory matches the same problem alone
        out = D.DeviceArray((m, 2000))
        b.state(big, out)
        got = out.to_host()
        out.free()
    with rs.DirectionSet(xyz, w, F, normalize=True) as ds, rs.Batch(2) as b2:
        s = b2.add_seeded(ds, 2000, 5, 0.5)
        b2.run(3)
        out = D.DeviceArray((m, 2000))
        b2.state(s, out)
        assert np.array_equal(out.to_host(), got)
        out.free()


@pytest.mark.parametrize("form", ["bf16x2", "fp16x2", "bf16x3"])
def test_split_resident_kernel_equals_streaming(gpu, form):
    """The resident-panel grouped kernel (M = 200, 400, 800; 8 waves of 2 cell tiles
    or 4 waves of 2, two workgroups per CU; at M = 800 the panel streamed through LDS by K-step) and the
    streaming-panel one compute the same products in the same order: after
    several graph-replayed iterations of an uneven mix (ragged cell counts, one M
    without a resident plan, an M = 800 problem of 20 cell chunks), states and
    residuals are bit-identical."""
    cases = [(400, 2000, "hg", 0.6), (200, 1025, "thomson", 0.0), (800, 333, "hg", 0.9), (161, 257, "hg", 0.3),
             (400, 17, "hg", 0.8), (200, 5000, "hg", 0.3), (800, 5000, "thomson", 0.0)]
    probs = [make(m, n, ph, g, seed=90 + i) for i, (m, n, ph, g) in enumerate(cases)]
    out = {}
    for kind in ("streaming", "resident", "resident4", "auto"):
        with rs.Batch(8) as b:
            slots = []
            for p in probs:
                # the resident kinds where a resident plan exists (M = 161 has none), else auto
                b.set_split_kernel(kind if p["m"] != 161 or kind == "streaming" else "auto")
                slots.append(add_form(b, p, form))
            for s, p in zip(slots, probs):
                if kind in ("resident", "resident4", "streaming") and p["m"] != 161:
                    assert b.slot_config(s, form) == ("kernel", kind)
            b.run(5)
            r, it = b.residuals()
            out[kind] = ([state(b, s, p) for s, p in zip(slots, probs)], [r[s] for s in slots])
    for k in ("resident", "resident4", "auto"):
        for a, c in zip(out["streaming"][0], out[k][0]):
            assert np.array_equal(a, c)
        assert out["streaming"][1] == out[k][1]
    for kind in ("resident", "resident4"):
        with rs.Batch(2) as b:
            b.set_split_kernel(kind)
            with pytest.raises(rs.N.RshError):
                add_form(b, probs[3], form)    # M = 161: no specialised resident plan
    for p in probs:
        p["ds"].close()


def test_fp64_batch_cell_limit(gpu):
    """FP64 batch problems hold fewer than 2^24 cells (the iteration kernel's
    32-bit buffer offsets span one row block of 16 rows); larger ones are refused
    before any allocation, and the batch stays usable."""
    m = 26
    xyz, w, F = rs.fibonacci_directions(m)
    with rs.DirectionSet(xyz, w, F, normalize=True) as ds, rs.Batch(2) as b:
        with pytest.raises(rs.N.RshError, match="2\\^24"):
            b.add_seeded(ds, 1 << 24, 1, 0.5)
        s = b.add_seeded(ds, 5000, 2, 0.5)
        b.run(2)
        r, it = b.residuals()
        assert it[s] == 2 and np.isfinite(r[s])


@pytest.mark.parametrize("form", ["fp64", "bf16x3"])
def test_batch_keeps_its_panel_through_cache_eviction(gpu, form):
    """A batch slot holds the phase panel its device problem table points at.
    Filling the direction set's panel cache past its capacity (one entry per
    single-direction output range of scatter_f64) evicts only entries nobody
    else holds; device memory allocated and overwritten afterwards (where a freed
    panel could land) must not change the batch's iterations: the state equals
    that of a batch of the same problem run without the evictions."""
    m, n = 200, 3000
    p = make(m, n, "hg", 0.6, seed=41)
    q = make(m, n, "hg", 0.6, seed=41)
    with rs.Batch(2) as b, rs.Batch(2) as ref:
        dI, dB = D.DeviceArray.from_host(p["I0"]), D.DeviceArray.from_host(p["B"])
        s = b.add(p["ds"], n, dI, dB, 0.5, form=form)
        r = ref.add(q["ds"], n, dI, dB, 0.5, form=form)
        b.run(2)
        b.sync()
        illum, out = D.DeviceArray((m, n)), D.DeviceArray((m, n))
        illum.fill(1.0)
        for k in range(80):                      # 80 distinct panels > the cache's 64
            p["ds"].scatter_f64(illum, 1, n, k, k + 1, out)
        D.synchronize()
        junk = [D.DeviceArray((m * m,)) for _ in range(96)]
        for a in junk:
            a.fill(float("nan"))
        D.synchronize()
        b.run(4)
        ref.run(6)
        got, want = state(b, s, p), state(ref, r, q)
        for a in junk + [illum, out, dI, dB]:
            a.free()
    assert np.array_equal(got, want)
    p["ds"].close()
    q["ds"].close()


def test_batch_autotune_choice_is_used_and_every_tile_agrees(gpu):
    """rsh_batch_autotune times the grouped FP64 tiles on the device and records the
    fastest for the class; batches created afterwards run it (slot_config), and
    every tile gives bit-identical states (same per-output summation order).
    Split forms: the tuner chooses between the resident- and streaming-panel
    grouped kernels per M, and new slots take that kernel."""
    choice, ms = rs.batch_autotune("fp64", 400, n_cells=20000, iterations=4)
    assert choice[0] == "tile" and choice[1] in [(4, 2), (4, 4), (8, 2)] and ms > 0
    assert rs.batch_tuned("fp64", 400) == choice
    p = make(400, 3000, "hg", 0.6, seed=51)
    states = {}
    for tile in [None, (4, 2), (4, 4), (8, 2)]:
        with rs.Batch(2) as b:
            b.set_f64_tile(tile)
            s = add(b, p)
            b.run(3)
            assert b.slot_config(s, "fp64") == ("tile", tile or choice[1])
            states[tile] = state(b, s, p)
    for tile in [(4, 2), (4, 4), (8, 2)]:
        assert np.array_equal(states[tile], states[None])
    for form in ("bf16x3", "fp16x2"):
        kchoice, kms = rs.batch_autotune(form, 400, n_cells=20000, iterations=4)
        assert kchoice[0] == "kernel" and kchoice[1] in ("resident", "resident4", "streaming") and kms > 0
        with rs.Batch(2) as b:
            s = add_form(b, p, form)
            assert b.slot_config(s, form) == kchoice
    p["ds"].close()
