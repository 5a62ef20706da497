#!/usr/bin/env python3
"""Benchmark of the scattering integral on one MI355X.

Workload (the configuration the headline metric is quoted on): 64^3 hexes x 6
tets = 1,572,864 cells x 4 faces, M = 400 Fibonacci directions with equal
weights, Thomson phase function, face intensities log-normal(0, 1.5) generated
on the device (seeded). One step = one evaluation of
S[k][c] = sum_j P[k][j] * mean_f I[j][c][f] for all 400 output directions, FP64
in and out, inputs resident in HBM before the timed region.

Prints ONE JSON line. Extra keys:
  roofline      dominant kernel (the FP64 MFMA GEMM) vs the FP64 MFMA peak,
                algorithmic flops / HIP-event kernel time on its own stream
  parity        max relative error on sampled cells vs the CPU oracle
  cpu_baseline  the oracle's restatement of the reference CPU path
                (GetS + CalculateIntCPU), timed on this host on a slice of
                output directions over all cells, scaled to a full evaluation
  moments       radiation moments of the same intensities (HBM roofline)
  split         split-precision forms (HG g = 0.8 on the same intensities)
  batch, serve  batched source iteration of uneven problems; serving trace
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "cell·dir pairs/s + %roofline, FP64 vs split-BF16 @≤1e-6 rel; serve p99 ms"
FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(REPO, "profiles", "pmc_mfma_f64_1572864x400.json")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cpu_info() -> dict:
    """CPU model, logical CPUs (nproc), CPUs this process may run on, physical cores."""
    model, phys = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            pid = core = None
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    core = v
                elif not k and pid is not None:
                    phys.add((pid, core))
                    pid = core = None
            if pid is not None:
                phys.add((pid, core))
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "physical_cores": len(phys) or None}


def cpu_baseline(illum_dev, xyz, w, F, n, m, threads, k_slice, cell_frac):
    """Times the oracle's restatement of the reference CPU path (GetS + CalculateIntCPU,
    solve_module/illum/Illum_async.cpp:201-256: OpenMP over output directions, then cells, then
    input directions) on a slice: k_slice output directions x the first n / cell_frac cells of the
    full mesh (same array strides) x all m input directions; 1 warm-up and 5 timed runs, median,
    for the -O3 and -Ofast (the reference's host flag) builds on `threads` threads and the -O3
    build on one thread (one output direction). The cost is exactly linear in output directions
    and cells (every (cell, direction) pair is the same 4 m loads and m phase evaluations), so
    each median is scaled to one full evaluation by (m / directions) x (n / cells)."""
    from oracle import orc

    log(f"[bench] cpu baseline: copying {illum_dev.nbytes / 1e9:.1f} GB of faces to host")
    host = illum_dev.to_host()
    c1 = n // cell_frac
    S = np.zeros((k_slice, n))

    def median_s(variant, nthreads, kcount):
        orc.load(variant)
        times = []
        for r in range(6):  # 1 warm-up + 5 timed
            t0 = time.perf_counter()
            orc.scatter_reference(host, xyz, w, F, k_disp=0, k_count=kcount, cell_begin=0, cell_end=c1,
                                  threads=nthreads, variant=variant, out=S)
            if r:
                times.append(time.perf_counter() - t0)
        return float(np.median(times)), times

    res = {}
    for label, variant, nthreads, kcount in (("o3", "o3", threads, k_slice), ("ofast", "ofast", threads, k_slice),
                                             ("o3_1core", "o3", 1, 1)):
        med, times = median_s(variant, nthreads, kcount)
        full_s = med * (m / kcount) * (n / c1)
        res[label] = {"pairs_per_s": n * m / full_s, "threads": nthreads, "output_directions": kcount,
                      "cells": c1, "median_s": med, "runs_s": times, "full_evaluation_s": full_s}
        log(f"[bench] cpu baseline {label}: {nthreads} threads, {kcount} dirs x {c1} cells: median {med:.3f} s "
            f"-> {n * m / full_s:.3e} pairs/s")
    del host
    info = host_cpu_info()
    return {
        "value": res["o3"]["pairs_per_s"],
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "build": "g++ -O3 -march=x86-64-v3 -fopenmp (value); -Ofast and 1-core figures beside it",
        "host": info,
        "ofast": res["ofast"],
        "o3": res["o3"],
        "one_core": res["o3_1core"],
        "sample": f"{k_slice} of {m} output directions x cells [0, {c1}) of {n} (full-mesh strides) x {m} input "
                  f"directions on {threads} threads (one OpenMP thread per output direction, the reference's "
                  f"schedule); 1 warm-up + 5 timed runs, median, scaled x{m / k_slice:g} x{n / c1:g} to one full "
                  f"evaluation; one-core figure: 1 output direction, same cells",
    }


def split_section(D, rs, illum, out, xyz, w, F, n, m, steps, warmup, stream, cells):
    """Henyey-Greenstein workload: same mesh and intensities, HG g = 0.8 with FP64 row normalisation;
    FP64 MFMA vs the split forms (FP32 MFMA accumulation). The FP64 result is pinned to the oracle on
    sampled cells; every split result (all n x m of them) is compared with it on the device."""
    from oracle import orc

    res = {"workload": "same 1,572,864-cell x 4-face mesh and intensities, M=400, "
                       "Henyey-Greenstein g=0.8, FP64 row normalisation",
           "phase": "hg", "g": 0.8, "forms": {}}
    ds = rs.DirectionSet(xyz, w, F, phase="hg", g=0.8, normalize=True)
    ref64 = D.DeviceArray((m, n))
    e0, e1 = D.Event(), D.Event()
    for _ in range(warmup):
        ds.scatter_f64(illum, 4, n, 0, m, ref64, stream=stream)
    e0.record(stream)
    for _ in range(steps):
        ds.scatter_f64(illum, 4, n, 0, m, ref64, stream=stream)
    e1.record(stream)
    t64 = e0.elapsed_ms(e1) / steps
    res["fp64_ms"] = t64
    P = orc.phase_matrix(xyz, w, F, 1, 0.8, True)
    ref_or = orc.scatter_matrix(illum.gather_cells(cells), P, threads=16)
    r64 = ref64.gather_columns(cells)
    res["fp64_vs_oracle"] = {"sampled_cells": int(len(cells)),
                             "max_rel_err": float(np.max(np.abs(r64 - ref_or) / ref_or)), "tolerance": 1e-12}
    log(f"[bench] HG g=0.8 FP64: {t64:.3f} ms, sampled max rel err vs oracle "
        f"{res['fp64_vs_oracle']['max_rel_err']:.2e}")
    out32 = D.DeviceArray((m, n), np.float32)

    def timed(fn):
        for _ in range(warmup):
            fn()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        t_ms = e0.elapsed_ms(e1) / steps
        stream.synchronize()
        return t_ms

    def errors(o):
        full = rs.rel_error_dev(o, ref64, m, n, threshold=1e-6, stream=stream)
        got_or = o.gather_columns(cells).astype(np.float64)
        r_or = np.abs(got_or - ref_or) / ref_or
        return {"max_rel_err": full["max_rel_err"], "mean_rel_err": full["mean_rel_err"],
                "results_above_1e-6": full["count_above"], "results_compared": full["results"],
                "max_rel_err_vs_oracle_sampled": float(r_or.max())}

    # Resident split state: the face average packed once (as the source
    # iteration keeps it), contracted by the default split kernel.
    for form in ("bf16x2", "fp16x2", "bf16x3"):
        sb, cb = rs.split_sizes(form, 1, n, m)
        state = D.DeviceArray((sb,), np.uint8)
        scale = D.DeviceArray((cb // 4,), np.float32)
        rs.pack_split_avg_dev(form, illum, n, m, state, scale, stream=stream)  # warm-up
        pack_ms = timed(lambda: rs.pack_split_avg_dev(form, illum, n, m, state, scale, stream=stream))
        pack_bytes = m * n * 4 * 8 + sb + n * 4
        entry = {"state": "cell-level (face average packed once)", "state_bytes": sb, "pack_ms": pack_ms,
                 "pack_hbm_gbs_algorithmic": pack_bytes / (pack_ms * 1e-3) / 1e9,
                 "pack": "one pass: faces read once, FP64 average, cell scale, MFMA-layout parts",
                 "kernel": "resident-panel"}
        for label, o, ob in (("fp64_out", out, 8), ("fp32_out", out32, 4)):
            t = timed(lambda: ds.scatter_split(form, state, scale, 1, n, 0, m, o, out_f32=(ob == 4), stream=stream))
            byts = sb + n * m * ob
            entry[label] = {"ms": t, "speedup_vs_fp64": t64 / t, "pairs_per_s": n * m / (t * 1e-3),
                            "hbm_gbs_algorithmic": byts / (t * 1e-3) / 1e9,
                            "roofline_frac_hbm": byts / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": byts,
                            "end_to_end_ms_from_fp64_faces": pack_ms + t,
                            "end_to_end_speedup_vs_fp64": t64 / (pack_ms + t), **errors(o)}
            log(f"[bench] {form} cell state {label}: {t:.3f} ms ({t64 / t:.2f}x FP64), pack {pack_ms:.3f} ms; "
                f"all {n * m} results: max rel err {entry[label]['max_rel_err']:.2e}, "
                f"mean {entry[label]['mean_rel_err']:.2e}, above 1e-6: {entry[label]['results_above_1e-6']}")
        entry["err"] = (f"every one of the {n * m} results vs the FP64 MFMA result (itself within "
                        f"{res['fp64_vs_oracle']['max_rel_err']:.1e} of the oracle on {len(cells)} sampled cells); "
                        f"also vs the oracle on those sampled cells")
        res["forms"][form] = entry
        state.free()
        scale.free()
    # Face-level split state (4 faces per cell kept apart, summed in the kernel).
    res["face_state"] = {}
    for form in ("bf16x2", "fp16x2", "bf16x3"):
        sb, cb = rs.split_sizes(form, 4, n, m)
        state = D.DeviceArray((sb,), np.uint8)
        scale = D.DeviceArray((cb // 4,), np.float32)
        rs.pack_split_dev(form, illum, 4, n, m, state, scale, stream=stream)
        t = timed(lambda: ds.scatter_split(form, state, scale, 4, n, 0, m, out, stream=stream))
        byts = sb + n * m * 8
        res["face_state"][form] = {"ms": t, "speedup_vs_fp64": t64 / t, "state_bytes": sb,
                                   "hbm_gbs_algorithmic": byts / (t * 1e-3) / 1e9, **errors(out)}
        log(f"[bench] {form} face state: {t:.3f} ms ({t64 / t:.2f}x FP64)")
        state.free()
        scale.free()
    out32.free()
    ref64.free()
    ds.close()
    return res


def moments_section(D, rs, ds, illum, xyz, w, F, n, m, steps, stream, cells):
    """Radiation moments of the same resident face intensities (energy, stream,
    impuls, divstream, divimpuls; the reference's CalculateAllParam), with
    synthetic per-cell geometry; HBM-bound, so its roofline is the HBM peak."""
    from oracle import orc

    rng = np.random.default_rng(0x5EED + 6)
    nrm = rng.normal(size=(n, 4, 3))
    nrm /= np.linalg.norm(nrm, axis=2, keepdims=True)
    areas = rng.uniform(0.5, 1.5, size=(n, 4))
    volume = rng.uniform(0.5, 1.5, size=n)
    geo = [D.DeviceArray.from_host(a) for a in (nrm, areas, volume)]
    width = {"energy": 1, "stream": 3, "impuls": 9, "divstream": 1, "divimpuls": 3}
    outs = {k: D.DeviceArray((n, v)) for k, v in width.items()}

    def go():
        ds.moments(illum, 4, n, energy=outs["energy"], stream_out=outs["stream"], impuls=outs["impuls"],
                   divstream=outs["divstream"], divimpuls=outs["divimpuls"], normals=geo[0], areas=geo[1],
                   volume=geo[2], stream=stream)

    go()
    e0, e1 = D.Event(), D.Event()
    e0.record(stream)
    for _ in range(steps):
        go()
    e1.record(stream)
    t = e0.elapsed_ms(e1) / steps
    stream.synchronize()
    byts = 8 * (4 * m * n + 17 * n + 17 * n)  # faces read once, geometry read, five outputs written
    ref = orc.moments(illum.gather_cells(cells), xyz, w, F, nrm[cells], areas[cells], volume[cells], threads=16)
    sub = illum.gather_cells(cells)
    e = np.max(np.einsum("j,jcf->cf", w, np.abs(sub)) / F, axis=1)
    d = e * areas[cells].sum(axis=1) / volume[cells]
    worst = 0.0
    for k, v in width.items():
        got = outs[k].to_host()[cells]
        got = got[:, 0] if v == 1 else got
        err = np.abs(got - ref[k])
        err = err if err.ndim == 1 else err.max(axis=1)
        worst = max(worst, float(np.max(err / (d if k.startswith("div") else e))))
    for a in [*geo, *outs.values()]:
        a.free()
    traffic = None
    pmc_path = os.path.join(REPO, "profiles", "pmc_moments_f64_1572864x400.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            pmc = json.load(f)
        if pmc.get("n_cells") == n and pmc.get("n_dirs") == m:
            traffic = pmc.get("hbm_bytes_per_launch")
    log(f"[bench] moments: {t:.3f} ms, {byts / (t * 1e-3) / 1e9:.0f} GB/s, scaled max err {worst:.2e}")
    return {"workload": "radiation moments (energy, stream, impuls, divstream, divimpuls) of the same "
                        "1,572,864-cell x 4-face intensities, M=400, synthetic cell geometry",
            "kernel": "moments_f64 (FP64 VALU, HBM-bound)", "ms": t, "cells_per_s": n / (t * 1e-3),
            "roofline": {"bound": "hbm", "achieved": byts / (t * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": byts / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
                         "algorithmic_bytes_per_launch": byts},
            "parity": {"sampled_cells": int(len(cells)), "max_scaled_err_vs_oracle": worst, "tolerance": 1e-12,
                       "scale": "per cell: max_f sum_j w_j |I_jf| / F (x sum_f area_f / volume for divergences)"}}


def sweep_section(D, rs, edge=32, m=400, steps=5, cpu=True):
    """Transport sweep (short characteristics, next row after the moments): a 6 edge^3-cell
    tetrahedral cube, M Fibonacci directions, the test-task class (Q = 1, alpha = 1, beta =
    1/2), synthetic log-normal scattering; one launch sweeps every direction and computes
    ReCalcIllum's residual. Also one full transport iteration (sweep + FP64 scattering)."""
    from oracle import orc
    from relativistic_solver_hllc_amd import sweep as sw

    v, pairs = sw.mesh_cube_tetra(edge)
    n = v.shape[0]
    xyz, w, F = rs.fibonacci_directions(m)
    t0 = time.perf_counter()
    per = sw.prepare(v, pairs, xyz, threads=16)
    prep_s = time.perf_counter() - t0
    sp = sw.Sweep(n, m, pairs)
    t0 = time.perf_counter()
    sp.set_all(per)
    upload_s = time.perf_counter() - t0
    levels = np.array([sp.levels(k) for k in range(m)])
    rng = np.random.default_rng(0x5EED + 7)
    S = rng.lognormal(size=(m, n))
    old = np.ones((m, n, 4))
    d_S, d_old = D.DeviceArray.from_host(S), D.DeviceArray.from_host(old)
    d_new, d_norm = D.DeviceArray((m, n, 4)), D.DeviceArray((m,))
    stream = D.Stream(0)
    sp.run(1, d_S, d_new, d_old, d_norm, stream=stream)
    stream.synchronize()
    e0, e1 = D.Event(), D.Event()
    e0.record(stream)
    for _ in range(steps):
        sp.run(1, d_S, d_new, d_old, d_norm, stream=stream)
    e1.record(stream)
    stream.synchronize()
    t = e0.elapsed_ms(e1) / steps
    got = d_new.to_host()
    norm = float(d_norm.to_host().max())
    # parity: the sequential oracle on a slice of directions
    sub = list(range(0, m, max(1, m // 16)))[:16]
    ref, ref_norm = orc.sweep([per[k] for k in sub], pairs, S[sub], old[sub], cls=1, threads=16)
    err = float(np.max(np.abs(got[sub] - ref) / np.maximum(np.abs(ref), 1e-300)))
    # full transport iteration: sweep, then the FP64 scattering integral of the new intensities
    ds = rs.DirectionSet(xyz, w, F)
    d_Sit = D.DeviceArray((m, n))
    ds.scatter_f64(d_new, 4, n, 0, m, d_Sit, stream=stream)
    stream.synchronize()
    e0.record(stream)
    for _ in range(steps):
        sp.run(1, d_Sit, d_new, d_old, d_norm, stream=stream)
        ds.scatter_f64(d_new, 4, n, 0, m, d_Sit, stream=stream)
    e1.record(stream)
    stream.synchronize()
    t_iter = e0.elapsed_ms(e1) / steps
    # algorithmic bytes: per (cell, direction) the level list entry, record offset, face state and
    # S value; 9 bytes per record; one FP64 read per interior in face; 4 face values written; the
    # residual reads 4 new and 4 old values
    records = sum(len(d[2]) for d in per)
    in_interior = sum(int(np.sum(((d[1][:, None] >> np.arange(4)) & 1).astype(bool) & (pairs >= 0))) for d in per)
    byts = m * n * (4 + 8 + 1 + 8) + records * 9 + in_interior * 8 + m * n * 4 * 8 * 3
    cpu_res = None
    if cpu:
        threads = min(16, host_cpu_info()["affinity_cpus"])
        c0 = time.perf_counter()
        _, _ = orc.sweep(per, pairs, S, old, cls=1, threads=threads, variant="o3")
        c_s = time.perf_counter() - c0
        cpu_res = {"value": n * m / c_s, "unit": "cell-direction updates/s", "cores": threads, "kind": "port",
                   "seconds": c_s, "sample": f"the whole sweep ({m} directions x {n} cells), oracle -O3, "
                                             "OpenMP over directions as the reference"}
    for a in (d_S, d_old, d_new, d_norm, d_Sit):
        a.free()
    ds.close()
    sp.close()
    gbs = byts / (t * 1e-3) / 1e9
    traffic = None
    pmc_path = os.path.join(REPO, "profiles", f"pmc_sweep_{n}x{m}.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            pmc = json.load(f)
        if pmc.get("n_cells") == n and pmc.get("n_dirs") == m:
            traffic = pmc.get("hbm_bytes_per_launch")
    log(f"[bench] sweep {n} cells x {m} dirs: {t:.3f} ms ({n * m / (t * 1e-3):.3e} cell-dir/s, {gbs:.0f} GB/s), "
        f"levels {levels.min()}-{levels.max()}, max rel err vs oracle {err:.2e}; iteration {t_iter:.3f} ms"
        + (f"; cpu {cpu_res['seconds']:.2f} s" if cpu_res else ""))
    return {"workload": f"transport sweep: unit cube in {edge}^3 x 6 = {n} tetrahedra, M={m} Fibonacci directions, "
                        "class_vtk 1 (test task), log-normal scattering, ReCalcIllum residual",
            "kernel": "sweep_kernel<1> (one workgroup per direction, level-parallel)", "ms": t,
            "cell_directions_per_s": n * m / (t * 1e-3), "levels_min": int(levels.min()),
            "levels_mean": float(levels.mean()), "levels_max": int(levels.max()),
            "prepare_s": prep_s, "upload_s": upload_s, "records": records,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes_per_launch": byts},
            "iteration_ms": t_iter, "iteration": "sweep + FP64 scattering integral (mfma_f64) of the new faces",
            "residual": norm,
            "parity": {"directions_checked": len(sub), "max_rel_err_vs_oracle": err, "tolerance": 1e-12,
                       "oracle_residual": ref_norm},
            "cpu_baseline": cpu_res}


def pcie_section(D, rs, illum, xyz, w, F, n, m):
    """The reference's host-buffer path through the problem API: the faces go up from pinned host memory
    (CudaSendIllumAsync -> rsh_upload_illum on stream 0), S is computed on streams 1 and 2 in two halves of
    the output directions and each half is copied back into pinned host memory as soon as it is done
    (CalculateIntScatteringAsyncMPIStream -> rsh_scatter with host output), timed with the host clock.
    This is the PCIe-inclusive rate; the headline value keeps the inputs resident in HBM."""
    host = D.PinnedArray((m, n, 4))
    s_host = D.PinnedArray((m, n))
    rs.N.call("rsh_dev_memcpy", host.data_ptr(), illum.data_ptr(), m * n * 4 * 8, 1, None)  # device -> pinned host
    D.synchronize()
    with rs.Problem(n, xyz, w, F, 0, m) as p:
        def once():
            t0 = time.perf_counter()
            p.send_illum_async(host, m * n * 4, stream=rs.STREAM_SCATTERING_1)
            p.sync_stream(rs.STREAM_SCATTERING_1)
            t1 = time.perf_counter()
            p.calculate_int_scattering(0, m // 2, rs.STREAM_SCATTERING_2, host_out=s_host)
            p.calculate_int_scattering(m // 2, m, rs.STREAM_PARAMS, host_out=s_host)
            p.sync_stream(rs.STREAM_SCATTERING_2)
            p.sync_stream(rs.STREAM_PARAMS)
            t2 = time.perf_counter()
            return t1 - t0, t2 - t1
        once()
        up, rest = once()
        ok = bool(np.isfinite(s_host.a[:: max(1, m // 7), :: max(1, n // 1000)]).all())
        # CalculateIntScattering's sequence pipelined by cell range (rsh_scatter_from_host)
        chunks = 32
        p.calculate_int_scattering_from_host(host, 0, m, host_out=s_host, chunks=chunks)
        t0 = time.perf_counter()
        p.calculate_int_scattering_from_host(host, 0, m, host_out=s_host, chunks=chunks)
        piped = time.perf_counter() - t0
        ok = ok and bool(np.isfinite(s_host.a[:: max(1, m // 7), :: max(1, n // 1000)]).all())
    host.free()
    s_host.free()
    total = up + rest
    log(f"[bench] pcie path: upload {up * 1e3:.1f} ms ({m * n * 32 / up / 1e9:.1f} GB/s), scatter + D2H "
        f"{rest * 1e3:.1f} ms, total {total * 1e3:.1f} ms = {n * m / total:.3e} pairs/s; pipelined over {chunks} "
        f"cell ranges {piped * 1e3:.1f} ms = {n * m / piped:.3e} pairs/s")
    return {"workload": "same evaluation through the reference's host-buffer API: 20.1 GB of faces uploaded from "
                        "pinned host memory, S (5.0 GB) copied back to pinned host memory in two halves",
            "upload_ms": up * 1e3, "upload_gbs": m * n * 32 / up / 1e9, "scatter_and_d2h_ms": rest * 1e3,
            "total_ms": total * 1e3, "pairs_per_s_pcie_inclusive": n * m / total,
            "pipelined": {"chunks": chunks, "total_ms": piped * 1e3, "pairs_per_s_pcie_inclusive": n * m / piped,
                          "how": "upload, integral and copy-back of 32 cell ranges overlapped on three streams "
                                 "(rsh_scatter_from_host)"},
            "finite_output": ok}


def uneven_problems(seed=0x5EED + 4, count=32, n_lo=1e4, n_hi=5e5):
    """Uneven problems: N log-uniform in [n_lo, n_hi] rounded to a structured cube 6 n^3,
    M in {200, 400, 800}, phase Thomson or HG g in {0.3, 0.6, 0.9}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n_target = float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        edge = max(1, int(round((n_target / 6.0) ** (1.0 / 3.0))))
        m = int(rng.choice([200, 400, 800]))
        ph = int(rng.integers(0, 4))
        phase, g = ("thomson", 0.0) if ph == 0 else ("hg", (0.3, 0.6, 0.9)[ph - 1])
        out.append(dict(n=6 * edge ** 3, m=m, phase=phase, g=g, seed=int(rng.integers(0, 1 << 30))))
    return out


def batch_section(D, rs, iters=20):
    """Batch workload: 32 uneven problems, `iters` source iterations I <- 0.5 B + 0.5 S each,
    state resident on the device, grouped graph-replayed launches; FP64 and split forms."""
    probs = uneven_problems()
    dirsets = {}
    for p in probs:
        key = (p["m"], p["phase"], p["g"])
        if key not in dirsets:
            xyz, w, F = rs.fibonacci_directions(p["m"])
            dirsets[key] = rs.DirectionSet(xyz, w, F, phase=p["phase"], g=p["g"], normalize=True)
    pairs = sum(p["n"] * p["m"] for p in probs)
    flops = sum(2.0 * p["m"] ** 2 * p["n"] for p in probs)
    res = {"workload": "32 problems, N log-uniform in [1e4, 5e5] as 6 n^3 cells, M in {200,400,800}, "
                       "Thomson or HG g in {0.3,0.6,0.9} (FP64 row-normalised), I0 ~ log-normal(0,1), "
                       "B ~ U(0.5,2), albedo 0.5, 20 iterations",
           "problems": len(probs), "cells_total": sum(p["n"] for p in probs), "pairs_per_iteration": pairs,
           "iterations": iters,
           "fp64_grouped_tile": "autotuned (below), <= 13 row blocks (208 rows) per pass",
           "forms": {}}
    # Grouped-launch configurations chosen by the autotuner on this device (untimed):
    # FP64 per row-block class (M = 200, 400, 800 share one), split forms per M.
    t0 = time.perf_counter()
    tuned = {"fp64": {"400": rs.batch_autotune("fp64", 400, n_cells=300000, iterations=4)[0][1]}}
    for form in ("bf16x2", "fp16x2", "bf16x3"):
        tuned[form] = {str(mm): rs.batch_autotune(form, mm, n_cells=300000, iterations=4)[0][1]
                       for mm in sorted({p["m"] for p in probs})}
    res["autotune"] = {"choices": tuned, "seconds": time.perf_counter() - t0,
                       "candidates": "FP64: MFMA tiles (waves, j-steps per LDS stage) (4,2), (4,4), (8,2); "
                                     "split: resident-panel grouped kernel with 8 or 4 waves vs streaming-panel grouped kernel"}
    log(f"[bench] batch autotune: {tuned}")
    states = {}
    for form in ("fp64", "bf16x2", "fp16x2", "bf16x3"):
        b = rs.Batch(len(probs))
        slots = []
        for p in probs:
            I0, B = D.DeviceArray((p["m"], p["n"])), D.DeviceArray((p["m"], p["n"]))
            rs.fill_lognormal_dev(I0, 0.0, 1.0, seed=p["seed"] * 2 + 1)
            rs.fill_uniform_dev(B, 0.5, 2.0, seed=p["seed"] * 2 + 2)
            slots.append(b.add(dirsets[(p["m"], p["phase"], p["g"])], p["n"], I0, B, 0.5, form=form))
            I0.free()
            B.free()
        b.run(2)  # warm-up: graph capture and first launches
        b.sync()
        t0 = time.perf_counter()
        b.run(iters)
        b.sync()
        dt = time.perf_counter() - t0
        r, it = b.residuals()
        probe = []
        for i in (0, 7, 19):
            out = D.DeviceArray((probs[i]["m"], probs[i]["n"]))
            b.state(slots[i], out)
            probe.append(out.to_host())
            out.free()
        states[form] = probe
        entry = {"seconds": dt, "ms_per_iteration": dt * 1e3 / iters, "pairs_per_s": pairs * iters / dt,
                 "tflops_useful": flops * iters / dt / 1e12, "max_last_residual": float(np.max(r[: len(probs)]))}
        if form != "fp64":
            errs = [float(np.max(np.abs(a - f) / np.abs(f))) for a, f in zip(probe, states["fp64"])]
            entry["max_rel_err_vs_fp64_state"] = max(errs)
            entry["speedup_vs_fp64"] = res["forms"]["fp64"]["seconds"] / dt
        res["forms"][form] = entry
        log(f"[bench] batch {form}: {dt * 1e3 / iters:.2f} ms/iteration, {pairs * iters / dt:.3e} pairs/s")
        b.close()
    for ds in dirsets.values():
        ds.close()
    return res


def serve_section(rs, rate=10.0, count=64, budget_gb=200):
    """Serving workload: real-time Poisson arrivals of seeded requests drawn as in the batch workload with
    N in [1e4, 1e6]; each iterates to its FP64 residual tolerance; continuous batching."""
    res = {"workload": f"Poisson arrivals at {rate:g}/s, {count} requests, N log-uniform in [1e4, 1e6] "
                       "as 6 n^3, M in {200,400,800}, Thomson or HG; tolerance 1e-8 (fp64) / 1e-6 (split); "
                       f"HBM budget {budget_gb} GB, 64 slots, 2 iterations per graph launch",
           "forms": {}}
    probs = uneven_problems(seed=0x5EED + 5, count=count, n_lo=1e4, n_hi=1e6)
    rng = np.random.default_rng(0x5EED + 55)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    for form, tol in (("fp64", 1e-8), ("bf16x3", 1e-6)):
        descs = [rs.request(p["n"], p["m"], p["phase"], p["g"], tolerance=tol, form=form, seed=p["seed"])
                 for p in probs]
        srv = rs.Server(int(budget_gb * (1 << 30)), 64, iters_per_step=2)
        lat, wall = srv.run_trace(descs, arrivals)
        infos = [srv.poll(i) for i in range(count)]
        srv.close()
        pairs_it = sum(p["n"] * p["m"] * inf["iterations"] for p, inf in zip(probs, infos))
        res["forms"][form] = {
            "tolerance": tol, "wall_s": wall, "requests_per_s": count / wall,
            "pair_iterations_per_s": pairs_it / wall,
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3), "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "latency_max_ms": float(lat.max() * 1e3),
            "converged": sum(inf["status"] == "converged" for inf in infos),
            "mean_iterations": float(np.mean([inf["iterations"] for inf in infos])),
        }
        log(f"[bench] serve {form}: p50 {res['forms'][form]['latency_p50_ms']:.1f} ms, "
            f"p99 {res['forms'][form]['latency_p99_ms']:.1f} ms, {count / wall:.1f} req/s")
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=6 * 64 ** 3)
    ap.add_argument("--dirs", type=int, default=400)
    ap.add_argument("--tile", type=str, default="", help="waves,chunk_steps (default: library default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-slice", type=int, default=40, help="output directions of the CPU slice (at least)")
    ap.add_argument("--cpu-cell-frac", type=int, default=8, help="the CPU slice covers n / this many cells")
    ap.add_argument("--no-moments", action="store_true", help="skip the radiation-moments section")
    ap.add_argument("--no-split", action="store_true", help="skip the split-precision (HG g=0.8) section")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven problems section")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving section")
    ap.add_argument("--no-sweep", action="store_true", help="skip the transport-sweep section")
    ap.add_argument("--no-pcie", action="store_true", help="skip the host-buffer (PCIe-inclusive) path")
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="before the benchmark, run the plain copy kernel 3x over this many GB "
                         "(reference for FETCH_SIZE / WRITE_SIZE counter scaling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding "
              "of this path is described, not built.", file=sys.stderr)
        return 2

    from relativistic_solver_hllc_amd import device as D
    from relativistic_solver_hllc_amd import scatter as rs

    if args.calibrate_copy > 0:
        cnt = int(args.calibrate_copy * 1e9 / 8) // 2 * 2
        a, b = D.DeviceArray((cnt,)), D.DeviceArray((cnt,))
        a.fill(1.0)
        for _ in range(3):
            b.copy_kernel_from(a)
        D.synchronize()
        log(f"[bench] copy calibration: 3 x {cnt * 8} bytes read and written")
        a.free()
        b.free()

    n, m = args.cells, args.dirs
    xyz, w, F = rs.fibonacci_directions(m)
    log(f"[bench] N={n} M={m}: allocating {4 * m * n * 8 / 1e9:.1f} GB faces + {m * n * 8 / 1e9:.1f} GB output")
    illum = D.DeviceArray((m, n, 4))
    out = D.DeviceArray((m, n))
    stream = D.Stream(0)
    rs.fill_lognormal_dev(illum, 0.0, 1.5, seed=0x5EED + 2, stream=stream)
    stream.synchronize()
    tile = tuple(int(v) for v in args.tile.split(",")) if args.tile else None
    kernel = "mfma" if tile else "auto"

    ds = rs.DirectionSet(xyz, w, F)
    for _ in range(args.warmup):
        ds.scatter_f64(illum, 4, n, 0, m, out, stream=stream, kernel=kernel, tile=tile)
    stream.synchronize()
    log(f"[bench] warmup done ({args.warmup} steps); timing {args.steps} steps")

    e0, e1 = D.Event(), D.Event()
    D.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        ds.scatter_f64(illum, 4, n, 0, m, out, stream=stream, kernel=kernel, tile=tile)
    e1.record(stream)
    stream.synchronize()
    D.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    kern_ms = e0.elapsed_ms(e1) / args.steps
    log(f"[bench] kernel {kern_ms:.3f} ms/step (events), wall {wall_ms:.3f} ms/step")

    # parity on sampled cells against the oracle (checker only)
    from oracle import orc

    rng = np.random.default_rng(12)
    cells = np.unique(np.concatenate([[0, n - 1], rng.integers(0, n, 30)]))
    ref = orc.scatter_reference(illum.gather_cells(cells), xyz, w, F, threads=16)
    got = out.gather_columns(cells)
    parity = float(np.max(np.abs(got - ref) / ref))
    log(f"[bench] sampled-cell max rel err vs oracle: {parity:.3e}")

    pairs = n * m
    flops = 2.0 * m * m * n
    achieved_tflops = flops / (kern_ms * 1e-3) / 1e12
    # the matrix-core rate this device reaches in a register-only loop (no memory traffic)
    peak64 = rs.mfma_peak("fp64")
    peak_bf16 = rs.mfma_peak("bf16")
    log(f"[bench] register-only MFMA loops: FP64 16x16x4 {peak64:.1f} TFLOP/s, BF16 16x16x32 {peak_bf16:.0f} TFLOP/s")
    traffic = None
    if os.path.exists(PMC_SUMMARY):
        with open(PMC_SUMMARY) as f:
            pmc = json.load(f)
        if pmc.get("n_cells") == n and pmc.get("n_dirs") == m:
            traffic = pmc.get("hbm_bytes_per_launch")

    moments = None
    if not args.no_moments:
        moments = moments_section(D, rs, ds, illum, xyz, w, F, n, m, max(args.steps // 2, 2), stream, cells)

    pcie = None
    if not args.no_pcie:
        pcie = pcie_section(D, rs, illum, xyz, w, F, n, m)

    split = None
    if not args.no_split:
        split = split_section(D, rs, illum, out, xyz, w, F, n, m, max(args.steps // 2, 2), 1, stream, cells)

    cpu = None
    if not args.no_cpu:
        # This job's CPU share of the box: the CPUs it may run on, at most 16 (one GPU's share).
        threads = args.cpu_threads or min(16, host_cpu_info()["affinity_cpus"])
        # at least max(40, threads) output directions, a whole number per thread
        k_slice = min(m, -(-max(args.cpu_slice, threads) // threads) * threads)
        cpu = cpu_baseline(illum, xyz, w, F, n, m, threads, k_slice, args.cpu_cell_frac)
        log(f"[bench] cpu baseline: {cpu['value']:.3e} pairs/s on {threads} threads")

    sweep = None
    if not args.no_sweep:
        sweep = sweep_section(D, rs, steps=max(args.steps // 4, 2), cpu=not args.no_cpu)

    batch = None
    serve = None
    if not args.no_batch or not args.no_serve:
        illum.free()
        out.free()
    if not args.no_batch:
        batch = batch_section(D, rs)
    if not args.no_serve:
        serve = serve_section(rs)

    result = {
        "metric": METRIC,
        "value": pairs / (kern_ms * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kern_ms,
        "wall_ms_per_step": wall_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (log-normal(0,1.5) face intensities generated on device, seed 0x5EED+2)",
        "config": {
            "workload": "scattering integral, 64^3x6 = 1,572,864 tet cells x 4 faces, M=400 Fibonacci "
                        "directions, Thomson phase, FP64",
            "n_cells": n,
            "n_dirs": m,
            "faces": 4,
            "phase": "thomson",
            "kernel": "mfma_f64 (v_mfma_f64_16x16x4f64)",
            "tile": list(tile) if tile else "default",
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved_tflops,
            "peak": FP64_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_MFMA_PEAK_TFLOPS,
            "traffic": traffic,
            "peak_measured": peak64,
            "frac_of_measured": achieved_tflops / peak64,
            "peak_measured_how": "register-only v_mfma_f64_16x16x4f64 loop, 8 waves per CU, 8 independent "
                                 "accumulators per wave (rsh_mfma_peak)",
            "bf16_peak_measured": peak_bf16,
            "algorithmic_bytes_per_launch": pairs * 40,
            "algorithmic_flops_per_launch": flops,
        },
        "parity": {"sampled_cells": int(len(cells)), "max_rel_err_vs_oracle": parity, "tolerance": 1e-12},
        "cpu_baseline": cpu,
        "moments": moments,
        "split": split,
        "sweep": sweep,
        "pcie": pcie,
        "batch": batch,
        "serve": serve,
    }
    ds.close()
    illum.free()
    out.free()
    print(json.dumps(result), flush=True)
    return 0 if parity < 1e-12 else 1


if __name__ == "__main__":
    sys.exit(main())
