#!/usr/bin/env python3
"""Writes the small golden fixtures under tests/golden/.

The expected outputs are computed here by brute force in exact or very high
precision arithmetic, independently of the C++ oracle and of the HIP library:

* Thomson scattering integral (reference semantics, solver_hllc_illum_cuda/
  solve_module/illum/Illum_async.cpp:201-256 and cuda/cuda_kernel.cu:23-50):
  S[k][c] = (sum_j Gamma(w_j . w_k) * Ibar[c][j] * w_j) / F, with
  Gamma(mu) = 3 (1 + mu^2) / 4 (cuda/cuda_utils.cu:7-16) and Ibar the mean of the
  4 face values (illum_utils.cpp:11-28). Evaluated with fractions.Fraction on the
  exact binary values of the double inputs, then rounded once to double.
* Henyey-Greenstein with FP64-style row normalisation (task extension):
  P[k][j] = p_g(mu_kj) w_j / sum_j' p_g(mu_kj') w_j', evaluated in decimal with
  60 significant digits.
* A 5-step source iteration on cell-level state, I <- (1 - a) B + a S(I), with the
  residual max |I_old - I_new| / I_new of Illum_async.cpp:472-510, in decimal.
* The direction partition GetSend / GetDisp (solve_module/solve_utils.cpp:177-202)
  written as "n // np each, first n % np get one more, displacements = exclusive
  prefix sum" -- a formulation independent of the reference's construction.
* A direction text file in the format of file_module/reader_txt.cpp:25-47.
* Radiation moments (energy, flux, pressure tensor and the two divergences;
  cuda/cuda_kernel.cu:113-226, cuda/cuda_utils.cu:18-167) of random tetrahedra,
  with fractions.Fraction, rounded once.

Inputs are drawn with numpy's seeded generator and stored in the fixtures, so the
fixtures are self-contained. Run from the repo root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import os
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
getcontext().prec = 60


def fibonacci_dirs(m: int):
    """Fibonacci sphere: z_i = 1 - (2i+1)/m, phi_i = i * pi (3 - sqrt 5); w = 4 pi / m."""
    ga = math.pi * (3.0 - math.sqrt(5.0))
    xyz = np.empty((m, 3))
    for i in range(m):
        z = 1.0 - (2.0 * i + 1.0) / m
        r = math.sqrt(max(0.0, 1.0 - z * z))
        ph = ga * i
        xyz[i] = (r * math.cos(ph), r * math.sin(ph), z)
    w = np.full(m, 4.0 * math.pi / m)
    F = 0.0
    for v in w:
        F += float(v)
    return xyz, w, F


def exact_thomson_scatter(illum, xyz, w, F):
    """illum [M][N][4] -> S [M][N], exact rational arithmetic, rounded once."""
    M, N, faces = illum.shape
    Fx = [[Fraction(float(v)) for v in row] for row in xyz]
    Wx = [Fraction(float(v)) for v in w]
    Ff = Fraction(float(F))
    # cell averages, exact
    Ibar = [[sum(Fraction(float(v)) for v in illum[j, c]) / faces for c in range(N)] for j in range(M)]
    gam = [[None] * M for _ in range(M)]
    for k in range(M):
        for j in range(M):
            mu = Fx[j][0] * Fx[k][0] + Fx[j][1] * Fx[k][1] + Fx[j][2] * Fx[k][2]
            gam[k][j] = Fraction(3, 4) * (1 + mu * mu) * Wx[j]
    S = np.empty((M, N))
    for k in range(M):
        gk = gam[k]
        for c in range(N):
            acc = Fraction(0)
            for j in range(M):
                acc += gk[j] * Ibar[j][c]
            S[k, c] = float(acc / Ff)
    return S


def dec(v: float) -> Decimal:
    return Decimal(Fraction(float(v)).numerator) / Decimal(Fraction(float(v)).denominator)


def hg_phase_matrix(xyz, w, g):
    """Row-normalised HG matrix in 60-digit decimal: P[k][j] = p(mu) w_j / sum_j p(mu) w_j."""
    M = xyz.shape[0]
    X = [[dec(v) for v in row] for row in xyz]
    W = [dec(v) for v in w]
    G = dec(g)
    P = []
    for k in range(M):
        row = []
        for j in range(M):
            mu = X[j][0] * X[k][0] + X[j][1] * X[k][1] + X[j][2] * X[k][2]
            d = 1 + G * G - 2 * G * mu
            row.append((1 - G * G) / (d * d.sqrt()) * W[j])
        s = sum(row)
        P.append([v / s for v in row])
    return P


def thomson_matrix_normalised(xyz, w):
    M = xyz.shape[0]
    X = [[dec(v) for v in row] for row in xyz]
    W = [dec(v) for v in w]
    P = []
    for k in range(M):
        row = []
        for j in range(M):
            mu = X[j][0] * X[k][0] + X[j][1] * X[k][1] + X[j][2] * X[k][2]
            row.append(Decimal(3) / 4 * (1 + mu * mu) * W[j])
        s = sum(row)
        P.append([v / s for v in row])
    return P


def dec_scatter(illum, P):
    """illum [M][N][faces] (float), P decimal [R][M] -> S decimal [R][N]."""
    M, N, faces = illum.shape
    Ibar = [[sum(dec(v) for v in illum[j, c]) / faces for c in range(N)] for j in range(M)]
    return [[sum(P[k][j] * Ibar[j][c] for j in range(M)) for c in range(N)] for k in range(len(P))]


def tetra_geometry(rng, n):
    """Random tetrahedra: outward unit normals [n][4][3] (face f opposite vertex
    f), face areas [n][4], volumes [n]."""
    v = rng.uniform(-5, 5, size=(n, 1, 3)) + rng.normal(size=(n, 4, 3))
    normals, areas = np.empty((n, 4, 3)), np.empty((n, 4))
    for f, (a, b, c) in enumerate(((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))):
        cr = np.cross(v[:, b] - v[:, a], v[:, c] - v[:, a])
        ln = np.linalg.norm(cr, axis=1)
        nrm = cr / ln[:, None]
        flip = np.einsum("ij,ij->i", nrm, v[:, f] - (v[:, a] + v[:, b] + v[:, c]) / 3.0) > 0
        nrm[flip] *= -1.0
        normals[:, f], areas[:, f] = nrm, 0.5 * ln
    e1, e2, e3 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]
    volume = np.abs(np.einsum("ij,ij->i", e1, np.cross(e2, e3))) / 6.0
    return normals, areas, volume


def exact_moments(illum, xyz, w, F, normals, areas, volume):
    """Energy, face-averaged flux and pressure tensor, and their discrete
    divergences over the cell faces, in exact rational arithmetic:
      T_f = sum_j w_j phi(omega_j) I_jf / F for phi = 1, omega, omega omega^T;
      energy = mean_f T_f[1]; stream = mean_f T_f[omega]; impuls = mean_f T_f[omega omega^T];
      divstream = sum_f (T_f[omega] . n_f) A_f / V; divimpuls = sum_f T_f[omega omega^T] n_f A_f / V."""
    M, N, faces = illum.shape
    X = [[Fraction(float(v)) for v in r] for r in xyz]
    W = [Fraction(float(v)) for v in w]
    Ff = Fraction(float(F))
    out = {k: np.empty(s) for k, s in (("energy", N), ("stream", (N, 3)), ("impuls", (N, 9)),
                                       ("divstream", N), ("divimpuls", (N, 3)))}
    for c in range(N):
        I = [[Fraction(float(illum[j, c, f])) for f in range(faces)] for j in range(M)]
        e = [sum(W[j] * I[j][f] for j in range(M)) / Ff for f in range(faces)]
        s = [[sum(W[j] * X[j][k] * I[j][f] for j in range(M)) / Ff for k in range(3)] for f in range(faces)]
        t = [[[sum(W[j] * X[j][h] * X[j][k] * I[j][f] for j in range(M)) / Ff for k in range(3)]
              for h in range(3)] for f in range(faces)]
        nr = [[Fraction(float(normals[c, f, k])) for k in range(3)] for f in range(faces)]
        A = [Fraction(float(areas[c, f])) for f in range(faces)]
        V = Fraction(float(volume[c]))
        out["energy"][c] = float(sum(e) / faces)
        for k in range(3):
            out["stream"][c, k] = float(sum(s[f][k] for f in range(faces)) / faces)
        for h in range(3):
            for k in range(3):
                out["impuls"][c, 3 * h + k] = float(sum(t[f][h][k] for f in range(faces)) / faces)
        out["divstream"][c] = float(sum(sum(s[f][k] * nr[f][k] for k in range(3)) * A[f]
                                        for f in range(faces)) / V)
        for h in range(3):
            out["divimpuls"][c, h] = float(sum(sum(t[f][h][k] * nr[f][k] for k in range(3)) * A[f]
                                               for f in range(faces)) / V)
    return out


def main():
    rng = np.random.default_rng(0x5EED)

    # 1. Thomson, reference semantics, N = 6 * 2^3 + 16 (ragged) = 64, M = 26.
    m, n = 26, 64
    xyz, w, F = fibonacci_dirs(m)
    illum = rng.lognormal(0.0, 1.0, size=(m, n, 4))
    S = exact_thomson_scatter(illum, xyz, w, F)
    np.savez(os.path.join(HERE, "tiny_thomson_N64_M26.npz"), illum=illum, xyz=xyz, w=w,
             F=np.array(F), S=S)

    # 1b. Thomson with directions that are not a Fibonacci set (random unit vectors,
    # unequal weights, F read "from a file" != sum w), ragged N = 37, M = 11.
    m2, n2 = 11, 37
    v = rng.normal(size=(m2, 3))
    xyz2 = v / np.linalg.norm(v, axis=1, keepdims=True)
    w2 = rng.uniform(0.5, 1.5, size=m2)
    F2 = 12.5
    illum2 = rng.lognormal(0.0, 1.0, size=(m2, n2, 4))
    S2 = exact_thomson_scatter(illum2, xyz2, w2, F2)
    np.savez(os.path.join(HERE, "tiny_thomson_random_N37_M11.npz"), illum=illum2, xyz=xyz2, w=w2,
             F=np.array(F2), S=S2)

    # 2. HG g = 0.8, FP64 row normalisation, N = 48, M = 26.
    n3 = 48
    illum3 = rng.lognormal(0.0, 1.5, size=(m, n3, 4))
    P3 = hg_phase_matrix(xyz, w, 0.8)
    S3 = dec_scatter(illum3, P3)
    np.savez(os.path.join(HERE, "tiny_hg08_norm_N48_M26.npz"), illum=illum3, xyz=xyz, w=w, F=np.array(F),
             g=np.array(0.8), P=np.array([[float(x) for x in r] for r in P3]),
             S=np.array([[float(x) for x in r] for r in S3]))

    # 3. Source iteration, cell-level state, Thomson normalised, N = 32, M = 14, a = 0.5.
    m4, n4, a = 14, 32, 0.5
    xyz4, w4, F4 = fibonacci_dirs(m4)
    B = rng.uniform(0.5, 2.0, size=(m4, n4))
    I0 = rng.lognormal(0.0, 1.0, size=(m4, n4))
    P4 = thomson_matrix_normalised(xyz4, w4)
    A = Decimal(1) - dec(a)
    ad = dec(a)
    Bd = [[dec(x) for x in r] for r in B]
    I = [[dec(x) for x in r] for r in I0]
    trace, resid = [], []
    for _ in range(5):
        S4 = [[sum(P4[k][j] * I[j][c] for j in range(m4)) for c in range(n4)] for k in range(m4)]
        In = [[A * Bd[k][c] + ad * S4[k][c] for c in range(n4)] for k in range(m4)]
        r = max(abs(I[k][c] - In[k][c]) / In[k][c] for k in range(m4) for c in range(n4))
        resid.append(float(r))
        I = In
        trace.append([[float(x) for x in row] for row in I])
    np.savez(os.path.join(HERE, "iter5_thomson_norm_N32_M14.npz"), xyz=xyz4, w=w4, F=np.array(F4),
             B=B, I0=I0, albedo=np.array(a), trace=np.array(trace), resid=np.array(resid))

    # 4. Direction partition (send counts and displacements).
    cases = [(1, 10), (4, 10), (5, 400), (3, 7), (8, 5), (7, 400), (6, 126), (5, 364), (9, 9), (2, 1)]
    rows = []
    for np_, nn in cases:
        send = [nn // np_ + (1 if i < nn % np_ else 0) for i in range(np_)]
        disp = [sum(send[:i]) for i in range(np_)]
        rows.append((np_, nn, send, disp))
    maxp = max(c[0] for c in cases)
    np.savez(os.path.join(HERE, "partition.npz"),
             np=np.array([r[0] for r in rows], dtype=np.int32),
             n=np.array([r[1] for r in rows], dtype=np.int32),
             send=np.array([r[2] + [-1] * (maxp - r[0]) for r in rows], dtype=np.int32),
             disp=np.array([r[3] + [-1] * (maxp - r[0]) for r in rows], dtype=np.int32))

    # 5. Direction text file (reader_txt.cpp:25-47): M, M lines "area x y z", full_area.
    with open(os.path.join(HERE, "dirs_M26.txt"), "w") as f:
        f.write(f"{m}\n")
        for i in range(m):
            f.write(f"{float(w[i])!r} {float(xyz[i, 0])!r} {float(xyz[i, 1])!r} {float(xyz[i, 2])!r}\n")
        f.write(f"{F!r}\n")
    # 6. Radiation moments (cuda/cuda_kernel.cu:113-226), exact rational
    # arithmetic, N = 7 random tetrahedra, M = 14 random directions with unequal
    # weights, F != sum w. Own generator, so the fixtures above do not move.
    rng6 = np.random.default_rng(0x3017)
    m6, n6 = 14, 7
    v6 = rng6.normal(size=(m6, 3))
    xyz6 = v6 / np.linalg.norm(v6, axis=1, keepdims=True)
    w6 = rng6.uniform(0.5, 1.5, size=m6)
    F6 = 13.25
    illum6 = rng6.lognormal(0.0, 1.0, size=(m6, n6, 4))
    normals6, areas6, volume6 = tetra_geometry(rng6, n6)
    mom = exact_moments(illum6, xyz6, w6, F6, normals6, areas6, volume6)
    np.savez(os.path.join(HERE, "moments_N7_M14.npz"), illum=illum6, xyz=xyz6, w=w6, F=np.array(F6),
             normals=normals6, areas=areas6, volume=volume6, **mom)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
