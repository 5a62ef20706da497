// CPU oracle for the reference's preparation of the transport sweep (its MAKE
// step), which writes the face nodes (X.bin), face states and the per-node
// characteristic records (LocX0<k>.bin) the sweep consumes.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the library's
// rsh_make_nodes / rsh_make_direction are compared against; nothing in the
// product path links or calls it. It is used by tests/.
//
// It restates, in the reference's step order and with its formulas (paths
// relative to solver_hllc_illum_cuda/ of Lartsev-Artem/relativistic_solver_hllc):
//
//   orc_make_nodes      MakeArrayX  make_illum_struct/struct_short_characteristics_logic_function.cpp:10-72
//                       with the BaseTetra_t constants  struct_short_characteristics_calculations.cpp:146-183
//                       and FromLocalToGlobalTetra / FromPlaneToTetra  (same file :102-128)
//   orc_make_direction  the direction loop body of RunMakeModule
//                       make_illum_struct/struct_short_characteristics_main.cpp:128-165:
//                       GetInAndOutFaces  utils/grid_geometry/geometry_solve.cpp:212-227
//                       GetNodes / CalculateNodeValue  struct_short_characteristics_logic_function.cpp:222-296
//                       IntersectionWithPlane, InTriangle (SetBasis, Make2dPoint)  geometry_solve.cpp:8-56,109-159
//                       CalculateIllumeOnInnerFace  struct_short_characteristics_logic_function.cpp:75-220
//                       FromGlobalToLocalTetra, FromTetraToPlane  struct_short_characteristics_calculations.cpp:84-123
//
// Eigen's 4x4 inverse is restated as Gauss-Jordan elimination with partial
// pivoting (the inverse is the same to rounding). Where the reference finds no
// in face for a node it stops the program ("Not found inner face") when face 3
// was an in face, and otherwise writes no record for that node; the oracle
// reports both as status 1 so that a test sees them.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off).

#include <cmath>
#include <cstdint>

namespace {

constexpr double kEps = 1e-10;  // global_value.h:13

struct Base {
    double straight[3][3], inclined[3][3], T[3][3], Tinv[3][3], start[3];
    Base() {
        const double s2 = std::sqrt(2.0), s6 = std::sqrt(6.0), s3 = std::sqrt(3.0), s23 = std::sqrt(2. / 3);
        const double st[3][3] = {{1. / 6, 1. / 6, 1}, {2. / 3, 1. / 6, 1}, {1. / 6, 2. / 3, 1}};
        const double in[3][3] = {{0, s23, 1}, {s2 / 4, 1. / (2 * s6), 1}, {-s2 / 4, 1. / (2 * s6), 1}};
        const double t[3][3] = {{-1. / s2, 1. / s2, 0}, {-1. / s6, -1. / s6, s23}, {1. / s3, 1. / s3, 1. / s3}};
        const double ti[3][3] = {{-1. / s2, -1. / s6, 1. / s3}, {1. / s2, -1. / s6, 1. / s3}, {0, s23, 1. / s3}};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                straight[i][j] = st[i][j];
                inclined[i][j] = in[i][j];
                T[i][j] = t[i][j];
                Tinv[i][j] = ti[i][j];
            }
        start[0] = 0.5;
        start[1] = 0.5;
        start[2] = 0;
    }
};
const Base kBase;

// V(i, j) of the cell's vertex matrix (Eigen column-major: column j = point j).
inline double V(const double* vm, int i, int j) { return vm[j * 4 + i]; }

void local_to_global(const double* vm, const double* l, double* g) {
    const double eta4 = 1 - l[0] - l[1] - l[2];
    for (int i = 0; i < 3; ++i) {
        g[i] = 0;
        for (int j = 0; j < 3; ++j) g[i] += V(vm, i, j) * l[j];
        g[i] += V(vm, i, 3) * eta4;
    }
}

void plane_to_tetra(const double* p, double* t) {
    double r[3];
    for (int i = 0; i < 3; i++) r[i] = kBase.Tinv[i][0] * p[0] + kBase.Tinv[i][1] * p[1] + kBase.Tinv[i][2] * p[2];
    for (int i = 0; i < 3; i++) t[i] = r[i] + kBase.start[i];
}

void tetra_to_plane(const double* t, double* p) {
    double d[3];
    for (int i = 0; i < 3; i++) d[i] = t[i] - kBase.start[i];
    for (int i = 0; i < 3; i++) p[i] = kBase.T[i][0] * d[0] + kBase.T[i][1] * d[1] + kBase.T[i][2] * d[2];
}

bool invert4(const double* vm, double inv[4][4]) {
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            a[i][j] = V(vm, i, j);
            a[i][4 + j] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) return false;
        if (p != c)
            for (int j = 0; j < 8; j++) {
                const double t = a[c][j];
                a[c][j] = a[p][j];
                a[p][j] = t;
            }
        const double d = a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] /= d;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            const double f = a[r][c];
            if (f != 0.0)
                for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) inv[i][j] = a[i][4 + j];
    return true;
}

void global_to_local(const double* vm, const double* g, double* l) {
    double inv[4][4];
    invert4(vm, inv);
    for (int i = 0; i < 3; ++i) {
        l[i] = 0;
        for (int j = 0; j < 3; ++j) l[i] += inv[i][j] * g[j];
        l[i] += inv[i][3];
    }
}

// MakeArrayX for one cell: x[f][node][3].
void cell_nodes(const double* vm, double* x) {
    double node[3];
    for (int f = 0; f < 4; f++)
        for (int k = 0; k < 3; k++) {
            double* out = x + (f * 3 + k) * 3;
            switch (f) {
                case 1:
                    node[0] = 0;
                    node[1] = kBase.straight[k][0];
                    node[2] = kBase.straight[k][1];
                    local_to_global(vm, node, out);
                    break;
                case 2:
                    node[0] = kBase.straight[k][0];
                    node[1] = 0;
                    node[2] = kBase.straight[k][1];
                    local_to_global(vm, node, out);
                    break;
                case 0:
                    node[0] = kBase.straight[k][0];
                    node[1] = kBase.straight[k][1];
                    node[2] = 0;
                    local_to_global(vm, node, out);
                    break;
                default: {
                    node[0] = kBase.inclined[k][0];
                    node[1] = kBase.inclined[k][1];
                    node[2] = 0;
                    double t[3];
                    plane_to_tetra(node, t);
                    local_to_global(vm, t, out);
                    break;
                }
            }
        }
}

// IntersectionWithPlane: face = A, B, C (9 doubles).
void intersect(const double* F, const double* x, const double* w, double* res) {
    const double *A = F, *B = F + 3, *C = F + 6;
    const double a = A[1] * (B[2] - C[2]) + B[1] * (C[2] - A[2]) + C[1] * (A[2] - B[2]);
    const double b = A[0] * (C[2] - B[2]) + B[0] * (A[2] - C[2]) + C[0] * (B[2] - A[2]);
    const double c = A[0] * (B[1] - C[1]) + B[0] * (C[1] - A[1]) + C[0] * (A[1] - B[1]);
    const double d = A[0] * (C[1] * B[2] - B[1] * C[2]) + B[0] * (A[1] * C[2] - C[1] * A[2]) +
                     C[0] * (B[1] * A[2] - A[1] * B[2]);
    const double t = -(a * x[0] + b * x[1] + c * x[2] + d) / (a * w[0] + b * w[1] + c * w[2]);
    for (int i = 0; i < 3; ++i) res[i] = w[i] * t + x[i];
}

// SetBasis: rows vec_1, vec_2, normal.
void set_basis(const double* n, double basis[3][3]) {
    double v1[3], v2[3];
    if (std::fabs(n[1]) < 1e-20) {
        v1[0] = 0;
        v1[1] = 1;
        v1[2] = 0;
    } else {
        v1[0] = 1;
        v1[2] = 0;
        v1[1] = -(n[0] * v1[0] + n[2] * v1[2]) / n[1];
    }
    if (n[1] < 0)
        for (int i = 0; i < 3; ++i) v1[i] *= -1;
    const double c[3] = {n[1] * v1[2] - n[2] * v1[1], n[2] * v1[0] - n[0] * v1[2], n[0] * v1[1] - n[1] * v1[0]};
    for (int i = 0; i < 3; ++i) v2[i] = -c[i];
    const double l1 = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
    const double l2 = std::sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
    for (int i = 0; i < 3; ++i) {
        basis[0][i] = v1[i] / l1;
        basis[1][i] = v2[i] / l2;
        basis[2][i] = n[i];
    }
}

void make_2d(const double* start, const double basis[3][3], const double* p, double* q) {
    q[0] = q[1] = 0;
    for (int k = 0; k < 3; k++) {
        q[0] += (p[k] - start[k]) * basis[0][k];
        q[1] += (p[k] - start[k]) * basis[1][k];
    }
}

bool in_triangle(const double* F, const double* n, const double* X) {
    double basis[3][3], A[2], B[2], C[2], P[2];
    set_basis(n, basis);
    make_2d(F, basis, F, A);
    make_2d(F, basis, F + 3, B);
    make_2d(F, basis, F + 6, C);
    make_2d(F, basis, X, P);
    const double r1 = (A[0] - P[0]) * (B[1] - A[1]) - (B[0] - A[0]) * (A[1] - P[1]);
    const double r2 = (B[0] - P[0]) * (C[1] - B[1]) - (C[0] - B[0]) * (B[1] - P[1]);
    const double r3 = (C[0] - P[0]) * (A[1] - C[1]) - (A[0] - C[0]) * (C[1] - P[1]);
    if (r1 < 0 && r2 < 0 && r3 < 0) return true;
    if (r1 > 0 && r2 > 0 && r3 > 0) return true;
    return false;
}

// CalculateIllumeOnInnerFace's returned local point.
void inner_face_point(int face, int neib, const double* vm, const double* x0, double* out) {
    out[0] = out[1] = 0;
    if (neib == -1 || neib == -2 || neib == -3 || neib == -4) return;
    double l[3], p[3];
    global_to_local(vm, x0, l);
    switch (face) {
        case 3:
            tetra_to_plane(l, p);
            out[0] = p[0];
            out[1] = p[1];
            return;
        case 1: out[0] = l[1]; out[1] = l[2]; return;
        case 2: out[0] = l[0]; out[1] = l[2]; return;
        default: out[0] = l[0]; out[1] = l[1]; return;
    }
}

}  // namespace

extern "C" {

void orc_make_nodes(int64_t n, const double* vm, double* nodes) {
    for (int64_t c = 0; c < n; c++) cell_nodes(vm + c * 16, nodes + c * 36);
}

// Returns 0, 1 (a node found no in face) or 2 (more records than cap; *count
// holds the records written so far and the count needed).
int orc_make_direction(int64_t n, const double* vm_all, const double* faces, const double* normals,
                       const int32_t* pairs, const int32_t* order, const double* w, int32_t* face_state, double* s,
                       double* x0_local, uint8_t* in_face, int64_t cap, int64_t* count) {
    int64_t r = 0;
    int status = 0;
    double x[36];
    for (int64_t h = 0; h < n; h++) {
        const int64_t c = order[h];
        const double* nc = normals + c * 12;
        int st = 0;
        for (int i = 0; i < 4; ++i) {
            const double* ni = nc + 3 * i;
            if (ni[0] * w[0] + ni[1] * w[1] + ni[2] * w[2] < -kEps) st |= 1 << i;
        }
        face_state[c] = st;
        const double* vm = vm_all + c * 16;
        cell_nodes(vm, x);
        for (int f = 0; f < 4; f++) {
            if ((st >> f) & 1) continue;
            for (int k = 0; k < 3; k++) {
                const double* xp = x + (f * 3 + k) * 3;
                bool found = false;
                for (int g = 0; g < 4; ++g) {
                    if (!((st >> g) & 1)) continue;
                    const double* F = faces + (c * 4 + g) * 9;
                    double x0[3];
                    intersect(F, xp, w, x0);
                    if (in_triangle(F, nc + 3 * g, x0)) {
                        if (r < cap) {
                            const double d[3] = {xp[0] - x0[0], xp[1] - x0[1], xp[2] - x0[2]};
                            s[r] = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                            in_face[r] = (uint8_t)g;
                            inner_face_point(g, pairs[c * 4 + g], vm, x0, x0_local + 2 * r);
                        }
                        r++;
                        found = true;
                        break;
                    }
                }
                if (!found) status = 1;
            }
        }
    }
    *count = r;
    if (status == 0 && r > cap) status = 2;
    return status;
}

}  // extern "C"
