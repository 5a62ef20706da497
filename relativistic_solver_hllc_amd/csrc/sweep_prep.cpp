// Host-side preparation of the transport sweep: a structured tetrahedral mesh,
// cell geometry, and the per-direction sweep structures the reference's
// make_illum_struct step writes to disk (make_illum_struct/
// struct_short_characteristics_main.cpp:128-165): face states, a topological
// cell order ("graph"), and for every node of every out face the distance s
// back to the cell's in face and that face's id (vec_x0 records).
//
// Conventions (this build's, stated because the reference takes them from its
// VTK mesh): face f of a tetrahedron is the face opposite vertex f; its three
// vertices are the other three in increasing order; node j of a face lies at
// barycentric weight 2/3 on face vertex j and 1/6 on the other two.
//
// The reference's own preparation, in its VTK face numbering and node order
// (rsh_mesh_vtk_arrays, rsh_make_nodes, rsh_make_direction), is here as well:
// from its input arrays (vertex.bin, grid.bin, normals.bin, pairs.bin and the
// graph order) it gives the face states, X.bin's nodes and the LocX0 records
// (s, in face, local point on the in face) that its MAKE step writes. Both
// placements put a face's nodes at the same three points, in another order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <string>
#include <unordered_map>
#include <vector>

#include "rsh_scatter.h"

namespace rsh {
int fail(int status, const std::string& msg);
}

namespace {

constexpr int kFaceVerts[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
// eps of GetInAndOutFaces (utils/grid_geometry/geometry_solve.cpp:212-227, global_value.h:13)
constexpr double kInEps = 1e-10;

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 vert(const double* v, int64_t c, int i) {
    const double* p = v + (c * 4 + i) * 3;
    return {p[0], p[1], p[2]};
}

// Outward unit normal, area of face f of cell c.
void face_normal(const double* v, int64_t c, int f, V3* n, double* area) {
    const V3 a = vert(v, c, kFaceVerts[f][0]), b = vert(v, c, kFaceVerts[f][1]), d = vert(v, c, kFaceVerts[f][2]);
    V3 nn = cross(sub(b, a), sub(d, a));
    const double len = std::sqrt(dot(nn, nn));
    if (dot(nn, sub(vert(v, c, f), a)) > 0) nn = {-nn.x, -nn.y, -nn.z};
    *n = {nn.x / len, nn.y / len, nn.z / len};
    *area = 0.5 * len;
}

V3 face_node(const double* v, int64_t c, int f, int j) {
    V3 p[3];
    for (int i = 0; i < 3; i++) p[i] = vert(v, c, kFaceVerts[f][i]);
    const V3 a = p[j], b = p[(j + 1) % 3], d = p[(j + 2) % 3];
    return {(4 * a.x + b.x + d.x) / 6, (4 * a.y + b.y + d.y) / 6, (4 * a.z + b.z + d.z) / 6};
}

// The reference's own numbering, taken from its VTK meshes (vtkTetra): face f
// holds points kVtkFace[f] in this order (grid.bin's A, B, C) and is the face
// opposite point kVtkOpp[f]; a face g of this library's numbering is VTK face
// kMineToVtk[g].
constexpr int kVtkFace[4][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {0, 2, 1}};
constexpr int kVtkOpp[4] = {2, 0, 1, 3};
constexpr int kMineToVtk[4] = {1, 2, 0, 3};
// The reference's node placement (MakeArrayX, make_illum_struct/
// struct_short_characteristics_logic_function.cpp:10-72, with base_tetra_geo's
// straight_face and inclined_face nodes, struct_short_characteristics_
// calculations.cpp:146-183): node j of VTK face f lies at weight 2/3 on point
// kNodeMajor[f][j] and 1/6 on the face's other two points. Faces 0-2 have the
// nodes (1/6, 1/6), (2/3, 1/6), (1/6, 2/3) in the face's two local axes (the
// first node nearest point 3); face 3's inclined-plane nodes map to points 2,
// 1, 0 in turn.
constexpr int kNodeMajor[4][3] = {{3, 0, 1}, {3, 1, 2}, {3, 0, 2}, {2, 1, 0}};

// Point j of cell c from a vertex matrix [N][4][4] (vertex.bin: Eigen Matrix4d,
// column-major, column j = (x, y, z, 1) of point j).
inline V3 vm_point(const double* vm, int64_t c, int j) {
    const double* p = vm + c * 16 + j * 4;
    return {p[0], p[1], p[2]};
}

V3 ref_face_node(const double* vm, int64_t c, int f, int j) {
    const int major = kNodeMajor[f][j];
    V3 o[2];
    int k = 0;
    for (int i = 0; i < 3; i++)
        if (kVtkFace[f][i] != major) o[k++] = vm_point(vm, c, kVtkFace[f][i]);
    const V3 a = vm_point(vm, c, major);
    return {(4 * a.x + o[0].x + o[1].x) / 6, (4 * a.y + o[0].y + o[1].y) / 6, (4 * a.z + o[0].z + o[1].z) / 6};
}

inline double det3(V3 a, V3 b, V3 d) { return dot(a, cross(b, d)); }

}  // namespace

extern "C" {

int rsh_mesh_cube_tetra(int32_t n_edge, double* vertices, int32_t* pairs) {
    if (n_edge < 1 || n_edge > 100 || !vertices || !pairs) return rsh::fail(RSH_E_ARG, "n_edge must be 1..100");
    const int64_t n = n_edge, cells = 6 * n * n * n;
    const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    std::vector<int64_t> lattice((size_t)cells * 4);  // lattice point index of each vertex
    const double h = 1.0 / (double)n;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++)
            for (int64_t k = 0; k < n; k++)
                for (int p = 0; p < 6; p++) {
                    const int64_t c = ((i * n + j) * n + k) * 6 + p;
                    int64_t q[3] = {i, j, k};
                    for (int vtx = 0; vtx < 4; vtx++) {
                        if (vtx > 0) q[perms[p][vtx - 1]] += 1;
                        double* out = vertices + (c * 4 + vtx) * 3;
                        out[0] = q[0] * h;
                        out[1] = q[1] * h;
                        out[2] = q[2] * h;
                        lattice[(size_t)(c * 4 + vtx)] = (q[0] * (n + 1) + q[1]) * (n + 1) + q[2];
                    }
                }
    std::unordered_map<uint64_t, int64_t> open;
    open.reserve((size_t)cells * 2);
    for (int64_t c = 0; c < cells; c++)
        for (int f = 0; f < 4; f++) {
            int64_t key3[3];
            for (int i = 0; i < 3; i++) key3[i] = lattice[(size_t)(c * 4 + kFaceVerts[f][i])];
            std::sort(key3, key3 + 3);
            const uint64_t key = ((uint64_t)key3[0] << 42) | ((uint64_t)key3[1] << 21) | (uint64_t)key3[2];
            auto it = open.find(key);
            if (it == open.end()) {
                open.emplace(key, c * 4 + f);
            } else {
                pairs[c * 4 + f] = (int32_t)it->second;
                pairs[it->second] = (int32_t)(c * 4 + f);
                open.erase(it);
            }
        }
    for (const auto& kv : open) {  // boundary faces: the x = 0 plane emits (eBound_OutSource), the rest is free
        const int64_t id = kv.second, c = id / 4;
        const int f = (int)(id % 4);
        bool x0 = true;
        for (int i = 0; i < 3; i++) x0 = x0 && vertices[(c * 4 + kFaceVerts[f][i]) * 3] == 0.0;
        pairs[id] = x0 ? RSH_BOUND_OUT_SOURCE : RSH_BOUND_FREE;
    }
    return RSH_OK;
}

int rsh_tetra_geometry(int64_t n, const double* vertices, double* normals, double* areas, double* volume,
                       double* nodes) {
    if (n < 0 || !vertices) return rsh::fail(RSH_E_ARG, "invalid arguments");
    for (int64_t c = 0; c < n; c++) {
        for (int f = 0; f < 4; f++) {
            V3 nn;
            double ar;
            face_normal(vertices, c, f, &nn, &ar);
            if (normals) {
                normals[(c * 4 + f) * 3] = nn.x;
                normals[(c * 4 + f) * 3 + 1] = nn.y;
                normals[(c * 4 + f) * 3 + 2] = nn.z;
            }
            if (areas) areas[c * 4 + f] = ar;
            if (nodes)
                for (int j = 0; j < 3; j++) {
                    const V3 x = face_node(vertices, c, f, j);
                    double* o = nodes + ((c * 4 + f) * 3 + j) * 3;
                    o[0] = x.x;
                    o[1] = x.y;
                    o[2] = x.z;
                }
        }
        if (volume) {
            const V3 a = vert(vertices, c, 0);
            volume[c] = std::fabs(dot(sub(vert(vertices, c, 1), a),
                                      cross(sub(vert(vertices, c, 2), a), sub(vert(vertices, c, 3), a)))) / 6;
        }
    }
    return RSH_OK;
}

int rsh_sweep_prepare_direction(int64_t n, const double* vertices, const int32_t* pairs, const double* omega,
                                int32_t* order, int32_t* face_state, double* x0_s, uint8_t* x0_face, int64_t cap,
                                int64_t* n_records) {
    if (n <= 0 || !vertices || !pairs || !omega || !order || !face_state || !n_records || cap < 0 ||
        (cap > 0 && (!x0_s || !x0_face)))
        return rsh::fail(RSH_E_ARG, "invalid arguments");
    const V3 w = {omega[0], omega[1], omega[2]};
    std::vector<V3> nrm((size_t)n * 4);
    for (int64_t c = 0; c < n; c++)
        for (int f = 0; f < 4; f++) {
            double ar;
            face_normal(vertices, c, f, &nrm[(size_t)(c * 4 + f)], &ar);
        }
    // Face states (GetInAndOutFaces): bit f set when n_f . omega < -eps. An
    // interior face is classified once, from its lower-numbered cell's normal,
    // so that the two sides always agree.
    for (int64_t c = 0; c < n; c++) {
        int st = 0;
        for (int f = 0; f < 4; f++) {
            const int32_t p = pairs[c * 4 + f];
            double d = dot(nrm[(size_t)(c * 4 + f)], w);
            if (p >= 0 && p / 4 < c) d = -dot(nrm[(size_t)p], w);
            if (d < -kInEps) st |= 1 << f;
        }
        face_state[c] = st;
    }
    // Topological order (Kahn, level by level): an in face depends on the
    // neighbour across it; a face that is "out" on both sides (|n . omega| <=
    // eps) is written by both cells, and the higher-numbered one goes later.
    std::vector<int32_t> indeg((size_t)n, 0);
    auto depends = [&](int64_t c, int f, int64_t* u) -> bool {
        const int32_t p = pairs[c * 4 + f];
        if (p < 0) return false;
        *u = p / 4;
        if ((face_state[c] >> f) & 1) return true;
        return !((face_state[*u] >> (p % 4)) & 1) && *u < c;
    };
    for (int64_t c = 0; c < n; c++)
        for (int f = 0; f < 4; f++) {
            int64_t u;
            if (depends(c, f, &u)) indeg[(size_t)c]++;
        }
    std::deque<int64_t> ready;
    for (int64_t c = 0; c < n; c++)
        if (indeg[(size_t)c] == 0) ready.push_back(c);
    int64_t placed = 0;
    while (!ready.empty()) {
        const int64_t c = ready.front();
        ready.pop_front();
        order[placed++] = (int32_t)c;
        for (int f = 0; f < 4; f++) {
            const int32_t p = pairs[c * 4 + f];
            if (p < 0) continue;
            const int64_t v = p / 4;
            int64_t u;
            if (depends(v, p % 4, &u) && u == c && --indeg[(size_t)v] == 0) ready.push_back(v);
        }
    }
    if (placed != n) return rsh::fail(RSH_E_ARG, "cyclic upwind dependencies for this direction");
    // Records: every node of every out face, in traversal order; the node is
    // traced back along -omega to the nearest in-face plane of its cell.
    int64_t r = 0;
    for (int64_t h = 0; h < n; h++) {
        const int64_t c = order[h];
        for (int f = 0; f < 4; f++) {
            if ((face_state[c] >> f) & 1) continue;
            for (int j = 0; j < 3; j++) {
                const V3 x = face_node(vertices, c, f, j);
                double best = INFINITY;
                int best_face = -1;
                for (int g = 0; g < 4; g++) {
                    if (!((face_state[c] >> g) & 1)) continue;
                    const V3 ng = nrm[(size_t)(c * 4 + g)];
                    const double den = dot(ng, w);
                    if (den >= 0) continue;
                    double t = dot(ng, sub(x, vert(vertices, c, kFaceVerts[g][0]))) / den;
                    if (t < 0) t = 0;
                    if (t < best) {
                        best = t;
                        best_face = g;
                    }
                }
                if (best_face < 0) return rsh::fail(RSH_E_ARG, "an out-face node has no in face behind it");
                if (r < cap) {
                    const V3 x0 = {x.x - best * w.x, x.y - best * w.y, x.z - best * w.z};
                    const V3 dx = sub(x, x0);
                    x0_s[r] = std::sqrt(dot(dx, dx));
                    x0_face[r] = (uint8_t)best_face;
                }
                r++;
            }
        }
    }
    *n_records = r;
    if (r > cap) return rsh::fail(RSH_E_NOMEM, "record buffer too small (n_records holds the size needed)");
    return RSH_OK;
}

int rsh_mesh_vtk_arrays(int64_t n, const double* vertices, const int32_t* pairs, double* vertex_matrix, double* faces,
                        double* normals, int32_t* pairs_vtk) {
    if (n <= 0 || !vertices || (pairs_vtk && !pairs)) return rsh::fail(RSH_E_ARG, "invalid arguments");
    for (int64_t c = 0; c < n; c++) {
        if (vertex_matrix)
            for (int j = 0; j < 4; j++) {
                const V3 p = vert(vertices, c, j);
                double* o = vertex_matrix + c * 16 + j * 4;
                o[0] = p.x;
                o[1] = p.y;
                o[2] = p.z;
                o[3] = 1.0;
            }
        for (int f = 0; f < 4; f++) {
            if (faces)
                for (int i = 0; i < 3; i++) {
                    const V3 p = vert(vertices, c, kVtkFace[f][i]);
                    double* o = faces + ((c * 4 + f) * 3 + i) * 3;
                    o[0] = p.x;
                    o[1] = p.y;
                    o[2] = p.z;
                }
            if (normals) {
                V3 nn;
                double ar;
                face_normal(vertices, c, kVtkOpp[f], &nn, &ar);
                double* o = normals + (c * 4 + f) * 3;
                o[0] = nn.x;
                o[1] = nn.y;
                o[2] = nn.z;
            }
            if (pairs_vtk) {
                const int32_t p = pairs[c * 4 + kVtkOpp[f]];
                if (p >= n * 4) return rsh::fail(RSH_E_ARG, "pairs names a face beyond the mesh");
                pairs_vtk[c * 4 + f] = p >= 0 ? (p / 4) * 4 + kMineToVtk[p % 4] : p;
            }
        }
    }
    return RSH_OK;
}

int rsh_make_nodes(int64_t n, const double* vertex_matrix, double* nodes) {
    if (n <= 0 || !vertex_matrix || !nodes) return rsh::fail(RSH_E_ARG, "invalid arguments");
    for (int64_t c = 0; c < n; c++)
        for (int f = 0; f < 4; f++)
            for (int j = 0; j < 3; j++) {
                const V3 x = ref_face_node(vertex_matrix, c, f, j);
                double* o = nodes + ((c * 4 + f) * 3 + j) * 3;
                o[0] = x.x;
                o[1] = x.y;
                o[2] = x.z;
            }
    return RSH_OK;
}

int rsh_make_direction(int64_t n, const double* vertex_matrix, const double* faces, const double* normals,
                       const int32_t* pairs, const int32_t* order, const double* omega, int32_t* face_state,
                       double* x0_s, double* x0_local, uint8_t* x0_face, int64_t cap, int64_t* n_records) {
    if (n <= 0 || !vertex_matrix || !faces || !normals || !pairs || !order || !omega || !face_state || !n_records ||
        cap < 0 || (cap > 0 && (!x0_s || !x0_local || !x0_face)))
        return rsh::fail(RSH_E_ARG, "invalid arguments");
    std::vector<uint8_t> seen((size_t)n, 0);
    for (int64_t h = 0; h < n; h++) {
        const int32_t c = order[h];
        if (c < 0 || c >= n || seen[(size_t)c]) return rsh::fail(RSH_E_ARG, "order is not a permutation of the cells");
        seen[(size_t)c] = 1;
    }
    const V3 w = {omega[0], omega[1], omega[2]};
    int64_t r = 0;
    for (int64_t h = 0; h < n; h++) {
        const int64_t c = order[h];
        // GetInAndOutFaces: bit f = in face, from the cell's own normal.
        int st = 0;
        for (int f = 0; f < 4; f++) {
            const double* nf = normals + (c * 4 + f) * 3;
            if (nf[0] * w.x + nf[1] * w.y + nf[2] * w.z < -kInEps) st |= 1 << f;
        }
        face_state[c] = st;
        // Point 3 is the origin of the cell's local coordinates (l0, l1, l2)
        // along points 0, 1, 2 (FromGlobalToLocalTetra), by Cramer's rule.
        const V3 p3 = vm_point(vertex_matrix, c, 3);
        const V3 e0 = sub(vm_point(vertex_matrix, c, 0), p3), e1 = sub(vm_point(vertex_matrix, c, 1), p3),
                 e2 = sub(vm_point(vertex_matrix, c, 2), p3);
        const double vol = det3(e0, e1, e2);
        for (int f = 0; f < 4; f++) {
            if ((st >> f) & 1) continue;
            for (int j = 0; j < 3; j++) {
                const V3 x = ref_face_node(vertex_matrix, c, f, j);
                // The first in face (in face order) whose plane the line
                // through x along omega meets strictly inside its triangle.
                int g = 0;
                V3 x0 = {0, 0, 0};
                for (; g < 4; g++) {
                    if (!((st >> g) & 1)) continue;
                    const double* t = faces + (c * 4 + g) * 9;
                    const V3 A = {t[0], t[1], t[2]}, B = {t[3], t[4], t[5]}, D = {t[6], t[7], t[8]};
                    const V3 nn = cross(sub(B, A), sub(D, A));
                    const double tt = dot(nn, sub(A, x)) / dot(nn, w);
                    x0 = {x.x + tt * w.x, x.y + tt * w.y, x.z + tt * w.z};
                    const double la = dot(nn, cross(sub(B, x0), sub(D, x0)));
                    const double lb = dot(nn, cross(sub(D, x0), sub(A, x0)));
                    const double lc = dot(nn, cross(sub(A, x0), sub(B, x0)));
                    if (la > 0 && lb > 0 && lc > 0) break;
                }
                if (g == 4)
                    return rsh::fail(RSH_E_ARG, "an out-face node meets no in face strictly inside its triangle "
                                                "(the reference stops with 'Not found inner face')");
                if (r < cap) {
                    const V3 dx = sub(x, x0);
                    x0_s[r] = std::sqrt(dot(dx, dx));
                    x0_face[r] = (uint8_t)g;
                    // CalculateIllumeOnInnerFace's local point on the in face:
                    // (0, 0) behind a boundary face, else two of the local
                    // coordinates (face 3: the inclined plane's coordinates
                    // ((l1 - l0) / sqrt 2, (1 - l0 - l1 + 2 l2) / sqrt 6)).
                    double* xl = x0_local + 2 * r;
                    const int32_t nb = pairs[c * 4 + g];
                    if (nb <= RSH_BOUND_FREE && nb >= RSH_BOUND_INNER_SOURCE) {
                        xl[0] = xl[1] = 0.0;
                    } else {
                        const V3 q = sub(x0, p3);
                        const double l0 = det3(q, e1, e2) / vol, l1 = det3(e0, q, e2) / vol,
                                     l2 = det3(e0, e1, q) / vol;
                        switch (g) {
                            case 0: xl[0] = l0; xl[1] = l1; break;
                            case 1: xl[0] = l1; xl[1] = l2; break;
                            case 2: xl[0] = l0; xl[1] = l2; break;
                            default:
                                xl[0] = (l1 - l0) / std::sqrt(2.0);
                                xl[1] = (1.0 - l0 - l1 + 2.0 * l2) / std::sqrt(6.0);
                                break;
                        }
                    }
                }
                r++;
            }
        }
    }
    *n_records = r;
    if (r > cap) return rsh::fail(RSH_E_NOMEM, "record buffer too small (n_records holds the size needed)");
    return RSH_OK;
}

}  // extern "C"
