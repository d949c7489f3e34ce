// Structured multigrid hierarchy with ugcore's LEVEL DoF numbering.
//
// This is the stand-in for lib_grid on the hot path: it reproduces the order in
// which GlobalMultiGridRefiner creates vertices
// (lib_grid/refinement/global_multi_grid_refiner.cpp:135-455) and the
// per-subset DoF numbering of DoFDistribution::reinit
// (lib_disc/dof_manager/dof_distribution.cpp:1401-1421), for lattice-topology
// hierarchies (one unit hexahedron refined regularly; the unit square as two
// triangles refined regularly).
//
// Objects (vertices, edges, faces, hexahedra) are keyed by their centre on the
// doubled lattice of their level, so every lookup is an array access; each
// object keeps its own local vertex order because that order decides the
// creation order of its children (grid_objects_1d.cpp:39-54,
// grid_objects_2d.cpp:230-236 and 650-662, hexahedron_rules.cpp:385-440,
// Grid::register_face/register_volume autogeneration,
// grid_connection_managment.cpp:922-1010 and 1440-1560).
//
// Output per level: dof_of_lattice (lattice vertex -> DoF) and the number of
// DoFs in subset 0 ("Inner"); DoFs >= n_inner belong to subset 1 ("Boundary").
#include <algorithm>
#include <array>
#include <climits>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "hierarchy.hpp"
#include "mesh_tables.hpp"

namespace ugamd {

namespace {

using namespace mesh;

struct P3 {
    int32_t x, y, z;
};

// One level of the simulated multigrid. Coordinates are on the level's own
// lattice (0..n); object keys live on the doubled lattice (0..2n).
struct Lvl {
    int dim = 3;
    int n = 0;
    std::vector<P3> vpos;
    std::vector<uint8_t> vsub;
    std::vector<int32_t> e_v;      // 2 per edge
    std::vector<uint8_t> e_sub;
    std::vector<int32_t> f_v;      // 4 per face (-1 padded for triangles)
    std::vector<uint8_t> f_nv, f_sub;
    std::vector<int32_t> f_e;      // 4 per face
    std::vector<int32_t> h_v;      // 8 per hexahedron
    std::vector<uint8_t> h_sub;
    std::vector<int32_t> h_f;      // 6 per hexahedron
    std::vector<int32_t> h_e;      // 12 per hexahedron
    std::vector<int32_t> t_v;      // 4 per tetrahedron
    std::vector<int32_t> t_e;      // 6 per tetrahedron
    std::vector<uint8_t> t_sub;
    std::vector<uint8_t> vpar;     // per vertex: 0 copy, else parent edge direction code
    std::vector<int32_t> key;      // doubled lattice -> edge or quadrilateral id (-1 none)
    std::vector<int32_t> key3;     // tripled lattice (vertex sum) -> triangle id (tetrahedral levels)
    int64_t W = 0;                 // doubled lattice width (2n+1)
    int64_t W3 = 0;                // tripled lattice width (3n+1)
    bool tets = false;

    int64_t kidx(int64_t X, int64_t Y, int64_t Z) const { return (Z * W + Y) * W + X; }
    void init_keys() {
        W = 2 * (int64_t)n + 1;
        key.assign(dim == 3 ? W * W * W : W * W, -1);
        if (tets) {
            W3 = 3 * (int64_t)n + 1;
            key3.assign(W3 * W3 * W3, -1);
        }
    }
    int64_t tri_key(const int32_t* v) const {
        int64_t s[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i) {
            s[0] += vpos[v[i]].x;
            s[1] += vpos[v[i]].y;
            s[2] += vpos[v[i]].z;
        }
        return (s[2] * W3 + s[1]) * W3 + s[0];
    }
    int64_t center_key(const int32_t* v, int nv) const {
        int64_t s[3] = {0, 0, 0};
        for (int i = 0; i < nv; ++i) {
            s[0] += vpos[v[i]].x;
            s[1] += vpos[v[i]].y;
            s[2] += vpos[v[i]].z;
        }
        // centre * 2 on the doubled lattice; for the objects used here the
        // doubled centre is integral (edge midpoints, quad/hex centres; the
        // diagonal edges of the triangle mesh)
        return kidx(2 * s[0] / nv, 2 * s[1] / nv, 2 * s[2] / nv);
    }
    int add_vertex(P3 p, uint8_t si) {
        vpos.push_back(p);
        vsub.push_back(si);
        return (int)vpos.size() - 1;
    }
    int edge_of(int a, int b) const {
        int32_t v[2] = {a, b};
        return key[center_key(v, 2)];
    }
    int add_edge(int a, int b, uint8_t si) {
        int id = (int)e_sub.size();
        e_v.push_back(a);
        e_v.push_back(b);
        e_sub.push_back(si);
        int32_t v[2] = {a, b};
        key[center_key(v, 2)] = id;
        return id;
    }
    int face_of(const int32_t* v, int nv) const { return nv == 3 ? key3[tri_key(v)] : key[center_key(v, nv)]; }
    // face registration with edge autogeneration in local edge order
    int add_face(const int32_t* v, int nv, uint8_t si) {
        int id = (int)f_sub.size();
        for (int i = 0; i < 4; ++i) f_v.push_back(i < nv ? v[i] : -1);
        f_nv.push_back((uint8_t)nv);
        f_sub.push_back(si);
        // only quadrilaterals are ever looked up (by the volumes that own
        // them); a triangle's centroid is not a doubled-lattice point
        if (nv == 4) key[center_key(v, nv)] = id;
        else if (tets) key3[tri_key(v)] = id;
        for (int i = 0; i < 4; ++i) {
            if (i >= nv) { f_e.push_back(-1); continue; }
            int a = v[i], b = v[(i + 1) % nv];
            int e = edge_of(a, b);
            if (e < 0) e = add_edge(a, b, si);
            f_e.push_back(e);
        }
        return id;
    }
    // tetrahedron registration: missing faces in local face order, created
    // with vertices (f0, f2, f1) (Tetrahedron::create_face), then edges
    void add_tet(const int32_t* t, uint8_t si) {
        for (int i = 0; i < 4; ++i) t_v.push_back(t[i]);
        t_sub.push_back(si);
        for (int i = 0; i < 4; ++i) {
            const int* f = kTetFace[i];
            int32_t fd[3] = {t[f[0]], t[f[2]], t[f[1]]};
            if (face_of(fd, 3) < 0) add_face(fd, 3, si);
        }
        for (int i = 0; i < 6; ++i) {
            int e = edge_of(t[kTetEdge[i][0]], t[kTetEdge[i][1]]);
            if (e < 0) e = add_edge(t[kTetEdge[i][0]], t[kTetEdge[i][1]], si);
            t_e.push_back(e);
        }
    }
    // volume registration: faces first (face_desc order f0,f3,f2,f1), then edges
    void add_hex(const int32_t* h, uint8_t si) {
        for (int i = 0; i < 8; ++i) h_v.push_back(h[i]);
        h_sub.push_back(si);
        for (int i = 0; i < 6; ++i) {
            const int* f = kHexFace[i];
            int32_t fd[4] = {h[f[0]], h[f[3]], h[f[2]], h[f[1]]};
            int id = face_of(fd, 4);
            if (id < 0) id = add_face(fd, 4, si);
            h_f.push_back(id);
        }
        for (int i = 0; i < 12; ++i) {
            int e = edge_of(h[kHexEdge[i][0]], h[kHexEdge[i][1]]);
            if (e < 0) e = add_edge(h[kHexEdge[i][0]], h[kHexEdge[i][1]], si);
            h_e.push_back(e);
        }
    }
};

P3 mid(const Lvl& L, const int32_t* v, int nv) {
    // centre of old-level vertices expressed on the new level's lattice
    int64_t s[3] = {0, 0, 0};
    for (int i = 0; i < nv; ++i) {
        s[0] += L.vpos[v[i]].x;
        s[1] += L.vpos[v[i]].y;
        s[2] += L.vpos[v[i]].z;
    }
    return P3{(int32_t)(2 * s[0] / nv), (int32_t)(2 * s[1] / nv), (int32_t)(2 * s[2] / nv)};
}

// Refine level C into a new level. With `vertices_only` only the vertex list is
// produced (enough for the DoF numbering of the top level).
Lvl refine(const Lvl& C, bool vertices_only) {
    Lvl F;
    F.dim = C.dim;
    F.n = 2 * C.n;
    F.tets = C.tets;
    if (!vertices_only) F.init_keys();
    const int nv = (int)C.vpos.size();
    const int ne = (int)C.e_sub.size();
    const int nf = (int)C.f_sub.size();
    const int nh = (int)C.h_sub.size();
    F.vpos.reserve((size_t)nv + ne + nf + nh);
    std::vector<int32_t> vc(nv), em(ne), fc(nf, -1);
    for (int v = 0; v < nv; ++v) {
        P3 p = C.vpos[v];
        vc[v] = F.add_vertex(P3{2 * p.x, 2 * p.y, 2 * p.z}, C.vsub[v]);
        if (F.tets) F.vpar.push_back(0);
    }
    for (int e = 0; e < ne; ++e) {
        em[e] = F.add_vertex(mid(C, &C.e_v[2 * e], 2), C.e_sub[e]);
        if (F.tets) {
            const P3 a = C.vpos[C.e_v[2 * e]], b = C.vpos[C.e_v[2 * e + 1]];
            F.vpar.push_back((uint8_t)edge_direction_code(b.x - a.x, b.y - a.y, b.z - a.z));
        }
        if (!vertices_only) {
            F.add_edge(vc[C.e_v[2 * e]], em[e], C.e_sub[e]);
            F.add_edge(em[e], vc[C.e_v[2 * e + 1]], C.e_sub[e]);
        }
    }
    for (int f = 0; f < nf; ++f) {
        const int k = C.f_nv[f];
        const uint8_t si = C.f_sub[f];
        int32_t vr[4], ev[4];
        for (int j = 0; j < k; ++j) {
            vr[j] = vc[C.f_v[4 * f + j]];
            ev[j] = em[C.f_e[4 * f + j]];
        }
        if (k == 4) {
            const int32_t c = F.add_vertex(mid(C, &C.f_v[4 * f], 4), si);
            fc[f] = c;
            if (vertices_only) continue;
            for (int j = 0; j < 4; ++j) {
                int32_t q[4] = {vr[j], ev[j], c, ev[(j + 3) % 4]};
                F.add_face(q, 4, si);
            }
        } else {
            if (vertices_only) continue;
            int32_t t[4][3] = {{vr[0], ev[0], ev[2]}, {vr[1], ev[1], ev[0]},
                               {vr[2], ev[2], ev[1]}, {ev[0], ev[1], ev[2]}};
            for (auto& tt : t) F.add_face(tt, 3, si);
        }
    }
    const int nt = (int)C.t_sub.size();
    for (int t = 0; t < nt; ++t) {
        int32_t all[10];
        for (int j = 0; j < 4; ++j) all[j] = vc[C.t_v[4 * t + j]];
        for (int j = 0; j < 6; ++j) all[4 + j] = em[C.t_e[6 * t + j]];
        // shortest inner diagonal between midpoints of opposite edges, from the
        // corner positions: |(p_a + p_b) - (p_c + p_d)|^2 on the level lattice
        auto d2 = [&](int a, int b, int c, int d) {
            const P3 A = C.vpos[C.t_v[4 * t + a]], B = C.vpos[C.t_v[4 * t + b]];
            const P3 Cc = C.vpos[C.t_v[4 * t + c]], D = C.vpos[C.t_v[4 * t + d]];
            const int64_t x = (int64_t)A.x + B.x - Cc.x - D.x, y = (int64_t)A.y + B.y - Cc.y - D.y,
                          z = (int64_t)A.z + B.z - Cc.z - D.z;
            return x * x + y * y + z * z;
        };
        int best = 2;
        int64_t dbest = d2(0, 2, 1, 3);
        const int64_t d13 = d2(1, 2, 0, 3), d05 = d2(0, 1, 2, 3);
        if (d13 < dbest) {
            best = 1;
            dbest = d13;
        }
        if (d05 < dbest) best = 0;
        for (int c = 0; c < 8; ++c) {
            const int* r = c < 4 ? kTetCorner[c] : kTetInner[best][c - 4];
            int32_t tv[4] = {all[r[0]], all[r[1]], all[r[2]], all[r[3]]};
            if (vertices_only) {
                for (int j = 0; j < 4; ++j) F.t_v.push_back(tv[j]);
                F.t_sub.push_back(C.t_sub[t]);
            } else {
                F.add_tet(tv, C.t_sub[t]);
            }
        }
    }
    for (int h = 0; h < nh; ++h) {
        int32_t all[27];
        for (int j = 0; j < 8; ++j) all[j] = vc[C.h_v[8 * h + j]];
        for (int j = 0; j < 12; ++j) all[8 + j] = em[C.h_e[12 * h + j]];
        for (int j = 0; j < 6; ++j) all[20 + j] = fc[C.h_f[6 * h + j]];
        all[26] = F.add_vertex(mid(C, &C.h_v[8 * h], 8), C.h_sub[h]);
        if (vertices_only) continue;
        for (int c = 0; c < 8; ++c) {
            int32_t hv[8];
            for (int j = 0; j < 8; ++j) hv[j] = all[kHexChild[c][j]];
            F.add_hex(hv, C.h_sub[h]);
        }
    }
    return F;
}

LevelNumbering number(const Lvl& L) {
    LevelNumbering out;
    out.dim = L.dim;
    out.n = L.n;
    const int64_t W = (int64_t)L.n + 1;
    out.dof_of_lattice.assign(L.dim == 3 ? W * W * W : W * W, -1);
    int32_t k = 0;
    for (int si = 0; si < 2; ++si) {
        for (size_t v = 0; v < L.vpos.size(); ++v) {
            if (L.vsub[v] != si) continue;
            const P3 p = L.vpos[v];
            out.dof_of_lattice[((int64_t)p.z * W + p.y) * W + p.x] = k++;
        }
        if (si == 0) out.n_inner = k;
    }
    out.n_dof = k;
    if (L.tets) {
        const int64_t n = L.n;
        out.cell_tets.assign(n * n * n * 6, 0);
        std::vector<uint8_t> fill(n * n * n, 0);
        const int64_t ntet = (int64_t)L.t_v.size() / 4;
        for (int64_t t = 0; t < ntet; ++t) {
            P3 q[4];
            int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX};
            for (int j = 0; j < 4; ++j) {
                q[j] = L.vpos[L.t_v[4 * t + j]];
                mn[0] = std::min(mn[0], q[j].x);
                mn[1] = std::min(mn[1], q[j].y);
                mn[2] = std::min(mn[2], q[j].z);
            }
            uint16_t code = 0;
            for (int j = 0; j < 4; ++j) {
                const int cx = q[j].x - mn[0], cy = q[j].y - mn[1], cz = q[j].z - mn[2];
                if (cx > 1 || cy > 1 || cz > 1) throw std::runtime_error("hierarchy: tetrahedron spans two cells");
                code |= (uint16_t)((cx | (cy << 1) | (cz << 2)) << (3 * j));
            }
            const int64_t cell = ((int64_t)mn[2] * n + mn[1]) * n + mn[0];
            if (fill[cell] >= 6) throw std::runtime_error("hierarchy: more than 6 tetrahedra in a cell");
            out.cell_tets[cell * 6 + fill[cell]++] = code;
        }
        if (!L.vpar.empty()) {
            out.parent_edge.assign(W * W * W, 0);
            for (size_t v = 0; v < L.vpos.size(); ++v) {
                const P3 p = L.vpos[v];
                out.parent_edge[((int64_t)p.z * W + p.y) * W + p.x] = L.vpar[v];
            }
        }
    }
    return out;
}

Lvl base_level(int kind) {
    Lvl L;
    L.n = 1;
    if (kind == 0) {
        L.dim = 3;
        L.init_keys();
        const int c[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                             {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        for (auto& p : c) L.add_vertex(P3{p[0], p[1], p[2]}, 1);
        int32_t h[8] = {0, 1, 2, 3, 4, 5, 6, 7};
        L.add_hex(h, 0);
        std::fill(L.f_sub.begin(), L.f_sub.end(), 1);
        std::fill(L.e_sub.begin(), L.e_sub.end(), 1);
    } else if (kind == 1) {
        L.dim = 2;
        L.init_keys();
        const int c[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (auto& p : c) L.add_vertex(P3{p[0], p[1], 0}, 1);
        int32_t t0[3] = {0, 1, 2}, t1[3] = {0, 2, 3};
        L.add_face(t0, 3, 0);
        L.add_face(t1, 3, 0);
        // the diagonal (0,2) is the only inner edge of the unit square
        for (size_t e = 0; e < L.e_sub.size(); ++e) {
            const int a = L.e_v[2 * e], b = L.e_v[2 * e + 1];
            L.e_sub[e] = (std::min(a, b) == 0 && std::max(a, b) == 2) ? 0 : 1;
        }
    } else if (kind == 2) {
        // 3x3x3 cells, vertices and cells in lexicographic order, per cell
        // the tetrahedra p0, p0+e_a, p0+e_a+e_b, p0+(1,1,1) for the axis
        // permutations (a,b,c) in lexicographic order; odd permutations
        // stored as (P0,P3,P2,P1) (positive orientation)
        L.dim = 3;
        L.n = 3;
        L.tets = true;
        L.init_keys();
        const int M = 3;
        auto vid = [&](int i, int j, int k) { return (k * (M + 1) + j) * (M + 1) + i; };
        for (int k = 0; k <= M; ++k)
            for (int j = 0; j <= M; ++j)
                for (int i = 0; i <= M; ++i) {
                    const bool bnd = i == 0 || j == 0 || k == 0 || i == M || j == M || k == M;
                    L.add_vertex(P3{i, j, k}, bnd ? 1 : 0);
                }
        const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        const bool odd[6] = {false, true, true, false, false, true};
        for (int k = 0; k < M; ++k)
            for (int j = 0; j < M; ++j)
                for (int i = 0; i < M; ++i)
                    for (int p = 0; p < 6; ++p) {
                        int c[3] = {i, j, k};
                        int32_t P[4];
                        P[0] = vid(c[0], c[1], c[2]);
                        for (int q = 0; q < 3; ++q) {
                            c[perms[p][q]] += 1;
                            P[q + 1] = vid(c[0], c[1], c[2]);
                        }
                        int32_t t[4] = {P[0], odd[p] ? P[3] : P[1], P[2], odd[p] ? P[1] : P[3]};
                        L.add_tet(t, 0);
                    }
        // faces and edges on a boundary plane belong to subset 1
        auto plane = [&](const int32_t* v, int nv) {
            for (int ax = 0; ax < 3; ++ax)
                for (int side = 0; side < 2; ++side) {
                    bool all = true;
                    for (int q = 0; q < nv; ++q) {
                        const P3 pp = L.vpos[v[q]];
                        const int c = ax == 0 ? pp.x : (ax == 1 ? pp.y : pp.z);
                        all = all && c == (side ? M : 0);
                    }
                    if (all) return true;
                }
            return false;
        };
        for (size_t f = 0; f < L.f_sub.size(); ++f) L.f_sub[f] = plane(&L.f_v[4 * f], 3) ? 1 : 0;
        for (size_t e = 0; e < L.e_sub.size(); ++e) L.e_sub[e] = plane(&L.e_v[2 * e], 2) ? 1 : 0;
    } else {
        throw std::runtime_error("hierarchy: unknown base grid kind");
    }
    return L;
}

}  // namespace

BaseMesh base_mesh(int kind) {
    const Lvl L = base_level(kind);
    BaseMesh B;
    B.dim = L.dim;
    B.n = L.n;
    B.tets = L.tets;
    B.face_nv = L.f_nv.empty() ? 4 : L.f_nv[0];
    for (uint8_t k : L.f_nv)
        if (k != B.face_nv) throw std::runtime_error("hierarchy: mixed face types");
    for (const P3& p : L.vpos) {
        B.vpos.push_back(p.x);
        B.vpos.push_back(p.y);
        B.vpos.push_back(p.z);
    }
    B.vsub = L.vsub;
    B.e_v = L.e_v;
    B.e_sub = L.e_sub;
    B.f_v = L.f_v;
    B.f_e = L.f_e;
    B.f_sub = L.f_sub;
    B.h_v = L.h_v;
    B.h_f = L.h_f;
    B.h_e = L.h_e;
    B.h_sub = L.h_sub;
    B.t_v = L.t_v;
    B.t_e = L.t_e;
    B.t_sub = L.t_sub;
    return B;
}

std::vector<LevelNumbering> build_level_numbering(int kind, int num_refs) {
    std::vector<LevelNumbering> out;
    Lvl cur = base_level(kind);
    out.push_back(number(cur));
    for (int r = 1; r <= num_refs; ++r) {
        Lvl nxt = refine(cur, r == num_refs);
        out.push_back(number(nxt));
        cur = std::move(nxt);
    }
    return out;
}

}  // namespace ugamd
