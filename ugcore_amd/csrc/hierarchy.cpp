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
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "hierarchy.hpp"

namespace ugamd {

namespace {

// local index tables of the reference hexahedron (hexahedron_rules.h)
constexpr int kHexEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                                 {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
constexpr int kHexFace[6][4] = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2},
                                {2, 6, 7, 3}, {3, 7, 4, 0}, {4, 7, 6, 5}};
// regular 1:8 rule in the index space [corners 0-7 | edge vertices 8-19 |
// face vertices 20-25 | centre 26]
constexpr int kHexChild[8][8] = {
    {0, 8, 20, 11, 12, 21, 26, 24}, {8, 1, 9, 20, 21, 13, 22, 26},
    {20, 9, 2, 10, 26, 22, 14, 23}, {11, 20, 10, 3, 24, 26, 23, 15},
    {12, 21, 26, 24, 4, 16, 25, 19}, {21, 13, 22, 26, 16, 5, 17, 25},
    {26, 22, 14, 23, 25, 17, 6, 18}, {24, 26, 23, 15, 19, 25, 18, 7}};

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
    std::vector<int32_t> key;      // doubled lattice -> edge or face id (-1 none)
    int64_t W = 0;                 // doubled lattice width (2n+1)

    int64_t kidx(int64_t X, int64_t Y, int64_t Z) const { return (Z * W + Y) * W + X; }
    void init_keys() {
        W = 2 * (int64_t)n + 1;
        key.assign(dim == 3 ? W * W * W : W * W, -1);
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
    int face_of(const int32_t* v, int nv) const { return key[center_key(v, nv)]; }
    // face registration with edge autogeneration in local edge order
    int add_face(const int32_t* v, int nv, uint8_t si) {
        int id = (int)f_sub.size();
        for (int i = 0; i < 4; ++i) f_v.push_back(i < nv ? v[i] : -1);
        f_nv.push_back((uint8_t)nv);
        f_sub.push_back(si);
        // only quadrilaterals are ever looked up (by the volumes that own
        // them); a triangle's centroid is not a doubled-lattice point
        if (nv == 4) key[center_key(v, nv)] = id;
        for (int i = 0; i < 4; ++i) {
            if (i >= nv) { f_e.push_back(-1); continue; }
            int a = v[i], b = v[(i + 1) % nv];
            int e = edge_of(a, b);
            if (e < 0) e = add_edge(a, b, si);
            f_e.push_back(e);
        }
        return id;
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
    }
    for (int e = 0; e < ne; ++e) {
        em[e] = F.add_vertex(mid(C, &C.e_v[2 * e], 2), C.e_sub[e]);
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
    } else {
        throw std::runtime_error("hierarchy: unknown base grid kind");
    }
    return L;
}

}  // namespace

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
