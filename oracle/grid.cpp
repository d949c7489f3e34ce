// Oracle (test infrastructure only): restatement of lib_grid's global multigrid
// refinement order and of the LEVEL DoF numbering.
//
// Rules followed:
//  * GlobalMultiGridRefiner::perform_refinement
//    (lib_grid/refinement/global_multi_grid_refiner.cpp:135-455): on the new
//    level, vertices are created as clones of the old level's vertices (storage
//    order), then one midpoint per old edge (edge order; each edge is split into
//    (v0', mid), (mid, v1'), grid_objects_1d.cpp:39-54), then one centre per old
//    quadrilateral (face order), then one centre per old hexahedron (volume
//    order). New faces/volumes are registered in creation order.
//  * Grid::register_face / register_volume
//    (lib_grid/grid/grid_connection_managment.cpp:922-1010, 1440-1560): a
//    registered volume first autogenerates its missing faces in local face order
//    (Hexahedron::create_face uses vertices f0,f3,f2,f1 of FACE_VRT_INDS,
//    grid_objects_3d.cpp:566-601); every registered face autogenerates its
//    missing edges (w_i, w_{i+1}) in local edge order.
//  * Child rules: quadrilateral (grid_objects_2d.cpp:650-662), triangle
//    (grid_objects_2d.cpp:230-236), hexahedron regular rule
//    (hexahedron_rules.cpp:385-440) with the local index space
//    [8 corners, 12 edge vertices, 6 face vertices, centre]
//    (grid_objects_3d.cpp:162-237).
//  * Subset inheritance: every new object takes its parent's subset
//    (tools/subset_handler_interface.cpp:975-1005, non-strict inheritance).
//  * LEVEL DoF numbering: per subset ascending, in subset-list (= creation)
//    order (lib_disc/dof_manager/dof_distribution.cpp:1401-1421).
//
// The base grids (one unit hexahedron; the unit square as two triangles) are
// defined here as if their elements were registered with autogeneration, with
// subset 0 = "Inner" (volume/faces of the domain and inner edges) and
// subset 1 = "Boundary" (all boundary vertices, edges, faces).
#include <algorithm>
#include <cassert>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

namespace {

// hexahedron_rules.h: EDGE_VRT_INDS, FACE_VRT_INDS
const int HEX_EDGE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                             {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
const int HEX_FACE[6][4] = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2},
                            {2, 6, 7, 3}, {3, 7, 4, 0}, {4, 7, 6, 5}};

// tetrahedron_rules.h: EDGE_VRT_INDS, FACE_VRT_INDS
const int TET_EDGE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
const int TET_FACE[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {0, 3, 1}};

uint64_t edge_key(int a, int b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

Key4 face_key(const int* v, int nv) {
    Key4 k;
    for (int i = 0; i < 4; ++i) k.v[i] = i < nv ? v[i] : -1;
    std::sort(k.v, k.v + nv);
    return k;
}

struct Builder {
    Level& L;
    // the coarse object whose children are being created (recorded as the
    // parent of every new edge, face and volume)
    int8_t cur_ptype = -1;
    int cur_parent = -1;
    explicit Builder(Level& l) : L(l) {}
    void parent_of(int8_t t, int id) {
        cur_ptype = t;
        cur_parent = id;
    }

    int add_vertex(const std::array<int64_t, 3>& p, int si, int8_t ptype, int parent) {
        L.vpos.push_back(p);
        L.vsub.push_back(si);
        L.vparent_type.push_back(ptype);
        L.vparent.push_back(parent);
        return (int)L.vpos.size() - 1;
    }
    int find_edge(int a, int b) const {
        auto it = L.edge_map.find(edge_key(a, b));
        return it == L.edge_map.end() ? -1 : it->second;
    }
    int add_edge(int a, int b, int si) {
        int id = (int)L.edges.size();
        L.edges.push_back({a, b});
        L.esub.push_back(si);
        L.eparent_type.push_back(cur_ptype);
        L.eparent.push_back(cur_parent);
        L.edge_map.emplace(edge_key(a, b), id);
        return id;
    }
    // Grid::register_face with FACEOPT_AUTOGENERATE_EDGES
    int register_face(const int* v, int nv, int si) {
        int id = (int)L.faces.size();
        std::array<int, 4> fv{-1, -1, -1, -1};
        for (int i = 0; i < nv; ++i) fv[i] = v[i];
        L.faces.push_back(fv);
        L.fnv.push_back((int8_t)nv);
        L.fsub.push_back(si);
        L.fparent_type.push_back(cur_ptype);
        L.fparent.push_back(cur_parent);
        L.face_map.emplace(face_key(v, nv), id);
        std::array<int, 4> fe{-1, -1, -1, -1};
        for (int i = 0; i < nv; ++i) {
            int a = v[i], b = v[(i + 1) % nv];
            int e = find_edge(a, b);
            if (e < 0) e = add_edge(a, b, si);
            fe[i] = e;
        }
        L.face_edges.push_back(fe);
        return id;
    }
    int find_face(const int* v, int nv) const {
        auto it = L.face_map.find(face_key(v, nv));
        return it == L.face_map.end() ? -1 : it->second;
    }
    // Grid::register_volume with VOLOPT_AUTOGENERATE_FACES (hexahedron)
    int register_hex(const int* h, int si) {
        int id = (int)L.vols.size();
        std::array<int, 8> hv;
        for (int i = 0; i < 8; ++i) hv[i] = h[i];
        L.vols.push_back(hv);
        L.volnv.push_back(8);
        L.volsub.push_back(si);
        L.volparent_type.push_back(cur_ptype);
        L.volparent.push_back(cur_parent);
        std::array<int, 6> vf;
        for (int i = 0; i < 6; ++i) {
            const int* f = HEX_FACE[i];
            int fd[4] = {h[f[0]], h[f[3]], h[f[2]], h[f[1]]};   // Hexahedron::face_desc
            int fid = find_face(fd, 4);
            if (fid < 0) fid = register_face(fd, 4, si);
            vf[i] = fid;
        }
        L.vol_faces.push_back(vf);
        std::array<int, 12> ve;
        for (int i = 0; i < 12; ++i) {
            int e = find_edge(h[HEX_EDGE[i][0]], h[HEX_EDGE[i][1]]);
            if (e < 0) e = add_edge(h[HEX_EDGE[i][0]], h[HEX_EDGE[i][1]], si);
            ve[i] = e;
        }
        L.vol_edges.push_back(ve);
        return id;
    }
    // Grid::register_volume with VOLOPT_AUTOGENERATE_FACES (tetrahedron):
    // missing faces in local face order, created as Tetrahedron::create_face
    // does (vertices f0, f2, f1; grid_objects_3d.cpp:361-368)
    int register_tet(const int* t, int si) {
        int id = (int)L.vols.size();
        std::array<int, 8> tv{t[0], t[1], t[2], t[3], -1, -1, -1, -1};
        L.vols.push_back(tv);
        L.volnv.push_back(4);
        L.volsub.push_back(si);
        L.volparent_type.push_back(cur_ptype);
        L.volparent.push_back(cur_parent);
        std::array<int, 6> vf{-1, -1, -1, -1, -1, -1};
        for (int i = 0; i < 4; ++i) {
            const int* f = TET_FACE[i];
            int fd[3] = {t[f[0]], t[f[2]], t[f[1]]};
            int fid = find_face(fd, 3);
            if (fid < 0) fid = register_face(fd, 3, si);
            vf[i] = fid;
        }
        L.vol_faces.push_back(vf);
        std::array<int, 12> ve;
        ve.fill(-1);
        for (int i = 0; i < 6; ++i) {
            int e = find_edge(t[TET_EDGE[i][0]], t[TET_EDGE[i][1]]);
            if (e < 0) e = add_edge(t[TET_EDGE[i][0]], t[TET_EDGE[i][1]], si);
            ve[i] = e;
        }
        L.vol_edges.push_back(ve);
        return id;
    }
};

std::array<int64_t, 3> avg(const Level& L, const int* v, int n) {
    std::array<int64_t, 3> s{0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) s[d] += L.vpos[v[i]][d];
    for (int d = 0; d < 3; ++d) {
        if (s[d] % n != 0) throw std::runtime_error("oracle grid: non-lattice midpoint");
        s[d] /= n;
    }
    return s;
}

// Regular refinement of a tetrahedron, rule STANDARD
// (tetrahedron_rules.cpp:487-609): four corner children, then the inner
// octahedron split along the shortest of the diagonals between the midpoints
// of opposite edges (0-5, 1-3, 2-4), 2-4 unless another is strictly shorter
// (1-3 checked first, then 0-5). Children are registered in rule order. The
// corner positions are the refiner's projector geometry
// (global_multi_grid_refiner.cpp:388-396); lattice coordinates make the
// squared distances exact.
void refine_tet(const Level& C, Builder& B, const std::vector<int>& cv, const std::vector<int>& me, int v) {
    const int NV = 4;
    int all[10];
    for (int j = 0; j < 4; ++j) all[j] = cv[C.vols[v][j]];
    for (int j = 0; j < 6; ++j) all[NV + j] = me[C.vol_edges[v][j]];
    const int si = C.volsub[v];
    auto P = [&](int j) { return C.vpos[C.vols[v][j]]; };
    // squared distance between midpoints of (a,b) and (c,d), times 4
    auto d2 = [&](int a, int b, int c, int d) {
        int64_t s = 0;
        for (int k = 0; k < 3; ++k) {
            const int64_t x = (P(a)[k] + P(b)[k]) - (P(c)[k] + P(d)[k]);
            s += x * x;
        }
        return s;
    };
    const int64_t d05 = d2(0, 1, 2, 3), d13 = d2(1, 2, 0, 3);
    int64_t d = d2(0, 2, 1, 3);
    int best = 2;
    if (d13 < d) {
        best = 1;
        d = d13;
    }
    if (d05 < d) best = 0;
    const int E = NV;
    const int corner[4][4] = {{0, E + 0, E + 2, E + 3},
                              {1, E + 1, E + 0, E + 4},
                              {2, E + 2, E + 1, E + 5},
                              {E + 3, E + 4, E + 5, 3}};
    const int inner[3][4][4] = {
        {{E + 0, E + 1, E + 2, E + 5}, {E + 1, E + 4, E + 5, E + 0}, {E + 2, E + 5, E + 3, E + 0}, {E + 0, E + 3, E + 4, E + 5}},
        {{E + 0, E + 1, E + 2, E + 3}, {E + 1, E + 4, E + 5, E + 3}, {E + 2, E + 5, E + 3, E + 1}, {E + 0, E + 3, E + 4, E + 1}},
        {{E + 0, E + 2, E + 3, E + 4}, {E + 1, E + 2, E + 0, E + 4}, {E + 2, E + 3, E + 4, E + 5}, {E + 4, E + 1, E + 2, E + 5}}};
    for (auto& c : corner) {
        int t[4] = {all[c[0]], all[c[1]], all[c[2]], all[c[3]]};
        B.register_tet(t, si);
    }
    for (auto& c : inner[best]) {
        int t[4] = {all[c[0]], all[c[1]], all[c[2]], all[c[3]]};
        B.register_tet(t, si);
    }
}

Level refine(const Level& C) {
    Level F;
    F.dim = C.dim;
    Builder B(F);
    const int nv = (int)C.vpos.size();
    std::vector<int> cv(nv), me(C.edges.size()), cf(C.faces.size(), -1);
    // 1. clones of old vertices
    for (int v = 0; v < nv; ++v) cv[v] = B.add_vertex(C.vpos[v], C.vsub[v], OT_VERTEX, v);
    // 2. edge midpoints and child edges
    for (size_t e = 0; e < C.edges.size(); ++e) {
        int ab[2] = {C.edges[e][0], C.edges[e][1]};
        int m = B.add_vertex(avg(C, ab, 2), C.esub[e], OT_EDGE, (int)e);
        me[e] = m;
        B.parent_of(OT_EDGE, (int)e);
        B.add_edge(cv[ab[0]], m, C.esub[e]);
        B.add_edge(m, cv[ab[1]], C.esub[e]);
    }
    // 3. faces
    for (size_t f = 0; f < C.faces.size(); ++f) {
        const int n = C.fnv[f];
        int vr[4], ev[4];
        for (int j = 0; j < n; ++j) {
            vr[j] = cv[C.faces[f][j]];
            ev[j] = me[C.face_edges[f][j]];
        }
        const int si = C.fsub[f];
        B.parent_of(OT_FACE, (int)f);
        if (n == 4) {
            int c = B.add_vertex(avg(C, C.faces[f].data(), 4), si, OT_FACE, (int)f);
            cf[f] = c;
            int q[4][4] = {{vr[0], ev[0], c, ev[3]},
                           {vr[1], ev[1], c, ev[0]},
                           {vr[2], ev[2], c, ev[1]},
                           {vr[3], ev[3], c, ev[2]}};
            for (auto& qq : q) B.register_face(qq, 4, si);
        } else {
            int t[4][3] = {{vr[0], ev[0], ev[2]},
                           {vr[1], ev[1], ev[0]},
                           {vr[2], ev[2], ev[1]},
                           {ev[0], ev[1], ev[2]}};
            for (auto& tt : t) B.register_face(tt, 3, si);
        }
    }
    // 4. volumes (hexahedra, regular rule)
    for (size_t v = 0; v < C.vols.size(); ++v) {
        B.parent_of(OT_VOLUME, (int)v);
        if (C.volnv[v] == 4) {
            refine_tet(C, B, cv, me, (int)v);
            continue;
        }
        if (C.volnv[v] != 8) throw std::runtime_error("oracle grid: unsupported volume");
        int all[27];
        for (int j = 0; j < 8; ++j) all[j] = cv[C.vols[v][j]];
        for (int j = 0; j < 12; ++j) all[8 + j] = me[C.vol_edges[v][j]];
        for (int j = 0; j < 6; ++j) all[20 + j] = cf[C.vol_faces[v][j]];
        const int si = C.volsub[v];
        all[26] = B.add_vertex(avg(C, C.vols[v].data(), 8), si, OT_VOLUME, (int)v);
        const int E = 8, Fo = 20, V = 26;
        const int rule[8][8] = {
            {0, E, Fo, E + 3, E + 4, Fo + 1, V, Fo + 4},
            {E, 1, E + 1, Fo, Fo + 1, E + 5, Fo + 2, V},
            {Fo, E + 1, 2, E + 2, V, Fo + 2, E + 6, Fo + 3},
            {E + 3, Fo, E + 2, 3, Fo + 4, V, Fo + 3, E + 7},
            {E + 4, Fo + 1, V, Fo + 4, 4, E + 8, Fo + 5, E + 11},
            {Fo + 1, E + 5, Fo + 2, V, E + 8, 5, E + 9, Fo + 5},
            {V, Fo + 2, E + 6, Fo + 3, Fo + 5, E + 9, 6, E + 10},
            {Fo + 4, V, Fo + 3, E + 7, E + 11, Fo + 5, E + 10, 7}};
        for (int c = 0; c < 8; ++c) {
            int h[8];
            for (int j = 0; j < 8; ++j) h[j] = all[rule[c][j]];
            B.register_hex(h, si);
        }
    }
    return F;
}

}  // namespace

void number_dofs(Level& L, int num_subsets) {
    const int nv = (int)L.vpos.size();
    L.dof.assign(nv, -1);
    L.vtx_of_dof.assign(nv, -1);
    int k = 0;
    for (int si = 0; si < num_subsets; ++si)
        for (int v = 0; v < nv; ++v)
            if (L.vsub[v] == si) {
                L.dof[v] = k;
                L.vtx_of_dof[k] = v;
                ++k;
            }
    L.n_dof = k;
}

MultiGrid build_hierarchy(int kind, int num_refs) {
    MultiGrid mg;
    mg.num_subsets = 2;
    mg.finest_scale = (int64_t)1 << num_refs;
    const int64_t S = mg.finest_scale;
    Level L0;
    Builder B(L0);
    if (kind == 0) {
        mg.dim = L0.dim = 3;
        const int64_t c[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                 {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        for (auto& p : c) B.add_vertex({p[0] * S, p[1] * S, p[2] * S}, 1, -1, -1);
        int h[8] = {0, 1, 2, 3, 4, 5, 6, 7};
        B.register_hex(h, 0);
        // every face, edge and vertex of the single cube lies on the boundary
        for (auto& s : L0.fsub) s = 1;
        for (auto& s : L0.esub) s = 1;
    } else if (kind == 1) {
        mg.dim = L0.dim = 2;
        const int64_t c[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (auto& p : c) B.add_vertex({p[0] * S, p[1] * S, 0}, 1, -1, -1);
        int t0[3] = {0, 1, 2}, t1[3] = {0, 2, 3};
        B.register_face(t0, 3, 0);
        B.register_face(t1, 3, 0);
        for (size_t e = 0; e < L0.edges.size(); ++e) {
            int a = L0.edges[e][0], b = L0.edges[e][1];
            bool diag = (std::min(a, b) == 0 && std::max(a, b) == 2);
            L0.esub[e] = diag ? 0 : 1;
        }
    } else if (kind == 2) {
        // unit cube as 3x3x3 hexahedra, each split into 6 Kuhn tetrahedra
        // around its main diagonal; vertices in lexicographic order (x
        // fastest), cells likewise, per cell the tetrahedra
        // p0, p0+e_a, p0+e_a+e_b, p0+(1,1,1) for (a,b,c) in lexicographic
        // permutation order. Odd permutations are stored as (P0,P3,P2,P1) so
        // that every tetrahedron is positively oriented while its two face
        // diagonals stay the local edges (0,2) and (1,3).
        mg.dim = L0.dim = 3;
        const int M = 3;
        const int64_t s = S;
        mg.finest_scale = M * S;
        auto vid = [&](int i, int j, int k) { return (k * (M + 1) + j) * (M + 1) + i; };
        for (int k = 0; k <= M; ++k)
            for (int j = 0; j <= M; ++j)
                for (int i = 0; i <= M; ++i) {
                    const bool bnd = i == 0 || j == 0 || k == 0 || i == M || j == M || k == M;
                    B.add_vertex({i * s, j * s, k * s}, bnd ? 1 : 0, -1, -1);
                }
        const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        const bool odd[6] = {false, true, true, false, false, true};
        for (int k = 0; k < M; ++k)
            for (int j = 0; j < M; ++j)
                for (int i = 0; i < M; ++i)
                    for (int p = 0; p < 6; ++p) {
                        int c[3] = {i, j, k};
                        int P[4];
                        P[0] = vid(c[0], c[1], c[2]);
                        for (int q = 0; q < 3; ++q) {
                            c[perms[p][q]] += 1;
                            P[q + 1] = vid(c[0], c[1], c[2]);
                        }
                        int t[4] = {P[0], P[1], P[2], P[3]};
                        if (odd[p]) {
                            t[1] = P[3];
                            t[3] = P[1];
                        }
                        B.register_tet(t, 0);
                    }
        // boundary faces/edges: all vertices on one boundary plane
        auto on_plane = [&](const int* v, int n) {
            for (int ax = 0; ax < 3; ++ax)
                for (int side = 0; side < 2; ++side) {
                    const int64_t val = side ? M * s : 0;
                    bool all = true;
                    for (int q = 0; q < n; ++q) all = all && L0.vpos[v[q]][ax] == val;
                    if (all) return true;
                }
            return false;
        };
        for (size_t f = 0; f < L0.faces.size(); ++f) L0.fsub[f] = on_plane(L0.faces[f].data(), 3) ? 1 : 0;
        for (size_t e = 0; e < L0.edges.size(); ++e) L0.esub[e] = on_plane(L0.edges[e].data(), 2) ? 1 : 0;
    } else {
        throw std::runtime_error("oracle grid: unknown base grid kind");
    }
    mg.levels.push_back(std::move(L0));
    for (int r = 0; r < num_refs; ++r) mg.levels.push_back(refine(mg.levels.back()));
    for (auto& L : mg.levels) number_dofs(L, mg.num_subsets);
    return mg;
}

}  // namespace orc
