// Oracle (test infrastructure only): Q2 Lagrange on the hexahedral hierarchy.
//
// Rules followed:
//  * DoFDistribution::reinit (lib_disc/dof_manager/dof_distribution.cpp:
//    1401-1444): reinit<Vertex>, <Edge>, <Face>, <Volume> in turn, each over
//    the subsets in ascending order and the subset's objects in list
//    (= creation) order; Q2 places one DoF on every object.
//  * Element loop (elem_disc_assemble_util.h:902-1003) over the level's
//    hexahedra of subset "Inner"; local shape a of LagrangeLSFS<Hexahedron, 2>
//    (q2_hex_multi_index) belongs to a vertex, an edge, a face or the volume
//    of the element, whose DoF it takes (DoFDistribution::indices,
//    dof_distribution.cpp:1117-1206). Dirichlet rows as the Q1 assembly.
//  * StdTransfer::assemble_prolongation, LAGRANGE branch
//    (lib_disc/operator/linear_operator/std_transfer_impl.h:214-388): every
//    fine object takes the DoFs of its parent's closure (dof_indices) and the
//    parent's Lagrange shape values at the object's DoF position
//    (ShapesAtGlobalPosition), entries created even where the value is 0.
//  * adjust_prolongation (lagrange_dirichlet_boundary_impl.h:485-614): a fine
//    Dirichlet row is zeroed and its parent's own DoF set to 1;
//    adjust_restriction (:617-740): R = P^T keeping zeros; for every fine
//    Dirichlet object in type order (vertex, edge, face, volume) and list
//    order, the row of its parent's own DoF is zeroed and the entry of the
//    object set to 1 (so the last such child of a parent wins).
#include <algorithm>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

namespace {

// reference hexahedron edges and faces in the order of the Q2 shapes
// (hexahedron_rules.h EDGE_VRT_INDS; faces as in q2_hex_multi_index)
const int QE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5}, {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
const int QF[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};

uint64_t ekey(int a, int b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

int find_edge(const Level& L, int a, int b) {
    auto it = L.edge_map.find(ekey(a, b));
    if (it == L.edge_map.end()) throw std::runtime_error("oracle q2: missing edge");
    return it->second;
}

int find_face(const Level& L, const int* v) {
    Key4 k;
    for (int i = 0; i < 4; ++i) k.v[i] = v[i];
    std::sort(k.v, k.v + 4);
    auto it = L.face_map.find(k);
    if (it == L.face_map.end()) throw std::runtime_error("oracle q2: missing face");
    return it->second;
}

// the object's vertices (positions, doubled units are 2 * vpos)
struct Obj {
    int8_t type;
    int id;
};

void obj_vertices(const Level& L, Obj o, std::vector<int>& vs) {
    vs.clear();
    switch (o.type) {
        case OT_VERTEX: vs.push_back(o.id); break;
        case OT_EDGE: vs = {L.edges[o.id][0], L.edges[o.id][1]}; break;
        case OT_FACE: vs.assign(L.faces[o.id].begin(), L.faces[o.id].begin() + 4); break;
        default: vs.assign(L.vols[o.id].begin(), L.vols[o.id].end()); break;
    }
}

// closure DoFs of a coarse object: its vertices, edges, faces and itself;
// the object's own DoF is the last entry
void closure_dofs(const Level& L, const Q2Numbering& Q, Obj o, std::vector<int>& dofs) {
    dofs.clear();
    switch (o.type) {
        case OT_VERTEX: dofs.push_back(Q.vdof[o.id]); break;
        case OT_EDGE:
            dofs = {Q.vdof[L.edges[o.id][0]], Q.vdof[L.edges[o.id][1]], Q.edof[o.id]};
            break;
        case OT_FACE:
            for (int i = 0; i < 4; ++i) dofs.push_back(Q.vdof[L.faces[o.id][i]]);
            for (int i = 0; i < 4; ++i) dofs.push_back(Q.edof[L.face_edges[o.id][i]]);
            dofs.push_back(Q.fdof[o.id]);
            break;
        default:
            for (int i = 0; i < 8; ++i) dofs.push_back(Q.vdof[L.vols[o.id][i]]);
            for (int i = 0; i < 12; ++i) dofs.push_back(Q.edof[L.vol_edges[o.id][i]]);
            for (int i = 0; i < 6; ++i) dofs.push_back(Q.fdof[L.vol_faces[o.id][i]]);
            dofs.push_back(Q.hdof[o.id]);
            break;
    }
}

double lag2(int m, double t) {
    return m == 0 ? (1.0 - t) * (1.0 - 2.0 * t) : m == 1 ? 4.0 * t * (1.0 - t) : t * (2.0 * t - 1.0);
}

// value at doubled position x of the Q2 shape of the parent object (axis-
// aligned box of the lattice) that belongs to its closure node at doubled
// position c: the product over the box's extended axes of the 1D quadratic
// Lagrange polynomial of the node at the local coordinate of x
double parent_shape(const Level& C, Obj parent, const std::array<int64_t, 3>& x, const std::array<int64_t, 3>& c) {
    std::vector<int> vs;
    obj_vertices(C, parent, vs);
    double v = 1.0;
    for (int d = 0; d < 3; ++d) {
        int64_t lo = 2 * C.vpos[vs[0]][d], hi = lo;
        for (int w : vs) {
            lo = std::min<int64_t>(lo, 2 * C.vpos[w][d]);
            hi = std::max<int64_t>(hi, 2 * C.vpos[w][d]);
        }
        if (lo == hi) continue;
        const double t = (double)(x[d] - lo) / (double)(hi - lo);
        const int64_t m2 = 2 * (c[d] - lo);
        if (m2 % (hi - lo) != 0) throw std::runtime_error("oracle q2: closure node off the parent's nodes");
        v *= lag2((int)(m2 / (hi - lo)), t);
    }
    return v;
}

Obj parent_of(const Level& F, int8_t type, int id) {
    switch (type) {
        case OT_VERTEX: return {F.vparent_type[id], F.vparent[id]};
        case OT_EDGE: return {F.eparent_type[id], F.eparent[id]};
        case OT_FACE: return {F.fparent_type[id], F.fparent[id]};
        default: return {F.volparent_type[id], F.volparent[id]};
    }
}

int dof_of(const Q2Numbering& Q, int8_t type, int id) {
    return type == OT_VERTEX ? Q.vdof[id] : type == OT_EDGE ? Q.edof[id] : type == OT_FACE ? Q.fdof[id] : Q.hdof[id];
}

size_t count_of(const Level& L, int8_t type) {
    return type == OT_VERTEX ? L.vpos.size()
           : type == OT_EDGE ? L.edges.size()
           : type == OT_FACE ? L.faces.size()
                             : L.vols.size();
}

int subset_of(const Level& L, int8_t type, int id) {
    return type == OT_VERTEX ? L.vsub[id] : type == OT_EDGE ? L.esub[id] : type == OT_FACE ? L.fsub[id] : L.volsub[id];
}

}  // namespace

Q2Numbering q2_numbering(const Level& L) {
    Q2Numbering Q;
    Q.vdof.assign(L.vpos.size(), -1);
    Q.edof.assign(L.edges.size(), -1);
    Q.fdof.assign(L.faces.size(), -1);
    Q.hdof.assign(L.vols.size(), -1);
    for (size_t h = 0; h < L.vols.size(); ++h)
        if (L.volnv[h] != 8) throw std::runtime_error("oracle q2: hexahedra only");
    int k = 0;
    for (int8_t t = OT_VERTEX; t <= OT_VOLUME; ++t)
        for (int si = 0; si < 2; ++si)
            for (size_t i = 0; i < count_of(L, t); ++i) {
                if (subset_of(L, t, (int)i) != si) continue;
                std::vector<int> vs;
                obj_vertices(L, {t, (int)i}, vs);
                std::array<int64_t, 3> s{0, 0, 0};
                for (int w : vs)
                    for (int d = 0; d < 3; ++d) s[d] += L.vpos[w][d];
                // doubled position: 2 * (sum / count)
                for (int d = 0; d < 3; ++d) s[d] = 2 * s[d] / (int64_t)vs.size();
                (t == OT_VERTEX ? Q.vdof : t == OT_EDGE ? Q.edof : t == OT_FACE ? Q.fdof : Q.hdof)[i] = k;
                Q.pos2.push_back(s);
                Q.dir.push_back(si == 1 ? 1 : 0);
                ++k;
            }
    Q.n_dof = k;
    return Q;
}

void assemble_q2_hex(const MultiGrid& mg, int lev, const Q2Numbering& Q, const Coefficient& k, double f_const,
                     CSR& A, std::vector<double>& b) {
    const Level& L = mg.levels[lev];
    RowBuilder B;
    B.resize_and_clear(Q.n_dof, Q.n_dof);
    b.assign(Q.n_dof, 0.0);
    const double scale = 1.0 / (double)mg.finest_scale;
    for (size_t e = 0; e < L.vols.size(); ++e) {
        if (L.volsub[e] != 0) continue;
        const auto& hv = L.vols[e];
        double X[8][3];
        int ind[27];
        for (int a = 0; a < 8; ++a) {
            ind[a] = Q.vdof[hv[a]];
            for (int d = 0; d < 3; ++d) X[a][d] = (double)L.vpos[hv[a]][d] * scale;
        }
        for (int j = 0; j < 12; ++j) ind[8 + j] = Q.edof[find_edge(L, hv[QE[j][0]], hv[QE[j][1]])];
        for (int j = 0; j < 6; ++j) {
            const int fv[4] = {hv[QF[j][0]], hv[QF[j][1]], hv[QF[j][2]], hv[QF[j][3]]};
            ind[20 + j] = Q.fdof[find_face(L, fv)];
        }
        ind[26] = Q.hdof[e];
        double Ke[27][27], fe[27];
        q2_hex_element_coef(X, k, f_const, Ke, fe);
        for (int i = 0; i < 27; ++i) {
            for (int j = 0; j < 27; ++j) B.at(ind[i], ind[j]) += Ke[i][j];
            b[ind[i]] += fe[i];
        }
    }
    A = defragment(B);
    for (int64_t r = 0; r < A.n_rows; ++r)
        if (Q.dir[r]) {
            set_row(A, r, 0.0);
            double* d = find_entry(A, r, r);
            if (!d) throw std::runtime_error("oracle q2: Dirichlet row without diagonal");
            *d = 1.0;
            b[r] = 0.0;
        }
}

CSR prolongation_q2(const MultiGrid& mg, int fine_lev, const Q2Numbering& Qc, const Q2Numbering& Qf) {
    const Level& F = mg.levels[fine_lev];
    const Level& C = mg.levels[fine_lev - 1];
    RowBuilder B;
    B.resize_and_clear(Qf.n_dof, Qc.n_dof);
    std::vector<int> dofs;
    for (int8_t t = OT_VERTEX; t <= OT_VOLUME; ++t)
        for (size_t i = 0; i < count_of(F, t); ++i) {
            const Obj par = parent_of(F, t, (int)i);
            if (par.type < 0) throw std::runtime_error("oracle q2: fine object without parent");
            const int row = dof_of(Qf, t, (int)i);
            closure_dofs(C, Qc, par, dofs);
            for (int cd : dofs) B.at(row, cd) = parent_shape(C, par, Qf.pos2[row], Qc.pos2[cd]);
        }
    CSR P = defragment(B);
    for (int8_t t = OT_VERTEX; t <= OT_VOLUME; ++t)
        for (size_t i = 0; i < count_of(F, t); ++i) {
            const int row = dof_of(Qf, t, (int)i);
            if (!Qf.dir[row]) continue;
            const Obj par = parent_of(F, t, (int)i);
            set_row(P, row, 0.0);
            *find_entry(P, row, dof_of(Qc, par.type, par.id)) = 1.0;
        }
    return P;
}

CSR restriction_q2(const MultiGrid& mg, int fine_lev, const Q2Numbering& Qc, const Q2Numbering& Qf, const CSR& P) {
    const Level& F = mg.levels[fine_lev];
    CSR R = transpose_keep_zeros(P);
    for (int8_t t = OT_VERTEX; t <= OT_VOLUME; ++t)
        for (size_t i = 0; i < count_of(F, t); ++i) {
            if (subset_of(F, t, (int)i) != 1) continue;
            const Obj par = parent_of(F, t, (int)i);
            const int crow = dof_of(Qc, par.type, par.id);
            set_row(R, crow, 0.0);
            *find_entry(R, crow, dof_of(Qf, t, (int)i)) = 1.0;
        }
    return R;
}

}  // namespace orc
