// Device refinement of the lattice hierarchies and the LEVEL DoF numbering.
//
// The entity lists of each level are built on the device in the creation order
// of GlobalMultiGridRefiner::perform_refinement
// (lib_grid/refinement/global_multi_grid_refiner.cpp:135-455), the same order
// the host hierarchy (hierarchy.cpp) replays:
//   vertices  [copies of the coarse vertices | edge midpoints | quadrilateral
//              centres | hexahedron centres], each part in coarse order;
//   edges     [2 children per coarse edge (the half at its first vertex
//              first) | the new edges inside each coarse face | inside each
//              coarse volume];
//   faces     [4 children per coarse face | the new faces inside each coarse
//              volume];
//   volumes   [8 children per coarse volume].
// Inside one coarse face or volume the new edges and faces appear in the order
// in which the children register them (Grid::register_face / register_volume
// create the missing sides in local order, grid_connection_managment.cpp:
// 922-1010, 1440-1560). That order is the same for every face or volume of a
// type (tetrahedra: one order per inner diagonal), so it is derived once on the
// host by replaying the registration on one element (derive_pattern). Each
// list entry is then a closed form of its parent's entry and is written by one
// thread; the sides of new faces and volumes are found by their centre on the
// doubled lattice of the new level, as the host hierarchy finds them.
//
// DoF numbering (DoFDistribution::reinit, lib_disc/dof_manager/
// dof_distribution.cpp:1401-1421): the vertices of subset "Inner" first, then
// those of "Boundary", each in vertex order -- a stable partition, computed
// with one exclusive scan over the vertex list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hierarchy.hpp"
#include "mesh_tables.hpp"
#include "ug_gpu/gpu_algebra.hpp"

namespace ug {
namespace gpu {

using namespace ugamd;
using namespace ugamd::mesh;

namespace {

// local creation patterns, passed to the kernels by value
struct RefinePattern {
    int8_t quad_child[4][4];   // [corners 0-3 | edge vertices 4-7 | centre 8]
    int8_t quad_edge[4][2];
    int8_t tri_child[4][3];    // [corners 0-2 | edge vertices 3-5]
    int8_t tri_edge[3][2];
    int8_t hex_child[8][8];    // [corners 0-7 | edges 8-19 | faces 20-25 | centre 26]
    int8_t hex_face[12][4];
    int8_t hex_edge[6][2];
    int8_t tet_child[3][8][4]; // per inner diagonal; [corners 0-3 | edges 4-9]
    int8_t tet_face[3][8][3];
    int8_t tet_edge[3][2];
    int8_t hex_side_face[6][4];    // kHexFace
    int8_t hex_side_edge[12][2];   // kHexEdge
    int8_t tet_side_edge[6][2];    // kTetEdge
};

// Replays the registration of the children of one face or volume in a local
// index space. mask[i] holds the boundary pieces of the parent (edges of a
// face, faces of a volume) that local vertex i lies on; a side whose vertices
// share a boundary piece lies in the parent's boundary and exists already
// (the boundary was refined before the interior), any other side is created on
// first registration, with its vertices in the registering order.
struct Replay {
    std::vector<uint32_t> mask;
    std::vector<std::vector<int>> faces, edges;

    bool on_boundary(const std::vector<int>& v) const {
        uint32_t m = 0xffffffffu;
        for (int i : v) m &= mask[i];
        return m != 0;
    }
    static bool same_set(const std::vector<int>& a, const std::vector<int>& b) {
        if (a.size() != b.size()) return false;
        for (int x : a)
            if (std::find(b.begin(), b.end(), x) == b.end()) return false;
        return true;
    }
    static bool known(const std::vector<std::vector<int>>& list, const std::vector<int>& v) {
        for (const auto& e : list)
            if (same_set(e, v)) return true;
        return false;
    }
    void edge(int a, int b) {
        const std::vector<int> v{a, b};
        if (on_boundary(v) || known(edges, v)) return;
        edges.push_back(v);
    }
    // Grid::register_face: the face, then its missing edges in local order
    void face(const std::vector<int>& v) {
        if (on_boundary(v) || known(faces, v)) return;
        faces.push_back(v);
        const int k = (int)v.size();
        for (int i = 0; i < k; ++i) edge(v[i], v[(i + 1) % k]);
    }
};

RefinePattern derive_pattern() {
    RefinePattern P{};
    auto need = [](bool ok, const char* what) {
        if (!ok) throw UGError(std::string("device hierarchy: unexpected refinement pattern (") + what + ")");
    };
    {   // quadrilateral: children in the face pass, edges j = (j, j+1)
        Replay R;
        R.mask.assign(9, 0);
        for (int j = 0; j < 4; ++j) {
            R.mask[j] = (1u << j) | (1u << ((j + 3) % 4));
            R.mask[4 + j] = 1u << j;
        }
        for (int c = 0; c < 4; ++c) {
            for (int k = 0; k < 4; ++k) P.quad_child[c][k] = (int8_t)kQuadChild[c][k];
            R.face({kQuadChild[c][0], kQuadChild[c][1], kQuadChild[c][2], kQuadChild[c][3]});
        }
        need(R.edges.size() == 4, "quadrilateral");
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 2; ++k) P.quad_edge[i][k] = (int8_t)R.edges[i][k];
    }
    {   // triangle
        Replay R;
        R.mask.assign(6, 0);
        for (int j = 0; j < 3; ++j) {
            R.mask[j] = (1u << j) | (1u << ((j + 2) % 3));
            R.mask[3 + j] = 1u << j;
        }
        for (int c = 0; c < 4; ++c) {
            for (int k = 0; k < 3; ++k) P.tri_child[c][k] = (int8_t)kTriChild[c][k];
            R.face({kTriChild[c][0], kTriChild[c][1], kTriChild[c][2]});
        }
        need(R.edges.size() == 3, "triangle");
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 2; ++k) P.tri_edge[i][k] = (int8_t)R.edges[i][k];
    }
    {   // hexahedron: Hexahedron faces are registered as (f0, f3, f2, f1)
        Replay R;
        R.mask.assign(27, 0);
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < 4; ++k) R.mask[kHexFace[i][k]] |= 1u << i;
        for (int j = 0; j < 12; ++j) R.mask[8 + j] = R.mask[kHexEdge[j][0]] & R.mask[kHexEdge[j][1]];
        for (int i = 0; i < 6; ++i) R.mask[20 + i] = 1u << i;
        for (int c = 0; c < 8; ++c) {
            const int* h = kHexChild[c];
            for (int k = 0; k < 8; ++k) P.hex_child[c][k] = (int8_t)h[k];
            for (int i = 0; i < 6; ++i) {
                const int* f = kHexFace[i];
                R.face({h[f[0]], h[f[3]], h[f[2]], h[f[1]]});
            }
            for (int i = 0; i < 12; ++i) R.edge(h[kHexEdge[i][0]], h[kHexEdge[i][1]]);
        }
        need(R.faces.size() == 12 && R.edges.size() == 6, "hexahedron");
        for (int i = 0; i < 12; ++i)
            for (int k = 0; k < 4; ++k) P.hex_face[i][k] = (int8_t)R.faces[i][k];
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < 2; ++k) P.hex_edge[i][k] = (int8_t)R.edges[i][k];
    }
    for (int best = 0; best < 3; ++best) {
        // tetrahedron: faces registered as (f0, f2, f1) (Tetrahedron::create_face)
        Replay R;
        R.mask.assign(10, 0);
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 3; ++k) R.mask[kTetFace[i][k]] |= 1u << i;
        for (int j = 0; j < 6; ++j) R.mask[4 + j] = R.mask[kTetEdge[j][0]] & R.mask[kTetEdge[j][1]];
        for (int c = 0; c < 8; ++c) {
            const int* t = c < 4 ? kTetCorner[c] : kTetInner[best][c - 4];
            for (int k = 0; k < 4; ++k) P.tet_child[best][c][k] = (int8_t)t[k];
            for (int i = 0; i < 4; ++i) {
                const int* f = kTetFace[i];
                R.face({t[f[0]], t[f[2]], t[f[1]]});
            }
            for (int i = 0; i < 6; ++i) R.edge(t[kTetEdge[i][0]], t[kTetEdge[i][1]]);
        }
        need(R.faces.size() == 8 && R.edges.size() == 1, "tetrahedron");
        for (int i = 0; i < 8; ++i)
            for (int k = 0; k < 3; ++k) P.tet_face[best][i][k] = (int8_t)R.faces[i][k];
        for (int k = 0; k < 2; ++k) P.tet_edge[best][k] = (int8_t)R.edges[0][k];
    }
    for (int i = 0; i < 6; ++i)
        for (int k = 0; k < 4; ++k) P.hex_side_face[i][k] = (int8_t)kHexFace[i][k];
    for (int i = 0; i < 12; ++i)
        for (int k = 0; k < 2; ++k) P.hex_side_edge[i][k] = (int8_t)kHexEdge[i][k];
    for (int i = 0; i < 6; ++i)
        for (int k = 0; k < 2; ++k) P.tet_side_edge[i][k] = (int8_t)kTetEdge[i][k];
    return P;
}

// error bits of the device refinement
enum : int {
    kErrLookup = 1,       // a side of a new face or volume was not found
    kErrCover = 2,        // the vertices do not cover the lattice once
    kErrTetCell = 4,      // a tetrahedron spans two lattice cells
    kErrCellCount = 8,    // a lattice cell does not hold 6 tetrahedra
    kErrParentEdge = 16,  // an edge midpoint without a unit lattice parent edge
};

// one level's entity lists (device pointers); positions packed x | y<<21 | z<<42
struct MeshDev {
    int64_t nv = 0, ne = 0, nf = 0, nh = 0, nt = 0;
    int n = 0, dim = 3, fnv = 4;
    int64_t* vpos = nullptr;
    uint8_t* vsub = nullptr;
    int32_t* e_v = nullptr;
    uint8_t* e_sub = nullptr;
    int32_t* f_v = nullptr;
    int32_t* f_e = nullptr;
    uint8_t* f_sub = nullptr;
    int32_t* h_v = nullptr;
    int32_t* h_f = nullptr;
    int32_t* h_e = nullptr;
    uint8_t* h_sub = nullptr;
    int32_t* t_v = nullptr;
    int32_t* t_e = nullptr;
    uint8_t* t_sub = nullptr;
};

__host__ __device__ __forceinline__ int64_t pk(int64_t x, int64_t y, int64_t z) { return x | (y << 21) | (z << 42); }
__device__ __forceinline__ int px(int64_t p) { return (int)(p & 0x1FFFFF); }
__device__ __forceinline__ int py(int64_t p) { return (int)((p >> 21) & 0x1FFFFF); }
__device__ __forceinline__ int pz(int64_t p) { return (int)((p >> 42) & 0x1FFFFF); }

// global vertex id of a fine vertex given by its local index inside coarse
// face f / hexahedron h / tetrahedron t
__device__ __forceinline__ int32_t face_local(const MeshDev& C, int64_t f, int loc) {
    const int k = C.fnv;
    if (loc < k) return C.f_v[4 * f + loc];
    if (loc < 2 * k) return (int32_t)(C.nv + C.f_e[4 * f + loc - k]);
    return (int32_t)(C.nv + C.ne + f);   // quadrilateral centre
}
__device__ __forceinline__ int32_t hex_local(const MeshDev& C, int64_t h, int loc) {
    if (loc < 8) return C.h_v[8 * h + loc];
    if (loc < 20) return (int32_t)(C.nv + C.h_e[12 * h + loc - 8]);
    if (loc < 26) return (int32_t)(C.nv + C.ne + C.h_f[6 * h + loc - 20]);
    return (int32_t)(C.nv + C.ne + C.nf + h);
}
__device__ __forceinline__ int32_t tet_local(const MeshDev& C, int64_t t, int loc) {
    if (loc < 4) return C.t_v[4 * t + loc];
    return (int32_t)(C.nv + C.t_e[6 * t + loc - 4]);
}

// inner diagonal of the regular tetrahedron refinement: the shortest one
// between midpoints of opposite edges (ties: 0-5 over 1-3 over 2-4), from the
// corner positions as in the host hierarchy
__device__ int tet_diagonal(const MeshDev& C, int64_t t) {
    int64_t x[4], y[4], z[4];
    for (int j = 0; j < 4; ++j) {
        const int64_t p = C.vpos[C.t_v[4 * t + j]];
        x[j] = px(p);
        y[j] = py(p);
        z[j] = pz(p);
    }
    auto d2 = [&](int a, int b, int c, int d) {
        const int64_t dx = x[a] + x[b] - x[c] - x[d], dy = y[a] + y[b] - y[c] - y[d], dz = z[a] + z[b] - z[c] - z[d];
        return dx * dx + dy * dy + dz * dz;
    };
    int best = 2;
    int64_t dbest = d2(0, 2, 1, 3);
    const int64_t d13 = d2(1, 2, 0, 3), d05 = d2(0, 1, 2, 3);
    if (d13 < dbest) {
        best = 1;
        dbest = d13;
    }
    if (d05 < dbest) best = 0;
    return best;
}

__device__ __forceinline__ int64_t gid() { return blockIdx.x * (int64_t)blockDim.x + threadIdx.x; }

__global__ __launch_bounds__(256) void k_refine_vertices(MeshDev C, MeshDev F) {
    const int64_t v = gid();
    if (v >= F.nv) return;
    const int64_t nq = C.fnv == 4 ? C.nf : 0;
    int64_t p;
    uint8_t s;
    int64_t i = v;
    if (i < C.nv) {
        const int64_t q = C.vpos[i];
        p = pk(2 * px(q), 2 * py(q), 2 * pz(q));
        s = C.vsub[i];
    } else if ((i -= C.nv) < C.ne) {
        const int64_t a = C.vpos[C.e_v[2 * i]], b = C.vpos[C.e_v[2 * i + 1]];
        p = pk(px(a) + px(b), py(a) + py(b), pz(a) + pz(b));
        s = C.e_sub[i];
    } else if ((i -= C.ne) < nq) {
        int64_t sx = 0, sy = 0, sz = 0;
        for (int k = 0; k < 4; ++k) {
            const int64_t q = C.vpos[C.f_v[4 * i + k]];
            sx += px(q);
            sy += py(q);
            sz += pz(q);
        }
        p = pk(sx / 2, sy / 2, sz / 2);
        s = C.f_sub[i];
    } else {
        i -= nq;
        int64_t sx = 0, sy = 0, sz = 0;
        for (int k = 0; k < 8; ++k) {
            const int64_t q = C.vpos[C.h_v[8 * i + k]];
            sx += px(q);
            sy += py(q);
            sz += pz(q);
        }
        p = pk(sx / 4, sy / 4, sz / 4);
        s = C.h_sub[i];
    }
    F.vpos[v] = p;
    F.vsub[v] = s;
}

__global__ __launch_bounds__(256) void k_refine_edges(MeshDev C, MeshDev F, RefinePattern P) {
    const int64_t e = gid();
    if (e >= F.ne) return;
    int32_t a, b;
    uint8_t s;
    int64_t i = e;
    if (i < 2 * C.ne) {
        const int64_t k = i >> 1;
        const int32_t m = (int32_t)(C.nv + k);
        if ((i & 1) == 0) {
            a = C.e_v[2 * k];
            b = m;
        } else {
            a = m;
            b = C.e_v[2 * k + 1];
        }
        s = C.e_sub[k];
    } else if ((i -= 2 * C.ne) < (int64_t)C.fnv * C.nf) {
        const int64_t f = i / C.fnv;
        const int j = (int)(i % C.fnv);
        const int8_t* loc = C.fnv == 4 ? P.quad_edge[j] : P.tri_edge[j];
        a = face_local(C, f, loc[0]);
        b = face_local(C, f, loc[1]);
        s = C.f_sub[f];
    } else if ((i -= (int64_t)C.fnv * C.nf) < 6 * C.nh) {
        const int64_t h = i / 6;
        const int j = (int)(i % 6);
        a = hex_local(C, h, P.hex_edge[j][0]);
        b = hex_local(C, h, P.hex_edge[j][1]);
        s = C.h_sub[h];
    } else {
        i -= 6 * C.nh;
        const int best = tet_diagonal(C, i);
        a = tet_local(C, i, P.tet_edge[best][0]);
        b = tet_local(C, i, P.tet_edge[best][1]);
        s = C.t_sub[i];
    }
    F.e_v[2 * e] = a;
    F.e_v[2 * e + 1] = b;
    F.e_sub[e] = s;
}

__global__ __launch_bounds__(256) void k_refine_faces(MeshDev C, MeshDev F, RefinePattern P) {
    const int64_t f = gid();
    if (f >= F.nf) return;
    int32_t v[4] = {-1, -1, -1, -1};
    uint8_t s;
    int64_t i = f;
    if (i < 4 * C.nf) {
        const int64_t pf = i >> 2;
        const int c = (int)(i & 3);
        for (int k = 0; k < C.fnv; ++k)
            v[k] = face_local(C, pf, C.fnv == 4 ? P.quad_child[c][k] : P.tri_child[c][k]);
        s = C.f_sub[pf];
    } else if ((i -= 4 * C.nf) < 12 * C.nh) {
        const int64_t h = i / 12;
        const int j = (int)(i % 12);
        for (int k = 0; k < 4; ++k) v[k] = hex_local(C, h, P.hex_face[j][k]);
        s = C.h_sub[h];
    } else {
        i -= 12 * C.nh;
        const int64_t t = i >> 3;
        const int j = (int)(i & 7);
        const int best = tet_diagonal(C, t);
        for (int k = 0; k < 3; ++k) v[k] = tet_local(C, t, P.tet_face[best][j][k]);
        s = C.t_sub[t];
    }
    for (int k = 0; k < 4; ++k) F.f_v[4 * f + k] = v[k];
    F.f_sub[f] = s;
}

__global__ __launch_bounds__(256) void k_refine_volumes(MeshDev C, MeshDev F, RefinePattern P) {
    const int64_t i = gid();
    if (i < F.nh) {
        const int64_t h = i >> 3;
        const int c = (int)(i & 7);
        for (int k = 0; k < 8; ++k) F.h_v[8 * i + k] = hex_local(C, h, P.hex_child[c][k]);
        F.h_sub[i] = C.h_sub[h];
    }
    if (i < F.nt) {
        const int64_t t = i >> 3;
        const int c = (int)(i & 7);
        const int best = tet_diagonal(C, t);
        for (int k = 0; k < 4; ++k) F.t_v[4 * i + k] = tet_local(C, t, P.tet_child[best][c][k]);
        F.t_sub[i] = C.t_sub[t];
    }
}

__global__ __launch_bounds__(256) void k_fill_i32(int64_t n, int32_t* p, int32_t v) {
    const int64_t i = gid();
    if (i < n) p[i] = v;
}

// key of a side on the doubled lattice of its level: twice its centre
__device__ __forceinline__ int64_t side_key(const MeshDev& F, const int32_t* v, int k) {
    int64_t sx = 0, sy = 0, sz = 0;
    for (int j = 0; j < k; ++j) {
        const int64_t q = F.vpos[v[j]];
        sx += px(q);
        sy += py(q);
        sz += pz(q);
    }
    const int64_t W2 = 2 * (int64_t)F.n + 1;
    const int64_t d = k / 2;   // 2 * centre = sum * 2 / k
    return ((sz / d) * W2 + sy / d) * W2 + sx / d;
}

__global__ __launch_bounds__(256) void k_key_sides(MeshDev F, int32_t* key) {
    const int64_t i = gid();
    if (i < F.ne) key[side_key(F, F.e_v + 2 * i, 2)] = (int32_t)i;
    // only quadrilaterals are looked up (by the hexahedra that own them)
    if (F.fnv == 4 && i < F.nf) key[side_key(F, F.f_v + 4 * i, 4)] = (int32_t)i;
}

__device__ __forceinline__ int32_t lookup(const MeshDev& F, const int32_t* key, const int32_t* v, int k, int* err) {
    const int32_t id = key[side_key(F, v, k)];
    if (id < 0) atomicOr(err, kErrLookup);
    return id;
}

__global__ __launch_bounds__(256) void k_resolve_sides(MeshDev F, const int32_t* key, RefinePattern P, int* err) {
    const int64_t i = gid();
    if (i < F.nf) {
        const int k = F.fnv;
        for (int j = 0; j < 4; ++j) {
            int32_t id = -1;
            if (j < k) {
                const int32_t e[2] = {F.f_v[4 * i + j], F.f_v[4 * i + (j + 1) % k]};
                id = lookup(F, key, e, 2, err);
            }
            F.f_e[4 * i + j] = id;
        }
    }
    if (i < F.nh) {
        const int32_t* h = F.h_v + 8 * i;
        for (int j = 0; j < 6; ++j) {
            const int32_t q[4] = {h[P.hex_side_face[j][0]], h[P.hex_side_face[j][1]], h[P.hex_side_face[j][2]],
                                  h[P.hex_side_face[j][3]]};
            F.h_f[6 * i + j] = lookup(F, key, q, 4, err);
        }
        for (int j = 0; j < 12; ++j) {
            const int32_t e[2] = {h[P.hex_side_edge[j][0]], h[P.hex_side_edge[j][1]]};
            F.h_e[12 * i + j] = lookup(F, key, e, 2, err);
        }
    }
    if (i < F.nt) {
        const int32_t* t = F.t_v + 4 * i;
        for (int j = 0; j < 6; ++j) {
            const int32_t e[2] = {t[P.tet_side_edge[j][0]], t[P.tet_side_edge[j][1]]};
            F.t_e[6 * i + j] = lookup(F, key, e, 2, err);
        }
    }
}

// ---- stable partition (DoF numbering) ------------------------------------
// items are the level's vertices (reference numbering: flag = subset Inner,
// target = the vertex's lattice point) or its lattice points in lexicographic
// order (lexicographic numbering: flag = not on the lattice boundary)
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanItems = kScanThreads * kScanPer;

struct PartitionItems {
    int64_t n = 0;
    int lex = 0;
    int dim = 3;
    int W = 0;                       // lattice points per side
    const int64_t* vpos = nullptr;   // reference numbering
    const uint8_t* vsub = nullptr;
};

__device__ __forceinline__ bool item_flag(const PartitionItems& I, int64_t i) {
    if (!I.lex) return I.vsub[i] == 0;
    const int64_t W = I.W;
    const int64_t x = i % W, y = (i / W) % W, z = I.dim == 3 ? i / (W * W) : 1;
    return x > 0 && x < W - 1 && y > 0 && y < W - 1 && (I.dim == 2 || (z > 0 && z < W - 1));
}
__device__ __forceinline__ int64_t item_target(const PartitionItems& I, int64_t i) {
    if (I.lex) return i;
    const int64_t p = I.vpos[i], W = I.W;
    return ((int64_t)pz(p) * W + py(p)) * W + px(p);
}

// exclusive scan of 256 per-thread values in LDS; returns the thread's offset
// and leaves the block total in *total
__device__ int block_exclusive(int v, int* lds, int* total) {
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const int add = t >= d ? lds[t - d] : 0;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    const int incl = lds[t];
    *total = lds[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_partition_count(PartitionItems I, int32_t* bsum) {
    __shared__ int lds[kScanThreads];
    const int64_t base = blockIdx.x * (int64_t)kScanItems + threadIdx.x * (int64_t)kScanPer;
    int c = 0;
    for (int k = 0; k < kScanPer; ++k)
        if (base + k < I.n && item_flag(I, base + k)) ++c;
    int total;
    block_exclusive(c, lds, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the block counts in place; total -> *out
__global__ __launch_bounds__(1024) void k_partition_blocks(int64_t nb, int32_t* bsum, int64_t* out) {
    __shared__ int lds[1024];
    __shared__ int64_t carry;
    const int t = threadIdx.x;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < nb; b0 += 1024) {
        const int v = b0 + t < nb ? bsum[b0 + t] : 0;
        lds[t] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int add = t >= d ? lds[t - d] : 0;
            __syncthreads();
            lds[t] += add;
            __syncthreads();
        }
        if (b0 + t < nb) bsum[b0 + t] = (int32_t)(carry + lds[t] - v);
        __syncthreads();
        if (t == 0) carry += lds[1023];
        __syncthreads();
    }
    if (t == 0) *out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_partition_assign(PartitionItems I, const int32_t* bsum,
                                                                   const int64_t* n_inner, int32_t* dof_of_lat) {
    __shared__ int lds[kScanThreads];
    const int64_t base = blockIdx.x * (int64_t)kScanItems + threadIdx.x * (int64_t)kScanPer;
    int c = 0;
    for (int k = 0; k < kScanPer; ++k)
        if (base + k < I.n && item_flag(I, base + k)) ++c;
    int total;
    int64_t rank = bsum[blockIdx.x] + block_exclusive(c, lds, &total);
    const int64_t ni = *n_inner;
    for (int k = 0; k < kScanPer; ++k) {
        const int64_t i = base + k;
        if (i >= I.n) break;
        const bool inner = item_flag(I, i);
        const int64_t dof = inner ? rank : ni + (i - rank);
        if (inner) ++rank;
        dof_of_lat[item_target(I, i)] = (int32_t)dof;
    }
}

__global__ __launch_bounds__(256) void k_check_cover(int64_t n, const int32_t* dof_of_lat, int* err) {
    const int64_t i = gid();
    if (i < n && dof_of_lat[i] < 0) atomicOr(err, kErrCover);
}

// ---- tetrahedra per lattice cell -----------------------------------------
__device__ __forceinline__ uint16_t tet_code(const MeshDev& F, int64_t t, int64_t* cell, bool* ok) {
    int x[4], y[4], z[4];
    for (int j = 0; j < 4; ++j) {
        const int64_t p = F.vpos[F.t_v[4 * t + j]];
        x[j] = px(p);
        y[j] = py(p);
        z[j] = pz(p);
    }
    const int mx = min(min(x[0], x[1]), min(x[2], x[3]));
    const int my = min(min(y[0], y[1]), min(y[2], y[3]));
    const int mz = min(min(z[0], z[1]), min(z[2], z[3]));
    uint16_t code = 0;
    *ok = true;
    for (int j = 0; j < 4; ++j) {
        const int cx = x[j] - mx, cy = y[j] - my, cz = z[j] - mz;
        if (cx > 1 || cy > 1 || cz > 1) *ok = false;
        code |= (uint16_t)(((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2)) << (3 * j));
    }
    *ok = *ok && mx < F.n && my < F.n && mz < F.n;
    *cell = ((int64_t)mz * F.n + my) * F.n + mx;
    return code;
}

__global__ __launch_bounds__(256) void k_cell_slots(MeshDev F, int32_t* fill, int32_t* ids, int* err) {
    const int64_t t = gid();
    if (t >= F.nt) return;
    int64_t cell;
    bool ok;
    tet_code(F, t, &cell, &ok);
    if (!ok) {
        atomicOr(err, kErrTetCell);
        return;
    }
    const int s = atomicAdd(fill + cell, 1);
    if (s < 6)
        ids[cell * 6 + s] = (int32_t)t;
    else
        atomicOr(err, kErrCellCount);
}

// the tetrahedra of a cell in creation order
__global__ __launch_bounds__(256) void k_cell_sort(MeshDev F, const int32_t* fill, const int32_t* ids,
                                                   uint16_t* cell_tets, int* err) {
    const int64_t cell = gid();
    const int64_t cells = (int64_t)F.n * F.n * F.n;
    if (cell >= cells) return;
    if (fill[cell] != 6) {
        atomicOr(err, kErrCellCount);
        return;
    }
    int32_t t[6];
    for (int k = 0; k < 6; ++k) t[k] = ids[cell * 6 + k];
    for (int a = 1; a < 6; ++a)
        for (int b = a; b > 0 && t[b - 1] > t[b]; --b) {
            const int32_t x = t[b];
            t[b] = t[b - 1];
            t[b - 1] = x;
        }
    for (int k = 0; k < 6; ++k) {
        int64_t c;
        bool ok;
        cell_tets[cell * 6 + k] = tet_code(F, t[k], &c, &ok);
    }
}

// direction code of a lattice edge vector (edge_direction_code): 1..13,
// equal for d and -d
__device__ __forceinline__ int direction_code(int dx, int dy, int dz) {
    int c = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
    if (c < 13) c = 26 - c;
    return c - 13;
}

// parent-edge direction code of each fine lattice point (0: copy of a coarse
// vertex); the fine level has only copies and edge midpoints (no quadrilaterals
// or hexahedra in the tetrahedral hierarchy)
__global__ __launch_bounds__(256) void k_parent_codes(MeshDev C, MeshDev F, uint8_t* pcode, int* err) {
    const int64_t e = gid();
    if (e >= C.ne) return;
    const int64_t a = C.vpos[C.e_v[2 * e]], b = C.vpos[C.e_v[2 * e + 1]];
    const int dx = px(b) - px(a), dy = py(b) - py(a), dz = pz(b) - pz(a);
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1 || (dx == 0 && dy == 0 && dz == 0)) {
        atomicOr(err, kErrParentEdge);
        return;
    }
    const int64_t p = F.vpos[C.nv + e], W = (int64_t)F.n + 1;
    pcode[((int64_t)pz(p) * W + py(p)) * W + px(p)] = (uint8_t)direction_code(dx, dy, dz);
}

// ---- host side -------------------------------------------------------------
unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

#define UGX_MESH_HIP(cmd)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (cmd);                                                                              \
        if (e_ != hipSuccess) throw UGError(std::string("device hierarchy: HIP error ") + hipGetErrorString(e_)); \
    } while (0)
#define UGX_MESH_LAUNCH(...)                 \
    do {                                     \
        hipLaunchKernelGGL(__VA_ARGS__);     \
        UGX_MESH_HIP(hipGetLastError());     \
    } while (0)

// owner of one level's entity lists
struct MeshLevel {
    MeshDev d;
    bool tets = false;
    DeviceBuffer<int64_t> vpos;
    DeviceBuffer<uint8_t> vsub, e_sub, f_sub, h_sub, t_sub;
    DeviceBuffer<int32_t> e_v, f_v, f_e, h_v, h_f, h_e, t_v, t_e;

    // allocate the lists for the counts in d; vertices_only keeps the edges,
    // faces and hexahedra empty (tetrahedra are kept for the per-cell lists)
    void allocate(bool vertices_only) {
        vpos.resize(std::max<int64_t>(1, d.nv));
        vsub.resize(std::max<int64_t>(1, d.nv));
        d.vpos = vpos.get();
        d.vsub = vsub.get();
        if (!vertices_only) {
            e_v.resize(std::max<int64_t>(1, 2 * d.ne));
            e_sub.resize(std::max<int64_t>(1, d.ne));
            f_v.resize(std::max<int64_t>(1, 4 * d.nf));
            f_e.resize(std::max<int64_t>(1, 4 * d.nf));
            f_sub.resize(std::max<int64_t>(1, d.nf));
            d.e_v = e_v.get();
            d.e_sub = e_sub.get();
            d.f_v = f_v.get();
            d.f_e = f_e.get();
            d.f_sub = f_sub.get();
            if (d.nh) {
                h_v.resize(8 * d.nh);
                h_f.resize(6 * d.nh);
                h_e.resize(12 * d.nh);
                h_sub.resize(d.nh);
                d.h_v = h_v.get();
                d.h_f = h_f.get();
                d.h_e = h_e.get();
                d.h_sub = h_sub.get();
            }
        }
        if (d.nt) {
            t_v.resize(4 * d.nt);
            t_sub.resize(d.nt);
            d.t_v = t_v.get();
            d.t_sub = t_sub.get();
            if (!vertices_only) {
                t_e.resize(6 * d.nt);
                d.t_e = t_e.get();
            }
        }
    }
};

MeshLevel upload_base(const BaseMesh& B) {
    MeshLevel L;
    L.tets = B.tets;
    L.d.n = B.n;
    L.d.dim = B.dim;
    L.d.fnv = B.face_nv;
    L.d.nv = (int64_t)B.vsub.size();
    L.d.ne = (int64_t)B.e_sub.size();
    L.d.nf = (int64_t)B.f_sub.size();
    L.d.nh = (int64_t)B.h_sub.size();
    L.d.nt = (int64_t)B.t_sub.size();
    L.allocate(false);
    std::vector<int64_t> pos(L.d.nv);
    for (int64_t v = 0; v < L.d.nv; ++v) pos[v] = pk(B.vpos[3 * v], B.vpos[3 * v + 1], B.vpos[3 * v + 2]);
    L.vpos.upload(pos.data(), L.d.nv);
    L.vsub.upload(B.vsub.data(), L.d.nv);
    L.e_v.upload(B.e_v.data(), 2 * L.d.ne);
    L.e_sub.upload(B.e_sub.data(), L.d.ne);
    L.f_v.upload(B.f_v.data(), 4 * L.d.nf);
    L.f_e.upload(B.f_e.data(), 4 * L.d.nf);
    L.f_sub.upload(B.f_sub.data(), L.d.nf);
    if (L.d.nh) {
        L.h_v.upload(B.h_v.data(), 8 * L.d.nh);
        L.h_f.upload(B.h_f.data(), 6 * L.d.nh);
        L.h_e.upload(B.h_e.data(), 12 * L.d.nh);
        L.h_sub.upload(B.h_sub.data(), L.d.nh);
    }
    if (L.d.nt) {
        L.t_v.upload(B.t_v.data(), 4 * L.d.nt);
        L.t_e.upload(B.t_e.data(), 6 * L.d.nt);
        L.t_sub.upload(B.t_sub.data(), L.d.nt);
    }
    return L;
}

MeshLevel refine_level(const MeshLevel& C, bool vertices_only, const RefinePattern& P, int* err, hipStream_t s) {
    const MeshDev& c = C.d;
    MeshLevel F;
    F.tets = C.tets;
    MeshDev& f = F.d;
    f.n = 2 * c.n;
    f.dim = c.dim;
    f.fnv = c.fnv;
    f.nv = c.nv + c.ne + (c.fnv == 4 ? c.nf : 0) + c.nh;
    f.ne = 2 * c.ne + (int64_t)c.fnv * c.nf + 6 * c.nh + c.nt;
    f.nf = 4 * c.nf + 12 * c.nh + 8 * c.nt;
    f.nh = 8 * c.nh;
    f.nt = 8 * c.nt;
    if (f.n >= (1 << 20) || f.ne >= INT32_MAX || f.nf >= INT32_MAX || f.nv >= INT32_MAX)
        throw UGError("device hierarchy: level too large for 32-bit entity indices");
    // the top level needs its vertices (and tetrahedra, for the per-cell lists)
    if (vertices_only) f.ne = f.nf = f.nh = 0;
    F.allocate(vertices_only);
    UGX_MESH_LAUNCH(k_refine_vertices, dim3(blocks_for(f.nv)), dim3(256), 0, s, c, f);
    const int64_t nvol = std::max(f.nh, f.nt);
    if (nvol) UGX_MESH_LAUNCH(k_refine_volumes, dim3(blocks_for(nvol)), dim3(256), 0, s, c, f, P);
    if (vertices_only) return F;
    UGX_MESH_LAUNCH(k_refine_edges, dim3(blocks_for(f.ne)), dim3(256), 0, s, c, f, P);
    UGX_MESH_LAUNCH(k_refine_faces, dim3(blocks_for(f.nf)), dim3(256), 0, s, c, f, P);
    const int64_t W2 = 2 * (int64_t)f.n + 1;
    const int64_t nkey = f.dim == 3 ? W2 * W2 * W2 : W2 * W2;
    DeviceBuffer<int32_t> key(nkey);
    UGX_MESH_LAUNCH(k_fill_i32, dim3(blocks_for(nkey)), dim3(256), 0, s, nkey, key.get(), -1);
    const int64_t nside = std::max(f.ne, f.nf);
    UGX_MESH_LAUNCH(k_key_sides, dim3(blocks_for(nside)), dim3(256), 0, s, f, key.get());
    const int64_t nres = std::max(std::max(f.nf, f.nh), f.nt);
    UGX_MESH_LAUNCH(k_resolve_sides, dim3(blocks_for(nres)), dim3(256), 0, s, f, key.get(), P, err);
    // the key array is released when this returns; keep the stream ordered
    UGX_MESH_HIP(hipStreamSynchronize(s));
    return F;
}

// DoF numbering of one level into L (dof_of_lat, N); the Inner count is
// written to *n_inner_dev
void number_level(const MeshLevel& M, int ordering, LevelGrid& L, int64_t* n_inner_dev, int* err, hipStream_t s) {
    PartitionItems I;
    I.lex = ordering == 1;
    I.dim = M.d.dim;
    I.W = L.W;
    I.n = I.lex ? L.N : M.d.nv;
    I.vpos = M.d.vpos;
    I.vsub = M.d.vsub;
    const int64_t nb = (I.n + kScanItems - 1) / kScanItems;
    DeviceBuffer<int32_t> bsum(std::max<int64_t>(1, nb));
    L.dof_of_lat.resize(L.N);
    UGX_MESH_LAUNCH(k_fill_i32, dim3(blocks_for(L.N)), dim3(256), 0, s, L.N, L.dof_of_lat.get(), -1);
    UGX_MESH_LAUNCH(k_partition_count, dim3((unsigned)nb), dim3(kScanThreads), 0, s, I, bsum.get());
    UGX_MESH_LAUNCH(k_partition_blocks, dim3(1), dim3(1024), 0, s, nb, bsum.get(), n_inner_dev);
    UGX_MESH_LAUNCH(k_partition_assign, dim3((unsigned)nb), dim3(kScanThreads), 0, s, I, bsum.get(), n_inner_dev,
                    L.dof_of_lat.get());
    UGX_MESH_LAUNCH(k_check_cover, dim3(blocks_for(L.N)), dim3(256), 0, s, L.N, L.dof_of_lat.get(), err);
    UGX_MESH_HIP(hipStreamSynchronize(s));   // bsum is released on return
}

void tet_cells(const MeshLevel& M, LevelGrid& L, int* err, hipStream_t s) {
    const int64_t cells = (int64_t)M.d.n * M.d.n * M.d.n;
    DeviceBuffer<int32_t> fill(cells), ids(cells * 6);
    UGX_MESH_LAUNCH(k_fill_i32, dim3(blocks_for(cells)), dim3(256), 0, s, cells, fill.get(), 0);
    UGX_MESH_LAUNCH(k_cell_slots, dim3(blocks_for(M.d.nt)), dim3(256), 0, s, M.d, fill.get(), ids.get(), err);
    L.cell_tets.resize(cells * 6);
    UGX_MESH_LAUNCH(k_cell_sort, dim3(blocks_for(cells)), dim3(256), 0, s, M.d, fill.get(), ids.get(),
                    L.cell_tets.get(), err);
    UGX_MESH_HIP(hipStreamSynchronize(s));
}

}  // namespace

void build_device_levels(int kind, int num_refs, int ordering, std::vector<std::unique_ptr<LevelGrid>>& levels) {
    static const RefinePattern P = derive_pattern();
    auto& ctx = context();
    hipStream_t s = ctx.stream;
    const BaseMesh B = base_mesh(kind);
    DeviceBuffer<int32_t> err(1);
    DeviceBuffer<int64_t> n_inner(num_refs + 1);
    UGX_MESH_LAUNCH(k_fill_i32, dim3(1), dim3(256), 0, s, (int64_t)1, err.get(), 0);
    MeshLevel cur = upload_base(B);
    std::vector<std::unique_ptr<LevelGrid>> out;
    for (int lev = 0; lev <= num_refs; ++lev) {
        if (lev > 0) {
            MeshLevel nxt = refine_level(cur, lev == num_refs, P, err.get(), s);
            auto L = std::make_unique<LevelGrid>();
            L->dim = nxt.d.dim;
            L->n = nxt.d.n;
            L->W = nxt.d.n + 1;
            L->N = L->dim == 3 ? (int64_t)L->W * L->W * L->W : (int64_t)L->W * L->W;
            if (nxt.d.nv != L->N) throw UGError("device hierarchy: vertex count does not match the lattice");
            if (nxt.tets) {
                L->pcode.resize(L->N);
                L->pcode.zero();
                UGX_MESH_LAUNCH(k_parent_codes, dim3(blocks_for(cur.d.ne)), dim3(256), 0, s, cur.d, nxt.d,
                                L->pcode.get(), err.get());
            }
            UGX_MESH_HIP(hipStreamSynchronize(s));   // the coarse lists are released next
            cur = std::move(nxt);
            out.push_back(std::move(L));
        } else {
            auto L = std::make_unique<LevelGrid>();
            L->dim = cur.d.dim;
            L->n = cur.d.n;
            L->W = cur.d.n + 1;
            L->N = L->dim == 3 ? (int64_t)L->W * L->W * L->W : (int64_t)L->W * L->W;
            if (cur.d.nv != L->N) throw UGError("device hierarchy: vertex count does not match the lattice");
            out.push_back(std::move(L));
        }
        LevelGrid& L = *out.back();
        number_level(cur, ordering, L, n_inner.get() + lev, err.get(), s);
        if (cur.tets) tet_cells(cur, L, err.get(), s);
    }
    int herr = 0;
    std::vector<int64_t> ni(num_refs + 1);
    err.download(&herr, 1);
    n_inner.download(ni.data(), num_refs + 1);
    UGX_MESH_HIP(hipStreamSynchronize(s));
    if (herr) throw UGError("device hierarchy: refinement check failed (error bits " + std::to_string(herr) + ")");
    for (int lev = 0; lev <= num_refs; ++lev) out[lev]->n_inner = ni[lev];
    levels = std::move(out);
}

}  // namespace gpu
}  // namespace ug
