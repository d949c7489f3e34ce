// Structured multigrid hierarchy with ugcore's LEVEL DoF numbering (host side).
#pragma once
#include <cstdint>
#include <vector>

namespace ugamd {

struct LevelNumbering {
    int dim = 3;
    int n = 0;                              // cells per side
    int64_t n_dof = 0;
    int64_t n_inner = 0;                    // DoFs [0, n_inner) are subset "Inner"
    std::vector<int32_t> dof_of_lattice;    // lattice vertex (x fastest) -> DoF
    // kind 2 (tetrahedra): per lattice cell (x fastest) its 6 tetrahedra in
    // creation order, each as 4 cell-corner indices (bit 0 x, bit 1 y, bit 2 z)
    // in the tetrahedron's local vertex order, packed c0 | c1<<3 | c2<<6 | c3<<9
    std::vector<uint16_t> cell_tets;
    // kind 2, levels >= 1: per lattice vertex 0 = copy of a coarse vertex,
    // otherwise 1..13 = direction code of the coarse edge it is the midpoint
    // of (edge_direction_code)
    std::vector<uint8_t> parent_edge;
};

// direction code of a lattice edge vector d (components in {-1,0,1}, d != 0):
// 1..13, equal for d and -d
inline int edge_direction_code(int dx, int dy, int dz) {
    int c = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
    if (c < 13) c = 26 - c;
    return c - 13;
}

// kind 0: unit cube from one hexahedron; kind 1: unit square from two triangles;
// kind 2: unit cube as 3x3x3 hexahedra, each split into 6 tetrahedra
std::vector<LevelNumbering> build_level_numbering(int kind, int num_refs);

// The coarsest level of a hierarchy as flat entity lists in creation order,
// the start of the device refinement. Vertex positions are lattice
// coordinates (x, y, z per vertex; z = 0 in 2D); faces hold face_nv vertices
// and sides padded to 4 with -1; subset 0 is "Inner", 1 "Boundary".
struct BaseMesh {
    int dim = 3;
    int n = 1;            // cells per side
    int face_nv = 4;      // 4: quadrilaterals, 3: triangles
    bool tets = false;
    std::vector<int32_t> vpos;
    std::vector<uint8_t> vsub;
    std::vector<int32_t> e_v;                  // 2 per edge
    std::vector<uint8_t> e_sub;
    std::vector<int32_t> f_v, f_e;             // 4 per face
    std::vector<uint8_t> f_sub;
    std::vector<int32_t> h_v, h_f, h_e;        // 8, 6, 12 per hexahedron
    std::vector<uint8_t> h_sub;
    std::vector<int32_t> t_v, t_e;             // 4, 6 per tetrahedron
    std::vector<uint8_t> t_sub;
};
BaseMesh base_mesh(int kind);

}  // namespace ugamd
