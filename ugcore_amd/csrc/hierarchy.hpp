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
};

// kind 0: unit cube from one hexahedron; kind 1: unit square from two triangles
std::vector<LevelNumbering> build_level_numbering(int kind, int num_refs);

}  // namespace ugamd
