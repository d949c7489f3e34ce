// Local index tables of ugcore's reference elements and of the regular
// refinement rules used by the lattice hierarchies; shared by the host
// hierarchy (hierarchy.cpp) and the device refiner (mesh_kernels.hip).
#pragma once

namespace ugamd {
namespace mesh {

// reference hexahedron (lib_grid/grid_objects/hexahedron_rules.h)
constexpr int kHexCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
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

// reference tetrahedron (lib_grid/grid_objects/tetrahedron_rules.h)
constexpr int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
constexpr int kTetFace[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {0, 3, 1}};
// regular rule STANDARD (tetrahedron_rules.cpp:487-609) in the index space
// [corners 0-3 | edge vertices 4-9]: four corner children, then the inner
// octahedron split along diagonal 0-5, 1-3 or 2-4 (edge vertices)
constexpr int kTetCorner[4][4] = {{0, 4, 6, 7}, {1, 5, 4, 8}, {2, 6, 5, 9}, {7, 8, 9, 3}};
constexpr int kTetInner[3][4][4] = {
    {{4, 5, 6, 9}, {5, 8, 9, 4}, {6, 9, 7, 4}, {4, 7, 8, 9}},
    {{4, 5, 6, 7}, {5, 8, 9, 7}, {6, 9, 7, 5}, {4, 7, 8, 5}},
    {{4, 6, 7, 8}, {5, 6, 4, 8}, {6, 7, 8, 9}, {8, 5, 6, 9}}};

// children of a refined face (GlobalMultiGridRefiner face pass): quadrilateral
// in [corners 0-3 | edge vertices 4-7 | centre 8], triangle in [corners 0-2 |
// edge vertices 3-5]; edge j joins corners j and j+1
constexpr int kQuadChild[4][4] = {{0, 4, 8, 7}, {1, 5, 8, 4}, {2, 6, 8, 5}, {3, 7, 8, 6}};
constexpr int kTriChild[4][3] = {{0, 3, 5}, {1, 4, 3}, {2, 5, 4}, {3, 4, 5}};

}  // namespace mesh
}  // namespace ugamd
