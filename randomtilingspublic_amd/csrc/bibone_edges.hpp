// bibone_edges.hpp -- the bibone tiling as one occupation bit per lattice
// edge: plane layout, edge bits and the slot tables of every derived cell
// state, shared by the host (start checks, merge) and the edge-form walk.
//
// A bibone tiling is a dimer cover (matching) of triangular-lattice
// vertices.  With vertices at plane-cell coordinates (i, j), the edge byte
// of cell (i, j) holds
//   bit 0  h(i, j) = (i, j) - (i, j+1)
//   bit 1  l(i, j) = (i, j) - (i+1, j)
//   bit 2  r(i, j) = (i, j+1) - (i+1, j)
// and every cell state of TriangleDimerTiler's eight planes is a function
// of these bits: the lozenge cells (TriangleDimerTiler.cpp:738-753 layout)
//   eH(i, j) = r(i-1, j) | l(i-1, j+1) << 1 | r(i, j) << 2 | l(i, j) << 3
//   eL(i, j) = h(i, j) | r(i, j) << 1 | h(i+1, j-1) << 2 | r(i, j-1) << 3
//   eR(i, j) = h(i, j) | l(i, j+1) << 1 | h(i+1, j) << 2 | l(i, j) << 3
// and the 2-bit fields of the triangle and butterfly cells, "which of
// three edges holds the dimer" (value 1, 2 or 3, 0 for none), as
// UpdateTriangleUFromLozenges / UpdateButterflys*FromLozenge
// (triangledimerkernel.cl:97-188) and UpdateTriangles (257-278) build them
// from the lozenge cells.
#pragma once
#include <stdint.h>

namespace rt {

// The eight planes of the bibone walk, M x M cells each (M = N / 2).
enum BiPlane { EH = 0, EL, ER, TU, TD, BH, BL, BR, NPLANE };

struct ESlot {
  int8_t di, dj;  // cell offset
  uint8_t bit;    // kEdgeH, kEdgeL or kEdgeR
};
constexpr uint8_t kEdgeH = 0, kEdgeL = 1, kEdgeR = 2;

// Bits 0..3 of eH, eL, eR.
constexpr ESlot kELoz[3][4] = {
    {{-1, 0, kEdgeR}, {-1, 1, kEdgeL}, {0, 0, kEdgeR}, {0, 0, kEdgeL}},
    {{0, 0, kEdgeH}, {0, 0, kEdgeR}, {1, -1, kEdgeH}, {0, -1, kEdgeR}},
    {{0, 0, kEdgeH}, {0, 1, kEdgeL}, {1, 0, kEdgeH}, {0, 0, kEdgeL}}};

// Fields of tU (up) and tD (down), slots for the values 1, 2, 3.  tU(i, j):
// the up triangles at (i, j), (i, j+1), (i+1, j), [h, l, r] of each.
// tD(i, j): tri_code of tU at (i, j-1), (i, j), (i-1, j), where
// tri_code(a, b) = [h(a+1, b), r(a, b), l(a, b+1)], the sides of the down
// triangle at (a, b).
constexpr ESlot kETri[2][3][3] = {
    {{{0, 0, kEdgeH}, {0, 0, kEdgeL}, {0, 0, kEdgeR}},
     {{0, 1, kEdgeH}, {0, 1, kEdgeL}, {0, 1, kEdgeR}},
     {{1, 0, kEdgeH}, {1, 0, kEdgeL}, {1, 0, kEdgeR}}},
    {{{1, -1, kEdgeH}, {0, -1, kEdgeR}, {0, 0, kEdgeL}},
     {{1, 0, kEdgeH}, {0, 0, kEdgeR}, {0, 1, kEdgeL}},
     {{0, 0, kEdgeH}, {-1, 0, kEdgeR}, {-1, 1, kEdgeL}}}};

// Wings f0..f3 of bH, bL, bR (slots for 1, 2, 3) and, in row 4, the flag
// bit (bit 8 of the cell; all three slots name the same edge).
constexpr ESlot kEBf[3][5][3] = {
    {{{-1, 0, kEdgeL}, {-1, 0, kEdgeH}, {-1, 0, kEdgeR}},
     {{-1, 1, kEdgeH}, {-1, 1, kEdgeR}, {-1, 1, kEdgeL}},
     {{0, 1, kEdgeL}, {1, 0, kEdgeH}, {0, 0, kEdgeR}},
     {{1, -1, kEdgeH}, {0, -1, kEdgeR}, {0, 0, kEdgeL}},
     {{0, 0, kEdgeH}, {0, 0, kEdgeH}, {0, 0, kEdgeH}}},
    {{{-1, 0, kEdgeR}, {-1, 1, kEdgeL}, {0, 0, kEdgeH}},
     {{0, 1, kEdgeL}, {1, 0, kEdgeH}, {0, 0, kEdgeR}},
     {{1, -1, kEdgeR}, {1, -1, kEdgeL}, {1, -1, kEdgeH}},
     {{0, -1, kEdgeL}, {0, -1, kEdgeH}, {0, -1, kEdgeR}},
     {{0, 0, kEdgeL}, {0, 0, kEdgeL}, {0, 0, kEdgeL}}},
    {{{-1, 0, kEdgeR}, {-1, 1, kEdgeL}, {0, 0, kEdgeH}},
     {{0, 1, kEdgeH}, {0, 1, kEdgeR}, {0, 1, kEdgeL}},
     {{1, 0, kEdgeR}, {1, 0, kEdgeL}, {1, 0, kEdgeH}},
     {{1, -1, kEdgeH}, {0, -1, kEdgeR}, {0, 0, kEdgeL}},
     {{0, 0, kEdgeR}, {0, 0, kEdgeR}, {0, 0, kEdgeR}}}};

// Slots a rotation exchanges, per field: 45 <-> 54 moves the triangle
// fields 1 <-> 2, 3 <-> 1, 2 <-> 3 (so both edges of each pair toggle);
// 85 <-> 170 moves every wing 1 <-> 2; 5 <-> 10 toggles all four lozenge
// bits.
constexpr int kETriTog[3][2] = {{0, 1}, {2, 0}, {1, 2}};

}  // namespace rt
