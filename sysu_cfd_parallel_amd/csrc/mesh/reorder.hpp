// Bandwidth-reducing row order of a sparse pattern (reverse Cuthill-McKee).
#pragma once
#include <vector>

namespace s355 {

// Reverse Cuthill-McKee order of the symmetric pattern (rowptr, col) of n
// rows: order[p] is the row placed at position p.  Each connected component
// starts from a pseudo-peripheral row (repeated breadth-first sweeps from the
// lowest-degree unvisited row); neighbours are visited by ascending degree,
// ties by row id, so the order is deterministic.
std::vector<int> rcm_order(int n, const int* rowptr, const int* col);

}  // namespace s355
