// geometry.hpp -- host-side geometry of a structured block (computed once, then uploaded).
#pragma once
#include <vector>

namespace gdtk {

struct BlockGeometry {
    // Active cells in the reference order (k*njc + j)*nic + i.
    std::vector<double> volume, iLen, jLen, kLen;
    // Cell centroids (pos[0] of the reference's cells): x, y, z planes of ncell values.
    std::vector<double> pos;
    // Per index direction: 10 planes (n.xyz, t1.xyz, t2.xyz, area) over the face array
    // of extents nf[d][0] x nf[d][1] x nf[d][2].
    std::vector<double> face[3];
    int nf[3][3];
};

// Throws std::runtime_error for degenerate cells (the reference throws FlowSolverException).
void compute_block_geometry(int dims, int nic, int njc, int nkc, const double* vertices, BlockGeometry& out);

}  // namespace gdtk
