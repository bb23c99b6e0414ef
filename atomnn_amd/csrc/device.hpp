// Device-side data layout shared by the HIP kernels and the host orchestration.
//
// HBM layout (one batch = any number of structures packed back to back):
//   pos        double[Ntot][3]      positions, structure after structure
//   type       int32 [Ntot]         species index
//   atom_struc int32 [Ntot]         owning structure of each atom
//   DevStruc   one per structure    cell, image table, neighbour-search mode
//   neighbour list, padded rows of `maxp` entries per centre atom:
//     cnt      int32 [Ntot]         entries of each centre
//     nbr_j    int32 [Ntot*maxp]    neighbour atom (index inside its structure)
//     nbr_meta int32 [Ntot*maxp]    image << 8 | species << 1 | min-image flag
//     nbr_r    double4[Ntot*maxp]   (r_i - r_j - T, |r|): points from neighbour to centre
//     partner  int32 [Ntot*maxp]    entry of the mirrored pair (j -> i), -1 if none applies
//   G          T[Ntot][ldg]          descriptors, dEdG same shape
//   X          TX4[Ntot*maxp]       per-entry force term  -dE_i/dr_ij
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace atomnn {

struct DevStruc {
    int atom0;     // first atom of the structure in the batch arrays
    int natoms;
    int periodic;  // 0: open structure (no images, no min-image flags)
    int nimg;      // (2 s0 + 1)(2 s1 + 1)(2 s2 + 1)
    int shell[3];  // image half-widths along the norm-sorted lattice vectors
    int mode;      // 0: all-pairs search, 1: cell list
    int nbin[3];   // cell-list bins along each fractional axis
    int bin0;      // first bin of this structure in the batch's bin arrays
    double R[9];     // R[r*3+c]; columns are the lattice vectors
    double RInv[9];
    double Rv[9];    // norm-sorted lattice vectors, Rv[q*3+r]
    double Gv[9];    // their dual vectors, Gv[q*3+r] (Gv_q . Rv_p = delta_qp)
    double reach[3]; // rc |Gv_q| (+ a rounding margin): images further off along q are out of range
    double dMax2;    // (half the shortest lattice vector)^2
    double fro;      // Frobenius norm of R
};

constexpr int kWave = 64;

}  // namespace atomnn
