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
// XCDs of an MI355X: a launch's blocks are dealt round-robin over them, each with its own L2
constexpr int kXcds = 8;

// XCD-local order of n items (centres): block b, which runs on XCD b mod 8, takes item
// (b mod 8) chunk + b / 8, so each XCD's blocks cover one contiguous run; -1 for the padding
// blocks past n.  Launch xcd_grid(n) blocks.
__device__ __forceinline__ int xcd_slot(int n) {
    const int chunk = (n + kXcds - 1) / kXcds;
    const int pos = (int)(blockIdx.x % kXcds) * chunk + (int)(blockIdx.x / kXcds);
    return pos < n ? pos : -1;
}
inline int xcd_grid(int n) { return kXcds * ((n + kXcds - 1) / kXcds); }

// Force assembly (k_assemble).  react points at two per-pass counters beside the neighbour
// maximum: react[0] = reactions the gather from the neighbours' rows missed (mode 0; cleared by
// the neighbour launch), react[1] = longest receiver row (mode 1; cleared by that mode's own
// launches).
// Launch configuration of the network kernels: wl = FP64 weights copied into LDS (once per block
// and species), grid = blocks, each taking every grid-th 16-atom tile (0: one block per tile).
// Every configuration computes the same results.
struct NetLaunch {
    int wl = 0;
    int grid = 0;
};

struct AssembleArgs {
    const DevStruc* S;
    const int* atom_struc;
    const int* cnt;
    const int* nbr_j;
    const int* meta;
    int maxp;
    int ntot;
    int all_images;
    double* F;
    int* react;
    int mode;      // 0: gather from the neighbours' rows, 1: receiver rows
    int* rcount;   // mode 1: [ntot] entries in each receiver row
    int* rlist;    // mode 1: [ntot][rcap] entry indices
    int rcap;
    // processing order of the atoms: the cell-list atoms in bin order (bin_atoms[0, ncell)), then
    // the other atoms (pair_centres[0, ntot - ncell)); each XCD's blocks take a contiguous chunk of it
    const int* bin_atoms;
    const int* pair_centres;
    int ncell;
};

// End-of-pass reductions, run by single-wavefront blocks appended to the assembly launch (or on
// their own when no forces are computed): the total energy of each structure, and the largest
// neighbour count of the pass (the row-overflow check of resident passes).
struct EnergyTail {
    const double* eatom;  // E_i, batch order
    const int* sstart;    // first atom of each structure, nstruc + 1 entries
    double* E;            // total energy of each structure
    int nstruc;
    int ncount;           // count blocks after the energy blocks (0: no count check)
    const int* cnt;
    int ntot;
    int* maxcnt;
};

// Block b of the tail (one wavefront).  Energy: lane t sums its contiguous chunk of the
// structure's E_i in order, then a fixed shuffle tree; the sum depends only on that structure's
// E_i, so batch and single passes agree bit for bit.  Count: strided slices, one atomicMax each.
__device__ __forceinline__ void energy_tail_block(const EnergyTail& t, int b, int lane) {
    if (b < t.nstruc) {
        const int a0 = t.sstart[b], a1 = t.sstart[b + 1];
        const int n = a1 - a0, chunk = (n + kWave - 1) / kWave;
        const int i0 = a0 + lane * chunk, i1 = a0 + min(n, (lane + 1) * chunk);
        double acc = 0.0;
        for (int i = i0; i < i1; ++i) acc += t.eatom[i];
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) acc += __shfl_down(acc, o);
        if (lane == 0) t.E[b] = acc;
        return;
    }
    const int c = b - t.nstruc;
    int m = 0;
    for (int i = c * kWave + lane; i < t.ntot; i += t.ncount * kWave) m = max(m, t.cnt[i]);
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, __shfl_down(m, o));
    if (lane == 0) atomicMax(t.maxcnt, m);
}

}  // namespace atomnn
