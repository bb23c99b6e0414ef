// Neighbour-list kernels.
//
// The list reproduces AtomNN's NeighborList::build (struc/neighbor.cpp:59-171) entry for
// entry: for centre i, neighbours j in ascending order; for each j the periodic images
// r_ij - R_n in the order of the image table (a-major over the norm-sorted lattice vectors,
// neighbor.cpp:87-95) relative to the minimum-image base Cell::diff(r_i, r_j)
// (struc/cell.cpp:103-128); inclusive cutoff d^2 <= rc^2; self images only when
// d^2 > 1e-6; the strictly closest image of each j != i (first on ties, and only below
// ||R||_F) carries the min flag (neighbor.cpp:131-141).  Open structures have no images
// and no min flags (neighbor.cpp:149-170).
//
// Two searches produce the same list:
//  * all-pairs: one wavefront per centre sweeps j in chunks of 64 lanes; each lane tests
//    the images of its j, a wave prefix sum places the kept entries in order.  Used for
//    small cells (any image shell > 0) and small structures.
//  * cell list: for structures whose image shells are all 0 (every cell height > 2 rc) the
//    entry of j is exactly diff(r_i, r_j) when within rc; candidates come from the 27
//    neighbouring fractional bins as one flat list; each wavefront marks its kept candidates
//    in an LDS bitmap over the structure's atoms and reads them out in index order (very
//    large structures: a bitonic sort of the kept candidates instead), so the order is the
//    reference's.
// The partner of an entry (the mirrored pair that carries the Newton reaction) is found
// by the force-assembly kernel.
#include "device.hpp"

namespace atomnn {

// math/special.hpp:46-47 with z = 1: fmod(x, 1) = x - trunc(x) exactly for finite x.
__device__ __forceinline__ double fmod1(double x) { return x - trunc(x); }
__device__ __forceinline__ double wrap_half(double x) { return fmod1(fmod1(x + 0.5) + 1.0) - 0.5; }

// Cell::diff (cell.cpp:103-128)
__device__ __forceinline__ void cell_diff(const DevStruc& s, double dx, double dy, double dz, double& ox, double& oy,
                                          double& oz) {
    double f0 = s.RInv[0] * dx + s.RInv[1] * dy + s.RInv[2] * dz;
    double f1 = s.RInv[3] * dx + s.RInv[4] * dy + s.RInv[5] * dz;
    double f2 = s.RInv[6] * dx + s.RInv[7] * dy + s.RInv[8] * dz;
    f0 = wrap_half(f0);
    f1 = wrap_half(f1);
    f2 = wrap_half(f2);
    ox = s.R[0] * f0 + s.R[1] * f1 + s.R[2] * f2;
    oy = s.R[3] * f0 + s.R[4] * f1 + s.R[5] * f2;
    oz = s.R[6] * f0 + s.R[7] * f1 + s.R[8] * f2;
    const double d2 = ox * ox + oy * oy + oz * oz;
    if (d2 > s.dMax2) {
        double bx = ox, by = oy, bz = oz, best = d2;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int c = q >> 1;
            const double sg = (q & 1) ? -1.0 : 1.0;
            const double tx = ox + sg * s.R[c], ty = oy + sg * s.R[3 + c], tz = oz + sg * s.R[6 + c];
            const double t2 = tx * tx + ty * ty + tz * tz;
            if (t2 < best) {
                best = t2;
                bx = tx;
                by = ty;
                bz = tz;
            }
        }
        ox = bx;
        oy = by;
        oz = bz;
    }
}

// image vector R_n = a Rv0 + b Rv1 + c Rv2 for linear index n (neighbor.cpp:87-95)
__device__ __forceinline__ void image_vec(const DevStruc& s, int n, double& x, double& y, double& z) {
    const int w2 = 2 * s.shell[2] + 1, w1 = 2 * s.shell[1] + 1;
    const int c = n % w2 - s.shell[2];
    const int b = (n / w2) % w1 - s.shell[1];
    const int a = n / (w2 * w1) - s.shell[0];
    x = (double)a * s.Rv[0] + (double)b * s.Rv[3] + (double)c * s.Rv[6];
    y = (double)a * s.Rv[1] + (double)b * s.Rv[4] + (double)c * s.Rv[7];
    z = (double)a * s.Rv[2] + (double)b * s.Rv[5] + (double)c * s.Rv[8];
}

__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& total) {
    int x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    total = __shfl(x, kWave - 1, kWave);
    return x - v;
}

__device__ __forceinline__ void put_entry(int64_t slot, int j, int img, int typ, bool mn, double x, double y, double z,
                                          double d, int* nbr_j, int* nbr_meta, double4* nbr_r) {
    nbr_j[slot] = j;
    nbr_meta[slot] = (img << 8) | (typ << 1) | (mn ? 1 : 0);
    nbr_r[slot] = make_double4(x, y, z, d);
}

// The last of the nb (<= 32) neighbouring bins whose first flat candidate is <= c: a binary
// search over the lanes' bin starts in five shuffles, executed by every lane.
__device__ __forceinline__ int owner_bin(int bfirst, int nb, int c) {
    int lo = 0;
#pragma unroll
    for (int step = 16; step > 0; step >>= 1) {
        const int mid = lo + step;
        const int fq = __shfl(bfirst, mid < nb ? mid : nb - 1, kWave);
        if (mid < nb && fq <= c) lo = mid;
    }
    return lo;
}

// Image index range that can hold r - R_n within rc along each norm-sorted lattice vector: with
// f_q = Gv_q . r, |f_q - n_q| = |(r - R_n) . Gv_q| <= rc |Gv_q|.  The ranges are sub-ranges of the
// image table, walked in its order, so the kept images and their order are unchanged.
struct ImgRange {
    int lo[3], hi[3];
};
__device__ __forceinline__ ImgRange img_range(const DevStruc& s, double bx, double by, double bz) {
    ImgRange g;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double f = s.Gv[3 * q] * bx + s.Gv[3 * q + 1] * by + s.Gv[3 * q + 2] * bz;
        g.lo[q] = max(-s.shell[q], (int)ceil(f - s.reach[q]));
        g.hi[q] = min(s.shell[q], (int)floor(f + s.reach[q]));
    }
    return g;
}

// ------------------------------------------------------------------ all-pairs search
__global__ __launch_bounds__(kWave) void k_neigh_allpairs(const DevStruc* __restrict__ S, const int* __restrict__ atom_struc,
                                                          const double* __restrict__ pos, const int* __restrict__ type,
                                                          const int* __restrict__ centres, double rc2, int maxp,
                                                          int* __restrict__ cnt, int* __restrict__ nbr_j,
                                                          int* __restrict__ nbr_meta, double4* __restrict__ nbr_r,
                                                          int* __restrict__ maxcnt) {
    const int i = centres[blockIdx.x];
    const DevStruc& s = S[atom_struc[i]];
    const int lane = threadIdx.x;
    const int li = i - s.atom0;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int64_t row = (int64_t)i * maxp;
    int written = 0;
    for (int j0 = 0; j0 < s.natoms; j0 += kWave) {
        const int j = j0 + lane;
        const bool valid = j < s.natoms;
        double bx = 0, by = 0, bz = 0;
        int tj = 0;
        int mine = 0;
        double mind = s.fro;
        int minn = -1;
        if (valid) {
            const int gj = s.atom0 + j;
            tj = type[gj];
            const double dx = xi - pos[3 * gj], dy = yi - pos[3 * gj + 1], dz = zi - pos[3 * gj + 2];
            if (s.periodic) {
                cell_diff(s, dx, dy, dz, bx, by, bz);
            } else {
                bx = dx;
                by = dy;
                bz = dz;
            }
        }
        // images of this j that can be within rc (open structures: the single image 0)
        ImgRange g{{0, 0, 0}, {-1, -1, -1}};
        if (valid) g = s.periodic ? img_range(s, bx, by, bz) : ImgRange{{0, 0, 0}, {0, 0, 0}};
        const int w1 = 2 * s.shell[1] + 1, w2 = 2 * s.shell[2] + 1;
        for (int a = g.lo[0]; a <= g.hi[0]; ++a)
            for (int b = g.lo[1]; b <= g.hi[1]; ++b)
                for (int c = g.lo[2]; c <= g.hi[2]; ++c) {
                    const int n = ((a + s.shell[0]) * w1 + (b + s.shell[1])) * w2 + (c + s.shell[2]);
                    double ix = 0, iy = 0, iz = 0;
                    if (s.periodic) image_vec(s, n, ix, iy, iz);
                    const double rx = bx - ix, ry = by - iy, rz = bz - iz;
                    const double d2 = rx * rx + ry * ry + rz * rz;
                    const bool keep = (j != li) ? (d2 <= rc2) : (1e-6 < d2 && d2 <= rc2);
                    if (keep) {
                        ++mine;
                        const double d = sqrt(d2);
                        if (j != li && s.periodic && d < mind) {
                            mind = d;
                            minn = n;
                        }
                    }
                }
        int tot;
        int at = written + wave_excl_scan(mine, lane, tot);
        if (mine > 0) {
            for (int a = g.lo[0]; a <= g.hi[0]; ++a)
                for (int b = g.lo[1]; b <= g.hi[1]; ++b)
                    for (int c = g.lo[2]; c <= g.hi[2]; ++c) {
                        const int n = ((a + s.shell[0]) * w1 + (b + s.shell[1])) * w2 + (c + s.shell[2]);
                        double ix = 0, iy = 0, iz = 0;
                        if (s.periodic) image_vec(s, n, ix, iy, iz);
                        const double rx = bx - ix, ry = by - iy, rz = bz - iz;
                        const double d2 = rx * rx + ry * ry + rz * rz;
                        const bool keep = (j != li) ? (d2 <= rc2) : (1e-6 < d2 && d2 <= rc2);
                        if (keep) {
                            if (at < maxp)
                                put_entry(row + at, j, n, tj, n == minn, rx, ry, rz, sqrt(d2), nbr_j, nbr_meta, nbr_r);
                            ++at;
                        }
                    }
        }
        written += tot;
    }
    if (lane == 0) {
        cnt[i] = written;
    }
}

// ------------------------------------------------------------------ cell list
// Bin of atom i (fractional coordinates wrapped into [0,1)), or -1 outside the cell list.
// The fields of a structure the binning reads (held in registers by the one-structure binning).
struct BinGeom {
    double RInv[9];
    int nbin[3];
    int bin0;
    int mode;
};
__device__ __forceinline__ BinGeom bin_geom(const DevStruc& s) {
    BinGeom g;
#pragma unroll
    for (int q = 0; q < 9; ++q) g.RInv[q] = s.RInv[q];
#pragma unroll
    for (int q = 0; q < 3; ++q) g.nbin[q] = s.nbin[q];
    g.bin0 = s.bin0;
    g.mode = s.mode;
    return g;
}
template <class SG>
__device__ __forceinline__ int bin_xyz(const SG& s, double x, double y, double z) {
    if (s.mode != 1) return -1;
    int b[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double f = s.RInv[3 * r] * x + s.RInv[3 * r + 1] * y + s.RInv[3 * r + 2] * z;
        f -= floor(f);
        int q = (int)(f * s.nbin[r]);
        q = q < 0 ? 0 : (q >= s.nbin[r] ? s.nbin[r] - 1 : q);
        b[r] = q;
    }
    return s.bin0 + (b[0] * s.nbin[1] + b[1]) * s.nbin[2] + b[2];
}
__device__ __forceinline__ int bin_of(const DevStruc* __restrict__ S, const int* __restrict__ atom_struc,
                                      const double* __restrict__ pos, int i) {
    return bin_xyz(S[atom_struc[i]], pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

__global__ void k_bin_atoms(const DevStruc* __restrict__ S, const int* __restrict__ atom_struc,
                            const double* __restrict__ pos, int natoms_total, int* __restrict__ atom_bin,
                            int* __restrict__ bin_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms_total) return;
    const int bin = bin_of(S, atom_struc, pos, i);
    atom_bin[i] = bin;
    if (bin >= 0) atomicAdd(&bin_count[bin], 1);
}

// Small batches: the counter reset, binning, bin scan and scatter of the four kernels below in
// one workgroup with the bin counters in LDS (one launch instead of four at a few thousand atoms).
// ONE: a single structure, whose cell is read once per wavefront (scalar loads) instead of per atom.
constexpr int kBinsSmallAtoms = 32768;
constexpr int kBinsSmallBins = 4096;
template <bool ONE>
__global__ __launch_bounds__(1024) void k_bins_small(const DevStruc* __restrict__ S, const int* __restrict__ atom_struc,
                                                     const double* __restrict__ pos, int ntot, int nbins,
                                                     int* __restrict__ maxcnt, int* __restrict__ work,
                                                     int* __restrict__ atom_bin, int* __restrict__ bin_count,
                                                     int* __restrict__ bin_start, int* __restrict__ bin_atoms) {
    __shared__ int cnt_l[kBinsSmallBins];
    __shared__ int start_l[kBinsSmallBins];
    __shared__ int part[1024 / kWave];
    const int t = threadIdx.x;
    if (t == 0) {
        maxcnt[0] = 0;  // neighbour maximum
        maxcnt[1] = 0;  // reactions missed by the force assembly's gather
        if (work) *work = 0;
    }
    for (int b = t; b < nbins; b += 1024) cnt_l[b] = 0;
    __syncthreads();
    // each thread's atoms i = t + 1024 k keep their bins in registers for the fill below
    int mybin[kBinsSmallAtoms / 1024];
#pragma unroll
    for (int k = 0; k < kBinsSmallAtoms / 1024; ++k) mybin[k] = -1;
    BinGeom g1{};
    if constexpr (ONE) g1 = bin_geom(S[0]);
    // groups of four atoms per thread; the positions of a group are loaded unconditionally at
    // clamped (in-bounds) indices, so the four loads go out together (a guarded load becomes a
    // branch with a full wait after it); the exit test is the same for the whole block
#pragma unroll
    for (int k0 = 0; k0 < kBinsSmallAtoms / 1024; k0 += 4) {
        if (k0 * 1024 >= ntot) break;
        double px[4], py[4], pz[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ic = min(t + 1024 * (k0 + q), ntot - 1);
            px[q] = pos[3 * ic];
            py[q] = pos[3 * ic + 1];
            pz[q] = pos[3 * ic + 2];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = t + 1024 * (k0 + q);
            if (i < ntot) {
                const int bin = ONE ? bin_xyz(g1, px[q], py[q], pz[q]) : bin_of(S, atom_struc, pos, i);
                atom_bin[i] = bin;
                mybin[k0 + q] = bin;
                if (bin >= 0) atomicAdd(&cnt_l[bin], 1);
            }
        }
    }
    __syncthreads();
    // exclusive scan: four consecutive bins per thread, then a scan of the 1024 thread sums
    int loc[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int b = 4 * t + k;
        loc[k] = sum;
        sum += b < nbins ? cnt_l[b] : 0;
    }
    // inclusive scan inside each wavefront (shuffles), then the 16 wavefront totals
    const int lane = t & (kWave - 1), wv = t / kWave;
    int inc = sum;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += y;
    }
    if (lane == kWave - 1) part[wv] = inc;
    __syncthreads();
    int excl = inc - sum;
    for (int k = 0; k < wv; ++k) excl += part[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int b = 4 * t + k;
        if (b < nbins) {
            const int st = excl + loc[k];
            bin_start[b] = st;
            bin_count[b] = cnt_l[b];
            start_l[b] = st;
        }
    }
    __syncthreads();
    for (int b = t; b < nbins; b += 1024) cnt_l[b] = 0;  // now the fill counters
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBinsSmallAtoms / 1024; ++k) {
        const int b = mybin[k];
        if (b < 0) continue;
        const int slot = atomicAdd(&cnt_l[b], 1);
        bin_atoms[start_l[b] + slot] = t + 1024 * k;
    }
}

// Scatter atoms into their bins.  Order inside a bin does not matter: candidates are sorted.
__global__ void k_fill_bins(const int* __restrict__ atom_bin, int natoms_total, const int* __restrict__ bin_start,
                            int* __restrict__ bin_fill, int* __restrict__ bin_atoms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms_total) return;
    const int b = atom_bin[i];
    if (b < 0) return;
    const int slot = atomicAdd(&bin_fill[b], 1);
    bin_atoms[bin_start[b] + slot] = i;
}

// One wavefront per centre: gather candidates from the 27 neighbouring bins, keep those with
// |diff(r_i, r_j)| <= rc, sort the kept entries by j in LDS, write them in order.
// LDS: vals double[cap][4] | keys int[cap] | skey int[n2cap] | perm int[n2cap], n2cap = 2^k >= cap.
__global__ __launch_bounds__(kWave) void k_neigh_cells(const DevStruc* __restrict__ S, const int* __restrict__ atom_struc,
                                                       const double* __restrict__ pos, const int* __restrict__ type,
                                                       const int* __restrict__ centres, const int* __restrict__ atom_bin,
                                                       const int* __restrict__ bin_start, const int* __restrict__ bin_count,
                                                       const int* __restrict__ bin_atoms, double rc2, int maxp, int n2cap,
                                                       int* __restrict__ cnt, int* __restrict__ nbr_j,
                                                       int* __restrict__ nbr_meta, double4* __restrict__ nbr_r,
                                                       int* __restrict__ maxcnt) {
    extern __shared__ double lds_cells[];
    const int cap = maxp;
    double* vals = lds_cells;
    int* keys = (int*)(vals + 4 * cap);
    int* skey = keys + cap;
    int* perm = skey + n2cap;
    const int i = centres[blockIdx.x];
    const DevStruc& s = S[atom_struc[i]];
    const int lane = threadIdx.x;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int myb = atom_bin[i] - s.bin0;
    const int bz = myb % s.nbin[2], by = (myb / s.nbin[2]) % s.nbin[1], bx = myb / (s.nbin[2] * s.nbin[1]);
    // neighbouring bins (along an axis with fewer than 3 bins every bin is visited once): lane o
    // owns bin o; a wave prefix sum over their counts gives one flat candidate list
    const int n0 = s.nbin[0] < 3 ? s.nbin[0] : 3, n1 = s.nbin[1] < 3 ? s.nbin[1] : 3, n2 = s.nbin[2] < 3 ? s.nbin[2] : 3;
    const int nb = n0 * n1 * n2;  // <= 27
    int bstart = 0, bcount = 0;
    if (lane < nb) {
        const int o2 = lane % n2, o1 = (lane / n2) % n1, o0 = lane / (n2 * n1);
        const int cx = s.nbin[0] < 3 ? o0 : (bx + o0 - 1 + s.nbin[0]) % s.nbin[0];
        const int cy = s.nbin[1] < 3 ? o1 : (by + o1 - 1 + s.nbin[1]) % s.nbin[1];
        const int cz = s.nbin[2] < 3 ? o2 : (bz + o2 - 1 + s.nbin[2]) % s.nbin[2];
        const int b = s.bin0 + (cx * s.nbin[1] + cy) * s.nbin[2] + cz;
        bstart = bin_start[b];
        bcount = bin_count[b];
    }
    int ncand;
    const int bfirst = wave_excl_scan(bcount, lane, ncand);  // first flat candidate of bin `lane`
    int kept = 0;
    for (int c0 = 0; c0 < ncand; c0 += kWave) {
        const int c = c0 + lane;
        const int ob = owner_bin(bfirst, nb, c);
        const int ostart = __shfl(bstart, ob, kWave), ofirst = __shfl(bfirst, ob, kWave);
        bool keep = false;
        double rx = 0, ry = 0, rz = 0, d2 = 0;
        int gj = 0;
        if (c < ncand) {
            gj = bin_atoms[ostart + (c - ofirst)];
            if (gj != i) {
                cell_diff(s, xi - pos[3 * gj], yi - pos[3 * gj + 1], zi - pos[3 * gj + 2], rx, ry, rz);
                d2 = rx * rx + ry * ry + rz * rz;
                keep = d2 <= rc2;
            }
        }
        int tot;
        const int at = kept + wave_excl_scan(keep ? 1 : 0, lane, tot);
        if (keep && at < cap) {
            keys[at] = gj - s.atom0;
            vals[4 * at] = rx;
            vals[4 * at + 1] = ry;
            vals[4 * at + 2] = rz;
            vals[4 * at + 3] = d2;
        }
        kept += tot;
    }
    __syncthreads();
    const int P = kept < cap ? kept : cap;
    int n2p = 1;
    while (n2p < P) n2p <<= 1;
    for (int q = lane; q < n2p; q += kWave) {
        skey[q] = q < P ? keys[q] : 0x7fffffff;
        perm[q] = q;
    }
    __syncthreads();
    // bitonic sort of (key, slot) by key; keys are distinct atoms
    for (int k = 2; k <= n2p; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int q = lane; q < n2p; q += kWave) {
                const int p = q ^ jj;
                if (p > q) {
                    const bool up = (q & k) == 0;
                    const int a = skey[q], b = skey[p];
                    if ((a > b) == up) {
                        skey[q] = b;
                        skey[p] = a;
                        const int t = perm[q];
                        perm[q] = perm[p];
                        perm[p] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int64_t row = (int64_t)i * maxp;
    for (int q = lane; q < P; q += kWave) {
        const int src = perm[q];
        const int j = keys[src];
        const double d = sqrt(vals[4 * src + 3]);
        // one image per j: it is the minimum image when below ||R||_F (neighbor.cpp:132-140)
        put_entry(row + q, j, 0, type[s.atom0 + j], d < s.fro, vals[4 * src], vals[4 * src + 1], vals[4 * src + 2], d,
                  nbr_j, nbr_meta, nbr_r);
    }
    if (lane == 0) {
        cnt[i] = kept;
    }
}

// Cell list without a sort: kept neighbours are marked in an LDS bitmap over the structure's
// atoms (one bit each, set with an LDS atomic OR, so the result does not depend on the order
// the candidates are visited in), then read out in ascending index order, which is the
// reference's order; each kept entry's displacement is recomputed by the same Cell::diff.
// LDS: one 32-bit word per 32 atoms of the largest cell-list structure of the batch.
__global__ __launch_bounds__(kWave) void k_neigh_cells_bits(
    const DevStruc* __restrict__ S, const int* __restrict__ atom_struc, const double* __restrict__ pos,
    const int* __restrict__ type, int ncell, const int* __restrict__ atom_bin,
    const int* __restrict__ bin_start, const int* __restrict__ bin_count, const int* __restrict__ bin_atoms,
    double rc2, int maxp, int* __restrict__ cnt, int* __restrict__ nbr_j, int* __restrict__ nbr_meta,
    double4* __restrict__ nbr_r) {
    extern __shared__ unsigned int lds_bits[];
    // centres in the XCD-local order of the cell-list atoms in bin order: the blocks of one XCD
    // cover one contiguous, spatially compact run of centres, whose candidates' positions and bin
    // lists they then share in that XCD's L2 (the centre a block computes does not change any row)
    const int slot = xcd_slot(ncell);
    if (slot < 0) return;
    const int i = bin_atoms[slot];
    const DevStruc& s = S[atom_struc[i]];
    const int lane = threadIdx.x;
    const int nwords = (s.natoms + 31) >> 5;
    for (int w = lane; w < nwords; w += kWave) lds_bits[w] = 0u;
    __syncthreads();
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int myb = atom_bin[i] - s.bin0;
    const int bz = myb % s.nbin[2], by = (myb / s.nbin[2]) % s.nbin[1], bx = myb / (s.nbin[2] * s.nbin[1]);
    const int n0 = s.nbin[0] < 3 ? s.nbin[0] : 3, n1 = s.nbin[1] < 3 ? s.nbin[1] : 3, n2 = s.nbin[2] < 3 ? s.nbin[2] : 3;
    const int nb = n0 * n1 * n2;  // <= 27
    int bstart = 0, bcount = 0;
    if (lane < nb) {
        const int o2 = lane % n2, o1 = (lane / n2) % n1, o0 = lane / (n2 * n1);
        const int cx = s.nbin[0] < 3 ? o0 : (bx + o0 - 1 + s.nbin[0]) % s.nbin[0];
        const int cy = s.nbin[1] < 3 ? o1 : (by + o1 - 1 + s.nbin[1]) % s.nbin[1];
        const int cz = s.nbin[2] < 3 ? o2 : (bz + o2 - 1 + s.nbin[2]) % s.nbin[2];
        const int b = s.bin0 + (cx * s.nbin[1] + cy) * s.nbin[2] + cz;
        bstart = bin_start[b];
        bcount = bin_count[b];
    }
    int ncand;
    const int bfirst = wave_excl_scan(bcount, lane, ncand);
    for (int c0 = 0; c0 < ncand; c0 += kWave) {
        const int c = c0 + lane;
        const int ob = owner_bin(bfirst, nb, c);
        const int ostart = __shfl(bstart, ob, kWave), ofirst = __shfl(bfirst, ob, kWave);
        if (c < ncand) {
            const int gj = bin_atoms[ostart + (c - ofirst)];
            if (gj != i) {
                double rx, ry, rz;
                cell_diff(s, xi - pos[3 * gj], yi - pos[3 * gj + 1], zi - pos[3 * gj + 2], rx, ry, rz);
                if (rx * rx + ry * ry + rz * rz <= rc2) {
                    const int lj = gj - s.atom0;
                    atomicOr(&lds_bits[lj >> 5], 1u << (lj & 31));
                }
            }
        }
    }
    __syncthreads();
    // lanes own contiguous word ranges; a wave prefix sum of their popcounts gives the slots
    const int per = (nwords + kWave - 1) / kWave;
    const int w0 = min(lane * per, nwords), w1 = min(w0 + per, nwords);
    int mine = 0;
    for (int w = w0; w < w1; ++w) mine += __popc(lds_bits[w]);
    int kept;
    int at = wave_excl_scan(mine, lane, kept);
    const int64_t row = (int64_t)i * maxp;
    for (int w = w0; w < w1; ++w) {
        unsigned int m = lds_bits[w];
        while (m) {
            const int bit = __ffs(m) - 1;
            m &= m - 1;
            if (at < maxp) {
                const int lj = (w << 5) + bit, gj = s.atom0 + lj;
                double rx, ry, rz;
                cell_diff(s, xi - pos[3 * gj], yi - pos[3 * gj + 1], zi - pos[3 * gj + 2], rx, ry, rz);
                const double d = sqrt(rx * rx + ry * ry + rz * rz);
                // one image per j: it is the minimum image when below ||R||_F (neighbor.cpp:132-140)
                put_entry(row + at, lj, 0, type[gj], d < s.fro, rx, ry, rz, d, nbr_j, nbr_meta, nbr_r);
            }
            ++at;
        }
    }
    if (lane == 0) cnt[i] = kept;
}

// Largest neighbour count of the pass (the row-overflow check).  A few hundred blocks each
// fold a strided slice and issue one atomicMax: the per-centre kernels do not contend for
// one address.
__global__ __launch_bounds__(256) void k_count_max(const int* __restrict__ cnt, int n, int* __restrict__ maxcnt) {
    int m = 0;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) m = max(m, cnt[t]);
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
    __shared__ int wm[256 / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) wm[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / kWave; ++w) m = max(m, wm[w]);
        atomicMax(maxcnt, m);
    }
}

// Clears the per-pass counters (max entry count, missed reactions and the cell-list bin counters)
// in one launch.
__global__ void k_zero(int* __restrict__ maxcnt, int* __restrict__ work, int* __restrict__ bin_count,
                       int* __restrict__ bin_fill, int nbins) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        maxcnt[0] = 0;  // neighbour maximum
        maxcnt[1] = 0;  // reactions missed by the force assembly's gather
        if (work) *work = 0;
    }
    if (t < nbins) {
        bin_count[t] = 0;
        bin_fill[t] = 0;
    }
}

// Exclusive scan of n ints by one block (bin starts of the cell list).
__global__ __launch_bounds__(1024) void k_scan_excl(const int* __restrict__ in, int* __restrict__ out, int n) {
    // blocks of 1024: inclusive scans inside each wavefront (shuffles), then the 16 wavefront
    // totals, then the running carry of the blocks before
    __shared__ int part[1024 / kWave];
    const int t = threadIdx.x, lane = t & (kWave - 1), wv = t / kWave;
    int carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int v = base + t < n ? in[base + t] : 0;
        int inc = v;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(inc, o, kWave);
            if (lane >= o) inc += y;
        }
        if (lane == kWave - 1) part[wv] = inc;
        __syncthreads();
        int off = carry;
        for (int k = 0; k < wv; ++k) off += part[k];
        if (base + t < n) out[base + t] = off + inc - v;
        for (int k = 0; k < 1024 / kWave; ++k) carry += part[k];
        __syncthreads();  // part is rewritten by the next block
    }
}

}  // namespace atomnn

#include "engine.hpp"

namespace atomnn {

hipError_t launch_count_max(const int* cnt, int n, int* maxcnt, hipStream_t st);

hipError_t launch_neighbors(const NeighborArgs& a, hipStream_t st) {
    hipError_t e;
    const bool small = a.ncell > 0 && a.ntot <= kBinsSmallAtoms && a.nbins <= kBinsSmallBins;
    if (small) {
        if (a.nstruc == 1)
            hipLaunchKernelGGL(k_bins_small<true>, dim3(1), dim3(1024), 0, st, a.S, a.atom_struc, a.pos, a.ntot,
                               a.nbins, a.maxcnt, a.work, a.atom_bin, a.bin_count, a.bin_start, a.bin_atoms);
        else
            hipLaunchKernelGGL(k_bins_small<false>, dim3(1), dim3(1024), 0, st, a.S, a.atom_struc, a.pos, a.ntot,
                               a.nbins, a.maxcnt, a.work, a.atom_bin, a.bin_count, a.bin_start, a.bin_atoms);
    } else {
        const int nz = a.ncell > 0 ? a.nbins : 0;
        hipLaunchKernelGGL(k_zero, dim3((nz + 255) / 256 + 1), dim3(256), 0, st, a.maxcnt, a.work, a.bin_count,
                           a.bin_fill, nz);
    }
    if (a.ncell > 0) {
        if (!small) {
            hipLaunchKernelGGL(k_bin_atoms, dim3((a.ntot + 255) / 256), dim3(256), 0, st, a.S, a.atom_struc, a.pos,
                               a.ntot, a.atom_bin, a.bin_count);
            hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, a.bin_count, a.bin_start, a.nbins);
            hipLaunchKernelGGL(k_fill_bins, dim3((a.ntot + 255) / 256), dim3(256), 0, st, a.atom_bin, a.ntot,
                               a.bin_start, a.bin_fill, a.bin_atoms);
        }
        const size_t bits_lds = (size_t)((a.cell_natoms_max + 31) / 32) * 4;
        if (bits_lds <= 32768) {
            hipLaunchKernelGGL(k_neigh_cells_bits, dim3(xcd_grid(a.ncell)), dim3(kWave), bits_lds, st, a.S,
                               a.atom_struc, a.pos, a.type, a.ncell, a.atom_bin, a.bin_start, a.bin_count,
                               a.bin_atoms, a.rc2, a.maxp, a.cnt, a.nbr_j, a.nbr_meta, a.nbr_r);
        } else {  // very large structures: sort the kept candidates instead
            int n2cap = 1;
            while (n2cap < a.maxp) n2cap <<= 1;
            const size_t lds = (size_t)a.maxp * 32 + (size_t)a.maxp * 4 + (size_t)n2cap * 8;
            hipLaunchKernelGGL(k_neigh_cells, dim3(a.ncell), dim3(kWave), lds, st, a.S, a.atom_struc, a.pos, a.type,
                               a.cell_centres, a.atom_bin, a.bin_start, a.bin_count, a.bin_atoms, a.rc2, a.maxp, n2cap,
                               a.cnt, a.nbr_j, a.nbr_meta, a.nbr_r, a.maxcnt);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (a.npair > 0) {
        hipLaunchKernelGGL(k_neigh_allpairs, dim3(a.npair), dim3(kWave), 0, st, a.S, a.atom_struc, a.pos, a.type,
                           a.pair_centres, a.rc2, a.maxp, a.cnt, a.nbr_j, a.nbr_meta, a.nbr_r, a.maxcnt);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (a.ntot > 0 && a.count_max) return launch_count_max(a.cnt, a.ntot, a.maxcnt, st);
    return hipGetLastError();
}

hipError_t launch_count_max(const int* cnt, int n, int* maxcnt, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_count_max, dim3(std::min((n + 255) / 256, 256)), dim3(256), 0, st, cnt, n, maxcnt);
    return hipGetLastError();
}

}  // namespace atomnn
