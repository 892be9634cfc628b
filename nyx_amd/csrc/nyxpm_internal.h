// nyxpm_internal.h -- device tables and launch helpers shared by the HIP kernels
// and the host orchestration of the MI355X particle-mesh library.
//
// Layout in HBM
//  * particles: structure of arrays (x, y, z, vx, vy, vz FP64, optional mass
//    FP64, int32 id), boxes stored one after another, each box's particles kept
//    grouped by deposit tile (double-buffered for the full sort).  Each tile owns
//    a slot range with slack; a slot whose x is NaN holds no particle, so kernels
//    that walk a box's slots skip those.
//  * cell fields: x-fastest, no ghost cells; periodic neighbours are found by
//    index wrapping.  All boxes share one pool per field; a box's level-l array
//    starts at a per-(box, level) offset.
//  * multigrid: one pool of residuals r and one of corrections e per precision,
//    every level of every box inside it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nyxpm {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // threads per block of the grid kernels
constexpr int kMaxLevels = 24;
constexpr int kBottomMaxCells = 4096;
constexpr int kReduceBlocks = 8192;  // blocks per box of the reducing level-0 kernels
constexpr int kNormSlots = 64;       // atomicMax slots per box for residual norms

// One periodic box as the particle kernels see it.
struct BoxDev {
    int n[3];
    int tile[3];           // tile shape (cells) used for binning/deposit
    int ntile[3];          // tiles per dimension
    int fx_exp;            // E: deposit fixed-point unit is 2^(E-31) (hi) and 2^(E-63) (lo)
    long long cell_off;    // level-0 offset in the cell pools
    long long tile_off;    // first global tile id
    long long part_off;    // first particle slot of this box
    long long np;          // particles in this box
    long long span;        // particle slots of this box (np plus the tile slack; a
                           // slot with x = NaN holds no particle)
    double plo[3];
    double dxi[3];
    double rhi[3];         // largest coordinate kept by the periodic shift
    double len[3];         // prob_hi - prob_lo
    double inv_vol;        // 1 / (dx*dy*dz)
    double mass_offset;    // sum(m) / ProbSize
    double m_const;        // mass of every particle when has_mass == 0
    int has_mass;
    int halo_stride;       // halo slots per tile
    unsigned tdiv[3];      // ceil(2^31 / tile[d]): i / tile[d] == (i * tdiv[d]) >> 31 for 0 <= i < 2^17
};

// A (box, level) piece of a launch over cells.  blk0..blk0+nblk-1 are the
// blocks of the launch that work on it.
struct CellSeg {
    int box;
    int n[3];
    int blk0, nblk;
    long long off;         // offset of this level in its pool
    long long coff;        // offset of the next coarser level (restrict/prolong),
                           // or of the bottom Green's function (bottom solve)
    long long foff;        // offset of the box's level-0 fields (phi, f)
    int zchunk;            // z-planes per block of the fused sweep
    int ebuf;              // which correction buffer (0: e, 1: e2) holds the coarse
                           // level's result (prolong) / the level's result (update)
    int lev;               // multigrid level of this piece (0: the box's own cells)
    double dh[3];          // 1/dx^2 per dimension on this level
    double gamma;          // -2 (dhx + dhy + dhz)
    double inv_gamma;      // 1/gamma
};

// A box's particles (or tiles) as one piece of a launch.
struct BoxSeg {
    int box;
    int blk0, nblk;
    int pad;
};

// XCD-aware block order.  Workgroups of a launch are dealt to the 8 XCDs round
// robin (block b runs on XCD b % 8), and each XCD has its own L2.  This bijection
// gives XCD x a contiguous range of logical blocks, so neighbouring tiles (which
// share halo cache lines) run under the same L2.
__device__ __forceinline__ int xcd_block(int b, int G) {
    const int q = G >> 3, rm = G & 7, x = b & 7, j = b >> 3;
    return (x < rm ? x * (q + 1) : rm * (q + 1) + (x - rm) * q) + j;
}

template <typename S>
__device__ __forceinline__ int find_seg(const S* segs, int nseg, int blk) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].blk0 <= blk) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int wrap1(int i, int n) {
    // one periodic image either side; anything further is clamped (and flagged
    // by the caller where it matters)
    i = i < 0 ? i + n : (i >= n ? i - n : i);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

__device__ __forceinline__ size_t cidx(int i, int j, int k, int nx, int ny) {
    return ((size_t)k * (size_t)ny + (size_t)j) * (size_t)nx + (size_t)i;
}

// Bit pattern order of non-negative doubles equals numeric order: max-norms are
// reduced with atomicMax on the pattern (exact and order independent).
__device__ __forceinline__ unsigned long long dbits(double v) {
    return (unsigned long long)__double_as_longlong(v);
}

}  // namespace nyxpm
