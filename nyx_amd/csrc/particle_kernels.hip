// particle_kernels.hip -- particle side of the PM step on gfx950:
//   * tile binning (counting sort by deposit tile, wave-aggregated atomics),
//   * deterministic CIC deposit (AssignDensity: NyxParticleContainer.H:90-97 ->
//     amrex_deposit_cic + SumBoundary + 1/vol),
//   * fold of tile halos + RHS preparation (Gravity.cpp:1207-1244, 283, 290),
//   * fused central-difference gradient + CIC gather + kick(-drift)
//     (get_new_grav_vector Gravity.cpp:817-850 + update_dm_particle_single,
//     DarkMatterParticleContainer.cpp:249-306), with the periodic shift of
//     Redistribute fused into the kick.
//
// Deposit determinism.  Each tile's particles are accumulated in LDS as exact
// fixed-point integers (two unsigned 64-bit words per cell: units 2^(E-31) and
// 2^(E-63) with 2^E <= max particle mass), so the sum does not depend on the order
// in which particles arrive; the per-tile sums are then converted to FP64 and the
// tile halos folded in a fixed order.  Runs are bitwise reproducible although the
// binning itself uses atomics.
#include "nyxpm_internal.h"

namespace nyxpm {

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------

// Cell index i = floor((x - plo)*dxi + 0.5) of a coordinate, as both the deposit
// and the interpolation compute it, plus the fractional weight.
__device__ __forceinline__ int cic_index(double x, double plo, double dxi, double* frac) {
    const double l = (x - plo) * dxi + 0.5;
    const int i = (int)floor(l);
    *frac = l - i;
    return i;
}

// Wrap an index known to lie within one period of [0, n); flags anything else.
__device__ __forceinline__ int wrap_checked(int i, int n, int* err) {
    if (i < -n || i >= 2 * n) { if (err) atomicOr(err, 1); }
    return wrap1(i, n);
}

// Wave-aggregated atomicAdd: lanes with equal keys share one atomic.  Returns, for
// each valid lane, the counter's previous value plus the lane's rank among the
// lanes with its key.  Must be called by every lane of the wave.
__device__ __forceinline__ int agg_atomic_add(int* counters, int key, bool valid) {
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned long long below = (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
    int result = 0;
    bool todo = valid;
    for (;;) {
        const unsigned long long act = __ballot(todo);
        if (act == 0ull) break;
        const int leader = __ffsll((unsigned long long)act) - 1;
        const int lkey = __shfl(key, leader);
        const bool mine = todo && key == lkey;
        const unsigned long long mm = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counters[lkey], (int)__popcll(mm));
        base = __shfl(base, leader);
        if (mine) {
            result = base + (int)__popcll(mm & below);
            todo = false;
        }
    }
    return result;
}

// Appends v to list (length counter n) for the active lanes with pred set: one
// atomic per wave, safe in divergent code (only active lanes take part).
__device__ __forceinline__ void wave_append(long long* list, int* n, long long v, bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n, (int)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) list[base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = v;
}

// x of an empty particle slot.
__device__ __forceinline__ double empty_slot() { return __longlong_as_double(0x7ff8000000000000LL); }

// Slots above which a tile's deposit block is launched ahead of the others.
constexpr int kHeavyTile = 2048;

// i / T for 0 <= i < 2^17 and T <= 32 with magic = ceil(2^31 / T) (exact over that
// range; the bound is checked when a context is created).
__device__ __forceinline__ int tile_div(int i, unsigned magic) {
    return (int)(((unsigned long long)(unsigned)i * magic) >> 31);
}

__device__ __forceinline__ int tile_of(const BoxDev& B, double x, double y, double z, int* err) {
    double f;
    const int ci = cic_index(x, B.plo[0], B.dxi[0], &f);
    const int cj = cic_index(y, B.plo[1], B.dxi[1], &f);
    const int ck = cic_index(z, B.plo[2], B.dxi[2], &f);
    // an index more than one period out: one flag branch for the three (rare)
    if ((unsigned)(ci + B.n[0]) >= (unsigned)(3 * B.n[0]) || (unsigned)(cj + B.n[1]) >= (unsigned)(3 * B.n[1]) ||
        (unsigned)(ck + B.n[2]) >= (unsigned)(3 * B.n[2]))
        atomicOr(err, 1);
    const int i = wrap1(ci, B.n[0]), j = wrap1(cj, B.n[1]), k = wrap1(ck, B.n[2]);
    const int tx = tile_div(i, B.tdiv[0]), ty = tile_div(j, B.tdiv[1]), tz = tile_div(k, B.tdiv[2]);
    return (int)(B.tile_off + ((long long)tz * B.ntile[1] + ty) * B.ntile[0] + tx);
}

// ---------------------------------------------------------------------------
// binning: keys + counts, exclusive scan, scatter
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_keys(const BoxSeg* segs, int nseg, const BoxDev* boxes,
                                                 const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, int* __restrict__ key,
                                                 int* counts, int* err) {
    const BoxSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    bool valid = false;
    int t = 0;
    if (q < B.span) {
        const long long p = B.part_off + q;
        const double px = x[p];
        valid = !isnan(px);  // NaN: an empty slot, keyed -1
        if (valid) t = tile_of(B, px, y[p], z[p], err);
        key[p] = valid ? t : -1;
    }
    agg_atomic_add(counts, t, valid);
}

// Slot capacity of each tile for a sort with slack: cnt + floor(cnt * num / 1024)
// + pad (integer arithmetic, so a box's capacities never exceed its slots).
__global__ void __launch_bounds__(kBlock) k_tile_caps(const int* __restrict__ counts, long long n, int num,
                                                      int pad, int* __restrict__ cap) {
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (t < n) {
        const long long c = counts[t];
        cap[t] = (int)(c + ((c * num) >> 10) + pad);
    }
}

// First slot of each tile: the scan of the capacities restarted at each box's
// first slot (boxes keep fixed slot ranges).
__global__ void __launch_bounds__(kBlock) k_tile_base(const BoxSeg* tsegs, int ntseg, const BoxDev* boxes,
                                                      const long long* __restrict__ start, long long n,
                                                      long long* __restrict__ tbase) {
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const BoxDev& B = boxes[tsegs[find_seg(tsegs, ntseg, (int)t)].box];
    tbase[t] = start[t] - start[B.tile_off] + B.part_off;
}

constexpr int kScanItems = 16;  // entries per thread in the scan passes

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int* __restrict__ counts, long long n,
                                                        long long* __restrict__ bsum) {
    const long long base = (long long)blockIdx.x * kBlock * kScanItems;
    long long s = 0;
    for (int it = 0; it < kScanItems; ++it) {
        const long long q = base + (long long)it * kBlock + threadIdx.x;
        if (q < n) s += counts[q];
    }
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ long long ws[kBlock / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < kBlock / kWave; ++w) t += ws[w];
        bsum[blockIdx.x] = t;
    }
}

// Exclusive scan of the block sums in place (one block), total in bsum[nb].
__global__ void __launch_bounds__(1024) k_scan_top(long long* bsum, int nb) {
    __shared__ long long sh[1024];
    long long carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int q = base + threadIdx.x;
        const long long v = q < nb ? bsum[q] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            long long t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (q < nb) bsum[q] = carry + sh[threadIdx.x] - v;
        const long long tot = sh[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

// start[q] = exclusive prefix of counts; cursor[q] = 0; start[n] = total.
__global__ void __launch_bounds__(kBlock) k_scan_down(const int* __restrict__ counts, long long n,
                                                      const long long* __restrict__ bsum,
                                                      long long* __restrict__ start, int* __restrict__ cursor) {
    __shared__ long long sh[kBlock];
    const long long base = (long long)blockIdx.x * kBlock * kScanItems;
    // each thread owns kScanItems consecutive entries
    long long vals[kScanItems];
    long long s = 0;
    for (int it = 0; it < kScanItems; ++it) {
        const long long q = base + (long long)threadIdx.x * kScanItems + it;
        vals[it] = q < n ? counts[q] : 0;
        s += vals[it];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        long long t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    long long run = bsum[blockIdx.x] + sh[threadIdx.x] - s;
    for (int it = 0; it < kScanItems; ++it) {
        const long long q = base + (long long)threadIdx.x * kScanItems + it;
        if (q < n) {
            start[q] = run;
            cursor[q] = 0;
        }
        run += vals[it];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[n] = bsum[gridDim.x];
}

struct PartPtrs {
    double *x, *y, *z, *vx, *vy, *vz, *m;
    int* id;
};

// Move every particle to its tile's slot range (order inside a tile is
// arbitrary); cursor[t] ends as the tile's particle count.
__global__ void __launch_bounds__(kBlock) k_scatter(const BoxSeg* segs, int nseg, const BoxDev* boxes,
                                                    const int* __restrict__ key,
                                                    const long long* __restrict__ tbase, int* cursor,
                                                    PartPtrs src, PartPtrs dst) {
    const BoxSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    const long long p = B.part_off + q;
    // occupied slots only (a kick-drift keys the slots it visits, not the free ones)
    bool valid = false;
    int t = 0;
    if (q < B.span) {
        const int k = key[p];
        valid = k >= 0 && !isnan(src.x[p]);
        t = valid ? k : 0;
    }
    const int r = agg_atomic_add(cursor, t, valid);
    if (valid) {
        const long long d = tbase[t] + r;
        dst.x[d] = src.x[p];
        dst.y[d] = src.y[p];
        dst.z[d] = src.z[p];
        dst.vx[d] = src.vx[p];
        dst.vy[d] = src.vy[p];
        dst.vz[d] = src.vz[p];
        if (src.m) dst.m[d] = src.m[p];
        dst.id[d] = src.id[p];
    }
}

// ---------------------------------------------------------------------------
// deposit
// ---------------------------------------------------------------------------

// Index of a low-halo cell (lx==0 || ly==0 || lz==0) of a tile whose LDS block is
// (bx+1) x (by+1) x (bz+1): the x face first (y fastest), then the y face without
// its x=0 column, then the z face without x=0 and y=0 (both x fastest).  The fold
// walks cells row-major, so the y- and z-face entries it reads for a row of a
// tile's top face are consecutive (a strided face order fetched a separate
// memory sector per entry).
__device__ __forceinline__ int halo_index(int lx, int ly, int lz, int bx, int by, int bz) {
    if (lx == 0) return lz * (by + 1) + ly;
    const int fx = (by + 1) * (bz + 1);
    if (ly == 0) return fx + lz * bx + (lx - 1);
    const int fy = bx * (bz + 1);
    return fx + fy + (ly - 1) * bx + (lx - 1);
}

// A non-negative FP32 value w < 2^(E+1) in units 2^(E-31), rounded to nearest:
// w * 2^(31-E) < 2^32 is exact in FP32 (a power-of-two scale, `scale`), and the
// native f32 -> u32 conversion (truncating) of it plus one half rounds it (ties
// as FP32 rounds the sum; a value that rounds up to 2^32 saturates at 2^32 - 1).
// Four instructions per contribution.
__device__ __forceinline__ unsigned fx32_units(float w, float scale) {
    return __float2uint_rz(w * scale + 0.5f);
}

// One block per tile.  Contributions m*sx[ii]*sy[jj]*sz[kk] (amrex_deposit_cic
// weights, product left to right) of the tile's particles go to LDS as fixed
// point; interior cells are written to `own`, the low halo layer to `halo`.
//
// F32 (the mixed-precision step): the fractional weights and the product are
// FP32 (the cell index and the fraction come from the FP64 position as in FP64
// mode), and each contribution is one 32-bit count of units 2^(E-31)
// (fx32_units) added into the cell's 64-bit word, so the LDS sum is still exact
// and order independent, with half the atomics: a cell's sum of at most 2^32
// contributions below 2^32 units each stays below 2^64.  Per contribution the
// FP32 rounding (~6e-8 relative) dominates the unit (2^-31 of the largest mass).
template <bool F32>
__global__ void __launch_bounds__(kBlock) k_deposit(const BoxSeg* segs, int nseg, const BoxDev* boxes,
                                                    const long long* __restrict__ tbase,
                                                    const int* __restrict__ thw,
                                                    const int* __restrict__ heavy, const int* nheavy,
                                                    int heavy_cap,
                                                    const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ m,
                                                    double* __restrict__ own, double* __restrict__ halo, int* err) {
    extern __shared__ unsigned long long acc[];
    // the first heavy_cap blocks take the heavy tiles (listed by k_make_items), so
    // the longest blocks start first; the rest take every other tile in order
    int tile;
    if ((int)blockIdx.x < heavy_cap) {
        if ((int)blockIdx.x >= *nheavy) return;
        tile = heavy[blockIdx.x];
    } else {
        tile = blockIdx.x - heavy_cap;
        if (thw[tile] > kHeavyTile) return;
    }
    const BoxSeg sg = segs[find_seg(segs, nseg, tile)];
    // the box's fields as block-uniform values in registers: read through a
    // reference they would be re-read from memory after every possible store
    // or atomic (the error flag) of the particle loop
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const int lt = tile - sg.blk0;
    const int tx = lt % B.ntile[0], ty = (lt / B.ntile[0]) % B.ntile[1], tz = lt / (B.ntile[0] * B.ntile[1]);
    const int o0 = tx * B.tile[0], o1 = ty * B.tile[1], o2 = tz * B.tile[2];
    const int bx = min(B.tile[0], B.n[0] - o0), by = min(B.tile[1], B.n[1] - o1), bz = min(B.tile[2], B.n[2] - o2);
    const int Lx = bx + 1, Ly = by + 1, Lz = bz + 1;
    const int LC = Lx * Ly * Lz;
    const long long gt = B.tile_off + lt;
    const long long p0 = tbase[gt], p1 = p0 + thw[gt];
    // the thread's first two particles are loaded before the accumulator is
    // cleared, so their latency overlaps the clearing
    double xa = 0.0, ya = 0.0, za = 0.0, ma = 0.0, xb = 0.0, yb = 0.0, zb = 0.0, mb = 0.0;
    auto load2 = [&](long long pa) {
        const long long pb = pa + kBlock;
        if (pa < p1) {
            xa = x[pa]; ya = y[pa]; za = z[pa];
            ma = B.has_mass ? m[pa] : B.m_const;
        }
        if (pb < p1) {
            xb = x[pb]; yb = y[pb]; zb = z[pb];
            mb = B.has_mass ? m[pb] : B.m_const;
        }
    };
    load2(p0 + threadIdx.x);
    for (int c = threadIdx.x; c < (F32 ? LC : 2 * LC); c += kBlock) acc[c] = 0ull;
    __syncthreads();
    const int E = B.fx_exp;
    // one particle's CIC contributions as two fixed-point words: w * 2^(31-E) =
    // hi + lo * 2^-32 (hi < 2^32 since w < 2^(E+1); lo rounded, a rounding up to
    // 2^32 carried into hi), both converted with the native f64 -> u32 instruction
    auto dep = [&](double px, double py, double pz, double mp) {
        if (isnan(px)) return;  // empty slot
        double xint, yint, zint;
        const int i = cic_index(px, B.plo[0], B.dxi[0], &xint);
        const int j = cic_index(py, B.plo[1], B.dxi[1], &yint);
        const int k = cic_index(pz, B.plo[2], B.dxi[2], &zint);
        // an index more than one period out (flag 1) or a cell outside the tile
        // the particle is binned in (flag 2): one branch for both, rarely taken
        const bool far = (unsigned)(i + B.n[0]) >= (unsigned)(3 * B.n[0]) ||
                         (unsigned)(j + B.n[1]) >= (unsigned)(3 * B.n[1]) ||
                         (unsigned)(k + B.n[2]) >= (unsigned)(3 * B.n[2]);
        const int iw = wrap1(i, B.n[0]) - o0;
        const int jw = wrap1(j, B.n[1]) - o1;
        const int kw = wrap1(k, B.n[2]) - o2;
        const bool outside = (unsigned)iw >= (unsigned)bx || (unsigned)jw >= (unsigned)by ||
                             (unsigned)kw >= (unsigned)bz;
        if (far || outside) {
            atomicOr(err, (far ? 1 : 0) | (outside ? 2 : 0));
            return;
        }
        if (F32) {
            const float scale = ldexpf(1.0f, 31 - E);
            const float fx = (float)xint, fy = (float)yint, fz = (float)zint, fm = (float)mp;
            const float sx[2] = {1.0f - fx, fx};
            const float sy[2] = {1.0f - fy, fy};
            const float sz[2] = {1.0f - fz, fz};
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii) {
                        const float w = sx[ii] * sy[jj] * sz[kk] * fm;
                        const int idx = ((kw + kk) * Ly + (jw + jj)) * Lx + (iw + ii);
                        atomicAdd(&acc[idx], (unsigned long long)fx32_units(w, scale));
                    }
            return;
        }
        const double sx[2] = {1.0 - xint, xint};
        const double sy[2] = {1.0 - yint, yint};
        const double sz[2] = {1.0 - zint, zint};
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int ii = 0; ii < 2; ++ii) {
                    const double w = sx[ii] * sy[jj] * sz[kk] * mp;
                    const double s = ldexp(w, 31 - E);
                    const double hf = floor(s);
                    const double lr = rint(ldexp(s - hf, 32));
                    const bool carry = lr >= 4294967296.0;
                    const unsigned long long hi = (unsigned long long)(unsigned)hf + (carry ? 1ull : 0ull);
                    const unsigned long long lo = carry ? 0ull : (unsigned long long)(unsigned)lr;
                    const int idx = ((kw + kk) * Ly + (jw + jj)) * Lx + (iw + ii);
                    atomicAdd(&acc[idx], hi);
                    atomicAdd(&acc[LC + idx], lo);
                }
    };
    // two particles per thread, both loaded before either is deposited
    for (long long pa = p0 + threadIdx.x; pa < p1; pa += 2 * kBlock) {
        if (pa != p0 + threadIdx.x) load2(pa);
        dep(xa, ya, za, ma);
        if (pa + kBlock < p1) dep(xb, yb, zb, mb);
    }
    __syncthreads();
    // write-out by rows of Lx cells (Lx <= 33 < 64): a wave takes rpw rows per
    // pass, so no cell needs a division for its coordinates
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int rpw = kWave / Lx;
    const int lrow = lane / Lx, lx = lane - lrow * Lx;
    if (lrow < rpw) {
        const int step = (kBlock / kWave) * rpw;
        const int r0 = wave * rpw + lrow;
        int ly = r0 % Ly, lz = r0 / Ly;
        const int dly = step % Ly, dlz = step / Ly;
        while (lz < Lz) {
            const int c = (lz * Ly + ly) * Lx + lx;
            const double v = F32 ? ldexp((double)acc[c], E - 31)
                                 : ldexp((double)acc[c], E - 31) + ldexp((double)acc[LC + c], E - 63);
            if (lx > 0 && ly > 0 && lz > 0)
                own[B.cell_off + cidx(o0 + lx - 1, o1 + ly - 1, o2 + lz - 1, B.n[0], B.n[1])] = v;
            else
                halo[gt * B.halo_stride + halo_index(lx, ly, lz, bx, by, bz)] = v;
            ly += dly;
            lz += dlz;
            if (ly >= Ly) {
                ly -= Ly;
                ++lz;
            }
        }
    }
}

// Momentum deposit for AssignDensityAndVels: one velocity component per launch,
// contributions (sx*sy*sz*m) * v_d as signed fixed point (two's complement in
// the u64 words, hi in units 2^(E-30) converted through int32, lo in 2^(E-62),
// 2^E <= max |m v_d| of the box, from maxbits), so the sums are exact and order
// independent like the mass deposit's; own cells to `own`, the low halo layer to
// `halo` (k_deposit's layout).
__global__ void __launch_bounds__(kBlock) k_deposit_mom(const BoxSeg* segs, int nseg, const BoxDev* boxes,
                                                        const long long* __restrict__ tbase,
                                                        const int* __restrict__ thw,
                                                        const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, const double* __restrict__ m,
                                                        const double* __restrict__ v,
                                                        const unsigned long long* maxbits, int comp,
                                                        double* __restrict__ own, double* __restrict__ halo,
                                                        int* err) {
    extern __shared__ unsigned long long acc[];
    const BoxSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const int lt = blockIdx.x - sg.blk0;
    const int tx = lt % B.ntile[0], ty = (lt / B.ntile[0]) % B.ntile[1], tz = lt / (B.ntile[0] * B.ntile[1]);
    const int o0 = tx * B.tile[0], o1 = ty * B.tile[1], o2 = tz * B.tile[2];
    const int bx = min(B.tile[0], B.n[0] - o0), by = min(B.tile[1], B.n[1] - o1), bz = min(B.tile[2], B.n[2] - o2);
    const int Lx = bx + 1, Ly = by + 1, Lz = bz + 1;
    const int LC = Lx * Ly * Lz;
    for (int c = threadIdx.x; c < 2 * LC; c += kBlock) acc[c] = 0ull;
    __syncthreads();
    const long long gt = B.tile_off + lt;
    const long long p0 = tbase[gt], p1 = p0 + thw[gt];
    const double vmax = __longlong_as_double((long long)maxbits[3 * sg.box + comp]);
    const int E = vmax > 0.0 ? ilogb(vmax) : 0;
    for (long long p = p0 + threadIdx.x; p < p1; p += kBlock) {
        const double px = x[p];
        if (isnan(px)) continue;  // empty slot
        double xint, yint, zint;
        const int i = cic_index(px, B.plo[0], B.dxi[0], &xint);
        const int j = cic_index(y[p], B.plo[1], B.dxi[1], &yint);
        const int k = cic_index(z[p], B.plo[2], B.dxi[2], &zint);
        const int iw = wrap_checked(i, B.n[0], err) - o0;
        const int jw = wrap_checked(j, B.n[1], err) - o1;
        const int kw = wrap_checked(k, B.n[2], err) - o2;
        if (iw < 0 || iw >= bx || jw < 0 || jw >= by || kw < 0 || kw >= bz) {
            atomicOr(err, 2);
            continue;
        }
        const double sx[2] = {1.0 - xint, xint};
        const double sy[2] = {1.0 - yint, yint};
        const double sz[2] = {1.0 - zint, zint};
        const double mp = B.has_mass ? m[p] : B.m_const;
        const double vp = v[p];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int ii = 0; ii < 2; ++ii) {
                    const double w = sx[ii] * sy[jj] * sz[kk] * mp;
                    const double s = ldexp(w * vp, 30 - E);  // |s| < 2^31
                    const double hf = floor(s);
                    const double lr = rint(ldexp(s - hf, 32));
                    const bool carry = lr >= 4294967296.0;
                    const long long hi = (long long)(int)hf + (carry ? 1 : 0);
                    const unsigned long long lo = carry ? 0ull : (unsigned long long)(unsigned)lr;
                    const int idx = ((kw + kk) * Ly + (jw + jj)) * Lx + (iw + ii);
                    atomicAdd(&acc[idx], (unsigned long long)hi);
                    atomicAdd(&acc[LC + idx], lo);
                }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < LC; c += kBlock) {
        const int lx = c % Lx, ly = (c / Lx) % Ly, lz = c / (Lx * Ly);
        const double val = ldexp((double)(long long)acc[c], E - 30) + ldexp((double)acc[LC + c], E - 62);
        if (lx > 0 && ly > 0 && lz > 0)
            own[B.cell_off + cidx(o0 + lx - 1, o1 + ly - 1, o2 + lz - 1, B.n[0], B.n[1])] = val;
        else
            halo[gt * B.halo_stride + halo_index(lx, ly, lz, bx, by, bz)] = val;
    }
}

// max |m v_d| per box and component (bit pattern, atomicMax) for k_deposit_mom.
__global__ void __launch_bounds__(kBlock) k_max_mom(const BoxSeg* segs, int nseg, const BoxDev* boxes,
                                                    const double* __restrict__ x, const double* __restrict__ m, const double* __restrict__ vx,
                                                    const double* __restrict__ vy, const double* __restrict__ vz,
                                                    unsigned long long* maxbits) {
    const BoxSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    if (q >= B.span) return;
    const long long p = B.part_off + q;
    if (isnan(x[p])) return;
    const double mp = B.has_mass ? m[p] : B.m_const;
    atomicMax(maxbits + 3 * sg.box + 0, dbits(fabs(mp * vx[p])));
    atomicMax(maxbits + 3 * sg.box + 1, dbits(fabs(mp * vy[p])));
    atomicMax(maxbits + 3 * sg.box + 2, dbits(fabs(mp * vz[p])));
}

// Sum of a block's values in a fixed order (wave butterfly, then waves in order).
__device__ __forceinline__ double block_sum_fixed(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __shared__ double ws[kBlock / kWave];
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) t += ws[w];
    return t;
}

// The edge and corner terms of cell (i, j, k) after SumBoundary: the low-halo
// entries of the upper neighbour tiles across two or three of the tile's top
// faces (neighbour in xy, xz, yz, xyz, in that order), added to v.
__device__ __forceinline__ double fold_edges(const BoxDev& B, const double* __restrict__ halo, double v, int tx,
                                             int ty, int tz, int li, int lj, int lk, bool topx, bool topy,
                                             bool topz) {
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int T0 = B.tile[0], T1 = B.tile[1], T2 = B.tile[2];
    const int nt0 = B.ntile[0], nt1 = B.ntile[1], nt2 = B.ntile[2];
    const int sel[4] = {3, 5, 6, 7};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int S = sel[s];
        const bool sx = S & 1, sy = S & 2, sz = S & 4;
        if ((sx && !topx) || (sy && !topy) || (sz && !topz)) continue;
        int ntx = tx + (sx ? 1 : 0), nty = ty + (sy ? 1 : 0), ntz = tz + (sz ? 1 : 0);
        if (ntx == nt0) ntx = 0;
        if (nty == nt1) nty = 0;
        if (ntz == nt2) ntz = 0;
        const int nbx = min(T0, nx - ntx * T0), nby = min(T1, ny - nty * T1), nbz = min(T2, nz - ntz * T2);
        const int hx = sx ? 0 : li + 1, hy = sy ? 0 : lj + 1, hz = sz ? 0 : lk + 1;
        const long long nt = B.tile_off + ((long long)ntz * nt1 + nty) * nt0 + ntx;
        v += halo[nt * B.halo_stride + halo_index(hx, hy, hz, nbx, nby, nbz)];
    }
    return v;
}

// Row-major walk over a block's share of a box's cells: block b of the piece takes
// cells [b*per, (b+1)*per) (per a multiple of kBlock, so whole cache lines), each
// thread every kBlock-th of them in order; body(i, j, k, c) per cell.  The cell
// order of each thread is fixed, so per-thread sums are reproducible.
template <class F>
__device__ __forceinline__ void for_block_cells(const CellSeg& sg, const BoxDev& B, F body) {
    const int nx = B.n[0], ny = B.n[1];
    const long long N = (long long)nx * ny * B.n[2];
    long long per = (N + sg.nblk - 1) / sg.nblk;
    per = (per + kBlock - 1) / kBlock * kBlock;
    const long long c0 = (long long)(blockIdx.x - sg.blk0) * per;
    const long long c1 = min(N, c0 + per);
    long long c = c0 + threadIdx.x;
    if (c >= c1) return;
    const long long plane = (long long)nx * ny;
    int k = (int)(c / plane);
    const long long rem = c - (long long)k * plane;
    int j = (int)(rem / nx);
    int i = (int)(rem - (long long)j * nx);
    for (; c < c1; c += kBlock) {
        body(i, j, k, c);
        i += kBlock;
        if (i >= nx) {
            const int t = i / nx;
            i -= t * nx;
            j += t;
            if (j >= ny) {
                const int u = j / ny;
                j -= u * ny;
                k += u;
            }
        }
    }
}

// The folded deposit (SumBoundary) of a block's share of a box's cells, in
// for_block_cells' order (the same cells per thread, in the same sequence):
// own deposit plus the low-halo entries of the up-to-7 upper neighbour tiles
// that hold the cell, added in a fixed order (neighbour in x, y, z, xy, xz, yz,
// xyz); emit(q, v) per cell in order.  U cells per thread at a time: their own
// values and face terms (x, y, z) are loaded before any is used, so a thread
// has up to 4U loads in flight; the edge and corner terms (cells on two or
// three top faces of a tile) follow them.  Coordinates advance by kBlock cells
// with block-uniform carries instead of divisions.
template <int U, class F>
__device__ __forceinline__ void fold_walk(const CellSeg& sg, const BoxDev& B, const double* own,
                                          const double* __restrict__ halo, F emit) {
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int T0 = B.tile[0], T1 = B.tile[1], T2 = B.tile[2];
    const int nt0 = B.ntile[0], nt1 = B.ntile[1], nt2 = B.ntile[2];
    const long long N = (long long)nx * ny * nz;
    long long per = (N + sg.nblk - 1) / sg.nblk;
    per = (per + kBlock - 1) / kBlock * kBlock;
    const long long c0 = (long long)(blockIdx.x - sg.blk0) * per;
    const long long c1 = min(N, c0 + per);
    long long c = c0 + threadIdx.x;
    if (c >= c1) return;
    const long long plane = (long long)nx * ny;
    int k = (int)(c / plane);
    const long long rem = c - (long long)k * plane;
    int j = (int)(rem / nx);
    int i = (int)(rem - (long long)j * nx);
    // kBlock cells = dk planes + dj rows + di cells (block-uniform)
    const int di = kBlock % nx, dq = kBlock / nx;
    const int dj = dq % ny, dk = dq / ny;
    const long long hbase = B.tile_off;
    const long long hs = B.halo_stride;
    for (; c < c1; c += (long long)U * kBlock) {
        double v[U], hx[U], hy[U], hz[U];
        int cx[U], cy[U], cz[U];
        int fl[U];  // bit 0, 1, 2: on the tile's top x, y, z face; -1: past the block's cells
#pragma unroll
        for (int u = 0; u < U; ++u) {
            cx[u] = i;
            cy[u] = j;
            cz[u] = k;
            fl[u] = -1;
            hx[u] = hy[u] = hz[u] = 0.0;
            if (c + (long long)u * kBlock < c1) {
                v[u] = own[B.cell_off + c + (long long)u * kBlock];
                const int tx = tile_div(i, B.tdiv[0]), ty = tile_div(j, B.tdiv[1]), tz = tile_div(k, B.tdiv[2]);
                const int li = i - tx * T0, lj = j - ty * T1, lk = k - tz * T2;
                const int bx = min(T0, nx - tx * T0), by = min(T1, ny - ty * T1), bz = min(T2, nz - tz * T2);
                const int lt = (tz * nt1 + ty) * nt0 + tx;
                fl[u] = (li == bx - 1 ? 1 : 0) | (lj == by - 1 ? 2 : 0) | (lk == bz - 1 ? 4 : 0);
                if (li == bx - 1) {  // x face of tile tx + 1 (same rows and planes)
                    const int nlt = tx + 1 == nt0 ? lt - tx : lt + 1;
                    hx[u] = halo[(hbase + nlt) * hs + (lk + 1) * (by + 1) + (lj + 1)];
                }
                if (lj == by - 1) {  // y face of tile ty + 1
                    const int nty = ty + 1 == nt1 ? 0 : ty + 1;
                    const int nby = min(T1, ny - nty * T1);
                    const int nlt = lt + (nty - ty) * nt0;
                    hy[u] = halo[(hbase + nlt) * hs + (nby + 1) * (bz + 1) + (lk + 1) * bx + li];
                }
                if (lk == bz - 1) {  // z face of tile tz + 1
                    const int ntz = tz + 1 == nt2 ? 0 : tz + 1;
                    const int nbz = min(T2, nz - ntz * T2);
                    const int nlt = lt + (ntz - tz) * nt0 * nt1;
                    hz[u] = halo[(hbase + nlt) * hs + (by + 1) * (nbz + 1) + bx * (nbz + 1) + lj * bx + li];
                }
            }
            i += di;
            j += dj;
            k += dk;
            if (i >= nx) {
                i -= nx;
                ++j;
            }
            if (j >= ny) {
                j -= ny;
                ++k;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int f = fl[u];
            if (f < 0) continue;
            double w = v[u];
            if (f & 1) w += hx[u];
            if (f & 2) w += hy[u];
            if (f & 4) w += hz[u];
            if (__builtin_popcount(f) >= 2) {
                const int tx = tile_div(cx[u], B.tdiv[0]), ty = tile_div(cy[u], B.tdiv[1]),
                          tz = tile_div(cz[u], B.tdiv[2]);
                w = fold_edges(B, halo, w, tx, ty, tz, cx[u] - tx * T0, cy[u] - ty * T1, cz[u] - tz * T2, f & 1,
                               f & 2, f & 4);
            }
            emit(B.cell_off + (size_t)(c + (long long)u * kBlock), w);
        }
    }
}

// Cells per thread in flight in the fold.
constexpr int kFoldU = 4;

// rho = (own + halos, fold_walk) * 1/vol; f = rho + (-mass_offset); per-block sum
// of f for the solvability correction.  Cells in row-major order (coalesced reads
// of own and writes of f, whole cache lines).  rho itself is written only when the
// caller wants the density (AssignDensity); the step needs only f.
template <bool WRITE_RHO>
__global__ void __launch_bounds__(kBlock) k_fold(const CellSeg* segs, int nseg, const BoxDev* boxes,
                                                 const double* own, const double* __restrict__ halo,
                                                 double* rho, double* __restrict__ f,
                                                 double* __restrict__ partial) {
    // rho may be own itself (each cell is read, then written, by one thread)
    const CellSeg& sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    double acc = 0.0;  // this thread's cells, in a fixed order
    fold_walk<kFoldU>(sg, B, own, halo, [&](size_t q, double v) {
        const double rv = v * B.inv_vol;
        if (WRITE_RHO) rho[q] = rv;
        const double fv = rv + (-B.mass_offset);
        f[q] = fv;
        acc += fv;
    });
    const double bs = block_sum_fixed(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// out = own + halos (fold_walk, no scaling): the folded raw sums of a deposit.
__global__ void __launch_bounds__(kBlock) k_fold_raw(const CellSeg* segs, int nseg, const BoxDev* boxes,
                                                     const double* __restrict__ own,
                                                     const double* __restrict__ halo, double* __restrict__ out) {
    const CellSeg& sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    fold_walk<kFoldU>(sg, B, own, halo, [&](size_t q, double v) { out[q] = v; });
}

// AssignDensityAndVels for box `box`: rho = mass * 1/vol, v_d = mom_d / mass (0
// without mass), written as 4 fields of the box's cells.
__global__ void __launch_bounds__(kBlock) k_vels_finish(const BoxDev* boxes, int box,
                                                        const double* __restrict__ mass,
                                                        const double* __restrict__ mom,
                                                        long long mom_stride, double* __restrict__ out) {
    const BoxDev B = boxes[box];
    const long long nc = (long long)B.n[0] * B.n[1] * B.n[2];
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= nc) return;
    const double ms = mass[B.cell_off + q];
    out[q] = ms * B.inv_vol;
    for (int d = 0; d < 3; ++d) {
        const double md = mom[d * mom_stride + B.cell_off + q];
        out[(d + 1) * nc + q] = ms != 0.0 ? md / ms : 0.0;
    }
}


// sum = (S*dV)/ProbSize, as vol_weight_sum computes it (sum_utils.cpp:85-93).
__global__ void __launch_bounds__(kBlock) k_rhs_sum(const CellSeg* segs, const double* __restrict__ partial,
                                                    const double* __restrict__ dv_size, double* __restrict__ solv) {
    const CellSeg& sg = segs[blockIdx.x];
    double s = 0.0;
    for (int b = threadIdx.x; b < sg.nblk; b += kBlock) s += partial[sg.blk0 + b];
    const double S = block_sum_fixed(s);
    if (threadIdx.x == 0) {
        double sum = S * dv_size[2 * sg.box];
        sum = sum / dv_size[2 * sg.box + 1];
        solv[sg.box] = sum;
    }
}

// f = ((f + (-sum)) * Ggravity) * (1/a), Ggravity = -4 pi G.
__global__ void __launch_bounds__(kBlock) k_rhs_final(const CellSeg* segs, int nseg, const BoxDev* boxes,
                                                      const double* __restrict__ solv, double* __restrict__ f,
                                                      double Ggravity, double inv_a) {
    const CellSeg& sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    if (q >= n) return;
    const double ns = -solv[sg.box];
    double v = f[B.cell_off + q] + ns;
    f[B.cell_off + q] = (v * Ggravity) * inv_a;
}

// ---------------------------------------------------------------------------
// gradient + CIC gather + kick (+ drift)
// ---------------------------------------------------------------------------

// Periodic position shift (Redistribute): above rhi -> subtract the length until
// not (clamp to lo on undershoot); below lo -> add until not (clamp to rhi).
// A coordinate more than two periods away (or not finite) is clamped and flagged
// instead of looping.
__device__ __forceinline__ double periodic_shift(double v, double lo, double rhi, double len, int* err) {
    if (!(v >= lo - 2.0 * len && v <= rhi + 2.0 * len)) {
        if (err) atomicOr(err, 8);
        return (v < lo) ? lo : rhi;  // NaN compares false: goes to rhi
    }
    if (v > rhi) {
        while (v > rhi) v = v - len;
        if (v < lo) v = lo;
    } else if (v < lo) {
        while (v < lo) v = v + len;
        if (v > rhi) v = rhi;
    }
    return v;
}

// val[d] = sum_{kk,jj,ii} sx[1-ii]*sy[1-jj]*sz[1-kk] * g_d(i-ii, j-jj, k-kk) with
// g_d(c) = 0.5*(dxi*(phi(c) - phi(c-e_d)) + dxi*(phi(c+e_d) - phi(c))).
// G is the interpolation arithmetic: double (the drop-in) or float (the
// mixed-precision step: g of each corner cell formed in FP64 and rounded to
// FP32, the fractions rounded to FP32, the weights and the sum in FP32, in the
// order of the staged kick, so the two agree bit for bit).  Forming g in FP64
// first keeps the difference of neighbouring phi values exact: phi rounded to
// FP32 first would lose ~n/2pi times more of g's digits to cancellation.
template <typename G>
__device__ __forceinline__ void gather_grad(const BoxDev& B, const double* __restrict__ P, double x, double y,
                                            double z, double val[3], int* err) {
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    double xint, yint, zint;
    const int i = cic_index(x, B.plo[0], B.dxi[0], &xint);
    const int j = cic_index(y, B.plo[1], B.dxi[1], &yint);
    const int k = cic_index(z, B.plo[2], B.dxi[2], &zint);
    if (i < -nx + 2 || i >= 2 * nx - 1 || j < -ny + 2 || j >= 2 * ny - 1 || k < -nz + 2 || k >= 2 * nz - 1)
        atomicOr(err, 4);
    int ix[4], iy[4], iz[4];
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) {
        ix[mm] = wrap1(i - 2 + mm, nx);
        iy[mm] = wrap1(j - 2 + mm, ny);
        iz[mm] = wrap1(k - 2 + mm, nz);
    }
    const G sx[2] = {G(1) - G(xint), G(xint)};
    const G sy[2] = {G(1) - G(yint), G(yint)};
    const G sz[2] = {G(1) - G(zint), G(zint)};
    auto ld = [&](int a, int b, int c) -> double { return P[cidx(ix[a], iy[b], iz[c], nx, ny)]; };
    const double dxi0 = B.dxi[0], dxi1 = B.dxi[1], dxi2 = B.dxi[2];
    G v0 = G(0), v1 = G(0), v2 = G(0);
#pragma unroll
    for (int kk = 0; kk <= 1; ++kk)
#pragma unroll
        for (int jj = 0; jj <= 1; ++jj)
#pragma unroll
            for (int ii = 0; ii <= 1; ++ii) {
                const int a = 2 - ii, b = 2 - jj, c = 2 - kk;
                const double pc = ld(a, b, c);
                const G gx = G(0.5 * (dxi0 * (pc - ld(a - 1, b, c)) + dxi0 * (ld(a + 1, b, c) - pc)));
                const G gy = G(0.5 * (dxi1 * (pc - ld(a, b - 1, c)) + dxi1 * (ld(a, b + 1, c) - pc)));
                const G gz = G(0.5 * (dxi2 * (pc - ld(a, b, c - 1)) + dxi2 * (ld(a, b, c + 1) - pc)));
                const G w = sx[1 - ii] * sy[1 - jj] * sz[1 - kk];
                v0 += w * gx;
                v1 += w * gy;
                v2 += w * gz;
            }
    val[0] = double(v0);
    val[1] = double(v1);
    val[2] = double(v2);
}

// moveKickDrift (DO_MOVE) or moveKick (!DO_MOVE) of update_dm_particle_single:
//   v = (a_prev*v + (0.5*dt)*val) * (1/a_cur);  DO_MOVE: x += (dt*(1/a_cur))*v
// DO_MOVE also bins the drifted particle (key + count for the next deposit);
// !DO_MOVE applies the periodic shift of Redistribute.
template <bool DO_MOVE, typename G>
__global__ void __launch_bounds__(kBlock) k_kick(const BoxSeg* segs, int nseg, const BoxDev* boxes,
                                                 const double* __restrict__ phi, double* __restrict__ x,
                                                 double* __restrict__ y, double* __restrict__ z,
                                                 double* __restrict__ vx, double* __restrict__ vy,
                                                 double* __restrict__ vz, double dt, double a_prev, double a_cur,
                                                 int* __restrict__ key, int* counts, int* err) {
    const BoxSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    const long long p = B.part_off + q;
    double px = 0.0;
    if (q < B.span) {
        px = x[p];
        if (DO_MOVE && isnan(px)) key[p] = -1;  // empty slot
    }
    const bool valid = q < B.span && !isnan(px);
    int t = 0;
    if (valid) {
        const double half_dt = 0.5 * dt;
        const double a_cur_inv = 1.0 / a_cur;
        const double dt_a_cur_inv = dt * a_cur_inv;
        double py = y[p], pz = z[p];
        double val[3];
        gather_grad<G>(B, phi + B.cell_off, px, py, pz, val, err);
        double ux = a_prev * vx[p] + half_dt * val[0];
        ux = ux * a_cur_inv;
        double uy = a_prev * vy[p] + half_dt * val[1];
        uy = uy * a_cur_inv;
        double uz = a_prev * vz[p] + half_dt * val[2];
        uz = uz * a_cur_inv;
        vx[p] = ux;
        vy[p] = uy;
        vz[p] = uz;
        if (DO_MOVE) {
            px = px + dt_a_cur_inv * ux;
            py = py + dt_a_cur_inv * uy;
            pz = pz + dt_a_cur_inv * uz;
            x[p] = px;
            y[p] = py;
            z[p] = pz;
            t = tile_of(B, px, py, pz, err);
            key[p] = t;
        } else {
            const double sxp = periodic_shift(px, B.plo[0], B.rhi[0], B.len[0], err);
            const double syp = periodic_shift(py, B.plo[1], B.rhi[1], B.len[1], err);
            const double szp = periodic_shift(pz, B.plo[2], B.rhi[2], B.len[2], err);
            if (sxp != px) x[p] = sxp;
            if (syp != py) y[p] = syp;
            if (szp != pz) z[p] = szp;
        }
    }
    if (DO_MOVE) agg_atomic_add(counts, t, valid);
}

// Work items for the tile-staged particle kernels: each tile's occupied slot
// range [tbase, tbase + thw) cut into chunks of at most kChunk slots (empty
// tiles give none).  Slots are
// reserved with one atomic per wave, so the item order varies between runs; every
// item is processed independently, so results do not.
constexpr int kChunk = 16 * kBlock;  // default particles per work item (fewer items: less phi staging)

struct TileItem {
    long long p0;   // first particle
    int np;         // particles in this chunk
    int tile;       // global tile id
};

__global__ void __launch_bounds__(kBlock) k_make_items(const long long* __restrict__ tbase,
                                                       const int* __restrict__ thw, long long ntiles,
                                                       TileItem* __restrict__ items, int* nitems, int chunk,
                                                       int* __restrict__ heavy, int* nheavy) {
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
    int nc = 0;
    long long p0 = 0, cnt = 0;
    if (t < ntiles) {
        p0 = tbase[t];
        cnt = thw[t];
        nc = (int)((cnt + chunk - 1) / chunk);
        // tiles the deposit starts first (one block per tile: the longest blocks)
        if (cnt > kHeavyTile) heavy[atomicAdd(nheavy, 1)] = (int)t;
    }
    // wave-inclusive prefix of nc, one atomic per wave
    const int lane = threadIdx.x & (kWave - 1);
    int incl = nc;
    for (int o = 1; o < kWave; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, kWave - 1);
    int base = 0;
    if (lane == kWave - 1 && total > 0) base = atomicAdd(nitems, total);
    base = __shfl(base, kWave - 1);
    int slot = base + incl - nc;
    for (int c = 0; c < nc; ++c) {
        TileItem it;
        it.p0 = p0 + (long long)c * chunk;
        it.np = (int)min((long long)chunk, cnt - (long long)c * chunk);
        it.tile = (int)t;
        items[slot + c] = it;
    }
}

// NyxParticleContainer::estTimestep (NyxParticleContainer.H:327-410) for one
// particle: cfl / max_d(|v_d| dxi_d / a), capped by 1 / sqrt(|g| dxi_x) with g the
// cell-centred gravity of the particle's cell (cell = floor((x - plo) dxi)).
__device__ __forceinline__ double particle_dt(double ux, double uy, double uz, double adxi0, double adxi1,
                                              double adxi2, double ax, double ay, double az, double dxi0,
                                              double cfl) {
    const double mx = fabs(ux) * adxi0;
    const double my = fabs(uy) * adxi1;
    const double mz = fabs(uz) * adxi2;
    double m = mx > my ? mx : my;
    m = m > mz ? m : mz;
    double dtp = (m > 0) ? (cfl / m) : 1e50;
    const double mag = sqrt(ax * ax + ay * ay + az * az);
    if (mag > 0) {
        const double da = 1 / sqrt(mag * dxi0);
        dtp = dtp < da ? dtp : da;
    }
    return dtp;
}

// Cell-centred gravity (k_grav_vector's arithmetic, rounded to G as the staged
// kick keeps it) at cell (i, j, k) of a box.
template <typename G>
__device__ __forceinline__ void grav_at(const BoxDev& B, const double* __restrict__ P, int i, int j, int k,
                                        double& gx, double& gy, double& gz) {
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
    const double pc = P[cidx(i, j, k, nx, ny)];
    gx = double(G(0.5 * (B.dxi[0] * (pc - P[cidx(im, j, k, nx, ny)]) + B.dxi[0] * (P[cidx(ip, j, k, nx, ny)] - pc))));
    gy = double(G(0.5 * (B.dxi[1] * (pc - P[cidx(i, jm, k, nx, ny)]) + B.dxi[1] * (P[cidx(i, jp, k, nx, ny)] - pc))));
    gz = double(G(0.5 * (B.dxi[2] * (pc - P[cidx(i, j, km, nx, ny)]) + B.dxi[2] * (P[cidx(i, j, kp, nx, ny)] - pc))));
}

// The estimate from global phi (any position; cells wrapped periodically).
template <typename G>
__device__ __forceinline__ double particle_dt_global(const BoxDev& B, const double* __restrict__ P, double px,
                                                     double py, double pz, double ux, double uy, double uz,
                                                     double a, double cfl) {
    const int i = wrap1((int)floor((px - B.plo[0]) * B.dxi[0]), B.n[0]);
    const int j = wrap1((int)floor((py - B.plo[1]) * B.dxi[1]), B.n[1]);
    const int k = wrap1((int)floor((pz - B.plo[2]) * B.dxi[2]), B.n[2]);
    double gx, gy, gz;
    grav_at<G>(B, P, i, j, k, gx, gy, gz);
    return particle_dt(ux, uy, uz, B.dxi[0] / a, B.dxi[1] / a, B.dxi[2] / a, gx, gy, gz, B.dxi[0], cfl);
}

// Minimum over a block into a per-box slot (bit pattern of a positive double:
// atomicMin is exact and order independent).
__device__ __forceinline__ void block_min_to(double v, unsigned long long* slot) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    __shared__ double wmin[kBlock / kWave];
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) wmin[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        for (int w = 1; w < kBlock / kWave; ++w) m = wmin[w] < m ? wmin[w] : m;
        if (m < 1e50) atomicMin(slot, dbits(m));
    }
}

// Standalone estimate over every particle (nyxpm_est_time_step when no fused
// estimate of the current state is kept).
template <typename G>
__global__ void __launch_bounds__(kBlock) k_est_dt(const BoxSeg* segs, int nseg, const BoxDev* boxes,
                                                   const double* __restrict__ phi, const double* __restrict__ x,
                                                   const double* __restrict__ y, const double* __restrict__ z,
                                                   const double* __restrict__ vx, const double* __restrict__ vy,
                                                   const double* __restrict__ vz, double a, double cfl,
                                                   unsigned long long* dtmin) {
    const BoxSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    double d = 1e50;
    if (q < B.span) {
        const long long p = B.part_off + q;
        const double px = x[p];
        if (!isnan(px)) d = particle_dt_global<G>(B, phi + B.cell_off, px, y[p], z[p], vx[p], vy[p], vz[p], a, cfl);
    }
    block_min_to(d, dtmin + sg.box);
}

// Kick variants of the tile-staged particle kernel.
// Cells of g per thread that the tile-staged kernel keeps in registers to
// write g over the staged phi (tiles of up to kBlock * kKickG g cells).
constexpr int kKickG = 8;
constexpr int kKick = 0;       // moveKick: kick, then the periodic shift of Redistribute
constexpr int kKickDrift = 1;  // moveKickDrift: kick, drift, new tile key
constexpr int kKickPair = 2;   // moveKick of one step followed by moveKickDrift of the next

// Particles the tile-staged kernel hands over to k_kick_deferred: index (bits
// 0-39), box (bits 40-59) and what is left to do (bit 62: the kick-drift half of
// a pair whose kick is already stored; bit 61: only the time-step estimate of
// the stored state; neither: the whole update of the launch).
constexpr long long kDeferPairKD = 1LL << 62;
constexpr long long kDeferDtOnly = 1LL << 61;  // update done; only the time-step estimate is left
__device__ __forceinline__ long long defer_entry(long long p, int box) {
    return p | ((long long)box << 40);
}

// Tile-staged gather + kick(-drift).  phi over the tile plus a halo of 2 is
// staged in LDS once per item (rows of the staged box spread over the lanes of a
// wave, so the index arithmetic is per row), and g (the average of the face
// gradients, the three components of a cell side by side) is formed there on the
// (tile+1)^3 cells the CIC stencils of the tile's particles touch.  Same
// arithmetic as k_kick, bit for bit.
//
// kKickPair runs the closing kick of step n (a_prev, a_cur) and the opening
// kick-drift of step n+1 (a_prev2, a_cur2) in one pass: both use g from the same
// phi at the same position.  A particle moved by the periodic shift in between
// has its kick stored and its kick-drift deferred (to be interpolated again at
// the shifted position, as the separate kernels would); so is any particle whose
// cell falls outside the item's tile (binning older than a shift).  Deferral
// keeps those rare paths out of this kernel's registers.
//
// Binning for the next deposit (kick-drifts): with inc, the particles whose new
// tile differs from the item's are listed for k_migrate; otherwise every
// particle gets its key, and the tile counts are summed per block for the
// particles that stay in the item's tile (one atomic), counted individually
// for the others.
//
// EST (closing kicks only): also NyxParticleContainer::estTimestep of the kicked
// particles (cfl, a = a_cur) into dtmin, from the staged g; a compile-time
// variant so that the kernels without it keep their register budget.
template <int MODE, bool EST, typename G>
__global__ void __launch_bounds__(kBlock) k_kick_tiled(const BoxSeg* tsegs, int ntseg, const BoxDev* boxes,
                                                       const TileItem* __restrict__ items, const int* nitems,
                                                       const double* __restrict__ phi, double* __restrict__ x,
                                                       double* __restrict__ y, double* __restrict__ z,
                                                       double* __restrict__ vx, double* __restrict__ vy,
                                                       double* __restrict__ vz, double dt, double a_prev,
                                                       double a_cur, double a_prev2, double a_cur2,
                                                       int* __restrict__ key, int* counts,
                                                       long long* __restrict__ defer, int* ndefer, double cfl,
                                                       unsigned long long* dtmin, int inc,
                                                       long long* __restrict__ moves, int* nmove, int* err) {
    extern __shared__ __align__(16) unsigned char kick_smem[];
    double* sphi = reinterpret_cast<double*>(kick_smem);
    __shared__ int s_same[kBlock / kWave];
    // items are in (roughly) tile order; giving each XCD a contiguous range of them
    // lets neighbouring tiles share the phi halo lines in that XCD's L2
    const int n_items = *nitems;
    if ((int)blockIdx.x >= n_items) return;
    const TileItem it = items[xcd_block(blockIdx.x, n_items)];
    const int sgi = find_seg(tsegs, ntseg, it.tile);
    const int box = __builtin_amdgcn_readfirstlane(tsegs[sgi].box);
    // the box's fields as block-uniform values in registers: read through a
    // reference they would be re-read from memory (and waited for, behind the
    // particle stores) after every store or atomic of the particle loop
    const BoxDev B = boxes[box];
    const int lt = it.tile - tsegs[sgi].blk0;
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int tx = lt % B.ntile[0], ty = (lt / B.ntile[0]) % B.ntile[1], tz = lt / (B.ntile[0] * B.ntile[1]);
    const int o0 = tx * B.tile[0], o1 = ty * B.tile[1], o2 = tz * B.tile[2];
    const int bx = min(B.tile[0], nx - o0), by = min(B.tile[1], ny - o1), bz = min(B.tile[2], nz - o2);
    const int SX = bx + 3, SY = by + 3, SZ = bz + 3;
    const double* P = phi + B.cell_off;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    // the item's particles through block-uniform base pointers and 32-bit
    // indices (scalar base + vector offset addressing)
    double* __restrict__ X = x + it.p0;
    double* __restrict__ Y = y + it.p0;
    double* __restrict__ Z = z + it.p0;
    double* __restrict__ VX = vx + it.p0;
    double* __restrict__ VY = vy + it.p0;
    double* __restrict__ VZ = vz + it.p0;
    int* __restrict__ KEY = key + it.p0;
    // one particle per thread at a time, the first loaded before the staging of
    // phi so that its latency overlaps the staging.  (Two per thread in flight
    // need 12-27 more VGPRs, one wave per SIMD fewer: 2% slower in mixed
    // precision, 14% in FP64.)
    double ax = 0.0, ay = 0.0, az = 0.0, aux = 0.0, auy = 0.0, auz = 0.0;
    auto load1 = [&](int q) {
        ax = X[q]; ay = Y[q]; az = Z[q];
        aux = VX[q]; auy = VY[q]; auz = VZ[q];
    };
    if ((int)threadIdx.x < it.np) load1(threadIdx.x);
    {
        // rows of SX cells: a wave takes rpw rows per pass (SX <= 35 < 64).  The
        // passes go in batches of kStageBatch: every load of a batch is issued
        // before any of them is stored, so their latencies overlap (one load,
        // wait, store per pass left 8 serial latencies per lane and tile).  Rows
        // past the last are clamped onto it (the extra lanes store the owner's
        // value to the same LDS word), so no branch separates the loads.
        constexpr int kStageBatch = 8;
        const int rpw = kWave / SX;
        const int lrow = lane / SX, lx = lane - lrow * SX;
        if (lrow < rpw) {
            const size_t gx = (size_t)wrap1(o0 - 2 + lx, nx);
            const int step = (kBlock / kWave) * rpw;
            const int r0 = wave * rpw + lrow;
            int ly = r0 % SY, lz = r0 / SY;
            const int dly = step % SY, dlz = step / SY;
            while (lz < SZ) {
                double v[kStageBatch];
                int o[kStageBatch];
#pragma unroll
                for (int u = 0; u < kStageBatch; ++u) {
                    const bool in = lz < SZ;
                    const int cy = in ? ly : SY - 1, cz = in ? lz : SZ - 1;
                    const int gy = wrap1(o1 - 2 + cy, ny), gz = wrap1(o2 - 2 + cz, nz);
                    v[u] = P[((size_t)gz * ny + gy) * nx + gx];
                    o[u] = (cz * SY + cy) * SX + lx;
                    ly += dly;
                    lz += dlz;
                    if (ly >= SY) {
                        ly -= SY;
                        ++lz;
                    }
                }
#pragma unroll
                for (int u = 0; u < kStageBatch; ++u) sphi[o[u]] = v[u];
            }
        }
    }
    __syncthreads();
    const double dxi0 = B.dxi[0], dxi1 = B.dxi[1], dxi2 = B.dxi[2];
    // g at the cells the tile's particles interpolate from (local -1 .. b-1 per
    // dimension), components of a cell adjacent; formed in FP64, kept as G.
    // Where every thread's share fits in registers (kKickG cells) g is formed
    // there and, after a barrier, written over the staged phi, which it no
    // longer needs: the block holds the larger of the two arrays instead of both
    // (more blocks per CU); otherwise g goes beside the staged phi.
    const int GX = bx + 1, GY = by + 1, GZ = bz + 1;
    const int ng = GX * GY * GZ;
    const bool alias = ng <= kBlock * kKickG;
    G* sg = reinterpret_cast<G*>(alias ? sphi : sphi + SX * SY * SZ);
    if (alias) {
        G gv[kKickG][3];
        // cell index q = tid + u * kBlock, as (cx, cy, cz), advanced incrementally
        int cx = threadIdx.x % GX, cy = (threadIdx.x / GX) % GY, cz = threadIdx.x / (GX * GY);
        const int dcx = kBlock % GX, dcy = (kBlock / GX) % GY, dcz = kBlock / (GX * GY);
#pragma unroll
        for (int u = 0; u < kKickG; ++u) {
            if ((int)threadIdx.x + u * kBlock < ng) {
                const int o = ((cz + 1) * SY + (cy + 1)) * SX + (cx + 1);  // staged index of the cell
                const double pc = sphi[o];
                gv[u][0] = G(0.5 * (dxi0 * (pc - sphi[o - 1]) + dxi0 * (sphi[o + 1] - pc)));
                gv[u][1] = G(0.5 * (dxi1 * (pc - sphi[o - SX]) + dxi1 * (sphi[o + SX] - pc)));
                gv[u][2] = G(0.5 * (dxi2 * (pc - sphi[o - SX * SY]) + dxi2 * (sphi[o + SX * SY] - pc)));
            }
            cx += dcx;
            cy += dcy;
            cz += dcz;
            if (cx >= GX) {
                cx -= GX;
                ++cy;
            }
            if (cy >= GY) {
                cy -= GY;
                ++cz;
            }
        }
        __syncthreads();  // every read of the staged phi is done
#pragma unroll
        for (int u = 0; u < kKickG; ++u) {
            const int q = (int)threadIdx.x + u * kBlock;
            if (q < ng) {
                sg[3 * q] = gv[u][0];
                sg[3 * q + 1] = gv[u][1];
                sg[3 * q + 2] = gv[u][2];
            }
        }
    } else {
        const int rpw = kWave / GX;
        const int lrow = lane / GX, cx = lane - lrow * GX;
        if (lrow < rpw) {
            const int step = (kBlock / kWave) * rpw;
            const int r0 = wave * rpw + lrow;
            int cy = r0 % GY, cz = r0 / GY;
            const int dcy = step % GY, dcz = step / GY;
            while (cz < GZ) {
                const int o = ((cz + 1) * SY + (cy + 1)) * SX + (cx + 1);  // staged index of the cell
                const double pc = sphi[o];
                G* g = sg + 3 * ((cz * GY + cy) * GX + cx);
                g[0] = G(0.5 * (dxi0 * (pc - sphi[o - 1]) + dxi0 * (sphi[o + 1] - pc)));
                g[1] = G(0.5 * (dxi1 * (pc - sphi[o - SX]) + dxi1 * (sphi[o + SX] - pc)));
                g[2] = G(0.5 * (dxi2 * (pc - sphi[o - SX * SY]) + dxi2 * (sphi[o + SX * SY] - pc)));
                cy += dcy;
                cz += dcz;
                if (cy >= GY) {
                    cy -= GY;
                    ++cz;
                }
            }
        }
    }
    __syncthreads();
    const double half_dt = 0.5 * dt;
    const double a_cur_inv = 1.0 / a_cur;
    const double dt_a_cur_inv = dt * a_cur_inv;
    const double a_cur2_inv = 1.0 / a_cur2;
    const double dt_a_cur2_inv = dt * a_cur2_inv;
    int same = 0;
    // time-step estimate of the state after the (closing) kick, a = a_cur
    constexpr bool est = EST && MODE != kKickDrift;
    const double adxi0 = dxi0 / a_cur, adxi1 = dxi1 / a_cur, adxi2 = dxi2 / a_cur;
    double dtmin_t = 1e50;
    // one particle: position (px, py, pz), velocity (qx, qy, qz), index q in the item
    auto one = [&](int q, double px, double py, double pz, double qx, double qy, double qz) {
        if (isnan(px)) return;  // empty slot
        double xint, yint, zint;
        const int i = cic_index(px, B.plo[0], dxi0, &xint);
        const int j = cic_index(py, B.plo[1], dxi1, &yint);
        const int k = cic_index(pz, B.plo[2], dxi2, &zint);
        const int li = wrap1(i, nx) - o0, lj = wrap1(j, ny) - o1, lk = wrap1(k, nz) - o2;
        if (li >= 0 && li < bx && lj >= 0 && lj < by && lk >= 0 && lk < bz) {
            // val_d = sum_{kk,jj,ii} sx[1-ii]*sy[1-jj]*sz[1-kk] * g_d(i-ii, j-jj, k-kk);
            // cell i-ii has g index li+1-ii
            const G fx = G(xint), fy = G(yint), fz = G(zint);
            const G mx = G(1) - fx, my = G(1) - fy, mz = G(1) - fz;
            G g0 = G(0), g1 = G(0), g2 = G(0);
#pragma unroll
            for (int kk = 0; kk <= 1; ++kk) {
                const G wz = kk ? mz : fz;
#pragma unroll
                for (int jj = 0; jj <= 1; ++jj) {
                    const G wy = jj ? my : fy;
                    const G* g = sg + 3 * (((lk + 1 - kk) * GY + (lj + 1 - jj)) * GX + li);
                    const G w0 = fx * wy * wz;  // ii = 0: cell i
                    g0 += w0 * g[3];
                    g1 += w0 * g[4];
                    g2 += w0 * g[5];
                    const G w1 = mx * wy * wz;  // ii = 1: cell i-1
                    g0 += w1 * g[0];
                    g1 += w1 * g[1];
                    g2 += w1 * g[2];
                }
            }
            const double v0 = double(g0), v1 = double(g1), v2 = double(g2);
            double ux = a_prev * qx + half_dt * v0;
            ux = ux * a_cur_inv;
            double uy = a_prev * qy + half_dt * v1;
            uy = uy * a_cur_inv;
            double uz = a_prev * qz + half_dt * v2;
            uz = uz * a_cur_inv;
            bool stay = true;  // the update of this launch completes here
            if (MODE == kKickDrift) {
                px = px + dt_a_cur_inv * ux;
                py = py + dt_a_cur_inv * uy;
                pz = pz + dt_a_cur_inv * uz;
            } else {
                // periodic_shift changes exactly the coordinates outside [plo, rhi]
                // (and flags the non-finite ones): the shifted values are formed
                // only on that rare path, so they hold no registers on the other
                const bool moved = !(px >= B.plo[0] && px <= B.rhi[0]) || !(py >= B.plo[1] && py <= B.rhi[1]) ||
                                   !(pz >= B.plo[2] && pz <= B.rhi[2]);
                if (est && !moved) {
                    // cell = floor((x - plo) dxi) is i or i-1 of the CIC index: in the staged g
                    const int ci = (int)floor((px - B.plo[0]) * dxi0) - (i - (li + o0));
                    const int cj = (int)floor((py - B.plo[1]) * dxi1) - (j - (lj + o1));
                    const int ck = (int)floor((pz - B.plo[2]) * dxi2) - (k - (lk + o2));
                    const G* gc = sg + 3 * (((ck - o2 + 1) * GY + (cj - o1 + 1)) * GX + (ci - o0 + 1));
                    const double d = particle_dt(ux, uy, uz, adxi0, adxi1, adxi2, double(gc[0]), double(gc[1]),
                                                 double(gc[2]), dxi0, cfl);
                    dtmin_t = d < dtmin_t ? d : dtmin_t;
                }
                if (moved) {
                    const double sxp = periodic_shift(px, B.plo[0], B.rhi[0], B.len[0], err);
                    const double syp = periodic_shift(py, B.plo[1], B.rhi[1], B.len[1], err);
                    const double szp = periodic_shift(pz, B.plo[2], B.rhi[2], B.len[2], err);
                    if (MODE == kKick) {
                        if (sxp != px) X[q] = sxp;
                        if (syp != py) Y[q] = syp;
                        if (szp != pz) Z[q] = szp;
                        if (est) defer[atomicAdd(ndefer, 1)] = defer_entry(it.p0 + q, box) | kDeferDtOnly;
                    } else {
                        // store the kick; the kick-drift follows in k_kick_deferred
                        X[q] = sxp;
                        Y[q] = syp;
                        Z[q] = szp;
                        stay = false;
                        defer[atomicAdd(ndefer, 1)] = defer_entry(it.p0 + q, box) | kDeferPairKD;
                    }
                } else if (MODE != kKick) {
                    double wx = a_prev2 * ux + half_dt * v0;
                    ux = wx * a_cur2_inv;
                    double wy = a_prev2 * uy + half_dt * v1;
                    uy = wy * a_cur2_inv;
                    double wz = a_prev2 * uz + half_dt * v2;
                    uz = wz * a_cur2_inv;
                    px = px + dt_a_cur2_inv * ux;
                    py = py + dt_a_cur2_inv * uy;
                    pz = pz + dt_a_cur2_inv * uz;
                }
            }
            // streaming stores (each particle is written once per pass; the
            // L2 keeps the phi lines the neighbouring tiles stage)
            __builtin_nontemporal_store(ux, &VX[q]);
            __builtin_nontemporal_store(uy, &VY[q]);
            __builtin_nontemporal_store(uz, &VZ[q]);
            if (MODE != kKick && stay) {
                __builtin_nontemporal_store(px, &X[q]);
                __builtin_nontemporal_store(py, &Y[q]);
                __builtin_nontemporal_store(pz, &Z[q]);
                const int t = tile_of(B, px, py, pz, err);
                if (inc) {
                    // incremental re-sort: only the particles that left the tile
                    wave_append(moves, nmove, defer_entry(it.p0 + q, box), t != it.tile);
                } else {
                    KEY[q] = t;
                    if (t == it.tile)
                        ++same;
                    else
                        atomicAdd(&counts[t], 1);
                }
            }
        } else {
            // binned in another tile: the whole update in k_kick_deferred
            defer[atomicAdd(ndefer, 1)] = defer_entry(it.p0 + q, box);
        }
    };
    for (int q = threadIdx.x; q < it.np; q += kBlock) {
        if (q != (int)threadIdx.x) load1(q);
        one(q, ax, ay, az, aux, auy, auz);
    }
    if (est) block_min_to(dtmin_t, dtmin + box);
    if (MODE != kKick) {
        for (int o = kWave / 2; o > 0; o >>= 1) same += __shfl_xor(same, o);
        if (lane == 0) s_same[wave] = same;
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int w = 0; w < kBlock / kWave; ++w) tot += s_same[w];
            if (tot) atomicAdd(&counts[it.tile], tot);
        }
    }
}

// The particles k_kick_tiled deferred, one thread each, phi read from global
// memory (gather_grad: same arithmetic as the staged path).
template <int MODE, typename G>
__global__ void __launch_bounds__(kBlock) k_kick_deferred(const long long* __restrict__ defer, const int* ndefer,
                                                          const BoxDev* boxes, const double* __restrict__ phi,
                                                          double* __restrict__ x, double* __restrict__ y,
                                                          double* __restrict__ z, double* __restrict__ vx,
                                                          double* __restrict__ vy, double* __restrict__ vz,
                                                          double dt, double a_prev, double a_cur, double a_prev2,
                                                          double a_cur2, int* __restrict__ key, int* counts,
                                                          double cfl, unsigned long long* dtmin, int inc,
                                                          long long* __restrict__ moves, int* nmove, int* err) {
    const int n = *ndefer;
    const bool est = MODE != kKickDrift && cfl > 0;
    const double half_dt = 0.5 * dt;
    // binning of a drifted particle (k_kick_tiled's rule; k_migrate finds out
    // whether a listed particle changed tile)
    auto bin = [&](long long p, int box, const BoxDev& B, double px, double py, double pz) {
        if (inc) {
            wave_append(moves, nmove, defer_entry(p, box), true);
        } else {
            const int t = tile_of(B, px, py, pz, err);
            key[p] = t;
            atomicAdd(&counts[t], 1);
        }
    };
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) {
        const long long e = defer[q];
        const long long p = e & ((1LL << 40) - 1);
        const int box = (int)((e >> 40) & 0xFFFFF);
        const BoxDev& B = boxes[box];
        const double* P = phi + B.cell_off;
        double px = x[p], py = y[p], pz = z[p];
        double ux = vx[p], uy = vy[p], uz = vz[p];
        double val[3];
        if (e & kDeferDtOnly) {
            atomicMin(dtmin + box, dbits(particle_dt_global<G>(B, P, px, py, pz, ux, uy, uz, a_cur, cfl)));
            continue;
        }
        if (!(e & kDeferPairKD)) {
            const double a_cur_inv = 1.0 / a_cur;
            gather_grad<G>(B, P, px, py, pz, val, err);
            ux = a_prev * ux + half_dt * val[0];
            ux = ux * a_cur_inv;
            uy = a_prev * uy + half_dt * val[1];
            uy = uy * a_cur_inv;
            uz = a_prev * uz + half_dt * val[2];
            uz = uz * a_cur_inv;
            if (MODE == kKickDrift) {
                const double dt_a_cur_inv = dt * a_cur_inv;
                px = px + dt_a_cur_inv * ux;
                py = py + dt_a_cur_inv * uy;
                pz = pz + dt_a_cur_inv * uz;
                vx[p] = ux; vy[p] = uy; vz[p] = uz;
                x[p] = px; y[p] = py; z[p] = pz;
                bin(p, box, B, px, py, pz);
                continue;
            }
            const double sxp = periodic_shift(px, B.plo[0], B.rhi[0], B.len[0], err);
            const double syp = periodic_shift(py, B.plo[1], B.rhi[1], B.len[1], err);
            const double szp = periodic_shift(pz, B.plo[2], B.rhi[2], B.len[2], err);
            if (est)
                atomicMin(dtmin + box, dbits(particle_dt_global<G>(B, P, sxp, syp, szp, ux, uy, uz, a_cur, cfl)));
            if (MODE == kKick) {
                vx[p] = ux; vy[p] = uy; vz[p] = uz;
                if (sxp != px) x[p] = sxp;
                if (syp != py) y[p] = syp;
                if (szp != pz) z[p] = szp;
                continue;
            }
            const bool moved = sxp != px || syp != py || szp != pz;
            px = sxp;
            py = syp;
            pz = szp;
            if (moved) gather_grad<G>(B, P, px, py, pz, val, err);
        } else {
            // pair whose kick is stored: its estimate, then the kick-drift at the
            // shifted position
            if (est) atomicMin(dtmin + box, dbits(particle_dt_global<G>(B, P, px, py, pz, ux, uy, uz, a_cur, cfl)));
            gather_grad<G>(B, P, px, py, pz, val, err);
        }
        const double a_cur2_inv = 1.0 / a_cur2;
        const double dt_a_cur2_inv = dt * a_cur2_inv;
        double wx = a_prev2 * ux + half_dt * val[0];
        ux = wx * a_cur2_inv;
        double wy = a_prev2 * uy + half_dt * val[1];
        uy = wy * a_cur2_inv;
        double wz = a_prev2 * uz + half_dt * val[2];
        uz = wz * a_cur2_inv;
        px = px + dt_a_cur2_inv * ux;
        py = py + dt_a_cur2_inv * uy;
        pz = pz + dt_a_cur2_inv * uz;
        vx[p] = ux; vy[p] = uy; vz[p] = uz;
        x[p] = px; y[p] = py; z[p] = pz;
        bin(p, box, B, px, py, pz);
    }
}

// Incremental re-sort.  Each listed particle whose position now lies in another
// tile than the one whose slot range holds it is copied into the free slack of
// its new tile (slot thw[t]++), and its old slot is emptied (x = NaN).  A tile
// with no free slot sets flags[0]: the particle stays where it is and the
// caller sorts in full.  flags[1] counts the particles moved (= slots emptied).
// Destinations lie at or above every tile's high-water mark at launch, so no
// destination is a listed source.
__global__ void __launch_bounds__(kBlock) k_migrate(const long long* __restrict__ moves, const int* nmove,
                                                    const BoxDev* boxes, const long long* __restrict__ tbase,
                                                    const int* __restrict__ tcap, int* thw, PartPtrs P, int* flags,
                                                    int* err) {
    const int n = *nmove;
    int moved = 0;
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) {
        const long long e = moves[q];
        const long long p = e & ((1LL << 40) - 1);
        const BoxDev& B = boxes[(int)((e >> 40) & 0xFFFFF)];
        const double px = P.x[p];
        if (isnan(px)) continue;
        const double py = P.y[p], pz = P.z[p];
        const int t = tile_of(B, px, py, pz, err);
        // the tile holding slot p: the last of the box whose range starts at or before p
        int lo = (int)B.tile_off, hi = (int)B.tile_off + B.ntile[0] * B.ntile[1] * B.ntile[2] - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tbase[mid] <= p) lo = mid; else hi = mid - 1;
        }
        if (lo == t) continue;
        const int slot = atomicAdd(&thw[t], 1);
        if (slot >= tcap[t]) {
            atomicOr(&flags[0], 1);
            continue;
        }
        const long long d = tbase[t] + slot;
        P.x[d] = px;
        P.y[d] = py;
        P.z[d] = pz;
        P.vx[d] = P.vx[p];
        P.vy[d] = P.vy[p];
        P.vz[d] = P.vz[p];
        if (P.m) P.m[d] = P.m[p];
        P.id[d] = P.id[p];
        P.x[p] = empty_slot();
        ++moved;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) moved += __shfl_xor(moved, o);
    if ((threadIdx.x & (kWave - 1)) == 0 && moved) atomicAdd(&flags[1], moved);
}

// Periodic shift alone (nyxpm_redistribute).
__global__ void __launch_bounds__(kBlock) k_wrap(const BoxSeg* segs, int nseg, const BoxDev* boxes,
                                                 double* __restrict__ x, double* __restrict__ y,
                                                 double* __restrict__ z, int* err) {
    const BoxSeg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    if (q >= B.span) return;
    const long long p = B.part_off + q;
    if (isnan(x[p])) return;
    x[p] = periodic_shift(x[p], B.plo[0], B.rhi[0], B.len[0], err);
    y[p] = periodic_shift(y[p], B.plo[1], B.rhi[1], B.len[1], err);
    z[p] = periodic_shift(z[p], B.plo[2], B.rhi[2], B.len[2], err);
}

// Cell-centred gravity field of one box (get_new_grav_vector), 3 components.
__global__ void __launch_bounds__(kBlock) k_grav_vector(const BoxDev* boxes, int box, const double* __restrict__ phi,
                                                        double* __restrict__ g) {
    const BoxDev B = boxes[box];
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const long long nc = (long long)nx * ny * nz;
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= nc) return;
    const int i = (int)(q % nx);
    const long long t = q / nx;
    const int j = (int)(t % ny), k = (int)(t / ny);
    const double* P = phi + B.cell_off;
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
    const double pc = P[q];
    const double fxl = B.dxi[0] * (pc - P[cidx(im, j, k, nx, ny)]);
    const double fxh = B.dxi[0] * (P[cidx(ip, j, k, nx, ny)] - pc);
    const double fyl = B.dxi[1] * (pc - P[cidx(i, jm, k, nx, ny)]);
    const double fyh = B.dxi[1] * (P[cidx(i, jp, k, nx, ny)] - pc);
    const double fzl = B.dxi[2] * (pc - P[cidx(i, j, km, nx, ny)]);
    const double fzh = B.dxi[2] * (P[cidx(i, j, kp, nx, ny)] - pc);
    g[q] = 0.5 * (fxl + fxh);
    g[nc + q] = 0.5 * (fyl + fyh);
    g[2 * nc + q] = 0.5 * (fzl + fzh);
}

// out[id] = in[p] for the occupied slots of one box (undo the tile order for
// download).
__global__ void __launch_bounds__(kBlock) k_unsort(long long p0, long long span, const double* __restrict__ x,
                                                   const int* __restrict__ id, const double* __restrict__ in,
                                                   double* __restrict__ out) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= span || isnan(x[p0 + q])) return;
    out[id[p0 + q]] = in[p0 + q];
}

__global__ void __launch_bounds__(kBlock) k_iota(long long p0, long long np, int* __restrict__ id) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < np) id[p0 + q] = (int)q;
}

__global__ void __launch_bounds__(kBlock) k_fill_int(int* a, long long n, int v) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < n) a[q] = v;
}

// Initial guess of a step's solve: *ok ? phi = 2 phi - prev : phi unchanged;
// prev = the old phi either way.
__global__ void __launch_bounds__(kBlock) k_phi_guess(double* __restrict__ phi, double* __restrict__ prev,
                                                      long long n, const int* ok) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= n) return;
    const double p = phi[q];
    if (*ok) phi[q] = 2.0 * p - prev[q];
    prev[q] = p;
}

__global__ void __launch_bounds__(kBlock) k_fill_dbl(double* a, long long n, double v) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < n) a[q] = v;
}

// f = rho + (-mass_offset) with per-block sums (used when the density was set
// from outside rather than by the fold).  Cells are visited exactly as k_fold
// visits them (same cells per block and per thread, same order), so the
// solvability sum of a density does not depend on which of the two produced f.
__global__ void __launch_bounds__(kBlock) k_rhs1(const CellSeg* segs, int nseg, const BoxDev* boxes,
                                                 const double* __restrict__ rho, double* __restrict__ f,
                                                 double* __restrict__ partial) {
    const CellSeg& sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const BoxDev B = boxes[__builtin_amdgcn_readfirstlane(sg.box)];
    double acc = 0.0;
    for_block_cells(sg, B, [&](int, int, int, long long c) {
        const size_t q = B.cell_off + (size_t)c;
        const double fv = rho[q] + (-B.mass_offset);
        f[q] = fv;
        acc += fv;
    });
    const double bs = block_sum_fixed(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// Accelerations at the particles of one box (the val of update_dm_particle_single)
// without touching the particles; written in upload (id) order, component-major.
template <typename G>
__global__ void __launch_bounds__(kBlock) k_gather(const BoxDev* boxes, int box, const double* __restrict__ phi,
                                                   const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, const int* __restrict__ id,
                                                   double* __restrict__ out, int* err) {
    const BoxDev B = boxes[box];
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q >= B.span) return;
    const long long p = B.part_off + q;
    const double px = x[p];
    if (isnan(px)) return;
    double val[3];
    gather_grad<G>(B, phi + B.cell_off, px, y[p], z[p], val, err);
    const long long o = id[p];
    out[o] = val[0];
    out[B.np + o] = val[1];
    out[2 * B.np + o] = val[2];
}

}  // namespace nyxpm
