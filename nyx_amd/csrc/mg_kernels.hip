// mg_kernels.hip -- geometric multigrid for the periodic 7-point Poisson operator
// (the MLMG/MLPoisson V-cycle Nyx calls from Gravity::solve_with_MLMG,
// Source/Gravity/Gravity.cpp:1318-1416), as gfx950 kernels.
//
// Every kernel takes a table of CellSeg pieces (one per box and level it works
// on) and an `active` flag per box, so one launch serves a batch of boxes and a
// box whose solve has converged costs nothing.  The arithmetic of each cell
// update is written in the same order as the CPU restatement so that, with FMA
// contraction off, the FP64 iterates agree bit for bit.
//
// Roofline: all of these are HBM-bound stencils (arithmetic intensity well below
// 1 flop/byte).  Compulsory bytes per fine cell (FP64 / FP32):
//   red-black half sweep  8 (e) + 8 (r) + 8 (e write)   / 4+4+4
//   residual + restrict   8 (e) + 8 (r) + 1 (coarse)    / 4+4+0.5
//   prolong + add         8 (e) + 8 (e write) + 1       / 4+4+0.5
#include "nyxpm_internal.h"
#include <type_traits>

namespace nyxpm {

template <typename T>
__device__ __forceinline__ T lap7(const T* a, int i, int j, int k, int nx, int ny, int nz,
                                  T dhx, T dhy, T dhz) {
    int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
    T c = a[cidx(i, j, k, nx, ny)];
    T lx = dhx * (a[cidx(im, j, k, nx, ny)] - T(2) * c + a[cidx(ip, j, k, nx, ny)]);
    T ly = dhy * (a[cidx(i, jm, k, nx, ny)] - T(2) * c + a[cidx(i, jp, k, nx, ny)]);
    T lz = dhz * (a[cidx(i, j, km, nx, ny)] - T(2) * c + a[cidx(i, j, kp, nx, ny)]);
    return lx + ly + lz;
}

// FP32 Gauss-Seidel update from the scaled residual rs = r * (1/gamma) and the
// scaled coefficients cx = dhx * (1/gamma), ...: gs_update<float>'s arithmetic.
__device__ __forceinline__ float gs_update_scaled(float rs, float sx, float sy, float sz, float cx, float cy,
                                                  float cz) {
    float t = __builtin_fmaf(-cx, sx, rs);
    t = __builtin_fmaf(-cy, sy, t);
    return __builtin_fmaf(-cz, sz, t);
}

// One Gauss-Seidel update of a cell with the periodic 7-point operator:
//   e + (r - gamma*e - dhx*sx - dhy*sy - dhz*sz) / gamma
// where sx, sy, sz are the sums of the cell's two neighbours along x, y, z.
// FP64: the operations in the written order, unfused, as the CPU restatement
// does them (Source/Gravity/Gravity.cpp:1318 hands the solve to AMReX's MLMG,
// whose red-black smoother this is).  FP32 (the correction V-cycles of mixed
// precision, which no test compares bit for bit with the CPU): the same value
// with gamma * (1/gamma) taken as 1, r/gamma - (dhx/gamma) sx - (dhy/gamma) sy -
// (dhz/gamma) sz, as one product and three fused multiply-adds after the sums
// (the cell's own value drops out).
template <typename T>
__device__ __forceinline__ T gs_update(T ec, T rr, T sx, T sy, T sz, T gamma, T inv_gamma, T dhx, T dhy,
                                       T dhz) {
    if constexpr (std::is_same<T, float>::value) {
        return gs_update_scaled(rr * inv_gamma, sx, sy, sz, dhx * inv_gamma, dhy * inv_gamma, dhz * inv_gamma);
    } else {
        const T res = rr - gamma * ec - dhx * sx - dhy * sy - dhz * sz;
        return ec + res * inv_gamma;
    }
}

// a[c] = v on every cell of each piece.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill(const CellSeg* segs, int nseg, const int* active, T* a, T v) {
    const int s = find_seg(segs, nseg, blockIdx.x);
    const CellSeg sg = segs[s];
    if (active && !active[sg.box]) return;
    const long long n = (long long)sg.n[0] * sg.n[1] * sg.n[2];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    if (q < n) a[sg.off + q] = v;
}

// One red-black Gauss-Seidel half sweep (cells with (i+j+k+colour) even):
//   res = r - gamma*e - dhx*(e(i-1)+e(i+1)) - dhy*(..) - dhz*(..);  e += res/gamma
// One thread per cell of the colour.  Needs even nx (true above the bottom).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gsrb_half(const CellSeg* segs, int nseg, const int* active,
                                                      T* __restrict__ e, const T* __restrict__ r, int colour) {
    const int s = find_seg(segs, nseg, blockIdx.x);
    const CellSeg& sg = segs[s];
    if (!active[sg.box]) return;
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const int hx = nx >> 1;
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    if (q >= (long long)hx * ny * nz) return;
    const int ih = (int)(q % hx);
    const long long t = q / hx;
    const int j = (int)(t % ny), k = (int)(t / ny);
    const int i = 2 * ih + ((j + k + colour) & 1);
    const T dhx = T(sg.dh[0]), dhy = T(sg.dh[1]), dhz = T(sg.dh[2]);
    const T gamma = T(sg.gamma), inv_gamma = T(sg.inv_gamma);
    T* E = e + sg.off;
    const T* R = r + sg.off;
    int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
    const size_t c = cidx(i, j, k, nx, ny);
    E[c] = gs_update(E[c], R[c], E[cidx(im, j, k, nx, ny)] + E[cidx(ip, j, k, nx, ny)],
                     E[cidx(i, jm, k, nx, ny)] + E[cidx(i, jp, k, nx, ny)],
                     E[cidx(i, j, km, nx, ny)] + E[cidx(i, j, kp, nx, ny)], gamma, inv_gamma, dhx, dhy, dhz);
}

// One full red-black Gauss-Seidel sweep (red half, then black half) in a single
// pass over the level: the block owns a TX x TY column of cells and a chunk of
// z-planes, and marches up in z keeping a ring of 4 planes of e (halo 2 in x, y)
// and 2 planes of r (halo 1) in LDS.  Per plane kr: red cells of plane kr are
// updated in LDS (including a one-cell ring around the column, recomputed
// redundantly), then the black cells of plane kr-1 are updated from the new red
// values and the plane's final values are written out.  Reads ein, writes eout (a
// different buffer), so neighbouring blocks never see half-updated values; the
// arithmetic per cell is exactly that of k_gsrb_half, so the result equals two
// half sweeps bit for bit.
// MODE bits:
//   kSweepZero    ein is identically zero (first sweep after the correction is
//                 reset), so it is not read;
//   kSweepProlong the input is ein + (coarse correction of the parent cell), the
//                 piecewise-constant prolongation of k_prolong done on load;
//   kSweepUpdate  the result is not stored: phi += result (level 0, FP64 phi at
//                 sg.foff), the update of k_update0 done on store;
//   kSweepResidual (with kSweepZero, level 0) r is not read but formed: the
//                 FP64 residual f - L phi of k_residual0 (same arithmetic, the
//                 rhs offset and scaling applied as there), computed per plane on
//                 the halo-1 region from phi planes kept in registers, stored to r
//                 for the block's own cells, max |r| (and max |f|) folded into the
//                 norm slots.  One pass replaces k_residual0 + the first sweep.
// Needs nx % TX == 0, ny % TY == 0 and even nx, ny, nz.  HBM bytes per cell:
// e read + r read + e write (3*sizeof(T)); halo re-reads are served by L2.
constexpr int kSweepZero = 1, kSweepProlong = 2, kSweepUpdate = 4, kSweepResidual = 8;

// Inputs of the residual formed inside the first sweep (kSweepResidual).
struct SweepRes {
    const double* f;           // level-0 rhs (at sg.foff)
    const double* solv;        // non-null: f holds rho - mass_offset (k_residual0's raw form)
    double G, inv_a;
    unsigned long long* nr;    // norm slots of max |r|
    unsigned long long* nf;    // and of max |f| (nullptr: not wanted)
};

template <typename T, int TX, int TY, int MODE>
__global__ void __launch_bounds__(TX * TY / 2) k_rb_sweep(const CellSeg* segs, int nseg, const int* active,
                                                          const T* __restrict__ ein, T* __restrict__ eout,
                                                          const T* __restrict__ r, const T* ec0, const T* ec1,
                                                          double* phi, SweepRes res) {
    constexpr bool ZERO_IN = (MODE & kSweepZero) != 0;
    constexpr bool PRO = (MODE & kSweepProlong) != 0;
    constexpr bool UPD = (MODE & kSweepUpdate) != 0;
    constexpr bool RES = (MODE & kSweepResidual) != 0;
    static_assert(!RES || ZERO_IN, "the residual is formed only in the first sweep");
    constexpr int NT = TX * TY / 2;
    constexpr int EX = TX + 4, EY = TY + 4;   // e ring planes: halo 2
    constexpr int RX = TX + 2, RY = TY + 2;   // r planes and red-update region: halo 1
    constexpr int NE = (EX * EY + NT - 1) / NT, NR = (RX * RY + NT - 1) / NT;
    // Planes are moved as pairs of x-adjacent cells (8-byte FP32 or 16-byte FP64
    // loads and LDS stores, half the instructions; one coarse load per pair in the
    // prolongation).  Pairs start at even x (nx is even; level offsets are 64-cell
    // aligned), so the r planes are kept with halo 2 in x as well (RSH shifts the
    // r index).  FP64 pairs only in the prolongation modes: elsewhere the FP64
    // sweep measured faster with single-cell moves.
    constexpr bool V2 = !RES && (sizeof(T) == 4 || PRO);
    using T2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    constexpr int RXL = V2 ? TX + 4 : RX, RSH = V2 ? 1 : 0;
    constexpr int EXP = EX / 2, RXP = RXL / 2;
    constexpr int NEP = (EXP * EY + NT - 1) / NT, NRP = (RXP * RY + NT - 1) / NT;
    __shared__ __align__(16) T se[4][EY][EX];
    __shared__ __align__(16) T sr[2][RY][RXL];
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const int s = find_seg(segs, nseg, bid);
    const CellSeg& sg = segs[s];
    if (!active[sg.box]) return;
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const int ntx = nx / TX, nty = ny / TY;
    const int zc = sg.zchunk;
    const int b = bid - sg.blk0;
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * zc;
    const int z1 = min(nz, z0 + zc);
    const T dhx = T(sg.dh[0]), dhy = T(sg.dh[1]), dhz = T(sg.dh[2]);
    const T gamma = T(sg.gamma), inv_gamma = T(sg.inv_gamma);
    const T* Ein = ein + sg.off;
    T* Eout = eout + sg.off;
    const T* R = r + sg.off;
    const T* Ec = (sg.ebuf ? ec1 : ec0) + sg.coff;
    const int cnx = nx >> 1, cny = ny >> 1;
    double* Ph = phi + sg.foff;
    const int tid = threadIdx.x;

    // Planes are fetched into registers DEPTH iterations ahead of their use
    // (DEPTH register sets, used by alternate iterations), so the global-memory
    // latency overlaps DEPTH planes' compute.  Every thread
    // loads and stores all of its elements: an element index past the region is
    // clamped onto the region's last element, so the extra lanes load that cell
    // and store the same value to the same LDS word as its owner.  With no
    // branch around them the loads issue back to back and only the LDS stores
    // wait for them.  A thread's in-plane offsets (periodic wrap included) are
    // the same for every plane and are computed once; the coarse values of the
    // prolongation change at even planes: an odd plane takes those of the plane
    // fetched before it (the other register set; fetches go in plane order), and
    // they are added when the plane is stored to LDS.
    // (two sets only where measured faster: the plain and zero-input sweeps; the
    // prolongation and update modes hold more registers per set)
    constexpr int DEPTH = RES || PRO || UPD ? 1 : 2;
    T pe[DEPTH][V2 ? 1 : NE], pr[DEPTH][V2 ? 1 : NR], pc[DEPTH][PRO && !V2 ? NE : 1];
    size_t eoff[V2 ? 1 : NE], coff_[PRO && !V2 ? NE : 1], roff[V2 ? 1 : NR];
    // the two cells of each pair in separate scalar arrays (an array of double2
    // under a conditional load is not split into registers)
    T pex[DEPTH][V2 ? NEP : 1], pey[DEPTH][V2 ? NEP : 1], prx[DEPTH][V2 ? NRP : 1], pry[DEPTH][V2 ? NRP : 1];
    T pc2[DEPTH][V2 && PRO ? NEP : 1];
    size_t eoff2[V2 ? NEP : 1], coff2[V2 && PRO ? NEP : 1], roff2[V2 ? NRP : 1];
    const size_t plane = (size_t)nx * ny, cplane = (size_t)cnx * cny;
    if (V2) {
#pragma unroll
        for (int u = 0; u < NEP; ++u) {
            const int pidx = min(tid + u * NT, EXP * EY - 1);
            const int ly = pidx / EXP, lxp = pidx - ly * EXP;
            int gx = x0 + 2 * lxp - 2, gy = y0 + ly - 2;
            gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
            gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
            eoff2[u] = (size_t)gy * nx + gx;
            if (PRO) coff2[u] = (size_t)(gy >> 1) * cnx + (gx >> 1);
        }
#pragma unroll
        for (int u = 0; u < NRP; ++u) {
            const int pidx = min(tid + u * NT, RXP * RY - 1);
            const int ly = pidx / RXP, lxp = pidx - ly * RXP;
            int gx = x0 + 2 * lxp - 2, gy = y0 + ly - 1;
            gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
            gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
            roff2[u] = (size_t)gy * nx + gx;
        }
    }
#pragma unroll
    for (int u = 0; u < (V2 ? 0 : NE); ++u) {
        const int idx = min(tid + u * NT, EX * EY - 1);
        const int ly = idx / EX, lx = idx - ly * EX;
        int gx = x0 + lx - 2, gy = y0 + ly - 2;
        gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
        gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        eoff[u] = (size_t)gy * nx + gx;
        if (PRO) coff_[u] = (size_t)(gy >> 1) * cnx + (gx >> 1);
    }
    // RES: in-plane position of each r element and whether it is one of the
    // block's own cells (stored, counted in the norms)
    int rgx[RES ? NR : 1], rgy[RES ? NR : 1];
    bool rown[RES ? NR : 1];
#pragma unroll
    for (int u = 0; u < (V2 ? 0 : NR); ++u) {
        const int idx = min(tid + u * NT, RX * RY - 1);
        const int ly = idx / RX, lx = idx - ly * RX;
        int gx = x0 + lx - 1, gy = y0 + ly - 1;
        gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
        gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        roff[u] = (size_t)gy * nx + gx;
        if (RES) {
            rgx[u] = gx;
            rgy[u] = gy;
            rown[u] = tid + u * NT < RX * RY && lx >= 1 && lx <= TX && ly >= 1 && ly <= TY;
        }
    }
    auto wrapz = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };
    // RES: phi of each r element at planes k-1 and k of the next plane k to form
    const double* Pf = phi + sg.foff;
    const double* Ff = RES ? res.f + sg.foff : nullptr;
    const bool raw = RES && res.solv != nullptr;
    const double rns = raw ? -res.solv[sg.box] : 0.0;
    const double rdx = sg.dh[0], rdy = sg.dh[1], rdz = sg.dh[2];
    double wm[RES ? NR : 1], wc[RES ? NR : 1];
    double ar = 0.0, af = 0.0;
    if (RES) {
#pragma unroll
        for (int u = 0; u < NR; ++u) {
            wm[u] = Pf[(size_t)wrapz(z0 - 2) * plane + roff[u]];
            wc[u] = Pf[(size_t)wrapz(z0 - 1) * plane + roff[u]];
        }
    }
    auto fetch_e = [&](int kz, int S) {
        const int kw = wrapz(kz);
        const bool new_coarse = !(kz & 1) || kz == z0 - 2;
        const int O = DEPTH - 1 - S;  // the set of the plane fetched before
        if (V2) {
            const T* Ep = Ein + (size_t)kw * plane;
#pragma unroll
            for (int u = 0; u < NEP; ++u) {
                if (ZERO_IN) {
                    pex[S][u] = T(0);
                    pey[S][u] = T(0);
                } else {
                    const T2 v = *reinterpret_cast<const T2*>(Ep + eoff2[u]);
                    // both cells of a pair have the same parent (added in put_e)
                    if (PRO) pc2[S][u] = new_coarse ? Ec[(size_t)(kw >> 1) * cplane + coff2[u]] : pc2[O][u];
                    pex[S][u] = v.x;
                    pey[S][u] = v.y;
                }
            }
            return;
        }
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            if (ZERO_IN) {
                pe[S][u] = T(0);
            } else {
                pe[S][u] = Ein[(size_t)kw * plane + eoff[u]];
                if (PRO) pc[S][u] = new_coarse ? Ec[(size_t)(kw >> 1) * cplane + coff_[u]] : pc[O][u];
            }
        }
    };
    auto put_e = [&](int kz, int S) {
        // zero input: the ring's input values are never read (the red step uses
        // literal zeros, the black step reads only red values written this
        // sweep), so the zero planes are not stored
        if (ZERO_IN) return;
        T* dst = &se[kz & 3][0][0];
        if (V2) {
            T2* d2 = reinterpret_cast<T2*>(dst);
#pragma unroll
            for (int u = 0; u < NEP; ++u) {
                T2 v;
                v.x = PRO ? pex[S][u] + pc2[S][u] : pex[S][u];
                v.y = PRO ? pey[S][u] + pc2[S][u] : pey[S][u];
                d2[min(tid + u * NT, EXP * EY - 1)] = v;
            }
            return;
        }
#pragma unroll
        for (int u = 0; u < NE; ++u) dst[min(tid + u * NT, EX * EY - 1)] = PRO ? pe[S][u] + pc[S][u] : pe[S][u];
    };
    auto fetch_r = [&](int kz, int S) {
        const int kw = wrapz(kz);
        if (RES) {
            // k_residual0's arithmetic on the planes kz-1, kz, kz+1 (called for
            // kz = z0-1, z0, ..., z1 in order)
            const size_t zo = (size_t)kw * plane, zp = (size_t)wrapz(kz + 1) * plane;
            const bool own_plane = kz >= z0 && kz < z1;
#pragma unroll
            for (int u = 0; u < NR; ++u) {
                const int gx = rgx[u], gy = rgy[u];
                const int xm = gx == 0 ? nx - 1 : gx - 1, xp = gx == nx - 1 ? 0 : gx + 1;
                const int ym = gy == 0 ? ny - 1 : gy - 1, yp = gy == ny - 1 ? 0 : gy + 1;
                const size_t row = zo + (size_t)gy * nx;
                const double pm = wm[u], pc = wc[u];
                const double pp = Pf[zp + roff[u]];
                const double lx = rdx * (Pf[row + xm] - 2.0 * pc + Pf[row + xp]);
                const double ly = rdy * (Pf[zo + (size_t)ym * nx + gx] - 2.0 * pc + Pf[zo + (size_t)yp * nx + gx]);
                const double lz = rdz * (pm - 2.0 * pc + pp);
                double fv = Ff[zo + roff[u]];
                if (raw) {
                    const double t = fv + rns;
                    fv = (t * res.G) * res.inv_a;
                }
                const double v = fv - (lx + ly + lz);
                pr[S][u] = T(v);
                if (own_plane && rown[u]) {
                    const_cast<T*>(R)[zo + roff[u]] = T(v);
                    ar = fmax(ar, fabs(v));
                    af = fmax(af, fabs(fv));
                }
                wm[u] = pc;
                wc[u] = pp;
            }
            return;
        }
        if (V2) {
            const T* Rp = R + (size_t)kw * plane;
#pragma unroll
            for (int u = 0; u < NRP; ++u)
                {
                    const T2 v = *reinterpret_cast<const T2*>(Rp + roff2[u]);
                    prx[S][u] = v.x;
                    pry[S][u] = v.y;
                }
            return;
        }
#pragma unroll
        for (int u = 0; u < NR; ++u) pr[S][u] = R[(size_t)kw * plane + roff[u]];
    };
    auto put_r = [&](int kz, int S) {
        T* dst = &sr[kz & 1][0][0];
        if (V2) {
            T2* d2 = reinterpret_cast<T2*>(dst);
#pragma unroll
            for (int u = 0; u < NRP; ++u) {
                T2 v;
                v.x = prx[S][u];
                v.y = pry[S][u];
                d2[min(tid + u * NT, RXP * RY - 1)] = v;
            }
            return;
        }
#pragma unroll
        for (int u = 0; u < NR; ++u) dst[min(tid + u * NT, RX * RY - 1)] = pr[S][u];
    };

    // iteration kr: stores e plane kr+1 and r plane kr (fetched DEPTH iterations
    // earlier into set S), refills set S, updates the red cells of plane kr and
    // the black cells of plane kr-1
    auto step = [&](int kr, int S) {
        put_e(kr + 1, S);
        put_r(kr, S);
        // No branch around the prefetch loads (the last iterations re-load planes
        // z1+1 and z1, already loaded), so the waits before the LDS stores and the
        // black step's store count only what they need.
        if (RES) {   // fetch_r forms and stores r: exactly once per plane
            if (kr < z1) {
                fetch_e(kr + 2, S);
                fetch_r(kr + 1, S);
            }
        } else {
            const int kf = min(kr + 1 + DEPTH, z1 + 1);
            fetch_e(kf, S);
            fetch_r(kf - 1, S);
        }
        __syncthreads();
        {   // red cells of plane kr on the halo-1 region: (i+j+k) even, which is
            // consistent across the periodic seams because every n is even
            const int sm = (kr - 1) & 3, sc = kr & 3, sp = (kr + 1) & 3, rs = kr & 1;
            constexpr int HX = RX / 2;  // red cells per row of the region
            for (int idx = tid; idx < HX * RY; idx += NT) {
                const int ly = idx / HX, mm = idx - ly * HX;
                const int lx = 2 * mm + ((x0 + y0 + ly + kr) & 1);  // gx + gy + kr even
                const int ex = lx + 1, ey = ly + 1;
                // with zero input every operand but r is an input zero (the
                // black cells of planes kr-1, kr, kr+1 are not yet updated):
                // the same operations on literal zeros, the same bits
                const T ec = ZERO_IN ? T(0) : se[sc][ey][ex];
                se[sc][ey][ex] =
                    gs_update(ec, sr[rs][ly][lx + RSH],
                              ZERO_IN ? T(0) + T(0) : se[sc][ey][ex - 1] + se[sc][ey][ex + 1],
                              ZERO_IN ? T(0) + T(0) : se[sc][ey - 1][ex] + se[sc][ey + 1][ex],
                              ZERO_IN ? T(0) + T(0) : se[sm][ey][ex] + se[sp][ey][ex], gamma, inv_gamma, dhx, dhy, dhz);
            }
        }
        __syncthreads();
        const int kb = kr - 1;
        if (kb >= z0) {   // black cells of plane kb on the interior, write plane kb
            const int sm = (kb - 1) & 3, sc = kb & 3, sp = (kb + 1) & 3, rs = kb & 1;
            const int ly = tid / (TX / 2), m = tid - ly * (TX / 2);
            const int lx = 2 * m;
            const int gy = y0 + ly;
            const int par = (x0 + lx + gy + kb) & 1;  // 0: cell lx is red, lx+1 black
            const int bxl = lx + (1 - par);            // local x of the black cell
            const int ex = bxl + 2, ey = ly + 2;
            const T ec = ZERO_IN ? T(0) : se[sc][ey][ex];   // the black cell's input
            const T eb = gs_update(ec, sr[rs][ly + 1][bxl + 1 + RSH], se[sc][ey][ex - 1] + se[sc][ey][ex + 1],
                                   se[sc][ey - 1][ex] + se[sc][ey + 1][ex], se[sm][ey][ex] + se[sp][ey][ex], gamma,
                                   inv_gamma, dhx, dhy, dhz);
            const T er = se[sc][ey][lx + par + 2];     // the red cell, updated this sweep
            const T e0 = par == 0 ? er : eb, e1 = par == 0 ? eb : er;
            const size_t o = cidx(x0 + lx, gy, wrapz(kb), nx, ny);
            if (UPD) {
                Ph[o] = Ph[o] + (double)e0;
                Ph[o + 1] = Ph[o + 1] + (double)e1;
            } else {
                Eout[o] = e0;
                Eout[o + 1] = e1;
            }
        }
        __syncthreads();
    };
    fetch_e(z0 - 2, 0);
    put_e(z0 - 2, 0);
    fetch_e(z0 - 1, DEPTH - 1);
    put_e(z0 - 1, DEPTH - 1);
    fetch_e(z0, 0);
    fetch_r(z0 - 1, 0);
    int kr = z0 - 1;
    if (DEPTH == 2) {
        fetch_e(z0 + 1, DEPTH - 1);
        fetch_r(z0, DEPTH - 1);
        for (; kr + 1 <= z1; kr += 2) {
            step(kr, 0);
            step(kr + 1, DEPTH - 1);
        }
        if (kr <= z1) step(kr, 0);
    } else {
        for (; kr <= z1; ++kr) step(kr, 0);
    }
    if (RES) {
        // k_residual0's norm slots: max over the block, atomicMax on the bit pattern
        for (int o = kWave / 2; o > 0; o >>= 1) {
            ar = fmax(ar, __shfl_xor(ar, o));
            af = fmax(af, __shfl_xor(af, o));
        }
        __shared__ double wr[NT / kWave], wf[NT / kWave];
        if ((tid & (kWave - 1)) == 0) {
            wr[tid / kWave] = ar;
            wf[tid / kWave] = af;
        }
        __syncthreads();
        if (tid == 0) {
            double mr = wr[0], mf = wf[0];
            for (int w = 1; w < NT / kWave; ++w) {
                mr = fmax(mr, wr[w]);
                mf = fmax(mf, wf[w]);
            }
            const int slot = sg.box * kNormSlots + (bid & (kNormSlots - 1));
            atomicMax(&res.nr[slot], dbits(mr));
            if (res.nf) atomicMax(&res.nf[slot], dbits(mf));
        }
    }
}

// Two full red-black Gauss-Seidel sweeps in one pass (temporal blocking): the
// pre-smoothing pair or the post-smoothing pair of a V(2,2) cycle on a fine
// level reads e and r once and writes e once, where two k_rb_sweep passes move
// each of them twice.
//
// A block owns a TX x TY column of cells and a chunk of z-planes [z0, z1) and
// marches up in z over the planes z0-4 .. z1+3.  It has one thread per pair of
// cells (x even, x+1) of the column widened by a halo of 4 in x and y.  The
// thread keeps its pair's e over the six planes k+1 .. k-4 and r over the four
// planes k .. k-3 in registers, each plane split into its red and its black
// cell (in a pair, the red cell is the first where y + z is even and the
// second otherwise).  Iteration k runs four updates of the pair, each with
// k_gsrb_half's arithmetic:
//   A  sweep 1, red, plane k        B  sweep 1, black, plane k-1
//   C  sweep 2, red, plane k-2      D  sweep 2, black, plane k-3
// after which plane k-3 is final and the block's own cells are written out.
// An update's z-neighbours are the same pair's cells in the planes above and
// below (registers); its in-plane neighbours have the other colour: one is the
// pair's other cell, three lie in the pairs left or right, above and below,
// and are read from LDS half-planes (one array per colour and plane, indexed
// by pair, so that a wavefront reads consecutive words).  Iteration k reads the
// half-planes black k (input values), red k-1, black k-2 and red k-3, each
// written in iteration k-1, and writes red k, black k-1, red k-2 and black k+1
// (input values), so one barrier per plane orders every write before its
// reads.  The region where the values are those of the two separate sweeps
// shrinks by one cell per step (halo 3 after A, 2 after B, 1 after C, 0 after
// D); every thread runs every step, and the values it computes outside that
// region never reach a cell inside it, so the block's own cells equal two
// k_rb_sweep passes (and four k_gsrb_half) bit for bit.
// MODE bits as k_rb_sweep's: kSweepZero (sweep 1 from e = 0), kSweepProlong
// (sweep 1's input is ein plus the coarse correction), kSweepUpdate (sweep 2's
// result is added to phi instead of stored).  Needs nx % TX == 0, ny % TY == 0,
// even nx, ny, nz, nz >= 8 and nx * ny < 2^31.  HBM bytes per cell:
// 3*sizeof(T) (e read, r read, e write) for both sweeps (+16 with
// kSweepUpdate: phi read and written).
template <typename T>
struct alignas(2 * sizeof(T)) CellPair {
    T a, b;
};
constexpr int kSw2TX = 32, kSw2TY = 16;
template <int TX, int TY>
constexpr int sw2_threads() {
    return (TX + 8) / 2 * (TY + 8);
}
template <typename T, int TX, int TY, int MODE>
__global__ void __launch_bounds__((sw2_threads<TX, TY>())) k_rb_sweep2(const CellSeg* segs, int nseg,
                                                                     const int* active, const T* __restrict__ ein,
                                                                     T* __restrict__ eout, const T* __restrict__ r,
                                                                     const T* ec0, const T* ec1, double* phi) {
    constexpr bool ZERO_IN = (MODE & kSweepZero) != 0;
    constexpr bool PRO = (MODE & kSweepProlong) != 0;
    constexpr bool UPD = (MODE & kSweepUpdate) != 0;
    constexpr int PX = (TX + 8) / 2, PY = TY + 8;  // pairs per row, rows
    // LDS half-plane: PY rows of PX pairs, with a row (plus one word) of
    // padding on either side for the out-of-region reads of the edge threads
    constexpr int LP = (PY + 2) * PX + 2;
    __shared__ T sred[4][LP];
    __shared__ T sblk[4][LP];
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const int s = find_seg(segs, nseg, bid);
    const CellSeg& sg = segs[s];
    if (!active[sg.box]) return;
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const int ntx = nx / TX, nty = ny / TY;
    const int zc = sg.zchunk;
    const int b = bid - sg.blk0;
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * zc;
    const int z1 = min(nz, z0 + zc);
    const T dhx = T(sg.dh[0]), dhy = T(sg.dh[1]), dhz = T(sg.dh[2]);
    const T gamma = T(sg.gamma), inv_gamma = T(sg.inv_gamma);
    // FP32: r is kept scaled by 1/gamma (gs_update_scaled)
    constexpr bool F32 = std::is_same<T, float>::value;
    const T cx = dhx * inv_gamma, cy = dhy * inv_gamma, cz = dhz * inv_gamma;
    auto upd = [&](T ec, T rr, T sx, T sy, T sz) {
        if constexpr (F32)
            return gs_update_scaled(rr, sx, sy, sz, cx, cy, cz);
        else
            return gs_update(ec, rr, sx, sy, sz, gamma, inv_gamma, dhx, dhy, dhz);
    };
    auto rscale = [&](T v) {
        if constexpr (F32)
            return v * inv_gamma;
        else
            return v;
    };
    const T* Ein = ein + sg.off;
    T* Eout = eout + sg.off;
    const T* R = r + sg.off;
    const T* Ec = (sg.ebuf ? ec1 : ec0) + sg.coff;
    double* Ph = phi + sg.foff;
    const int cnx = nx >> 1, cny = ny >> 1;
    const size_t plane = (size_t)nx * ny, cplane = (size_t)cnx * cny;
    auto wrapz = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };

    const int tid = threadIdx.x;
    const int ly = tid / PX, m = tid - ly * PX;
    int gx = x0 - 4 + 2 * m, gy = y0 - 4 + ly;
    gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
    gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
    const unsigned po = (unsigned)gy * (unsigned)nx + (unsigned)gx;      // pair's in-plane offset
    const unsigned cpo = (unsigned)(gy >> 1) * (unsigned)cnx + (unsigned)(gx >> 1);  // its parent's
    const bool own = ly >= 4 && ly < 4 + TY && m >= 2 && m < 2 + TX / 2;
    // the own pair nearest to this one (phi is read there by every thread, so
    // that no load is conditional: a conditional load costs the compiler its
    // count of loads in flight, and then every wait is a wait for all of them)
    const unsigned pown = (unsigned)(y0 + min(max(ly - 4, 0), TY - 1)) * (unsigned)nx +
                          (unsigned)(x0 + 2 * min(max(m - 2, 0), TX / 2 - 1));
    const int la = (ly + 1) * PX + m + 1;  // LDS index of the pair
    const int k0 = z0 - 3, kend = z1 + 2;  // iterations k0 .. kend
    // In iteration k every update is of the pair's cell at index (gy + k) & 1
    // (red of k and k-2, black of k-1 and k-3), whose outside x-neighbour lies
    // in pair m-1 for index 0 and m+1 for index 1.  c0: that index for
    // iterations of k0's parity.
    const int c0 = (gy + k0) & 1;
    const int lxe = la + (c0 ? 1 : -1), lxo = la + (c0 ? -1 : 1);
    // a wavefront whose rows all lie outside the region a step must get right
    // skips that step (B, C, D: rows 2 .. TY+5, 3 .. TY+4, 4 .. TY+3)
    const int wv = __builtin_amdgcn_readfirstlane(tid) / 64;
    const int wr0 = wv * 64 / PX, wr1 = min(PY - 1, (wv * 64 + 63) / PX);
    const bool doB = wr1 >= 2 && wr0 <= TY + 5, doC = wr1 >= 3 && wr0 <= TY + 4, doD = wr1 >= 4 && wr0 <= TY + 3;

    // e and r of the pair: index i <-> plane k+1-i (e), k-i (r) at the top of
    // iteration k
    T Er[6], Eb[6], Fr[4], Fb[4];
#pragma unroll
    for (int i = 0; i < 6; ++i) Er[i] = Eb[i] = T(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) Fr[i] = Fb[i] = T(0);
    // loads in flight: e of plane k+3 and r of plane k+2 issued in iteration k
    // (slot k & 1), moved into the window at the end of iteration k+1; phi of
    // plane k-1 for the own cells (kSweepUpdate), consumed in iteration k+2
    CellPair<T> qe[2], qr[2];
    T qc[2];
    CellPair<double> qp[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        qe[i] = qr[i] = CellPair<T>{T(0), T(0)};
        qc[i] = T(0);
        qp[i] = CellPair<double>{0.0, 0.0};
    }
    // (planes past the last one needed re-read that one: every load is issued
    // in every iteration)
    // plane of a load: u (unwrapped, held at lim, the last plane needed), w
    // (wrapped), its offset and its parent plane's offset.  The loop advances
    // its cursors one plane per iteration instead of recomputing the offsets
    // (a scalar multiply chain per load).
    struct Cursor {
        int u, lim, w;
        size_t off, coff;
    };
    auto cursor = [&](int u, int lim) {
        Cursor q;
        q.u = min(u, lim);
        q.lim = lim;
        q.w = wrapz(q.u);
        q.off = (size_t)q.w * plane;
        q.coff = (size_t)(q.w >> 1) * cplane;
        return q;
    };
    auto advance = [&](Cursor& q) {
        if (q.u < q.lim) {
            ++q.u;
            if (++q.w == nz) {
                q.w = 0;
                q.off = 0;
                q.coff = 0;
            } else {
                q.off += plane;
                if (!(q.w & 1)) q.coff += cplane;
            }
        }
    };
    auto load_e = [&](const Cursor& q, CellPair<T>& v, T& cv) {
        if (ZERO_IN) return;
        v = *reinterpret_cast<const CellPair<T>*>(Ein + q.off + po);
        if (PRO) cv = Ec[q.coff + cpo];
    };
    auto load_r = [&](const Cursor& q, CellPair<T>& v) {
        v = *reinterpret_cast<const CellPair<T>*>(R + q.off + po);
    };
    // red and black cell of a loaded e pair whose red index is cr
    auto split_e = [&](int cr, const CellPair<T>& v, T cv, T& red, T& blk) {
        if (ZERO_IN) {
            red = blk = T(0);
            return;
        }
        T a = v.a, bb = v.b;
        if (PRO) {
            a = a + cv;
            bb = bb + cv;
        }
        red = cr ? bb : a;
        blk = cr ? a : bb;
    };

    // prologue: e of planes k0-1 .. k0+1 into the window, plane k0+2 and r of
    // plane k0+1 in flight (slot of k0-1), r of plane k0 into the window, and
    // the black input of plane k0 into LDS
    {
        CellPair<T> v;
        T cv = T(0);
        v = CellPair<T>{T(0), T(0)};
        load_e(cursor(k0 - 1, z1 + 3), v, cv);
        split_e(c0 ^ 1, v, cv, Er[2], Eb[2]);
        load_e(cursor(k0, z1 + 3), v, cv);
        split_e(c0, v, cv, Er[1], Eb[1]);
        load_e(cursor(k0 + 1, z1 + 3), v, cv);
        split_e(c0 ^ 1, v, cv, Er[0], Eb[0]);
        v = CellPair<T>{T(0), T(0)};
        load_r(cursor(k0, z1 + 2), v);
        Fr[0] = rscale(c0 ? v.b : v.a);
        Fb[0] = rscale(c0 ? v.a : v.b);
        load_e(cursor(k0 + 2, z1 + 3), qe[1], qc[1]);
        load_r(cursor(k0 + 1, z1 + 2), qr[1]);
        if (!ZERO_IN) sblk[0][la] = Eb[1];
    }
    __syncthreads();
    // the loop's loads: e of plane k+3, r of plane k+2, phi of plane k-1 (held
    // in [z0, z1-1]); its stores: plane k-3 (from z0)
    Cursor ce = cursor(k0 + 3, z1 + 3), cr = cursor(k0 + 2, z1 + 2);
    size_t pho = (size_t)z0 * plane, so = (size_t)z0 * plane;

    auto body = [&](auto JJ, int k) {
        constexpr int J = decltype(JJ)::value;
        // LDS slot of plane p is (p - k0) & 3: planes k, k-1, k-2, k-3 (= k+1)
        constexpr int s0 = J & 3, s1 = (J + 3) & 3, s2 = (J + 2) & 3, s3 = (J + 1) & 3;
        constexpr int qn = J & 1, qo = (J + 1) & 1;  // queue slot loaded now / consumed now
        const int c = c0 ^ (J & 1);
        const int lx = (J & 1) ? lxo : lxe;
        load_e(ce, qe[qn], qc[qn]);
        load_r(cr, qr[qn]);
        advance(ce);
        advance(cr);
        // the in-plane neighbours of all four steps, issued together: each step
        // reads a half-plane written in the previous iteration and none that
        // this one writes
        T xa = T(0), ya0 = T(0), ya1 = T(0);
        if (!ZERO_IN) {
            xa = sblk[s0][lx];
            ya0 = sblk[s0][la - PX];
            ya1 = sblk[s0][la + PX];
        }
        const T xb = sred[s1][lx], yb0 = sred[s1][la - PX], yb1 = sred[s1][la + PX];
        const T xc = sblk[s2][lx], yc0 = sblk[s2][la - PX], yc1 = sblk[s2][la + PX];
        const T xd = sred[s3][lx], yd0 = sred[s3][la - PX], yd1 = sred[s3][la + PX];
        // A: sweep 1, red, plane k
        if (ZERO_IN) {
            // every e operand an input zero: the update is r/gamma exactly
            const T z = T(0);
            if constexpr (F32)
                Er[1] = Fr[0];
            else
                Er[1] = upd(z, Fr[0], z + z, z + z, z + z);
        } else {
            Er[1] = upd(Er[1], Fr[0], Eb[1] + xa, ya0 + ya1, Eb[2] + Eb[0]);
        }
        sred[s0][la] = Er[1];
        // B: sweep 1, black, plane k-1
        if (doB) {
            Eb[2] = upd(Eb[2], Fb[1], Er[2] + xb, yb0 + yb1, Er[3] + Er[1]);
            sblk[s1][la] = Eb[2];
        }
        // C: sweep 2, red, plane k-2
        if (doC) {
            Er[3] = upd(Er[3], Fr[2], Eb[3] + xc, yc0 + yc1, Eb[4] + Eb[2]);
            sred[s2][la] = Er[3];
        }
        // D: sweep 2, black, plane k-3, and the plane out
        if (doD) Eb[4] = upd(Eb[4], Fb[3], Er[4] + xd, yd0 + yd1, Er[5] + Er[3]);
        const int kb = k - 3;
        if (own && kb >= z0 && kb < z1) {
            // plane k-3: its red cell has index c ^ 1
            const T e0 = c ? Er[4] : Eb[4], e1 = c ? Eb[4] : Er[4];
            // streaming stores: the output is read again only by the next pass,
            // and lines kept in L2 would push out the input halos the
            // neighbouring blocks are about to read
            if (UPD) {
                typedef double v2d __attribute__((ext_vector_type(2)));
                v2d o = {qp[qn].a + (double)e0, qp[qn].b + (double)e1};
                __builtin_nontemporal_store(o, reinterpret_cast<v2d*>(Ph + so + po));
            } else {
                typedef T v2t __attribute__((ext_vector_type(2)));
                v2t o = {e0, e1};
                __builtin_nontemporal_store(o, reinterpret_cast<v2t*>(Eout + so + po));
            }
        }
        if (kb >= z0) so += plane;
        if (UPD) {
            qp[qn] = *reinterpret_cast<const CellPair<double>*>(Ph + pho + pown);
            if (k - 1 >= z0 && k - 1 < z1 - 1) pho += plane;
        }
        // the black input of plane k+1 for the next iteration's step A
        if (!ZERO_IN) sblk[s3][la] = Eb[0];
        __syncthreads();
        // the window moves down one plane: e of plane k+2 and r of plane k+1
        // come in from the loads issued in iteration k-1
#pragma unroll
        for (int i = 5; i > 0; --i) {
            Er[i] = Er[i - 1];
            Eb[i] = Eb[i - 1];
        }
        split_e(c, qe[qo], qc[qo], Er[0], Eb[0]);
#pragma unroll
        for (int i = 3; i > 0; --i) {
            Fr[i] = Fr[i - 1];
            Fb[i] = Fb[i - 1];
        }
        Fr[0] = rscale(c ? qr[qo].a : qr[qo].b);
        Fb[0] = rscale(c ? qr[qo].b : qr[qo].a);
    };
    // whole groups of four iterations (up to three past kend, whose loads
    // re-read planes already loaded and whose planes are not written out): one
    // exit from the loop, so that the loads in flight across its back edge keep
    // their registers
    for (int k = k0; k <= kend; k += 4) {
        body(std::integral_constant<int, 0>{}, k);
        body(std::integral_constant<int, 1>{}, k + 1);
        body(std::integral_constant<int, 2>{}, k + 2);
        body(std::integral_constant<int, 3>{}, k + 3);
    }
}

// Residual of the correction, restricted: r_c(I) = 0.125 * sum_children (r - L e),
// children summed k, then j, then i innermost.  One thread per coarse cell.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_resid_restrict(const CellSeg* segs, int nseg, const int* active,
                                                           const T* __restrict__ e, T* __restrict__ r) {
    const int s = find_seg(segs, nseg, blockIdx.x);
    const CellSeg& sg = segs[s];
    if (!active[sg.box]) return;
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    if (q >= (long long)cx * cy * cz) return;
    const int I = (int)(q % cx);
    const long long t = q / cx;
    const int J = (int)(t % cy), K = (int)(t / cy);
    const T dhx = T(sg.dh[0]), dhy = T(sg.dh[1]), dhz = T(sg.dh[2]);
    const T* E = e + sg.off;
    const T* R = r + sg.off;
    T acc = T(0);
    for (int kk = 0; kk < 2; ++kk)
        for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) {
                int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                T rc = R[cidx(i, j, k, nx, ny)] - lap7(E, i, j, k, nx, ny, nz, dhx, dhy, dhz);
                acc += rc;
            }
    r[sg.coff + q] = T(0.125) * acc;
}

// The same restriction marching up the coarse planes: a block owns a CX x CY tile
// of coarse columns (2CX x 2CY fine) and a chunk of coarse planes, with a ring of
// the 4 fine e planes 2K-1 .. 2K+2 (halo 1 in x and y) and the 2 fine r planes of
// coarse plane K in LDS, the next planes fetched into registers during the
// compute.  Each fine value is read once per block (plus the x-y halo), where the
// thread-per-cell kernel fetched the e planes between coarse planes twice.  Same
// arithmetic (lap7's, children summed k, j, i) bit for bit.  Needs nx % 2CX ==
// ny % 2CY == nz % 2 == 0.
template <typename T, int CX, int CY>
__global__ void __launch_bounds__(CX * CY) k_resid_restrict_z(const CellSeg* segs, int nseg, const int* active,
                                                              const T* __restrict__ e, T* __restrict__ r) {
    constexpr int NT = CX * CY;
    constexpr int FX = 2 * CX + 2, FY = 2 * CY + 2;  // fine e plane with halo 1
    constexpr int RX = 2 * CX, RY = 2 * CY;          // fine r plane
    constexpr int NEP = (FX * FY + NT - 1) / NT, NRP = (RX * RY) / NT;
    static_assert((RX * RY) % NT == 0, "r plane split");
    __shared__ T se[4][FY][FX];
    __shared__ T sr[2][RY][RX];
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const int s = find_seg(segs, nseg, bid);
    const CellSeg& sg = segs[s];
    if (!active[sg.box]) return;
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
    const int ntx = nx / RX, nty = ny / RY;
    const int b = bid - sg.blk0;
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
    const int x0 = tx * RX, y0 = ty * RY;
    const int K0 = tz * sg.zchunk, K1 = min(cz, K0 + sg.zchunk);
    const T dhx = T(sg.dh[0]), dhy = T(sg.dh[1]), dhz = T(sg.dh[2]);
    const T* E = e + sg.off;
    const T* R = r + sg.off;
    const int tid = threadIdx.x;
    const size_t plane = (size_t)nx * ny;
    size_t eoff[NEP], roff[NRP];
#pragma unroll
    for (int u = 0; u < NEP; ++u) {
        const int idx = min(tid + u * NT, FX * FY - 1);
        const int ly = idx / FX, lx = idx - ly * FX;
        int gx = x0 + lx - 1, gy = y0 + ly - 1;
        gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
        gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        eoff[u] = (size_t)gy * nx + gx;
    }
#pragma unroll
    for (int u = 0; u < NRP; ++u) {
        const int idx = tid + u * NT;
        const int ly = idx / RX, lx = idx - ly * RX;
        roff[u] = (size_t)(y0 + ly) * nx + (x0 + lx);
    }
    auto wrapz = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };
    auto fetch_e = [&](int kz, T* pe) {
        const size_t zo = (size_t)wrapz(kz) * plane;
#pragma unroll
        for (int u = 0; u < NEP; ++u)
            if (tid + u * NT < FX * FY) pe[u] = E[zo + eoff[u]];
    };
    auto put_e = [&](int kz, const T* pe) {
        T* dst = &se[kz & 3][0][0];
#pragma unroll
        for (int u = 0; u < NEP; ++u)
            if (tid + u * NT < FX * FY) dst[tid + u * NT] = pe[u];
    };
    auto fetch_r = [&](int kz, T* pr) {
        const size_t zo = (size_t)kz * plane;
#pragma unroll
        for (int u = 0; u < NRP; ++u) pr[u] = R[zo + roff[u]];
    };
    auto put_r = [&](int kz, const T* pr) {
        T* dst = &sr[kz & 1][0][0];
#pragma unroll
        for (int u = 0; u < NRP; ++u) dst[tid + u * NT] = pr[u];
    };
    T pa[NEP], pb[NEP], qa[NRP], qb[NRP];
    fetch_e(2 * K0 - 1, pa);
    fetch_e(2 * K0, pb);
    put_e(2 * K0 - 1, pa);
    put_e(2 * K0, pb);
    fetch_e(2 * K0 + 1, pa);
    fetch_e(2 * K0 + 2, pb);
    fetch_r(2 * K0, qa);
    fetch_r(2 * K0 + 1, qb);
    put_e(2 * K0 + 1, pa);
    put_e(2 * K0 + 2, pb);
    put_r(2 * K0, qa);
    put_r(2 * K0 + 1, qb);
    __syncthreads();
    const int lxc = tid % CX, lyc = tid / CX;
    T* Rc = r + sg.coff + ((size_t)(y0 >> 1) + lyc) * cx + (size_t)(x0 >> 1) + lxc;
    for (int K = K0; K < K1; ++K) {
        const bool more = K + 1 < K1;
        if (more) {
            fetch_e(2 * K + 3, pa);
            fetch_e(2 * K + 4, pb);
            fetch_r(2 * K + 2, qa);
            fetch_r(2 * K + 3, qb);
        }
        T acc = T(0);
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            const int k = 2 * K + kk;
            const int sc = k & 3, sm = (k - 1) & 3, sp = (k + 1) & 3;
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int ii = 0; ii < 2; ++ii) {
                    const int fy = 2 * lyc + jj + 1, fx = 2 * lxc + ii + 1;
                    const T ec = se[sc][fy][fx];
                    // lap7's arithmetic
                    const T lx = dhx * (se[sc][fy][fx - 1] - T(2) * ec + se[sc][fy][fx + 1]);
                    const T ly = dhy * (se[sc][fy - 1][fx] - T(2) * ec + se[sc][fy + 1][fx]);
                    const T lz = dhz * (se[sm][fy][fx] - T(2) * ec + se[sp][fy][fx]);
                    const T rc = sr[k & 1][fy - 1][fx - 1] - (lx + ly + lz);
                    acc += rc;
                }
        }
        Rc[(size_t)K * cy * cx] = T(0.125) * acc;
        __syncthreads();
        if (more) {
            put_e(2 * K + 3, pa);
            put_e(2 * K + 4, pb);
            put_r(2 * K + 2, qa);
            put_r(2 * K + 3, qb);
            __syncthreads();
        }
    }
}

// e_fine += e_coarse(parent): piecewise-constant prolongation (ec: the coarse
// level's buffer, addressed at sg.coff).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_prolong(const CellSeg* segs, int nseg, const int* active, T* e,
                                                    const T* ec0, const T* ec1) {
    const int s = find_seg(segs, nseg, blockIdx.x);
    const CellSeg& sg = segs[s];
    if (!active[sg.box]) return;
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    if (q >= (long long)nx * ny * nz) return;
    const int i = (int)(q % nx);
    const long long t = q / nx;
    const int j = (int)(t % ny), k = (int)(t / ny);
    const T* ecrs = sg.ebuf ? ec1 : ec0;
    const T ec = ecrs[sg.coff + cidx(i >> 1, j >> 1, k >> 1, nx >> 1, ny >> 1)];
    e[sg.off + q] = e[sg.off + q] + ec;
}

// Exact bottom solve: e[c] = sum_{c'} G[(c - c') mod n] r[c'], c' in natural order.
// One block per box; r and G staged in LDS (at most 4096 cells each).  The
// convolution is FP64 in both precisions (G and the sum in FP64, r widened), so
// the FP32 V-cycle of the mixed path keeps its coarsest solve in FP64.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bottom(const CellSeg* segs, int nseg, const int* active,
                                                   T* __restrict__ e, const T* __restrict__ r,
                                                   const double* __restrict__ G) {
    const CellSeg& sg = segs[blockIdx.x];
    if (!active[sg.box]) return;
    __shared__ T sr[kBottomMaxCells];
    __shared__ double sG[kBottomMaxCells];
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const int nc = nx * ny * nz;
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        sr[c] = r[sg.off + c];
        sG[c] = G[sg.coff + c];
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        double acc = 0.0;
        for (int qz = 0; qz < nz; ++qz) {
            int dz = cz - qz; dz = dz < 0 ? dz + nz : dz;
            for (int qy = 0; qy < ny; ++qy) {
                int dy = cy - qy; dy = dy < 0 ? dy + ny : dy;
                for (int qx = 0; qx < nx; ++qx) {
                    int dxx = cx - qx; dxx = dxx < 0 ? dxx + nx : dxx;
                    acc += sG[(dz * ny + dy) * nx + dxx] * (double)sr[(qz * ny + qy) * nx + qx];
                }
            }
        }
        e[sg.off + c] = T(acc);
    }
}

// The coarse end of a V-cycle in one workgroup per box: every level from the
// first one with at most kCoarseCells cells (segs[first[b]], "top") down to the
// bottom (segs[first[b+1]-1]), and back up.  Corrections e of all these levels
// and residuals r of the levels below the top live in LDS; the top's r is read
// from the pool and the top's final e is written to the pool (buffer e), where
// the prolongation of the next finer level finds it.  Per cell the arithmetic is
// that of k_fill, k_gsrb_half, k_resid_restrict, k_bottom and k_prolong, in the
// same order, so the result equals those launches bit for bit; what it saves is
// ~30 latency-bound launches per cycle.
constexpr int kCoarseCells = 4096;
constexpr int kCoarseThreads = 1024;

template <typename T>
__device__ __forceinline__ void coarse_half(const CellSeg& sg, T* E, const T* R, int colour) {
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const int hx = nx >> 1;
    const int nh = hx * ny * nz;
    const T dhx = T(sg.dh[0]), dhy = T(sg.dh[1]), dhz = T(sg.dh[2]);
    const T gamma = T(sg.gamma), inv_gamma = T(sg.inv_gamma);
    for (int q = threadIdx.x; q < nh; q += blockDim.x) {
        const int ih = q % hx;
        const int t = q / hx;
        const int j = t % ny, k = t / ny;
        const int i = 2 * ih + ((j + k + colour) & 1);
        int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
        int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
        int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
        const int c = (k * ny + j) * nx + i;
        E[c] = gs_update(E[c], R[c], E[(k * ny + j) * nx + im] + E[(k * ny + j) * nx + ip],
                         E[(k * ny + jm) * nx + i] + E[(k * ny + jp) * nx + i],
                         E[(km * ny + j) * nx + i] + E[(kp * ny + j) * nx + i], gamma, inv_gamma, dhx, dhy, dhz);
    }
}

template <typename T>
__global__ void __launch_bounds__(kCoarseThreads) k_coarse_vcycle(const CellSeg* segs, const int* first,
                                                                  const int* active, T* __restrict__ e,
                                                                  const T* __restrict__ r,
                                                                  const double* __restrict__ G, int nu1, int nu2,
                                                                  int keep_top) {
    extern __shared__ double coarse_smem[];
    T* lds = reinterpret_cast<T*>(coarse_smem);
    const int l0 = first[blockIdx.x], l1 = first[blockIdx.x + 1];
    if (!active[segs[l0].box]) return;
    const int nl = l1 - l0;
    auto ncell = [&](int l) { return segs[l].n[0] * segs[l].n[1] * segs[l].n[2]; };
    // LDS: e of levels l0..l1-1, then r of levels l0+1..l1-1
    auto eoff = [&](int l) {
        int o = 0;
        for (int q = l0; q < l; ++q) o += ncell(q);
        return o;
    };
    auto roff = [&](int l) {
        int o = eoff(l1);
        for (int q = l0 + 1; q < l; ++q) o += ncell(q);
        return o;
    };
    auto rptr = [&](int l) -> const T* { return l == l0 ? r + segs[l0].off : lds + roff(l); };
    const int tid = threadIdx.x, nth = blockDim.x;
    // down: zero guess (keep_top: on the box's own level, the e of the previous
    // V-cycle of the same refinement step), nu1 red-black sweeps, residual
    // restricted to the next level
    for (int l = l0; l + 1 < l1; ++l) {
        const CellSeg& sg = segs[l];
        T* E = lds + eoff(l);
        const T* R = rptr(l);
        const int nc = ncell(l);
        const bool keep = keep_top && l == l0 && sg.lev == 0;
        const T* Eg = e + sg.off;
        for (int q = tid; q < nc; q += nth) E[q] = keep ? Eg[q] : T(0);
        __syncthreads();
        for (int it = 0; it < nu1; ++it)
            for (int col = 0; col < 2; ++col) {
                coarse_half(sg, E, R, col);
                __syncthreads();
            }
        const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
        const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
        const T dhx = T(sg.dh[0]), dhy = T(sg.dh[1]), dhz = T(sg.dh[2]);
        T* Rc = lds + roff(l + 1);
        for (int q = tid; q < cx * cy * cz; q += nth) {
            const int I = q % cx;
            const int t = q / cx;
            const int J = t % cy, K = t / cy;
            T acc = T(0);
            for (int kk = 0; kk < 2; ++kk)
                for (int jj = 0; jj < 2; ++jj)
                    for (int ii = 0; ii < 2; ++ii) {
                        const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                        T rc = R[(k * ny + j) * nx + i] - lap7(E, i, j, k, nx, ny, nz, dhx, dhy, dhz);
                        acc += rc;
                    }
            Rc[q] = T(0.125) * acc;
        }
        __syncthreads();
    }
    // bottom: e = G * r (periodic convolution, natural order), FP64 in both
    // precisions as in k_bottom
    {
        const int lb = l1 - 1;
        const CellSeg& sg = segs[lb];
        const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
        const int nc = nx * ny * nz;
        const T* R = rptr(lb);
        const double* Gb = G + sg.coff;
        T* E = lds + eoff(lb);
        for (int c = tid; c < nc; c += nth) {
            const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
            double acc = 0.0;
            for (int qz = 0; qz < nz; ++qz) {
                int dz = cz - qz; dz = dz < 0 ? dz + nz : dz;
                for (int qy = 0; qy < ny; ++qy) {
                    int dy = cy - qy; dy = dy < 0 ? dy + ny : dy;
                    for (int qx = 0; qx < nx; ++qx) {
                        int dxx = cx - qx; dxx = dxx < 0 ? dxx + nx : dxx;
                        acc += Gb[(dz * ny + dy) * nx + dxx] * (double)R[(qz * ny + qy) * nx + qx];
                    }
                }
            }
            E[c] = T(acc);
        }
        __syncthreads();
    }
    // up: piecewise-constant prolongation, nu2 red-black sweeps
    for (int l = l1 - 2; l >= l0; --l) {
        const CellSeg& sg = segs[l];
        T* E = lds + eoff(l);
        const T* Ec = lds + eoff(l + 1);
        const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
        const int nc = nx * ny * nz;
        for (int q = tid; q < nc; q += nth) {
            const int i = q % nx;
            const int t = q / nx;
            const int j = t % ny, k = t / ny;
            E[q] = E[q] + Ec[((k >> 1) * (ny >> 1) + (j >> 1)) * (nx >> 1) + (i >> 1)];
        }
        __syncthreads();
        const T* R = rptr(l);
        for (int it = 0; it < nu2; ++it)
            for (int col = 0; col < 2; ++col) {
                coarse_half(sg, E, R, col);
                __syncthreads();
            }
    }
    {
        T* Eg = e + segs[l0].off;
        const T* E = lds + eoff(l0);
        const int nc = ncell(l0);
        for (int q = tid; q < nc; q += nth) Eg[q] = E[q];
    }
    (void)nl;
}

// Level-0 residual r = f - L phi (FP64), stored as T.  A block owns a TX x TY
// column of cells and marches up a chunk of zc planes; each phi plane (with a
// one-cell periodic halo in x and y) is loaded once into a ring of 4 LDS planes,
// one plane ahead of its use, and the 7-point stencil reads the ring.  Threads
// of a tile that runs past the box edge load wrapped halo values and store
// nothing.  The arithmetic is that of lap7.  Each block folds max |r| (and max
// |f| when nf != nullptr) into one of kNormSlots per-box slots with atomicMax
// on the bit pattern (exact and order independent); k_check reduces the slots.
constexpr int kResTX = 64, kResTY = 16;
template <typename T>
__global__ void __launch_bounds__(kResTX * kResTY) k_residual0(const CellSeg* segs, int nseg, const int* active,
                                                              const double* __restrict__ f,
                                                              const double* __restrict__ solv, double Ggravity,
                                                              double inv_a, const double* __restrict__ phi,
                                                              T* __restrict__ r,
                                                              unsigned long long* nr, unsigned long long* nf) {
    constexpr int TX = kResTX, TY = kResTY, NT = TX * TY;
    constexpr int HX = TX + 2, HY = TY + 2, NH = (HX * HY + NT - 1) / NT;
    __shared__ double sp[4][HY][HX];
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const int s = find_seg(segs, nseg, bid);
    const CellSeg& sg = segs[s];
    if (!active[sg.box]) return;
    const int nx = sg.n[0], ny = sg.n[1], nz = sg.n[2];
    const int nbx = (nx + TX - 1) / TX, nby = (ny + TY - 1) / TY;
    const int b = bid - sg.blk0;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int x0 = bx * TX, y0 = by * TY;
    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int i = x0 + lx, j = y0 + ly;
    const bool own = i < nx && j < ny;
    const int z0 = bz * sg.zchunk, z1 = min(nz, z0 + sg.zchunk);
    const double* P = phi + sg.coff;
    const double* F = f + sg.coff;
    T* Rr = r + sg.off;
    const double dhx = sg.dh[0], dhy = sg.dh[1], dhz = sg.dh[2];
    // with solv: f holds rho - mass_offset; the right-hand side is k_rhs_final's
    // ((f - solv) * G) * (1/a), formed here instead of in a pass of its own
    const bool raw = solv != nullptr;
    const double ns = raw ? -solv[sg.box] : 0.0;
    const size_t plane = (size_t)nx * ny;
    const size_t col = own ? (size_t)j * nx + i : 0;
    // in-plane offsets of this thread's elements of the halo region (periodic)
    size_t hoff[NH];
#pragma unroll
    for (int u = 0; u < NH; ++u) {
        const int idx = min(tid + u * NT, HX * HY - 1);
        const int hy = idx / HX, hx = idx - hy * HX;
        int gx = (x0 + hx - 1) % nx, gy = (y0 + hy - 1) % ny;
        gx = gx < 0 ? gx + nx : gx;
        gy = gy < 0 ? gy + ny : gy;
        hoff[u] = (size_t)gy * nx + gx;
    }
    auto wrapz = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };
    double hv[NH];
    auto fetch = [&](int k) {
        const double* Pk = P + (size_t)wrapz(k) * plane;
#pragma unroll
        for (int u = 0; u < NH; ++u)
            if (tid + u * NT < HX * HY) hv[u] = Pk[hoff[u]];
    };
    auto put = [&](int k) {
        double* dst = &sp[k & 3][0][0];
#pragma unroll
        for (int u = 0; u < NH; ++u)
            if (tid + u * NT < HX * HY) dst[tid + u * NT] = hv[u];
    };
    double ar = 0.0, af = 0.0;
    fetch(z0 - 1);
    put(z0 - 1);
    fetch(z0);
    put(z0);
    fetch(z0 + 1);
    double fn = own ? F[(size_t)z0 * plane + col] : 0.0;
    // iteration k: plane k+1 goes into the ring, the loads of plane k+2 (and f
    // of plane k+1) are issued, plane k is formed.  Plane k+1 takes the slot of
    // plane k-3, last read in iteration k-2, one barrier earlier.
    for (int k = z0; k < z1; ++k) {
        put(k + 1);
        const double fv0 = fn;
        if (k + 1 < z1) {
            fetch(k + 2);
            if (own) fn = F[(size_t)(k + 1) * plane + col];
        }
        __syncthreads();
        if (own) {
            const int ex = lx + 1, ey = ly + 1;
            const double pc = sp[k & 3][ey][ex];
            const double lxv = dhx * (sp[k & 3][ey][ex - 1] - 2.0 * pc + sp[k & 3][ey][ex + 1]);
            const double lyv = dhy * (sp[k & 3][ey - 1][ex] - 2.0 * pc + sp[k & 3][ey + 1][ex]);
            const double lzv = dhz * (sp[(k - 1) & 3][ey][ex] - 2.0 * pc + sp[(k + 1) & 3][ey][ex]);
            double fv = fv0;
            if (raw) {
                const double t = fv + ns;
                fv = (t * Ggravity) * inv_a;
            }
            const double v = fv - (lxv + lyv + lzv);
            __builtin_nontemporal_store(T(v), &Rr[(size_t)k * plane + col]);  // read by the next pass only
            ar = fmax(ar, fabs(v));
            af = fmax(af, fabs(fv));
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        ar = fmax(ar, __shfl_xor(ar, o));
        af = fmax(af, __shfl_xor(af, o));
    }
    __shared__ double wr[NT / kWave], wf[NT / kWave];
    if ((tid & (kWave - 1)) == 0) {
        wr[tid / kWave] = ar;
        wf[tid / kWave] = af;
    }
    __syncthreads();
    if (tid == 0) {
        double mr = wr[0], mf = wf[0];
        for (int w = 1; w < NT / kWave; ++w) {
            mr = fmax(mr, wr[w]);
            mf = fmax(mf, wf[w]);
        }
        const int slot = sg.box * kNormSlots + (bid & (kNormSlots - 1));
        atomicMax(&nr[slot], dbits(mr));
        if (nf) atomicMax(&nf[slot], dbits(mf));
    }
}

// phi += e (level 0).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_update0(const CellSeg* segs, int nseg, const int* active,
                                                    double* __restrict__ phi, const T* e0, const T* e1) {
    const int s = find_seg(segs, nseg, blockIdx.x);
    const CellSeg& sg = segs[s];
    if (!active[sg.box]) return;
    const T* __restrict__ e = sg.ebuf ? e1 : e0;
    const long long n = (long long)sg.n[0] * sg.n[1] * sg.n[2];
    const long long q = (long long)(blockIdx.x - sg.blk0) * kBlock + threadIdx.x;
    if (q < n) phi[sg.coff + q] = phi[sg.coff + q] + (double)e[sg.off + q];
}

// Convergence bookkeeping per box (MLMG stop rule), on the device so that a
// captured graph can run a data-dependent number of cycles.
//   init: target = max(abs_tol, max(rel_tol, 1e-16) * max(|r0|, |f|)),
//         active = |r0| > target (or fixed_iters > 0)
//   iter: iters++, active = |r| > target && iters < limit
// state: [0] resnorm0 [1] rhsnorm [2] final [3] target
__global__ void k_check(int nbox, int init, int fixed_iters, int max_iter, double rel_tol,
                        double abs_tol, int* active, int* iters, int* failed,
                        unsigned long long* nr, unsigned long long* nf, double* state, int inc) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    unsigned long long br = 0ull, bf = 0ull;
    for (int q = 0; q < kNormSlots; ++q) {
        br = max(br, nr[b * kNormSlots + q]);
        bf = max(bf, nf[b * kNormSlots + q]);
        nr[b * kNormSlots + q] = 0ull;
        nf[b * kNormSlots + q] = 0ull;
    }
    if (init) {
        const double r0 = __longlong_as_double((long long)br);
        const double fn = __longlong_as_double((long long)bf);
        const double mx = r0 > fn ? r0 : fn;
        const double rt = rel_tol > 1e-16 ? rel_tol : 1e-16;
        double target = rt * mx;
        if (abs_tol > target) target = abs_tol;
        state[4 * b + 0] = r0;
        state[4 * b + 1] = fn;
        state[4 * b + 2] = r0;
        state[4 * b + 3] = target;
        iters[b] = 0;
        failed[b] = 0;
        active[b] = fixed_iters > 0 ? 1 : (r0 > target ? 1 : 0);
    } else if (active[b]) {
        const double rn = __longlong_as_double((long long)br);
        const int it = iters[b] + inc;  // the V-cycles run since the last check
        iters[b] = it;
        state[4 * b + 2] = rn;
        if (fixed_iters > 0) {
            active[b] = it < fixed_iters ? 1 : 0;
        } else {
            const int more = rn > state[4 * b + 3];
            active[b] = (more && it < max_iter) ? 1 : 0;
            failed[b] = (more && it >= max_iter) ? 1 : 0;
        }
    }
}

// After the recorded V-cycles of a captured solve: a box still active has not
// converged (sticky over the solves of one replay), and the most cycles a box
// used.  The host clears gstat before each replay.
__global__ void k_graph_note(int nbox, const int* active, const int* iters, int* gstat) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    if (active[b]) gstat[b] = 1;
    gstat[nbox + b] = max(gstat[nbox + b], iters[b]);
}

}  // namespace nyxpm
