// nyxpm.hip -- host side of the MI355X particle-mesh library and its C ABI
// (include/nyxpm.h).  One translation unit: the kernel files are included here.
//
// The host code owns every device buffer, builds the launch tables for a batch
// of periodic boxes, and drives the step in the order of
// Nyx::advance_particles_only (Source/TimeStep/Nyx_advance_particles.cpp:137-268):
//   moveKickDrift -> AssignDensity -> solve_for_phi -> moveKick -> Redistribute.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <sys/stat.h>

#include <cerrno>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/nyxpm.h"
#include "nyxpm_internal.h"
#include "mg_kernels.hip"
#include "particle_kernels.hip"

using namespace nyxpm;

// Staging of nyxpm_gather_accel_noalloc: the accelerations of up to
// kNoallocMaxParticles particles (3 doubles each, 24 MiB), one copy of the array
// per GPU the library runs on, held until the process ends.  Host code never
// hands the array's name to a kernel or a copy (in host code it is a host-side
// stand-in, not a device address): it uses the device address
// hipGetSymbolAddress returns, looked up once per device (bind_scratch).
constexpr long long kNoallocMaxParticles = 1ll << 20;
static __device__ __align__(256) double scratch[3 * kNoallocMaxParticles];

namespace {

struct Level {
    int n[3];
    double dh[3];
    double gamma, inv_gamma;
    long long off;  // offset in the MG pools
    long long nc;
};

struct Box {
    int n[3];
    double plo[3], phi[3], dx[3], dxi[3];
    double rhi[3], len[3];
    double inv_vol, dv, prob_size;
    int nlev;
    Level lev[kMaxLevels];
    long long cell_off;
    int tile[3], ntile[3];
    long long tile_off, ntiles;
    long long part_off = 0, np = 0;
    long long span = 0;  // particle slots (np plus the tile slack)
    bool has_mass = false;
    bool mass_array = false;  // uploaded with per-particle masses (else m_const)
    double m_const = 0.0, mass_total = 0.0, mass_offset = 0.0;
    int fx_exp = 0;
    long long G_off;
    int dc = 0;  // first level of the one-workgroup coarse V-cycle (nlev-1: bottom only)
};

struct EventPair {
    hipEvent_t a, b;
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

// Largest double below hi whose cell index stays < n (upper end kept by the
// periodic shift).
double domain_rhi(double lo, double hi, int n) {
    const double dxi = 1.0 / ((hi - lo) / (double)n);
    double r = std::nextafter(hi, -INFINITY);
    while ((int)std::floor((r - lo) * dxi) >= n) r = std::nextafter(r, -INFINITY);
    return r;
}

// Coarsening chain: halve all dimensions while all are even and stay >= 2.
int mg_levels(const int n[3], int dims[][3]) {
    int m[3] = {n[0], n[1], n[2]};
    int L = 0;
    for (;;) {
        dims[L][0] = m[0];
        dims[L][1] = m[1];
        dims[L][2] = m[2];
        ++L;
        if (L >= kMaxLevels) break;
        bool ok = true;
        for (int d = 0; d < 3; ++d)
            if ((m[d] & 1) || m[d] / 2 < 2) ok = false;
        if (!ok) break;
        for (int d = 0; d < 3; ++d) m[d] /= 2;
    }
    return L;
}

// Pseudo-inverse of the periodic 7-point operator on the bottom level, as a
// convolution kernel: G[c] = (1/N) sum_{k!=0} cos(2 pi sum_d ((k_d c_d) mod n_d)/n_d)
// / lambda_k, lambda_k = sum_d dh_d (2 cos(2 pi k_d/n_d) - 2).
void bottom_green(const int n[3], const double dh[3], double* G) {
    const double twopi = 2.0 * M_PI;
    const int N = n[0] * n[1] * n[2];
    for (int cz = 0; cz < n[2]; ++cz)
        for (int cy = 0; cy < n[1]; ++cy)
            for (int cx = 0; cx < n[0]; ++cx) {
                double acc = 0.0;
                for (int kz = 0; kz < n[2]; ++kz)
                    for (int ky = 0; ky < n[1]; ++ky)
                        for (int kx = 0; kx < n[0]; ++kx) {
                            if (kx == 0 && ky == 0 && kz == 0) continue;
                            double lam = dh[0] * (2.0 * cos(twopi * kx / n[0]) - 2.0)
                                       + dh[1] * (2.0 * cos(twopi * ky / n[1]) - 2.0)
                                       + dh[2] * (2.0 * cos(twopi * kz / n[2]) - 2.0);
                            double th = twopi * ((double)((kx * cx) % n[0]) / n[0]
                                               + (double)((ky * cy) % n[1]) / n[1]
                                               + (double)((kz * cz) % n[2]) / n[2]);
                            acc += cos(th) / lam;
                        }
                G[((size_t)cz * n[1] + cy) * n[0] + cx] = acc / (double)N;
            }
}

inline int nblocks(long long n, int per = kBlock) { return (int)((n + per - 1) / per); }

// V-cycles a captured solve records beyond the most a solve of the same call
// needed when run eagerly (converged boxes skip the extra cycles on the device).
constexpr int kGraphMargin = 2;


}  // namespace

constexpr int kIrFirstDefault = 5;
// with the extrapolated guess (phi_guess = 2) the first residual is about ten
// times smaller, and the FP32 correction reaches its precision limit (about
// 1e-7 of the residual it starts from) one cycle earlier
constexpr int kIrFirstExtrapolated = 4;

struct nyxpm_ctx {
    std::string err;
    int verbose = 0;
    int precision = NYXPM_FP64;
    int nu1 = 2, nu2 = 2, max_iter = 200;
    // V-cycles recorded per solve in a captured step: fixed (> 0) or, with 0, the
    // most cycles of the solves since the call began plus kGraphMargin
    int graph_cycles = 0;
    // mixed precision: FP32 V-cycles of a solve's first refinement step before its
    // first FP64 residual check (the correction carried over, phi updated once)
    int ir_first = kIrFirstDefault;
    int iters_seen = 0;  // most V-cycles of a box in the solves of the current call
    bool no_fused = false;  // force half-sweep smoothing (for comparisons)
    bool untiled_kick = false;  // force the thread-per-particle kick (for comparisons)
    bool separate_kicks = false;  // nyxpm_pm_steps without the kick / kick-drift pairing
    bool per_level_coarse = false;  // coarse levels by per-level launches, not k_coarse_vcycle
    bool fuse_res = false;          // level-0 residual formed inside the first sweep
    int sweep_zchunk = 0;       // fixed z-chunk of the fused sweep (0: chosen per level)
    int force_tx[2] = {0, 0}, force_ty[2] = {8, 8};  // fixed sweep shape per precision (0: chosen)
    bool fuse_update = true;    // phi += e in the last finest-level sweep
    int kick_chunk = kChunk;    // particles per work item of the tile-staged kicks
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<Box> boxes;
    int dmax = 1;  // max levels over boxes
    long long ncell_total = 0, ntiles_total = 0, npart_total = 0, mg_total = 0, G_total = 0;
    int halo_stride = 0;
    int tile[3] = {16, 8, 8};  // measured best at 512^3 clustered (the autotuner may pick another)

    // particles (two buffers for the tile sort)
    PartPtrs P[2] = {};
    int cur = 0;
    int* key = nullptr;
    long long part_cap = 0;
    bool any_mass = false;
    bool keys_valid = false, sorted_valid = false, density_valid = false;
    bool phi_valid = false;
    bool f1_valid = false;  // f holds rho - mass_offset and partial holds its block sums
    // f_raw: f holds rho - mass_offset and the residual kernels apply the
    // solvability offset and scaling on the fly, ((f - solv) * f_G) * f_inv_a
    bool f_raw = false;
    double f_G = 0.0, f_inv_a = 0.0;
    // initial guess of the solves inside nyxpm_pm_steps: 0 (default) and 1 =
    // previous phi, as Nyx; 2 = 2 phi_n - phi_{n-1}.
    // phi_prev holds phi_{n-1} when the device flag guess_ok is 1 (a device flag,
    // so that captured steps extrapolate exactly as eager ones do).
    int phi_guess = 0;
    double* phi_prev = nullptr;
    int* guess_ok = nullptr;

    // tiles: counts of the last binning, and the slot layout of the current
    // buffer (tile t owns slots tbase[t] .. tbase[t] + tcap[t] - 1, of which the
    // first cursor[t] have been handed out)
    int* counts = nullptr;
    int* cursor = nullptr;
    long long* tstart = nullptr;
    long long* tbase = nullptr;
    int* tcap = nullptr;
    // incremental re-sort: tile capacity cnt + cnt * slack_num / 1024 + slack_pad
    // (slack_num == 0: dense tiles, a full sort after every kick-drift); after a
    // kick-drift the particles that changed tile wait in a list (in the idle
    // buffer's y array) for k_migrate
    int slack_num = 256, slack_pad = 16;
    int hole_num = 64;            // full sort once the emptied slots exceed np * hole_num / 1024
    bool moves_pending = false;
    bool inc_suspended = false;   // no incremental re-sort (graph capture)
    long long holes = 0;          // slots emptied since the last full sort
    long long n_full_sorts = 0, n_inc_sorts = 0, n_moved = 0;
    int* nmove = nullptr;
    int* mig = nullptr;           // k_migrate flags: overflow, moved
    int* h_mig = nullptr;         // pinned mirror
    long long* bsum = nullptr;
    int scan_blocks = 0;
    TileItem* items = nullptr;  // tile chunks of the last sort (tile-staged kernels)
    int* nitems = nullptr;
    int* heavy = nullptr;       // tiles of more than kHeavyTile slots (deposit order)
    int* nheavy = nullptr;
    int heavy_cap = 0;
    int* ndefer = nullptr;  // entries of the list handed from the tile-staged kick to k_kick_deferred
    // particle time step (estTimestep): per-box minimum as bits, kept by the closing
    // kicks when cfl > 0; valid for the state after the last such kick (a = dt_a)
    unsigned long long* dtmin = nullptr;
    double cfl = 0.0;
    bool dt_valid = false;
    double dt_a = 0.0;
    long long items_cap = 0;
    bool items_valid = false;

    // level-0 fields
    double *own = nullptr, *halo = nullptr, *rho = nullptr, *f = nullptr, *phi = nullptr, *partial = nullptr;
    // MG pools
    double *r64 = nullptr, *e64 = nullptr, *e64b = nullptr;
    float *r32 = nullptr, *e32 = nullptr, *e32b = nullptr;
    double* G64 = nullptr;  // bottom Green's functions (FP64 in both precisions)
    // per-box device state
    int *active = nullptr, *iters = nullptr, *failed = nullptr, *err_flag = nullptr;
    unsigned long long *nr = nullptr, *nf = nullptr;
    double *state = nullptr, *solv = nullptr, *dv_size = nullptr;
    BoxDev* dboxes = nullptr;
    // host mirror of per-box flags
    int* h_flags = nullptr;  // pinned: active[nbox] iters[nbox] failed[nbox] err
    double* h_state = nullptr;

    // launch tables
    CellSeg* seg_pool = nullptr;
    std::vector<CellSeg> seg_host;
    struct Table {
        int first = 0, count = 0, blocks = 0;
    };
    std::vector<Table> t_full, t_half;  // per depth (non-bottom boxes)
    // residual + restriction per precision and depth: boxes the z-marching kernel
    // takes (t_rrz) and the rest (t_rrc, thread per coarse cell)
    std::vector<Table> t_rrz[2], t_rrc[2];
    // fused-sweep column shape and launch table per precision (0: FP64, 1: FP32)
    // and depth; tx == 0 means separate half sweeps at that depth
    struct SweepCfg {
        int tx = 0, ty = 8;
        int zc = 0;        // z-chunk chosen by the autotuner (0: occupancy heuristic)
        double us = 0.0;   // autotuned finest-launch time (0: not tuned)
        Table t;
    };
    std::vector<SweepCfg> sw[2];
    // FP32 levels whose pre- and post-smoothing pairs run as one pass each
    // (k_rb_sweep2): column shape, z-chunk and launch table per depth (count 0:
    // two k_rb_sweep passes)
    struct Sw2Cfg {
        int tx = kSw2TX, ty = kSw2TY;
        int zc = 0;        // z-chunk chosen by the autotuner (0: sw2_zc)
        double us = 0.0;   // autotuned launch time (0: not tuned)
        bool on = false;   // the pass is used (t: its pieces wherever the level fits)
        Table t;
    };
    std::vector<Sw2Cfg> sw2;
    bool sweep2 = true;  // opts.sweep2 >= 0
    int sw2_zc = 128;    // z-planes per block of k_rb_sweep2
    int sw2_tx = kSw2TX, sw2_ty = kSw2TY;  // its column (before autotuning)
    bool sw2_auto = true;  // default z-chunk (not set in the options)
    std::vector<int> pre_buf;    // buffer holding e after pre-smoothing, per depth
    std::vector<int> fin_buf;    // buffer holding e after post-smoothing, per depth
    Table t_bottom, t_l0, t_res;  // t_res: level 0, at most kReduceBlocks blocks per box
    // one-workgroup coarse V-cycle: per box, levels dc..nlev-1 (cc_first[b] ..
    // cc_first[b+1]-1 of table t_cc) and the LDS it needs per precision
    Table t_cc;
    int* cc_first = nullptr;
    size_t cc_lds[2] = {0, 0};
    Table t_resz;                 // level 0, z-marching residual columns
    BoxSeg* pseg = nullptr;
    int pseg_n = 0, pseg_blocks = 0;
    BoxSeg* tseg = nullptr;
    int tseg_n = 0, tseg_blocks = 0;

    // timing
    bool timing = false;
    std::set<std::string> timing_only;   // empty: every stage is timed
    std::map<std::string, std::vector<EventPair>> pending;
    std::map<std::string, std::pair<long long, double>> timed;
    std::vector<hipEvent_t> event_pool;

    // graph
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    double g_key[6] = {0, 0, 0, 0, 0, 0};
    int g_prec = -1;
    int g_mode = -1;    // 0 separate kicks, 1 faithful, 2 pair kicks
    int g_state = -1;   // host state the capture started from (particle buffer, binning flags)
    int g_cycles = 0;   // V-cycles recorded per solve
    // per box, over the solves of one replay: [b] 1 if a solve ended still above
    // its target, [nbox + b] the most V-cycles a solve used (k_graph_note)
    int* gstat = nullptr;
    int* h_gstat = nullptr;

    long long device_bytes = 0;
    // the device the context was created on and that device's `scratch` array
    // (hipGetSymbolAddress; shared by every context on the device)
    int device = 0;
    double* scratch_dev = nullptr;

    CellSeg* segs(const Table& t) const { return seg_pool + t.first; }
};

#define FAIL(ctx, code, ...)                         \
    do {                                             \
        char _b[512];                                \
        snprintf(_b, sizeof(_b), __VA_ARGS__);       \
        (ctx)->err = _b;                             \
        return (code);                               \
    } while (0)

#define HIPCK(ctx, call)                                                                      \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess) FAIL(ctx, NYXPM_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(_e)); \
    } while (0)

#define LAUNCHCK(ctx)                                                                          \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess) FAIL(ctx, NYXPM_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(_e)); \
    } while (0)

namespace {

template <typename T>
int dalloc(nyxpm_ctx* c, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) return NYXPM_OK;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        FAIL(c, NYXPM_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    }
    c->device_bytes += (long long)(n * sizeof(T));
    return NYXPM_OK;
}

template <typename T>
void dfree(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

// Timing: bracket a launch of class `name` with events when enabled.
struct Timed {
    nyxpm_ctx* c;
    std::string name;
    hipEvent_t a = nullptr, b = nullptr;
    hipEvent_t get() {
        if (!c->event_pool.empty()) {
            hipEvent_t e = c->event_pool.back();
            c->event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    Timed(nyxpm_ctx* c_, const char* n_) : c(c_), name(n_) {
        if (c->timing && !c->timing_only.empty() && !c->timing_only.count(name)) return;
        if (c->timing) {
            a = get();
            b = get();
            hipEventRecord(a, c->stream);
        }
    }
    ~Timed() {
        if (a) {   // bracketed when constructed
            hipEventRecord(b, c->stream);
            c->pending[name].push_back({a, b});
        }
    }
};

int collect_timing(nyxpm_ctx* c) {
    if (c->pending.empty()) return NYXPM_OK;
    HIPCK(c, hipStreamSynchronize(c->stream));
    for (auto& kv : c->pending) {
        auto& acc = c->timed[kv.first];
        for (auto& ep : kv.second) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ep.a, ep.b);
            acc.first += 1;
            acc.second += ms;
            c->event_pool.push_back(ep.a);
            c->event_pool.push_back(ep.b);
        }
    }
    c->pending.clear();
    return NYXPM_OK;
}

char g_name_buf[64];
const char* lvl_name(const char* base, int d) {
    snprintf(g_name_buf, sizeof(g_name_buf), "%s.d%d", base, d);
    return g_name_buf;
}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------

void tile_boxes(nyxpm_ctx* c);

int setup_boxes(nyxpm_ctx* c, int nbox, const nyxpm_geom* geoms) {
    c->boxes.resize(nbox);
    long long cell = 0, tiles = 0, mg = 0, Gtot = 0;
    c->dmax = 1;
    for (int b = 0; b < nbox; ++b) {
        Box& B = c->boxes[b];
        const nyxpm_geom& g = geoms[b];
        for (int d = 0; d < 3; ++d) {
            if (g.n_cell[d] < 2) FAIL(c, NYXPM_ERR_ARG, "box %d: n_cell[%d]=%d must be >= 2", b, d, g.n_cell[d]);
            if (g.n_cell[d] > 65536)
                FAIL(c, NYXPM_ERR_UNSUPPORTED, "box %d: n_cell[%d]=%d above 65536", b, d, g.n_cell[d]);
            if (!(g.prob_hi[d] > g.prob_lo[d])) FAIL(c, NYXPM_ERR_ARG, "box %d: prob_hi <= prob_lo", b);
            B.n[d] = g.n_cell[d];
            B.plo[d] = g.prob_lo[d];
            B.phi[d] = g.prob_hi[d];
            B.dx[d] = (B.phi[d] - B.plo[d]) / (double)B.n[d];
            B.dxi[d] = 1.0 / B.dx[d];
            B.rhi[d] = domain_rhi(B.plo[d], B.phi[d], B.n[d]);
            B.len[d] = B.phi[d] - B.plo[d];
        }
        B.dv = B.dx[0] * B.dx[1] * B.dx[2];
        B.inv_vol = 1.0 / B.dv;
        B.prob_size = (B.phi[0] - B.plo[0]) * (B.phi[1] - B.plo[1]) * (B.phi[2] - B.plo[2]);
        int dims[kMaxLevels][3];
        B.nlev = mg_levels(B.n, dims);
        long long nb = (long long)dims[B.nlev - 1][0] * dims[B.nlev - 1][1] * dims[B.nlev - 1][2];
        if (nb > kBottomMaxCells)
            FAIL(c, NYXPM_ERR_UNSUPPORTED,
                 "box %d: %dx%dx%d coarsens only to %dx%dx%d (%lld cells); the exact bottom solve "
                 "takes at most %d cells -- use dimensions of the form m*2^p with small m",
                 b, B.n[0], B.n[1], B.n[2], dims[B.nlev - 1][0], dims[B.nlev - 1][1], dims[B.nlev - 1][2], nb,
                 kBottomMaxCells);
        for (int l = 0; l < B.nlev; ++l) {
            Level& L = B.lev[l];
            for (int d = 0; d < 3; ++d) {
                L.n[d] = dims[l][d];
                double dxl = B.dx[d] * (double)(B.n[d] / dims[l][d]);
                double dxil = 1.0 / dxl;
                L.dh[d] = dxil * dxil;
            }
            L.gamma = -2.0 * (L.dh[0] + L.dh[1] + L.dh[2]);
            L.inv_gamma = 1.0 / L.gamma;
            L.nc = (long long)L.n[0] * L.n[1] * L.n[2];
            L.off = mg;
            mg += (L.nc + 63) / 64 * 64;  // every level starts 64-cell aligned (paired FP32 loads)
        }
        B.G_off = Gtot;
        Gtot += nb;
        B.cell_off = cell;
        cell += B.lev[0].nc;
        c->dmax = std::max(c->dmax, B.nlev);
    }
    c->ncell_total = cell;
    c->mg_total = mg;
    c->G_total = Gtot;
    tile_boxes(c);
    return NYXPM_OK;
}

// Particle tiles of every box for c->tile (a box smaller than the tile is one
// tile): counts, offsets, and the halo slots a tile's deposit leaves.
void tile_boxes(nyxpm_ctx* c) {
    long long tiles = 0;
    for (auto& B : c->boxes) {
        for (int d = 0; d < 3; ++d) {
            B.tile[d] = std::min(c->tile[d], B.n[d]);
            B.ntile[d] = (B.n[d] + B.tile[d] - 1) / B.tile[d];
        }
        B.tile_off = tiles;
        B.ntiles = (long long)B.ntile[0] * B.ntile[1] * B.ntile[2];
        tiles += B.ntiles;
    }
    c->ntiles_total = tiles;
    c->halo_stride = (c->tile[0] + 1) * (c->tile[1] + 1) * (c->tile[2] + 1) - c->tile[0] * c->tile[1] * c->tile[2];
}

// Blocks and z-chunk of one box's fused-sweep piece for the current shape of
// (precision pr, depth d): the autotuned z-chunk, else about four rounds of
// resident blocks (LDS- and thread-limited residency per CU), chunks of at least
// 8 planes.
void sweep_seg_shape(nyxpm_ctx* c, int pr, int d, const Level& L, CellSeg& w) {
    const int tx = c->sw[pr][d].tx, ty = c->sw[pr][d].ty;
    const long long cols = (long long)(L.n[0] / tx) * (L.n[1] / ty);
    const int sz = pr == 0 ? 8 : 4;
    const long long lds = (4LL * (ty + 4) * (tx + 4) + 2LL * (ty + 2) * (tx + 4)) * sz;
    const long long per_cu = std::max(1LL, std::min(2048LL / (tx * ty / 2), (160LL * 1024) / lds));
    const long long resident = 256LL * per_cu;
    long long nzc = (4 * resident + cols - 1) / std::max(1LL, cols);
    int zfix = c->sw[pr][d].zc > 0 ? c->sw[pr][d].zc : c->sweep_zchunk;
    if (zfix > 0) nzc = (L.n[2] + zfix - 1) / zfix;
    else nzc = std::max(1LL, std::min(nzc, (long long)std::max(1, L.n[2] / 8)));
    nzc = std::max(1LL, std::min(nzc, (long long)L.n[2]));
    w.zchunk = (int)((L.n[2] + nzc - 1) / nzc);
    nzc = (L.n[2] + w.zchunk - 1) / w.zchunk;
    w.nblk = (int)(cols * nzc);
}

// Whether a level suits k_rb_sweep2 with a tx x ty column.
bool sw2_fits(const Level& L, int tx, int ty) {
    return L.n[0] % tx == 0 && L.n[1] % ty == 0 && L.n[2] % 2 == 0 && L.n[2] >= 8 &&
           (long long)L.n[0] * L.n[1] < (1LL << 31);
}

// Blocks and z-chunk of one box's k_rb_sweep2 piece for depth d's current
// shape and z-chunk.
void sw2_seg_shape(nyxpm_ctx* c, int d, const Level& L, CellSeg& w) {
    const auto& s2 = c->sw2[d];
    w.zchunk = std::min(s2.zc > 0 ? s2.zc : c->sw2_zc, L.n[2]);
    w.nblk = (L.n[0] / s2.tx) * (L.n[1] / s2.ty) * ((L.n[2] + w.zchunk - 1) / w.zchunk);
}

// Coarse tile of k_resid_restrict_z: kRrCX x CY coarse columns, CY by precision
// (pr 0: FP64, 1: FP32) so that the LDS ring stays near 30 KB.
constexpr int kRrCX = 32, kRrCY64 = 4, kRrCY32 = 8;
inline int rrz_cy(int pr) { return pr == 0 ? kRrCY64 : kRrCY32; }

int build_tseg(nyxpm_ctx* c);

int build_tables(nyxpm_ctx* c) {
    std::vector<CellSeg>& S = c->seg_host;
    S.clear();
    const int nbox = (int)c->boxes.size();
    auto mk = [&](int b, int l, long long per) {
        const Box& B = c->boxes[b];
        const Level& L = B.lev[l];
        CellSeg s{};
        s.box = b;
        for (int d = 0; d < 3; ++d) {
            s.n[d] = L.n[d];
            s.dh[d] = L.dh[d];
        }
        s.gamma = L.gamma;
        s.inv_gamma = L.inv_gamma;
        s.off = L.off;
        s.coff = (l + 1 < B.nlev) ? B.lev[l + 1].off : 0;
        s.foff = B.cell_off;
        s.lev = l;
        s.nblk = nblocks(per);
        return s;
    };
    auto add_table = [&](std::vector<CellSeg> v) {
        nyxpm_ctx::Table t;
        t.first = (int)S.size();
        t.count = (int)v.size();
        int blk = 0;
        for (auto& s : v) {
            s.blk0 = blk;
            blk += s.nblk;
            S.push_back(s);
        }
        t.blocks = blk;
        return t;
    };
    // Fused sweeps where every box at the depth has a column shape dividing its
    // level and even dimensions; the buffer schedule follows from nu1/nu2 (it does
    // not depend on the column shape, so both precisions share it).
    c->t_full.assign(c->dmax, {});
    c->t_half.assign(c->dmax, {});
    c->sw2.assign(c->dmax, {});
    for (auto& s2 : c->sw2) {
        s2.tx = c->sw2_tx;
        s2.ty = c->sw2_ty;
    }
    for (int pr = 0; pr < 2; ++pr) {
        c->t_rrz[pr].assign(c->dmax, {});
        c->t_rrc[pr].assign(c->dmax, {});
    }
    c->pre_buf.assign(c->dmax, 0);
    c->fin_buf.assign(c->dmax, 0);
    static const int shapes64[][2] = {{64, 16}, {64, 8}, {32, 8}, {16, 8}};
    static const int shapes32[][2] = {{128, 8}, {64, 16}, {64, 8}, {32, 8}, {16, 8}};
    for (int pr = 0; pr < 2; ++pr) c->sw[pr].assign(c->dmax, {});
    // levels below the first with at most kCoarseCells cells run in one workgroup
    for (int b = 0; b < nbox; ++b) {
        Box& B = c->boxes[b];
        B.dc = B.nlev - 1;
        if (!c->per_level_coarse)
            for (int l = 0; l < B.nlev; ++l)
                if (B.lev[l].nc <= kCoarseCells) {
                    B.dc = l;
                    break;
                }
    }
    std::vector<bool> fusable(c->dmax, false);
    for (int d = 0; d + 1 < c->dmax; ++d) {
        bool any = false, even = true;
        for (int b = 0; b < nbox; ++b) {
            const Box& B = c->boxes[b];
            if (d >= B.dc) continue;
            any = true;
            const Level& L = B.lev[d];
            if (L.n[0] % 2 || L.n[1] % 2 || L.n[2] % 2) even = false;
        }
        fusable[d] = any && even && !c->no_fused;
        for (int pr = 0; pr < 2 && fusable[d]; ++pr) {
            const int ns = pr == 0 ? 4 : 5;
            for (int q = 0; q < ns; ++q) {
                const int tx = pr == 0 ? shapes64[q][0] : shapes32[q][0];
                const int ty = pr == 0 ? shapes64[q][1] : shapes32[q][1];
                if (c->force_tx[pr] && (tx != c->force_tx[pr] || ty != c->force_ty[pr])) continue;
                bool ok = true;
                for (int b = 0; b < nbox && ok; ++b) {
                    const Box& B = c->boxes[b];
                    if (d >= B.dc) continue;
                    const Level& L = B.lev[d];
                    if (L.n[0] % tx || L.n[1] % ty) ok = false;
                    // the wide shapes only where they still give >= 64 columns
                    if (!c->force_tx[pr] && (tx * ty > 512) && (long long)(L.n[0] / tx) * (L.n[1] / ty) < 64)
                        ok = false;
                }
                if (ok) {
                    c->sw[pr][d].tx = tx;
                    c->sw[pr][d].ty = ty;
                    break;
                }
            }
        }
        if (fusable[d] && (c->sw[0][d].tx == 0 || c->sw[1][d].tx == 0)) {
            c->sw[0][d].tx = c->sw[1][d].tx = 0;
            fusable[d] = false;
        }
        int cur = 0;
        if (fusable[d]) {
            for (int q = 1; q < c->nu1; ++q) cur ^= 1;
            c->pre_buf[d] = cur;
            for (int q = 0; q < c->nu2; ++q) cur ^= 1;
            c->fin_buf[d] = cur;
        }
    }
    // a box whose whole V-cycle runs in the coarse workgroup has no finest-level
    // sweep to carry the phi update
    c->fuse_update = true;
    for (int b = 0; b < nbox; ++b)
        if (c->boxes[b].dc == 0) c->fuse_update = false;
    auto coarse_buf = [&](const Box& B, int d) {  // buffer of level d+1's result
        return (d + 1 == B.dc) ? 0 : c->fin_buf[d + 1];
    };
    for (int d = 0; d + 1 < c->dmax; ++d) {
        std::vector<CellSeg> vf, vh, vs[2], vrz[2], vrc[2];
        for (int b = 0; b < nbox; ++b) {
            const Box& B = c->boxes[b];
            if (d >= B.dc) continue;
            const Level& L = B.lev[d];
            CellSeg f = mk(b, d, L.nc);
            f.ebuf = coarse_buf(B, d);
            vf.push_back(f);
            vh.push_back(mk(b, d, L.nc / 2));
            for (int pr = 0; pr < 2; ++pr) {
                // the z-marching kernel where the tile divides the level and the
                // level offers at least one block per CU (chunks of 16 coarse
                // planes, halved down to 2 while fewer than 1024 blocks)
                const int cy = rrz_cy(pr);
                long long cols = 0;
                int cz = 0, zc = 16;
                if (L.n[0] % (2 * kRrCX) == 0 && L.n[1] % (2 * cy) == 0 && L.n[2] % 2 == 0) {
                    cols = (long long)(L.n[0] / (2 * kRrCX)) * (L.n[1] / (2 * cy));
                    cz = L.n[2] / 2;
                    zc = std::min(zc, cz);
                    while (zc > 2 && cols * ((cz + zc - 1) / zc) < 1024) zc /= 2;
                }
                if (cols > 0 && cols * ((cz + zc - 1) / zc) >= 256) {
                    CellSeg z = mk(b, d, 1);
                    z.zchunk = zc;
                    z.nblk = (int)(cols * ((cz + zc - 1) / zc));
                    vrz[pr].push_back(z);
                } else {
                    vrc[pr].push_back(mk(b, d, L.nc / 8));
                }
            }
            for (int pr = 0; pr < 2; ++pr) {
                if (!c->sw[pr][d].tx) continue;
                CellSeg w = mk(b, d, 1);
                w.ebuf = coarse_buf(B, d);  // for the prolongation done on load
                sweep_seg_shape(c, pr, d, L, w);
                vs[pr].push_back(w);
            }
        }
        c->t_full[d] = add_table(vf);
        c->t_half[d] = add_table(vh);
        for (int pr = 0; pr < 2; ++pr) {
            c->sw[pr][d].t = add_table(vs[pr]);
            c->t_rrz[pr][d] = add_table(vrz[pr]);
            c->t_rrc[pr][d] = add_table(vrc[pr]);
        }
        // the two-sweep pass (FP32) where every box's level fits its column shape
        // and offers at least 1024 blocks, with V(2,2) cycles
        std::vector<CellSeg> v2;
        bool ok2 = c->sweep2 && c->sw[1][d].tx != 0 && c->nu1 == 2 && c->nu2 == 2 && !c->fuse_res;
        long long blocks2 = 0;
        for (int b = 0; b < nbox && ok2; ++b) {
            const Box& B = c->boxes[b];
            if (d >= B.dc) continue;
            const Level& L = B.lev[d];
            if (!sw2_fits(L, c->sw2_tx, c->sw2_ty)) {
                ok2 = false;
                break;
            }
            CellSeg w = mk(b, d, 1);
            w.ebuf = coarse_buf(B, d);
            sw2_seg_shape(c, d, L, w);
            blocks2 += w.nblk;
            v2.push_back(w);
        }
        // used with the default z-chunk where the level offers enough blocks (the
        // autotuner may switch it on elsewhere), with a z-chunk set in the
        // options wherever the shape fits
        if (!ok2) v2.clear();
        c->sw2[d].t = add_table(v2);
        c->sw2[d].on = ok2 && !v2.empty() && !(c->sw2_auto && blocks2 < 1024);
    }
    {
        std::vector<CellSeg> vb;
        for (int b = 0; b < nbox; ++b) {
            const Box& B = c->boxes[b];
            CellSeg s = mk(b, B.nlev - 1, 1);
            s.coff = B.G_off;
            s.nblk = 1;
            vb.push_back(s);
        }
        c->t_bottom = add_table(vb);
    }
    {
        std::vector<CellSeg> vcc;
        std::vector<int> firsts;
        c->cc_lds[0] = c->cc_lds[1] = 0;
        for (int b = 0; b < nbox; ++b) {
            const Box& B = c->boxes[b];
            firsts.push_back((int)vcc.size());
            long long cells = 0;
            for (int l = B.dc; l < B.nlev; ++l) {
                CellSeg s = mk(b, l, 1);
                if (l == B.nlev - 1) s.coff = B.G_off;
                vcc.push_back(s);
                cells += B.lev[l].nc * (l == B.dc ? 1 : 2);  // e of every level, r below the top
            }
            for (int pr = 0; pr < 2; ++pr)
                c->cc_lds[pr] = std::max(c->cc_lds[pr], (size_t)cells * (pr == 0 ? 8 : 4));
        }
        firsts.push_back((int)vcc.size());
        c->t_cc = add_table(vcc);
        int rc0 = dalloc(c, &c->cc_first, firsts.size());
        if (rc0) return rc0;
        HIPCK(c, hipMemcpy(c->cc_first, firsts.data(), firsts.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    {
        std::vector<CellSeg> v0;
        for (int b = 0; b < nbox; ++b) {
            CellSeg s = mk(b, 0, c->boxes[b].lev[0].nc);
            s.coff = c->boxes[b].cell_off;  // level-0 field pools
            s.ebuf = c->boxes[b].dc == 0 ? 0 : c->fin_buf[0];
            v0.push_back(s);
        }
        c->t_l0 = add_table(v0);
        std::vector<CellSeg> vr, vz;
        for (int b = 0; b < nbox; ++b) {
            CellSeg s = mk(b, 0, c->boxes[b].lev[0].nc);
            s.coff = c->boxes[b].cell_off;
            s.nblk = std::min(s.nblk, kReduceBlocks);
            vr.push_back(s);
            // z-marching residual: kResTX x kResTY columns, chunks of up to 64 planes
            const Level& L = c->boxes[b].lev[0];
            CellSeg z = mk(b, 0, 1);
            z.coff = c->boxes[b].cell_off;
            const int nbx = (L.n[0] + kResTX - 1) / kResTX, nby = (L.n[1] + kResTY - 1) / kResTY;
            z.zchunk = std::min(64, L.n[2]);
            z.nblk = nbx * nby * ((L.n[2] + z.zchunk - 1) / z.zchunk);
            vz.push_back(z);
        }
        c->t_res = add_table(vr);
        c->t_resz = add_table(vz);
    }
    int rc = dalloc(c, &c->seg_pool, S.size());
    if (rc) return rc;
    HIPCK(c, hipMemcpy(c->seg_pool, S.data(), S.size() * sizeof(CellSeg), hipMemcpyHostToDevice));

    return build_tseg(c);
}

// The tile table (one block per tile of each box).
int build_tseg(nyxpm_ctx* c) {
    const int nbox = (int)c->boxes.size();
    std::vector<BoxSeg> ts;
    int blk = 0;
    for (int b = 0; b < nbox; ++b) {
        BoxSeg s{b, blk, (int)c->boxes[b].ntiles, 0};
        blk += s.nblk;
        ts.push_back(s);
    }
    if (!c->tseg) {
        int rc = dalloc(c, &c->tseg, ts.size());
        if (rc) return rc;
    }
    HIPCK(c, hipMemcpy(c->tseg, ts.data(), ts.size() * sizeof(BoxSeg), hipMemcpyHostToDevice));
    c->tseg_n = nbox;
    c->tseg_blocks = blk;
    return NYXPM_OK;
}

int upload_boxdev(nyxpm_ctx* c) {
    std::vector<BoxDev> v(c->boxes.size());
    std::vector<double> dvs(2 * c->boxes.size());
    for (size_t b = 0; b < c->boxes.size(); ++b) {
        const Box& B = c->boxes[b];
        BoxDev& D = v[b];
        for (int d = 0; d < 3; ++d) {
            D.n[d] = B.n[d];
            D.tile[d] = B.tile[d];
            D.ntile[d] = B.ntile[d];
            D.plo[d] = B.plo[d];
            D.dxi[d] = B.dxi[d];
            D.rhi[d] = B.rhi[d];
            D.len[d] = B.len[d];
        }
        D.fx_exp = B.fx_exp;
        for (int d = 0; d < 3; ++d) D.tdiv[d] = (unsigned)(((1ull << 31) + B.tile[d] - 1) / B.tile[d]);
        D.cell_off = B.cell_off;
        D.tile_off = B.tile_off;
        D.part_off = B.part_off;
        D.np = B.np;
        D.span = B.span;
        D.inv_vol = B.inv_vol;
        D.mass_offset = B.mass_offset;
        D.m_const = B.m_const;
        D.has_mass = B.has_mass ? 1 : 0;
        D.halo_stride = c->halo_stride;
        dvs[2 * b] = B.dv;
        dvs[2 * b + 1] = B.prob_size;
    }
    HIPCK(c, hipMemcpyAsync(c->dboxes, v.data(), v.size() * sizeof(BoxDev), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(c->dv_size, dvs.data(), dvs.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return NYXPM_OK;
}

int build_pseg(nyxpm_ctx* c) {
    std::vector<BoxSeg> ps;
    int blk = 0;
    for (size_t b = 0; b < c->boxes.size(); ++b) {
        BoxSeg s{(int)b, blk, std::max(1, nblocks(c->boxes[b].span)), 0};
        blk += s.nblk;
        ps.push_back(s);
    }
    if (!c->pseg) {
        int rc = dalloc(c, &c->pseg, ps.size());
        if (rc) return rc;
    }
    HIPCK(c, hipMemcpy(c->pseg, ps.data(), ps.size() * sizeof(BoxSeg), hipMemcpyHostToDevice));
    c->pseg_n = (int)ps.size();
    c->pseg_blocks = blk;
    return NYXPM_OK;
}

// Buffers sized by the tile count: the deposit's halo slots, per-tile counts,
// cursors, starts, bases and capacities, the scan's block sums.
int alloc_tile_buffers(nyxpm_ctx* c) {
    int rc;
    if ((rc = dalloc(c, &c->halo, (size_t)c->ntiles_total * c->halo_stride))) return rc;
    if ((rc = dalloc(c, &c->counts, c->ntiles_total))) return rc;
    if ((rc = dalloc(c, &c->cursor, c->ntiles_total))) return rc;
    if ((rc = dalloc(c, &c->tstart, c->ntiles_total + 1))) return rc;
    if ((rc = dalloc(c, &c->tbase, c->ntiles_total))) return rc;
    if ((rc = dalloc(c, &c->tcap, c->ntiles_total))) return rc;
    c->scan_blocks = nblocks(c->ntiles_total, kBlock * kScanItems);
    if ((rc = dalloc(c, &c->bsum, c->scan_blocks + 1))) return rc;
    HIPCK(c, hipMemset(c->counts, 0, c->ntiles_total * sizeof(int)));
    return NYXPM_OK;
}

void free_tile_buffers(nyxpm_ctx* c) {
    dfree(c->halo); dfree(c->counts); dfree(c->cursor); dfree(c->tstart);
    dfree(c->tbase); dfree(c->tcap); dfree(c->bsum);
}

int alloc_fields(nyxpm_ctx* c) {
    const int nbox = (int)c->boxes.size();
    int rc;
    if ((rc = dalloc(c, &c->own, c->ncell_total))) return rc;
    c->rho = c->own;  // the fold turns the deposit's own cells into rho in place
    if ((rc = dalloc(c, &c->f, c->ncell_total))) return rc;
    if ((rc = dalloc(c, &c->phi, c->ncell_total))) return rc;
    if ((rc = alloc_tile_buffers(c))) return rc;
    if ((rc = dalloc(c, &c->partial, (size_t)c->t_l0.blocks))) return rc;
    if ((rc = dalloc(c, &c->nmove, 1))) return rc;
    if ((rc = dalloc(c, &c->mig, 2))) return rc;
    if ((rc = dalloc(c, &c->guess_ok, 1))) return rc;
    HIPCK(c, hipMemset(c->guess_ok, 0, sizeof(int)));
    HIPCK(c, hipHostMalloc((void**)&c->h_mig, 2 * sizeof(int), hipHostMallocDefault));
    if ((rc = dalloc(c, &c->r64, c->mg_total))) return rc;
    if ((rc = dalloc(c, &c->e64, c->mg_total))) return rc;
    if ((rc = dalloc(c, &c->e64b, c->mg_total))) return rc;
    if ((rc = dalloc(c, &c->G64, c->G_total))) return rc;
    if ((rc = dalloc(c, &c->active, nbox))) return rc;
    if ((rc = dalloc(c, &c->iters, nbox))) return rc;
    if ((rc = dalloc(c, &c->failed, nbox))) return rc;
    if ((rc = dalloc(c, &c->err_flag, 1))) return rc;
    if ((rc = dalloc(c, &c->nitems, 1))) return rc;
    if ((rc = dalloc(c, &c->nheavy, 1))) return rc;
    HIPCK(c, hipMemset(c->nheavy, 0, sizeof(int)));
    if ((rc = dalloc(c, &c->ndefer, 1))) return rc;
    if ((rc = dalloc(c, &c->dtmin, nbox))) return rc;
    if ((rc = dalloc(c, &c->nr, (size_t)nbox * kNormSlots))) return rc;
    if ((rc = dalloc(c, &c->nf, (size_t)nbox * kNormSlots))) return rc;
    if ((rc = dalloc(c, &c->state, 4 * nbox))) return rc;
    if ((rc = dalloc(c, &c->solv, nbox))) return rc;
    if ((rc = dalloc(c, &c->dv_size, 2 * nbox))) return rc;
    if ((rc = dalloc(c, &c->dboxes, nbox))) return rc;
    if ((rc = dalloc(c, &c->gstat, 2 * nbox))) return rc;
    HIPCK(c, hipHostMalloc((void**)&c->h_gstat, 2 * nbox * sizeof(int), hipHostMallocDefault));
    HIPCK(c, hipHostMalloc((void**)&c->h_flags, (3 * nbox + 1) * sizeof(int), hipHostMallocDefault));
    HIPCK(c, hipHostMalloc((void**)&c->h_state, 4 * nbox * sizeof(double), hipHostMallocDefault));
    HIPCK(c, hipMemset(c->phi, 0, c->ncell_total * sizeof(double)));
    HIPCK(c, hipMemset(c->nr, 0, (size_t)nbox * kNormSlots * sizeof(unsigned long long)));
    HIPCK(c, hipMemset(c->nf, 0, (size_t)nbox * kNormSlots * sizeof(unsigned long long)));
    HIPCK(c, hipMemset(c->err_flag, 0, sizeof(int)));
    // bottom Green's functions
    std::vector<double> G(c->G_total);
    for (auto& B : c->boxes) {
        const Level& L = B.lev[B.nlev - 1];
        bottom_green(L.n, L.dh, G.data() + B.G_off);
    }
    HIPCK(c, hipMemcpy(c->G64, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice));
    return NYXPM_OK;
}

int ensure_mixed_pools(nyxpm_ctx* c) {
    int rc;
    if (!c->r32 && (rc = dalloc(c, &c->r32, c->mg_total))) return rc;
    if (!c->e32 && (rc = dalloc(c, &c->e32, c->mg_total))) return rc;
    if (!c->e32b && (rc = dalloc(c, &c->e32b, c->mg_total))) return rc;
    return NYXPM_OK;
}

void free_particles(nyxpm_ctx* c) {
    for (int s = 0; s < 2; ++s) {
        dfree(c->P[s].x); dfree(c->P[s].y); dfree(c->P[s].z);
        dfree(c->P[s].vx); dfree(c->P[s].vy); dfree(c->P[s].vz);
        dfree(c->P[s].m); dfree(c->P[s].id);
    }
    dfree(c->key);
    dfree(c->items);
    dfree(c->heavy);
}

int alloc_particle_set(nyxpm_ctx* c, PartPtrs& P, long long n, bool mass) {
    int rc;
    if ((rc = dalloc(c, &P.x, n))) return rc;
    if ((rc = dalloc(c, &P.y, n))) return rc;
    if ((rc = dalloc(c, &P.z, n))) return rc;
    if ((rc = dalloc(c, &P.vx, n))) return rc;
    if ((rc = dalloc(c, &P.vy, n))) return rc;
    if ((rc = dalloc(c, &P.vz, n))) return rc;
    if (mass && (rc = dalloc(c, &P.m, n))) return rc;
    if ((rc = dalloc(c, &P.id, n))) return rc;
    return NYXPM_OK;
}

// The idle particle buffer, allocated on first use.
int ensure_idle_buffer(nyxpm_ctx* c) {
    PartPtrs& B = c->P[1 - c->cur];
    if (B.x) return NYXPM_OK;
    int rc = alloc_particle_set(c, B, std::max(c->part_cap, 1LL), c->any_mass);
    if (rc) {
        PartPtrs none = {};
        for (double** p : {&B.x, &B.y, &B.z, &B.vx, &B.vy, &B.vz, &B.m}) dfree(*p);
        dfree(B.id);
        B = none;
    }
    return rc;
}

// ---------------------------------------------------------------------------
// device phases
// ---------------------------------------------------------------------------

int err_check(nyxpm_ctx* c) {
    int e = 0;
    HIPCK(c, hipMemcpyAsync(&e, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (e) {
        hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream);
        FAIL(c, NYXPM_ERR_STATE,
             "particle outside one periodic image of its box or binning inconsistent (flags 0x%x)", e);
    }
    return NYXPM_OK;
}

int phase_keys(nyxpm_ctx* c) {
    PartPtrs& A = c->P[c->cur];
    {
        Timed t(c, "keys");
        hipLaunchKernelGGL(k_fill_int, dim3(nblocks(c->ntiles_total)), dim3(kBlock), 0, c->stream, c->counts,
                           c->ntiles_total, 0);
        hipLaunchKernelGGL(k_keys, dim3(c->pseg_blocks), dim3(kBlock), 0, c->stream, c->pseg, c->pseg_n, c->dboxes,
                           A.x, A.y, A.z, c->key, c->counts, c->err_flag);
    }
    LAUNCHCK(c);
    c->keys_valid = true;
    return NYXPM_OK;
}

int make_items(nyxpm_ctx* c) {
    Timed t(c, "items");
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->nitems, 1LL, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->nheavy, 1LL, 0);
    hipLaunchKernelGGL(k_make_items, dim3(nblocks(c->ntiles_total)), dim3(kBlock), 0, c->stream, c->tbase,
                       c->cursor, c->ntiles_total, c->items, c->nitems, c->kick_chunk, c->heavy, c->nheavy);
    LAUNCHCK(c);
    c->items_valid = true;
    return NYXPM_OK;
}

// Incremental re-sort after a kick-drift: the listed particles that changed tile
// move into their new tile's slack.  Returns in *done whether the layout is
// whole again (false: a tile ran out of slack, or the emptied slots passed
// their limit; the caller then sorts in full).  Reads two flags back.
int phase_migrate(nyxpm_ctx* c, bool* done) {
    PartPtrs& A = c->P[c->cur];
    const long long* moves = reinterpret_cast<const long long*>(c->P[1 - c->cur].y);
    {
        Timed t(c, "migrate");
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->mig, 2LL, 0);
        hipLaunchKernelGGL(k_migrate, dim3(1024), dim3(kBlock), 0, c->stream, moves, c->nmove, c->dboxes, c->tbase,
                           c->tcap, c->cursor, A, c->mig, c->err_flag);
    }
    LAUNCHCK(c);
    HIPCK(c, hipMemcpyAsync(c->h_mig, c->mig, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->holes += c->h_mig[1];
    c->n_moved += c->h_mig[1];
    *done = c->h_mig[0] == 0 && c->holes * 1024 <= c->npart_total * (long long)c->hole_num;
    if (*done) ++c->n_inc_sorts;
    if (*done) return make_items(c);
    return NYXPM_OK;
}

// Brings the particles into tile order for the deposit and the tile-staged
// kicks: incrementally after a kick-drift that listed its tile changes, else by
// a full counting sort into the other buffer (tiles with slack when slack_num > 0).
int phase_sort(nyxpm_ctx* c) {
    int rc;
    if (c->moves_pending) {
        c->moves_pending = false;
        bool done = false;
        if ((rc = phase_migrate(c, &done))) return rc;
        if (done) {
            c->sorted_valid = true;
            c->keys_valid = false;
            return NYXPM_OK;
        }
        c->keys_valid = false;
    }
    if (!c->keys_valid && (rc = phase_keys(c))) return rc;
    if ((rc = ensure_idle_buffer(c))) return rc;
    PartPtrs& A = c->P[c->cur];
    PartPtrs& B = c->P[1 - c->cur];
    const bool slack = c->slack_num > 0;
    {
        Timed t(c, "scan");
        const int* sizes = c->counts;
        if (slack) {
            hipLaunchKernelGGL(k_tile_caps, dim3(nblocks(c->ntiles_total)), dim3(kBlock), 0, c->stream, c->counts,
                               c->ntiles_total, c->slack_num, c->slack_pad, c->tcap);
            sizes = c->tcap;
        }
        hipLaunchKernelGGL(k_scan_reduce, dim3(c->scan_blocks), dim3(kBlock), 0, c->stream, sizes,
                           c->ntiles_total, c->bsum);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, c->stream, c->bsum, c->scan_blocks);
        hipLaunchKernelGGL(k_scan_down, dim3(c->scan_blocks), dim3(kBlock), 0, c->stream, sizes,
                           c->ntiles_total, c->bsum, c->tstart, c->cursor);
        hipLaunchKernelGGL(k_tile_base, dim3(nblocks(c->ntiles_total)), dim3(kBlock), 0, c->stream, c->tseg,
                           c->tseg_n, c->dboxes, c->tstart, c->ntiles_total, c->tbase);
    }
    {
        Timed t(c, "scatter");
        // slots the scatter leaves free start empty
        if (slack)
            hipLaunchKernelGGL(k_fill_dbl, dim3(nblocks(c->part_cap)), dim3(kBlock), 0, c->stream, B.x,
                               c->part_cap, std::nan(""));
        hipLaunchKernelGGL(k_scatter, dim3(c->pseg_blocks), dim3(kBlock), 0, c->stream, c->pseg, c->pseg_n,
                           c->dboxes, c->key, c->tbase, c->cursor, A, B);
    }
    LAUNCHCK(c);
    c->cur = 1 - c->cur;
    c->holes = 0;
    ++c->n_full_sorts;
    if ((rc = make_items(c))) return rc;
    c->sorted_valid = true;
    c->keys_valid = false;
    return NYXPM_OK;
}

// want_rho: also materialise the density (AssignDensity); the step needs only f.
// f32: FP32 per-particle weights (the mixed-precision step); AssignDensity
// itself, the drop-in, always deposits with FP64 weights.
int phase_deposit(nyxpm_ctx* c, bool want_rho, bool f32 = false) {
    int rc;
    if (!c->sorted_valid && (rc = phase_sort(c))) return rc;
    PartPtrs& A = c->P[c->cur];
    const size_t lds = (f32 ? 1 : 2) * sizeof(unsigned long long) * (c->tile[0] + 1) * (c->tile[1] + 1) *
                       (c->tile[2] + 1);
    {
        Timed t(c, "deposit");
        hipLaunchKernelGGL(f32 ? k_deposit<true> : k_deposit<false>, dim3(c->heavy_cap + c->tseg_blocks),
                           dim3(kBlock), lds, c->stream, c->tseg, c->tseg_n, c->dboxes, c->tbase, c->cursor,
                           c->heavy, c->nheavy, c->heavy_cap, A.x, A.y, A.z, A.m, c->own, c->halo, c->err_flag);
    }
    {
        Timed t(c, "fold");
        if (want_rho)
            hipLaunchKernelGGL(k_fold<true>, dim3(c->t_res.blocks), dim3(kBlock), 0, c->stream, c->segs(c->t_res),
                               c->t_res.count, c->dboxes, c->own, c->halo, c->rho, c->f, c->partial);
        else
            hipLaunchKernelGGL(k_fold<false>, dim3(c->t_res.blocks), dim3(kBlock), 0, c->stream,
                               c->segs(c->t_res), c->t_res.count, c->dboxes, c->own, c->halo, c->rho, c->f,
                               c->partial);
    }
    LAUNCHCK(c);
    c->density_valid = want_rho;
    c->f1_valid = true;
    c->f_raw = false;
    return NYXPM_OK;
}

// f from the current density: solvability offset and scaling.
int phase_rhs(nyxpm_ctx* c, double a, double four_pi_G) {
    const double Ggravity = -four_pi_G;
    const double inv_a = 1.0 / a;
    {
        Timed t(c, "rhs");
        if (!c->f1_valid)
            hipLaunchKernelGGL(k_rhs1, dim3(c->t_res.blocks), dim3(kBlock), 0, c->stream, c->segs(c->t_res),
                               c->t_res.count, c->dboxes, c->rho, c->f, c->partial);
        hipLaunchKernelGGL(k_rhs_sum, dim3(c->t_res.count), dim3(kBlock), 0, c->stream, c->segs(c->t_res),
                           c->partial, c->dv_size, c->solv);
    }
    LAUNCHCK(c);
    // the offset and scaling are applied where f is read (k_residual0), so f is
    // not rewritten; rhs_materialize writes the final values when asked for
    c->f1_valid = false;
    c->f_raw = true;
    c->f_G = Ggravity;
    c->f_inv_a = inv_a;
    return NYXPM_OK;
}

// Initial guess of a step's solve (nyxpm_pm_steps): phi_n as Nyx uses it, or on
// request the linear extrapolation 2 phi_n - phi_{n-1} in time.  The stopping rule
// is the same either way.
void invalidate_guess(nyxpm_ctx* c) {
    if (c->guess_ok) hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->guess_ok, 1LL, 0);
}

bool guess_wanted(const nyxpm_ctx* c) { return c->phi_guess == 2; }

// phi_{n-1}'s buffer; allocated before any graph capture (no hipMalloc inside one).
int ensure_guess_buffer(nyxpm_ctx* c) {
    if (c->phi_prev || !guess_wanted(c)) return NYXPM_OK;
    return dalloc(c, &c->phi_prev, c->ncell_total);
}

int prepare_guess(nyxpm_ctx* c) {
    if (!guess_wanted(c)) {
        invalidate_guess(c);
        return NYXPM_OK;
    }
    int rc;
    if ((rc = ensure_guess_buffer(c))) return rc;
    {
        Timed t(c, "guess");
        hipLaunchKernelGGL(k_phi_guess, dim3(nblocks(c->ncell_total)), dim3(kBlock), 0, c->stream, c->phi,
                           c->phi_prev, c->ncell_total, c->guess_ok);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->guess_ok, 1LL, 1);
    }
    LAUNCHCK(c);
    return NYXPM_OK;
}

// f = ((f - solv) * G) * (1/a) in place (the values the residual kernels use).
int rhs_materialize(nyxpm_ctx* c) {
    if (!c->f_raw) return NYXPM_OK;
    hipLaunchKernelGGL(k_rhs_final, dim3(c->t_l0.blocks), dim3(kBlock), 0, c->stream, c->segs(c->t_l0),
                       c->t_l0.count, c->dboxes, c->solv, c->f, c->f_G, c->f_inv_a);
    LAUNCHCK(c);
    c->f_raw = false;
    return NYXPM_OK;
}

template <typename T>
struct Pools;
template <>
struct Pools<double> {
    static double* r(nyxpm_ctx* c) { return c->r64; }
    static double* e(nyxpm_ctx* c) { return c->e64; }
    static double* e2(nyxpm_ctx* c) { return c->e64b; }
};
template <>
struct Pools<float> {
    static float* r(nyxpm_ctx* c) { return c->r32; }
    static float* e(nyxpm_ctx* c) { return c->e32; }
    static float* e2(nyxpm_ctx* c) { return c->e32b; }
};

template <typename T, int MODE>
void launch_sweep(nyxpm_ctx* c, int d, const T* ein, T* eout, const T* r, const T* ec0, const T* ec1,
                  SweepRes res = {}) {
    const auto& cfg = c->sw[sizeof(T) == 8 ? 0 : 1][d];
    const auto& t = cfg.t;
    const int key = cfg.tx * 100 + cfg.ty;
#define NYXPM_SWEEP(TX, TY)                                                                                  \
    case TX * 100 + TY:                                                                                      \
        hipLaunchKernelGGL((k_rb_sweep<T, TX, TY, MODE>), dim3(t.blocks), dim3(TX * TY / 2), 0, c->stream, \
                           c->segs(t), t.count, c->active, ein, eout, r, ec0, ec1, c->phi, res);            \
        break;
    switch (key) {
        NYXPM_SWEEP(128, 8)
        NYXPM_SWEEP(64, 16)
        NYXPM_SWEEP(64, 8)
        NYXPM_SWEEP(32, 8)
        NYXPM_SWEEP(16, 8)
        default:
            break;
    }
#undef NYXPM_SWEEP
}

// Two red-black sweeps in one pass (k_rb_sweep2, FP32) at depth d.
template <int MODE>
void launch_sweep2(nyxpm_ctx* c, int d, const float* ein, float* eout, const float* r, const float* ec0,
                   const float* ec1) {
    const auto& s2 = c->sw2[d];
    const auto& t = s2.t;
#define NYXPM_SW2(TX, TY)                                                                               \
    hipLaunchKernelGGL((k_rb_sweep2<float, TX, TY, MODE>), dim3(t.blocks), dim3(sw2_threads<TX, TY>()), 0, \
                       c->stream, c->segs(t), t.count, c->active, ein, eout, r, ec0, ec1, c->phi)
    if (s2.tx == 64 && s2.ty == 16)
        NYXPM_SW2(64, 16);
    else if (s2.tx == 32 && s2.ty == 32)
        NYXPM_SW2(32, 32);
    else
        NYXPM_SW2(32, 16);
#undef NYXPM_SW2
}

// Whether depth d runs its smoothing pairs as k_rb_sweep2 passes.
template <typename T>
bool two_sweep(const nyxpm_ctx* c, int d) {
    return sizeof(T) == 4 && d < (int)c->sw2.size() && c->sw2[d].on && c->sw2[d].t.count > 0;
}

// One V-cycle on every active box.  Returns in *updated whether phi += e has
// already been applied (fused into the last finest-level sweep).  With
// sweep0_done the first level-0 sweep was already run (by residual_sweep0).
// first = false: a further V-cycle on the same residual (the FP32 cycles of one
// refinement step in mixed precision): the level-0 correction starts from the
// previous cycle's instead of zero.  last = false: phi += e is left for a later
// cycle (no update fused into this one).
template <typename T>
int vcycle(nyxpm_ctx* c, bool* updated, bool sweep0_done = false, bool first = true, bool last = true) {
    T* r = Pools<T>::r(c);
    T* eb[2] = {Pools<T>::e(c), Pools<T>::e2(c)};
    const int pr = sizeof(T) == 8 ? 0 : 1;
    const int D = c->dmax;
    *updated = false;
    for (int d = 0; d + 1 < D; ++d) {
        auto& tf = c->t_full[d];
        auto& th = c->t_half[d];
        if (tf.count == 0) continue;
        int cur = 0;
        const bool keep = !first && d == 0;
        if (two_sweep<T>(c, d) && !(sweep0_done && d == 0)) {
            // both pre-smoothing sweeps in one pass, into the buffer the two
            // post-smoothing sweeps then read (so that they end in fin_buf)
            cur = 1 - c->pre_buf[d];
            float* e0 = reinterpret_cast<float*>(eb[0]);
            float* e1 = reinterpret_cast<float*>(eb[1]);
            float* ebf[2] = {e0, e1};
            const float* rf = reinterpret_cast<const float*>(r);
            if (keep) {
                Timed t(c, lvl_name("sweep2", d));
                launch_sweep2<0>(c, d, ebf[c->fin_buf[d]], ebf[cur], rf, e0, e1);
            } else {
                Timed t(c, lvl_name("sweep2_0", d));
                launch_sweep2<kSweepZero>(c, d, ebf[cur], ebf[cur], rf, e0, e1);
            }
        } else if (c->sw[pr][d].tx) {
            if (keep) {
                // the previous cycle left e in fin_buf; nu1 + nu2 is even, so the
                // nu1 sweeps end in pre_buf as after a cycle from zero
                cur = c->fin_buf[0];
                for (int q = 0; q < c->nu1; ++q) {
                    Timed t(c, lvl_name("sweep", d));
                    launch_sweep<T, 0>(c, d, eb[cur], eb[1 - cur], r, eb[0], eb[1]);
                    cur ^= 1;
                }
            } else {
                if (!(sweep0_done && d == 0)) {
                    Timed t(c, lvl_name("sweep0", d));
                    launch_sweep<T, kSweepZero>(c, d, eb[0], eb[0], r, eb[0], eb[1]);
                }
                for (int q = 1; q < c->nu1; ++q) {
                    Timed t(c, lvl_name("sweep", d));
                    launch_sweep<T, 0>(c, d, eb[cur], eb[1 - cur], r, eb[0], eb[1]);
                    cur ^= 1;
                }
            }
        } else {
            if (!keep) {
                Timed t(c, lvl_name("fill", d));
                hipLaunchKernelGGL(k_fill<T>, dim3(tf.blocks), dim3(kBlock), 0, c->stream, c->segs(tf), tf.count,
                                   c->active, eb[0], T(0));
            }
            for (int q = 0; q < c->nu1; ++q)
                for (int col = 0; col < 2; ++col) {
                    Timed t(c, lvl_name("smooth", d));
                    hipLaunchKernelGGL(k_gsrb_half<T>, dim3(th.blocks), dim3(kBlock), 0, c->stream, c->segs(th),
                                       th.count, c->active, eb[0], r, col);
                }
        }
        {
            Timed t(c, lvl_name("resid_restrict", d));
            const auto& tz = c->t_rrz[pr][d];
            const auto& tr = c->t_rrc[pr][d];
            constexpr int CY = sizeof(T) == 8 ? kRrCY64 : kRrCY32;
            if (tz.count)
                hipLaunchKernelGGL((k_resid_restrict_z<T, kRrCX, CY>), dim3(tz.blocks), dim3(kRrCX * CY), 0,
                                   c->stream, c->segs(tz), tz.count, c->active, eb[cur], r);
            if (tr.count)
                hipLaunchKernelGGL(k_resid_restrict<T>, dim3(tr.blocks), dim3(kBlock), 0, c->stream, c->segs(tr),
                                   tr.count, c->active, eb[cur], r);
        }
    }
    if (c->per_level_coarse) {
        Timed t(c, "bottom");
        hipLaunchKernelGGL(k_bottom<T>, dim3(c->t_bottom.count), dim3(kBlock), 0, c->stream, c->segs(c->t_bottom),
                           c->t_bottom.count, c->active, eb[0], r, c->G64);
    } else {
        Timed t(c, "coarse");
        hipLaunchKernelGGL(k_coarse_vcycle<T>, dim3((unsigned)c->boxes.size()), dim3(kCoarseThreads),
                           c->cc_lds[sizeof(T) == 8 ? 0 : 1], c->stream, c->segs(c->t_cc), c->cc_first, c->active,
                           eb[0], r, c->G64, c->nu1, c->nu2, first ? 0 : 1);
    }
    for (int d = D - 2; d >= 0; --d) {
        auto& tf = c->t_full[d];
        auto& th = c->t_half[d];
        if (tf.count == 0) continue;
        int cur = c->pre_buf[d];
        if (two_sweep<T>(c, d) && !(sweep0_done && d == 0)) {
            // both post-smoothing sweeps in one pass (prolongation on load), from
            // the pre-smoothing pass's buffer into fin_buf (or into phi)
            cur = 1 - c->pre_buf[d];
            float* e0 = reinterpret_cast<float*>(eb[0]);
            float* e1 = reinterpret_cast<float*>(eb[1]);
            float* ebf[2] = {e0, e1};
            const float* rf = reinterpret_cast<const float*>(r);
            const bool upd = last && d == 0 && c->fuse_update;
            if (upd) {
                Timed t(c, lvl_name("sweep2_u", d));
                launch_sweep2<kSweepProlong | kSweepUpdate>(c, d, ebf[cur], ebf[1 - cur], rf, e0, e1);
                *updated = true;
            } else {
                Timed t(c, lvl_name("sweep2_p", d));
                launch_sweep2<kSweepProlong>(c, d, ebf[cur], ebf[1 - cur], rf, e0, e1);
            }
        } else if (c->sw[pr][d].tx && c->nu2 > 0) {
            // prolongation fused into the first sweep; phi += e into the last one
            // on the finest level
            for (int q = 0; q < c->nu2; ++q) {
                const bool firstq = q == 0, lastq = q + 1 == c->nu2;
                const bool upd = lastq && last && d == 0 && c->fuse_update;
                Timed t(c, lvl_name(firstq ? "sweep_p" : (upd ? "sweep_u" : "sweep"), d));
                if (firstq && upd)
                    launch_sweep<T, kSweepProlong | kSweepUpdate>(c, d, eb[cur], eb[1 - cur], r, eb[0], eb[1]);
                else if (firstq)
                    launch_sweep<T, kSweepProlong>(c, d, eb[cur], eb[1 - cur], r, eb[0], eb[1]);
                else if (upd)
                    launch_sweep<T, kSweepUpdate>(c, d, eb[cur], eb[1 - cur], r, eb[0], eb[1]);
                else
                    launch_sweep<T, 0>(c, d, eb[cur], eb[1 - cur], r, eb[0], eb[1]);
                cur ^= 1;
                if (upd) *updated = true;
            }
        } else {
            {
                Timed t(c, lvl_name("prolong", d));
                hipLaunchKernelGGL(k_prolong<T>, dim3(tf.blocks), dim3(kBlock), 0, c->stream, c->segs(tf),
                                   tf.count, c->active, eb[cur], eb[0], eb[1]);
            }
            if (c->sw[pr][d].tx) {
                for (int q = 0; q < c->nu2; ++q) {
                    Timed t(c, lvl_name("sweep", d));
                    launch_sweep<T, 0>(c, d, eb[cur], eb[1 - cur], r, eb[0], eb[1]);
                    cur ^= 1;
                }
            } else {
                for (int q = 0; q < c->nu2; ++q)
                    for (int col = 0; col < 2; ++col) {
                        Timed t(c, lvl_name("smooth", d));
                        hipLaunchKernelGGL(k_gsrb_half<T>, dim3(th.blocks), dim3(kBlock), 0, c->stream,
                                           c->segs(th), th.count, c->active, eb[0], r, col);
                    }
            }
        }
    }
    LAUNCHCK(c);
    return NYXPM_OK;
}

// Whether the level-0 residual is formed inside the first sweep of the V-cycle:
// every box's level 0 has a fused sweep (no box runs its whole V-cycle in the
// coarse workgroup) and fused sweeps are not switched off.
template <typename T>
bool residual_in_sweep(const nyxpm_ctx* c) {
    const int pr = sizeof(T) == 8 ? 0 : 1;
    return c->fuse_res && c->fuse_update && !c->no_fused && c->dmax > 1 && c->sw[pr][0].tx != 0;
}

// r = f - L phi at level 0 with its norms, and the first sweep of the next
// V-cycle from it (k_rb_sweep with kSweepResidual): k_residual0 + sweep0 in one pass.
template <typename T>
int residual_sweep0(nyxpm_ctx* c, bool with_f) {
    Timed t(c, "res_sweep0");
    T* r = Pools<T>::r(c);
    T* e0 = Pools<T>::e(c);
    T* e1 = Pools<T>::e2(c);
    SweepRes res{c->f, c->f_raw ? c->solv : nullptr, c->f_G, c->f_inv_a, c->nr, with_f ? c->nf : nullptr};
    launch_sweep<T, kSweepZero | kSweepResidual>(c, 0, e0, e0, r, e0, e1, res);
    return NYXPM_OK;
}

template <typename T>
int residual(nyxpm_ctx* c, bool with_f) {
    Timed t(c, "residual");
    hipLaunchKernelGGL(k_residual0<T>, dim3(c->t_resz.blocks), dim3(kResTX * kResTY), 0, c->stream,
                       c->segs(c->t_resz), c->t_resz.count, c->active, c->f, c->f_raw ? c->solv : nullptr,
                       c->f_G, c->f_inv_a, c->phi, Pools<T>::r(c), c->nr, with_f ? c->nf : nullptr);
    return NYXPM_OK;
}

template <typename T>
int update(nyxpm_ctx* c) {
    Timed t(c, "update");
    hipLaunchKernelGGL(k_update0<T>, dim3(c->t_l0.blocks), dim3(kBlock), 0, c->stream, c->segs(c->t_l0),
                       c->t_l0.count, c->active, c->phi, Pools<T>::e(c), Pools<T>::e2(c));
    return NYXPM_OK;
}

// inc: the V-cycles run since the previous check (added to the per-box count)
int check(nyxpm_ctx* c, bool init, int fixed_iters, double rel, double abs_tol, int inc = 1) {
    const int nbox = (int)c->boxes.size();
    hipLaunchKernelGGL(k_check, dim3(nblocks(nbox, 64)), dim3(64), 0, c->stream, nbox, init ? 1 : 0, fixed_iters,
                       c->max_iter, rel, abs_tol, c->active, c->iters, c->failed, c->nr, c->nf, c->state, inc);
    LAUNCHCK(c);
    return NYXPM_OK;
}

int read_flags(nyxpm_ctx* c, bool* any_active) {
    const int nbox = (int)c->boxes.size();
    HIPCK(c, hipMemcpyAsync(c->h_flags, c->active, nbox * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    bool a = false;
    for (int b = 0; b < nbox; ++b) a = a || c->h_flags[b];
    *any_active = a;
    return NYXPM_OK;
}

// One multigrid solve of L phi = f on every box, from the current phi.  With
// `graph_cycles` > 0 the number of cycles issued is fixed (device flags stop
// converged boxes) and nothing is read back, so the sequence can be captured.
// V-cycles of the first refinement step of a solve before its first FP64
// residual check: c->ir_first FP32 cycles in mixed precision (the correction
// carried from cycle to cycle, phi updated once), one otherwise and with a fixed
// iteration count.  Later steps run one cycle each.
template <typename T>
int first_group(const nyxpm_ctx* c, int fixed_iters) {
    if (sizeof(T) == 8 || fixed_iters > 0 || c->ir_first <= 1 || (c->nu1 + c->nu2) % 2) return 1;
    return std::min(c->ir_first, c->max_iter);
}

// g V-cycles on the current residual (the later ones continue from the
// correction of the earlier ones), then phi += e; the caller forms the next
// residual.  rs: the first level-0 sweep of the first cycle was run already.
template <typename T>
int vcycle_group(nyxpm_ctx* c, int g, bool rs) {
    int rc;
    bool upd = false;
    for (int q = 0; q < g; ++q)
        if ((rc = vcycle<T>(c, &upd, rs && q == 0, q == 0, q + 1 == g))) return rc;
    if (!upd && (rc = update<T>(c))) return rc;
    return NYXPM_OK;
}

template <typename T>
int solve_impl(nyxpm_ctx* c, double rel, double abs_tol, int fixed_iters, int graph_cycles) {
    const int nbox = (int)c->boxes.size();
    int rc;
    // with the residual formed in the first sweep, every residual also starts the
    // next V-cycle (wasted only after the last one)
    const bool rs = residual_in_sweep<T>(c);
    auto res = [&](bool with_f) { return rs ? residual_sweep0<T>(c, with_f) : residual<T>(c, with_f); };
    const int g0 = first_group<T>(c, fixed_iters);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->active, (long long)nbox, 1);
    if ((rc = res(true))) return rc;
    if ((rc = check(c, true, fixed_iters, rel, abs_tol))) return rc;
    if (graph_cycles > 0) {
        // the same schedule as eagerly, graph_cycles V-cycles in all
        for (int done = 0, g = std::min(g0, graph_cycles); done < graph_cycles; done += g, g = 1) {
            if ((rc = vcycle_group<T>(c, g, rs))) return rc;
            if ((rc = res(false))) return rc;
            if ((rc = check(c, false, fixed_iters, rel, abs_tol, g))) return rc;
        }
        // a box still active now did not converge within the recorded cycles:
        // kept in gstat, which the next solve of the replay does not reset
        hipLaunchKernelGGL(k_graph_note, dim3(nblocks(nbox, 64)), dim3(64), 0, c->stream, nbox, c->active,
                           c->iters, c->gstat);
        LAUNCHCK(c);
        return NYXPM_OK;
    }
    const int limit = fixed_iters > 0 ? fixed_iters : c->max_iter;
    for (int it = 0, g = g0; it <= limit; it += g, g = 1) {
        bool any;
        if ((rc = read_flags(c, &any))) return rc;
        if (!any) break;
        if ((rc = vcycle_group<T>(c, g, rs))) return rc;
        if ((rc = res(false))) return rc;
        if ((rc = check(c, false, fixed_iters, rel, abs_tol, g))) return rc;
    }
    return NYXPM_OK;
}

int solve(nyxpm_ctx* c, double rel, double abs_tol, int fixed_iters, int graph_cycles) {
    c->dt_valid = false;  // phi changes
    if (c->precision == NYXPM_MIXED) {
        int rc = ensure_mixed_pools(c);
        if (rc) return rc;
        return solve_impl<float>(c, rel, abs_tol, fixed_iters, graph_cycles);
    }
    return solve_impl<double>(c, rel, abs_tol, fixed_iters, graph_cycles);
}

// Per-box results of the last solve; NYXPM_ERR_NOT_CONVERGED if any box failed.
int solve_results(nyxpm_ctx* c, int* iters, double* final_res, bool allow_active) {
    const int nbox = (int)c->boxes.size();
    HIPCK(c, hipMemcpyAsync(c->h_flags, c->active, nbox * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(c->h_flags + nbox, c->iters, nbox * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(c->h_flags + 2 * nbox, c->failed, nbox * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(c->h_state, c->state, 4 * nbox * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    int bad = -1;
    for (int b = 0; b < nbox; ++b) {
        c->iters_seen = std::max(c->iters_seen, c->h_flags[nbox + b]);
        if (iters) iters[b] = c->h_flags[nbox + b];
        if (final_res) final_res[b] = c->h_state[4 * b + 2];
        if (c->h_flags[2 * nbox + b] || (!allow_active && c->h_flags[b])) bad = b;
    }
    if (bad >= 0)
        FAIL(c, NYXPM_ERR_NOT_CONVERGED, "box %d: multigrid stopped at |r|=%g > target %g after %d cycles", bad,
             c->h_state[4 * bad + 2], c->h_state[4 * bad + 3], c->h_flags[nbox + bad]);
    return NYXPM_OK;
}

template <int MODE>
int phase_kick(nyxpm_ctx* c, double dt, double a_prev, double a_cur, double a_prev2 = 1.0, double a_cur2 = 1.0);

// MODE is kKick (a_prev, a_cur), kKickDrift (a_prev, a_cur) or kKickPair: the
// kick (a_prev, a_cur) of one step and the kick-drift (a_prev2, a_cur2) of the next.
// G is the interpolation arithmetic: double, or float in the mixed-precision
// mode (g formed in FP64 and kept in FP32, FP32 weights and sums; positions,
// velocities and the update stay FP64).
template <int MODE, typename G>
int phase_kick_g(nyxpm_ctx* c, double dt, double a_prev, double a_cur, double a_prev2, double a_cur2) {
    int rc;
    // a kick-drift's list of tile changes is used up before the particles move again
    if (c->moves_pending && (rc = phase_sort(c))) return rc;
    if ((rc = ensure_idle_buffer(c))) return rc;
    PartPtrs& A = c->P[c->cur];
    constexpr bool MOVE = MODE != kKick;
    // staged phi tile with a halo of 2, and the three components of g on the
    // (tile+1)^3 cells its particles interpolate from: written over the staged
    // phi where a thread's share of g fits in registers, else beside it
    const size_t lds_phi = sizeof(double) * (c->tile[0] + 3) * (c->tile[1] + 3) * (c->tile[2] + 3);
    const long long ng = (long long)(c->tile[0] + 1) * (c->tile[1] + 1) * (c->tile[2] + 1);
    const size_t lds_g = sizeof(G) * 3 * ng;
    const size_t lds = ng <= (long long)kBlock * kKickG ? std::max(lds_phi, lds_g) : lds_phi + lds_g;
    if (c->items_valid && !c->untiled_kick && lds <= 64 * 1024) {
        Timed t(c, MODE == kKickPair ? "kick_pair" : MOVE ? "kick_drift" : "kick");
        // incremental re-sort: the kick-drift lists the particles that change tile
        // (the list lives in the idle buffer's y array, like the deferred list in x)
        const int inc = MOVE && c->slack_num > 0 && !c->inc_suspended ? 1 : 0;
        long long* moves = reinterpret_cast<long long*>(c->P[1 - c->cur].y);
        if (inc)
            hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->nmove, 1LL, 0);
        else if (MOVE)
            hipLaunchKernelGGL(k_fill_int, dim3(nblocks(c->ntiles_total)), dim3(kBlock), 0, c->stream, c->counts,
                               c->ntiles_total, 0);
        // the deferred-particle list lives in the sort target's x array (as many
        // 8-byte slots as particles), idle until the next sort overwrites it
        long long* defer = reinterpret_cast<long long*>(c->P[1 - c->cur].x);
        // the closing kick (alone or in a pair) also estimates the next time step
        const double cfl = MODE != kKickDrift ? c->cfl : 0.0;
        if (cfl > 0)
            hipLaunchKernelGGL(k_fill_dbl, dim3(1), dim3(kBlock), 0, c->stream, reinterpret_cast<double*>(c->dtmin),
                               (long long)c->boxes.size(), 1e50);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->ndefer, 1LL, 0);
        auto kern = cfl > 0 ? k_kick_tiled<MODE, MODE != kKickDrift, G> : k_kick_tiled<MODE, false, G>;
        hipLaunchKernelGGL(kern, dim3((unsigned)c->items_cap), dim3(kBlock), lds, c->stream, c->tseg, c->tseg_n,
                           c->dboxes, c->items, c->nitems, c->phi, A.x, A.y, A.z, A.vx, A.vy, A.vz, dt, a_prev,
                           a_cur, a_prev2, a_cur2, c->key, c->counts, defer, c->ndefer, cfl, c->dtmin, inc, moves,
                           c->nmove, c->err_flag);
        hipLaunchKernelGGL((k_kick_deferred<MODE, G>), dim3(256), dim3(kBlock), 0, c->stream, defer, c->ndefer,
                           c->dboxes, c->phi, A.x, A.y, A.z, A.vx, A.vy, A.vz, dt, a_prev, a_cur, a_prev2, a_cur2,
                           c->key, c->counts, cfl, c->dtmin, inc, moves, c->nmove, c->err_flag);
        LAUNCHCK(c);
        c->dt_valid = cfl > 0;
        c->dt_a = a_cur;
        c->keys_valid = MOVE && !inc;
        c->moves_pending = inc != 0;
        c->sorted_valid = false;
        c->density_valid = false;
        return NYXPM_OK;
    }
    if (MODE == kKickPair) {
        int rc = phase_kick<kKick>(c, dt, a_prev, a_cur);
        return rc ? rc : phase_kick<kKickDrift>(c, dt, a_prev2, a_cur2);
    }
    c->dt_valid = false;
    if (MOVE) {
        Timed t(c, "kick_drift");
        hipLaunchKernelGGL(k_fill_int, dim3(nblocks(c->ntiles_total)), dim3(kBlock), 0, c->stream, c->counts,
                           c->ntiles_total, 0);
        hipLaunchKernelGGL((k_kick<true, G>), dim3(c->pseg_blocks), dim3(kBlock), 0, c->stream, c->pseg, c->pseg_n,
                           c->dboxes, c->phi, A.x, A.y, A.z, A.vx, A.vy, A.vz, dt, a_prev, a_cur, c->key,
                           c->counts, c->err_flag);
        c->keys_valid = true;
    } else {
        Timed t(c, "kick");
        hipLaunchKernelGGL((k_kick<false, G>), dim3(c->pseg_blocks), dim3(kBlock), 0, c->stream, c->pseg, c->pseg_n,
                           c->dboxes, c->phi, A.x, A.y, A.z, A.vx, A.vy, A.vz, dt, a_prev, a_cur, c->key,
                           c->counts, c->err_flag);
        c->keys_valid = false;
    }
    LAUNCHCK(c);
    c->sorted_valid = false;
    c->density_valid = false;
    return NYXPM_OK;
}

template <int MODE>
int phase_kick(nyxpm_ctx* c, double dt, double a_prev, double a_cur, double a_prev2, double a_cur2) {
    if (c->precision == NYXPM_MIXED) return phase_kick_g<MODE, float>(c, dt, a_prev, a_cur, a_prev2, a_cur2);
    return phase_kick_g<MODE, double>(c, dt, a_prev, a_cur, a_prev2, a_cur2);
}

int check_box(nyxpm_ctx* c, int box) {
    if (box < 0 || box >= (int)c->boxes.size()) FAIL(c, NYXPM_ERR_ARG, "box %d out of range", box);
    return NYXPM_OK;
}

int require_particles(nyxpm_ctx* c) {
    if (!c->P[0].x) FAIL(c, NYXPM_ERR_STATE, "no particles uploaded");
    return NYXPM_OK;
}

// Whether the deposits of a step (and of nyxpm_pm_init) use FP32 per-particle
// weights: in the mixed-precision mode.
bool mixed_deposit(const nyxpm_ctx* c) { return c->precision == NYXPM_MIXED; }

// The canonical KDK step, in stream order; with graph_cycles > 0 nothing is
// read back (capturable).
int step_body(nyxpm_ctx* c, double dt, double a_old, double a_new, double four_pi_G, double rel, double abs_tol,
              int nyx_faithful, int graph_cycles) {
    int rc;
    const double a_half = 0.5 * (a_old + a_new);
    if (nyx_faithful) {
        // Nyx_advance_particles.cpp:148-170: old phi re-solved from a zero guess
        if ((rc = phase_keys(c))) return rc;
        if ((rc = phase_sort(c))) return rc;
        if ((rc = phase_deposit(c, false, mixed_deposit(c)))) return rc;
        if ((rc = phase_rhs(c, a_old, four_pi_G))) return rc;
        hipLaunchKernelGGL(k_fill_dbl, dim3(nblocks(c->ncell_total)), dim3(kBlock), 0, c->stream, c->phi,
                           c->ncell_total, 0.0);
        if ((rc = solve(c, rel, abs_tol, 0, graph_cycles))) return rc;
        if (graph_cycles == 0 && (rc = solve_results(c, nullptr, nullptr, false))) return rc;
    }
    if ((rc = phase_kick<kKickDrift>(c, dt, a_old, a_half))) return rc;
    if ((rc = phase_sort(c))) return rc;
    if ((rc = phase_deposit(c, false, mixed_deposit(c)))) return rc;
    if ((rc = phase_rhs(c, a_new, four_pi_G))) return rc;
    if (nyx_faithful)
        invalidate_guess(c);  // Nyx's guess: the old phi just solved
    else if ((rc = prepare_guess(c)))
        return rc;
    if ((rc = solve(c, rel, abs_tol, 0, graph_cycles))) return rc;
    if ((rc = phase_kick<kKick>(c, dt, a_half, a_new))) return rc;
    return NYXPM_OK;
}

// A single-solve KDK step whose opening kick-drift was already done (by the
// previous step's pair kick or by the prologue of nyxpm_pm_steps).  With
// pair_next the closing kick is fused with the next step's kick-drift.
// next_a_new > 0: pair the closing kick with the next step's kick-drift, whose
// scale factors are next_a_old -> next_a_new (a_half = their mean).
int step_tail(nyxpm_ctx* c, double dt, double a_old, double a_new, double four_pi_G, double rel, double abs_tol,
              int graph_cycles, double next_a_old, double next_a_new) {
    int rc;
    const double a_half = 0.5 * (a_old + a_new);
    if ((rc = phase_sort(c))) return rc;
    if ((rc = phase_deposit(c, false, mixed_deposit(c)))) return rc;
    if ((rc = phase_rhs(c, a_new, four_pi_G))) return rc;
    if ((rc = prepare_guess(c))) return rc;
    if ((rc = solve(c, rel, abs_tol, 0, graph_cycles))) return rc;
    if (next_a_new > 0.0)
        return phase_kick<kKickPair>(c, dt, a_half, a_new, next_a_old, 0.5 * (next_a_old + next_a_new));
    return phase_kick<kKick>(c, dt, a_half, a_new);
}

// Nyx::integrate_comoving_a (comoving.cpp:630-689): RK2 (Heun) on
// da/dt = H0 sqrt(OmM/a + OmR/a^2 + OmL a^2), H0 = 100 h, OmL = 1 - OmM - OmR
// (comoving_type > 0; otherwise da/dt = h), with 2^(iter-1) substeps, doubling
// until two successive results agree to 1e-8 relative (at most 11 passes).
double integrate_comoving_a(double old_a, double dt, double om_m, double om_r, double h, int comoving_type) {
    if (h == 0.0) return old_a;
    const double H0 = h * 100.0;
    const double om_l = 1.0 - om_m - om_r;
    double prev = 2.0, out = old_a;
    for (int iter = 1; iter <= 11; ++iter) {
        const int nsteps = 1 << (iter - 1);
        const double Dt = dt / nsteps;
        double end_a = old_a;
        for (int j = 1; j <= nsteps; ++j) {
            const double start_a = end_a;
            const double s0 = comoving_type > 0
                                  ? H0 * std::sqrt(om_m / start_a + om_r / (start_a * start_a) + om_l * start_a * start_a)
                                  : h;
            end_a = start_a + s0 * Dt;
            const double s1 = comoving_type > 0
                                  ? H0 * std::sqrt(om_m / end_a + om_r / (end_a * end_a) + om_l * end_a * end_a)
                                  : h;
            end_a = start_a + 0.5 * (s0 + s1) * Dt;
        }
        out = end_a;
        if (std::fabs(1.0 - out / prev) <= 1e-8) break;
        prev = out;
    }
    return out;
}

// Whether p is device memory (pointer attributes; pageable host memory is not
// known to HIP, which reports an error that is cleared here).
bool is_device_ptr(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

// Device address of the current device's `scratch` array, looked up once per
// device with that device current and kept for the process; the context keeps it.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int bind_scratch(nyxpm_ctx* c) {
    int dev = 0;
    HIPCK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) FAIL(c, NYXPM_ERR_UNSUPPORTED, "device %d above %d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            FAIL(c, NYXPM_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    c->device = dev;
    c->scratch_dev = g_scratch[dev];
    return NYXPM_OK;
}

// Accelerations of box `box`'s particles (k_gather) into `out`, device memory of
// 3*np doubles; stream-ordered, no synchronisation.
int launch_gather(nyxpm_ctx* c, int box, double* out) {
    const Box& B = c->boxes[box];
    PartPtrs& A = c->P[c->cur];
    hipLaunchKernelGGL(c->precision == NYXPM_MIXED ? k_gather<float> : k_gather<double>, dim3(nblocks(B.span)),
                       dim3(kBlock), 0, c->stream, c->dboxes, box, c->phi, A.x, A.y, A.z, A.id, out, c->err_flag);
    LAUNCHCK(c);
    return NYXPM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int nyxpm_create_batch(nyxpm_ctx** out, int nbox, const nyxpm_geom* geoms, const nyxpm_opts* opts) {
    if (!out) return NYXPM_ERR_ARG;
    *out = nullptr;
    if (nbox < 1 || !geoms) return NYXPM_ERR_ARG;
    nyxpm_ctx* c = new nyxpm_ctx();
    if (opts) {
        if (opts->precision == NYXPM_MIXED) c->precision = NYXPM_MIXED;
        if (opts->nu1 > 0) c->nu1 = opts->nu1;
        if (opts->nu2 > 0) c->nu2 = opts->nu2;
        if (opts->max_iter > 0) c->max_iter = opts->max_iter;
        for (int d = 0; d < 3; ++d)
            if (opts->tile[d] > 0) c->tile[d] = opts->tile[d];
        c->verbose = opts->verbose;
        c->no_fused = opts->smoother == 1;
        c->untiled_kick = opts->kick == 1;
        c->separate_kicks = opts->separate_kicks != 0;
        if (opts->particle_cfl > 0) c->cfl = opts->particle_cfl;
        c->per_level_coarse = opts->coarse == 1;
        if (opts->sweep_tx > 0) {
            c->force_tx[0] = c->force_tx[1] = opts->sweep_tx;
            c->force_ty[0] = c->force_ty[1] = opts->sweep_ty > 0 ? opts->sweep_ty : 8;
        }
        if (opts->sweep_zchunk > 0) c->sweep_zchunk = opts->sweep_zchunk;
        if (opts->kick_chunk > 0) c->kick_chunk = std::max(kBlock, opts->kick_chunk);
        if (opts->phi_guess >= 0 && opts->phi_guess <= 2) c->phi_guess = opts->phi_guess;
        c->fuse_res = opts->residual == 1;
        if (opts->graph_cycles > 0) c->graph_cycles = opts->graph_cycles;
        if (opts->ir_first > 0)
            c->ir_first = opts->ir_first;
        else if (c->phi_guess == 2)
            c->ir_first = kIrFirstExtrapolated;
        if (opts->sweep2_tx > 0 || opts->sweep2_ty > 0) {
            const int tx = opts->sweep2_tx > 0 ? opts->sweep2_tx : kSw2TX;
            const int ty = opts->sweep2_ty > 0 ? opts->sweep2_ty : kSw2TY;
            if (!((tx == 32 && ty == 16) || (tx == 64 && ty == 16) || (tx == 32 && ty == 32))) {
                delete c;
                return NYXPM_ERR_ARG;
            }
            c->sw2_tx = tx;
            c->sw2_ty = ty;
        }
        if (opts->sweep2 < 0) c->sweep2 = false;
        if (opts->sweep2 > 1) {
            c->sw2_zc = std::max(8, opts->sweep2);
            c->sw2_auto = false;
        }
        if (opts->resort_slack < 0) {
            c->slack_num = 0;
            c->slack_pad = 0;
        } else if (opts->resort_slack > 0) {
            c->slack_num = std::max(1, (int)std::lround(std::min(opts->resort_slack, 4.0) * 1024.0));
        }
    }
    *out = c;
    if (opts && opts->resort_slack > 4.0)
        FAIL(c, NYXPM_ERR_ARG, "resort_slack %g above 4 (slots per particle)", opts->resort_slack);
    for (int d = 0; d < 3; ++d)
        if (c->tile[d] < 1 || c->tile[d] > 32) FAIL(c, NYXPM_ERR_ARG, "tile[%d]=%d outside [1,32]", d, c->tile[d]);
    if ((c->tile[0] + 1) * (c->tile[1] + 1) * (c->tile[2] + 1) * 16 > 64 * 1024)
        FAIL(c, NYXPM_ERR_ARG, "tile %dx%dx%d needs more than 64 KiB of LDS", c->tile[0], c->tile[1], c->tile[2]);
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) FAIL(c, NYXPM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    c->own_stream = true;
    int rc;
    if ((rc = bind_scratch(c))) return rc;
    if ((rc = setup_boxes(c, nbox, geoms))) return rc;
    if ((rc = build_tables(c))) return rc;
    if ((rc = alloc_fields(c))) return rc;
    if ((rc = build_pseg(c))) return rc;
    if ((rc = upload_boxdev(c))) return rc;
    return NYXPM_OK;
}

int nyxpm_create(nyxpm_ctx** out, const nyxpm_geom* geom, const nyxpm_opts* opts) {
    return nyxpm_create_batch(out, 1, geom, opts);
}

int nyxpm_destroy(nyxpm_ctx* c) {
    if (!c) return NYXPM_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    free_particles(c);
    dfree(c->counts); dfree(c->cursor); dfree(c->tstart); dfree(c->bsum);
    dfree(c->tbase); dfree(c->tcap); dfree(c->nmove); dfree(c->mig); dfree(c->phi_prev); dfree(c->guess_ok);
    if (c->h_mig) hipHostFree(c->h_mig);
    c->rho = nullptr;
    dfree(c->own); dfree(c->halo); dfree(c->f); dfree(c->phi); dfree(c->partial);

    dfree(c->r64); dfree(c->e64); dfree(c->e64b); dfree(c->G64);
    dfree(c->r32); dfree(c->e32); dfree(c->e32b);
    dfree(c->active); dfree(c->iters); dfree(c->failed); dfree(c->err_flag); dfree(c->nitems); dfree(c->nr); dfree(c->nf);
    dfree(c->ndefer); dfree(c->dtmin); dfree(c->nheavy);
    dfree(c->state); dfree(c->solv); dfree(c->dv_size); dfree(c->dboxes);
    dfree(c->seg_pool); dfree(c->pseg); dfree(c->tseg); dfree(c->cc_first);
    dfree(c->gstat);
    if (c->h_gstat) hipHostFree(c->h_gstat);
    if (c->h_flags) hipHostFree(c->h_flags);
    if (c->h_state) hipHostFree(c->h_state);
    for (auto& kv : c->pending)
        for (auto& ep : kv.second) { hipEventDestroy(ep.a); hipEventDestroy(ep.b); }
    for (auto ev : c->event_pool) hipEventDestroy(ev);
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    delete c;
    return NYXPM_OK;
}

int nyxpm_set_stream(nyxpm_ctx* c, void* s) {
    if (!c) return NYXPM_ERR_ARG;
    if (c->own_stream && c->stream) {
        hipStreamSynchronize(c->stream);
        hipStreamDestroy(c->stream);
    }
    c->own_stream = false;
    c->stream = (hipStream_t)s;
    if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
    return NYXPM_OK;
}

int nyxpm_sync(nyxpm_ctx* c) {
    if (!c) return NYXPM_ERR_ARG;
    HIPCK(c, hipStreamSynchronize(c->stream));
    return err_check(c);
}

const char* nyxpm_last_error(const nyxpm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int nyxpm_num_boxes(const nyxpm_ctx* c) { return c ? (int)c->boxes.size() : 0; }

int nyxpm_set_precision(nyxpm_ctx* c, int precision) {
    if (!c || (precision != NYXPM_FP64 && precision != NYXPM_MIXED)) return NYXPM_ERR_ARG;
    c->precision = precision;
    return NYXPM_OK;
}

int64_t nyxpm_num_particles(const nyxpm_ctx* c, int box) {
    if (!c || box < 0 || box >= (int)c->boxes.size()) return -1;
    return c->boxes[box].np;
}

int64_t nyxpm_device_bytes(const nyxpm_ctx* c) { return c ? c->device_bytes : 0; }

int nyxpm_sort_stats(const nyxpm_ctx* c, int64_t* out4) {
    if (!c || !out4) return NYXPM_ERR_ARG;
    out4[0] = c->n_full_sorts;
    out4[1] = c->n_inc_sorts;
    out4[2] = c->n_moved;
    out4[3] = c->holes;
    return NYXPM_OK;
}

// Particle slots of a box of np particles and ntiles tiles: np plus the largest
// total slack k_tile_caps can hand out.
static long long slot_count(const nyxpm_ctx* c, long long np, long long ntiles) {
    if (c->slack_num == 0) return np;
    return np + ((np * c->slack_num + 1023) >> 10) + (long long)c->slack_pad * ntiles;
}

static int upload_impl(nyxpm_ctx* c, int box, int64_t np, const double* x, const double* y, const double* z,
                       const double* vx, const double* vy, const double* vz, const double* m, double m_const,
                       hipMemcpyKind kind) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    if (np < 0 || np > (int64_t)0x7fffffff) FAIL(c, NYXPM_ERR_ARG, "np=%lld out of range", (long long)np);
    if (np > 0 && (!x || !y || !z || !vx || !vy || !vz)) FAIL(c, NYXPM_ERR_ARG, "null particle array");
    if (!m && !(m_const >= 0.0)) FAIL(c, NYXPM_ERR_ARG, "m_const must be >= 0");
    HIPCK(c, hipStreamSynchronize(c->stream));
    const int nbox = (int)c->boxes.size();
    // new layout: each box keeps np plus the tile slack as slots
    // (boxes without particles get their empty slack slots)
    std::vector<long long> new_off(nbox), old_off(nbox), old_span(nbox), new_span(nbox);
    long long tot = 0, slots = 0;
    for (int b = 0; b < nbox; ++b) {
        const Box& X = c->boxes[b];
        old_off[b] = X.part_off;
        old_span[b] = X.np > 0 ? X.span : 0;
        new_off[b] = slots;
        new_span[b] = b == box ? slot_count(c, np, X.ntiles) : X.np > 0 ? X.span : slot_count(c, 0, X.ntiles);
        tot += (b == box) ? np : X.np;
        slots += new_span[b];
    }
    const bool need_mass = c->any_mass || (m != nullptr);
    // the second buffer (sort target, kick lists, download staging) is allocated
    // when first needed, after the caller may have released its upload source
    PartPtrs N[2] = {};
    int* nkey = nullptr;
    long long cap = std::max(slots, 1LL);
    if ((rc = alloc_particle_set(c, N[0], cap, need_mass))) return rc;
    if ((rc = dalloc(c, &nkey, cap))) return rc;
    // keep the other boxes' slots as they are (empty slots included)
    PartPtrs& O = c->P[c->cur];
    for (int b = 0; b < nbox; ++b) {
        if (b == box || old_span[b] == 0) continue;
        const size_t nb8 = old_span[b] * sizeof(double);
        const long long so = old_off[b], d = new_off[b];
        HIPCK(c, hipMemcpyAsync(N[0].x + d, O.x + so, nb8, hipMemcpyDeviceToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].y + d, O.y + so, nb8, hipMemcpyDeviceToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].z + d, O.z + so, nb8, hipMemcpyDeviceToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].vx + d, O.vx + so, nb8, hipMemcpyDeviceToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].vy + d, O.vy + so, nb8, hipMemcpyDeviceToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].vz + d, O.vz + so, nb8, hipMemcpyDeviceToDevice, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].id + d, O.id + so, old_span[b] * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
        if (need_mass) {
            if (O.m) {
                HIPCK(c, hipMemcpyAsync(N[0].m + d, O.m + so, nb8, hipMemcpyDeviceToDevice, c->stream));
            } else {
                std::vector<double> mm(old_span[b], c->boxes[b].m_const);
                HIPCK(c, hipMemcpyAsync(N[0].m + d, mm.data(), nb8, hipMemcpyHostToDevice, c->stream));
                HIPCK(c, hipStreamSynchronize(c->stream));
            }
        }
    }
    const long long d = new_off[box];
    const size_t n8 = (size_t)np * sizeof(double);
    if (np > 0) {
        HIPCK(c, hipMemcpyAsync(N[0].x + d, x, n8, kind, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].y + d, y, n8, kind, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].z + d, z, n8, kind, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].vx + d, vx, n8, kind, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].vy + d, vy, n8, kind, c->stream));
        HIPCK(c, hipMemcpyAsync(N[0].vz + d, vz, n8, kind, c->stream));
        if (need_mass) {
            if (m) {
                HIPCK(c, hipMemcpyAsync(N[0].m + d, m, n8, kind, c->stream));
            } else {
                std::vector<double> mm(np, m_const);
                HIPCK(c, hipMemcpyAsync(N[0].m + d, mm.data(), n8, hipMemcpyHostToDevice, c->stream));
                HIPCK(c, hipStreamSynchronize(c->stream));
            }
        }
        hipLaunchKernelGGL(k_iota, dim3(nblocks(np)), dim3(kBlock), 0, c->stream, d, (long long)np, N[0].id);
        LAUNCHCK(c);
    }
    for (int b = 0; b < nbox; ++b) {
        // slots past the particles start empty
        const long long used = b == box ? np : old_span[b];
        if (new_span[b] > used) {
            hipLaunchKernelGGL(k_fill_dbl, dim3(nblocks(new_span[b] - used)), dim3(kBlock), 0, c->stream,
                               N[0].x + new_off[b] + used, new_span[b] - used, std::nan(""));
            LAUNCHCK(c);
        }
    }
    HIPCK(c, hipStreamSynchronize(c->stream));
    const long long icap = c->ntiles_total + slots / c->kick_chunk + 1;
    if (icap > 0x7fffffffLL) FAIL(c, NYXPM_ERR_ARG, "too many tile chunks");
    TileItem* nitm = nullptr;
    if ((rc = dalloc(c, &nitm, icap))) return rc;
    // a tile above kHeavyTile slots holds more than kHeavyTile particles or holes
    const long long hcap = std::min<long long>(c->ntiles_total, slots / kHeavyTile + 1);
    int* nheavy_list = nullptr;
    if ((rc = dalloc(c, &nheavy_list, hcap))) return rc;
    free_particles(c);
    c->P[0] = N[0];
    c->P[1] = N[1];
    c->key = nkey;
    c->items = nitm;
    c->items_cap = icap;
    dfree(c->heavy);
    c->heavy = nheavy_list;
    c->heavy_cap = (int)hcap;
    c->items_valid = false;
    c->dt_valid = false;
    c->cur = 0;
    c->part_cap = cap;
    c->any_mass = need_mass;
    for (int b = 0; b < nbox; ++b) {
        c->boxes[b].part_off = new_off[b];
        c->boxes[b].span = new_span[b];
        if (b == box) c->boxes[b].np = np;
    }
    c->npart_total = tot;
    c->moves_pending = false;
    invalidate_guess(c);
    // mass bookkeeping for the box (sumParticleMass and the fixed-point unit)
    Box& B = c->boxes[box];
    double mref = 0.0, total = 0.0;
    if (m) {
        std::vector<double> hm;
        const double* mp = m;
        if (kind == hipMemcpyDeviceToDevice) {
            hm.resize(np);
            HIPCK(c, hipMemcpy(hm.data(), m, n8, hipMemcpyDeviceToHost));
            mp = hm.data();
        }
        for (int64_t p = 0; p < np; ++p) {
            if (!(mp[p] >= 0.0)) FAIL(c, NYXPM_ERR_ARG, "particle %lld has negative or NaN mass", (long long)p);
            total += mp[p];
            mref = std::max(mref, mp[p]);
        }
        B.has_mass = true;
    } else {
        total = (double)np * m_const;
        mref = m_const;
        B.has_mass = false;
    }
    // boxes uploaded earlier without masses keep m_const; with a mass array present
    // every box reads the array
    for (auto& BB : c->boxes) BB.has_mass = need_mass;
    B.m_const = m ? 0.0 : m_const;
    B.mass_array = m != nullptr;
    B.mass_total = total;
    B.mass_offset = total / B.prob_size;
    B.fx_exp = mref > 0.0 ? std::ilogb(mref) : 0;
    if ((rc = upload_boxdev(c))) return rc;
    if ((rc = build_pseg(c))) return rc;
    c->keys_valid = c->sorted_valid = c->density_valid = false;
    if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
    return NYXPM_OK;
}

int nyxpm_particles_upload(nyxpm_ctx* c, int box, int64_t np, const double* x, const double* y, const double* z,
                           const double* vx, const double* vy, const double* vz, const double* m, double m_const) {
    return upload_impl(c, box, np, x, y, z, vx, vy, vz, m, m_const, hipMemcpyHostToDevice);
}

int nyxpm_particles_upload_device(nyxpm_ctx* c, int box, int64_t np, const double* x, const double* y,
                                  const double* z, const double* vx, const double* vy, const double* vz,
                                  const double* m, double m_const) {
    return upload_impl(c, box, np, x, y, z, vx, vy, vz, m, m_const, hipMemcpyDeviceToDevice);
}

// Nyx's binary particle file (particle_init_type = BinaryFile, read by
// DMPC->InitFromBinaryFile(file, AMREX_SPACEDIM + 1), NyxParticles.cpp:525-545):
// int64 NP, int32 DM, int32 NX, then NP records of DM positions followed by NX
// extra values, of which the first DM + 1 are the mass and the velocity (the
// particle's rdata(0..3), DarkMatterParticleContainer.H:14-18).  Values are
// FP64 or FP32, told apart by the file size.
int nyxpm_particles_load_binary(nyxpm_ctx* c, int box, const char* path, int64_t* np_out) {
    if (!c || !path) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    FILE* f = std::fopen(path, "rb");
    if (!f) FAIL(c, NYXPM_ERR_ARG, "cannot open %s", path);
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    long long np = 0;
    int dm = 0, nx = 0;
    if (std::fread(&np, sizeof(np), 1, f) != 1 || std::fread(&dm, sizeof(dm), 1, f) != 1 ||
        std::fread(&nx, sizeof(nx), 1, f) != 1)
        FAIL(c, NYXPM_ERR_ARG, "%s: short header", path);
    if (dm != 3 || nx < 4 || np < 0 || np > (long long)0x7fffffff)
        FAIL(c, NYXPM_ERR_ARG, "%s: header NP=%lld DM=%d NX=%d (need DM = 3, NX >= 4)", path, np, dm, nx);
    if (std::fseek(f, 0, SEEK_END) != 0) FAIL(c, NYXPM_ERR_ARG, "%s: cannot seek", path);
    const long long bytes = std::ftell(f) - 16;
    const long long vals = np * (long long)(dm + nx);
    const int vsize = vals > 0 && bytes == vals * 8 ? 8 : (vals > 0 && bytes == vals * 4 ? 4 : (vals == 0 ? 8 : 0));
    if (!vsize)
        FAIL(c, NYXPM_ERR_ARG, "%s: %lld data bytes do not match %lld values of 4 or 8 bytes", path, bytes, vals);
    if (std::fseek(f, 16, SEEK_SET) != 0) FAIL(c, NYXPM_ERR_ARG, "%s: cannot seek", path);
    std::vector<double> a(7 * (size_t)np);
    const int rec = dm + nx;
    std::vector<unsigned char> buf((size_t)rec * vsize * 65536);
    for (long long p0 = 0; p0 < np; p0 += 65536) {
        const long long cnt = std::min(65536LL, np - p0);
        if (std::fread(buf.data(), (size_t)rec * vsize, (size_t)cnt, f) != (size_t)cnt)
            FAIL(c, NYXPM_ERR_ARG, "%s: short read", path);
        for (long long q = 0; q < cnt; ++q) {
            for (int v = 0; v < 7; ++v) {  // x y z m vx vy vz
                const unsigned char* src = buf.data() + ((size_t)q * rec + v) * vsize;
                double d;
                if (vsize == 8) {
                    std::memcpy(&d, src, 8);
                } else {
                    float fl;
                    std::memcpy(&fl, src, 4);
                    d = fl;
                }
                a[(size_t)v * np + p0 + q] = d;
            }
        }
    }
    const double* A = a.data();
    rc = upload_impl(c, box, np, A, A + np, A + 2 * np, A + 4 * np, A + 5 * np, A + 6 * np, A + 3 * np, 0.0,
                     hipMemcpyHostToDevice);
    if (!rc && np_out) *np_out = np;
    return rc;
}

int nyxpm_particles_download(nyxpm_ctx* c, int box, double* x, double* y, double* z, double* vx, double* vy,
                             double* vz) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    const Box& B = c->boxes[box];
    if (B.np == 0) return NYXPM_OK;
    if ((rc = ensure_idle_buffer(c))) return rc;
    PartPtrs& A = c->P[c->cur];
    double* src[6] = {A.x, A.y, A.z, A.vx, A.vy, A.vz};
    double* dst[6] = {x, y, z, vx, vy, vz};
    // staging: the other particle buffer is free between phases
    double* tmp = c->P[1 - c->cur].x;
    for (int q = 0; q < 6; ++q) {
        if (!dst[q]) continue;
        hipLaunchKernelGGL(k_unsort, dim3(nblocks(B.span)), dim3(kBlock), 0, c->stream, B.part_off, B.span, A.x,
                           A.id, src[q], tmp);
        LAUNCHCK(c);
        HIPCK(c, hipMemcpyAsync(dst[q], tmp, B.np * sizeof(double), hipMemcpyDefault, c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
    }
    return err_check(c);
}

// ---------------------------------------------------------------------------
// particle files
// ---------------------------------------------------------------------------

// One component of box `box`'s particles (an array of the live buffer) into a
// host or device array, in upload (id) order.
static int download_component(nyxpm_ctx* c, int box, const double* src, double* dst) {
    const Box& B = c->boxes[box];
    double* tmp = c->P[1 - c->cur].x;  // the other particle buffer is free between phases
    hipLaunchKernelGGL(k_unsort, dim3(nblocks(B.span)), dim3(kBlock), 0, c->stream, B.part_off, B.span,
                       c->P[c->cur].x, c->P[c->cur].id, src, tmp);
    LAUNCHCK(c);
    HIPCK(c, hipMemcpyAsync(dst, tmp, B.np * sizeof(double), hipMemcpyDefault, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return NYXPM_OK;
}

// After an upload of positions that may lie outside the box: the periodic shift
// of RedistributeLocal (positions more than two periods away are flagged).
static int shift_into_box(nyxpm_ctx* c) {
    PartPtrs& A = c->P[c->cur];
    hipLaunchKernelGGL(k_wrap, dim3(c->pseg_blocks), dim3(kBlock), 0, c->stream, c->pseg, c->pseg_n, c->dboxes, A.x,
                       A.y, A.z, c->err_flag);
    LAUNCHCK(c);
    c->keys_valid = c->sorted_valid = c->density_valid = false;
    HIPCK(c, hipStreamSynchronize(c->stream));
    return err_check(c);
}

// Nyx's ASCII particle file (particle_init_type = AsciiFile, read by
// DMPC->InitFromAsciiFile(file, AMREX_SPACEDIM + 1), NyxParticles.cpp:506-521):
// the particle count, then per particle its position and the AMREX_SPACEDIM + 1
// extra values, mass and velocity (rdata(0..3), DarkMatterParticleContainer.H:
// 14-18), all whitespace separated.  A particle outside the box is shifted by
// whole periods into it, as AMReX's reader does before it places a particle.
int nyxpm_particles_load_ascii(nyxpm_ctx* c, int box, const char* path, int64_t* np_out) {
    if (!c || !path) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    FILE* f = std::fopen(path, "rb");
    if (!f) FAIL(c, NYXPM_ERR_ARG, "cannot open %s", path);
    std::string text;
    {
        char buf[1 << 16];
        size_t got;
        while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
        std::fclose(f);
    }
    const char* q = text.c_str();
    char* end = nullptr;
    const long long np = std::strtoll(q, &end, 10);
    if (end == q || np < 0 || np > (long long)0x7fffffff) FAIL(c, NYXPM_ERR_ARG, "%s: no particle count", path);
    q = end;
    std::vector<double> a(7 * (size_t)np);
    for (long long p = 0; p < np; ++p)
        for (int v = 0; v < 7; ++v) {  // x y z m vx vy vz
            const double d = std::strtod(q, &end);
            if (end == q)
                FAIL(c, NYXPM_ERR_ARG, "%s: particle %lld of %lld: value %d is missing or not a number", path, p,
                     np, v);
            q = end;
            a[(size_t)v * np + p] = d;
        }
    const double* A = a.data();
    rc = upload_impl(c, box, np, A, A + np, A + 2 * np, A + 4 * np, A + 5 * np, A + 6 * np, A + 3 * np, 0.0,
                     hipMemcpyHostToDevice);
    if (rc) return rc;
    if (np > 0 && (rc = shift_into_box(c))) return rc;
    if (np_out) *np_out = np;
    return NYXPM_OK;
}

static int make_dir(nyxpm_ctx* c, const std::string& d) {
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) FAIL(c, NYXPM_ERR_ARG, "cannot create directory %s", d.c_str());
    return NYXPM_OK;
}

// AMReX's particle checkpoint / plotfile of the dark-matter container, as Nyx
// writes it through NyxParticleContainer::NyxCheckpoint (is_checkpoint = 1) and
// WriteNyxPlotFile (0) (NyxParticleContainer.H:497-517; called by
// Nyx::particle_check_point and Nyx::particle_plot_file, Nyx_output.cpp:652,
// 735), one level, one grid:
//   <dir>/<name>/Header            text: version (real type as its last "_"
//                                  token), dim, extra real count and names
//                                  (mass xvel yvel zvel), extra int count,
//                                  is_checkpoint, particle count, next id,
//                                  finest level, grids per level, then per grid
//                                  "file count offset"
//   <dir>/<name>/Level_0/Particle_H  the BoxArray of the level
//   <dir>/<name>/Level_0/DATA_00000  checkpoint: count x (id, cpu) int32 first;
//                                  then count records of x y z mass xvel yvel
//                                  zvel, native-endian FP64
// and for a plotfile Nyx's <dir>/<name>/precision ("NATIVE", Nyx_output.cpp:
// 677-688).  Particles are written in upload order with ids 1..np, cpu 0.
int nyxpm_particles_write_amrex(nyxpm_ctx* c, int box, const char* dir, const char* name, int is_checkpoint) {
    if (!c || !dir || !name) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    if ((rc = require_particles(c))) return rc;
    const Box& B = c->boxes[box];
    const long long np = B.np;
    std::vector<double> cols(7 * (size_t)np);  // x y z m vx vy vz
    if (np > 0) {
        if ((rc = ensure_idle_buffer(c))) return rc;
        PartPtrs& A = c->P[c->cur];
        const double* src[7] = {A.x, A.y, A.z, A.m, A.vx, A.vy, A.vz};
        for (int v = 0; v < 7; ++v) {
            double* dst = cols.data() + (size_t)v * np;
            if (v == 3 && !A.m) {
                std::fill(dst, dst + np, B.m_const);
                continue;
            }
            if ((rc = download_component(c, box, src[v], dst))) return rc;
        }
        if ((rc = err_check(c))) return rc;
    }
    const std::string top = std::string(dir), base = top + "/" + name, lev = base + "/Level_0";
    if ((rc = make_dir(c, top)) || (rc = make_dir(c, base)) || (rc = make_dir(c, lev))) return rc;
    {
        FILE* h = std::fopen((base + "/Header").c_str(), "w");
        if (!h) FAIL(c, NYXPM_ERR_ARG, "cannot write %s/Header", base.c_str());
        std::fprintf(h, "Version_Two_Dot_Zero_double\n3\n4\nmass\nxvel\nyvel\nzvel\n0\n%d\n%lld\n%lld\n0\n1\n%d %lld %d\n",
                     is_checkpoint ? 1 : 0, np, np + 1, 0, np, 0);
        if (std::fclose(h) != 0) FAIL(c, NYXPM_ERR_ARG, "cannot write %s/Header", base.c_str());
    }
    {
        FILE* h = std::fopen((lev + "/Particle_H").c_str(), "w");
        if (!h) FAIL(c, NYXPM_ERR_ARG, "cannot write %s/Particle_H", lev.c_str());
        std::fprintf(h, "(1 0\n((0,0,0) (%d,%d,%d) (0,0,0))\n)\n", B.n[0] - 1, B.n[1] - 1, B.n[2] - 1);
        std::fclose(h);
    }
    {
        FILE* d = std::fopen((lev + "/DATA_00000").c_str(), "wb");
        if (!d) FAIL(c, NYXPM_ERR_ARG, "cannot write %s/DATA_00000", lev.c_str());
        bool ok = true;
        if (is_checkpoint) {
            std::vector<int> ids(2 * (size_t)np);
            for (long long p = 0; p < np; ++p) {
                ids[2 * p] = (int)(p + 1);
                ids[2 * p + 1] = 0;
            }
            ok = std::fwrite(ids.data(), sizeof(int), ids.size(), d) == ids.size();
        }
        std::vector<double> rec(7 * 65536);
        for (long long p0 = 0; ok && p0 < np; p0 += 65536) {
            const long long cnt = std::min(65536LL, np - p0);
            for (long long q = 0; q < cnt; ++q)
                for (int v = 0; v < 7; ++v) rec[7 * q + v] = cols[(size_t)v * np + p0 + q];
            ok = std::fwrite(rec.data(), sizeof(double), 7 * cnt, d) == (size_t)(7 * cnt);
        }
        if (std::fclose(d) != 0) ok = false;
        if (!ok) FAIL(c, NYXPM_ERR_ARG, "cannot write %s/DATA_00000", lev.c_str());
    }
    if (!is_checkpoint) {
        FILE* h = std::fopen((base + "/precision").c_str(), "w");
        if (!h) FAIL(c, NYXPM_ERR_ARG, "cannot write %s/precision", base.c_str());
        std::fprintf(h, "NATIVE\n");
        std::fclose(h);
    }
    return NYXPM_OK;
}

// DMPC->Restart(dir, name) (NyxParticles.cpp:586, 932): reads the AMReX particle
// checkpoint or plotfile written as above (any number of grids and DATA files on
// level 0, FP64 or FP32 reals, extra int components skipped) into box `box`.
// The first four extra reals are taken as mass, xvel, yvel, zvel, the order
// Nyx's containers name them (DarkMatterParticleContainer.H:14-18).
int nyxpm_particles_read_amrex(nyxpm_ctx* c, int box, const char* dir, const char* name, int64_t* np_out) {
    if (!c || !dir || !name) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    const std::string base = std::string(dir) + "/" + name;
    FILE* h = std::fopen((base + "/Header").c_str(), "r");
    if (!h) FAIL(c, NYXPM_ERR_ARG, "cannot open %s/Header", base.c_str());
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{h};
    char ver[256] = {0};
    int dim = 0, nre = 0, nie = 0, chk = 0, finest = 0, ngrids = 0;
    long long np = 0, next_id = 0;
    if (std::fscanf(h, "%255s %d %d", ver, &dim, &nre) != 3) FAIL(c, NYXPM_ERR_ARG, "%s/Header: bad start", base.c_str());
    const char* us = std::strrchr(ver, '_');
    const int rsize = us && !std::strcmp(us, "_double") ? 8 : (us && !std::strcmp(us, "_single") ? 4 : 0);
    if (!rsize) FAIL(c, NYXPM_ERR_ARG, "%s/Header: real type of '%s' not recognised", base.c_str(), ver);
    if (dim != 3 || nre < 4) FAIL(c, NYXPM_ERR_ARG, "%s/Header: dim %d with %d extra reals (need 3, >= 4)",
                                  base.c_str(), dim, nre);
    for (int q = 0; q < nre; ++q) {
        char nm[256];
        if (std::fscanf(h, "%255s", nm) != 1) FAIL(c, NYXPM_ERR_ARG, "%s/Header: real names", base.c_str());
    }
    if (std::fscanf(h, "%d", &nie) != 1) FAIL(c, NYXPM_ERR_ARG, "%s/Header: int count", base.c_str());
    for (int q = 0; q < nie; ++q) {
        char nm[256];
        if (std::fscanf(h, "%255s", nm) != 1) FAIL(c, NYXPM_ERR_ARG, "%s/Header: int names", base.c_str());
    }
    if (std::fscanf(h, "%d %lld %lld %d", &chk, &np, &next_id, &finest) != 4)
        FAIL(c, NYXPM_ERR_ARG, "%s/Header: counts", base.c_str());
    if (finest != 0) FAIL(c, NYXPM_ERR_UNSUPPORTED, "%s: particles on %d levels; one level only", base.c_str(),
                          finest + 1);
    if (np < 0 || np > (long long)0x7fffffff) FAIL(c, NYXPM_ERR_ARG, "%s/Header: %lld particles", base.c_str(), np);
    if (std::fscanf(h, "%d", &ngrids) != 1 || ngrids < 0) FAIL(c, NYXPM_ERR_ARG, "%s/Header: grids", base.c_str());
    const int nint = chk ? 2 + nie : 0, nreal = 3 + nre;
    std::vector<double> a(7 * (size_t)np);
    long long filled = 0;
    std::vector<unsigned char> buf;
    for (int g = 0; g < ngrids; ++g) {
        int which = 0;
        long long cnt = 0, where = 0;
        if (std::fscanf(h, "%d %lld %lld", &which, &cnt, &where) != 3)
            FAIL(c, NYXPM_ERR_ARG, "%s/Header: grid %d", base.c_str(), g);
        if (cnt == 0) continue;
        if (cnt < 0 || filled + cnt > np) FAIL(c, NYXPM_ERR_ARG, "%s/Header: grid counts exceed %lld", base.c_str(), np);
        char fn[64];
        std::snprintf(fn, sizeof(fn), "/Level_0/DATA_%05d", which);
        FILE* d = std::fopen((base + fn).c_str(), "rb");
        if (!d) FAIL(c, NYXPM_ERR_ARG, "cannot open %s%s", base.c_str(), fn);
        bool ok = std::fseek(d, (long)(where + cnt * nint * (long long)sizeof(int)), SEEK_SET) == 0;
        buf.resize((size_t)cnt * nreal * rsize);
        ok = ok && std::fread(buf.data(), 1, buf.size(), d) == buf.size();
        std::fclose(d);
        if (!ok) FAIL(c, NYXPM_ERR_ARG, "%s%s: short read (grid %d, %lld particles)", base.c_str(), fn, g, cnt);
        for (long long q = 0; q < cnt; ++q)
            for (int v = 0; v < 7; ++v) {
                const unsigned char* src = buf.data() + ((size_t)q * nreal + v) * rsize;
                double dv;
                if (rsize == 8) {
                    std::memcpy(&dv, src, 8);
                } else {
                    float fv;
                    std::memcpy(&fv, src, 4);
                    dv = fv;
                }
                a[(size_t)v * np + filled + q] = dv;
            }
        filled += cnt;
    }
    if (filled != np) FAIL(c, NYXPM_ERR_ARG, "%s: grids hold %lld of %lld particles", base.c_str(), filled, np);
    const double* A = a.data();
    rc = upload_impl(c, box, np, A, A + np, A + 2 * np, A + 4 * np, A + 5 * np, A + 6 * np, A + 3 * np, 0.0,
                     hipMemcpyHostToDevice);
    if (rc) return rc;
    if (np > 0 && (rc = shift_into_box(c))) return rc;
    if (np_out) *np_out = np;
    return NYXPM_OK;
}

int nyxpm_assign_density(nyxpm_ctx* c, int box, double* rho_out) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = require_particles(c))) return rc;
    if (box >= 0 && (rc = check_box(c, box))) return rc;
    if (rho_out && box < 0) FAIL(c, NYXPM_ERR_ARG, "rho_out needs a box index");
    if ((rc = phase_deposit(c, true))) return rc;
    if (rho_out) {
        const Box& B = c->boxes[box];
        HIPCK(c, hipMemcpyAsync(rho_out, c->rho + B.cell_off, B.lev[0].nc * sizeof(double), hipMemcpyDefault,
                                c->stream));
    }
    HIPCK(c, hipStreamSynchronize(c->stream));
    return err_check(c);
}

// AssignDensityAndVels into `out` (4 fields of box `box`'s cells, device memory
// when dst_dev, else written through the device buffer dst) with the given
// scratch: sums (4 N: folded mass and momenta of every box), mown (N) and mhalo
// (H) for the momentum deposit, maxbits (3 per box).  Stream-ordered; the
// caller synchronises.
static int density_and_vels_impl(nyxpm_ctx* c, int box, double* dst, double* sums, double* mown, double* mhalo,
                                 unsigned long long* maxbits) {
    PartPtrs& A = c->P[c->cur];
    const long long N = c->ncell_total;
    const int nbox = (int)c->boxes.size();
    const size_t lds = 2 * sizeof(unsigned long long) * (c->tile[0] + 1) * (c->tile[1] + 1) * (c->tile[2] + 1);
    hipLaunchKernelGGL(k_deposit<false>, dim3(c->heavy_cap + c->tseg_blocks), dim3(kBlock), lds, c->stream,
                       c->tseg, c->tseg_n, c->dboxes, c->tbase, c->cursor, c->heavy, c->nheavy, c->heavy_cap, A.x,
                       A.y, A.z, A.m, c->own, c->halo, c->err_flag);
    hipLaunchKernelGGL(k_fold_raw, dim3(c->t_res.blocks), dim3(kBlock), 0, c->stream, c->segs(c->t_res),
                       c->t_res.count, c->dboxes, c->own, c->halo, sums);
    // the per-box maxima start at zero on every call (the scratch keeps the last
    // call's contents)
    hipLaunchKernelGGL(k_fill_dbl, dim3(1), dim3(kBlock), 0, c->stream, reinterpret_cast<double*>(maxbits),
                       (long long)(3 * nbox), 0.0);
    if (c->pseg_blocks > 0)
        hipLaunchKernelGGL(k_max_mom, dim3(c->pseg_blocks), dim3(kBlock), 0, c->stream, c->pseg, c->pseg_n,
                           c->dboxes, A.x, A.m, A.vx, A.vy, A.vz, maxbits);
    const double* vel[3] = {A.vx, A.vy, A.vz};
    for (int d = 0; d < 3; ++d) {
        hipLaunchKernelGGL(k_deposit_mom, dim3(c->tseg_blocks), dim3(kBlock), lds, c->stream, c->tseg, c->tseg_n,
                           c->dboxes, c->tbase, c->cursor, A.x, A.y, A.z, A.m, vel[d], maxbits, d, mown, mhalo,
                           c->err_flag);
        hipLaunchKernelGGL(k_fold_raw, dim3(c->t_res.blocks), dim3(kBlock), 0, c->stream, c->segs(c->t_res),
                           c->t_res.count, c->dboxes, mown, mhalo, sums + (d + 1) * N);
    }
    const long long nc = c->boxes[box].lev[0].nc;
    hipLaunchKernelGGL(k_vels_finish, dim3(nblocks(nc)), dim3(kBlock), 0, c->stream, c->dboxes, box, sums,
                       sums + N, N, dst);
    LAUNCHCK(c);
    return NYXPM_OK;
}

int nyxpm_assign_density_and_vels(nyxpm_ctx* c, int box, double* out) {
    if (!c || !out) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = require_particles(c))) return rc;
    if ((rc = check_box(c, box))) return rc;
    if (!c->sorted_valid && (rc = phase_sort(c))) return rc;
    const long long N = c->ncell_total;
    const long long H = (long long)c->ntiles_total * c->halo_stride;
    const int nbox = (int)c->boxes.size();
    // scratch: mass and three momentum sums (all boxes), the momentum deposit's
    // own/halo cells, the per-box maxima
    double *sums = nullptr, *mown = nullptr, *mhalo = nullptr;
    unsigned long long* maxbits = nullptr;
    auto release = [&]() {
        const long long bytes = (sums ? 4 * N * 8 : 0) + (mown ? N * 8 : 0) + (mhalo ? std::max(H, 1LL) * 8 : 0) +
                                (maxbits ? 3LL * nbox * 8 : 0);
        dfree(sums);
        dfree(mown);
        dfree(mhalo);
        dfree(maxbits);
        c->device_bytes -= bytes;
    };
    if ((rc = dalloc(c, &sums, 4 * N)) || (rc = dalloc(c, &mown, N)) || (rc = dalloc(c, &mhalo, std::max(H, 1LL))) ||
        (rc = dalloc(c, &maxbits, 3 * nbox))) {
        release();
        return rc;
    }
    const long long nc = c->boxes[box].lev[0].nc;
    const bool dev = is_device_ptr(out);
    double* dst = out;
    if (!dev && (rc = dalloc(c, &dst, 4 * nc))) {
        release();
        return rc;
    }
    rc = density_and_vels_impl(c, box, dst, sums, mown, mhalo, maxbits);
    hipError_t e = rc ? hipSuccess : hipGetLastError();
    if (!rc && !dev && e == hipSuccess)
        e = hipMemcpyAsync(out, dst, 4 * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (!dev) {
        dfree(dst);
        c->device_bytes -= 4 * nc * (long long)sizeof(double);
    }
    release();
    // the deposit buffers now hold raw mass, not rho
    c->density_valid = false;
    c->f1_valid = false;
    if (rc) return rc;
    if (e != hipSuccess) FAIL(c, NYXPM_ERR_HIP, "assign_density_and_vels: %s", hipGetErrorString(e));
    return err_check(c);
}

int nyxpm_sum_particle_mass(nyxpm_ctx* c, int box, double* mass_out) {
    if (!c || !mass_out) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    *mass_out = c->boxes[box].mass_total;
    return NYXPM_OK;
}

int nyxpm_set_density(nyxpm_ctx* c, int box, const double* rho_in) {
    if (!c || !rho_in) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    const Box& B = c->boxes[box];
    HIPCK(c, hipMemcpyAsync(c->rho + B.cell_off, rho_in, B.lev[0].nc * sizeof(double), hipMemcpyDefault,
                            c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->density_valid = true;
    c->f1_valid = false;
    return NYXPM_OK;
}

int nyxpm_solve_for_phi(nyxpm_ctx* c, double a, double four_pi_G, double rel_tol, double abs_tol, int use_guess,
                        int* iters, double* final_resnorm) {
    if (!c) return NYXPM_ERR_ARG;
    if (!(a > 0.0)) FAIL(c, NYXPM_ERR_ARG, "a must be > 0");
    if (!c->density_valid) FAIL(c, NYXPM_ERR_STATE, "no density: call nyxpm_assign_density first");
    int rc;
    if ((rc = phase_rhs(c, a, four_pi_G))) return rc;
    if (!use_guess)
        hipLaunchKernelGGL(k_fill_dbl, dim3(nblocks(c->ncell_total)), dim3(kBlock), 0, c->stream, c->phi,
                           c->ncell_total, 0.0);
    if ((rc = solve(c, rel_tol, abs_tol, 0, 0))) return rc;
    if ((rc = solve_results(c, iters, final_resnorm, false))) return rc;
    c->phi_valid = true;
    invalidate_guess(c);
    return err_check(c);
}

int nyxpm_solve_poisson(nyxpm_ctx* c, int box, const double* rhs, double* phi_io, double rel_tol, double abs_tol,
                        int fixed_iters, int* iters, double* norms) {
    if (!c || !rhs || !phi_io) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    if (c->boxes.size() != 1) FAIL(c, NYXPM_ERR_ARG, "nyxpm_solve_poisson works on single-box contexts");
    const Box& B = c->boxes[box];
    HIPCK(c, hipMemcpyAsync(c->f + B.cell_off, rhs, B.lev[0].nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->f_raw = false;  // the caller's f is final
    HIPCK(c, hipMemcpyAsync(c->phi + B.cell_off, phi_io, B.lev[0].nc * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
    if ((rc = solve(c, rel_tol, abs_tol, fixed_iters, 0))) return rc;
    int it = 0;
    double fin = 0;
    rc = solve_results(c, &it, &fin, false);
    if (iters) *iters = it;
    if (norms) for (int q = 0; q < 4; ++q) norms[q] = c->h_state[q];
    HIPCK(c, hipMemcpyAsync(phi_io, c->phi + B.cell_off, B.lev[0].nc * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->phi_valid = true;
    invalidate_guess(c);
    return rc;
}

int nyxpm_get_grav_vector(nyxpm_ctx* c, int box, double* g_out) {
    if (!c || !g_out) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    const Box& B = c->boxes[box];
    const bool dev = is_device_ptr(g_out);
    double* out = g_out;
    if (!dev && (rc = dalloc(c, &out, 3 * B.lev[0].nc))) return rc;
    hipLaunchKernelGGL(k_grav_vector, dim3(nblocks(B.lev[0].nc)), dim3(kBlock), 0, c->stream, c->dboxes, box, c->phi,
                       out);
    hipError_t e = hipGetLastError();
    if (!dev && e == hipSuccess)
        e = hipMemcpyAsync(g_out, out, 3 * B.lev[0].nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (!dev) {
        hipFree(out);
        c->device_bytes -= 3 * B.lev[0].nc * (long long)sizeof(double);
    }
    if (e != hipSuccess) FAIL(c, NYXPM_ERR_HIP, "get_grav_vector: %s", hipGetErrorString(e));
    return NYXPM_OK;
}

int nyxpm_gather_accel(nyxpm_ctx* c, int box, double* g_out) {
    if (!c || !g_out) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    if ((rc = require_particles(c))) return rc;
    const Box& B = c->boxes[box];
    if (B.np == 0) return NYXPM_OK;
    const bool dev = is_device_ptr(g_out);
    double* out = g_out;
    if (!dev && (rc = dalloc(c, &out, 3 * B.np))) return rc;
    rc = launch_gather(c, box, out);
    hipError_t e = rc ? hipErrorLaunchFailure : hipSuccess;
    if (!dev && e == hipSuccess)
        e = hipMemcpyAsync(g_out, out, 3 * B.np * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (!dev) {
        hipFree(out);
        c->device_bytes -= 3 * B.np * (long long)sizeof(double);
    }
    if (rc) return rc;
    if (e != hipSuccess) FAIL(c, NYXPM_ERR_HIP, "gather_accel: %s", hipGetErrorString(e));
    return err_check(c);
}

int nyxpm_gather_accel_noalloc(nyxpm_ctx* c, int box, double* g_out) {
    if (!c || !g_out) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    if ((rc = require_particles(c))) return rc;
    const Box& B = c->boxes[box];
    if (B.np > kNoallocMaxParticles)
        FAIL(c, NYXPM_ERR_ARG, "gather_accel_noalloc: box %d holds %lld particles, above the %lld its staging array takes",
             box, B.np, kNoallocMaxParticles);
    if (B.np == 0) return NYXPM_OK;
    if (!c->scratch_dev) FAIL(c, NYXPM_ERR_STATE, "gather_accel_noalloc: no staging array on device %d", c->device);
    // device memory: written directly; host memory: through this device's array
    // (every element of its first 3*np doubles is written before the copy reads it)
    const bool dev = is_device_ptr(g_out);
    if ((rc = launch_gather(c, box, dev ? g_out : c->scratch_dev))) return rc;
    if (!dev)
        HIPCK(c, hipMemcpyAsync(g_out, c->scratch_dev, 3 * B.np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return err_check(c);
}

int nyxpm_est_time_step(nyxpm_ctx* c, double a, double cfl, double* dt_out) {
    if (!c || !dt_out) return NYXPM_ERR_ARG;
    if (!(a > 0.0) || !(cfl > 0.0)) FAIL(c, NYXPM_ERR_ARG, "a and cfl must be > 0");
    int rc;
    if ((rc = require_particles(c))) return rc;
    const int nbox = (int)c->boxes.size();
    // the estimate the last closing kick kept, if it is of this state, a and cfl
    if (!(c->dt_valid && a == c->dt_a && cfl == c->cfl)) {
        PartPtrs& A = c->P[c->cur];
        hipLaunchKernelGGL(k_fill_dbl, dim3(1), dim3(kBlock), 0, c->stream, reinterpret_cast<double*>(c->dtmin),
                           (long long)nbox, 1e50);
        if (c->pseg_blocks > 0)
            hipLaunchKernelGGL(c->precision == NYXPM_MIXED ? k_est_dt<float> : k_est_dt<double>, dim3(c->pseg_blocks), dim3(kBlock), 0, c->stream, c->pseg, c->pseg_n,
                               c->dboxes, c->phi, A.x, A.y, A.z, A.vx, A.vy, A.vz, a, cfl, c->dtmin);
        LAUNCHCK(c);
    }
    std::vector<double> v(nbox);
    HIPCK(c, hipMemcpyAsync(v.data(), c->dtmin, nbox * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < nbox; ++b) dt_out[b] = c->boxes[b].np > 0 ? v[b] : -1e50;
    return err_check(c);
}

static int field_io(nyxpm_ctx* c, int box, double* field, double* host, bool to_host) {
    if (!c || !host) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = check_box(c, box))) return rc;
    const Box& B = c->boxes[box];
    // host or device memory on the caller's side (unified addressing)
    if (to_host)
        HIPCK(c, hipMemcpyAsync(host, field + B.cell_off, B.lev[0].nc * sizeof(double), hipMemcpyDefault,
                                c->stream));
    else
        HIPCK(c, hipMemcpyAsync(field + B.cell_off, host, B.lev[0].nc * sizeof(double), hipMemcpyDefault,
                                c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return NYXPM_OK;
}

int nyxpm_get_phi(nyxpm_ctx* c, int box, double* out) { return field_io(c, box, c ? c->phi : nullptr, out, true); }
int nyxpm_set_phi(nyxpm_ctx* c, int box, const double* in) {
    int rc = field_io(c, box, c ? c->phi : nullptr, const_cast<double*>(in), false);
    if (!rc) {
        c->phi_valid = true;
        invalidate_guess(c);
        c->dt_valid = false;
    }
    return rc;
}
int nyxpm_get_rhs(nyxpm_ctx* c, int box, double* out) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = rhs_materialize(c))) return rc;
    return field_io(c, box, c->f, out, true);
}

int nyxpm_move_kick_drift(nyxpm_ctx* c, double dt, double a_old, double a_half) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = require_particles(c))) return rc;
    if (!(a_half > 0.0)) FAIL(c, NYXPM_ERR_ARG, "a_half must be > 0");
    if ((rc = phase_kick<kKickDrift>(c, dt, a_old, a_half))) return rc;
    return nyxpm_sync(c);
}

int nyxpm_move_kick(nyxpm_ctx* c, double dt, double a_new, double a_half) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = require_particles(c))) return rc;
    if (!(a_new > 0.0)) FAIL(c, NYXPM_ERR_ARG, "a_new must be > 0");
    if ((rc = phase_kick<kKick>(c, dt, a_half, a_new))) return rc;
    return nyxpm_sync(c);
}

int nyxpm_redistribute(nyxpm_ctx* c) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = require_particles(c))) return rc;
    PartPtrs& A = c->P[c->cur];
    hipLaunchKernelGGL(k_wrap, dim3(c->pseg_blocks), dim3(kBlock), 0, c->stream, c->pseg, c->pseg_n, c->dboxes, A.x,
                       A.y, A.z, c->err_flag);
    LAUNCHCK(c);
    // a shift can change a particle's tile (rounding at the upper boundary): the
    // next deposit sorts in full
    c->keys_valid = c->sorted_valid = c->density_valid = false;
    c->moves_pending = false;
    c->dt_valid = false;
    return nyxpm_sync(c);
}

int nyxpm_pm_init(nyxpm_ctx* c, double a, double four_pi_G, double rel_tol, double abs_tol) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = require_particles(c))) return rc;
    if ((rc = phase_deposit(c, true, mixed_deposit(c)))) return rc;
    if ((rc = phase_rhs(c, a, four_pi_G))) return rc;
    hipLaunchKernelGGL(k_fill_dbl, dim3(nblocks(c->ncell_total)), dim3(kBlock), 0, c->stream, c->phi,
                       c->ncell_total, 0.0);
    if ((rc = solve(c, rel_tol, abs_tol, 0, 0))) return rc;
    if ((rc = solve_results(c, nullptr, nullptr, false))) return rc;
    c->phi_valid = true;
    invalidate_guess(c);
    return err_check(c);
}

int nyxpm_pm_steps(nyxpm_ctx* c, int nsteps, double dt, double a_old, double a_new, double four_pi_G,
                   double rel_tol, double abs_tol, int nyx_faithful, int use_graph, int* iters) {
    if (!c || nsteps < 0) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = require_particles(c))) return rc;
    if (!c->phi_valid && !nyx_faithful) FAIL(c, NYXPM_ERR_STATE, "call nyxpm_pm_init before stepping");
    if (!(a_old > 0.0) || !(a_new > 0.0)) FAIL(c, NYXPM_ERR_ARG, "scale factors must be > 0");
    // Without the faithful re-solve, the closing kick of step s and the opening
    // kick-drift of step s+1 read g from the same phi at the same positions and
    // run as one particle pass (bit-identical to the separate kernels).
    const bool fused = !nyx_faithful && !c->separate_kicks;
    const double a_half = 0.5 * (a_old + a_new);
    // a captured step cannot read back k_migrate's flags: full sorts throughout
    struct Suspend {
        nyxpm_ctx* c;
        bool old;
        ~Suspend() { c->inc_suspended = old; }
    } suspend{c, c->inc_suspended};
    if (use_graph) {
        if (c->moves_pending && (rc = phase_sort(c))) return rc;
        if ((rc = ensure_guess_buffer(c))) return rc;
        c->inc_suspended = true;
    }
    c->iters_seen = 0;
    if (fused && nsteps > 0 && (rc = phase_kick<kKickDrift>(c, dt, a_old, a_half))) return rc;
    auto one_step = [&](int s) -> int {
        // every step of this call runs from a_old to a_new
        if (fused)
            return step_tail(c, dt, a_old, a_new, four_pi_G, rel_tol, abs_tol, 0, a_old, s + 1 < nsteps ? a_new : 0.0);
        return step_body(c, dt, a_old, a_new, four_pi_G, rel_tol, abs_tol, nyx_faithful, 0);
    };
    for (int s = 0; s < nsteps; ++s) {
        // a replay covers two steps; in the fused form both must end in a pair
        // kick.  The first step of a call runs eagerly: its V-cycle count sets
        // the cycles recorded in the graph.
        const bool replay = use_graph && s >= 1 && (fused ? s + 2 < nsteps : s + 1 < nsteps);
        if (!replay) {
            if ((rc = one_step(s))) return rc;
            if ((rc = solve_results(c, iters, nullptr, false))) return rc;
            continue;
        }
        const double key[6] = {dt, a_old, a_new, four_pi_G, rel_tol, abs_tol};
        const int mode = fused ? 2 : nyx_faithful ? 1 : 0;
        // the launches a capture records depend on which particle buffer is live
        // and on the binning flags; a replay is valid only from the same state
        const int state = c->cur | (c->keys_valid ? 2 : 0) | (c->items_valid ? 4 : 0) | (c->sorted_valid ? 8 : 0) |
                          (c->moves_pending ? 16 : 0);
        const int cycles = c->graph_cycles > 0 ? c->graph_cycles
                                               : std::min(c->max_iter, std::max(2, c->iters_seen + kGraphMargin));
        bool same = c->gexec && c->g_prec == c->precision && c->g_mode == mode && c->g_state == state &&
                    c->g_cycles >= cycles && std::memcmp(key, c->g_key, sizeof(key)) == 0;
        if (!same) {
            if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
            if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
            if (c->precision == NYXPM_MIXED && (rc = ensure_mixed_pools(c))) return rc;
            const bool was_timing = c->timing;
            c->timing = false;
            // the step's host-side bookkeeping (buffer swap) must be replayed
            // consistently: capture two steps so the particle buffers return to
            // their starting roles
            HIPCK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int cur0 = c->cur;
            for (int r = 0; r < 2 && !rc; ++r)
                rc = fused ? step_tail(c, dt, a_old, a_new, four_pi_G, rel_tol, abs_tol, cycles, a_old, a_new)
                           : step_body(c, dt, a_old, a_new, four_pi_G, rel_tol, abs_tol, nyx_faithful, cycles);
            hipGraph_t g = nullptr;
            hipError_t ee = hipStreamEndCapture(c->stream, &g);
            c->timing = was_timing;
            if (rc) return rc;
            if (ee != hipSuccess) FAIL(c, NYXPM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
            c->cur = cur0;
            c->graph = g;
            HIPCK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
            std::memcpy(c->g_key, key, sizeof(key));
            c->g_prec = c->precision;
            c->g_mode = mode;
            c->g_state = state;
            c->g_cycles = cycles;
        }
        const int nbox = (int)c->boxes.size();
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->gstat, 2LL * nbox, 0);
        HIPCK(c, hipGraphLaunch(c->gexec, c->stream));
        ++s;
        // every solve of the replay: a box still active after the recorded cycles
        // has not converged (the last solve's flags alone would hide an earlier one)
        HIPCK(c, hipMemcpyAsync(c->h_gstat, c->gstat, 2 * nbox * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        if ((rc = solve_results(c, iters, nullptr, true))) return rc;
        for (int b = 0; b < nbox; ++b) {
            c->iters_seen = std::max(c->iters_seen, c->h_gstat[nbox + b]);
            if (c->h_gstat[b])
                FAIL(c, NYXPM_ERR_NOT_CONVERGED,
                     "box %d: multigrid still above its target after the %d V-cycles recorded per solve in the "
                     "captured step", b, c->g_cycles);
        }
    }
    c->sorted_valid = false;
    c->keys_valid = false;
    c->density_valid = false;
    return err_check(c);
}

int nyxpm_integrate_comoving_a(double old_a, double dt, double omega_m, double omega_r, double h,
                               int comoving_type, double* new_a) {
    if (!new_a || !(old_a > 0.0)) return NYXPM_ERR_ARG;
    *new_a = integrate_comoving_a(old_a, dt, omega_m, omega_r, h, comoving_type);
    return NYXPM_OK;
}

int nyxpm_pm_steps_comoving(nyxpm_ctx* c, int nsteps, double dt, double* a_io, double omega_m, double omega_r,
                            double h, int comoving_type, double four_pi_G, double rel_tol, double abs_tol,
                            int nyx_faithful, int* iters) {
    if (!c || nsteps < 0 || !a_io) return NYXPM_ERR_ARG;
    int rc;
    if ((rc = require_particles(c))) return rc;
    if (!c->phi_valid && !nyx_faithful) FAIL(c, NYXPM_ERR_STATE, "call nyxpm_pm_init before stepping");
    if (!(*a_io > 0.0)) FAIL(c, NYXPM_ERR_ARG, "scale factor must be > 0");
    // a at the step boundaries (Nyx::integrate_comoving_a, once per step)
    std::vector<double> a(nsteps + 1);
    a[0] = *a_io;
    for (int s = 0; s < nsteps; ++s) {
        a[s + 1] = integrate_comoving_a(a[s], dt, omega_m, omega_r, h, comoving_type);
        if (!(a[s + 1] > 0.0)) FAIL(c, NYXPM_ERR_ARG, "scale factor became %g", a[s + 1]);
    }
    const bool fused = !nyx_faithful && !c->separate_kicks;
    if (fused && nsteps > 0 && (rc = phase_kick<kKickDrift>(c, dt, a[0], 0.5 * (a[0] + a[1])))) return rc;
    for (int s = 0; s < nsteps; ++s) {
        if (fused)
            rc = step_tail(c, dt, a[s], a[s + 1], four_pi_G, rel_tol, abs_tol, 0, a[s + 1],
                           s + 1 < nsteps ? a[s + 2] : 0.0);
        else
            rc = step_body(c, dt, a[s], a[s + 1], four_pi_G, rel_tol, abs_tol, nyx_faithful, 0);
        if (rc) return rc;
        if ((rc = solve_results(c, iters, nullptr, false))) return rc;
    }
    *a_io = a[nsteps];
    c->sorted_valid = false;
    c->keys_valid = false;
    c->density_valid = false;
    return err_check(c);
}

// Re-derive the fused-sweep pieces of (pr, d) in place (same boxes, new shape or
// z-chunk) and upload them.
static int rebuild_sweep_table(nyxpm_ctx* c, int pr, int d) {
    auto& t = c->sw[pr][d].t;
    int blk = 0;
    for (int q = 0; q < t.count; ++q) {
        CellSeg& w = c->seg_host[t.first + q];
        const Level& L = c->boxes[w.box].lev[d];
        sweep_seg_shape(c, pr, d, L, w);
        w.blk0 = blk;
        blk += w.nblk;
    }
    t.blocks = blk;
    HIPCK(c, hipMemcpy(c->seg_pool + t.first, c->seg_host.data() + t.first, t.count * sizeof(CellSeg),
                       hipMemcpyHostToDevice));
    return NYXPM_OK;
}

// Re-derive the residual-restriction pieces of (pr, d) in place with zc coarse
// planes per block, and upload them.
static int rebuild_rrz_table(nyxpm_ctx* c, int pr, int d, int zc) {
    auto& t = c->t_rrz[pr][d];
    const int cy = rrz_cy(pr);
    int blk = 0;
    for (int q = 0; q < t.count; ++q) {
        CellSeg& w = c->seg_host[t.first + q];
        const Level& L = c->boxes[w.box].lev[d];
        const long long cols = (long long)(L.n[0] / (2 * kRrCX)) * (L.n[1] / (2 * cy));
        const int cz = L.n[2] / 2;
        w.zchunk = std::min(zc, cz);
        w.nblk = (int)(cols * ((cz + w.zchunk - 1) / w.zchunk));
        w.blk0 = blk;
        blk += w.nblk;
    }
    t.blocks = blk;
    HIPCK(c, hipMemcpy(c->seg_pool + t.first, c->seg_host.data() + t.first, t.count * sizeof(CellSeg),
                       hipMemcpyHostToDevice));
    return NYXPM_OK;
}

// Re-derive the k_rb_sweep2 pieces of depth d in place for its current shape
// and z-chunk, and upload them.
static int rebuild_sw2_table(nyxpm_ctx* c, int d) {
    auto& t = c->sw2[d].t;
    int blk = 0;
    for (int q = 0; q < t.count; ++q) {
        CellSeg& w = c->seg_host[t.first + q];
        sw2_seg_shape(c, d, c->boxes[w.box].lev[d], w);
        w.blk0 = blk;
        blk += w.nblk;
    }
    t.blocks = blk;
    HIPCK(c, hipMemcpy(c->seg_pool + t.first, c->seg_host.data() + t.first, t.count * sizeof(CellSeg),
                       hipMemcpyHostToDevice));
    return NYXPM_OK;
}

// Every box's particles in upload (id) order, in device arrays: x, y, z, vx,
// vy, vz and (for a box uploaded with per-particle masses) m, 7 per box.
struct ParticleSnapshot {
    std::vector<double*> a;
};

static void snapshot_free(ParticleSnapshot& s) {
    for (double*& p : s.a) dfree(p);
    s.a.clear();
}

static int snapshot_take(nyxpm_ctx* c, ParticleSnapshot& s) {
    int rc;
    HIPCK(c, hipStreamSynchronize(c->stream));
    if ((rc = ensure_idle_buffer(c))) return rc;
    const int nbox = (int)c->boxes.size();
    s.a.assign(7 * (size_t)nbox, nullptr);
    const PartPtrs& A = c->P[c->cur];
    for (int b = 0; b < nbox; ++b) {
        const Box& B = c->boxes[b];
        if (B.np == 0) continue;
        const double* src[7] = {A.x, A.y, A.z, A.vx, A.vy, A.vz, B.mass_array ? A.m : nullptr};
        for (int q = 0; q < 7; ++q) {
            if (!src[q]) continue;
            if ((rc = dalloc(c, &s.a[7 * b + q], B.np))) return rc;
            if ((rc = download_component(c, b, src[q], s.a[7 * b + q]))) return rc;
        }
    }
    return err_check(c);
}

// Re-tiles every box with particle tile t and puts the snapshot's particles back
// through the upload path, which lays out the slots for the new tile count.
// Positions, velocities and phi come back bit for bit; the next deposit adds a
// tile-boundary cell's parts in the new tiles' FP64 fold, so the density may
// differ in its last bit from the old tiling's; the phi extrapolation history
// starts over, as after any upload.
static int retile_from(nyxpm_ctx* c, const int t[3], const ParticleSnapshot& s) {
    int rc;
    HIPCK(c, hipStreamSynchronize(c->stream));
    const int nbox = (int)c->boxes.size();
    std::vector<long long> np(nbox);
    std::vector<double> mc(nbox);
    for (int b = 0; b < nbox; ++b) {
        np[b] = c->boxes[b].np;
        mc[b] = c->boxes[b].m_const;
    }
    for (int d = 0; d < 3; ++d) c->tile[d] = t[d];
    tile_boxes(c);
    free_tile_buffers(c);
    if ((rc = alloc_tile_buffers(c))) return rc;
    if ((rc = build_tseg(c))) return rc;
    // every box's slots laid out for the new tile counts, each box copied once
    // from the snapshot (the mass bookkeeping of the boxes is unchanged)
    std::vector<long long> off(nbox), span(nbox);
    long long slots = 0;
    for (int b = 0; b < nbox; ++b) {
        off[b] = slots;
        span[b] = slot_count(c, np[b], c->boxes[b].ntiles);
        slots += span[b];
    }
    PartPtrs N = {};
    int* nkey = nullptr;
    const long long cap = std::max(slots, 1LL);
    if ((rc = alloc_particle_set(c, N, cap, c->any_mass))) return rc;
    if ((rc = dalloc(c, &nkey, cap))) return rc;
    for (int b = 0; b < nbox; ++b) {
        const double* const* a = s.a.data() + 7 * (size_t)b;
        const long long d = off[b];
        const size_t n8 = (size_t)np[b] * sizeof(double);
        if (np[b] > 0) {
            double* dst[6] = {N.x, N.y, N.z, N.vx, N.vy, N.vz};
            for (int q = 0; q < 6; ++q)
                HIPCK(c, hipMemcpyAsync(dst[q] + d, a[q], n8, hipMemcpyDeviceToDevice, c->stream));
            if (c->any_mass) {
                if (a[6]) {
                    HIPCK(c, hipMemcpyAsync(N.m + d, a[6], n8, hipMemcpyDeviceToDevice, c->stream));
                } else {
                    hipLaunchKernelGGL(k_fill_dbl, dim3(nblocks(np[b])), dim3(kBlock), 0, c->stream, N.m + d,
                                       (long long)np[b], mc[b]);
                }
            }
            hipLaunchKernelGGL(k_iota, dim3(nblocks(np[b])), dim3(kBlock), 0, c->stream, d, (long long)np[b],
                               N.id);
            LAUNCHCK(c);
        }
        if (span[b] > np[b]) {
            hipLaunchKernelGGL(k_fill_dbl, dim3(nblocks(span[b] - np[b])), dim3(kBlock), 0, c->stream,
                               N.x + d + np[b], span[b] - np[b], std::nan(""));
            LAUNCHCK(c);
        }
    }
    HIPCK(c, hipStreamSynchronize(c->stream));
    const long long icap = c->ntiles_total + slots / c->kick_chunk + 1;
    if (icap > 0x7fffffffLL) FAIL(c, NYXPM_ERR_ARG, "too many tile chunks");
    TileItem* nitm = nullptr;
    if ((rc = dalloc(c, &nitm, icap))) return rc;
    const long long hcap = std::min<long long>(c->ntiles_total, slots / kHeavyTile + 1);
    int* nheavy_list = nullptr;
    if ((rc = dalloc(c, &nheavy_list, hcap))) return rc;
    free_particles(c);
    c->P[0] = N;
    c->P[1] = PartPtrs{};
    c->key = nkey;
    c->items = nitm;
    c->items_cap = icap;
    c->heavy = nheavy_list;
    c->heavy_cap = (int)hcap;
    c->items_valid = false;
    c->dt_valid = false;
    c->cur = 0;
    c->part_cap = cap;
    for (int b = 0; b < nbox; ++b) {
        c->boxes[b].part_off = off[b];
        c->boxes[b].span = span[b];
    }
    c->moves_pending = false;
    invalidate_guess(c);
    if ((rc = upload_boxdev(c))) return rc;
    if ((rc = build_pseg(c))) return rc;
    c->keys_valid = c->sorted_valid = c->density_valid = false;
    if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
    return NYXPM_OK;
}

int nyxpm_autotune(nyxpm_ctx* c, int precision_mask, char* report, int report_len) {
    if (!c) return NYXPM_ERR_ARG;
    int rc;
    HIPCK(c, hipStreamSynchronize(c->stream));
    if ((precision_mask & 2) && (rc = ensure_mixed_pools(c))) return rc;
    const int nbox = (int)c->boxes.size();
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, c->stream, c->active, (long long)nbox, 1);
    static const int shapes[][2] = {{128, 8}, {64, 16}, {64, 8}, {32, 8}, {16, 8}};
    static const int zcs[] = {2, 4, 8, 16, 32, 64, 128, 256};
    hipEvent_t e0, e1;
    HIPCK(c, hipEventCreate(&e0));
    HIPCK(c, hipEventCreate(&e1));
    std::string rep = "{";
    bool first_entry = true;
    for (int pr = 0; pr < 2; ++pr) {
        if (!(precision_mask & (1 << pr))) continue;
        for (int d = 0; d + 1 < c->dmax; ++d) {
            auto& cfg = c->sw[pr][d];
            if (!cfg.tx || cfg.t.count == 0) continue;
            if (pr == 1 && two_sweep<float>(c, d)) continue;  // k_rb_sweep2 smooths this depth
            const int tx0 = cfg.tx, ty0 = cfg.ty, zc0 = cfg.zc;
            double best = 1e30;
            int btx = tx0, bty = ty0, bzc = zc0;
            int maxnz = 0;
            for (int q = 0; q < cfg.t.count; ++q)
                maxnz = std::max(maxnz, c->boxes[c->seg_host[cfg.t.first + q].box].lev[d].n[2]);
            for (auto& sh : shapes) {
                bool ok = true;
                for (int q = 0; q < cfg.t.count && ok; ++q) {
                    const Level& L = c->boxes[c->seg_host[cfg.t.first + q].box].lev[d];
                    if (L.n[0] % sh[0] || L.n[1] % sh[1]) ok = false;
                }
                if (!ok) continue;
                for (int zc : zcs) {
                    if (zc > maxnz) continue;
                    cfg.tx = sh[0];
                    cfg.ty = sh[1];
                    cfg.zc = zc;
                    if ((rc = rebuild_sweep_table(c, pr, d))) return rc;
                    float ms = 0.f;
                    for (int rep_i = 0; rep_i < 2; ++rep_i) {  // warm-up, then a timed batch of 4
                        HIPCK(c, hipEventRecord(e0, c->stream));
                        for (int k = 0; k < (rep_i ? 4 : 1); ++k) {
                            if (pr == 0)
                                launch_sweep<double, 0>(c, d, c->e64, c->e64b, c->r64, c->e64, c->e64b);
                            else
                                launch_sweep<float, 0>(c, d, c->e32, c->e32b, c->r32, c->e32, c->e32b);
                        }
                        HIPCK(c, hipEventRecord(e1, c->stream));
                        HIPCK(c, hipEventSynchronize(e1));
                        HIPCK(c, hipEventElapsedTime(&ms, e0, e1));
                    }
                    LAUNCHCK(c);
                    const double us = ms * 1e3 / 4.0;
                    if (us < best) {
                        best = us;
                        btx = sh[0];
                        bty = sh[1];
                        bzc = zc;
                    }
                }
            }
            cfg.tx = btx;
            cfg.ty = bty;
            cfg.zc = bzc;
            cfg.us = best;
            if ((rc = rebuild_sweep_table(c, pr, d))) return rc;
            char buf[160];
            snprintf(buf, sizeof(buf), "%s\"%s.d%d\": {\"shape\": [%d, %d], \"zchunk\": %d, \"us\": %.2f}",
                     first_entry ? "" : ", ", pr == 0 ? "fp64" : "fp32", d, btx, bty, bzc, best);
            rep += buf;
            first_entry = false;
        }
        // the z-chunk (coarse planes per block) of the marching residual-restriction
        for (int d = 0; d + 1 < c->dmax; ++d) {
            auto& tz = c->t_rrz[pr][d];
            if (tz.count == 0) continue;
            static const int rzcs[] = {2, 4, 8, 16, 32, 64, 128};
            int maxcz = 0;
            for (int q = 0; q < tz.count; ++q)
                maxcz = std::max(maxcz, c->boxes[c->seg_host[tz.first + q].box].lev[d].n[2] / 2);
            const int zc0 = c->seg_host[tz.first].zchunk;
            double best = 1e30;
            int bzc = zc0;
            for (int zc : rzcs) {
                if (zc != rzcs[0] && zc >= 2 * maxcz) continue;
                if ((rc = rebuild_rrz_table(c, pr, d, zc))) return rc;
                float ms = 0.f;
                for (int rep_i = 0; rep_i < 2; ++rep_i) {
                    HIPCK(c, hipEventRecord(e0, c->stream));
                    for (int k = 0; k < (rep_i ? 4 : 1); ++k) {
                        if (pr == 0)
                            hipLaunchKernelGGL((k_resid_restrict_z<double, kRrCX, kRrCY64>), dim3(tz.blocks),
                                               dim3(kRrCX * kRrCY64), 0, c->stream, c->segs(tz), tz.count, c->active,
                                               c->e64, c->r64);
                        else
                            hipLaunchKernelGGL((k_resid_restrict_z<float, kRrCX, kRrCY32>), dim3(tz.blocks),
                                               dim3(kRrCX * kRrCY32), 0, c->stream, c->segs(tz), tz.count, c->active,
                                               c->e32, c->r32);
                    }
                    HIPCK(c, hipEventRecord(e1, c->stream));
                    HIPCK(c, hipEventSynchronize(e1));
                    HIPCK(c, hipEventElapsedTime(&ms, e0, e1));
                }
                LAUNCHCK(c);
                const double us = ms * 1e3 / 4.0;
                if (c->verbose)
                    std::fprintf(stderr, "nyxpm autotune: resid_restrict %s.d%d z-chunk %d: %.2f us\n",
                                 pr == 0 ? "fp64" : "fp32", d, zc, us);
                if (us < best) {
                    best = us;
                    bzc = zc;
                }
            }
            if ((rc = rebuild_rrz_table(c, pr, d, bzc))) return rc;
            char buf[160];
            snprintf(buf, sizeof(buf), "%s\"rrz.%s.d%d\": {\"zchunk\": %d, \"us\": %.2f}", first_entry ? "" : ", ",
                     pr == 0 ? "fp64" : "fp32", d, bzc, best);
            rep += buf;
            first_entry = false;
        }
    }
    // the particle tile (deposit, fold, re-sort and paired kick of the current
    // precision; the trial kicks have dt = 0, and the particles come back from a
    // snapshot at the end)
    if ((precision_mask & 4) && c->npart_total > 0) {
        static const int tiles[][3] = {{16, 8, 8}, {12, 8, 8}, {24, 8, 8}, {16, 16, 8},
                                       {16, 8, 4}, {32, 8, 4}, {8, 8, 8}};
        const bool f32 = c->precision == NYXPM_MIXED;
        ParticleSnapshot snap;
        if ((rc = snapshot_take(c, snap))) {
            snapshot_free(snap);
            return rc;
        }
        int best_t[3] = {c->tile[0], c->tile[1], c->tile[2]};
        double best = 1e30;
        for (auto& t : tiles) {
            if ((t[0] + 1) * (t[1] + 1) * (t[2] + 1) * 16 > 64 * 1024) continue;
            if ((rc = retile_from(c, t, snap))) break;
            float ms = 0.f;
            for (int rep_i = 0; rep_i < 2 && !rc; ++rep_i) {  // warm-up, then a timed pair of steps
                HIPCK(c, hipEventRecord(e0, c->stream));
                for (int k = 0; k < (rep_i ? 2 : 1) && !rc; ++k) {
                    if ((rc = phase_deposit(c, false, f32))) break;
                    rc = phase_kick<kKickPair>(c, 0.0, 1.0, 1.0, 1.0, 1.0);
                }
                HIPCK(c, hipEventRecord(e1, c->stream));
                HIPCK(c, hipEventSynchronize(e1));
                HIPCK(c, hipEventElapsedTime(&ms, e0, e1));
            }
            if (rc) break;
            const double us = ms * 1e3 / 2.0;
            if (c->verbose)
                std::fprintf(stderr, "nyxpm autotune: particle tile %dx%dx%d: %.2f us\n", t[0], t[1], t[2], us);
            if (us < best) {
                best = us;
                for (int d = 0; d < 3; ++d) best_t[d] = t[d];
            }
        }
        if (!rc) rc = retile_from(c, best_t, snap);
        snapshot_free(snap);
        if (rc) return rc;
        // the trial deposits overwrote the density and rhs buffers
        c->f1_valid = c->f_raw = c->density_valid = false;
        char buf[160];
        snprintf(buf, sizeof(buf), "%s\"tile\": {\"shape\": [%d, %d, %d], \"us\": %.2f}",
                 first_entry ? "" : ", ", best_t[0], best_t[1], best_t[2], best);
        rep += buf;
        first_entry = false;
    }
    // the two-sweep pass: column shape and z-chunk per depth (FP32)
    static const int sw2_shapes[][2] = {{32, 16}, {64, 16}, {32, 32}};
    static const int sw2_zcs[] = {32, 64, 128, 256, 512};
    for (int d = 0; (precision_mask & 2) && d + 1 < c->dmax; ++d) {
        auto& s2 = c->sw2[d];
        if (s2.t.count == 0) continue;
        const int tx0 = s2.tx, ty0 = s2.ty;
        int maxnz = 0;
        for (int q = 0; q < s2.t.count; ++q)
            maxnz = std::max(maxnz, c->boxes[c->seg_host[s2.t.first + q].box].lev[d].n[2]);
        double best = 1e30;
        int btx = tx0, bty = ty0, bzc = s2.zc;
        for (auto& sh : sw2_shapes) {
            bool ok = true;
            for (int q = 0; q < s2.t.count && ok; ++q)
                ok = sw2_fits(c->boxes[c->seg_host[s2.t.first + q].box].lev[d], sh[0], sh[1]);
            if (!ok) continue;
            for (int zc : sw2_zcs) {
                if (zc != sw2_zcs[0] && zc >= 2 * maxnz) continue;
                s2.tx = sh[0];
                s2.ty = sh[1];
                s2.zc = zc;
                if ((rc = rebuild_sw2_table(c, d))) return rc;
                float ms = 0.f;
                for (int rep_i = 0; rep_i < 2; ++rep_i) {  // warm-up, then a timed batch of 4
                    HIPCK(c, hipEventRecord(e0, c->stream));
                    for (int k = 0; k < (rep_i ? 4 : 1); ++k)
                        launch_sweep2<0>(c, d, c->e32, c->e32b, c->r32, c->e32, c->e32b);
                    HIPCK(c, hipEventRecord(e1, c->stream));
                    HIPCK(c, hipEventSynchronize(e1));
                    HIPCK(c, hipEventElapsedTime(&ms, e0, e1));
                }
                LAUNCHCK(c);
                const double us = ms * 1e3 / 4.0;
                if (c->verbose)
                    std::fprintf(stderr, "nyxpm autotune: sweep2.d%d %dx%d z-chunk %d: %.2f us\n", d, sh[0], sh[1],
                                 zc, us);
                if (us < best) {
                    best = us;
                    btx = sh[0];
                    bty = sh[1];
                    bzc = zc;
                }
            }
        }
        s2.tx = btx;
        s2.ty = bty;
        s2.zc = bzc;
        s2.us = best;
        // where the default left the pass off (too few blocks), it replaces the
        // two k_rb_sweep passes on a level of at least 2^22 cells if it is faster
        // than two of them (on smaller levels the launches take a few
        // microseconds, and a timing is mostly launch overhead, which a tracer
        // inflates)
        const double us1 = c->sw[1][d].us;
        long long cells = 0;
        for (int q = 0; q < s2.t.count; ++q) cells += c->boxes[c->seg_host[s2.t.first + q].box].lev[d].nc;
        if (!s2.on && us1 > 0.0 && best < 2.0 * us1 && cells >= (1LL << 22)) s2.on = true;
        if ((rc = rebuild_sw2_table(c, d))) return rc;
        char buf[200];
        snprintf(buf, sizeof(buf),
                 "%s\"sweep2.d%d\": {\"shape\": [%d, %d], \"zchunk\": %d, \"us\": %.2f, \"on\": %s}",
                 first_entry ? "" : ", ", d, btx, bty, bzc, best, s2.on ? "true" : "false");
        rep += buf;
        first_entry = false;
    }
    rep += "}";
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
    if (report && report_len > 0) {
        std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = 0;
    }
    return NYXPM_OK;
}

int nyxpm_timing_enable(nyxpm_ctx* c, int on) {
    if (!c) return NYXPM_ERR_ARG;
    c->timing = on != 0;
    return NYXPM_OK;
}

int nyxpm_timing_select(nyxpm_ctx* c, const char* names) {
    if (!c) return NYXPM_ERR_ARG;
    c->timing_only.clear();
    if (!names) return NYXPM_OK;
    std::string s(names);
    size_t p = 0;
    while (p <= s.size()) {
        const size_t q = std::min(s.find(',', p), s.size());
        if (q > p) c->timing_only.insert(s.substr(p, q - p));
        p = q + 1;
    }
    return NYXPM_OK;
}

int nyxpm_timing_reset(nyxpm_ctx* c) {
    if (!c) return NYXPM_ERR_ARG;
    int rc = collect_timing(c);
    c->timed.clear();
    return rc;
}

int nyxpm_timing_get(nyxpm_ctx* c, const char* name, int64_t* count, double* mean_ms, double* total_ms) {
    if (!c || !name) return NYXPM_ERR_ARG;
    int rc = collect_timing(c);
    if (rc) return rc;
    auto it = c->timed.find(name);
    long long n = it == c->timed.end() ? 0 : it->second.first;
    double t = it == c->timed.end() ? 0.0 : it->second.second;
    if (count) *count = n;
    if (total_ms) *total_ms = t;
    if (mean_ms) *mean_ms = n ? t / n : 0.0;
    return NYXPM_OK;
}

}  // extern "C"
