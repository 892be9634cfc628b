/*
 * nyxpm.h -- C ABI of the MI355X particle-mesh gravity library.
 *
 * The library replaces, on one periodic level, the hot path Nyx drives through
 * its particle-container and Gravity interfaces.  Each entry point names the
 * reference interface it stands in for (paths relative to the Nyx source tree):
 *
 *   nyxpm_assign_density      NyxParticleContainerBase::AssignDensitySingleLevel /
 *                             AssignDensity        (Source/Particle/NyxParticleContainer.H:43-46, 90-97)
 *   nyxpm_sum_particle_mass   NyxParticleContainerBase::sumParticleMass
 *                                                  (NyxParticleContainer.H:38, 70-79)
 *   nyxpm_solve_for_phi       Gravity::solve_for_phi with Gravity::CorrectRhsUsingOffset and
 *                             Gravity::solve_with_MLMG (Source/Gravity/Gravity.H:54-56;
 *                             Gravity.cpp:267-315, 1207-1244, 1318-1416)
 *   nyxpm_solve_poisson       Gravity::solve_with_MLMG alone (Gravity.H:43-48)
 *   nyxpm_get_grav_vector     Gravity::get_new_grav_vector / get_old_grav_vector
 *                                                  (Gravity.H:84-85; Gravity.cpp:781-850)
 *   nyxpm_gather_accel        the acceleration update_dm_particle_single interpolates
 *                                                  (DarkMatterParticleContainer.cpp:276-292)
 *   nyxpm_move_kick_drift     DarkMatterParticleContainer::moveKickDrift
 *                                                  (DarkMatterParticleContainer.H:31-32; .cpp:75-181)
 *   nyxpm_move_kick           DarkMatterParticleContainer::moveKick
 *                                                  (DarkMatterParticleContainer.H:33-34; .cpp:183-247)
 *   nyxpm_redistribute        NyxParticleContainerBase::RedistributeLocal (periodic shift only)
 *                                                  (NyxParticleContainer.H:123-136)
 *   nyxpm_pm_init             the phi solve of Nyx::post_init (Source/Driver/Nyx.cpp:1958-2006)
 *   nyxpm_pm_steps            the particle-only advance of Nyx::advance_particles_only
 *   nyxpm_pm_steps_comoving   the same with a(t) from Nyx::integrate_comoving_a
 *                             (Source/Driver/comoving.cpp:630-689)
 *                             followed by post_timestep's redistribute
 *                                                  (Source/TimeStep/Nyx_advance_particles.cpp:9-272;
 *                                                   Source/Driver/Nyx.cpp:1421-1436)
 *
 * Conventions
 *  - Every function returns 0 on success or a NYXPM_ERR_* code; the library never
 *    aborts the process.  nyxpm_last_error() gives a message for the last failure.
 *  - Host pointers passed in are borrowed for the duration of the call.  The
 *    library owns every device buffer; particle and field data stay resident in
 *    HBM between calls.  Entry points with the _device suffix take device
 *    pointers instead (stream-ordered on the context's stream).
 *  - One context per host thread; a context is not internally synchronised.
 *    All work is ordered on the context's stream; nyxpm_sync() waits for it.
 *  - Cell fields are x-fastest, no ghost cells: index (k*ny + j)*nx + i.  Vector
 *    fields are component-major (3 consecutive cell fields).
 *  - A context may hold several independent periodic boxes ("batch"); box-indexed
 *    entry points take the box number, the others act on every box.
 */
#ifndef NYXPM_H
#define NYXPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NYXPM_OK 0
#define NYXPM_ERR_ARG 1          /* invalid argument */
#define NYXPM_ERR_HIP 2          /* a HIP runtime call failed */
#define NYXPM_ERR_NOMEM 3        /* device allocation failed */
#define NYXPM_ERR_NOT_CONVERGED 4/* multigrid did not reach its tolerance within max_iter */
#define NYXPM_ERR_UNSUPPORTED 5  /* geometry the solver does not handle (see message) */
#define NYXPM_ERR_STATE 6        /* call out of order (e.g. no particles uploaded) */

#define NYXPM_FP64 0   /* FP64 everywhere: the drop-in path */
#define NYXPM_MIXED 1  /* FP32 V-cycles inside FP64 iterative refinement */

typedef struct nyxpm_ctx nyxpm_ctx;

/* One periodic box: Geometry's n_cell, prob_lo, prob_hi (all directions periodic). */
typedef struct nyxpm_geom {
    int n_cell[3];
    double prob_lo[3];
    double prob_hi[3];
} nyxpm_geom;

typedef struct nyxpm_opts {
    int precision;     /* NYXPM_FP64 or NYXPM_MIXED */
    int nu1, nu2;      /* pre/post smoothing sweeps (MLMG defaults 2, 2); 0 = default */
    int max_iter;      /* V-cycle limit (MLMG default 200); 0 = default */
    int tile[3];       /* deposit/gather tile shape in cells; 0 = default (16,8,8) */
    int verbose;       /* 0 = silent */
    int smoother;      /* 0: fused red-black sweeps where the level allows (default);
                          1: separate red and black half sweeps everywhere */
    int kick;          /* 0: tile-staged gather + kick (default); 1: one thread per
                          particle reading phi from global memory */
    int sweep_tx, sweep_ty; /* fused-sweep column shape for every level that allows
                          it (128x8, 64x16, 64x8, 32x8, 16x8); 0 = chosen per level */
    int sweep_zchunk;  /* z-planes per fused-sweep block; 0 = chosen per level */
    int kick_chunk;    /* particles per work item of the tile-staged kick (>= 256);
                          0 = default (4096) */
    int separate_kicks; /* 0 = nyxpm_pm_steps runs the closing kick of a step and
                           the next step's kick-drift as one particle pass
                           (bit-identical); 1 = separate passes */
    int coarse;        /* 0: the levels of at most 4096 cells (down to the bottom
                          solve) run as one workgroup per box (default);
                          1: one launch per level and step (same results) */
    double particle_cfl; /* > 0: every closing kick (nyxpm_move_kick, the end of each
                          nyxpm_pm_steps step) also computes estTimestep of the
                          kicked particles with this cfl and a = a_new, which
                          nyxpm_est_time_step then returns without another pass;
                          0 = off */
    double resort_slack; /* incremental re-sort: each tile of the full sort gets
                          cnt * resort_slack + 16 free slots, a kick-drift lists
                          only the particles that change tile and they move into
                          that slack; a full sort runs when a tile's slack is
                          used up or the emptied slots pass np/16.  0 = default
                          (0.25), < 0 = off (full sort after every kick-drift),
                          at most 4.  Same results either way. */
    int phi_guess;     /* initial guess of the solve inside each nyxpm_pm_steps step:
                          0 = default = 1 = phi_n, as Nyx; 2 = 2 phi_n - phi_{n-1}
                          (one extra pass over phi; saves a V-cycle where the
                          previous guess sits just above a cycle boundary).  The
                          stopping rule is the same.  Not used with nyx_faithful. */
    int residual;      /* 0: the level-0 residual is a pass of its own (default);
                          1: it is formed inside the first sweep of the next
                          V-cycle where the fused sweep runs (same results; one
                          pass fewer but half the occupancy, slower at 512^3) */
    int graph_cycles;  /* V-cycles recorded per solve in a hipGraph-captured step
                          (nyxpm_pm_steps with use_graph); 0 = default: the most
                          V-cycles a solve of the same call needed eagerly plus 2,
                          re-captured when a replay uses up that margin */
    int ir_first;      /* mixed precision: FP32 V-cycles of each solve's first
                          refinement step, run back to back on one FP64 residual
                          (each continuing from the correction of the one before,
                          phi += e once) before the first FP64 residual and
                          stopping test; later steps run one V-cycle each.  Saves
                          the FP64 residual and update of the cycles in between;
                          a solve that would stop earlier runs up to ir_first
                          cycles.  0 = default: 5, or 4 with phi_guess = 2 (a
                          smaller first residual: the FP32 correction reaches
                          its limit, ~1e-7 of the residual it starts from, a
                          cycle earlier); 1 = a test after every cycle.
                          Needs nu1 + nu2 even (else 1); FP64 solves and fixed
                          iteration counts always test after every cycle. */
    int sweep2;        /* FP32 levels with V(2,2) cycles: each pair of smoothing
                          sweeps as one pass over the level (e and r read once, e
                          written once; same results bit for bit).  0 = default
                          (on, 128 z-planes per block), > 1 = on with that many
                          z-planes per block, < 0 = off. */
    int sweep2_tx, sweep2_ty; /* column of the one-pass smoothing pair: 32 x 16,
                          64 x 16 or 32 x 32 cells (0 = default 32 x 16; the
                          autotuner may pick another per level) */
} nyxpm_opts;

/* ---- lifecycle ---------------------------------------------------------- */
int nyxpm_create(nyxpm_ctx **ctx, const nyxpm_geom *geom, const nyxpm_opts *opts);
int nyxpm_create_batch(nyxpm_ctx **ctx, int nbox, const nyxpm_geom *geoms, const nyxpm_opts *opts);
int nyxpm_destroy(nyxpm_ctx *ctx);
int nyxpm_set_stream(nyxpm_ctx *ctx, void *hip_stream);
int nyxpm_sync(nyxpm_ctx *ctx);
const char *nyxpm_last_error(const nyxpm_ctx *ctx);
int nyxpm_num_boxes(const nyxpm_ctx *ctx);
int nyxpm_set_precision(nyxpm_ctx *ctx, int precision);

/* ---- particles ---------------------------------------------------------- */
/* Replace the particles of box `box` (host arrays of length np).  m may be NULL:
 * then every particle has mass m_const.  Particles keep their upload index as id. */
int nyxpm_particles_upload(nyxpm_ctx *ctx, int box, int64_t np,
                           const double *x, const double *y, const double *z,
                           const double *vx, const double *vy, const double *vz,
                           const double *m, double m_const);
/* Same, device pointers. */
int nyxpm_particles_upload_device(nyxpm_ctx *ctx, int box, int64_t np,
                                  const double *x, const double *y, const double *z,
                                  const double *vx, const double *vy, const double *vz,
                                  const double *m, double m_const);
/* DMPC->InitFromBinaryFile(file, AMREX_SPACEDIM + 1) (NyxParticles.cpp:525-545)
 * for box `box`: reads Nyx's binary particle file (int64 NP, int32 DM = 3, int32
 * NX >= 4, then NP records x y z m vx vy vz [more extras, skipped], FP64 or FP32
 * by file size) and uploads it as nyxpm_particles_upload does (per-particle
 * mass).  np_out (may be NULL) receives NP. */
int nyxpm_particles_load_binary(nyxpm_ctx *ctx, int box, const char *path, int64_t *np_out);
/* DMPC->InitFromAsciiFile(file, AMREX_SPACEDIM + 1) (NyxParticles.cpp:506-521)
 * for box `box`: Nyx's ASCII particle file (the count, then x y z m vx vy vz per
 * particle, whitespace separated), uploaded with per-particle masses; particles
 * outside the box are shifted by whole periods into it.  np_out may be NULL. */
int nyxpm_particles_load_ascii(nyxpm_ctx *ctx, int box, const char *path, int64_t *np_out);
/* NyxParticleContainer::NyxCheckpoint (is_checkpoint != 0) / WriteNyxPlotFile
 * (NyxParticleContainer.H:497-517, via Nyx_output.cpp:652, 735): box `box`'s
 * particles as AMReX's particle checkpoint or plotfile <dir>/<name>/{Header,
 * Level_0/Particle_H, Level_0/DATA_00000} (one level, one grid, FP64 reals x y z
 * mass xvel yvel zvel, checkpoints with int32 id and cpu first; plotfiles also
 * get Nyx's <dir>/<name>/precision file), in upload order with ids 1..np. */
int nyxpm_particles_write_amrex(nyxpm_ctx *ctx, int box, const char *dir, const char *name, int is_checkpoint);
/* DMPC->Restart(dir, name) (NyxParticles.cpp:586, 932) for box `box`: reads an
 * AMReX particle checkpoint or plotfile of one level (any grids and DATA files,
 * FP64 or FP32 reals; the first four extra reals are mass and velocity) and
 * uploads it; particles outside the box are shifted into it.  np_out may be NULL. */
int nyxpm_particles_read_amrex(nyxpm_ctx *ctx, int box, const char *dir, const char *name, int64_t *np_out);
/* Copy box `box`'s particles to host or device arrays (unified addressing), in
 * upload (id) order.  Any pointer may be NULL to skip that component. */
int nyxpm_particles_download(nyxpm_ctx *ctx, int box,
                             double *x, double *y, double *z,
                             double *vx, double *vy, double *vz);
int64_t nyxpm_num_particles(const nyxpm_ctx *ctx, int box);

/* ---- AssignDensity / sumParticleMass ------------------------------------ */
/* Deposit every box's particles with CIC weights (deterministic), periodic,
 * times 1/cell volume.  rho_out (n cells of box `box`, host or device memory) may
 * be NULL to keep the result in the context only; box = -1 with rho_out NULL
 * deposits all boxes. */
int nyxpm_assign_density(nyxpm_ctx *ctx, int box, double *rho_out);
int nyxpm_sum_particle_mass(nyxpm_ctx *ctx, int box, double *mass_out);
/* DarkMatterParticleContainer::AssignDensityAndVels (DarkMatterParticleContainer.cpp:
 * 456-460, AssignDensity with AMREX_SPACEDIM + 1 components; used by
 * ParticleDerive.cpp:126-168 and the Santa Barbara initialisation,
 * NyxParticles.cpp:771-773) for box `box`: out (4 fields of n cells, host or
 * device memory) receives the CIC mass density (= nyxpm_assign_density) and the
 * CIC mass-weighted cell velocity sum(w m v_d) / sum(w m) for d = x, y, z, 0 in
 * cells without mass.  Deterministic (exact fixed-point sums). */
int nyxpm_assign_density_and_vels(nyxpm_ctx *ctx, int box, double *out);

/* ---- Gravity ------------------------------------------------------------ */
/* solve_for_phi on every box: Rhs = density of the last assign_density (or
 * nyxpm_set_density); Rhs -= mass_offset; Rhs -= vol_weight_sum(Rhs);
 * Rhs *= -four_pi_G; Rhs *= 1/a; then V-cycles from phi (use_guess != 0) or from
 * zero until |Rhs - L phi|_inf <= max(abs_tol, rel_tol*max(|Rhs|_inf, |r0|_inf)).
 * iters / final_resnorm (length nbox each, may be NULL) receive per-box results. */
int nyxpm_solve_for_phi(nyxpm_ctx *ctx, double a, double four_pi_G, double rel_tol,
                        double abs_tol, int use_guess, int *iters, double *final_resnorm);
/* solve_with_MLMG on box `box`: L phi = rhs for a given right-hand side (host
 * array, mean-free).  phi_io (host) holds the guess on entry, the solution on exit. */
int nyxpm_solve_poisson(nyxpm_ctx *ctx, int box, const double *rhs, double *phi_io,
                        double rel_tol, double abs_tol, int fixed_iters, int *iters,
                        double *norms /* 4: resnorm0, rhsnorm, final, target; may be NULL */);
/* Accelerations interpolated to box `box`'s particles from the current phi, as
 * moveKick/moveKickDrift use them (update_dm_particle_single's val,
 * DarkMatterParticleContainer.cpp:276-292), without changing the particles:
 * g_out holds 3*np doubles, component-major, in upload (id) order. */
int nyxpm_gather_accel(nyxpm_ctx *ctx, int box, double *g_out);
/* nyxpm_gather_accel that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same results as
 * nyxpm_gather_accel (same kernel).  A host g_out is staged through a fixed
 * device array of 3 * 2^20 doubles (24 MiB per GPU, held until the process ends)
 * instead of a buffer allocated per call, so the box may hold at most 2^20
 * particles (NYXPM_ERR_ARG above that); a device g_out is written directly.
 * Every context on a GPU shares that GPU's array: calls that overlap on one GPU
 * (contexts on different host threads or streams) race. */
int nyxpm_gather_accel_noalloc(nyxpm_ctx *ctx, int box, double *g_out);
/* NyxParticleContainer::estTimestep(grav, a, 0, cfl)
 * (NyxParticleContainer.H:327-410, called by Nyx::particle_est_time_step,
 * NyxParticles.cpp:1028-1036) with grav = the cell-centred gravity of the current
 * phi: per particle cfl / max_d(|v_d| dxi_d / a), capped by 1/sqrt(|g| dxi_x) with g
 * at the particle's cell floor((x - plo) dxi); minimum over the box's particles,
 * -1e50 for a box without particles.  dt_out holds nbox values. */
int nyxpm_est_time_step(nyxpm_ctx *ctx, double a, double cfl, double *dt_out);
/* Cell-centred gravity g = average of face gradients of phi (3 fields). */
int nyxpm_get_grav_vector(nyxpm_ctx *ctx, int box, double *g_out);
/* Field access for box `box` (arrays of n cells).  The array arguments of
 * assign_density, gather_accel, get_grav_vector and these may be host or device
 * memory (HIP unified addressing): device memory is written or read directly. */
int nyxpm_get_phi(nyxpm_ctx *ctx, int box, double *phi_out);
int nyxpm_set_phi(nyxpm_ctx *ctx, int box, const double *phi_in);
int nyxpm_set_density(nyxpm_ctx *ctx, int box, const double *rho_in);
int nyxpm_get_rhs(nyxpm_ctx *ctx, int box, double *rhs_out);

/* ---- particle update ---------------------------------------------------- */
/* moveKickDrift: v = (a_old v + dt/2 g(x)) / a_half; x += dt/a_half v, g from the
 * current phi (gradient fused into the interpolation). */
int nyxpm_move_kick_drift(nyxpm_ctx *ctx, double dt, double a_old, double a_half);
/* moveKick: v = (a_half v + dt/2 g(x)) / a_new, then the periodic position shift
 * of Redistribute. */
int nyxpm_move_kick(nyxpm_ctx *ctx, double dt, double a_new, double a_half);
/* Periodic position shift alone (RedistributeLocal on one level). */
int nyxpm_redistribute(nyxpm_ctx *ctx);

/* ---- whole step --------------------------------------------------------- */
/* Before the first step: phi for the initial positions (from a zero guess). */
int nyxpm_pm_init(nyxpm_ctx *ctx, double a, double four_pi_G, double rel_tol, double abs_tol);
/* nsteps KDK steps with fixed dt: KD with the stored phi, deposit, solve (previous
 * phi as guess), K, wrap.  nyx_faithful != 0 re-solves the old phi from a zero
 * guess at the start of each step, as Nyx_advance_particles.cpp:148-170 does.
 * use_graph != 0 runs the first step eagerly, then captures two steps into a
 * hipGraph once and replays them (the V-cycle count is decided on the device;
 * opts.graph_cycles cycles are recorded per solve, by default the most the
 * eager solves needed plus 2).  A solve of a replay that is still above its
 * target after the recorded cycles returns NYXPM_ERR_NOT_CONVERGED.
 * iters (length nbox, may be NULL) receives the V-cycles of the last step. */
int nyxpm_pm_steps(nyxpm_ctx *ctx, int nsteps, double dt, double a_old, double a_new,
                   double four_pi_G, double rel_tol, double abs_tol, int nyx_faithful,
                   int use_graph, int *iters);
/* Nyx::integrate_comoving_a (comoving.cpp:630-689): a after dt from old_a, RK2 on
 * da/dt = 100 h sqrt(omega_m/a + omega_r/a^2 + (1-omega_m-omega_r) a^2) with step
 * doubling to 1e-8 (comoving_type > 0; otherwise da/dt = h).  Host arithmetic only. */
int nyxpm_integrate_comoving_a(double old_a, double dt, double omega_m, double omega_r,
                               double h, int comoving_type, double *new_a);
/* nsteps KDK steps in comoving coordinates: the scale factor is integrated by
 * nyxpm_integrate_comoving_a each step (a_old = *a_io at the start, *a_io = a at
 * the end); step s runs from a_s to a_(s+1), a_half = their mean
 * (Nyx_advance_particles.cpp:180,245), the Rhs is scaled by 1/a_(s+1). */
int nyxpm_pm_steps_comoving(nyxpm_ctx *ctx, int nsteps, double dt, double *a_io,
                            double omega_m, double omega_r, double h, int comoving_type,
                            double four_pi_G, double rel_tol, double abs_tol, int nyx_faithful,
                            int *iters);

/* ---- tuning ------------------------------------------------------------- */
/* Autotune launch shapes on this context's own arrays and keep the fastest.
 * precision_mask bit 0 (FP64) and bit 1 (FP32): for every multigrid depth, the
 * fused smoother's column shape the level allows (128x8, 64x16, 64x8, 32x8,
 * 16x8) and z-chunk (2..256 planes); with bit 1 also the one-pass smoothing
 * pair's column (32x16, 64x16, 32x32) and z-chunk (32..512), switched on at a
 * level of 2^22 cells or more where it beats two single sweeps.  Bit 2, with
 * particles uploaded: the particle tile of the deposit and the kicks (16x8x8,
 * 12x8x8, 24x8x8, 16x16x8, 16x8x4, 32x8x4, 8x8x8), timed on a deposit, fold,
 * re-sort and paired kick of dt = 0 in the context's precision; the particles
 * are restored from a copy taken first, so the state is unchanged (the phi
 * extrapolation history starts over).  The smoother choices never change a
 * result (every shape gives the same bits); the particle tile changes the
 * density in its last bit at tile boundaries (the fold adds a cell's parts
 * from the tiles that hold them in FP64), as the tile option does.  report
 * (may be NULL) receives a JSON
 * object {"fp64.d0": {"shape": [tx, ty], "zchunk": z, "us": t}, ...,
 * "tile": {"shape": [tx, ty, tz], "us": t}, "sweep2.d0": {...}}.  No
 * counterpart in the reference (AMReX's MLMG and particle tiling have fixed
 * launch shapes). */
int nyxpm_autotune(nyxpm_ctx *ctx, int precision_mask, char *report, int report_len);

/* ---- measurement -------------------------------------------------------- */
/* Kernel-class timing with HIP events on the context stream.  When enabled, each
 * launch of a timed class is bracketed by events; nyxpm_timing_get returns the
 * number of launches and the mean duration in milliseconds for `name`:
 * particle stages "keys", "scan", "scatter" (full sort), "migrate" (incremental
 * re-sort), "items", "deposit", "fold", "rhs", "kick_drift", "kick",
 * "kick_pair"; multigrid stages "residual", "update", "coarse", "bottom" and,
 * per depth d, "sweep0.dD", "sweep.dD", "sweep_p.dD", "sweep_u.dD",
 * "smooth.dD", "resid_restrict.dD", "prolong.dD", "fill.dD". */
int nyxpm_timing_enable(nyxpm_ctx *ctx, int on);
/* Restricts the timing to the stages named in `names` (comma-separated, e.g.
 * "kick_pair"); NULL or "" times every stage.  Each timed launch adds two event
 * markers to the stream, so a measurement of whole steps times only the stages
 * it reports.  No counterpart in the reference (measurement only). */
int nyxpm_timing_select(nyxpm_ctx *ctx, const char *names);
int nyxpm_timing_get(nyxpm_ctx *ctx, const char *name, int64_t *count, double *mean_ms,
                     double *total_ms);
int nyxpm_timing_reset(nyxpm_ctx *ctx);
/* Bytes of device memory the context holds. */
int64_t nyxpm_device_bytes(const nyxpm_ctx *ctx);
/* Particle re-sort counters since the context was created: out[0] full counting
 * sorts, out[1] incremental re-sorts that kept the slot layout, out[2] particles
 * those moved to another tile, out[3] emptied slots in the current layout.
 * No counterpart in the reference (Redistribute re-bins every particle). */
int nyxpm_sort_stats(const nyxpm_ctx *ctx, int64_t *out4);

#ifdef __cplusplus
}
#endif
#endif /* NYXPM_H */
