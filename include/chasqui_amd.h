/* chasqui_amd — C ABI of the MI355X-native per-RK-stage MHD pipeline.
 *
 * Plain C types only (no torch, no Kokkos, no C++ types cross this boundary). Every call
 * returns 0 on success or a negative CHQ_E* code; chq_last_error() gives the message of the
 * last failure on the calling thread. Nothing throws across the ABI.
 *
 * A context (chq_ctx) owns one mesh: one HIP stream and one device arena holding every
 * field of the reference's Mesh (mesh.hpp:13-63), in a padded device layout. Host buffers
 * passed to chq_upload/chq_download use the reference layout (LayoutRight, [var][k][j][i],
 * cell arrays nk*nj*ni, face/edge arrays (nk+1)*(nj+1)*(ni+1)), so they compare 1:1 with the
 * reference's arrays. One context = one stream; a context is not thread-safe, distinct
 * contexts may be used concurrently. All kernel entry points are asynchronous on the
 * context's stream unless they return a host value.
 *
 * Buffer sets: `cur`/`out`/`buf` = 0 selects u0 and b?f0, 1 selects u1 and b?f1 (the
 * reference's ping-pong pair, main.cpp:48-50, 69-74).
 */
#ifndef CHASQUI_AMD_H
#define CHASQUI_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHQ_OK 0
#define CHQ_EINVAL -1   /* bad argument or configuration */
#define CHQ_EHIP -2     /* HIP runtime error */
#define CHQ_ENOMEM -3   /* device allocation failed */
#define CHQ_EUNSUP -4   /* option outside what this build implements */

typedef struct chq_ctx chq_ctx;

/* Mesh description: Mesh::nx1..3, ng, x?min/x?max (mesh.hpp:14-20; main.cpp:165-175). */
typedef struct {
  int nx1, nx2, nx3, ng;
  double x1min, x1max, x2min, x2max, x3min, x3max;
} chq_mesh_desc;

/* Physics configuration: Config (physics.hpp:22-33) + the driver switches of main.cpp:112-163.
 * recon: 0 dc, 1 plm, 2 ppm, 3 wenoz, 4 ppmx (defs.hpp:35). rsolver: 0 hlle, 1 hlld,
 * 2 lhlld (physics.hpp:20). nstage: 1..4 = rk1, rk2, rk3, rk4/ssprk43 (defs.hpp:36).
 * bc[d]: 0 periodic, 1 outflow, 2 reflect (defs.hpp:37).
 * precision: 0 = FP64 throughout (the reference's arithmetic, bitwise); 1 = mixed: primitives,
 * reconstruction, Riemann solve, fluxes and EMFs in FP32, con2prim, RK combine, CT and all sums
 * in FP64 (conserved state and face B stay FP64). Mixed needs dual_energy = 0. No counterpart
 * in the reference, which is FP64 only (defs.hpp:11 Real = double).
 * characteristic: reconstruct in the characteristic variables of each side's cell
 * (Config::characteristic, physics.hpp:30, 71, 110-150); ignored with dc; FP64 only, and not
 * in batches or served runs.
 * nu, kappa, eta: viscosity, thermal conduction and Ohmic resistivity (diffusion::Coeffs,
 * diffusion.hpp:18-21; main.cpp:144-146), added in every stage, or, with sts = 1, advanced by
 * RKL2 super-time-stepping after each step (sts.hpp:111-157; main.cpp:142, 278-290). FP64
 * single-mesh path only.
 * flux_limiter: the Hu-Adams-Shu positivity flux limiter (positivity.hpp:24-89, main.cpp:60-64).
 * grav_*, g1..g3, gm, gx0..gz0, rsoft: uniform and point-mass gravity (srcterms::Gravity,
 * srcterms.hpp:11-40, main.cpp:148-157); cooling, lambda0, cool_alpha, tfloor: exact power-law
 * cooling after each step (srcterms::Cooling, 45-71, main.cpp:158-162, 282-286). FP64 single-mesh
 * path only. */
typedef struct {
  int recon, rsolver, nstage;
  double gamma, dfloor, pfloor, cfl;
  int flatten, low_mach, fofc, dual_energy;
  double de_eta, lhlld_phi;
  int bc[3];
  int precision;
  int characteristic;
  double nu, kappa, eta;
  int sts;
  int flux_limiter;
  int grav_uniform, grav_point;
  double g1, g2, g3, gm, gx0, gy0, gz0, rsoft;
  int cooling;
  double lambda0, cool_alpha, tfloor;
} chq_config;

/* Array identifiers for chq_upload / chq_download / chq_array_count. */
enum chq_array {
  CHQ_U0 = 0, CHQ_U1, CHQ_W, CHQ_BCC, CHQ_ECC,
  CHQ_B1F0, CHQ_B2F0, CHQ_B3F0, CHQ_B1F1, CHQ_B2F1, CHQ_B3F1,
  CHQ_F1, CHQ_F2, CHQ_F3,
  CHQ_E2X1, CHQ_E3X1, CHQ_E1X2, CHQ_E3X2, CHQ_E1X3, CHQ_E2X3,
  CHQ_E1, CHQ_E2, CHQ_E3,
  CHQ_FLAG, /* int32 */
  CHQ_NARRAYS
};

/* Time-loop state kept on the device (read with chq_get_status). */
typedef struct {
  double t, dt, tlim, min_dt;
  int cycle, nlim, active, nbad_last;
  long long fofc_total;
  long long sts_stages;   /* RKL2 stages taken (super-time-stepping, main.cpp:279, 308) */
  long long limiter_total;   /* positivity flux-limiter activations (main.cpp:307) */
} chq_status;

const char *chq_last_error(void);
const char *chq_version(void);

/* Mesh::Allocate (mesh.hpp:37-63) + Config; selects the current device. */
int chq_create(const chq_mesh_desc *mesh, const chq_config *cfg, chq_ctx **out);
int chq_destroy(chq_ctx *ctx);
/* geometry: iv[16] = nx1 nx2 nx3 ng ni nj nk is ie js je ks ke two_d three_d pitch */
int chq_geometry(const chq_ctx *ctx, int *iv, double *dx3);
/* bytes of device memory held by the context */
size_t chq_device_bytes(const chq_ctx *ctx);
/* device bytes chq_create would allocate for this mesh and configuration (0 if invalid); the
 * serving loop admits runs against an HBM budget with it */
size_t chq_bytes_for(const chq_mesh_desc *mesh, const chq_config *cfg);

/* Host <-> device copies in the reference layout (pgen::Initialize's deep_copy,
 * pgen.hpp:302-306; outputs, output.hpp:33-34; restart, restart.hpp:146-149). `count` is
 * the number of elements of the host buffer and must equal chq_array_count(). */
size_t chq_array_count(const chq_ctx *ctx, int array);
int chq_upload(chq_ctx *ctx, int array, const void *host, size_t count);
int chq_download(chq_ctx *ctx, int array, void *host, size_t count);

/* ---- stage kernels with the reference's semantics (all six conserved variables) ---- */
/* kernels::FluxesDispatch<DIR> (physics.hpp:505-514): writes f{dir+1} and its two face
 * EMFs from w/bcc and the face-normal field of buffer set `cur`. */
int chq_fluxes(chq_ctx *ctx, int dir, int cur);
/* kernels::CornerEMF (physics.hpp:242-297): writes e1, e2, e3. */
int chq_corner_emf(chq_ctx *ctx);
/* kernels::Update (physics.hpp:307-348) with ua = u0/b?f0, ub = u1/b?f1, out = set `out`. */
int chq_update(chq_ctx *ctx, int out, double g0, double g1, double beta_dt);
/* kernels::Boundaries (physics.hpp:354-432) on buffer set `buf`. */
int chq_boundaries(chq_ctx *ctx, int buf);
/* kernels::ConsToPrim (physics.hpp:437-494) on set `buf`; when compute_dt, *min_dt receives
 * the CFL minimum (synchronises); min_dt may be NULL otherwise. */
int chq_cons_to_prim(chq_ctx *ctx, int buf, int compute_dt, double *min_dt);
/* fofc::FlagBadCells (fofc.hpp:17-44): writes flag, *count receives the number (synchronises). */
int chq_fofc_flag(chq_ctx *ctx, int buf, int *count);
/* fofc::RepairFluxes<DIR> (fofc.hpp:46-96) from w/bcc and face B of set `cur`. */
int chq_fofc_repair(chq_ctx *ctx, int dir, int cur);
/* diffusion::Fluxes<DIR> (diffusion.hpp:131-138): adds the viscous, conductive and resistive
 * fluxes of the context's coefficients to f{dir+1} on the active faces, from w/bcc and the
 * face B of set `cur`. */
int chq_diffusion_fluxes(chq_ctx *ctx, int dir, int cur);
/* diffusion::ResistiveEMF (diffusion.hpp:141-156): adds eta J to e1, e2, e3. */
int chq_resistive_emf(chq_ctx *ctx, int cur);
/* diffusion::TimeStep (diffusion.hpp:159-174) from the current primitives (synchronises). */
int chq_diffusion_dt(chq_ctx *ctx, double *dt);
/* positivity::LimitFluxes<DIR> (positivity.hpp:24-89): blends the current dir fluxes toward the
 * local Lax-Friedrichs flux where the half-updates with beta_dt would leave rho or p below the
 * floors; *limited (may be null; then no synchronisation) = faces limited. */
int chq_limit_fluxes(chq_ctx *ctx, int dir, double beta_dt, long *limited);
/* srcterms::ApplyGravity (srcterms.hpp:21-40): gravity sources into u[out] with beta_dt from the
 * current primitives (no-op unless [gravity] is on). */
int chq_apply_gravity(chq_ctx *ctx, int out, double beta_dt);
/* srcterms::ApplyCooling (srcterms.hpp:55-71): exact power-law cooling of u[buf] over dt from
 * the current primitives (no-op unless [cooling] is on). */
int chq_apply_cooling(chq_ctx *ctx, int buf, double dt);
/* Stage() (main.cpp:46-93): fluxes, corner EMF, update, fofc, boundaries; *nbad (may be
 * NULL) receives the fofc count (synchronises when non-NULL). */
int chq_stage(chq_ctx *ctx, int cur, int out, double g0, double g1, double beta_dt, long *nbad);

/* ---- the time loop with dt, t and the fofc count resident on the device ---- */
/* main.cpp:192-201: boundaries(u0) + cons_to_prim(u0, dt) and sets t, tlim, nlim. */
int chq_init(chq_ctx *ctx, double t, double tlim, int nlim);
/* chq_init from a restart: the loop continues at (t, cycle) (main.cpp:196-201) */
int chq_init_resume(chq_ctx *ctx, double t, int cycle, double tlim, int nlim);
/* enqueue `nsteps` RK steps (main.cpp:246-292); asynchronous, no host round trip. Steps
 * after t reaches tlim (or cycle reaches nlim) are no-ops on the device. Without dual energy,
 * u[IENT] = rho p / rho^gamma of the final state is stored by the next chq_download of
 * CHQ_U0 (or per-kernel entry point) rather than by every step. */
int chq_step(chq_ctx *ctx, int nsteps);
/* enqueue nsteps and time them with HIP events on the context stream; *ms = elapsed */
int chq_step_timed(chq_ctx *ctx, int nsteps, double *ms);
int chq_get_status(chq_ctx *ctx, chq_status *st);
int chq_sync(chq_ctx *ctx);

/* Per-kernel timing of chq_step (no reference counterpart: measurement support). When a
 * class is selected, HIP events are recorded on the context stream around each launch of
 * that class; chq_kernel_timing_read synchronises and returns the summed duration and the
 * number of launches since selection (then starts over). kernel_class -1 turns it off;
 * CHQ_K_ALL brackets every class, read per class with chq_kernel_timing_breakdown.
 * CHQ_K_REPAIR brackets the whole fofc branch of a stage (repair, edge EMFs, re-update,
 * boundaries and con2prim of the repaired state; its launches exit at once when nothing was
 * flagged). */
enum chq_kernel_class {
  CHQ_K_FLUX = 0, CHQ_K_CORNER, CHQ_K_UPDATE, CHQ_K_BC, CHQ_K_C2P, CHQ_K_REPAIR, CHQ_K_NCLASS,
  CHQ_K_ALL = 64
};
int chq_kernel_timing(chq_ctx *ctx, int kernel_class);
int chq_kernel_timing_read(chq_ctx *ctx, double *total_ms, int *launches);
/* per-class totals since selection (CHQ_K_ALL): ms[CHQ_K_NCLASS], launches[CHQ_K_NCLASS] */
int chq_kernel_timing_breakdown(chq_ctx *ctx, double *ms, int *launches);
void *chq_stream(chq_ctx *ctx);

/* ---- host layer: decks, problem setups, outputs (host-only, no device needed) ---- */
typedef struct chq_deck chq_deck;
/* Params::ReadFile + Override (params.hpp:19-44) on `text`, then the driver's configuration
 * (main.cpp:112-185) and the problem generator (pgen.hpp:269-307: cpaw, orszag_tang, blast,
 * shock_tube). `overrides`: newline-separated block.key=value lines, or NULL. */
int chq_deck_open(const char *text, const char *overrides, chq_deck **out);
/* the same, with the initial state built on first use (chq_deck_initial / chq_deck_upload), so
 * a serving loop can parse at arrival and set up at admission; problem-setup errors then
 * surface there */
int chq_deck_open_lazy(const char *text, const char *overrides, chq_deck **out);
/* build a lazily opened deck's initial state now (e.g. on a worker thread); not to be called
 * concurrently with other calls on the same deck */
int chq_deck_prepare(chq_deck *deck);
int chq_deck_close(chq_deck *deck);
const char *chq_deck_last_error(void);
int chq_deck_mesh(const chq_deck *deck, chq_mesh_desc *mesh);
int chq_deck_config(const chq_deck *deck, chq_config *cfg);
int chq_deck_run(const chq_deck *deck, double *tlim, int *nlim, int *ndiag);
/* value of block.key after defaults were recorded ("" if absent); valid until close */
const char *chq_deck_get(const chq_deck *deck, const char *block, const char *key);
/* initial state (CHQ_U0, CHQ_W, CHQ_B1F0, CHQ_B2F0, CHQ_B3F0) in the reference layout */
int chq_deck_initial(const chq_deck *deck, int array, void *host, size_t count);
/* Params::Print (params.hpp:82-90) into buf; returns the length needed (with the NUL) */
size_t chq_deck_print(const chq_deck *deck, char *buf, size_t len);
/* upload the initial state (u0, b?f0) into a context created from chq_deck_mesh/config; the
 * primitives follow from chq_init (con2prim over every cell, main.cpp:200-201) */
int chq_deck_upload(const chq_deck *deck, chq_ctx *ctx);

/* output::WriteVTK (output.hpp:32-85): legacy VTK, PLUTO layout, of the current primitives */
int chq_write_vtk(chq_ctx *ctx, const chq_mesh_desc *mesh, const char *path, double time, int cycle);
/* output::History (output.hpp:91-119) of set `buf`: out[10] = mass mom1 mom2 mom3 etot ekin
 * emag eint divB_max divB_L1 */
int chq_history(chq_ctx *ctx, const chq_mesh_desc *mesh, int buf, double out[10]);
/* chq_history for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); no counterpart in the reference. The
 * same kernels and the same bits as chq_history, with the partial sums in a static per-device
 * scratch (320 KiB, resolved by chq_create) instead of a buffer the library allocates: it calls
 * no hipMalloc / hipFree. Meshes of more than 131072 active (k, j) rows (nx2 * nx3) are refused
 * with CHQ_EINVAL. All calls on one GPU share the scratch: calls that overlap on one GPU (two
 * contexts on different streams or threads) race and must be serialised by the caller. */
int chq_history_noalloc(chq_ctx *ctx, const chq_mesh_desc *mesh, int buf, double out[10]);
/* restart::Write / restart::Read (restart.hpp:19-58): the reference's binary restart format
 * (header + u0 + b1f0..b3f0, FP64), readable by either code. After chq_restart_read, call
 * chq_init_resume with the returned time and cycle. */
int chq_restart_write(chq_ctx *ctx, const char *path, double time, int cycle);
int chq_restart_read(chq_ctx *ctx, const char *path, double *time, int *cycle);

/* ---- batches: many independent runs advanced by shared launches (no counterpart in the
 * reference, which advances one mesh with a host-driven loop, main.cpp:246-297) ----
 * Every run is a context created, filled and initialised (chq_init) as for chq_step; the batch
 * then advances all of them together: one launch per kernel class and RK stage walks a
 * device-resident list of (run, tile) items, each run with its own dt, t, cycle, tlim, CFL,
 * boundary conditions and size (1D, 2D and 3D runs mix). A batch holds runs of one scheme
 * (reconstruction, Riemann solver, integrator, precision; dual energy off), fixed by its
 * first run. Results are bitwise those of chq_step on the same context. The RK step is
 * captured in a hipGraph and replayed; runs joining or leaving only rewrite device lists.
 * While in a batch a context is not stepped by chq_step; its arrays and status may be read
 * (reads wait for the batch's stream). */
typedef struct chq_batch chq_batch;
int chq_batch_create(int capacity, chq_batch **out);
int chq_batch_destroy(chq_batch *batch);
int chq_batch_add(chq_batch *batch, chq_ctx *ctx);
int chq_batch_remove(chq_batch *batch, chq_ctx *ctx);   /* the last run moves into its slot */
int chq_batch_size(const chq_batch *batch);
chq_ctx *chq_batch_run(chq_batch *batch, int index);
int chq_batch_step(chq_batch *batch, int nsteps);      /* asynchronous */
int chq_batch_step_timed(chq_batch *batch, int nsteps, double *ms);
int chq_batch_sync(chq_batch *batch);
/* runs that have not reached tlim / nlim; flags[i] (optional, size() entries) = 1 if run i
 * has not */
int chq_batch_unfinished(chq_batch *batch, int *count, int *flags);
int chq_batch_use_graph(chq_batch *batch, int on);     /* default on */
/* Autotuner: times every kernel class over candidate grid sizes on the current runs (stage-1
 * configuration; touches only arrays a step rewrites) and keeps the fastest; best_ms = sum of
 * the best per-class launch times. Call between steps. */
int chq_batch_tune(chq_batch *batch, int reps, double *best_ms);
/* grid size (blocks) of kernel class `cls` (0..10) and its name */
int chq_batch_grid(chq_batch *batch, int cls, int *grid, const char **name);
int chq_batch_set_grid(chq_batch *batch, int cls, int grid);

/* ---- serving: continuous batching of arriving runs (no counterpart in the reference) ----
 * Runs arrive as decks (chq_deck_open text + overrides) and wait in arrival order; each loop
 * iteration (chq_server_pump) retires the runs that reached tlim / nlim in the chunk the
 * previous iteration launched, launches the next chunk (`chunk_steps` steps of every batch,
 * one batch per scheme, each on its own stream) and, while it runs, admits what fits the HBM
 * budget (chq_bytes_for) and the batch slots (first fit; the first run that does not fit
 * keeps its memory reserved). It returns without waiting for the chunk. A finished run keeps
 * its device state until chq_server_release; read it through chq_server_job_ctx.
 * Mixed-precision runs whose state stops being finite are re-run in FP64. */
typedef struct chq_server chq_server;
int chq_server_create(size_t hbm_budget, int slots_per_batch, int chunk_steps, chq_server **out);
int chq_server_destroy(chq_server *srv);
/* precision: -1 the deck's, 0 FP64, 1 mixed */
int chq_server_submit(chq_server *srv, const char *deck_text, const char *overrides, int precision, int *job_id);
int chq_server_pump(chq_server *srv, int max_done, int *done_ids, int *n_done);
/* info = state (0 queued 1 running 2 done 3 failed 4 released), precision, fell back to FP64,
 * cycle, t, submit / admit / done times (s, chq_server_clock), device bytes, fofc corrections */
int chq_server_job(chq_server *srv, int job_id, double info[10]);
chq_ctx *chq_server_job_ctx(chq_server *srv, int job_id);
const char *chq_server_job_error(chq_server *srv, int job_id);
int chq_server_release(chq_server *srv, int job_id);
/* stats = queued, running, finished, bytes in use, budget, batches, batch steps issued, host
 * seconds spent launching steps, admitting, waiting for polls, retiring */
int chq_server_stats(chq_server *srv, double stats[11]);
double chq_server_clock(void);

#ifdef __cplusplus
}
#endif
#endif
