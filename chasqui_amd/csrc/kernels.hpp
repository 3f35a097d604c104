// Kernel argument blocks and launch helpers shared by the stage kernels and the C-ABI.
//
// T is the type of the arrays the face-flux / corner-EMF side works in: double for the FP64
// drop-in, float for the mixed-precision path (FP32 primitives, fluxes and EMFs). Conserved
// variables and face-centred B are always double.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.hpp"

namespace chq {

// A face-flux launch: one direction, faces [il..iu] x [jl..ju] x [kl..ku] (inclusive).
// Division by a divisor fixed before launch, for numerators below 2^31 (Granlund and Montgomery,
// "Division by invariant integers using multiplication", 1994, Theorem 4.2): with
// l = ceil(log2 d) and m = ceil(2^(31+l) / d), 0 <= m d - 2^(31+l) < d <= 2^l, so
// floor(n / d) = floor(n m / 2^(31+l)) for every 0 <= n < 2^31, and m < 2^32. A 32 x 32 -> 64-bit
// multiply and a shift per lane instead of an integer division (about 100 instructions for a
// 64-bit one). 1 <= d < 2^31 (a d below 1 is taken as 1).
struct FastDiv {
  unsigned m;
  int s;
};
__host__ __device__ inline FastDiv fast_div(int d) {
  if (d < 1) d = 1;
  int l = 0;
  while ((1LL << l) < d) ++l;
  const unsigned long long p = 1ULL << (31 + l);
  return FastDiv{unsigned((p + (unsigned long long)d - 1) / (unsigned long long)d), 31 + l};
}
__host__ __device__ __forceinline__ unsigned fdiv(unsigned n, FastDiv f) {
  return unsigned(((unsigned long long)n * f.m) >> f.s);
}

template <class T>
struct FluxArgsT {
  Geom g;
  Eos eos;
  int flatten, low_mach;
  int entropy;        // reconstruct and flux the entropy variable (IENT)
  int charac;         // reconstruct in characteristic variables (FP64 single-mesh launches only)
  double lhlld_phi;
  int il, iu, jl, ju, kl, ku;
  int seg;            // faces per thread of the y/z march (fluxk::march_seg)
  FastDiv q1, q2;     // x faces: by the cells per row and by the rows per plane of the face
                      // range (x_shape); y/z faces: by nx and by nt (m_item)
  const T *w;         // primitives, 6 components (IENT used only when entropy is transported)
  const T *bcc;       // cell-centred B, 3 components
  const double *bn;   // face-normal field of this direction (the stage-input face B)
  T *flx;             // fluxes, 6 components
  T *et2, *et1;       // face EMFs: et2 = -F(B_t1), et1 = +F(B_t2)
  const int *flag;    // fofc flags (repair launches only)
  const Ctrl *ctrl;
};
using FluxArgs = FluxArgsT<double>;

template <class T>
struct CornerArgsT {
  Geom g;
  const T *f1, *f2, *f3;       // mass-flux components (IDN) of the face fluxes
  const T *e2x1, *e3x1, *e1x2, *e3x2, *e1x3, *e2x3;
  const T *ecc;
  T *e1, *e2, *e3;
  const Ctrl *ctrl;
  int only_if_bad;             // skip unless the current stage flagged cells (fofc re-run)
  int xg;                      // XCD tile grouping (xcd_tile)
};
using CornerArgs = CornerArgsT<double>;

// RK combine + divergence + CT: fluxes and edge EMFs of type T are widened to double and
// every sum, product and combination is done in double.
template <class T>
struct UpdateArgsT {
  Geom g;
  double g0, g1, beta;         // beta_dt = beta * dt, dt from ctrl (or beta_dt given)
  double beta_dt;              // used when use_ctrl_dt == 0
  int use_ctrl_dt;
  int nvar;                    // 5 (entropy not transported) or 6
  const T *f1, *f2, *f3;
  const double *ua, *ub;
  double *uo;
  const double *b1a, *b2a, *b3a, *b1b, *b2b, *b3b;
  double *b1o, *b2o, *b3o;
  const T *e1, *e2, *e3;
  const Ctrl *ctrl;
  int only_if_bad;
  // FP64 first pass of a stage that may re-update locally (output = ub, g1 = 0): count into
  // *zcount the points whose recomputation could differ from this pass, i.e. where g0 * ua is
  // -0 and the divergence (curl) term is 0, or where the result is not finite
  int zwatch;
  int *zcount;
  int xg;                      // XCD tile grouping (xcd_tile)
};
using UpdateArgs = UpdateArgsT<double>;

struct BcArgs {
  Geom g;
  int bc, nvar;
  double *u, *b1, *b2, *b3;
  const Ctrl *ctrl;
  int only_if_bad;
};

// con2prim computes in double and stores primitives, cell B and cell EMF as T
template <class T>
struct C2pArgsT {
  Geom g;
  Eos eos;
  const double *b1, *b2, *b3;
  double *u;
  T *w, *bcc, *ecc;
  int write_went;              // also store w[IENT] (needed when entropy is transported)
  int entropy_out;             // resynchronise u[IENT] = rho p / rho^gamma
  int compute_dt;
  int slot;                    // which running minimum receives the CFL dt (0 or 1)
  int flag_mode;               // 0: none; 1: flag array + count; 2: append flagged cells to badlist
  int *flag;
  int *badlist;
  Ctrl *ctrl;
  int only_if_bad;
  int xg;                      // XCD tile grouping (xcd_tile)
};
using C2pArgs = C2pArgsT<double>;

// ---------------------------------------------------------------- work decomposition
// The memory-bound kernels walk "virtual blocks": chunks of 256 consecutive points of one
// (j, i) plane of an nx x ny box (rows are linearised, so rows of 257 or 262 points waste no
// lanes), times nz planes.
// XCD-grouped tile order: the dispatcher deals consecutive workgroups of a launch to the 8 XCDs
// in turn, so neighbouring tiles (rows that read each other's lines) land in different L2s.
// Block b of n takes tile xcd_tile(b, n, grp): each XCD gets runs of grp consecutive tiles,
// groups of 8 * grp tiles in launch order (the tail past the last whole group keeps b). A
// bijection on [0, n); grp <= 1 is the identity.
__host__ __device__ __forceinline__ int xcd_tile(int b, int n, int grp) {
  if (grp <= 1) return b;
  const int span = 8 * grp, full = (n / span) * span;
  if (b >= full) return b;
  const int l = b >> 3;
  return (l / grp) * span + (b & 7) * grp + l % grp;
}
// Group sizes of the fused stage's kernels, measured at 256^3 (ms per RK3 step, FP64 / mixed):
// con2prim 32 (1.65 -> 1.58 / 1.28 -> 1.16; 4, 8, 16 and 64 gain less), corner EMF 4 (1.47 ->
// 1.44 / 0.90 -> 0.88; 8 and 16 the same), the float update march 8 (2.62 -> 2.60); the FP64
// update keeps the launch order (runs of 2, 3, 4, 8 or 16 made it 0.4-1.2% slower).
constexpr int XG_C2P = 32, XG_CORNER = 4, XG_UPDATE_F32 = 8;

struct VB {
  int nx, ny, cpp, n;   // plane size, chunks per plane, total chunks
  __device__ __forceinline__ bool at(int vb, int &k, int &j, int &i) const {
    k = vb / cpp;
    const int t = (vb - k * cpp) * 256 + int(threadIdx.x);
    j = t / nx;
    i = t - j * nx;
    return t < nx * ny;
  }
};
inline __host__ __device__ VB vgrid(int nx, int ny, int nz) {
  VB v;
  v.nx = nx;
  v.ny = ny;
  v.cpp = div_up(nx * ny, 256);
  v.n = v.cpp * nz;
  return v;
}
// corner EMF: the union box [is..ie+1] x [js..je+1] x [ks..ke+1] (one k plane below 3D)
inline __host__ __device__ VB corner_vb(const Geom &g) {
  return vgrid(g.ie - g.is + 2, g.je - g.js + 2, g.three_d ? g.ke - g.ks + 2 : 1);
}
// update: active cells plus the upper faces of the present directions
inline __host__ __device__ VB update_vb(const Geom &g) {
  return vgrid(g.ie - g.is + 2, g.two_d ? g.je - g.js + 2 : 1, g.three_d ? g.ke - g.ks + 2 : 1);
}
// con2prim: every cell, ghosts included
inline __host__ __device__ VB c2p_vb(const Geom &g) { return vgrid(g.ni, g.nj, g.nk); }
// boundaries of direction d: threads over (k, j, ghost), (k, ghost, i) or (ghost, j, i)
inline __host__ __device__ long long bc_points(const Geom &g, int d) {
  if (d == 0) return (long long)(g.nk + 1) * (g.nj + 1) * g.ng;
  if (d == 1) return (long long)(g.nk + 1) * g.ng * (g.ni + 1);
  return (long long)g.ng * (g.nj + 1) * (g.ni + 1);
}

// face-flux decomposition (flux_impl.hpp): x faces by wave64 windows over the rows' cells laid
// end to end, y/z faces by threads marching a.seg faces along the direction; both linearised so
// that no lane idles at the end of a row (a 258-wide row in 64-wide tiles would leave a fifth
// of the lanes idle)
namespace fluxk {
constexpr int BX = 64, BY = 4;
// Faces per thread in the y/z march. SEG where the launch's march threads fill the device
// (measured at 256^3: 24 beats 8, 16, 32 in FP64 and 13, 20, 26, 37, 52 in FP32); SEG_SMALL
// where they would leave it idle (small meshes, batches of small runs), and SEG_TINY where
// they would use a sixteenth of it, since one thread's serial march is then the launch's
// latency. The choice moves work between threads only: every
// face is computed by the same expressions either way.
constexpr int SEG = 24, SEG_SMALL = 8, SEG_TINY = 4;
constexpr long long SEG_FILL = 131072;   // march threads of two 256-thread blocks per CU, 256 CUs
// march threads of direction d (1 or 2) for segments of `seg` faces, from the face ranges
inline __host__ __device__ long long march_threads(int il, int iu, int jl, int ju, int kl, int ku, int d, int seg) {
  const long long nx = iu - il + 1, nt = d == 1 ? ku - kl + 1 : ju - jl + 1, nd = d == 1 ? ju - jl + 1 : ku - kl + 1;
  return nx * nt * ((nd + seg - 1) / seg);
}
inline __host__ __device__ int march_seg(long long threads_at_seg) {
  return threads_at_seg < SEG_FILL / 16 ? SEG_TINY : threads_at_seg < SEG_FILL ? SEG_SMALL : SEG;
}

template <Recon RC> constexpr bool wide_v = !(RC == Recon::dc || RC == Recon::plm);
inline __host__ __device__ int fpw_of(Recon rc) {
  const bool w = !(rc == Recon::dc || rc == Recon::plm);
  return 64 - (w ? 3 : 2) - (w ? 2 : 1);
}
template <Recon RC> constexpr int fpw_v = 64 - (wide_v<RC> ? 3 : 2) - (wide_v<RC> ? 2 : 1);   // faces per wave

// x-face work: the cells of all rows (ni per row, rows (j, k) of the face range) laid end to
// end; wave w holds cells w * fpw - H ... w * fpw - H + 63 and computes the faces of its middle
// fpw lanes; 4 waves per block. nseg = ni (cells per row).
template <class T>
__host__ __device__ inline void x_shape(const FluxArgsT<T> &a, int fpw, int &nseg, int &nwaves) {
  nseg = a.g.ni;
  const long long cells = (long long)nseg * (a.ju - a.jl + 1) * (a.ku - a.kl + 1);
  nwaves = int((cells + fpw - 1) / fpw);
}

// y/z-face work: one thread per (i, transverse row, segment of a.seg faces along DIR), i fastest,
// in blocks of BX x BY threads (gx blocks; gy = gz = 1).
template <class T, int DIR>
__host__ __device__ inline void m_shape(const FluxArgsT<T> &a, int &gx, int &gy, int &gz) {
  const int nx = a.iu - a.il + 1;
  const int nt = DIR == 1 ? a.ku - a.kl + 1 : a.ju - a.jl + 1;
  const int nd = DIR == 1 ? a.ju - a.jl + 1 : a.ku - a.kl + 1;
  gx = int(((long long)nx * nt * div_up(nd, a.seg) + BX * BY - 1) / (BX * BY));
  gy = 1;
  gz = 1;
}
// thread `lin` of the y/z decomposition: i, transverse index t and segment sg (false: no work)
template <class T, int DIR>
__host__ __device__ inline bool m_item(const FluxArgsT<T> &a, long long lin, int &i, int &t, int &sg) {
  const int nx = a.iu - a.il + 1;
  const int nt = DIR == 1 ? a.ku - a.kl + 1 : a.ju - a.jl + 1;
  const int nd = DIR == 1 ? a.ju - a.jl + 1 : a.ku - a.kl + 1;
  const unsigned r = fdiv(unsigned(lin), a.q1), rs = fdiv(r, a.q2);   // lin < 2^31 (mesh limit)
  i = a.il + int(unsigned(lin) - r * unsigned(nx));
  t = int(r - rs * unsigned(nt));
  sg = int(rs);
  return sg < div_up(nd, a.seg);
}

}  // namespace fluxk

// ---------------------------------------------------------------- batches
// Many independent meshes advanced by the same launches. Every run keeps its own arena and
// control block; the launches walk device-resident work lists of (run, tile) items, so a
// captured step graph stays valid while runs join and leave (only list contents change).
struct Item { int run, tile; };
struct ItemList {
  const Item *item;
  const int *n;        // number of items (device memory)
};

// every launch description of one run, for each (stage-input buffer, primitive set)
template <class T>
struct RunArgs {
  FluxArgsT<T> flux[2][2][3];    // [input buffer][primitive set][dir]
  CornerArgsT<T> corner[2];      // [primitive set]
  UpdateArgsT<T> update[2];      // [output buffer]; coefficients come with the launch
  C2pArgsT<T> c2p[2][2];         // [buffer][primitive set written]
  BcArgs bc[2][3];               // [buffer][dir]
  Ctrl *ctrl;
  int *badlist;                  // cells flagged by con2prim in the current stage
  int ndir;
  int cond_tiles[5];             // tiles of the whole-mesh fofc re-runs: update, bc x1..x3, c2p
};

// The fofc branch of a batch: the runs that flagged cells in the current stage, compacted on
// the device, with per-class tile prefixes (classes below).
enum : int { CL_REPAIR = 0, CL_CORNER, CL_UPDATE, CL_BC1, CL_BC2, CL_BC3, CL_C2P, CL_N };
struct CondList {
  int cap;             // capacity in runs
  int *base;           // [0] = number of flagged runs, then run[cap], then pre[CL_N][cap + 1]
  __host__ __device__ int *nrun() const { return base; }
  __host__ __device__ int *run() const { return base + 1; }
  __host__ __device__ int *pre(int cls) const { return base + 1 + cap + cls * (cap + 1); }
  static size_t ints(int cap) { return size_t(1) + cap + size_t(CL_N) * (cap + 1); }
};

// launches (defined in stage.hip / flux.hip / flux32.hip)
hipError_t launch_flux(const FluxArgsT<double> &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s);
hipError_t launch_flux(const FluxArgsT<float> &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s);
// characteristic-variable reconstruction (FP64; launch_flux forwards when a.charac is set)
hipError_t launch_flux_charac(const FluxArgsT<double> &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s);
hipError_t launch_repair(const FluxArgs &a, int dir, bool only_if_bad, hipStream_t s);
hipError_t launch_repair_list(const FluxArgsT<double> a[3], int ndir, const int *list, hipStream_t s);
hipError_t launch_repair_list(const FluxArgsT<float> a[3], int ndir, const int *list, hipStream_t s);
template <class T> hipError_t launch_corner(const CornerArgsT<T> &a, hipStream_t s);
template <class T> hipError_t launch_corner_list(const CornerArgsT<T> &a, const int *list, hipStream_t s);
template <class T> hipError_t launch_update(const UpdateArgsT<T> &a, hipStream_t s);
// u[IENT] = rho p / rho^gamma over every cell of buffer a.u with con2prim's expressions (the
// deferred resynchronisation of the fast path's last stage); writes nothing else
hipError_t launch_ient_sync(const C2pArgsT<double> &a, hipStream_t s);
// the local fofc re-update: stamp the points of the update box within one cell of each listed
// (flagged) cell with the control block's epoch, then recompute exactly the stamped points
hipError_t launch_mark_near(const Geom &g, const Ctrl *ctrl, const int *list, int *marks, hipStream_t s);
template <class T> hipError_t launch_update_near(const UpdateArgsT<T> &a, const int *marks, hipStream_t s);
hipError_t launch_bc(const BcArgs &a, int dir, hipStream_t s);

// explicit diffusion (FP64 single-mesh path), diffusion.hip
struct DiffArgs {
  Geom g;
  double nu, kappa, eta, gm1;
  const double *w, *bcc;                 // primitives, cell-centred B
  const double *b1, *b2, *b3;            // face B the current density is taken from
  double *f[3];                          // face fluxes (6 components each)
  double *e1, *e2, *e3;                  // edge EMFs
  const int *flag;                       // fofc flags (repair launches only)
  const Ctrl *ctrl;
  int only_if_bad;                       // skip unless the current stage flagged cells (fofc branch)
};
hipError_t launch_diff_flux(const DiffArgs &a, int dir, bool flagged_only, hipStream_t s);
hipError_t launch_resistive_emf(const DiffArgs &a, hipStream_t s);
// minimum of w[IDN] over the active cells into ctrl->rho_bits (reset first)
hipError_t launch_rho_min(const double *w, const Geom &g, Ctrl *c, bool reset, hipStream_t s);
// ctrl->min_dt = min(ctrl->min_dt, parabolic limit from ctrl->rho_bits) when the step ran
hipError_t launch_diff_dt_fold(Ctrl *c, const DiffArgs &a, hipStream_t s);
// super-time-stepping pieces
hipError_t launch_fill_zero(double *p, long long n, hipStream_t s);
hipError_t launch_sts_div(const Geom &g, const double *f1, const double *f2, const double *f3, const double *e1,
                          const double *e2, const double *e3, double *mu, double *mb1, double *mb2, double *mb3,
                          hipStream_t s);
struct StsTerm { const double *x; double a; };
hipError_t launch_sts_combine(const Geom &g, double *out, const StsTerm t[5], int ncomp, bool faces, hipStream_t s);
hipError_t launch_ctrl_take_min(Ctrl *c, int slot, hipStream_t s);   // ctrl->min_dt = running minimum `slot`

// positivity flux limiter and source terms (FP64 single-mesh path), srcterms.hip
struct SrcArgs {
  Geom g;
  Eos eos;
  double x1min, x2min, x3min;
  int grav_uniform, grav_point;
  double g1, g2, g3, gm, gx0, gy0, gz0, rsoft;
  double lambda0, cool_alpha, tfloor;
  double beta;                           // stage weight: beta * ctrl->dt, or beta itself
  int use_ctrl_dt;
  double *u;                             // conserved state the sources act on
  const double *w, *bcc;
  double *f[3];
  Ctrl *ctrl;
  int only_if_bad;                       // fofc branch: skip unless cells were flagged
};
hipError_t launch_limit_fluxes(const SrcArgs &a, int dir, hipStream_t s);
hipError_t launch_gravity(const SrcArgs &a, hipStream_t s);
hipError_t launch_cooling(const SrcArgs &a, hipStream_t s);
template <class T> hipError_t launch_c2p(const C2pArgsT<T> &a, hipStream_t s);
hipError_t launch_flag(const C2pArgs &a, hipStream_t s);   // standalone FlagBadCells
// control-block kernels
hipError_t launch_ctrl_begin_step(Ctrl *c, hipStream_t s);
// stage: RK stage index of the time loop (selects the adaptive fofc order), -1 otherwise
hipError_t launch_ctrl_begin_stage(Ctrl *c, int reset_slots, hipStream_t s, int stage = -1);
hipError_t launch_ctrl_end_stage(Ctrl *c, int last, int fused, hipStream_t s);
hipError_t launch_ctrl_end_step(Ctrl *c, hipStream_t s);

// batched launches: `grid` blocks walk the list (grid-stride); stage coefficients are
// launch arguments, everything per run comes from the RunArgs table
hipError_t launchb_flux(const RunArgs<double> *R, ItemList L, int dir, int cur, Recon rc, RSolver rs, int grid, hipStream_t s);
hipError_t launchb_flux(const RunArgs<float> *R, ItemList L, int dir, int cur, Recon rc, RSolver rs, int grid, hipStream_t s);
hipError_t launchb_repair(const RunArgs<double> *R, CondList C, int cur, int grid, hipStream_t s);
hipError_t launchb_repair(const RunArgs<float> *R, CondList C, int cur, int grid, hipStream_t s);
// the fofc branch of a batch of small runs: one block per run (cap blocks), the repair, then the
// edge EMFs, re-update, boundaries and con2prim of the run
hipError_t launchb_repair_runs(const RunArgs<double> *R, const int *nruns, int cap, int cur, hipStream_t s);
hipError_t launchb_repair_runs(const RunArgs<float> *R, const int *nruns, int cap, int cur, hipStream_t s);
template <class T> hipError_t launchb_fofc_runs(const RunArgs<T> *R, const int *nruns, int cap, int out, double g0,
                                                double g1, double beta, int last, hipStream_t s);
template <class T> hipError_t launchb_corner(const RunArgs<T> *R, ItemList L, int grid, hipStream_t s);
template <class T> hipError_t launchb_update(const RunArgs<T> *R, ItemList L, int out, double g0, double g1, double beta, int grid, hipStream_t s);
template <class T> hipError_t launchb_bc(const RunArgs<T> *R, ItemList L, int dir, int buf, int grid, hipStream_t s);
template <class T> hipError_t launchb_c2p(const RunArgs<T> *R, ItemList L, int buf, int last, int grid, hipStream_t s);
template <class T> hipError_t launchb_fofc_scan(const RunArgs<T> *R, const int *nruns, CondList C, hipStream_t s);
template <class T> hipError_t launchb_corner_cond(const RunArgs<T> *R, CondList C, int grid, hipStream_t s);
template <class T> hipError_t launchb_update_cond(const RunArgs<T> *R, CondList C, int out, double g0, double g1, double beta, int grid, hipStream_t s);
template <class T> hipError_t launchb_bc_cond(const RunArgs<T> *R, CondList C, int dir, int buf, int grid, hipStream_t s);
template <class T> hipError_t launchb_c2p_cond(const RunArgs<T> *R, CondList C, int buf, int last, int grid, hipStream_t s);
// control-block step of every run (one thread per table slot, `cap` slots); the end of a step
// also writes live[r] = run r has not reached tlim / nlim
template <class T> hipError_t launchb_ctrl(const RunArgs<T> *R, const int *nruns, int cap, int op, int arg, int *live, hipStream_t s);
// CTRL_STEP_FIRST: begin step + begin stage 1; CTRL_TURN: end a stage + begin the next (arg:
// the next one is the last); CTRL_END_LAST: end the last stage + end step (+ live flags)
enum : int { CTRL_BEGIN_STEP = 0, CTRL_BEGIN_STAGE, CTRL_END_STAGE, CTRL_END_STEP, CTRL_LIVE, CTRL_STEP_FIRST,
             CTRL_TURN, CTRL_END_LAST };

// position of `it` in an exclusive prefix pre[0..n] (largest p with pre[p] <= it)
__device__ __forceinline__ int locate(const int *pre, int n, int it) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= it) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace chq
