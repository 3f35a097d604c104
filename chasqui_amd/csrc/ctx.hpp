// The context behind the C ABI's chq_ctx handle: geometry, configuration, device arena and
// the launch descriptions of every stage kernel. Shared by the single-mesh entry points
// (capi.cpp) and the batch of many meshes (batch.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "../../include/chasqui_amd.h"
#include "kernels.hpp"

namespace chq {
// records the message returned by chq_last_error() and passes the status code through
int fail(int code, const std::string &msg);
#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(CHQ_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// history columns (output.hpp:91-119): 8 volume sums, max |div B| (slot HST_MAX), sum |div B|
constexpr int HST_NV = 10, HST_MAX = 8;

// RK coefficient table (g0, g1, beta, cur, out), main.cpp:250-268
struct Coef { double g0, g1, beta; int cur, out; };
inline Coef rk_coef(int nstage, int s) {
  if (nstage == 1) return {1, 0, 1, 0, 0};
  if (nstage == 2) return s == 1 ? Coef{1, 0, 1, 0, 1} : Coef{0.5, 0.5, 0.5, 1, 0};
  if (nstage == 3) {
    if (s == 1) return {1, 0, 1, 0, 1};
    if (s == 2) return {0.75, 0.25, 0.25, 1, 1};
    return {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 1, 0};
  }
  if (s == 1) return {1, 0, 0.5, 0, 1};
  if (s == 2) return {0, 1, 0.5, 1, 1};
  if (s == 3) return {2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 1, 1};
  return {0, 1, 0.5, 1, 0};
}
}  // namespace chq

namespace chq {
// waits for a batch advancing the context and refreshes the current primitive set (host reads)
int ctx_ready(chq_ctx *c);
// history reduction (history.hip): blocks of the first pass, and the passes over set `buf`
// into `work` (history_blocks(g) partials + HST_NV results), scaled into out[HST_NV]
int history_blocks(const Geom &g);
int history_scratch(int dev, double **ptr);
int history_on(chq_ctx *c, int buf, double *work, double out[HST_NV]);
}  // namespace chq

using namespace chq;

// The arrays the flux / corner-EMF side works in, per precision: double for the FP64 path,
// float for the mixed-precision path (the conserved state and face B are always double).
template <class T>
struct FieldSet {
  T *w[2]{}, *bcc[2]{}, *ecc[2]{};   // two primitive sets (current / tentative)
  T *flx[3]{}, *ef[6]{}, *ee[3]{};   // fluxes, face EMFs (e2x1 e3x1 e1x2 e3x2 e1x3 e2x3), edge EMFs
};

struct chq_ctx {
  Geom g{};
  Eos eos{};
  chq_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  char *arena = nullptr;
  size_t arena_bytes = 0;
  double *u[2]{}, *bf[2][3]{};
  FieldSet<double> f64;      // FP64 path
  FieldSet<float> f32;       // mixed precision (only one of the two sets is allocated)
  bool mixed = false;        // primitives, fluxes and EMFs are FP32
  int *flag = nullptr;
  int *badlist = nullptr;    // flagged cells of the current stage (fused fofc path); the area
                             // holds 2 cstride ints: the list, then one near-mark per point
  int *near_marks() const { return badlist + g.cstride; }
  Ctrl *ctrl = nullptr;      // time loop
  Ctrl *ctrl_api = nullptr;  // per-kernel entry points (always active)
  void *staging = nullptr;   // device staging buffer for layout conversion
  size_t staging_bytes = 0;
  double *hst_work = nullptr;   // history partials and results (chq_history)
  size_t hst_bytes = 0;
  double *hst_scratch = nullptr;   // this device's static history scratch (chq_history_noalloc)
  int wcur = 0;              // which primitive set (w, bcc, ecc) is current
  bool wcur_stale = false;   // the time loop may have stopped early: the device copy (Ctrl::wcur) rules
  bool ient_stale = false;   // u[0]'s u[IENT] awaits k_ient_sync: the fast path (entropy not
                             // transported) leaves it to the first host read of the state
  void *batch = nullptr;     // the batch advancing this context, if any
  hipStream_t batch_stream = nullptr;   // its stream (host reads wait for it)
  bool entropy = false;      // transport IENT in the time loop (only needed with dual energy)
  // super-time-stepping work arrays (sts.hpp:16-32), allocated at the first operator-split step:
  // y[0..2], m0, mj (NVAR components each) and the face sets yb[0..2], mb0, mbj (3 each)
  double *sts_buf = nullptr;
  double *sy[3]{}, *sm0 = nullptr, *smj = nullptr, *syb[3][3]{}, *smb0[3]{}, *smbj[3]{};
  long long sts_stages = 0;  // RKL2 stages taken so far
  double x1min = 0.0, x2min = 0.0, x3min = 0.0;   // domain origin (point-mass gravity)
  // per-kernel-class timing: event pairs recorded around launches of one class
  int timing_class = -1;          // a chq_kernel_class, CHQ_K_ALL, or -1 (off)
  std::vector<hipEvent_t> tev;   // begin/end pairs
  std::vector<int> tcls;         // class of each pair
  size_t tev_used = 0;

  void tmark(int cls, bool begin) {
    if (cls != timing_class && timing_class != CHQ_K_ALL) return;
    if (tev_used == tev.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      tev.push_back(e);
    }
    if (!begin && (tev_used % 2) == 0) return;
    if (begin && (tev_used % 2) == 1) return;   // unmatched begin: keep the open pair
    if (begin) {
      const size_t pair = tev_used / 2;
      if (tcls.size() <= pair) tcls.resize(pair + 1);
      tcls[pair] = cls;
    }
    (void)hipEventRecord(tev[tev_used++], stream);
  }

  // face-flux launch description for direction d from buffer set `cur`
  FluxArgs flux_args(int d, int cur, int wset, const Ctrl *c) const {
    FluxArgs a{};
    a.g = g;
    a.eos = eos;
    a.flatten = cfg.flatten;
    a.low_mach = cfg.low_mach;
    a.entropy = 1;
    a.charac = cfg.characteristic && cfg.recon != 0;   // physics.hpp:71
    a.lhlld_phi = cfg.lhlld_phi;
    // physics.hpp:85-88
    int il = g.is, iu = g.ie, jl = g.js, ju = g.je, kl = g.ks, ku = g.ke;
    if (d == 0) iu = g.ie + 1; else if (g.nx1 > 1) { il = g.is - 1; iu = g.ie + 1; }
    if (d == 1) ju = g.je + 1; else if (g.two_d) { jl = g.js - 1; ju = g.je + 1; }
    if (d == 2) ku = g.ke + 1; else if (g.three_d) { kl = g.ks - 1; ku = g.ke + 1; }
    a.il = il; a.iu = iu; a.jl = jl; a.ju = ju; a.kl = kl; a.ku = ku;
    a.seg = d == 0 ? fluxk::SEG : fluxk::march_seg(fluxk::march_threads(il, iu, jl, ju, kl, ku, d, fluxk::SEG));
    if (d == 0) { a.q1 = fast_div(g.ni); a.q2 = fast_div(ju - jl + 1); }
    else { a.q1 = fast_div(iu - il + 1); a.q2 = fast_div(d == 1 ? ku - kl + 1 : ju - jl + 1); }
    a.w = f64.w[wset];
    a.bcc = f64.bcc[wset];
    a.bn = bf[cur][d];
    a.flx = f64.flx[d];
    // et2 = e3x1 / e1x2 / e2x3 ; et1 = e2x1 / e3x2 / e1x3 (physics.hpp:93-94)
    static const int et2_id[3] = {1, 2, 5}, et1_id[3] = {0, 3, 4};   // ef = e2x1 e3x1 e1x2 e3x2 e1x3 e2x3
    a.et2 = f64.ef[et2_id[d]];
    a.et1 = f64.ef[et1_id[d]];
    a.flag = flag;
    a.ctrl = c;
    return a;
  }
  CornerArgs corner_args(int wset, const Ctrl *c, int only_if_bad) const {
    CornerArgs a{};
    a.g = g;
    a.f1 = f64.flx[0]; a.f2 = f64.flx[1]; a.f3 = f64.flx[2];
    a.e2x1 = f64.ef[0]; a.e3x1 = f64.ef[1]; a.e1x2 = f64.ef[2]; a.e3x2 = f64.ef[3]; a.e1x3 = f64.ef[4]; a.e2x3 = f64.ef[5];
    a.ecc = f64.ecc[wset];
    a.e1 = f64.ee[0]; a.e2 = f64.ee[1]; a.e3 = f64.ee[2];
    a.ctrl = c;
    a.only_if_bad = only_if_bad;
    return a;
  }
  UpdateArgs update_args(int out, double g0, double g1, int nvar, const Ctrl *c, int only_if_bad) const {
    UpdateArgs a{};
    a.g = g;
    a.g0 = g0; a.g1 = g1;
    a.nvar = nvar;
    a.f1 = f64.flx[0]; a.f2 = f64.flx[1]; a.f3 = f64.flx[2];
    a.ua = u[0]; a.ub = u[1]; a.uo = u[out];
    a.b1a = bf[0][0]; a.b2a = bf[0][1]; a.b3a = bf[0][2];
    a.b1b = bf[1][0]; a.b2b = bf[1][1]; a.b3b = bf[1][2];
    a.b1o = bf[out][0]; a.b2o = bf[out][1]; a.b3o = bf[out][2];
    a.e1 = f64.ee[0]; a.e2 = f64.ee[1]; a.e3 = f64.ee[2];
    a.ctrl = c;
    a.only_if_bad = only_if_bad;
    return a;
  }
  C2pArgs c2p_args(int buf, int wset, Ctrl *c) const {
    C2pArgs a{};
    a.g = g;
    a.eos = eos;
    a.b1 = bf[buf][0]; a.b2 = bf[buf][1]; a.b3 = bf[buf][2];
    a.u = u[buf];
    a.w = f64.w[wset]; a.bcc = f64.bcc[wset]; a.ecc = f64.ecc[wset];
    a.write_went = 1;
    a.entropy_out = 1;
    a.flag = flag;
    a.badlist = badlist;
    a.ctrl = c;
    return a;
  }
  hipError_t boundaries(int buf, int nvar, const Ctrl *c, int only_if_bad) {
    BcArgs a{};
    a.g = g;
    a.nvar = nvar;
    a.u = u[buf];
    a.b1 = bf[buf][0]; a.b2 = bf[buf][1]; a.b3 = bf[buf][2];
    a.ctrl = c;
    a.only_if_bad = only_if_bad;
    a.bc = cfg.bc[0];
    hipError_t e = launch_bc(a, 0, stream);
    if (e != hipSuccess) return e;
    if (g.two_d) { a.bc = cfg.bc[1]; e = launch_bc(a, 1, stream); if (e != hipSuccess) return e; }
    if (g.three_d) { a.bc = cfg.bc[2]; e = launch_bc(a, 2, stream); }
    return e;
  }
  int ndir() const { return 1 + (g.two_d ? 1 : 0) + (g.three_d ? 1 : 0); }
  bool diffusive() const { return cfg.nu > 0.0 || cfg.kappa > 0.0 || cfg.eta > 0.0; }
  // diffusion launch description: primitives of set `wset`, current density from face set `cur`
  // (or the given face arrays); in_stage = the coefficients the RK stage uses (none under STS)
  DiffArgs diff_args(const double *b1, const double *b2, const double *b3, int wset, const Ctrl *c, bool in_stage) const {
    DiffArgs a{};
    a.g = g;
    const bool on = !(in_stage && cfg.sts);
    a.nu = on ? cfg.nu : 0.0;
    a.kappa = on ? cfg.kappa : 0.0;
    a.eta = on ? cfg.eta : 0.0;
    a.gm1 = eos.gm1;
    a.w = f64.w[wset];
    a.bcc = f64.bcc[wset];
    a.b1 = b1; a.b2 = b2; a.b3 = b3;
    for (int d = 0; d < 3; ++d) a.f[d] = f64.flx[d];
    a.e1 = f64.ee[0]; a.e2 = f64.ee[1]; a.e3 = f64.ee[2];
    a.flag = flag;
    a.ctrl = c;
    return a;
  }
  DiffArgs diff_args(int cur, int wset, const Ctrl *c, bool in_stage) const {
    return diff_args(bf[cur][0], bf[cur][1], bf[cur][2], wset, c, in_stage);
  }
  bool diff_any(const DiffArgs &a) const { return a.nu > 0.0 || a.kappa > 0.0 || a.eta > 0.0; }
  bool gravity() const { return cfg.grav_uniform || cfg.grav_point; }
  // flux limiter / gravity / cooling launch description: sources act on u[out] with the
  // primitives of set `wset`; the step weight is beta * ctrl->dt (use_ctrl_dt) or beta itself
  SrcArgs src_args(int out, int wset, const Ctrl *c, double beta, bool use_ctrl_dt) const {
    SrcArgs a{};
    a.g = g;
    a.eos = eos;
    a.x1min = x1min; a.x2min = x2min; a.x3min = x3min;
    a.grav_uniform = cfg.grav_uniform; a.grav_point = cfg.grav_point;
    a.g1 = cfg.g1; a.g2 = cfg.g2; a.g3 = cfg.g3;   // added whenever gravity is on (srcterms.hpp:27)
    a.gm = cfg.gm; a.gx0 = cfg.gx0; a.gy0 = cfg.gy0; a.gz0 = cfg.gz0; a.rsoft = cfg.rsoft;
    a.lambda0 = cfg.lambda0; a.cool_alpha = cfg.cool_alpha; a.tfloor = cfg.tfloor;
    a.beta = beta;
    a.use_ctrl_dt = use_ctrl_dt ? 1 : 0;
    a.u = u[out];
    a.w = f64.w[wset];
    a.bcc = f64.bcc[wset];
    for (int d = 0; d < 3; ++d) a.f[d] = f64.flx[d];
    a.ctrl = const_cast<Ctrl *>(c);
    return a;
  }

  template <class T> FluxArgsT<T> flux_args_t(const FieldSet<T> &f, int d, int cur, int wset, const Ctrl *c) const {
    FluxArgs a0 = flux_args(d, cur, wset, c);
    FluxArgsT<T> a{};
    a.g = a0.g; a.eos = a0.eos; a.flatten = a0.flatten; a.low_mach = a0.low_mach; a.entropy = a0.entropy;
    a.charac = a0.charac;
    a.lhlld_phi = a0.lhlld_phi;
    a.il = a0.il; a.iu = a0.iu; a.jl = a0.jl; a.ju = a0.ju; a.kl = a0.kl; a.ku = a0.ku;
    a.seg = a0.seg;
    a.q1 = a0.q1; a.q2 = a0.q2;
    a.w = f.w[wset]; a.bcc = f.bcc[wset]; a.bn = a0.bn; a.flx = f.flx[d];
    static const int et2_id[3] = {1, 2, 5}, et1_id[3] = {0, 3, 4};
    a.et2 = f.ef[et2_id[d]];
    a.et1 = f.ef[et1_id[d]];
    a.flag = flag;
    a.ctrl = c;
    return a;
  }
  template <class T> CornerArgsT<T> corner_args_t(const FieldSet<T> &f, int wset, const Ctrl *c, int oib) const {
    CornerArgsT<T> a{};
    a.g = g;
    a.f1 = f.flx[0]; a.f2 = f.flx[1]; a.f3 = f.flx[2];
    a.e2x1 = f.ef[0]; a.e3x1 = f.ef[1]; a.e1x2 = f.ef[2]; a.e3x2 = f.ef[3]; a.e1x3 = f.ef[4]; a.e2x3 = f.ef[5];
    a.ecc = f.ecc[wset];
    a.e1 = f.ee[0]; a.e2 = f.ee[1]; a.e3 = f.ee[2];
    a.ctrl = c;
    a.only_if_bad = oib;
    a.xg = XG_CORNER;
    return a;
  }
  template <class T> UpdateArgsT<T> update_args_t(const FieldSet<T> &f, int out, double g0, double g1, int nvar,
                                                  const Ctrl *c, int oib) const {
    UpdateArgs a0 = update_args(out, g0, g1, nvar, c, oib);
    UpdateArgsT<T> a{};
    a.g = a0.g; a.g0 = g0; a.g1 = g1; a.beta = 0.0; a.beta_dt = 0.0; a.use_ctrl_dt = 0; a.nvar = nvar;
    a.f1 = f.flx[0]; a.f2 = f.flx[1]; a.f3 = f.flx[2];
    a.ua = a0.ua; a.ub = a0.ub; a.uo = a0.uo;
    a.b1a = a0.b1a; a.b2a = a0.b2a; a.b3a = a0.b3a; a.b1b = a0.b1b; a.b2b = a0.b2b; a.b3b = a0.b3b;
    a.b1o = a0.b1o; a.b2o = a0.b2o; a.b3o = a0.b3o;
    a.e1 = f.ee[0]; a.e2 = f.ee[1]; a.e3 = f.ee[2];
    a.ctrl = c;
    a.only_if_bad = oib;
    a.xg = std::is_same<T, float>::value ? XG_UPDATE_F32 : 0;
    return a;
  }
  template <class T> C2pArgsT<T> c2p_args_t(const FieldSet<T> &f, int buf, int wset, Ctrl *c) const {
    C2pArgs a0 = c2p_args(buf, wset, c);
    C2pArgsT<T> a{};
    a.g = a0.g; a.eos = a0.eos; a.b1 = a0.b1; a.b2 = a0.b2; a.b3 = a0.b3; a.u = a0.u;
    a.w = f.w[wset]; a.bcc = f.bcc[wset]; a.ecc = f.ecc[wset];
    a.write_went = 1; a.entropy_out = 1; a.compute_dt = 0; a.slot = 0; a.flag_mode = 0;
    a.flag = flag; a.badlist = badlist; a.ctrl = c; a.only_if_bad = 0;
    a.xg = XG_C2P;
    return a;
  }
};
