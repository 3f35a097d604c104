// Kernel argument blocks and launch helpers shared by the stage kernels and the C-ABI.
//
// T is the type of the arrays the face-flux / corner-EMF side works in: double for the FP64
// drop-in, float for the mixed-precision path (FP32 primitives, fluxes and EMFs). Conserved
// variables and face-centred B are always double.
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"

namespace chq {

// A face-flux launch: one direction, faces [il..iu] x [jl..ju] x [kl..ku] (inclusive).
template <class T>
struct FluxArgsT {
  Geom g;
  Eos eos;
  int flatten, low_mach;
  int entropy;        // reconstruct and flux the entropy variable (IENT)
  double lhlld_phi;
  int il, iu, jl, ju, kl, ku;
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
};
using C2pArgs = C2pArgsT<double>;

// launches (defined in stage.hip / flux.hip / flux32.hip)
hipError_t launch_flux(const FluxArgsT<double> &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s);
hipError_t launch_flux(const FluxArgsT<float> &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s);
hipError_t launch_repair(const FluxArgs &a, int dir, bool only_if_bad, hipStream_t s);
hipError_t launch_repair_list(const FluxArgsT<double> a[3], int ndir, const int *list, hipStream_t s);
hipError_t launch_repair_list(const FluxArgsT<float> a[3], int ndir, const int *list, hipStream_t s);
template <class T> hipError_t launch_corner(const CornerArgsT<T> &a, hipStream_t s);
template <class T> hipError_t launch_corner_list(const CornerArgsT<T> &a, const int *list, hipStream_t s);
template <class T> hipError_t launch_update(const UpdateArgsT<T> &a, hipStream_t s);
hipError_t launch_bc(const BcArgs &a, int dir, hipStream_t s);
template <class T> hipError_t launch_c2p(const C2pArgsT<T> &a, hipStream_t s);
hipError_t launch_flag(const C2pArgs &a, hipStream_t s);   // standalone FlagBadCells
// control-block kernels
hipError_t launch_ctrl_begin_step(Ctrl *c, hipStream_t s);
hipError_t launch_ctrl_begin_stage(Ctrl *c, int reset_slots, hipStream_t s);
hipError_t launch_ctrl_end_stage(Ctrl *c, int last, int fused, hipStream_t s);
hipError_t launch_ctrl_end_step(Ctrl *c, hipStream_t s);

}  // namespace chq
