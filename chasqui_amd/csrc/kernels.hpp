// Kernel argument blocks and launch helpers shared by the stage kernels and the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"

namespace chq {

// A face-flux launch: one direction, faces [il..iu] x [jl..ju] x [kl..ku] (inclusive).
struct FluxArgs {
  Geom g;
  Eos eos;
  int flatten, low_mach;
  int entropy;        // reconstruct and flux the entropy variable (IENT)
  double lhlld_phi;
  int il, iu, jl, ju, kl, ku;
  const double *w;    // primitives, 6 components (IENT used only when entropy is transported)
  const double *bcc;  // cell-centred B, 3 components
  const double *bn;   // face-normal field of this direction (the stage-input face B)
  double *flx;        // fluxes, 6 components
  double *et2, *et1;  // face EMFs: et2 = -F(B_t1), et1 = +F(B_t2)
  const int *flag;    // fofc flags (repair launches only)
  const Ctrl *ctrl;
};

struct CornerArgs {
  Geom g;
  const double *f1, *f2, *f3;  // mass-flux components (IDN) of the face fluxes
  const double *e2x1, *e3x1, *e1x2, *e3x2, *e1x3, *e2x3;
  const double *ecc;
  double *e1, *e2, *e3;
  const Ctrl *ctrl;
  int only_if_bad;             // skip unless the current stage flagged cells (fofc re-run)
};

struct UpdateArgs {
  Geom g;
  double g0, g1, beta;         // beta_dt = beta * dt, dt from ctrl (or beta_dt given)
  double beta_dt;              // used when use_ctrl_dt == 0
  int use_ctrl_dt;
  int nvar;                    // 5 (entropy not transported) or 6
  const double *f1, *f2, *f3;
  const double *ua, *ub;
  double *uo;
  const double *b1a, *b2a, *b3a, *b1b, *b2b, *b3b;
  double *b1o, *b2o, *b3o;
  const double *e1, *e2, *e3;
  const Ctrl *ctrl;
  int only_if_bad;
};

struct BcArgs {
  Geom g;
  int bc, nvar;
  double *u, *b1, *b2, *b3;
  const Ctrl *ctrl;
  int only_if_bad;
};

struct C2pArgs {
  Geom g;
  Eos eos;
  const double *b1, *b2, *b3;
  double *u;
  double *w, *bcc, *ecc;
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

// launches (defined in stage.hip / flux.hip)
hipError_t launch_flux(const FluxArgs &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s);
hipError_t launch_repair(const FluxArgs &a, int dir, bool only_if_bad, hipStream_t s);
hipError_t launch_repair_list(const FluxArgs a[3], int ndir, const int *list, hipStream_t s);
hipError_t launch_corner(const CornerArgs &a, hipStream_t s);
hipError_t launch_corner_list(const CornerArgs &a, const int *list, hipStream_t s);
hipError_t launch_update(const UpdateArgs &a, hipStream_t s);
hipError_t launch_bc(const BcArgs &a, int dir, hipStream_t s);
hipError_t launch_c2p(const C2pArgs &a, hipStream_t s);
hipError_t launch_flag(const C2pArgs &a, hipStream_t s);   // standalone FlagBadCells
// control-block kernels
hipError_t launch_ctrl_begin_step(Ctrl *c, hipStream_t s);
hipError_t launch_ctrl_begin_stage(Ctrl *c, int reset_slots, hipStream_t s);
hipError_t launch_ctrl_end_stage(Ctrl *c, int last, int fused, hipStream_t s);
hipError_t launch_ctrl_end_step(Ctrl *c, hipStream_t s);

}  // namespace chq
