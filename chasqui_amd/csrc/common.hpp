// Device data model of the MI355X stage pipeline: geometry, padded HBM layout, control block.
//
// Layout. Every field (cell, face or edge centred) lives in one padded 3-D grid of
// (nk+1) planes x (nj+1) rows x `pitch` doubles, the reference's +1-in-every-dimension face
// convention (mesh.hpp:1-10) applied to cells too, so one index function serves all
// arrays. Rows are padded to a multiple of 16 doubles (128 B, one L2 line) and shifted by
// `ioff` so that the first active cell of every row starts a cache line: a wave64 load of
// 64 active doubles touches exactly four lines. Multi-component fields (u, w, fluxes) are
// component-major with a `cstride` that is a multiple of 16 elements.
#pragma once
#include <cstdint>

namespace chq {

enum : int { IDN = 0, IM1 = 1, IM2 = 2, IM3 = 3, IEN = 4, IENT = 5, NVAR = 6 };   // defs.hpp:16
enum : int { IVX = 1, IVY = 2, IVZ = 3, IPR = 4 };                               // defs.hpp:18
enum class Recon : int { dc = 0, plm = 1, ppm = 2, wenoz = 3, ppmx = 4 };         // defs.hpp:35
enum class RSolver : int { hlle = 0, hlld = 1, lhlld = 2 };                       // physics.hpp:20

struct Geom {
  int nx1, nx2, nx3, ng;
  int ni, nj, nk;
  int is, ie, js, je, ks, ke;
  int two_d, three_d;
  int pitch;        // doubles per row
  int ioff;         // element offset of i = 0 inside a row
  long long plane;  // pitch * (nj + 1)
  long long cstride;  // plane * (nk + 1): distance between components
  double dx1, dx2, dx3, idx1, idx2, idx3;
  __host__ __device__ inline long long at(int k, int j, int i) const {
    return (long long)k * plane + (long long)j * pitch + ioff + i;
  }
};

struct Eos {
  double gamma, gm1, dfloor, pfloor;
  int dual_energy;
  double de_eta;
};

// Device-resident time-loop state: lets a whole RK step run without a host round trip
// (dt, t and the fofc count never leave the GPU), which is what makes the step capturable.
struct Ctrl {
  double t, dt, tlim, cfl;
  double min_dt;                 // CFL minimum of the last completed step
  unsigned long long mdt_bits[2];  // running minima (bit patterns of positive doubles)
  int active;                    // this step runs (t < tlim and cycle - cycle0 < nlim)
  int cycle, nlim;
  int cycle0;                    // cycle the loop started from (a restart's cycle): nlim counts from it
  int nbad;                      // fofc flags of the current stage
  int nbad_last;                 // flags of the last stage (selects the repaired minimum)
  int wcur;                      // current primitive set (batched runs; solo contexts track it on the host)
  long long fofc_total;
  unsigned long long rho_bits;   // minimum active density of the last con2prim (parabolic dt)
  long long limiter_total;       // positivity flux-limiter activations
  int early;                     // this stage flags before the boundary fill
  int stage;                     // RK stage index of the time loop (-1: a single-kernel entry point)
  int fired;                     // bit s: stage s of the last step flagged cells
  int epoch;                     // stages begun since the loop started (near-marks of the local re-update)
  int zsens;                     // first-pass points the local fofc re-update cannot skip (zero sign, non-finite)
};

inline __host__ __device__ int div_up(int a, int b) { return (a + b - 1) / b; }

}  // namespace chq
