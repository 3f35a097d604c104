// Memory-bound stage kernels for gfx950: upwind corner EMF, RK combine + flux divergence +
// CT update, boundaries, con2prim (+ cell-centred B/EMF, CFL minimum and, in the fused step,
// the fofc flag), the standalone fofc flag, and the device-resident time-loop control.
//
// Reference: kernels::CornerEMF / Update / Boundaries / ConsToPrim (physics.hpp:242-494),
// fofc::FlagBadCells (fofc.hpp:17-44) and the driver's dt logic (main.cpp:246-292). For
// FP64 every expression keeps the reference's operation order (bitwise parity).
//
// Launch shape: blocks of 64 x 4 threads (i fastest), one wave64 per row segment, so all
// loads and stores of a wave are contiguous 512 B segments.
#include <type_traits>

#include "kernels.hpp"
#include "numerics.hpp"

namespace chq {

namespace {

constexpr int BX = 64, BY = 4;

// only_if_bad: 0 always, 1 only when the stage flagged cells (the fofc branch), 2 only when it
// flagged none (the fused stage's first boundary fill and con2prim when the flux limiter is on),
// 3 only when the stage flags early (the flag pass before the fill), 4 unless it flagged early
// and found cells (the first fill and con2prim of the adaptive order), 5 / 6 when it flagged
// cells and its first pass counted no / some point the local re-update cannot skip
__device__ __forceinline__ bool gate(const Ctrl *c, int only_if_bad) {
  if (!c->active) return false;
  if (only_if_bad == 1 && c->nbad == 0) return false;
  if (only_if_bad == 2 && c->nbad != 0) return false;
  if (only_if_bad == 3 && !c->early) return false;
  if (only_if_bad == 4 && c->early && c->nbad != 0) return false;
  if (only_if_bad == 5 && (c->nbad == 0 || c->zsens != 0)) return false;
  if (only_if_bad == 6 && (c->nbad == 0 || c->zsens == 0)) return false;
  return true;
}

// Blocks iterate over virtual blocks (VB, kernels.hpp): a normal launch has one block per
// tile; most launches of the fofc branch, which usually exit at once, use a small
// grid-stride grid so an empty branch costs a few microseconds (con2prim keeps one tile per
// block: the re-run of a firing stage is a full pass, and a grid-stride form slows it 13%).
constexpr int COND_BLOCKS = 16384;   // grid of the conditional (fofc-branch) launches

// ---------------------------------------------------------------- corner EMF
template <class T>
__device__ __forceinline__ T upw(T flux, T left, T right) {   // physics.hpp:233-236
  const T tol = Lit<T>::upwind_tol;
  return (flux > tol) ? left : (flux < -tol) ? right : T(0.5) * (left + right);
}

// 3D: one thread per (k, j, i) of the union box [ks..ke+1] x [js..je+1] x [is..ie+1];
// e1 for i <= ie, e2 for j <= je, e3 for k <= ke (physics.hpp:265-296).
template <class T>
__device__ __forceinline__ void corner3_at(const CornerArgsT<T> &a, int k, int j, int i) {
  const Geom &g = a.g;
  const long long c = g.at(k, j, i), sj = g.pitch, sk = g.plane, cs = g.cstride;
  const T *ecc = a.ecc;
  if (i <= g.ie) {   // E1 at (k-1/2, j-1/2, i)
    const T *ex = ecc;
    const T l3 = upw(a.f2[c - sk], a.e1x3[c - sj] - ex[c - sk - sj], a.e1x3[c] - ex[c - sk]);
    const T r3 = upw(a.f2[c], a.e1x3[c - sj] - ex[c - sj], a.e1x3[c] - ex[c]);
    const T l2 = upw(a.f3[c - sj], a.e1x2[c - sk] - ex[c - sk - sj], a.e1x2[c] - ex[c - sj]);
    const T r2 = upw(a.f3[c], a.e1x2[c - sk] - ex[c - sk], a.e1x2[c] - ex[c]);
    a.e1[c] = T(0.25) * (((l3 + r3) + (l2 + r2)) + ((a.e1x2[c - sk] + a.e1x2[c]) + (a.e1x3[c - sj] + a.e1x3[c])));
  }
  if (j <= g.je) {   // E2 at (k-1/2, j, i-1/2)
    const T *ey = ecc + cs;
    const T l3 = upw(a.f1[c - sk], a.e2x3[c - 1] - ey[c - sk - 1], a.e2x3[c] - ey[c - sk]);
    const T r3 = upw(a.f1[c], a.e2x3[c - 1] - ey[c - 1], a.e2x3[c] - ey[c]);
    const T l1 = upw(a.f3[c - 1], a.e2x1[c - sk] - ey[c - sk - 1], a.e2x1[c] - ey[c - 1]);
    const T r1 = upw(a.f3[c], a.e2x1[c - sk] - ey[c - sk], a.e2x1[c] - ey[c]);
    a.e2[c] = T(0.25) * (((l3 + r3) + (l1 + r1)) + ((a.e2x3[c - 1] + a.e2x3[c]) + (a.e2x1[c - sk] + a.e2x1[c])));
  }
  if (k <= g.ke) {   // E3 at (k, j-1/2, i-1/2)
    const T *ez = ecc + 2 * cs;
    const T l2 = upw(a.f1[c - sj], a.e3x2[c - 1] - ez[c - sj - 1], a.e3x2[c] - ez[c - sj]);
    const T r2 = upw(a.f1[c], a.e3x2[c - 1] - ez[c - 1], a.e3x2[c] - ez[c]);
    const T l1 = upw(a.f2[c - 1], a.e3x1[c - sj] - ez[c - sj - 1], a.e3x1[c] - ez[c - 1]);
    const T r1 = upw(a.f2[c], a.e3x1[c - sj] - ez[c - sj], a.e3x1[c] - ez[c]);
    a.e3[c] = T(0.25) * (((l2 + r2) + (l1 + r1)) + ((a.e3x2[c - 1] + a.e3x2[c]) + (a.e3x1[c - sj] + a.e3x1[c])));
  }
}

// 2D and 1D: E1/E2 are copies of the face EMFs on the two k-planes (physics.hpp:250-256);
// E3 is the Gardiner-Stone average in 2D (288-296) or a copy in 1D (257-261).
template <class T>
__device__ __forceinline__ void corner12_at(const CornerArgsT<T> &a, int j, int i) {
  const Geom &g = a.g;
  const int k = g.ks;
  const long long c = g.at(k, j, i), sj = g.pitch, sk = g.plane, cs = g.cstride;
  a.e1[c] = a.e1x2[c];
  a.e1[c + sk] = a.e1x2[c];
  a.e2[c] = a.e2x1[c];
  a.e2[c + sk] = a.e2x1[c];
  if (g.two_d) {
    const T *ez = a.ecc + 2 * cs;
    const T l2 = upw(a.f1[c - sj], a.e3x2[c - 1] - ez[c - sj - 1], a.e3x2[c] - ez[c - sj]);
    const T r2 = upw(a.f1[c], a.e3x2[c - 1] - ez[c - 1], a.e3x2[c] - ez[c]);
    const T l1 = upw(a.f2[c - 1], a.e3x1[c - sj] - ez[c - sj - 1], a.e3x1[c] - ez[c - 1]);
    const T r1 = upw(a.f2[c], a.e3x1[c - sj] - ez[c - sj], a.e3x1[c] - ez[c]);
    a.e3[c] = T(0.25) * (((l2 + r2) + (l1 + r1)) + ((a.e3x2[c - 1] + a.e3x2[c]) + (a.e3x1[c - sj] + a.e3x1[c])));
  } else if (j == g.js) {
    a.e3[c] = a.e3x1[c];
    a.e3[c + sj] = a.e3x1[c];
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_corner(const CornerArgsT<T> a, const VB v) {
  if (!gate(a.ctrl, a.only_if_bad)) return;
  const Geom &g = a.g;
  for (int vb = blockIdx.x; vb < v.n; vb += gridDim.x) {
    int k, j, i;
    if (!v.at(vb, k, j, i)) continue;
    if (g.three_d) corner3_at(a, g.ks + k, g.js + j, g.is + i);
    else corner12_at(a, g.js + j, g.is + i);
  }
}

// Recompute the edge EMFs that can depend on a repaired face: the corners (k..k+1, j..j+1,
// i..i+1) of every listed (flagged) cell; every other edge keeps its value, which is what a
// whole-domain recompute would produce from unchanged inputs. Overlapping cells write the
// same value to a shared edge.
template <class T>
__device__ __forceinline__ void corner_list_entry(const CornerArgsT<T> &a, const int *list, int t) {
  const Geom &g = a.g;
  const long long c0 = list[t >> 3];
  const int o = t & 7;
  const long long r = c0 - g.ioff;
  const int k = int(r / g.plane), j = int((r % g.plane) / g.pitch), i = int(r % g.pitch);
  const int kk = k + ((o >> 2) & 1), jj = j + ((o >> 1) & 1), ii = i + (o & 1);
  if (g.three_d) {
    // the edge kernel's union box and per-component ranges (physics.hpp:266, 277, 288)
    if (kk > g.ke + 1 || jj > g.je + 1 || ii > g.ie + 1) return;
    corner3_at(a, kk, jj, ii);
  } else {
    if (kk != g.ks || jj > g.je + 1 || ii > g.ie + 1) return;
    if (!g.two_d && jj != g.js) return;
    corner12_at(a, jj, ii);
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_corner_list(const CornerArgsT<T> a, const int *list) {
  if (!gate(a.ctrl, 1)) return;
  const int n = a.ctrl->nbad * 8;
  for (int t = int(blockIdx.x) * 256 + int(threadIdx.x); t < n; t += int(gridDim.x) * 256)
    corner_list_entry(a, list, t);
}

// ---------------------------------------------------------------- update
// One thread per (k, j, i) of [ks..ke+1] x [js..je+1] x [is..ie+1]: the conserved update of
// the active cell (physics.hpp:314-322) and the CT update of its three lower faces
// (physics.hpp:327-347). In-place (uo aliasing ua or ub) is pointwise safe: an output aliases
// only the inputs of its own point, so every input is loaded first and every output stored
// after. With the dimensionality a template argument the loads of all variables form one
// straight run and are in flight together (per variable, behind runtime 2D / 3D tests, each
// round of loads waited for the one before).
template <class T, int DIM>
__device__ __forceinline__ void update_at_d(const UpdateArgsT<T> &a, const double bdt, int k, int j, int i) {
  constexpr bool two_d = DIM >= 2, three_d = DIM == 3;
  const Geom &g = a.g;
  const double g0 = a.g0, g1 = a.g1;
  const long long c = g.at(k, j, i), sj = g.pitch, sk = g.plane, cs = g.cstride;
  const double idx1 = g.idx1, idx2 = g.idx2, idx3 = g.idx3;
  // g0 * xa + g1 * xb; the mixed-precision path skips the read of xb when g1 == 0 (the first
  // RK stage), the FP64 path keeps the reference's expression (it fixes the sign of zeros)
  const bool skip_b = std::is_same<T, float>::value && g1 == 0.0;
  const int nvar = a.nvar;
  const bool cell = i <= g.ie && j <= g.je && k <= g.ke;
  const bool f1 = j <= g.je && k <= g.ke;                          // B1 on x1-faces
  const bool f2 = i <= g.ie && k <= g.ke && (two_d || j == g.js);  // B2 on x2-faces
  const bool f3 = i <= g.ie && j <= g.je && (three_d || k == g.ks);  // B3 on x3-faces
  // loads
  double xa[NVAR], xb[NVAR], d1[NVAR], d2[NVAR], d3[NVAR];
  if (cell) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
      if (n >= nvar) break;
      const long long o = n * cs + c;
      d1[n] = double(a.f1[o + 1]) - double(a.f1[o]);
      if (two_d) d2[n] = double(a.f2[o + sj]) - double(a.f2[o]);
      if (three_d) d3[n] = double(a.f3[o + sk]) - double(a.f3[o]);
      xa[n] = a.ua[o];
      xb[n] = skip_b ? 0.0 : a.ub[o];
    }
  }
  double e3j = 0.0, e2k = 0.0, e3i = 0.0, e1k = 0.0, e2i = 0.0, e1j = 0.0;
  double b1a = 0.0, b1b = 0.0, b2a = 0.0, b2b = 0.0, b3a = 0.0, b3b = 0.0;
  if (f1) {
    if (two_d) e3j = double(a.e3[c + sj]) - double(a.e3[c]);
    if (three_d) e2k = double(a.e2[c + sk]) - double(a.e2[c]);
    b1a = a.b1a[c];
    b1b = skip_b ? 0.0 : a.b1b[c];
  }
  if (f2) {
    e3i = double(a.e3[c + 1]) - double(a.e3[c]);
    if (three_d) e1k = double(a.e1[c + sk]) - double(a.e1[c]);
    b2a = a.b2a[c];
    b2b = skip_b ? 0.0 : a.b2b[c];
  }
  if (f3) {
    e2i = double(a.e2[c + 1]) - double(a.e2[c]);
    if (two_d) e1j = double(a.e1[c + sj]) - double(a.e1[c]);
    b3a = a.b3a[c];
    b3b = skip_b ? 0.0 : a.b3b[c];
  }
  auto comb = [&](double va, double vb) { return skip_b ? g0 * va : g0 * va + g1 * vb; };
  // the watch of UpdateArgsT::zwatch (FP64, g1 = 0): a recomputation reads this result as its
  // ub and adds 0 * (this result), which can only change a result whose g0 * ua is -0 with a
  // zero divergence (curl) term, or a result that is not finite
  bool zs = false;
  auto watch = [&](double va, double z, double r) {
    if (!std::is_same<T, double>::value || !a.zwatch) return;
    const double p = g0 * va;
    zs = zs || (p == 0.0 && signbit(p) && z == 0.0) || !isfinite(r);
  };
  // arithmetic and stores (the expressions and their order are the reference's)
  if (cell) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
      if (n >= nvar) break;
      double dv = d1[n] * idx1;
      if (two_d) dv += d2[n] * idx2;
      if (three_d) dv += d3[n] * idx3;
      const double z = bdt * dv;
      const double r = comb(xa[n], xb[n]) - z;
      a.uo[n * cs + c] = r;
      watch(xa[n], z, r);
    }
  }
  if (f1) {
    double cr = 0.0;
    if (two_d) cr += e3j * idx2;
    if (three_d) cr -= e2k * idx3;
    const double z = bdt * cr;
    const double r = comb(b1a, b1b) - z;
    a.b1o[c] = r;
    watch(b1a, z, r);
  }
  if (f2) {
    double cr = -e3i * idx1;
    if (three_d) cr += e1k * idx3;
    const double z = bdt * cr;
    const double v = comb(b2a, b2b) - z;
    a.b2o[c] = v;
    if (!two_d) a.b2o[c + sj] = v;
    watch(b2a, z, v);
  }
  if (f3) {
    double cr = e2i * idx1;
    if (two_d) cr -= e1j * idx2;
    const double z = bdt * cr;
    const double v = comb(b3a, b3b) - z;
    a.b3o[c] = v;
    if (!three_d) a.b3o[c + sk] = v;
    watch(b3a, z, v);
  }
  if (zs) atomicAdd(a.zcount, 1);
}
template <class T>
__device__ __forceinline__ void update_at(const UpdateArgsT<T> &a, const double bdt, int k, int j, int i) {
  if (a.g.three_d) update_at_d<T, 3>(a, bdt, k, j, i);
  else if (a.g.two_d) update_at_d<T, 2>(a, bdt, k, j, i);
  else update_at_d<T, 1>(a, bdt, k, j, i);
}

// The same update with the variables in a runtime loop and the 2D / 3D terms behind runtime
// tests: each variable's loads wait for the previous variable's, but the kernel needs 25 VGPRs
// (occupancy 8) instead of 120 (4). The FP64 3D update streams better this way (1.156 vs
// 1.171 ms per launch at 256^3); the 2D update, which the batched ensembles run, the other way.
template <class T>
__device__ __forceinline__ void update_at_loop(const UpdateArgsT<T> &a, const double bdt, int k, int j, int i) {
  const Geom &g = a.g;
  const double g0 = a.g0, g1 = a.g1;
  const long long c = g.at(k, j, i), sj = g.pitch, sk = g.plane, cs = g.cstride;
  const double idx1 = g.idx1, idx2 = g.idx2, idx3 = g.idx3;
  // g0 * xa + g1 * xb; the mixed-precision path skips the read of xb when g1 == 0 (the first
  // RK stage), the FP64 path keeps the reference's expression (it fixes the sign of zeros)
  const bool skip_b = std::is_same<T, float>::value && g1 == 0.0;
  auto comb = [&](const double *xa, const double *xb, long long o) {
    return skip_b ? g0 * xa[o] : g0 * xa[o] + g1 * xb[o];
  };
  // the watch of UpdateArgsT::zwatch (FP64, g1 = 0): a recomputation reads this result as its
  // ub and adds 0 * (this result), which can only change a result whose g0 * ua is -0 with a
  // zero divergence (curl) term, or a result that is not finite
  bool zs = false;
  auto watch = [&](const double *xa, long long o, double z, double r) {
    if (!std::is_same<T, double>::value || !a.zwatch) return;
    const double p = g0 * xa[o];
    zs = zs || (p == 0.0 && signbit(p) && z == 0.0) || !isfinite(r);
  };
  if (i <= g.ie && j <= g.je && k <= g.ke) {
    for (int n = 0; n < a.nvar; ++n) {
      const long long o = n * cs + c;
      double dv = (double(a.f1[o + 1]) - double(a.f1[o])) * idx1;
      if (g.two_d) dv += (double(a.f2[o + sj]) - double(a.f2[o])) * idx2;
      if (g.three_d) dv += (double(a.f3[o + sk]) - double(a.f3[o])) * idx3;
      const double z = bdt * dv;
      const double r = comb(a.ua, a.ub, o) - z;
      a.uo[o] = r;
      watch(a.ua, o, z, r);
    }
  }
  if (j <= g.je && k <= g.ke) {   // B1 on x1-faces
    double cr = 0.0;
    if (g.two_d) cr += (double(a.e3[c + sj]) - double(a.e3[c])) * idx2;
    if (g.three_d) cr -= (double(a.e2[c + sk]) - double(a.e2[c])) * idx3;
    const double z = bdt * cr;
    const double r = comb(a.b1a, a.b1b, c) - z;
    a.b1o[c] = r;
    watch(a.b1a, c, z, r);
  }
  if (i <= g.ie && k <= g.ke && (g.two_d || j == g.js)) {   // B2 on x2-faces
    double cr = -(double(a.e3[c + 1]) - double(a.e3[c])) * idx1;
    if (g.three_d) cr += (double(a.e1[c + sk]) - double(a.e1[c])) * idx3;
    const double z = bdt * cr;
    const double v = comb(a.b2a, a.b2b, c) - z;
    a.b2o[c] = v;
    if (!g.two_d) a.b2o[c + sj] = v;
    watch(a.b2a, c, z, v);
  }
  if (i <= g.ie && j <= g.je && (g.three_d || k == g.ks)) {   // B3 on x3-faces
    double cr = (double(a.e2[c + 1]) - double(a.e2[c])) * idx1;
    if (g.two_d) cr -= (double(a.e1[c + sj]) - double(a.e1[c])) * idx2;
    const double z = bdt * cr;
    const double v = comb(a.b3a, a.b3b, c) - z;
    a.b3o[c] = v;
    if (!g.three_d) a.b3o[c + sk] = v;
    watch(a.b3a, c, z, v);
  }
  if (zs) atomicAdd(a.zcount, 1);
}


// The local fofc re-update (main.cpp:86-88) of an RK stage u0 -> u1 with g1 = 0. The
// reference re-updates the whole domain from the repaired fluxes; its aliased input u1 (b?f1)
// is this stage's first-pass output and enters as 0 * u1, so away from the repaired faces and
// the edges around them the recomputation reproduces the first pass bit for bit, except where
// the first pass's watch counted a point (then the whole domain is recomputed instead). The
// repair changes the fluxes of the flagged cells' faces and the EMFs of their edges, and only
// the points of the update box within one cell of a flagged cell read those, so:
//   k_mark_near   - stamps those points (3x3x3 around each listed cell) with the stage's epoch;
//                   several cells stamp a shared point with the same value
//   k_update_near - one thread per point of the update box recomputes the stamped ones, each
//                   once, reading the first-pass value of the aliased buffer, as the reference
// Stamps of earlier stages hold smaller epochs and are ignored.
__global__ void __launch_bounds__(256) k_mark_near(const Geom g, const Ctrl *ctl, const int *list, int *marks) {
  if (!gate(ctl, 5)) return;
  const int epoch = ctl->epoch;
  const int n = ctl->nbad * 27;
  const int ku = g.three_d ? g.ke + 1 : g.ks, ju = g.two_d ? g.je + 1 : g.js;
  for (int t = int(blockIdx.x) * 256 + int(threadIdx.x); t < n; t += int(gridDim.x) * 256) {
    const long long r = (long long)list[t / 27] - g.ioff;   // k * plane + j * pitch + i
    const int o = t % 27;
    const int k = int(r / g.plane) + o / 9 - 1;
    const int j = int((r % g.plane) / g.pitch) + (o / 3) % 3 - 1;
    const int i = int(r % g.pitch) + o % 3 - 1;
    if (k < g.ks || k > ku || j < g.js || j > ju || i < g.is || i > g.ie + 1) continue;
    marks[g.at(k, j, i)] = epoch;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_update_near(const UpdateArgsT<T> a, const VB v, const int *marks) {
  if (!gate(a.ctrl, 5)) return;
  const Geom &g = a.g;
  const double bdt = a.use_ctrl_dt ? a.beta * a.ctrl->dt : a.beta_dt;
  const int epoch = a.ctrl->epoch;
  for (int vb = blockIdx.x; vb < v.n; vb += gridDim.x) {
    int k, j, i;
    if (!v.at(vb, k, j, i)) continue;
    k += g.ks; j += g.js; i += g.is;
    if (marks[g.at(k, j, i)] == epoch) update_at_loop(a, bdt, k, j, i);
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_update(const UpdateArgsT<T> a, const VB v) {
  if (!gate(a.ctrl, a.only_if_bad)) return;
  const Geom &g = a.g;
  const double bdt = a.use_ctrl_dt ? a.beta * a.ctrl->dt : a.beta_dt;
  const int grp = int(gridDim.x) == v.n ? a.xg : 0;   // one tile per block: XCD-grouped order
  for (int vb = blockIdx.x; vb < v.n; vb += gridDim.x) {
    int k, j, i;
    if (!v.at(xcd_tile(vb, v.n, grp), k, j, i)) continue;
    update_at_loop(a, bdt, g.ks + k, g.js + j, g.is + i);
  }
}
// 1D and 2D meshes: the loads-first form
template <class T>
__global__ void __launch_bounds__(256) k_update_lf(const UpdateArgsT<T> a, const VB v) {
  if (!gate(a.ctrl, a.only_if_bad)) return;
  const Geom &g = a.g;
  const double bdt = a.use_ctrl_dt ? a.beta * a.ctrl->dt : a.beta_dt;
  const int grp = int(gridDim.x) == v.n ? a.xg : 0;
  for (int vb = blockIdx.x; vb < v.n; vb += gridDim.x) {
    int k, j, i;
    if (!v.at(xcd_tile(vb, v.n, grp), k, j, i)) continue;
    update_at(a, bdt, g.ks + k, g.js + j, g.is + i);
  }
}


// ---------------------------------------------------------------- 3D k-marching forms
// 3D corner EMF and update with the (j, i) plane in 256-point chunks and each thread marching
// a few planes in k: the k-1 (corner) or k+1 (update) values a stencil needs were loaded by the
// same thread in the previous iteration, so every input plane is read once per segment instead
// of being fetched again by the blocks of the next plane. Same expressions, same order as
// corner3_at / update_at.
// planes per thread, measured at 256^3 (corner: 4 beat 8, 16, 32, 64 and tied with 2 in launch
// order; with the XCD-grouped order 2 is 1% faster than 4 and 1.3% faster than 1; mixed update: 2
// beats 3, 4, 8, 16 and ties with 1)
constexpr int KSEG_CORNER = 2, KSEG_UPDATE = 2;

template <class T>
__global__ void __launch_bounds__(256) k_corner_march(const CornerArgsT<T> a, const VB v, const int nk, const int kseg) {
  if (!gate(a.ctrl, a.only_if_bad)) return;
  const Geom &g = a.g;
  const int nb = int(gridDim.x * gridDim.y);
  const int t = xcd_tile(int(blockIdx.y * gridDim.x + blockIdx.x), nb, a.xg);
  const int bx = t % int(gridDim.x), by = t / int(gridDim.x);
  int kk, jj, ii;
  if (!v.at(bx, kk, jj, ii)) return;   // v spans one plane
  const int i = g.is + ii, j = g.js + jj;
  const int k0 = g.ks + by * kseg;
  const int k1 = min(k0 + kseg - 1, g.ks + nk - 1);
  const long long sj = g.pitch, cs = g.cstride;
  const T *ex = a.ecc, *ey = a.ecc + cs, *ez = a.ecc + 2 * cs;
  const bool do1 = i <= g.ie, do2 = j <= g.je;
  // plane k-1 values carried through the march
  T p_f2 = T(0), p_e1x2 = T(0), p_exm = T(0), p_ex = T(0);   // E1: f2, e1x2, ex(j-1), ex
  T p_f1 = T(0), p_e2x1 = T(0), p_eym = T(0), p_ey = T(0);   // E2: f1, e2x1, ey(i-1), ey
  {
    const long long c = g.at(k0 - 1, j, i);
    if (do1) { p_f2 = a.f2[c]; p_e1x2 = a.e1x2[c]; p_exm = ex[c - sj]; p_ex = ex[c]; }
    if (do2) { p_f1 = a.f1[c]; p_e2x1 = a.e2x1[c]; p_eym = ey[c - 1]; p_ey = ey[c]; }
  }
  for (int k = k0; k <= k1; ++k) {
    const long long c = g.at(k, j, i);
    if (do1) {   // E1 at (k-1/2, j-1/2, i)
      const T f2 = a.f2[c], e1x2 = a.e1x2[c], e1x3m = a.e1x3[c - sj], e1x3 = a.e1x3[c];
      const T exm = ex[c - sj], exc = ex[c], f3m = a.f3[c - sj], f3 = a.f3[c];
      const T l3 = upw(p_f2, e1x3m - p_exm, e1x3 - p_ex);
      const T r3 = upw(f2, e1x3m - exm, e1x3 - exc);
      const T l2 = upw(f3m, p_e1x2 - p_exm, e1x2 - exm);
      const T r2 = upw(f3, p_e1x2 - p_ex, e1x2 - exc);
      a.e1[c] = T(0.25) * (((l3 + r3) + (l2 + r2)) + ((p_e1x2 + e1x2) + (e1x3m + e1x3)));
      p_f2 = f2; p_e1x2 = e1x2; p_exm = exm; p_ex = exc;
    }
    if (do2) {   // E2 at (k-1/2, j, i-1/2)
      const T f1 = a.f1[c], e2x1 = a.e2x1[c], e2x3m = a.e2x3[c - 1], e2x3 = a.e2x3[c];
      const T eym = ey[c - 1], eyc = ey[c], f3m = a.f3[c - 1], f3 = a.f3[c];
      const T l3 = upw(p_f1, e2x3m - p_eym, e2x3 - p_ey);
      const T r3 = upw(f1, e2x3m - eym, e2x3 - eyc);
      const T l1 = upw(f3m, p_e2x1 - p_eym, e2x1 - eym);
      const T r1 = upw(f3, p_e2x1 - p_ey, e2x1 - eyc);
      a.e2[c] = T(0.25) * (((l3 + r3) + (l1 + r1)) + ((e2x3m + e2x3) + (p_e2x1 + e2x1)));
      p_f1 = f1; p_e2x1 = e2x1; p_eym = eym; p_ey = eyc;
    }
    if (k <= g.ke) {   // E3 at (k, j-1/2, i-1/2): this plane only
      const T l2 = upw(a.f1[c - sj], a.e3x2[c - 1] - ez[c - sj - 1], a.e3x2[c] - ez[c - sj]);
      const T r2 = upw(a.f1[c], a.e3x2[c - 1] - ez[c - 1], a.e3x2[c] - ez[c]);
      const T l1 = upw(a.f2[c - 1], a.e3x1[c - sj] - ez[c - sj - 1], a.e3x1[c] - ez[c - 1]);
      const T r1 = upw(a.f2[c], a.e3x1[c - sj] - ez[c - sj], a.e3x1[c] - ez[c]);
      a.e3[c] = T(0.25) * (((l2 + r2) + (l1 + r1)) + ((a.e3x2[c - 1] + a.e3x2[c]) + (a.e3x1[c - sj] + a.e3x1[c])));
    }
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_update_march(const UpdateArgsT<T> a, const VB v, const int nk, const int kseg) {
  if (!gate(a.ctrl, a.only_if_bad)) return;
  const Geom &g = a.g;
  const int nb = int(gridDim.x * gridDim.y);
  const int t = xcd_tile(int(blockIdx.y * gridDim.x + blockIdx.x), nb, a.xg);
  const int bx = t % int(gridDim.x), by = t / int(gridDim.x);
  int kk, jj, ii;
  if (!v.at(bx, kk, jj, ii)) return;
  const int i = g.is + ii, j = g.js + jj;
  const int k0 = g.ks + by * kseg;
  const int k1 = min(k0 + kseg - 1, g.ks + nk - 1);
  const double bdt = a.use_ctrl_dt ? a.beta * a.ctrl->dt : a.beta_dt;
  const double g0 = a.g0, g1 = a.g1;
  const long long sj = g.pitch, sk = g.plane, cs = g.cstride;
  const double idx1 = g.idx1, idx2 = g.idx2, idx3 = g.idx3;
  const bool skip_b = std::is_same<T, float>::value && g1 == 0.0;
  auto comb = [&](const double *xa, const double *xb, long long o) {
    return skip_b ? g0 * xa[o] : g0 * xa[o] + g1 * xb[o];
  };
  const bool cell_ij = i <= g.ie && j <= g.je;
  // plane k values of f3 and of the edge EMFs e1, e2, loaded as the previous iteration's k+1
  T f3c[NVAR];
  T e1c = T(0), e2c = T(0);
  {
    const long long c = g.at(k0, j, i);
    if (cell_ij && k0 <= g.ke)
      for (int n = 0; n < a.nvar; ++n) f3c[n] = a.f3[n * cs + c];
    e1c = a.e1[c];
    e2c = a.e2[c];
  }
  for (int k = k0; k <= k1; ++k) {
    const long long c = g.at(k, j, i);
    T e1n = T(0), e2n = T(0);
    if (k <= g.ke) { e1n = a.e1[c + sk]; e2n = a.e2[c + sk]; }
    if (cell_ij && k <= g.ke) {
      for (int n = 0; n < a.nvar; ++n) {
        const long long o = n * cs + c;
        const T f3n = a.f3[o + sk];
        double dv = (double(a.f1[o + 1]) - double(a.f1[o])) * idx1;
        dv += (double(a.f2[o + sj]) - double(a.f2[o])) * idx2;
        dv += (double(f3n) - double(f3c[n])) * idx3;
        a.uo[o] = comb(a.ua, a.ub, o) - bdt * dv;
        f3c[n] = f3n;
      }
    }
    if (j <= g.je && k <= g.ke) {   // B1 on x1-faces
      double cr = 0.0;
      cr += (double(a.e3[c + sj]) - double(a.e3[c])) * idx2;
      cr -= (double(e2n) - double(e2c)) * idx3;
      a.b1o[c] = comb(a.b1a, a.b1b, c) - bdt * cr;
    }
    if (i <= g.ie && k <= g.ke) {   // B2 on x2-faces
      double cr = -(double(a.e3[c + 1]) - double(a.e3[c])) * idx1;
      cr += (double(e1n) - double(e1c)) * idx3;
      a.b2o[c] = comb(a.b2a, a.b2b, c) - bdt * cr;
    }
    if (i <= g.ie && j <= g.je) {   // B3 on x3-faces
      double cr = (double(a.e2[c + 1]) - double(e2c)) * idx1;
      cr -= (double(a.e1[c + sj]) - double(e1c)) * idx2;
      a.b3o[c] = comb(a.b3a, a.b3b, c) - bdt * cr;
    }
    e1c = e1n;
    e2c = e2n;
  }
}

// ---------------------------------------------------------------- boundaries
// physics.hpp:354-432. x1: threads over (k, j, g) with g fastest; x2: (k, g, i); x3: (g, j, i).
__device__ __forceinline__ void bc1_at(const BcArgs &a, const long long t) {
  const Geom &g = a.g;
  const int ng = g.ng;
  const long long n_all = (long long)(g.nk + 1) * (g.nj + 1) * ng;
  if (t >= n_all) return;
  const int gg = int(t % ng), j = int((t / ng) % (g.nj + 1)), k = int(t / ((long long)ng * (g.nj + 1)));
  const int bc = a.bc, nx = g.nx1, is = g.is, ie = g.ie;
  const int il = gg, ir = ie + 1 + gg;
  const int sl = bc == 0 ? il + nx : bc == 1 ? is : 2 * is - 1 - il;
  const int sr = bc == 0 ? ir - nx : bc == 1 ? ie : 2 * ie + 1 - ir;
  const double sg = bc == 2 ? -1.0 : 1.0;
  const long long row = g.at(k, j, 0), cs = g.cstride;
  // every load of this thread first, then every store: none of its stores feeds one of its
  // loads, and its two ghosts of each row lie in lines of their own, so the loads overlap
  const bool cells = k < g.nk && j < g.nj;
  double vl[NVAR], vr[NVAR], b2l = 0.0, b2r = 0.0, b3l = 0.0, b3r = 0.0;
  if (cells) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n)
      if (n < a.nvar) { vl[n] = a.u[n * cs + row + sl]; vr[n] = a.u[n * cs + row + sr]; }
    b2l = a.b2[row + sl]; b2r = a.b2[row + sr];
    b3l = a.b3[row + sl]; b3r = a.b3[row + sr];
  }
  double *b1 = a.b1 + row;
  const int fl = gg, fr = ie + 2 + gg;
  double f1l, f1r;
  if (bc == 0) { f1l = b1[fl + nx]; f1r = b1[fr - nx]; }
  else if (bc == 1) { f1l = b1[is]; f1r = b1[ie + 1]; }
  else { f1l = -b1[2 * is - fl]; f1r = -b1[2 * (ie + 1) - fr]; }
  const bool shared = bc == 0 && gg == 0;   // the periodic shared face takes the lower copy
  const double f1s = shared ? b1[is] : 0.0;
  if (cells) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n)
      if (n < a.nvar) {
        const double f = n == IM1 ? sg : 1.0;
        a.u[n * cs + row + il] = f * vl[n];
        a.u[n * cs + row + ir] = f * vr[n];
      }
    a.b2[row + il] = b2l; a.b2[row + ir] = b2r;
    a.b3[row + il] = b3l; a.b3[row + ir] = b3r;
  }
  b1[fl] = f1l;
  b1[fr] = f1r;
  if (shared) b1[ie + 1] = f1s;
}

__device__ __forceinline__ void bc2_at(const BcArgs &a, const long long t) {
  const Geom &g = a.g;
  const int ng = g.ng;
  const long long n_all = (long long)(g.nk + 1) * ng * (g.ni + 1);
  if (t >= n_all) return;
  const int i = int(t % (g.ni + 1)), gg = int((t / (g.ni + 1)) % ng), k = int(t / ((long long)ng * (g.ni + 1)));
  const int bc = a.bc, ny = g.nx2, js = g.js, je = g.je;
  const int jl = gg, jr = je + 1 + gg;
  const int sl = bc == 0 ? jl + ny : bc == 1 ? js : 2 * js - 1 - jl;
  const int sr = bc == 0 ? jr - ny : bc == 1 ? je : 2 * je + 1 - jr;
  const double sg = bc == 2 ? -1.0 : 1.0;
  const long long sj = g.pitch, cs = g.cstride;
  const long long col = g.at(k, 0, i);
  // loads first, then stores (as in bc1_at)
  const bool cells = k < g.nk && i < g.ni, faces1 = k < g.nk;
  double vl[NVAR], vr[NVAR], b3l = 0.0, b3r = 0.0, b1l = 0.0, b1r = 0.0;
  if (cells) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n)
      if (n < a.nvar) { vl[n] = a.u[n * cs + col + sl * sj]; vr[n] = a.u[n * cs + col + sr * sj]; }
    b3l = a.b3[col + sl * sj]; b3r = a.b3[col + sr * sj];
  }
  if (faces1) { b1l = a.b1[col + sl * sj]; b1r = a.b1[col + sr * sj]; }
  double *b2 = a.b2 + col;
  const int fl = gg, fr = je + 2 + gg;
  double f2l, f2r;
  if (bc == 0) { f2l = b2[(fl + ny) * sj]; f2r = b2[(fr - ny) * sj]; }
  else if (bc == 1) { f2l = b2[js * sj]; f2r = b2[(je + 1) * sj]; }
  else { f2l = -b2[(2 * js - fl) * sj]; f2r = -b2[(2 * (je + 1) - fr) * sj]; }
  const bool shared = bc == 0 && gg == 0;
  const double f2s = shared ? b2[js * sj] : 0.0;
  if (cells) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n)
      if (n < a.nvar) {
        const double f = n == IM2 ? sg : 1.0;
        a.u[n * cs + col + jl * sj] = f * vl[n];
        a.u[n * cs + col + jr * sj] = f * vr[n];
      }
    a.b3[col + jl * sj] = b3l; a.b3[col + jr * sj] = b3r;
  }
  if (faces1) { a.b1[col + jl * sj] = b1l; a.b1[col + jr * sj] = b1r; }
  b2[fl * sj] = f2l;
  b2[fr * sj] = f2r;
  if (shared) b2[(je + 1) * sj] = f2s;
}

__device__ __forceinline__ void bc3_at(const BcArgs &a, const long long t) {
  const Geom &g = a.g;
  const int ng = g.ng;
  const long long n_all = (long long)ng * (g.nj + 1) * (g.ni + 1);
  if (t >= n_all) return;
  const int i = int(t % (g.ni + 1)), j = int((t / (g.ni + 1)) % (g.nj + 1)), gg = int(t / ((long long)(g.nj + 1) * (g.ni + 1)));
  const int bc = a.bc, nz = g.nx3, ks = g.ks, ke = g.ke;
  const int kl = gg, kr = ke + 1 + gg;
  const int sl = bc == 0 ? kl + nz : bc == 1 ? ks : 2 * ks - 1 - kl;
  const int sr = bc == 0 ? kr - nz : bc == 1 ? ke : 2 * ke + 1 - kr;
  const double sg = bc == 2 ? -1.0 : 1.0;
  const long long sk = g.plane, cs = g.cstride;
  const long long col = g.at(0, j, i);
  // loads first, then stores (as in bc1_at)
  const bool cells = j < g.nj && i < g.ni;
  double vl[NVAR], vr[NVAR];
  if (cells) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n)
      if (n < a.nvar) { vl[n] = a.u[n * cs + col + sl * sk]; vr[n] = a.u[n * cs + col + sr * sk]; }
  }
  const double b1l = a.b1[col + sl * sk], b1r = a.b1[col + sr * sk];
  const double b2l = a.b2[col + sl * sk], b2r = a.b2[col + sr * sk];
  double *b3 = a.b3 + col;
  const int fl = gg, fr = ke + 2 + gg;
  double f3l, f3r;
  if (bc == 0) { f3l = b3[(fl + nz) * sk]; f3r = b3[(fr - nz) * sk]; }
  else if (bc == 1) { f3l = b3[ks * sk]; f3r = b3[(ke + 1) * sk]; }
  else { f3l = -b3[(2 * ks - fl) * sk]; f3r = -b3[(2 * (ke + 1) - fr) * sk]; }
  const bool shared = bc == 0 && gg == 0;
  const double f3s = shared ? b3[ks * sk] : 0.0;
  if (cells) {
#pragma unroll
    for (int n = 0; n < NVAR; ++n)
      if (n < a.nvar) {
        const double f = n == IM3 ? sg : 1.0;
        a.u[n * cs + col + kl * sk] = f * vl[n];
        a.u[n * cs + col + kr * sk] = f * vr[n];
      }
  }
  a.b1[col + kl * sk] = b1l; a.b1[col + kr * sk] = b1r;
  a.b2[col + kl * sk] = b2l; a.b2[col + kr * sk] = b2r;
  b3[fl * sk] = f3l;
  b3[fr * sk] = f3r;
  if (shared) b3[(ke + 1) * sk] = f3s;
}

template <int DIR>
__device__ __forceinline__ void bc_at(const BcArgs &a, const long long t) {
  if (DIR == 0) bc1_at(a, t);
  else if (DIR == 1) bc2_at(a, t);
  else bc3_at(a, t);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_bc(const BcArgs a) {
  if (!gate(a.ctrl, a.only_if_bad)) return;
  bc_at<DIR>(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// cons -> prim over all cells, physics.hpp:437-494, with the CFL minimum reduced by wave64
// shuffles, one LDS pass and one 64-bit atomic per block. flag_mode 1 adds the fofc flag of
// the (tentative) active cells with the expressions of fofc.hpp:24-40.
// flag_mode 1 writes the flag array (reference semantics); 2 appends flagged cells to a list.
// DT: reduce the CFL minimum. ENT: resynchronise u[IENT] (and w[IENT] if write_went); the
// lean variant (no pow) serves RK stages whose u[IENT] nothing reads.
// bz: the cell-centred x3 field (the caller averages the two x3 faces). DE: the dual
// energy branches are compiled in (without them the kernel holds no pow() for the flag and
// needs far fewer registers).
// IENT_ONLY: store u[IENT] and nothing else (the deferred resynchronisation, k_ient_sync).
template <class T, bool DT, bool ENT, bool DE, bool IENT_ONLY = false>
__device__ __forceinline__ void c2p_core(const C2pArgsT<T> &a, int k, int j, int i, const double bz, double &mdt,
                                         int &nbad_local) {
  const Geom &g = a.g;
  const Eos e = a.eos;
  const long long c = g.at(k, j, i), sj = g.pitch, cs = g.cstride;
  const double bx = 0.5 * (a.b1[c] + a.b1[c + 1]);
  const double by = g.two_d ? 0.5 * (a.b2[c] + a.b2[c + sj]) : a.b2[c];
  double *u = a.u + c;
  const double u0 = u[IDN * cs], u1 = u[IM1 * cs], u2 = u[IM2 * cs], u3 = u[IM3 * cs], u4 = u[IEN * cs];
  const bool active = i >= g.is && i <= g.ie && j >= g.js && j <= g.je && k >= g.ks && k <= g.ke;
  // flag_mode 3: append (as 2) unless this stage already flagged before the fill
  const int fm = IENT_ONLY ? 0 : a.flag_mode == 3 ? (a.ctrl->early ? 0 : 2) : a.flag_mode;
  if (fm && active) {   // fofc.hpp:24-40 on the tentative state
    bool bad = u0 <= e.dfloor;
    if (!bad) {
      const double kin = 0.5 * (sq(u1) + sq(u2) + sq(u3)) / u0;
      const double ei = u4 - kin - 0.5 * (bx * bx + by * by + bz * bz);
      double p = e.gm1 * ei;
      if (DE && e.dual_energy && ei < e.de_eta * fabs(u4)) p = (u[IENT * cs] / u0) * pow(u0, e.gamma);
      bad = p <= e.pfloor;
    }
    if (fm == 1) {
      a.flag[c] = bad ? 1 : 0;
      nbad_local += bad ? 1 : 0;
    } else if (bad) {
      const int pos = atomicAdd(&a.ctrl->nbad, 1);
      a.badlist[pos] = int(c);
    }
  }
  const double d = fmax(u0, e.dfloor);
  const double id = 1.0 / d;
  const double vx = u1 * id, vy = u2 * id, vz = u3 * id;
  const double bsq = bx * bx + by * by + bz * bz;
  const double kin = 0.5 * d * (vx * vx + vy * vy + vz * vz);
  const double ei = u4 - kin - 0.5 * bsq;
  double p = e.gm1 * ei;
  double se = 0.0;
  if (ENT && DE && e.dual_energy && ei < e.de_eta * fabs(u4)) {
    se = u[IENT * cs] * id;
    p = se * pow(d, e.gamma);
    p = fmax(p, e.pfloor);
    u[IEN * cs] = kin + 0.5 * bsq + p / e.gm1;
  } else {
    p = fmax(p, e.pfloor);
    if constexpr (ENT) {
      se = p / pow(d, e.gamma);
      u[IENT * cs] = d * se;
    }
  }
  if constexpr (IENT_ONLY) return;
  T *w = a.w + c;
  w[IDN * cs] = T(d);
  w[IVX * cs] = T(vx);
  w[IVY * cs] = T(vy);
  w[IVZ * cs] = T(vz);
  w[IPR * cs] = T(p);
  if (ENT && a.write_went) w[IENT * cs] = T(se);
  a.bcc[c] = T(bx);
  a.bcc[c + cs] = T(by);
  a.bcc[c + 2 * cs] = T(bz);
  a.ecc[c] = T(vz * by - vy * bz);
  a.ecc[c + cs] = T(vx * bz - vz * bx);
  a.ecc[c + 2 * cs] = T(vy * bx - vx * by);
  if (DT && active) {
    const double gp = e.gamma * p;
    const double aa = (gp + bsq) * id;
    auto cf = [&](double b) {
      const double disc = aa * aa - 4.0 * gp * b * b * id * id;
      return sqrt(0.5 * (aa + sqrt(fmax(disc, 0.0))));
    };
    double dt = g.dx1 / (fabs(vx) + cf(bx));
    if (g.two_d) dt = fmin(dt, g.dx2 / (fabs(vy) + cf(by)));
    if (g.three_d) dt = fmin(dt, g.dx3 / (fabs(vz) + cf(bz)));
    mdt = fmin(mdt, dt);
  }
}

template <class T, bool DT, bool ENT, bool DE, bool IENT_ONLY = false>
__device__ __forceinline__ void c2p_at(const C2pArgsT<T> &a, int k, int j, int i, double &mdt, int &nbad_local) {
  const Geom &g = a.g;
  const long long c = g.at(k, j, i);
  const double bz = g.three_d ? 0.5 * (a.b3[c] + a.b3[c + g.plane]) : a.b3[c];
  c2p_core<T, DT, ENT, DE, IENT_ONLY>(a, k, j, i, bz, mdt, nbad_local);
}

// The deferred resynchronisation of u[IENT] = rho p / rho^gamma over every cell, with the
// expressions of con2prim: the fast path's last RK stage leaves it to the first host read of
// the state (u[IENT] is read by nothing in that path when entropy is not transported), and this
// computes the value that con2prim would have stored from the same u and face B.
__global__ void __launch_bounds__(256) k_ient_sync(const C2pArgsT<double> a, const VB v) {
  double mdt = 1e300;
  int nb = 0;
  int k, j, i;
  if (v.at(int(blockIdx.x), k, j, i)) c2p_at<double, false, true, false, true>(a, k, j, i, mdt, nb);
}

// Block merge of the CFL minimum and the flag count into the control block (the block's
// minimum goes out only when it beats the current one). Every thread of the block calls it.
template <bool DT, class T>
__device__ __forceinline__ void c2p_block_merge(const C2pArgsT<T> &a, double mdt, int nbad_local, double *smin, int *ssum) {
  const int lane = int(threadIdx.x) & 63, wid = int(threadIdx.x) >> 6;
  const double wm = wave_min(mdt);
  const int ws = wave_sum(nbad_local);
  if (lane == 0) { smin[wid] = wm; ssum[wid] = ws; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = smin[0];
    int s = ssum[0];
    for (int q = 1; q < BX * BY / 64; ++q) { m = fmin(m, smin[q]); s += ssum[q]; }
    if (DT && m < 1e300) {
      const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
      if (bits < __atomic_load_n(&a.ctrl->mdt_bits[a.slot], __ATOMIC_RELAXED)) atomicMin(&a.ctrl->mdt_bits[a.slot], bits);
    }
    if (a.flag_mode == 1 && s) atomicAdd(&a.ctrl->nbad, s);
  }
  __syncthreads();   // smin / ssum are reused by the next tile of a grid-stride loop
}

// cons -> prim over all cells, physics.hpp:437-494. The CFL minimum is reduced with wave64
// shuffles and one LDS pass per block, then merged with a 64-bit atomicMin on the bit
// pattern, issued only when the block's minimum beats the current one.
// One tile per block, also for the fofc-branch re-run (gate 1): a grid-stride form with a capped
// grid exits sooner when the gate is closed (about 5 vs 29 us at 256^3) but ran the re-run of a
// firing stage 13% slower (749 vs 660 us: blocks of four or five tiles leave a long tail), and
// in the benchmark stage 1 fires every step. One tile per block also keeps the loop-invariant
// constants of pow() out of the registers of the whole kernel.
template <class T, bool DT, bool ENT, bool DE>
__global__ void __launch_bounds__(256) k_c2p(const C2pArgsT<T> a, const VB v) {
  if (!gate(a.ctrl, a.only_if_bad)) return;
  double mdt = 1e300;
  int nbad_local = 0;
  int k, j, i;
  if (v.at(xcd_tile(int(blockIdx.x), int(gridDim.x), a.xg), k, j, i)) c2p_at<T, DT, ENT, DE>(a, k, j, i, mdt, nbad_local);
  if (!DT && a.flag_mode != 1) return;
  __shared__ double smin[BX * BY / 64];
  __shared__ int ssum[BX * BY / 64];
  c2p_block_merge<DT>(a, mdt, nbad_local, smin, ssum);
}

// standalone FlagBadCells over the active cells (fofc.hpp:17-44); count into ctrl->nbad, with
// the flag array (flag_mode 1 or 0) or appending to the flagged-cell list (flag_mode 2)
__global__ void __launch_bounds__(BX * BY) k_flag(const C2pArgs a) {
  if (!gate(a.ctrl, a.only_if_bad)) return;
  const Geom &g = a.g;
  const int i = g.is + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = g.js + int(blockIdx.y) * BY + int(threadIdx.y);
  int bad_count = 0;
  // planes ks + blockIdx.z, + gridDim.z, ...: the grid is capped (launch_flag), so a launch
  // whose gate is closed costs a few microseconds
  for (int k = g.ks + int(blockIdx.z); k <= g.ke && i <= g.ie && j <= g.je; k += int(gridDim.z)) {
    const Eos e = a.eos;
    const long long c = g.at(k, j, i), sj = g.pitch, sk = g.plane, cs = g.cstride;
    const double bx = 0.5 * (a.b1[c] + a.b1[c + 1]);
    const double by = g.two_d ? 0.5 * (a.b2[c] + a.b2[c + sj]) : a.b2[c];
    const double bz = g.three_d ? 0.5 * (a.b3[c] + a.b3[c + sk]) : a.b3[c];
    const double *u = a.u + c;
    const double d = u[IDN * cs];
    bool bad = d <= e.dfloor;
    if (!bad) {
      const double kin = 0.5 * (sq(u[IM1 * cs]) + sq(u[IM2 * cs]) + sq(u[IM3 * cs])) / d;
      const double ei = u[IEN * cs] - kin - 0.5 * (bx * bx + by * by + bz * bz);
      double p = e.gm1 * ei;
      if (e.dual_energy && ei < e.de_eta * fabs(u[IEN * cs])) p = (u[IENT * cs] / d) * pow(d, e.gamma);
      bad = p <= e.pfloor;
    }
    if (a.flag_mode == 2) {   // append to the flagged-cell list (the fused stage's repair reads it)
      if (bad) a.badlist[atomicAdd(&a.ctrl->nbad, 1)] = int(c);
    } else {
      a.flag[c] = bad ? 1 : 0;
      bad_count += bad;
    }
  }
  if (a.flag_mode == 2) return;
  const int ws = wave_sum(bad_count);
  if ((int(threadIdx.x) & 63) == 0 && ws) atomicAdd(&a.ctrl->nbad, ws);
}

// ---------------------------------------------------------------- control
__device__ __forceinline__ unsigned long long bits_of(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double from_bits(unsigned long long b) { return __longlong_as_double((long long)b); }

// main.cpp:243-248 (nlim counts the cycles run since the loop started, main.cpp:243)
__device__ __forceinline__ bool ctrl_running(const Ctrl *c) {
  return c->t < c->tlim && (c->nlim < 0 || c->cycle - c->cycle0 < c->nlim);
}
__device__ __forceinline__ void ctrl_begin_step(Ctrl *c) {
  if (!ctrl_running(c)) { c->active = 0; return; }
  c->active = 1;
  double dt = c->cfl * c->min_dt;
  if (c->t + dt > c->tlim) dt = c->tlim - c->t;
  c->dt = dt;
}
// The adaptive fofc order: an RK stage whose counterpart in the previous step flagged cells
// flags right after the update, before the boundary fill (the reference's order), so when it
// fires again it fills and converts once; other stages keep the flag inside con2prim and pay a
// second fill and con2prim only when they fire. Firing clusters by stage (the first stage's
// full forward step, mostly), not from one stage to the next. Both orders give the same bits:
// boundaries never change the active cells the flag reads. stage -1 (the single-kernel entry
// points) never flags early.
__device__ __forceinline__ void ctrl_begin_stage(Ctrl *c, int reset_slots, int stage) {
  c->epoch += 1;
  c->zsens = 0;
  c->stage = stage;
  c->early = stage >= 0 && ((c->fired >> stage) & 1);
  c->nbad = 0;
  if (reset_slots) { c->mdt_bits[0] = bits_of(1e300); c->mdt_bits[1] = bits_of(1e300); }
}
// the CFL minimum of the final (possibly fofc-repaired) con2prim of the last stage
// (fused order: slot 1 holds the minimum of the repaired state when cells were flagged).
// The fused order alternates the primitive sets every stage that ran.
__device__ __forceinline__ void ctrl_end_stage(Ctrl *c, int last, int fused) {
  if (!c->active) return;
  if (fused) c->wcur ^= 1;
  c->fofc_total += c->nbad;
  c->nbad_last = c->nbad;
  if (c->stage >= 0) c->fired = c->nbad ? (c->fired | (1 << c->stage)) : (c->fired & ~(1 << c->stage));
  if (last) c->min_dt = fmin(from_bits(c->mdt_bits[(fused && c->nbad) ? 1 : 0]), 1e300);
}
// main.cpp:291-292
__device__ __forceinline__ void ctrl_end_step(Ctrl *c) {
  if (!c->active) return;
  c->t += c->dt;
  c->cycle += 1;
}
__global__ void k_begin_step(Ctrl *c) { ctrl_begin_step(c); }
__global__ void k_begin_stage(Ctrl *c, int reset_slots, int stage) { ctrl_begin_stage(c, reset_slots, stage); }
__global__ void k_end_stage(Ctrl *c, int last, int fused) { ctrl_end_stage(c, last, fused); }
__global__ void k_end_step(Ctrl *c) { ctrl_end_step(c); }

// ---------------------------------------------------------------- batched kernels
// Grid-stride walks over device-resident (run, tile) lists; each item runs the same tile body
// as the single-mesh kernels, with that run's launch description (RunArgs) and control block.
// The primitive set a run reads is its Ctrl::wcur; con2prim writes the other one.

template <class T>
__global__ void __launch_bounds__(256) kb_corner(const RunArgs<T> *R, const ItemList L) {
  const int n = *L.n;
  for (int it = int(blockIdx.x); it < n; it += int(gridDim.x)) {
    const Item m = L.item[xcd_tile(it, n, XG_CORNER)];   // (the grids are multiples of 8 blocks)
    const Ctrl *ctl = R[m.run].ctrl;
    if (!ctl->active) continue;
    const CornerArgsT<T> a = R[m.run].corner[ctl->wcur];
    const Geom &g = a.g;
    int k, j, i;
    if (!corner_vb(g).at(m.tile, k, j, i)) continue;
    if (g.three_d) corner3_at(a, g.ks + k, g.js + j, g.is + i);
    else corner12_at(a, g.js + j, g.is + i);
  }
}

template <class T>
__device__ __forceinline__ void update_tile(const RunArgs<T> &run, const Ctrl *ctl, int tile, int out, double g0,
                                            double g1, double beta) {
  UpdateArgsT<T> a = run.update[out];
  a.g0 = g0;
  a.g1 = g1;
  const double bdt = beta * ctl->dt;
  const Geom &g = a.g;
  int k, j, i;
  if (update_vb(g).at(tile, k, j, i)) update_at(a, bdt, g.ks + k, g.js + j, g.is + i);
}

template <class T>
__global__ void __launch_bounds__(256) kb_update(const RunArgs<T> *R, const ItemList L, int out, double g0, double g1,
                                                 double beta) {
  const int n = *L.n;
  const int grp = std::is_same<T, float>::value ? XG_UPDATE_F32 : 0;
  for (int it = int(blockIdx.x); it < n; it += int(gridDim.x)) {
    const Item m = L.item[xcd_tile(it, n, grp)];
    const Ctrl *ctl = R[m.run].ctrl;
    if (!ctl->active) continue;
    update_tile(R[m.run], ctl, m.tile, out, g0, g1, beta);
  }
}

template <class T, int DIR>
__global__ void __launch_bounds__(256) kb_bc(const RunArgs<T> *R, const ItemList L, int buf) {
  const int n = *L.n;
  for (int it = int(blockIdx.x); it < n; it += int(gridDim.x)) {
    const Item m = L.item[it];
    if (!R[m.run].ctrl->active) continue;
    const BcArgs a = R[m.run].bc[buf][DIR];
    bc_at<DIR>(a, (long long)m.tile * 256 + threadIdx.x);
  }
}

// one con2prim tile of a run (the block's CFL minimum and flags merge into that run's
// control block); `cond` selects the fofc re-run (no flags, second minimum slot)
template <class T, bool LAST>
__device__ __forceinline__ void c2p_tile(const RunArgs<T> &run, const Ctrl *ctl, int tile, int buf, bool cond,
                                         double *smin, int *ssum) {
  C2pArgsT<T> a = run.c2p[buf][ctl->wcur ^ 1];
  a.compute_dt = LAST;
  a.entropy_out = 0;   // u[IENT]: deferred to the first host read (k_ient_sync)
  a.slot = cond ? 1 : 0;
  if (cond) a.flag_mode = 0;
  double mdt = 1e300;
  int nb = 0;
  int k, j, i;
  if (c2p_vb(a.g).at(tile, k, j, i)) c2p_at<T, LAST, false, false>(a, k, j, i, mdt, nb);   // batches: no dual energy
  if (LAST) c2p_block_merge<LAST>(a, mdt, nb, smin, ssum);
}

template <class T, bool LAST>
__global__ void __launch_bounds__(256) kb_c2p(const RunArgs<T> *R, const ItemList L, int buf) {
  __shared__ double smin[BX * BY / 64];
  __shared__ int ssum[BX * BY / 64];
  const int n = *L.n;
  for (int it = int(blockIdx.x); it < n; it += int(gridDim.x)) {
    const Item m = L.item[xcd_tile(it, n, XG_C2P)];
    const Ctrl *ctl = R[m.run].ctrl;
    if (!ctl->active) continue;
    c2p_tile<T, LAST>(R[m.run], ctl, m.tile, buf, false, smin, ssum);
  }
}

// Compaction of the runs that flagged cells this stage (one block): run list and, per class,
// the exclusive prefix of their tiles.
template <class T>
__global__ void __launch_bounds__(1024) kb_fofc_scan(const RunArgs<T> *R, const int *nruns, const CondList C) {
  constexpr int NC = CL_N + 1;   // flagged-run count + the classes
  __shared__ int s_tot[NC];
  __shared__ int s_wave[16][NC];
  const int tid = int(threadIdx.x), lane = tid & 63, wid = tid >> 6;
  if (tid < NC) s_tot[tid] = 0;
  __syncthreads();
  const int nr = *nruns;
  for (int base = 0; base < nr; base += 1024) {
    const int r = base + tid;
    int cnt[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) cnt[q] = 0;
    if (r < nr) {
      const Ctrl *c = R[r].ctrl;
      if (c->active && c->nbad > 0) {
        const int nb = c->nbad;
        cnt[0] = 1;
        cnt[1 + CL_REPAIR] = div_up(nb * 2 * R[r].ndir, 256);
        cnt[1 + CL_CORNER] = div_up(nb * 8, 256);
        cnt[1 + CL_UPDATE] = R[r].cond_tiles[0];
        cnt[1 + CL_BC1] = R[r].cond_tiles[1];
        cnt[1 + CL_BC2] = R[r].cond_tiles[2];
        cnt[1 + CL_BC3] = R[r].cond_tiles[3];
        cnt[1 + CL_C2P] = R[r].cond_tiles[4];
      }
    }
    // inclusive wave scans, then the wave offsets
    int inc[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      int v = cnt[q];
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
      }
      inc[q] = v;
      if (lane == 63) s_wave[wid][q] = v;
    }
    __syncthreads();
    if (cnt[0]) {
      int off[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        int o = s_tot[q];
        for (int w = 0; w < wid; ++w) o += s_wave[w][q];
        off[q] = o + inc[q] - cnt[q];
      }
      const int pos = off[0];
      C.run()[pos] = r;
#pragma unroll
      for (int q = 0; q < CL_N; ++q) C.pre(q)[pos] = off[1 + q];
    }
    __syncthreads();
    if (tid < NC) {
      int add = 0;
      for (int w = 0; w < 16; ++w) add += s_wave[w][tid];
      s_tot[tid] += add;
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int m = s_tot[0];
    *C.nrun() = m;
    for (int q = 0; q < CL_N; ++q) C.pre(q)[m] = s_tot[1 + q];
  }
}

// walk of one class of the compacted fofc list: fn(run, tile)
template <class F>
__device__ __forceinline__ void cond_walk(const CondList &C, int cls, F fn) {
  const int nrun = *C.nrun();
  if (nrun == 0) return;
  const int *pre = C.pre(cls);
  const int total = pre[nrun];
  for (int it = int(blockIdx.x); it < total; it += int(gridDim.x)) {
    const int p = locate(pre, nrun, it);
    fn(C.run()[p], it - pre[p]);
  }
}

template <class T>
__global__ void __launch_bounds__(256) kb_corner_cond(const RunArgs<T> *R, const CondList C) {
  cond_walk(C, CL_CORNER, [&](int r, int tile) {
    const Ctrl *ctl = R[r].ctrl;
    const CornerArgsT<T> a = R[r].corner[ctl->wcur];
    const int t = tile * 256 + int(threadIdx.x);
    if (t < ctl->nbad * 8) corner_list_entry(a, R[r].badlist, t);
  });
}

template <class T>
__global__ void __launch_bounds__(256) kb_update_cond(const RunArgs<T> *R, const CondList C, int out, double g0,
                                                      double g1, double beta) {
  cond_walk(C, CL_UPDATE, [&](int r, int tile) { update_tile(R[r], R[r].ctrl, tile, out, g0, g1, beta); });
}

template <class T, int DIR>
__global__ void __launch_bounds__(256) kb_bc_cond(const RunArgs<T> *R, const CondList C, int buf) {
  cond_walk(C, CL_BC1 + DIR, [&](int r, int tile) {
    const BcArgs a = R[r].bc[buf][DIR];
    bc_at<DIR>(a, (long long)tile * 256 + threadIdx.x);
  });
}

template <class T, bool LAST>
__global__ void __launch_bounds__(256) kb_c2p_cond(const RunArgs<T> *R, const CondList C, int buf) {
  __shared__ double smin[BX * BY / 64];
  __shared__ int ssum[BX * BY / 64];
  cond_walk(C, CL_C2P, [&](int r, int tile) { c2p_tile<T, LAST>(R[r], R[r].ctrl, tile, buf, true, smin, ssum); });
}

// The rest of the fofc branch of a batch of small runs, after kb_repair_runs: one block per
// run walks the tiles of its run phase by phase (edge EMFs around the listed cells, the
// whole-run re-update, boundaries x1 -> x2 -> x3, con2prim), with block barriers between the
// phases; a run that flagged nothing exits at once. The tile bodies are those of the
// multi-launch branch, so every value is the same.
template <class T, bool LAST>
__global__ void __launch_bounds__(256) kb_fofc_runs(const RunArgs<T> *R, const int *nruns, int out, double g0,
                                                    double g1, double beta) {
  __shared__ double smin[BX * BY / 64];
  __shared__ int ssum[BX * BY / 64];
  const int r = int(blockIdx.x);
  if (r >= *nruns) return;
  const RunArgs<T> &run = R[r];
  const Ctrl *ctl = run.ctrl;
  if (!ctl->active || ctl->nbad == 0) return;   // the same for every thread of the block
  const int nbad = ctl->nbad;
  {
    const CornerArgsT<T> a = run.corner[ctl->wcur];
    for (int t = int(threadIdx.x); t < nbad * 8; t += 256) corner_list_entry(a, run.badlist, t);
  }
  __syncthreads();
  for (int tile = 0; tile < run.cond_tiles[0]; ++tile) update_tile(run, ctl, tile, out, g0, g1, beta);
  __syncthreads();
  for (int tile = 0; tile < run.cond_tiles[1]; ++tile) bc_at<0>(run.bc[out][0], (long long)tile * 256 + threadIdx.x);
  __syncthreads();
  for (int tile = 0; tile < run.cond_tiles[2]; ++tile) bc_at<1>(run.bc[out][1], (long long)tile * 256 + threadIdx.x);
  __syncthreads();
  for (int tile = 0; tile < run.cond_tiles[3]; ++tile) bc_at<2>(run.bc[out][2], (long long)tile * 256 + threadIdx.x);
  __syncthreads();
  for (int tile = 0; tile < run.cond_tiles[4]; ++tile) c2p_tile<T, LAST>(run, ctl, tile, out, true, smin, ssum);
}

// live[r] (end of step) = 1 while run r has not reached tlim / nlim
template <class T>
__global__ void __launch_bounds__(256) kb_ctrl(const RunArgs<T> *R, const int *nruns, int op, int arg, int *live) {
  const int r = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (r >= *nruns) return;
  Ctrl *c = R[r].ctrl;
  if (op == CTRL_BEGIN_STEP) ctrl_begin_step(c);
  else if (op == CTRL_BEGIN_STAGE) ctrl_begin_stage(c, arg, -1);
  else if (op == CTRL_END_STAGE) ctrl_end_stage(c, arg, 1);
  else if (op == CTRL_STEP_FIRST) {
    ctrl_begin_step(c);
    ctrl_begin_stage(c, arg, -1);
  } else if (op == CTRL_TURN) {
    ctrl_end_stage(c, 0, 1);
    ctrl_begin_stage(c, arg, -1);
  } else {
    if (op == CTRL_END_LAST) ctrl_end_stage(c, 1, 1);
    if (op == CTRL_END_STEP || op == CTRL_END_LAST) ctrl_end_step(c);
    live[r] = ctrl_running(c) ? 1 : 0;
  }
}

}  // namespace

namespace {
int grid_of(const VB &v, bool conditional) { return conditional ? (v.n < COND_BLOCKS ? v.n : COND_BLOCKS) : v.n; }
}  // namespace

template <class T>
hipError_t launch_corner(const CornerArgsT<T> &a, hipStream_t s) {
  const Geom &g = a.g;
  if (g.three_d && !a.only_if_bad) {   // union box [ks..ke+1] in KSEG-plane segments
    const int nk = g.ke - g.ks + 2;
    const VB v = vgrid(g.ie - g.is + 2, g.je - g.js + 2, 1);
    hipLaunchKernelGGL(k_corner_march<T>, dim3(v.n, div_up(nk, KSEG_CORNER)), dim3(256), 0, s, a, v, nk,
                       KSEG_CORNER);
    return hipGetLastError();
  }
  const VB v = corner_vb(g);
  hipLaunchKernelGGL(k_corner<T>, dim3(grid_of(v, a.only_if_bad)), dim3(256), 0, s, a, v);
  return hipGetLastError();
}

template <class T>
hipError_t launch_corner_list(const CornerArgsT<T> &a, const int *list, hipStream_t s) {
  hipLaunchKernelGGL(k_corner_list<T>, dim3(256), dim3(256), 0, s, a, list);
  return hipGetLastError();
}

template <class T>
hipError_t launch_update(const UpdateArgsT<T> &a, hipStream_t s) {
  const Geom &g = a.g;
  // the k march pays off for the float field set; with FP64 fluxes the plane-ordered tiles of
  // k_update stream HBM better (measured at 256^3: 3.34 vs 3.79 ms per step)
  if (g.three_d && std::is_same<T, float>::value) {   // also the fofc re-run (gated in the kernel)
    const int nk = g.ke - g.ks + 2;
    const VB v = vgrid(g.ie - g.is + 2, g.je - g.js + 2, 1);
    hipLaunchKernelGGL(k_update_march<T>, dim3(v.n, div_up(nk, KSEG_UPDATE)), dim3(256), 0, s, a, v, nk,
                       KSEG_UPDATE);
    return hipGetLastError();
  }
  const VB v = update_vb(g);
  if (g.three_d) hipLaunchKernelGGL(k_update<T>, dim3(grid_of(v, a.only_if_bad)), dim3(256), 0, s, a, v);
  else hipLaunchKernelGGL(k_update_lf<T>, dim3(grid_of(v, a.only_if_bad)), dim3(256), 0, s, a, v);
  return hipGetLastError();
}

template hipError_t launch_corner<double>(const CornerArgsT<double> &, hipStream_t);
template hipError_t launch_corner<float>(const CornerArgsT<float> &, hipStream_t);
template hipError_t launch_corner_list<double>(const CornerArgsT<double> &, const int *, hipStream_t);
template hipError_t launch_corner_list<float>(const CornerArgsT<float> &, const int *, hipStream_t);
template hipError_t launch_update<double>(const UpdateArgsT<double> &, hipStream_t);
template hipError_t launch_update<float>(const UpdateArgsT<float> &, hipStream_t);

hipError_t launch_mark_near(const Geom &g, const Ctrl *ctrl, const int *list, int *marks, hipStream_t s) {
  hipLaunchKernelGGL(k_mark_near, dim3(256), dim3(256), 0, s, g, ctrl, list, marks);
  return hipGetLastError();
}

template <class T>
hipError_t launch_update_near(const UpdateArgsT<T> &a, const int *marks, hipStream_t s) {
  const VB v = update_vb(a.g);   // a capped grid-stride grid: most points only read their mark
  hipLaunchKernelGGL(k_update_near<T>, dim3(grid_of(v, true)), dim3(256), 0, s, a, v, marks);
  return hipGetLastError();
}
template hipError_t launch_update_near<double>(const UpdateArgsT<double> &, const int *, hipStream_t);
template hipError_t launch_update_near<float>(const UpdateArgsT<float> &, const int *, hipStream_t);

hipError_t launch_bc(const BcArgs &a, int dir, hipStream_t s) {
  const int nb = int((bc_points(a.g, dir) + 255) / 256);
  if (dir == 0) hipLaunchKernelGGL(k_bc<0>, dim3(nb), dim3(256), 0, s, a);
  else if (dir == 1) hipLaunchKernelGGL(k_bc<1>, dim3(nb), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_bc<2>, dim3(nb), dim3(256), 0, s, a);
  return hipGetLastError();
}

namespace {
template <class T, bool DE>
void launch_c2p_de(const C2pArgsT<T> &a, hipStream_t s) {
  const Geom &g = a.g;
  const bool ent = a.entropy_out || a.eos.dual_energy;
  const VB v = c2p_vb(g);
  const dim3 grid(v.n), block(256);
  if (a.compute_dt) {
    if (ent) hipLaunchKernelGGL((k_c2p<T, true, true, DE>), grid, block, 0, s, a, v);
    else hipLaunchKernelGGL((k_c2p<T, true, false, DE>), grid, block, 0, s, a, v);
  } else {
    if (ent) hipLaunchKernelGGL((k_c2p<T, false, true, DE>), grid, block, 0, s, a, v);
    else hipLaunchKernelGGL((k_c2p<T, false, false, DE>), grid, block, 0, s, a, v);
  }
}
}  // namespace

hipError_t launch_ient_sync(const C2pArgsT<double> &a, hipStream_t s) {
  const VB v = c2p_vb(a.g);
  hipLaunchKernelGGL(k_ient_sync, dim3(v.n), dim3(256), 0, s, a, v);
  return hipGetLastError();
}

template <class T>
hipError_t launch_c2p(const C2pArgsT<T> &a, hipStream_t s) {
  // dual energy only on the FP64 path (mixed precision and batches reject it)
  if constexpr (std::is_same<T, double>::value) {
    if (a.eos.dual_energy) {
      launch_c2p_de<T, true>(a, s);
      return hipGetLastError();
    }
  }
  launch_c2p_de<T, false>(a, s);
  return hipGetLastError();
}

template hipError_t launch_c2p<double>(const C2pArgsT<double> &, hipStream_t);
template hipError_t launch_c2p<float>(const C2pArgsT<float> &, hipStream_t);

hipError_t launch_flag(const C2pArgs &a, hipStream_t s) {
  const Geom &g = a.g;
  const int gx = div_up(g.ie - g.is + 1, BX), gy = div_up(g.je - g.js + 1, BY), nk = g.ke - g.ks + 1;
  const int gz = std::max(1, std::min(nk, COND_BLOCKS / (gx * gy)));   // planes strided by gz
  hipLaunchKernelGGL(k_flag, dim3(gx, gy, gz), dim3(BX, BY, 1), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ctrl_begin_step(Ctrl *c, hipStream_t s) {
  hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(1), 0, s, c);
  return hipGetLastError();
}
hipError_t launch_ctrl_begin_stage(Ctrl *c, int reset_slots, hipStream_t s, int stage) {
  hipLaunchKernelGGL(k_begin_stage, dim3(1), dim3(1), 0, s, c, reset_slots, stage);
  return hipGetLastError();
}
hipError_t launch_ctrl_end_stage(Ctrl *c, int last, int fused, hipStream_t s) {
  hipLaunchKernelGGL(k_end_stage, dim3(1), dim3(1), 0, s, c, last, fused);
  return hipGetLastError();
}
hipError_t launch_ctrl_end_step(Ctrl *c, hipStream_t s) {
  hipLaunchKernelGGL(k_end_step, dim3(1), dim3(1), 0, s, c);
  return hipGetLastError();
}

// ---------------------------------------------------------------- batched launches
template <class T>
hipError_t launchb_corner(const RunArgs<T> *R, ItemList L, int grid, hipStream_t s) {
  hipLaunchKernelGGL(kb_corner<T>, dim3(grid), dim3(256), 0, s, R, L);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_update(const RunArgs<T> *R, ItemList L, int out, double g0, double g1, double beta, int grid,
                          hipStream_t s) {
  hipLaunchKernelGGL(kb_update<T>, dim3(grid), dim3(256), 0, s, R, L, out, g0, g1, beta);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_bc(const RunArgs<T> *R, ItemList L, int dir, int buf, int grid, hipStream_t s) {
  if (dir == 0) hipLaunchKernelGGL((kb_bc<T, 0>), dim3(grid), dim3(256), 0, s, R, L, buf);
  else if (dir == 1) hipLaunchKernelGGL((kb_bc<T, 1>), dim3(grid), dim3(256), 0, s, R, L, buf);
  else hipLaunchKernelGGL((kb_bc<T, 2>), dim3(grid), dim3(256), 0, s, R, L, buf);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_c2p(const RunArgs<T> *R, ItemList L, int buf, int last, int grid, hipStream_t s) {
  if (last) hipLaunchKernelGGL((kb_c2p<T, true>), dim3(grid), dim3(256), 0, s, R, L, buf);
  else hipLaunchKernelGGL((kb_c2p<T, false>), dim3(grid), dim3(256), 0, s, R, L, buf);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_fofc_scan(const RunArgs<T> *R, const int *nruns, CondList C, hipStream_t s) {
  hipLaunchKernelGGL(kb_fofc_scan<T>, dim3(1), dim3(1024), 0, s, R, nruns, C);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_corner_cond(const RunArgs<T> *R, CondList C, int grid, hipStream_t s) {
  hipLaunchKernelGGL(kb_corner_cond<T>, dim3(grid), dim3(256), 0, s, R, C);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_update_cond(const RunArgs<T> *R, CondList C, int out, double g0, double g1, double beta, int grid,
                               hipStream_t s) {
  hipLaunchKernelGGL(kb_update_cond<T>, dim3(grid), dim3(256), 0, s, R, C, out, g0, g1, beta);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_bc_cond(const RunArgs<T> *R, CondList C, int dir, int buf, int grid, hipStream_t s) {
  if (dir == 0) hipLaunchKernelGGL((kb_bc_cond<T, 0>), dim3(grid), dim3(256), 0, s, R, C, buf);
  else if (dir == 1) hipLaunchKernelGGL((kb_bc_cond<T, 1>), dim3(grid), dim3(256), 0, s, R, C, buf);
  else hipLaunchKernelGGL((kb_bc_cond<T, 2>), dim3(grid), dim3(256), 0, s, R, C, buf);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_fofc_runs(const RunArgs<T> *R, const int *nruns, int cap, int out, double g0, double g1,
                             double beta, int last, hipStream_t s) {
  const dim3 grid(cap > 0 ? cap : 1), block(256);
  if (last) hipLaunchKernelGGL((kb_fofc_runs<T, true>), grid, block, 0, s, R, nruns, out, g0, g1, beta);
  else hipLaunchKernelGGL((kb_fofc_runs<T, false>), grid, block, 0, s, R, nruns, out, g0, g1, beta);
  return hipGetLastError();
}

template <class T>
hipError_t launchb_c2p_cond(const RunArgs<T> *R, CondList C, int buf, int last, int grid, hipStream_t s) {
  if (last) hipLaunchKernelGGL((kb_c2p_cond<T, true>), dim3(grid), dim3(256), 0, s, R, C, buf);
  else hipLaunchKernelGGL((kb_c2p_cond<T, false>), dim3(grid), dim3(256), 0, s, R, C, buf);
  return hipGetLastError();
}
template <class T>
hipError_t launchb_ctrl(const RunArgs<T> *R, const int *nruns, int cap, int op, int arg, int *live, hipStream_t s) {
  hipLaunchKernelGGL(kb_ctrl<T>, dim3(div_up(cap > 0 ? cap : 1, 256)), dim3(256), 0, s, R, nruns, op, arg, live);
  return hipGetLastError();
}

#define CHQ_BATCH_INST(T)                                                                                     \
  template hipError_t launchb_corner<T>(const RunArgs<T> *, ItemList, int, hipStream_t);                     \
  template hipError_t launchb_update<T>(const RunArgs<T> *, ItemList, int, double, double, double, int,      \
                                        hipStream_t);                                                         \
  template hipError_t launchb_bc<T>(const RunArgs<T> *, ItemList, int, int, int, hipStream_t);               \
  template hipError_t launchb_c2p<T>(const RunArgs<T> *, ItemList, int, int, int, hipStream_t);              \
  template hipError_t launchb_fofc_scan<T>(const RunArgs<T> *, const int *, CondList, hipStream_t);          \
  template hipError_t launchb_corner_cond<T>(const RunArgs<T> *, CondList, int, hipStream_t);                \
  template hipError_t launchb_update_cond<T>(const RunArgs<T> *, CondList, int, double, double, double, int, \
                                             hipStream_t);                                                    \
  template hipError_t launchb_bc_cond<T>(const RunArgs<T> *, CondList, int, int, int, hipStream_t);          \
  template hipError_t launchb_c2p_cond<T>(const RunArgs<T> *, CondList, int, int, int, hipStream_t);         \
  template hipError_t launchb_fofc_runs<T>(const RunArgs<T> *, const int *, int, int, double, double, double, int, \
                                           hipStream_t);         \
  template hipError_t launchb_ctrl<T>(const RunArgs<T> *, const int *, int, int, int, int *, hipStream_t);
CHQ_BATCH_INST(double)
CHQ_BATCH_INST(float)
#undef CHQ_BATCH_INST

}  // namespace chq
