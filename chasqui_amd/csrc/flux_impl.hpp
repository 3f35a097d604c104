// Face-flux kernel templates (reconstruction + flattening + positivity fallback + Riemann
// solve + flux/face-EMF store) and the fofc donor-cell repair, for gfx950. Included by
// flux.hip (T = double, strict IEEE arithmetic: the FP64 drop-in) and flux32.hip (T = float:
// the mixed-precision path's FP32 reconstruction, Riemann solve and EMFs).
//
// Reference: kernels::Fluxes<DIR,R,S> (physics.hpp:65-226) and fofc::RepairFluxes<DIR>
// (fofc.hpp:46-96). L/R states never leave registers. Each cell's two half-states are
// computed once (recon_cell.hpp) and shared by its two faces:
//   * x faces (DIR 0): lane = cell over the rows' cells laid end to end; neighbour values
//     move between lanes (DPP whole-wave shifts in FP64); a wave yields 64 - 5 faces (halo 3 + 2)
//     and may span the end of one row and the start of the next.
//   * y and z faces (DIR 1, 2): lane = i (coalesced loads), each thread marches along j or
//     k over a segment of a.seg faces with a register window: every primitive is loaded
//     once per segment and each cell is reconstructed once. Threads are numbered with i
//     fastest over the exact face range, so no lane idles at a row end.
// For T = double the arithmetic and its order are the reference's: fluxes are bitwise equal.
#pragma once
#include "kernels.hpp"
#include "numerics.hpp"
#include "recon_cell.hpp"

namespace chq {
namespace fluxk {

// The value of lane l - S (lane_up) or l + S (lane_dn) across the wave, S = 1 or 2. FP64, S = 1:
// DPP whole-wave shifts (wave_shr:1 / wave_shl:1, one VALU move per 32-bit half instead of two
// LDS permutes and their wait; the FP64 x kernels ran 4% faster). The first lane (lane_up) or
// the last (lane_dn) receives 0; the x-face kernels take no result from those halo lanes (their
// faces need the shifted values of lanes H - 1 .. H + FPW only). Every lane of the wave must be
// active. A shift by 2 (the wide stencils' outer cells) stays an LDS permute: as two DPP
// shifts it made the PPMX x kernel slower. FP32 keeps __shfl_up / __shfl_down: at eight waves per
// SIMD the FP32 x kernel is issue-bound, and the extra VALU moves made it slower.
template <bool UP>
__device__ __forceinline__ double wave_shift1(double v) {
  constexpr int ctrl = UP ? 0x138 : 0x130;
  const long long b = __double_as_longlong(v);
  const int l2 = __builtin_amdgcn_mov_dpp(int(b), ctrl, 0xF, 0xF, true);
  const int h2 = __builtin_amdgcn_mov_dpp(int(b >> 32), ctrl, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)h2 << 32) | (unsigned)l2));
}
template <int S = 1, class T>
__device__ __forceinline__ T lane_up(T v) {
  if constexpr (sizeof(T) == 8 && S == 1) return wave_shift1<true>(v);
  else return __shfl_up(v, S, 64);
}
template <int S = 1, class T>
__device__ __forceinline__ T lane_dn(T v) {
  if constexpr (sizeof(T) == 8 && S == 1) return wave_shift1<false>(v);
  else return __shfl_down(v, S, 64);
}


// Riemann solve of one face from its two half-states (physics.hpp:167-211) into f[NFLX].
// c = face index (device layout); sd = stride along DIR.
template <class T, int DIR, RSolver RS, int NV>
__device__ __forceinline__ void solve_face(const FluxArgsT<T> &a, long long c, long long sd, const T *Ls,
                                           const T *Rs_, T f[NFLX]) {
  using V = Rsv<T>;
  const Geom &g = a.g;
  const T gam = T(a.eos.gamma), gm1 = T(a.eos.gm1);
  const long long cs = g.cstride;
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  St<T> L{Ls[0], Ls[1], Ls[2], Ls[3], Ls[4], Ls[5], Ls[6]};
  St<T> R{Rs_[0], Rs_[1], Rs_[2], Rs_[3], Rs_[4], Rs_[5], Rs_[6]};
  const T bn = T(a.bn[c]);
  if (a.low_mach) {                                                // physics.hpp:170-178
    const T cl = V::cfast(L, bn, gam), cr = V::cfast(R, bn, gam);
    const T z = vmin(T(1), vmax(vabs(L.u) / cl, vabs(R.u) / cr));
    const T um = T(0.5) * (L.u + R.u), ud = T(0.5) * (z * (L.u - R.u));
    L.u = um + ud;
    R.u = um - ud;
  }
  if constexpr (RS == RSolver::hlld) {
    V::template hlld<false>(L, R, bn, gam, gm1, f);
  } else if constexpr (RS == RSolver::lhlld) {                     // physics.hpp:182-211
    constexpr int t1 = (DIR + 1) % 3, t2 = (DIR + 2) % 3;
    const T *wu = a.w + IU * cs, *wv = a.w + IV * cs, *ww = a.w + IW * cs;
    const long long cL = c - sd;
    const long long s1 = t1 == 0 ? 1 : t1 == 1 ? g.pitch : g.plane;
    const long long s2 = t2 == 0 ? 1 : t2 == 1 ? g.pitch : g.plane;
    const bool has1 = t1 == 0 || (t1 == 1 && g.two_d) || (t1 == 2 && g.three_d);
    const bool has2 = t2 == 0 || (t2 == 1 && g.two_d) || (t2 == 2 && g.three_d);
    const T cfm = vmax(V::cfast(L, bn, gam), V::cfast(R, bn, gam));
    const T du = wu[c] - wu[cL];
    T dvw = T(0);
    if (has1) {
      const T d1 = vmin(vmin(wv[cL] - wv[cL - s1], wv[cL + s1] - wv[cL]), vmin(wv[c] - wv[c - s1], wv[c + s1] - wv[c]));
      dvw = vmin(dvw, d1);
    }
    if (has2) {
      const T d2 = vmin(vmin(ww[cL] - ww[cL - s2], ww[cL + s2] - ww[cL]), vmin(ww[c] - ww[c - s2], ww[c + s2] - ww[c]));
      dvw = vmin(dvw, d2);
    }
    const T ratio = vmin(T(1), (-vmin(du, T(0)) + cfm) / (-dvw + cfm));
    const T r2 = ratio * ratio;
    const T theta = r2 * r2;
    const T chi = vmin(T(1), vmax(V::cflow(L, bn), V::cflow(R, bn)) / cfm);
    const T phi = (a.lhlld_phi >= 0.0) ? T(a.lhlld_phi) : chi * (T(2) - chi);
    V::template hlld<true>(L, R, bn, gam, gm1, f, theta, phi);
  } else {
    V::hlle(L, R, bn, gam, gm1, f);
  }
}

// Riemann solve of one face and store of its fluxes and face EMFs (physics.hpp:167-224).
template <class T, int DIR, RSolver RS, bool ENT, int NV>
__device__ __forceinline__ void solve_store(const FluxArgsT<T> &a, long long c, long long sd, const T *Ls,
                                            const T *Rs_) {
  const long long cs = a.g.cstride;
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  T f[NFLX];
  solve_face<T, DIR, RS, NV>(a, c, sd, Ls, Rs_, f);
  T *fx = a.flx + c;
  fx[IDN * cs] = f[FD];
  fx[IU * cs] = f[FMU];
  fx[IV * cs] = f[FMV];
  fx[IW * cs] = f[FMW];
  fx[IEN * cs] = f[FEN];
  if constexpr (ENT) {
    const T sl = Ls[7], sr = Rs_[7];
    fx[IENT * cs] = (f[FD] > T(0)) ? f[FD] * sl : (f[FD] < T(0)) ? f[FD] * sr : T(0.5) * (f[FD] * (sl + sr));
  }
  a.et2[c] = -f[FBY];
  a.et1[c] = f[FBZ];
}

// the NV reconstructed variables in the rotated frame: d, u, v, w, p (from w), B_t1, B_t2
// (from bcc), IENT
template <class T, int DIR, int NV>
__device__ __forceinline__ void var_ptrs(const FluxArgsT<T> &a, const T *p[NV]) {
  const long long cs = a.g.cstride;
  p[0] = a.w + IDN * cs;
  p[1] = a.w + (IVX + DIR) * cs;
  p[2] = a.w + (IVX + (DIR + 1) % 3) * cs;
  p[3] = a.w + (IVX + (DIR + 2) % 3) * cs;
  p[4] = a.w + IPR * cs;
  p[5] = a.bcc + ((DIR + 1) % 3) * cs;
  p[6] = a.bcc + ((DIR + 2) % 3) * cs;
  if constexpr (NV == 8) p[7] = a.w + IENT * cs;
}

// Lane = cell along the face direction: the two half-states of this lane's cell from the
// cells of the neighbouring lanes (flattening and the donor fallback included) and, in
// Lface, the left state of the face below this lane's cell (the lower lane's Lf).
template <class T, Recon RC, int NV>
__device__ __forceinline__ void recon_lanes(const T q[NV], const bool flatten, const T dfl, const T pfl, T Lface[NV],
                                            T Rf[NV]) {
  constexpr bool W = wide_v<RC>;
  using R = Rec<T>;
  T Lf[NV];
  T fco = T(0);
  if constexpr (RC == Recon::ppm) {
    T d[NV], e[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      const T qm = lane_up(q[n]), qp = lane_dn(q[n]);
      d[n] = R::mc(qm, q[n], qp);
    }
#pragma unroll
    for (int n = 0; n < NV; ++n) {   // e[n] = edge at c+1/2
      const T qp = lane_dn(q[n]), dp = lane_dn(d[n]);
      e[n] = R::edge(q[n], qp, d[n], dp);
    }
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      T lo = lane_up(e[n]), hi = e[n];
      R::mono(q[n], lo, hi);
      Rf[n] = lo;
      Lf[n] = hi;
    }
  } else {
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      const T qm1 = lane_up(q[n]), qp1 = lane_dn(q[n]);
      T qm2 = qm1, qp2 = qp1;
      if constexpr (W) {
        qm2 = lane_up<2>(q[n]);
        qp2 = lane_dn<2>(q[n]);
      }
      cell_halves<T, RC>(qm2, qm1, q[n], qp1, qp2, Rf[n], Lf[n]);
    }
  }
  if constexpr (W) {
    if (flatten) {
      const T p = q[4];
      const T pm2 = lane_up<2>(p), pm1 = lane_up(p);
      const T pp1 = lane_dn(p), pp2 = lane_dn<2>(p);
      const T um1 = lane_up(q[1]), up1 = lane_dn(q[1]);
      fco = R::flat(pm2, pm1, p, pp1, pp2, um1, up1);
    }
  }
  finish_cell<T, NV>(Lf, Rf, q, fco, W && flatten, dfl, pfl);
  // face i = cell index of this lane: left state = Lf of the lane below
#pragma unroll
  for (int n = 0; n < NV; ++n) Lface[n] = lane_up(Lf[n]);
}

// ---------------------------------------------------------------- x faces: lane = cell
// One wave's row segment of x faces (lane = cell).
template <class T, Recon RC, RSolver RS, bool ENT>
__device__ __forceinline__ void flux_x_wave(const FluxArgsT<T> &a, const int nseg, const int wave) {
  constexpr int NV = ENT ? 8 : 7;
  constexpr bool W = wide_v<RC>;
  constexpr int H = W ? 3 : 2;            // lanes before the first face: cell halo + 1
  constexpr int FPW = fpw_v<RC>;
  const int lane = int(threadIdx.x) & 63;
  const Geom &g = a.g;
  // this lane's cell in the rows laid end to end (nseg = cells per row); lanes before the
  // first or past the last cell read any valid cell and compute no face
  const int nrow_j = a.ju - a.jl + 1;
  const int ncell = nseg * nrow_j * (a.ku - a.kl + 1);   // below 2^30 (the mesh size limit)
  const int v = wave * FPW - H + lane;
  const unsigned vc = unsigned(v < 0 ? 0 : (v >= ncell ? ncell - 1 : v));
  const unsigned row = fdiv(vc, a.q1), kr = fdiv(row, a.q2);   // vc / nseg, row / nrow_j
  const int ci = int(vc - row * unsigned(nseg));
  const int j = a.jl + int(row - kr * unsigned(nrow_j)), k = a.kl + int(kr);
  const long long row0 = g.at(k, j, 0);
  const T *src[NV];
  var_ptrs<T, 0, NV>(a, src);
  T q[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) q[n] = src[n][row0 + ci];
  T Lface[NV], Rf[NV];
  recon_lanes<T, RC, NV>(q, a.flatten, T(a.eos.dfloor), T(a.eos.pfloor), Lface, Rf);
  // face ci (between cells ci - 1 and ci): its stencil ci - H .. ci + 2 lies in this row and in
  // this wave's lanes
  if (lane < H || lane >= H + FPW || v != int(vc) || ci < a.il || ci > a.iu) return;
  solve_store<T, 0, RS, ENT, NV>(a, row0 + ci, 1, Lface, Rf);
}


template <class T, Recon RC, RSolver RS, bool ENT>
__global__ void __launch_bounds__(256) k_flux_x(const FluxArgsT<T> a, const int nseg, const int nwaves) {
  if (!a.ctrl->active) return;
  const int wave = int(blockIdx.x) * 4 + int(threadIdx.x >> 6);
  if (wave >= nwaves) return;
  flux_x_wave<T, RC, RS, ENT>(a, nseg, wave);
}

// ---------------------------------------------------------------- y/z faces: register march
template <class T, int DIR, Recon RC, RSolver RS, bool ENT>
__device__ __forceinline__ void flux_m_tile(const FluxArgsT<T> &a, const int bx) {
  constexpr int NV = ENT ? 8 : 7;
  constexpr bool W = wide_v<RC>;
  const Geom &g = a.g;
  // i, transverse index t (k for DIR 1, j for DIR 2) and segment along DIR
  int i, t, sg;
  if (!m_item<T, DIR>(a, (long long)bx * (BX * BY) + int(threadIdx.y) * BX + int(threadIdx.x), i, t, sg)) return;
  const int lo_d = DIR == 1 ? a.jl : a.kl, hi_d = DIR == 1 ? a.ju : a.ku;
  const int F0 = lo_d + sg * a.seg;
  const int F1 = min(F0 + a.seg - 1, hi_d);
  int j, k;
  if constexpr (DIR == 1) { k = a.kl + t; j = 0; }
  else { j = a.jl + t; k = 0; }
  const long long sd = DIR == 1 ? g.pitch : g.plane;
  const long long base = DIR == 1 ? g.at(k, 0, i) : g.at(0, j, i);   // index of (d = 0)
  const T *src[NV];
  var_ptrs<T, DIR, NV>(a, src);
  using R = Rec<T>;
  const bool flat = W && a.flatten;
  const T dfl = T(a.eos.dfloor), pfl = T(a.eos.pfloor);
  // window around the cell being finished, c: x1 = q_{c-2}, x2 = q_{c-1}, w0 = q_c, wp1 = q_{c+1}
  // (PPM keeps x1 = d_c and x2 = e_{c-1/2} instead, plus p_{c-2}, p_{c-1}, u_{c-1})
  T w0[NV], wp1[NV], x1[NV], x2[NV];
  T pm2 = T(0), pm1 = T(0), um1 = T(0);
  int c = F0 - 1;
  {
    const long long cc = base + (long long)c * sd;
    T qm2[NV], qm1[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      qm1[n] = src[n][cc - sd];
      w0[n] = src[n][cc];
      wp1[n] = src[n][cc + sd];
      qm2[n] = W ? src[n][cc - 2 * sd] : qm1[n];
    }
    if constexpr (RC == Recon::ppm) {
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        const T dm = R::mc(qm2[n], qm1[n], w0[n]);
        x1[n] = R::mc(qm1[n], w0[n], wp1[n]);            // d_c
        x2[n] = R::edge(qm1[n], w0[n], dm, x1[n]);       // e_{c-1/2}
      }
      pm2 = qm2[4];
      pm1 = qm1[4];
      um1 = qm1[1];
    } else {
#pragma unroll
      for (int n = 0; n < NV; ++n) { x1[n] = qm2[n]; x2[n] = qm1[n]; }
    }
  }
  T Lprev[NV];
  // wide stencils: q_{c+2} of the next iteration is loaded one iteration ahead, so its latency
  // overlaps the current cell's arithmetic (the same addresses as loading it in place)
  T qpre[NV];
  if constexpr (W) {
    const long long cc = base + (long long)c * sd;
#pragma unroll
    for (int n = 0; n < NV; ++n) qpre[n] = src[n][cc + 2 * sd];
  }
  for (; c <= F1; ++c) {
    // finish cell c: needs q_{c+2} (wide) or q_{c+1} (narrow, already in wp1)
    const long long cc = base + (long long)c * sd;
    T qn[NV];
    if constexpr (W) {
#pragma unroll
      for (int n = 0; n < NV; ++n) qn[n] = qpre[n];
      if (c < F1) {
#pragma unroll
        for (int n = 0; n < NV; ++n) qpre[n] = src[n][cc + 3 * sd];
      }
    } else {
#pragma unroll
      for (int n = 0; n < NV; ++n) qn[n] = wp1[n];
    }
    T Lf[NV], Rf[NV], fco = T(0);
    if constexpr (RC == Recon::ppm) {
      T dn[NV], en[NV];
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        dn[n] = R::mc(w0[n], wp1[n], qn[n]);             // d_{c+1}
        en[n] = R::edge(w0[n], wp1[n], x1[n], dn[n]);    // e_{c+1/2}
        T lo = x2[n], hi = en[n];
        R::mono(w0[n], lo, hi);
        Rf[n] = lo;
        Lf[n] = hi;
      }
      if (flat) fco = R::flat(pm2, pm1, w0[4], wp1[4], qn[4], um1, wp1[1]);
      pm2 = pm1;
      pm1 = w0[4];
      um1 = w0[1];
#pragma unroll
      for (int n = 0; n < NV; ++n) { x1[n] = dn[n]; x2[n] = en[n]; }
    } else {
#pragma unroll
      for (int n = 0; n < NV; ++n) cell_halves<T, RC>(x1[n], x2[n], w0[n], wp1[n], qn[n], Rf[n], Lf[n]);
      if (flat) fco = R::flat(x1[4], x2[4], w0[4], wp1[4], qn[4], x2[1], wp1[1]);
#pragma unroll
      for (int n = 0; n < NV; ++n) { x1[n] = x2[n]; x2[n] = w0[n]; }
    }
    T qc[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) qc[n] = w0[n];
    finish_cell<T, NV>(Lf, Rf, qc, fco, flat, dfl, pfl);
    if (c >= F0) solve_store<T, DIR, RS, ENT, NV>(a, cc, sd, Lprev, Rf);   // face c (between c-1 and c)
    // advance the window
    if constexpr (W) {
#pragma unroll
      for (int n = 0; n < NV; ++n) { w0[n] = wp1[n]; wp1[n] = qn[n]; }
    } else if (c < F1) {   // narrow stencil: the next cell's right neighbour
#pragma unroll
      for (int n = 0; n < NV; ++n) { w0[n] = wp1[n]; wp1[n] = src[n][cc + 2 * sd]; }
    }
#pragma unroll
    for (int n = 0; n < NV; ++n) Lprev[n] = Lf[n];
  }
}

template <class T, int DIR, Recon RC, RSolver RS, bool ENT>
__global__ void __launch_bounds__(BX * BY) k_flux_m(const FluxArgsT<T> a) {
  if (!a.ctrl->active) return;
  flux_m_tile<T, DIR, RC, RS, ENT>(a, int(blockIdx.x));
}

// Donor-cell HLLE flux and face EMFs of one face from the stage-input primitives
// (fofc.hpp:76-88).
template <class T, int DIR>
__device__ __forceinline__ void repair_face(const FluxArgsT<T> &a, long long c, long long sd) {
  const Geom &g = a.g;
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  constexpr int B1 = (DIR + 1) % 3, B2 = (DIR + 2) % 3;
  const long long cL = c - sd, cs = g.cstride;
  const T *w = a.w, *b = a.bcc;
  St<T> L{w[IDN * cs + cL], w[IU * cs + cL], w[IV * cs + cL], w[IW * cs + cL], w[IPR * cs + cL],
          b[B1 * cs + cL], b[B2 * cs + cL]};
  St<T> R{w[IDN * cs + c], w[IU * cs + c], w[IV * cs + c], w[IW * cs + c], w[IPR * cs + c],
          b[B1 * cs + c], b[B2 * cs + c]};
  T f[NFLX];
  Rsv<T>::hlle(L, R, T(a.bn[c]), T(a.eos.gamma), T(a.eos.gm1), f);
  T *fx = a.flx + c;
  fx[IDN * cs] = f[FD];
  fx[IU * cs] = f[FMU];
  fx[IV * cs] = f[FMV];
  fx[IW * cs] = f[FMW];
  fx[IEN * cs] = f[FEN];
  if (a.entropy) {  // upwind IENT flux with the donor entropies (fofc.hpp:85-86)
    const T sL = w[IENT * cs + cL], sR = w[IENT * cs + c];
    fx[IENT * cs] = (f[FD] > T(0)) ? f[FD] * sL : (f[FD] < T(0)) ? f[FD] * sR : T(0.5) * (f[FD] * (sL + sR));
  }
  a.et2[c] = -f[FBY];
  a.et1[c] = f[FBZ];
}

// The repair driven by the list of flagged (active) cells that con2prim builds: every face
// touching a listed cell — its two faces in each direction — is recomputed. A face shared by
// two listed cells is written twice with the same value.
template <class T>
__global__ void __launch_bounds__(256) k_repair_list(const FluxArgsT<T> a0, const FluxArgsT<T> a1,
                                                     const FluxArgsT<T> a2, const int *list, const int ndir) {
  const Ctrl *ctl = a0.ctrl;
  if (!ctl->active || ctl->nbad == 0) return;
  const Geom &g = a0.g;
  const int per = 2 * ndir;
  const int n = ctl->nbad * per;
  for (int t = int(blockIdx.x) * 256 + int(threadIdx.x); t < n; t += int(gridDim.x) * 256) {
    const long long c0 = list[t / per];
    const int r = t % per, dir = r >> 1, side = r & 1;
    if (dir == 0) repair_face<T, 0>(a0, c0 + side, 1);
    else if (dir == 1) repair_face<T, 1>(a1, c0 + side * (long long)g.pitch, g.pitch);
    else repair_face<T, 2>(a2, c0 + side * g.plane, g.plane);
  }
}

template <class T, int DIR, Recon RC, RSolver RS, bool ENT>
hipError_t go_flux(const FluxArgsT<T> &a, hipStream_t s) {
  const int nx = a.iu - a.il + 1, ny = a.ju - a.jl + 1, nz = a.ku - a.kl + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
  if constexpr (DIR == 0) {
    int nseg, nwaves;
    x_shape(a, fpw_v<RC>, nseg, nwaves);
    hipLaunchKernelGGL((k_flux_x<T, RC, RS, ENT>), dim3(div_up(nwaves, 4)), dim3(256), 0, s, a, nseg, nwaves);
  } else {
    int gx, gy, gz;
    m_shape<T, DIR>(a, gx, gy, gz);
    hipLaunchKernelGGL((k_flux_m<T, DIR, RC, RS, ENT>), dim3(gx, gy, gz), dim3(BX, BY, 1), 0, s, a);
  }
  return hipGetLastError();
}

template <class T, int DIR, Recon RC, bool ENT>
hipError_t go_flux_rs(const FluxArgsT<T> &a, RSolver rs, hipStream_t s) {
  switch (rs) {
    case RSolver::hlld: return go_flux<T, DIR, RC, RSolver::hlld, ENT>(a, s);
    case RSolver::lhlld: return go_flux<T, DIR, RC, RSolver::lhlld, ENT>(a, s);
    default: return go_flux<T, DIR, RC, RSolver::hlle, ENT>(a, s);
  }
}

template <class T, int DIR, bool ENT>
hipError_t go_flux_rc(const FluxArgsT<T> &a, Recon rc, RSolver rs, hipStream_t s) {
  switch (rc) {
    case Recon::dc: return go_flux_rs<T, DIR, Recon::dc, ENT>(a, rs, s);
    case Recon::plm: return go_flux_rs<T, DIR, Recon::plm, ENT>(a, rs, s);
    case Recon::ppm: return go_flux_rs<T, DIR, Recon::ppm, ENT>(a, rs, s);
    case Recon::ppmx: return go_flux_rs<T, DIR, Recon::ppmx, ENT>(a, rs, s);
    default: return go_flux_rs<T, DIR, Recon::wenoz, ENT>(a, rs, s);
  }
}

template <class T>
hipError_t launch_flux_t(const FluxArgsT<T> &a, int dir, Recon rc, RSolver rs, bool ent, hipStream_t s) {
  if (dir == 0) return ent ? go_flux_rc<T, 0, true>(a, rc, rs, s) : go_flux_rc<T, 0, false>(a, rc, rs, s);
  if (dir == 1) return ent ? go_flux_rc<T, 1, true>(a, rc, rs, s) : go_flux_rc<T, 1, false>(a, rc, rs, s);
  return ent ? go_flux_rc<T, 2, true>(a, rc, rs, s) : go_flux_rc<T, 2, false>(a, rc, rs, s);
}

template <class T>
hipError_t launch_repair_list_t(const FluxArgsT<T> a[3], int ndir, const int *list, hipStream_t s) {
  hipLaunchKernelGGL(k_repair_list<T>, dim3(256), dim3(256), 0, s, a[0], a[ndir > 1 ? 1 : 0], a[ndir > 2 ? 2 : 0],
                     list, ndir);
  return hipGetLastError();
}


// ---------------------------------------------------------------- batched launches
// Grid-stride walks over (run, tile) lists of a batch; the tile bodies are the ones above,
// with the run's launch description for its current primitive set (Ctrl::wcur).
template <class T, Recon RC, RSolver RS>
__global__ void __launch_bounds__(256) kb_flux_x(const RunArgs<T> *R, const ItemList L, const int cur) {
  const int n = *L.n;
  for (int it = int(blockIdx.x); it < n; it += int(gridDim.x)) {
    const Item m = L.item[it];
    const Ctrl *ctl = R[m.run].ctrl;
    if (!ctl->active) continue;
    const FluxArgsT<T> a = R[m.run].flux[cur][ctl->wcur][0];
    int nseg, nwaves;
    x_shape(a, fpw_v<RC>, nseg, nwaves);
    const int wave = m.tile * 4 + int(threadIdx.x >> 6);
    if (wave < nwaves) flux_x_wave<T, RC, RS, false>(a, nseg, wave);
  }
}

template <class T, int DIR, Recon RC, RSolver RS>
__global__ void __launch_bounds__(BX * BY) kb_flux_m(const RunArgs<T> *R, const ItemList L, const int cur) {
  const int n = *L.n;
  for (int it = int(blockIdx.x); it < n; it += int(gridDim.x)) {
    const Item m = L.item[it];
    const Ctrl *ctl = R[m.run].ctrl;
    if (!ctl->active) continue;
    const FluxArgsT<T> a = R[m.run].flux[cur][ctl->wcur][DIR];
    flux_m_tile<T, DIR, RC, RS, false>(a, m.tile);
  }
}

// fofc repair of the flagged runs (compacted list, tiles of 256 (cell, face) entries)
// entry t of run r's repair: face `side` of direction `dir` of listed cell t / (2 ndir)
template <class T>
__device__ __forceinline__ void repair_entry(const RunArgs<T> &run, const Ctrl *ctl, int t, int cur) {
  const int per = 2 * run.ndir;
  if (t >= ctl->nbad * per) return;
  const FluxArgsT<T> *fa = run.flux[cur][ctl->wcur];
  const Geom &g = fa[0].g;
  const long long c0 = run.badlist[t / per];
  const int q = t % per, dir = q >> 1, side = q & 1;
  if (dir == 0) repair_face<T, 0>(fa[0], c0 + side, 1);
  else if (dir == 1) repair_face<T, 1>(fa[1], c0 + side * (long long)g.pitch, g.pitch);
  else repair_face<T, 2>(fa[2], c0 + side * g.plane, g.plane);
}

template <class T>
__global__ void __launch_bounds__(256) kb_repair(const RunArgs<T> *R, const CondList C, const int cur) {
  const int nrun = *C.nrun();
  if (nrun == 0) return;
  const int *pre = C.pre(CL_REPAIR);
  const int total = pre[nrun];
  for (int it = int(blockIdx.x); it < total; it += int(gridDim.x)) {
    const int p = locate(pre, nrun, it);
    const int r = C.run()[p];
    repair_entry(R[r], R[r].ctrl, (it - pre[p]) * 256 + int(threadIdx.x), cur);
  }
}

// The repair of a batch of small runs: one block per run, which walks its run's entries
// (runs that flagged nothing exit at once; no compaction pass is needed).
template <class T>
__global__ void __launch_bounds__(256) kb_repair_runs(const RunArgs<T> *R, const int *nruns, const int cur) {
  const int r = int(blockIdx.x);
  if (r >= *nruns) return;
  const Ctrl *ctl = R[r].ctrl;
  if (!ctl->active || ctl->nbad == 0) return;
  const int n = ctl->nbad * 2 * R[r].ndir;
  for (int t = int(threadIdx.x); t < n; t += 256) repair_entry(R[r], ctl, t, cur);
}

template <class T, int DIR, Recon RC>
hipError_t gob_flux_rs(const RunArgs<T> *R, ItemList L, int cur, RSolver rs, int grid, hipStream_t s) {
  if constexpr (DIR == 0) {
    if (rs == RSolver::hlld) hipLaunchKernelGGL((kb_flux_x<T, RC, RSolver::hlld>), dim3(grid), dim3(256), 0, s, R, L, cur);
    else if (rs == RSolver::lhlld) hipLaunchKernelGGL((kb_flux_x<T, RC, RSolver::lhlld>), dim3(grid), dim3(256), 0, s, R, L, cur);
    else hipLaunchKernelGGL((kb_flux_x<T, RC, RSolver::hlle>), dim3(grid), dim3(256), 0, s, R, L, cur);
  } else {
    const dim3 b(BX, BY, 1);
    if (rs == RSolver::hlld) hipLaunchKernelGGL((kb_flux_m<T, DIR, RC, RSolver::hlld>), dim3(grid), b, 0, s, R, L, cur);
    else if (rs == RSolver::lhlld) hipLaunchKernelGGL((kb_flux_m<T, DIR, RC, RSolver::lhlld>), dim3(grid), b, 0, s, R, L, cur);
    else hipLaunchKernelGGL((kb_flux_m<T, DIR, RC, RSolver::hlle>), dim3(grid), b, 0, s, R, L, cur);
  }
  return hipGetLastError();
}

template <class T, int DIR>
hipError_t gob_flux_rc(const RunArgs<T> *R, ItemList L, int cur, Recon rc, RSolver rs, int grid, hipStream_t s) {
  switch (rc) {
    case Recon::dc: return gob_flux_rs<T, DIR, Recon::dc>(R, L, cur, rs, grid, s);
    case Recon::plm: return gob_flux_rs<T, DIR, Recon::plm>(R, L, cur, rs, grid, s);
    case Recon::ppm: return gob_flux_rs<T, DIR, Recon::ppm>(R, L, cur, rs, grid, s);
    case Recon::ppmx: return gob_flux_rs<T, DIR, Recon::ppmx>(R, L, cur, rs, grid, s);
    default: return gob_flux_rs<T, DIR, Recon::wenoz>(R, L, cur, rs, grid, s);
  }
}

template <class T>
hipError_t launchb_flux_t(const RunArgs<T> *R, ItemList L, int dir, int cur, Recon rc, RSolver rs, int grid, hipStream_t s) {
  if (dir == 0) return gob_flux_rc<T, 0>(R, L, cur, rc, rs, grid, s);
  if (dir == 1) return gob_flux_rc<T, 1>(R, L, cur, rc, rs, grid, s);
  return gob_flux_rc<T, 2>(R, L, cur, rc, rs, grid, s);
}

template <class T>
hipError_t launchb_repair_t(const RunArgs<T> *R, CondList C, int cur, int grid, hipStream_t s) {
  hipLaunchKernelGGL(kb_repair<T>, dim3(grid), dim3(256), 0, s, R, C, cur);
  return hipGetLastError();
}

template <class T>
hipError_t launchb_repair_runs_t(const RunArgs<T> *R, const int *nruns, int cap, int cur, hipStream_t s) {
  hipLaunchKernelGGL(kb_repair_runs<T>, dim3(cap > 0 ? cap : 1), dim3(256), 0, s, R, nruns, cur);
  return hipGetLastError();
}

}  // namespace fluxk
}  // namespace chq
