// Fused face-flux kernel (reconstruction + flattening + positivity fallback + Riemann solve
// + flux/face-EMF store) and the fofc donor-cell repair kernel, for gfx950.
//
// Reference: kernels::Fluxes<DIR,R,S> (physics.hpp:65-226) and fofc::RepairFluxes<DIR>
// (fofc.hpp:46-96). L/R states live only in registers. One thread per face; a block is a
// 64 x 4 tile of faces (i fastest), so every stencil load of a wave64 is one coalesced
// 512 B row segment, and the six loads of a stencil along j or k re-hit the same rows in
// L1/L2 from the neighbouring waves of the block.
#include "kernels.hpp"
#include "numerics.hpp"

namespace chq {

namespace {

constexpr int BX = 64, BY = 4;

template <int DIR>
__device__ __forceinline__ long long dir_stride(const Geom &g) {
  if constexpr (DIR == 0) return 1;
  else if constexpr (DIR == 1) return g.pitch;
  else return g.plane;
}

template <int DIR, Recon RC, RSolver RS, bool ENT>
__global__ void __launch_bounds__(BX * BY) k_flux(const FluxArgs a) {
  if (!a.ctrl->active) return;
  const int i = a.il + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = a.jl + int(blockIdx.y) * BY + int(threadIdx.y);
  const int k = a.kl + int(blockIdx.z);
  if (i > a.iu || j > a.ju) return;
  const Geom &g = a.g;
  using R = Rec<double>;
  using V = Rsv<double>;
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  constexpr int B1 = (DIR + 1) % 3, B2 = (DIR + 2) % 3;
  const long long c = g.at(k, j, i);
  const long long sd = dir_stride<DIR>(g);
  const long long cs = g.cstride;
  constexpr bool wide = !(RC == Recon::dc || RC == Recon::plm);

  // stencils of d, u, v, w, p, B_t1, B_t2 in the rotated frame (physics.hpp:99-106)
  double st[7][6];
  const double *src[7] = {a.w + IDN * cs, a.w + IU * cs, a.w + IV * cs, a.w + IW * cs,
                          a.w + IPR * cs, a.bcc + B1 * cs, a.bcc + B2 * cs};
#pragma unroll
  for (int n = 0; n < 7; ++n) {
    const double *p = src[n] + c;
    if constexpr (wide) {
#pragma unroll
      for (int s = 0; s < 6; ++s) st[n][s] = p[(s - 3) * sd];
    } else {
#pragma unroll
      for (int s = 1; s < 5; ++s) st[n][s] = p[(s - 3) * sd];
      st[n][0] = st[n][1];
      st[n][5] = st[n][4];
    }
  }
  double wl[7], wr[7];
#pragma unroll
  for (int n = 0; n < 7; ++n) R::template face<RC>(st[n], wl[n], wr[n]);

  double sl = 0.0, sr = 0.0, sl0 = 0.0, sr0 = 0.0;   // entropy (physics.hpp:152-154)
  if constexpr (ENT) {
    double q[6];
    const double *p = a.w + IENT * cs + c;
    if constexpr (wide) {
#pragma unroll
      for (int s = 0; s < 6; ++s) q[s] = p[(s - 3) * sd];
    } else {
#pragma unroll
      for (int s = 1; s < 5; ++s) q[s] = p[(s - 3) * sd];
      q[0] = q[1];
      q[5] = q[4];
    }
    R::template face<RC>(q, sl, sr);
    sl0 = q[2];
    sr0 = q[3];
  }
  if (wide && a.flatten) {                                         // physics.hpp:156-162
    const double fL = R::flat(st[4][0], st[4][1], st[4][2], st[4][3], st[4][4], st[1][1], st[1][3]);
    const double fR = R::flat(st[4][1], st[4][2], st[4][3], st[4][4], st[4][5], st[1][2], st[1][4]);
    if (fL > 0.0) {
#pragma unroll
      for (int n = 0; n < 7; ++n) wl[n] += fL * (st[n][2] - wl[n]);
      if constexpr (ENT) sl += fL * (sl0 - sl);
    }
    if (fR > 0.0) {
#pragma unroll
      for (int n = 0; n < 7; ++n) wr[n] += fR * (st[n][3] - wr[n]);
      if constexpr (ENT) sr += fR * (sr0 - sr);
    }
  }
  const Eos e = a.eos;
  if (wl[0] <= e.dfloor || wl[4] <= e.pfloor) {                    // physics.hpp:163-165
#pragma unroll
    for (int n = 0; n < 7; ++n) wl[n] = st[n][2];
    sl = sl0;
  }
  if (wr[0] <= e.dfloor || wr[4] <= e.pfloor) {
#pragma unroll
    for (int n = 0; n < 7; ++n) wr[n] = st[n][3];
    sr = sr0;
  }
  St<double> L{wl[0], wl[1], wl[2], wl[3], wl[4], wl[5], wl[6]};
  St<double> Rs{wr[0], wr[1], wr[2], wr[3], wr[4], wr[5], wr[6]};
  const double bn = a.bn[c];
  if (a.low_mach) {                                                // physics.hpp:170-178
    const double cl = V::cfast(L, bn, e.gamma), cr = V::cfast(Rs, bn, e.gamma);
    const double z = fmin(1.0, fmax(fabs(L.u) / cl, fabs(Rs.u) / cr));
    const double um = 0.5 * (L.u + Rs.u), ud = 0.5 * (z * (L.u - Rs.u));
    L.u = um + ud;
    Rs.u = um - ud;
  }
  double f[NFLX];
  if constexpr (RS == RSolver::hlld) {
    V::template hlld<false>(L, Rs, bn, e.gamma, e.gm1, f);
  } else if constexpr (RS == RSolver::lhlld) {                     // physics.hpp:182-211
    constexpr int t1 = (DIR + 1) % 3, t2 = (DIR + 2) % 3;
    const double *wu = a.w + IU * cs, *wv = a.w + IV * cs, *ww = a.w + IW * cs;
    const long long cL = c - sd;
    const long long s1 = t1 == 0 ? 1 : t1 == 1 ? g.pitch : g.plane;
    const long long s2 = t2 == 0 ? 1 : t2 == 1 ? g.pitch : g.plane;
    const bool has1 = t1 == 0 || (t1 == 1 && g.two_d) || (t1 == 2 && g.three_d);
    const bool has2 = t2 == 0 || (t2 == 1 && g.two_d) || (t2 == 2 && g.three_d);
    const double cfm = fmax(V::cfast(L, bn, e.gamma), V::cfast(Rs, bn, e.gamma));
    const double du = wu[c] - wu[cL];
    double dvw = 0.0;
    if (has1) {
      const double d1 = fmin(fmin(wv[cL] - wv[cL - s1], wv[cL + s1] - wv[cL]), fmin(wv[c] - wv[c - s1], wv[c + s1] - wv[c]));
      dvw = fmin(dvw, d1);
    }
    if (has2) {
      const double d2 = fmin(fmin(ww[cL] - ww[cL - s2], ww[cL + s2] - ww[cL]), fmin(ww[c] - ww[c - s2], ww[c + s2] - ww[c]));
      dvw = fmin(dvw, d2);
    }
    const double ratio = fmin(1.0, (-fmin(du, 0.0) + cfm) / (-dvw + cfm));
    const double r2 = ratio * ratio;
    const double theta = r2 * r2;
    const double chi = fmin(1.0, fmax(V::cflow(L, bn), V::cflow(Rs, bn)) / cfm);
    const double phi = (a.lhlld_phi >= 0.0) ? a.lhlld_phi : chi * (2.0 - chi);
    V::template hlld<true>(L, Rs, bn, e.gamma, e.gm1, f, theta, phi);
  } else {
    V::hlle(L, Rs, bn, e.gamma, e.gm1, f);
  }
  // store (physics.hpp:215-224)
  double *fx = a.flx + c;
  fx[IDN * cs] = f[FD];
  fx[IU * cs] = f[FMU];
  fx[IV * cs] = f[FMV];
  fx[IW * cs] = f[FMW];
  fx[IEN * cs] = f[FEN];
  if constexpr (ENT)
    fx[IENT * cs] = (f[FD] > 0.0) ? f[FD] * sl : (f[FD] < 0.0) ? f[FD] * sr : 0.5 * (f[FD] * (sl + sr));
  a.et2[c] = -f[FBY];
  a.et1[c] = f[FBZ];
}

// Donor-cell HLLE on faces that touch a flagged active cell (fofc.hpp:66-94). In the
// fused step it runs after every stage and exits at once unless the stage flagged cells.
template <int DIR>
__global__ void __launch_bounds__(BX * BY) k_repair(const FluxArgs a, const int only_if_bad) {
  if (!a.ctrl->active) return;
  if (only_if_bad && a.ctrl->nbad == 0) return;
  const int i = a.il + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = a.jl + int(blockIdx.y) * BY + int(threadIdx.y);
  const int k = a.kl + int(blockIdx.z);
  if (i > a.iu || j > a.ju) return;
  const Geom &g = a.g;
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  constexpr int B1 = (DIR + 1) % 3, B2 = (DIR + 2) % 3;
  const int kL = k - dk, jL = j - dj, iL = i - di;
  auto inside = [&](int kk, int jj, int ii) {
    return ii >= g.is && ii <= g.ie && jj >= g.js && jj <= g.je && kk >= g.ks && kk <= g.ke;
  };
  const long long c = g.at(k, j, i), cL = g.at(kL, jL, iL), cs = g.cstride;
  const bool fl = inside(k, j, i) && a.flag[c];
  const bool fr = inside(kL, jL, iL) && a.flag[cL];
  if (!(fl || fr)) return;
  const double *w = a.w, *b = a.bcc;
  St<double> L{w[IDN * cs + cL], w[IU * cs + cL], w[IV * cs + cL], w[IW * cs + cL], w[IPR * cs + cL],
               b[B1 * cs + cL], b[B2 * cs + cL]};
  St<double> R{w[IDN * cs + c], w[IU * cs + c], w[IV * cs + c], w[IW * cs + c], w[IPR * cs + c],
               b[B1 * cs + c], b[B2 * cs + c]};
  double f[NFLX];
  Rsv<double>::hlle(L, R, a.bn[c], a.eos.gamma, a.eos.gm1, f);
  double *fx = a.flx + c;
  fx[IDN * cs] = f[FD];
  fx[IU * cs] = f[FMU];
  fx[IV * cs] = f[FMV];
  fx[IW * cs] = f[FMW];
  fx[IEN * cs] = f[FEN];
  if (a.entropy) {  // upwind IENT flux with the donor entropies (fofc.hpp:85-86)
    const double sL = w[IENT * cs + cL], sR = w[IENT * cs + c];
    fx[IENT * cs] = (f[FD] > 0.0) ? f[FD] * sL : (f[FD] < 0.0) ? f[FD] * sR : 0.5 * (f[FD] * (sL + sR));
  }
  a.et2[c] = -f[FBY];
  a.et1[c] = f[FBZ];
}

template <int DIR, Recon RC, RSolver RS, bool ENT>
hipError_t go_flux(const FluxArgs &a, hipStream_t s) {
  const int nx = a.iu - a.il + 1, ny = a.ju - a.jl + 1, nz = a.ku - a.kl + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
  dim3 grid(div_up(nx, BX), div_up(ny, BY), nz), block(BX, BY, 1);
  hipLaunchKernelGGL((k_flux<DIR, RC, RS, ENT>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <int DIR, Recon RC, bool ENT>
hipError_t go_flux_rs(const FluxArgs &a, RSolver rs, hipStream_t s) {
  switch (rs) {
    case RSolver::hlld: return go_flux<DIR, RC, RSolver::hlld, ENT>(a, s);
    case RSolver::lhlld: return go_flux<DIR, RC, RSolver::lhlld, ENT>(a, s);
    default: return go_flux<DIR, RC, RSolver::hlle, ENT>(a, s);
  }
}

template <int DIR, bool ENT>
hipError_t go_flux_rc(const FluxArgs &a, Recon rc, RSolver rs, hipStream_t s) {
  switch (rc) {
    case Recon::dc: return go_flux_rs<DIR, Recon::dc, ENT>(a, rs, s);
    case Recon::plm: return go_flux_rs<DIR, Recon::plm, ENT>(a, rs, s);
    case Recon::ppm: return go_flux_rs<DIR, Recon::ppm, ENT>(a, rs, s);
    case Recon::ppmx: return go_flux_rs<DIR, Recon::ppmx, ENT>(a, rs, s);
    default: return go_flux_rs<DIR, Recon::wenoz, ENT>(a, rs, s);
  }
}

template <int DIR>
hipError_t go_flux_dir(const FluxArgs &a, Recon rc, RSolver rs, bool ent, hipStream_t s) {
  return ent ? go_flux_rc<DIR, true>(a, rc, rs, s) : go_flux_rc<DIR, false>(a, rc, rs, s);
}

}  // namespace

hipError_t launch_flux(const FluxArgs &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s) {
  if (dir == 0) return go_flux_dir<0>(a, rc, rs, entropy, s);
  if (dir == 1) return go_flux_dir<1>(a, rc, rs, entropy, s);
  return go_flux_dir<2>(a, rc, rs, entropy, s);
}

hipError_t launch_repair(const FluxArgs &a, int dir, bool only_if_bad, hipStream_t s) {
  const int nx = a.iu - a.il + 1, ny = a.ju - a.jl + 1, nz = a.ku - a.kl + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
  dim3 grid(div_up(nx, BX), div_up(ny, BY), nz), block(BX, BY, 1);
  const int oib = only_if_bad ? 1 : 0;
  if (dir == 0) hipLaunchKernelGGL(k_repair<0>, grid, block, 0, s, a, oib);
  else if (dir == 1) hipLaunchKernelGGL(k_repair<1>, grid, block, 0, s, a, oib);
  else hipLaunchKernelGGL(k_repair<2>, grid, block, 0, s, a, oib);
  return hipGetLastError();
}

}  // namespace chq
