// CPU oracle — pointwise numerics (reconstruction, flattening, Riemann solvers).
//
// TEST INFRASTRUCTURE ONLY. This directory is the parity oracle for the HIP stage
// pipeline: a plain C++ restatement of the reference algorithm, used by tests/, by
// __graft_entry__.smoke() and by bench.py's CPU-baseline leg. The product library never
// includes, links or calls anything here.
//
// Every routine restates the reference with the same floating-point operation order (the
// order is what makes bitwise mirror symmetry and bitwise CPU/GPU parity possible); each
// one cites the reference file:line it follows. Build with -ffp-contract=off.
//
// Pins (see tests/test_oracle_*.py): bitwise mirror symmetry of every routine (the
// reference's own invariant test, tests/symmetry_test.cpp:20-70, re-implemented with the
// same sample classes, plus LHLLD and DC which the reference leaves untested); closed
// forms (constant/linear data, extremum -> donor, WENO-Z ideal weights, F(U,U)=F(U),
// supersonic upwinding, isolated contact); and, through the whole pipeline, the CPAW L1
// table of README.md:90-96.
#pragma once
#include <cmath>

namespace orc {

using Real = double;

inline Real sqr(Real x) { return x * x; }
inline Real sgn(Real x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); }  // defs.hpp:33

enum class Recon : int { dc = 0, plm = 1, ppm = 2, wenoz = 3, ppmx = 4 };   // defs.hpp:35
enum class RSolver : int { hlle = 0, hlld = 1, lhlld = 2 };                 // physics.hpp:20

struct Eos {                       // defs.hpp:39-46
  Real gamma = 5.0 / 3.0, gm1 = 2.0 / 3.0, dfloor = 1e-10, pfloor = 1e-12;
  bool dual_energy = false;
  Real de_eta = 1e-3;
};

// ---------------------------------------------------------------- reconstruction
// Stencil convention (recon.hpp:3-7): s[0..5] = cells i-3..i+2 around face i-1/2;
// L = right-face value of cell i-1, R = left-face value of cell i.

// van Leer harmonic slope, recon.hpp:14-18
inline Real slope_vl(Real a, Real b, Real c) {
  const Real l = b - a, r = c - b;
  const Real p = l * r;
  return (p > 0.0) ? 2.0 * p / (l + r) : 0.0;
}

// monotonized-central slope, recon.hpp:21-27
inline Real slope_mc(Real a, Real b, Real c) {
  const Real l = b - a, r = c - b;
  if (l * r <= 0.0) return 0.0;
  const Real cen = 0.5 * (c - a);
  const Real cap = 2.0 * std::fmin(std::fabs(l), std::fabs(r));
  return sgn(cen) * std::fmin(std::fabs(cen), cap);
}

// recon.hpp:29-37
inline void rec_dc(const Real s[6], Real &L, Real &R) { L = s[2]; R = s[3]; }
inline void rec_plm(const Real s[6], Real &L, Real &R) {
  L = s[2] + 0.5 * slope_vl(s[1], s[2], s[3]);
  R = s[3] - 0.5 * slope_vl(s[2], s[3], s[4]);
}

// CW84 interface value, recon.hpp:42-44
inline Real ppm_edge(Real qa, Real qb, Real da, Real db) { return 0.5 * (qa + qb) + (da - db) / 6.0; }

// CW84 parabola monotonization, recon.hpp:47-57
inline void ppm_mono(Real q, Real &lo, Real &hi) {
  if ((hi - q) * (q - lo) <= 0.0) { lo = q; hi = q; return; }
  const Real d = hi - lo;
  const Real c = q - 0.5 * (lo + hi);
  const Real t = d * d / 6.0;
  if (d * c > t) lo = 3.0 * q - 2.0 * hi;
  else if (d * c < -t) hi = 3.0 * q - 2.0 * lo;
}

// recon.hpp:60-78
inline void rec_ppm(const Real s[6], Real &L, Real &R) {
  const Real m1 = slope_mc(s[0], s[1], s[2]);
  const Real m2 = slope_mc(s[1], s[2], s[3]);
  const Real m3 = slope_mc(s[2], s[3], s[4]);
  const Real m4 = slope_mc(s[3], s[4], s[5]);
  const Real e0 = ppm_edge(s[1], s[2], m1, m2);
  const Real e1 = ppm_edge(s[2], s[3], m2, m3);
  const Real e2 = ppm_edge(s[3], s[4], m3, m4);
  Real a = e0, b = e1;
  ppm_mono(s[2], a, b);
  L = b;
  Real c = e1, d = e2;
  ppm_mono(s[3], c, d);
  R = c;
}

// extremum-preserving 4th-order interface, recon.hpp:85-100
inline Real ppmx_edge(Real qm1, Real q0, Real qp1, Real qp2) {
  const Real C = 1.25;
  Real f = (7.0 / 12.0) * (q0 + qp1) - (1.0 / 12.0) * (qm1 + qp2);
  if ((f - q0) * (qp1 - f) < 0.0) {
    const Real d2 = 3.0 * ((q0 + qp1) - 2.0 * f);
    const Real dl = (qm1 + qp1) - 2.0 * q0;
    const Real dr = (q0 + qp2) - 2.0 * qp1;
    Real lim = 0.0;
    if (d2 * dl > 0.0 && d2 * dr > 0.0)
      lim = sgn(d2) * std::fmin(std::fmin(C * std::fabs(dl), C * std::fabs(dr)), std::fabs(d2));
    f = 0.5 * (q0 + qp1) - lim / 6.0;
  }
  return f;
}

// recon.hpp:104-127
inline void ppmx_mono(Real qm2, Real qm1, Real q, Real qp1, Real qp2, Real &lo, Real &hi) {
  const Real C = 1.25;
  const Real dm = q - lo, dp = hi - q;
  const bool ext = (dm * dp <= 0.0) || ((qp1 - q) * (q - qm1) <= 0.0);
  if (ext) {
    const Real d2 = 6.0 * ((lo + hi) - 2.0 * q);
    const Real dc = (qm1 + qp1) - 2.0 * q;
    const Real dl = (qm2 + q) - 2.0 * qm1;
    const Real dr = (q + qp2) - 2.0 * qp1;
    Real lim = 0.0;
    if (d2 * dc > 0.0 && d2 * dl > 0.0 && d2 * dr > 0.0)
      lim = sgn(d2) * std::fmin(std::fmin(C * std::fabs(dl), C * std::fabs(dr)),
                                std::fmin(C * std::fabs(dc), std::fabs(d2)));
    const Real ratio = (std::fabs(d2) > 1e-12 * (std::fabs(q) + std::fabs(qm1) + std::fabs(qp1)) + 1e-300)
                           ? lim / d2 : 0.0;
    if (ratio < 1.0 - 1e-12) { lo = q - ratio * dm; hi = q + ratio * dp; }
  } else {
    if (std::fabs(dm) >= 2.0 * std::fabs(dp)) lo = q - 2.0 * dp;
    if (std::fabs(dp) >= 2.0 * std::fabs(dm)) hi = q + 2.0 * dm;
  }
}

// recon.hpp:129-140
inline void rec_ppmx(const Real s[6], Real &L, Real &R) {
  const Real e0 = ppmx_edge(s[0], s[1], s[2], s[3]);
  const Real e1 = ppmx_edge(s[1], s[2], s[3], s[4]);
  const Real e2 = ppmx_edge(s[2], s[3], s[4], s[5]);
  Real a = e0, b = e1;
  ppmx_mono(s[0], s[1], s[2], s[3], s[4], a, b);
  L = b;
  Real c = e1, d = e2;
  ppmx_mono(s[1], s[2], s[3], s[4], s[5], c, d);
  R = c;
}

// WENO-Z (Borges et al. 2008) right-interface value, recon.hpp:144-164 (FP64 eps 1e-36)
inline Real wenoz_right(Real a, Real b, Real c, Real d, Real e) {
  const Real eps = 1e-36;
  const Real k13 = 13.0 / 12.0;
  const Real s0 = k13 * sqr((a + c) - 2.0 * b) + 0.25 * sqr((3.0 * c + a) - 4.0 * b);
  const Real s1 = k13 * sqr((b + d) - 2.0 * c) + 0.25 * sqr(b - d);
  const Real s2 = k13 * sqr((e + c) - 2.0 * d) + 0.25 * sqr((3.0 * c + e) - 4.0 * d);
  const Real tau = std::fabs(s0 - s2);
  const Real w0 = 0.1 * (1.0 + tau / (s0 + eps));
  const Real w1 = 0.6 * (1.0 + tau / (s1 + eps));
  const Real w2 = 0.3 * (1.0 + tau / (s2 + eps));
  const Real inv = 1.0 / (w0 + w1 + w2);
  const Real p0 = (2.0 * a - 7.0 * b + 11.0 * c) / 6.0;
  const Real p1 = (-b + 5.0 * c + 2.0 * d) / 6.0;
  const Real p2 = (2.0 * c + 5.0 * d - e) / 6.0;
  return inv * (w0 * p0 + w1 * p1 + w2 * p2);
}

// recon.hpp:166-171
inline void rec_wenoz(const Real s[6], Real &L, Real &R) {
  L = wenoz_right(s[0], s[1], s[2], s[3], s[4]);
  R = wenoz_right(s[5], s[4], s[3], s[2], s[1]);
}

inline void reconstruct(Recon r, const Real s[6], Real &L, Real &R) {   // recon.hpp:189-196
  switch (r) {
    case Recon::dc: rec_dc(s, L, R); break;
    case Recon::plm: rec_plm(s, L, R); break;
    case Recon::ppm: rec_ppm(s, L, R); break;
    case Recon::ppmx: rec_ppmx(s, L, R); break;
    default: rec_wenoz(s, L, R);
  }
}

// CW84 shock flattening coefficient of a cell, recon.hpp:179-187.
// p[0..4] = pressures j-2..j+2; um1, up1 = normal velocities of j-1, j+1.
inline Real flatten_coef(const Real p[5], Real um1, Real up1) {
  const Real dp = p[3] - p[1];
  const Real dp2 = p[4] - p[0];
  const Real pm = std::fmin(p[1], p[3]);
  if (!((std::fabs(dp) > 0.33 * pm) && (um1 > up1))) return 0.0;
  const Real r = (std::fabs(dp2) > 1e-300) ? dp / dp2 : 0.0;
  return std::fmin(1.0, std::fmax(0.0, 10.0 * (r - 0.75)));
}

// ---------------------------------------------------------------- Riemann solvers
// Rotated-frame state (riemann.hpp:11-13): u normal, (v,w) transverse, (by,bz) transverse B.
struct Prim { Real d, u, v, w, p, by, bz; };
enum { F_D = 0, F_MU, F_MV, F_MW, F_E, F_BY, F_BZ, NF = 7 };   // riemann.hpp:17

// riemann.hpp:19-25
inline Real fast_speed(const Prim &s, Real bn, Real gamma) {
  const Real b2 = bn * bn + s.by * s.by + s.bz * s.bz;
  const Real gp = gamma * s.p;
  const Real a = (gp + b2) / s.d;
  const Real disc = a * a - 4.0 * gp * bn * bn / (s.d * s.d);
  return std::sqrt(0.5 * (a + std::sqrt(std::fmax(disc, 0.0))));
}

// riemann.hpp:27-40
inline void phys_flux(const Prim &s, Real bn, Real gm1, Real &e, Real &pt, Real f[NF]) {
  const Real v2 = s.u * s.u + s.v * s.v + s.w * s.w;
  const Real b2 = bn * bn + s.by * s.by + s.bz * s.bz;
  e = s.p / gm1 + 0.5 * (s.d * v2 + b2);
  pt = s.p + 0.5 * b2;
  const Real vb = s.u * bn + s.v * s.by + s.w * s.bz;
  f[F_D] = s.d * s.u;
  f[F_MU] = s.d * s.u * s.u + pt - bn * bn;
  f[F_MV] = s.d * s.u * s.v - bn * s.by;
  f[F_MW] = s.d * s.u * s.w - bn * s.bz;
  f[F_E] = (e + pt) * s.u - bn * vb;
  f[F_BY] = s.by * s.u - bn * s.v;
  f[F_BZ] = s.bz * s.u - bn * s.w;
}

// Minoshima-Miyoshi modified fast speed, riemann.hpp:44-51
inline Real flow_speed_u(const Prim &s, Real bn) {
  const Real v2 = s.u * s.u + s.v * s.v + s.w * s.w;
  const Real ca2 = (bn * bn + s.by * s.by + s.bz * s.bz) / s.d;
  const Real cax2 = bn * bn / s.d;
  const Real a = ca2 + v2;
  const Real disc = a * a - 4.0 * v2 * cax2;
  return std::sqrt(0.5 * (a + std::sqrt(std::fmax(disc, 0.0))));
}

// HLLE, riemann.hpp:55-68
inline void hlle(const Prim &L, const Prim &R, Real bn, const Eos &eos, Real f[NF]) {
  Real fl[NF], fr[NF], el, er, pl, pr;
  phys_flux(L, bn, eos.gm1, el, pl, fl);
  phys_flux(R, bn, eos.gm1, er, pr, fr);
  const Real cl = fast_speed(L, bn, eos.gamma), cr = fast_speed(R, bn, eos.gamma);
  const Real sl = std::fmin(L.u - cl, R.u - cr);
  const Real sr = std::fmax(L.u + cl, R.u + cr);
  if (sl >= 0.0) { for (int n = 0; n < NF; ++n) f[n] = fl[n]; return; }
  if (sr <= 0.0) { for (int n = 0; n < NF; ++n) f[n] = fr[n]; return; }
  const Real ul[NF] = {L.d, L.d * L.u, L.d * L.v, L.d * L.w, el, L.by, L.bz};
  const Real ur[NF] = {R.d, R.d * R.u, R.d * R.v, R.d * R.w, er, R.by, R.bz};
  const Real inv = 1.0 / (sr - sl);
  for (int n = 0; n < NF; ++n) f[n] = (sr * fl[n] - sl * fr[n] + sl * sr * (ur[n] - ul[n])) * inv;
}

// flux of an intermediate state with total pressure pts, riemann.hpp:81-91
inline void star_flux(Real d, Real sm, Real v, Real w, Real e, Real by, Real bz, Real bn, Real pts, Real f[NF]) {
  const Real vb = (sm * bn + v * by) + w * bz;
  f[F_D] = d * sm;
  f[F_MU] = (d * sm) * sm + (pts - bn * bn);
  f[F_MV] = (d * sm) * v - bn * by;
  f[F_MW] = (d * sm) * w - bn * bz;
  f[F_E] = (e + pts) * sm - bn * vb;
  f[F_BY] = by * sm - bn * v;
  f[F_BZ] = bz * sm - bn * w;
}

// HLLD (Miyoshi & Kusano 2005) with the LHLLD factors of Minoshima & Miyoshi 2021,
// riemann.hpp:96-234. theta = phi = 1 and lhlld = false is plain HLLD.
inline void hlld(const Prim &L, const Prim &R, Real bn, const Eos &eos, Real f[NF],
                 Real theta = 1.0, Real phi = 1.0, bool lhlld = false) {
  Real fl[NF], fr[NF], el, er, ptl, ptr;
  phys_flux(L, bn, eos.gm1, el, ptl, fl);
  phys_flux(R, bn, eos.gm1, er, ptr, fr);
  const Real cl = fast_speed(L, bn, eos.gamma), cr = fast_speed(R, bn, eos.gamma);
  Real sl, sr;
  if (lhlld) {
    const Real cm = std::fmax(cl, cr);
    sl = std::fmin(0.0, std::fmin(L.u, R.u) - cm);
    sr = std::fmax(0.0, std::fmax(L.u, R.u) + cm);
  } else {
    sl = std::fmin(L.u - cl, R.u - cr);
    sr = std::fmax(L.u + cl, R.u + cr);
  }
  if (sl >= 0.0) { for (int n = 0; n < NF; ++n) f[n] = fl[n]; return; }
  if (sr <= 0.0) { for (int n = 0; n < NF; ++n) f[n] = fr[n]; return; }

  const Real bn2 = bn * bn;
  const Real al = sl - L.u, ar = sr - R.u;
  const Real ml = L.d * al, mr = R.d * ar;
  const Real iden = 1.0 / (mr - ml);
  const Real sm = ((mr * R.u - ml * L.u) + theta * (ptl - ptr)) * iden;
  const Real pts = ((mr * ptl - ml * ptr) + (phi * ((L.d * R.d) * (ar * al))) * (R.u - L.u)) * iden;

  // left star state (riemann.hpp:127-147)
  const Real bl = sl - sm;
  const Real dls = ml / bl;
  Real vls, wls, byls, bzls;
  {
    const Real den = (L.d * al) * bl - bn2;
    const Real scale = std::fabs((L.d * al) * bl) + bn2;
    if (std::fabs(den) <= 1e-12 * scale + 1e-300) {
      vls = L.v; wls = L.w; byls = L.by; bzls = L.bz;
    } else {
      const Real nv = (bn * (sm - L.u)) / den;
      vls = L.v - L.by * nv;
      wls = L.w - L.bz * nv;
      const Real nb = ((L.d * al) * al - bn2) / den;
      byls = L.by * nb;
      bzls = L.bz * nb;
    }
  }
  const Real vbl = (L.u * bn + L.v * L.by) + L.w * L.bz;
  const Real vbls = (sm * bn + vls * byls) + wls * bzls;
  const Real els = (((al * el - ptl * L.u) + pts * sm) + bn * (vbl - vbls)) / bl;

  // right star state (riemann.hpp:149-169)
  const Real br = sr - sm;
  const Real drs = mr / br;
  Real vrs, wrs, byrs, bzrs;
  {
    const Real den = (R.d * ar) * br - bn2;
    const Real scale = std::fabs((R.d * ar) * br) + bn2;
    if (std::fabs(den) <= 1e-12 * scale + 1e-300) {
      vrs = R.v; wrs = R.w; byrs = R.by; bzrs = R.bz;
    } else {
      const Real nv = (bn * (sm - R.u)) / den;
      vrs = R.v - R.by * nv;
      wrs = R.w - R.bz * nv;
      const Real nb = ((R.d * ar) * ar - bn2) / den;
      byrs = R.by * nb;
      bzrs = R.bz * nb;
    }
  }
  const Real vbr = (R.u * bn + R.v * R.by) + R.w * R.bz;
  const Real vbrs = (sm * bn + vrs * byrs) + wrs * bzrs;
  const Real ers = (((ar * er - ptr * R.u) + pts * sm) + bn * (vbr - vbrs)) / br;

  // Alfven waves bounding the double-star region (riemann.hpp:171-175)
  const Real qdl = std::sqrt(dls), qdr = std::sqrt(drs);
  const Real abn = std::fabs(bn);
  const Real sls = sm - abn / qdl;
  const Real srs = sm + abn / qdr;

  const Real uls[NF] = {dls, dls * sm, dls * vls, dls * wls, els, byls, bzls};
  const Real urs[NF] = {drs, drs * sm, drs * vrs, drs * wrs, ers, byrs, bzrs};
  const Real ul[NF] = {L.d, L.d * L.u, L.d * L.v, L.d * L.w, el, L.by, L.bz};
  const Real ur[NF] = {R.d, R.d * R.u, R.d * R.v, R.d * R.w, er, R.by, R.bz};
  Real fls[NF], frs[NF];
  if (lhlld) {
    star_flux(dls, sm, vls, wls, els, byls, bzls, bn, pts, fls);
    star_flux(drs, sm, vrs, wrs, ers, byrs, bzrs, bn, pts, frs);
  } else {
    for (int n = 0; n < NF; ++n) { fls[n] = fl[n] + sl * (uls[n] - ul[n]); frs[n] = fr[n] + sr * (urs[n] - ur[n]); }
  }
  if (sls > 0.0) { for (int n = 0; n < NF; ++n) f[n] = fls[n]; return; }
  if (srs < 0.0) { for (int n = 0; n < NF; ++n) f[n] = frs[n]; return; }

  // degenerate bn -> single contact (riemann.hpp:199-207)
  const Real bsmall = 1e-12 * (ptl + ptr);
  if (bn2 <= bsmall) {
    if (sm > 0.0) { for (int n = 0; n < NF; ++n) f[n] = fls[n]; }
    else if (sm < 0.0) { for (int n = 0; n < NF; ++n) f[n] = frs[n]; }
    else { for (int n = 0; n < NF; ++n) f[n] = 0.5 * (fls[n] + frs[n]); }
    return;
  }
  // double-star states (riemann.hpp:208-219)
  const Real s = sgn(bn);
  const Real isum = 1.0 / (qdl + qdr);
  const Real vss = ((qdl * vls + qdr * vrs) + (byrs - byls) * s) * isum;
  const Real wss = ((qdl * wls + qdr * wrs) + (bzrs - bzls) * s) * isum;
  const Real byss = ((qdl * byrs + qdr * byls) + (qdl * qdr) * ((vrs - vls) * s)) * isum;
  const Real bzss = ((qdl * bzrs + qdr * bzls) + (qdl * qdr) * ((wrs - wls) * s)) * isum;
  const Real vbss = (sm * bn + vss * byss) + wss * bzss;
  const Real elss = els - qdl * ((vbls - vbss) * s);
  const Real erss = ers + qdr * ((vbrs - vbss) * s);
  const Real ulss[NF] = {dls, dls * sm, dls * vss, dls * wss, elss, byss, bzss};
  const Real urss[NF] = {drs, drs * sm, drs * vss, drs * wss, erss, byss, bzss};
  Real flss[NF], frss[NF];
  if (lhlld) {
    star_flux(dls, sm, vss, wss, elss, byss, bzss, bn, pts, flss);
    star_flux(drs, sm, vss, wss, erss, byss, bzss, bn, pts, frss);
  } else {
    for (int n = 0; n < NF; ++n) {
      flss[n] = fls[n] + sls * (ulss[n] - uls[n]);
      frss[n] = frs[n] + srs * (urss[n] - urs[n]);
    }
  }
  if (sm > 0.0) { for (int n = 0; n < NF; ++n) f[n] = flss[n]; }
  else if (sm < 0.0) { for (int n = 0; n < NF; ++n) f[n] = frss[n]; }
  else { for (int n = 0; n < NF; ++n) f[n] = 0.5 * (flss[n] + frss[n]); }
}

// ---------------------------------------------------------------- characteristic variables
// Eigensystem of the 1D primitive ideal-MHD system in W = (rho, u, v, w, p, By, Bz) with the
// normal field bx (Roe & Balsara 1996; Stone et al. 2008 App. A), eigen_mhd.hpp:15-82.
// Waves ordered u-cf, u-ca, u-cs, u, u+cs, u+ca, u+cf; R holds the right eigenvectors as
// columns, L the left ones as rows (L R = I). The velocity does not enter. Every entry is the
// reference's expression with its association.
inline void eigensystem(Real d, Real p, Real by, Real bz, Real bx, Real gamma, Real L[7][7], Real R[7][7]) {
  const Real sqrtd = std::sqrt(d);
  const Real isqrtd = 1.0 / sqrtd;
  const Real asq = gamma * p / d;
  const Real a = std::sqrt(asq);
  const Real casq = bx * bx / d;
  const Real bperpsq = by * by + bz * bz;
  const Real ctsq = bperpsq / d;
  const Real tsum = asq + casq + ctsq;
  const Real tdif = asq + ctsq - casq;
  const Real disc = std::sqrt(std::fmax(tdif * tdif + 4.0 * ctsq * casq, 0.0));
  const Real cfsq = 0.5 * (tsum + disc);
  const Real cssq = std::fmax(0.5 * (tsum - disc), 0.0);
  const Real cf = std::sqrt(cfsq), cs = std::sqrt(cssq);
  Real alf, als;                                                  // eigen_mhd.hpp:33-46
  if (cfsq - cssq <= 1e-12 * cfsq) { alf = 1.0; als = 0.0; }
  else if (ctsq <= 1e-12 * (asq + casq)) { alf = asq >= casq ? 1.0 : 0.0; als = asq >= casq ? 0.0 : 1.0; }
  else if (asq - cssq <= 0.0) { alf = 0.0; als = 1.0; }
  else if (cfsq - asq <= 0.0) { alf = 1.0; als = 0.0; }
  else {
    alf = std::sqrt((asq - cssq) / (cfsq - cssq));
    als = std::sqrt((cfsq - asq) / (cfsq - cssq));
  }
  Real bety, betz;                                                // eigen_mhd.hpp:48-55
  if (bperpsq > 0.0) {
    const Real ib = 1.0 / std::sqrt(bperpsq);
    bety = by * ib;
    betz = bz * ib;
  } else {
    bety = 1.0 / std::sqrt(2.0);
    betz = bety;
  }
  const Real s = bx >= 0.0 ? 1.0 : -1.0;
  const Real cff = cf * alf, css = cs * als;
  const Real qf = cff * s, qs = css * s;
  const Real af = a * alf * sqrtd, as = a * als * sqrtd;
  for (int r = 0; r < 7; ++r)
    for (int c = 0; c < 7; ++c) { L[r][c] = 0.0; R[r][c] = 0.0; }
  // right eigenvectors, eigen_mhd.hpp:63-70 (row = variable, column = wave)
  R[0][0] = d * alf; R[0][2] = d * als; R[0][3] = 1.0; R[0][4] = d * als; R[0][6] = d * alf;
  R[1][0] = -cff; R[1][2] = -css; R[1][4] = css; R[1][6] = cff;
  R[2][0] = qs * bety; R[2][1] = -betz; R[2][2] = -qf * bety; R[2][4] = qf * bety; R[2][5] = betz; R[2][6] = -qs * bety;
  R[3][0] = qs * betz; R[3][1] = bety; R[3][2] = -qf * betz; R[3][4] = qf * betz; R[3][5] = -bety; R[3][6] = -qs * betz;
  R[4][0] = d * alf * asq; R[4][2] = d * als * asq; R[4][4] = d * als * asq; R[4][6] = d * alf * asq;
  R[5][0] = as * bety; R[5][1] = -betz * s * sqrtd; R[5][2] = -af * bety; R[5][4] = -af * bety;
  R[5][5] = -betz * s * sqrtd; R[5][6] = as * bety;
  R[6][0] = as * betz; R[6][1] = bety * s * sqrtd; R[6][2] = -af * betz; R[6][4] = -af * betz;
  R[6][5] = bety * s * sqrtd; R[6][6] = as * betz;
  // left eigenvectors, eigen_mhd.hpp:72-81 (row = wave, column = variable)
  const Real norm = 0.5 / asq;
  const Real ncff = norm * cff, ncss = norm * css, nqf = norm * qf, nqs = norm * qs;
  const Real naf = norm * af / d, nas = norm * as / d;
  L[0][1] = -ncff; L[0][2] = nqs * bety; L[0][3] = nqs * betz; L[0][4] = norm * alf / d; L[0][5] = nas * bety; L[0][6] = nas * betz;
  L[1][2] = -0.5 * betz; L[1][3] = 0.5 * bety; L[1][5] = -0.5 * betz * s * isqrtd; L[1][6] = 0.5 * bety * s * isqrtd;
  L[2][1] = -ncss; L[2][2] = -nqf * bety; L[2][3] = -nqf * betz; L[2][4] = norm * als / d; L[2][5] = -naf * bety; L[2][6] = -naf * betz;
  L[3][0] = 1.0; L[3][4] = -1.0 / asq;
  L[4][1] = ncss; L[4][2] = nqf * bety; L[4][3] = nqf * betz; L[4][4] = norm * als / d; L[4][5] = -naf * bety; L[4][6] = -naf * betz;
  L[5][2] = 0.5 * betz; L[5][3] = -0.5 * bety; L[5][5] = -0.5 * betz * s * isqrtd; L[5][6] = 0.5 * bety * s * isqrtd;
  L[6][1] = ncff; L[6][2] = -nqs * bety; L[6][3] = -nqs * betz; L[6][4] = norm * alf / d; L[6][5] = nas * bety; L[6][6] = nas * betz;
}

// Reconstruction of the 7 primitive variables in characteristic variables for one side of a
// face, physics.hpp:110-150: project the 6-point stencils with the left eigenvectors of that
// side's cell (stencil point `cell`: 2 for the left side, 3 for the right), reconstruct each
// characteristic variable, take the requested side (left: 0, right: 1) and map back with R.
inline void charac_side(Recon rc, const Real st[7][6], int cell, Real bx, Real gamma, int side, Real out[7]) {
  Real Lm[7][7], Rm[7][7];
  eigensystem(st[0][cell], st[4][cell], st[5][cell], st[6][cell], bx, gamma, Lm, Rm);
  Real ch[7];
  for (int kk = 0; kk < 7; ++kk) {
    Real c[6];
    for (int p = 0; p < 6; ++p) {
      Real acc = 0.0;
      for (int n = 0; n < 7; ++n) acc += Lm[kk][n] * st[n][p];
      c[p] = acc;
    }
    Real cl, cr;
    reconstruct(rc, c, cl, cr);
    ch[kk] = side == 0 ? cl : cr;
  }
  for (int n = 0; n < 7; ++n) {
    Real acc = 0.0;
    for (int kk = 0; kk < 7; ++kk) acc += Rm[n][kk] * ch[kk];
    out[n] = acc;
  }
}

}  // namespace orc
