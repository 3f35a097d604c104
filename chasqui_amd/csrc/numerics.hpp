// Pointwise numerics of the face-flux kernel, compiled for gfx950 (and for the host, so the
// same code can be unit-tested there): reconstruction, shock flattening, HLLE and
// HLLD/LHLLD. Templated on the arithmetic type T (double for the FP64 drop-in, float for
// the mixed-precision path). For T = double every expression keeps the reference's
// floating-point operation order, so with -ffp-contract=off the results are bitwise equal
// to the reference method; the comments cite the reference lines each routine follows.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#include "common.hpp"

#define CHQ_FN __host__ __device__ __forceinline__

namespace chq {

// Literal thresholds of the reference, per arithmetic type. FP64 values are the
// reference's; FP32 values are what its single-precision build evaluates them to
// (recon.hpp:145-149 for the WENO epsilon; Real(1e-300) rounds to 0 in float).
template <class T> struct Lit;
template <> struct Lit<double> {
  static constexpr double tiny = 1e-300, weno_eps = 1e-36, rel = 1e-12, upwind_tol = 1e-14;
};
template <> struct Lit<float> {
  static constexpr float tiny = 0.0f, weno_eps = 1e-20f, rel = 1e-12f, upwind_tol = 1e-14f;
};

template <class T> CHQ_FN T sq(T x) { return x * x; }
template <class T> CHQ_FN T sgn(T x) { return (x > T(0)) ? T(1) : ((x < T(0)) ? T(-1) : T(0)); }
// math by type: the double forms are the C library's; float calls the single-precision forms
CHQ_FN double vmin(double a, double b) { return fmin(a, b); }
CHQ_FN double vmax(double a, double b) { return fmax(a, b); }
CHQ_FN double vabs(double a) { return fabs(a); }
CHQ_FN double vsqrt(double a) { return sqrt(a); }
CHQ_FN float vmin(float a, float b) { return fminf(a, b); }
CHQ_FN float vmax(float a, float b) { return fmaxf(a, b); }
CHQ_FN float vabs(float a) { return fabsf(a); }
CHQ_FN float vsqrt(float a) { return sqrtf(a); }

// x / 6, rounded exactly as the division is (the reconstructions divide by 6 up to eight times
// per variable and cell). Markstein's theorem: with y = RN(1/6) (relative error < 2^-53) the
// product q = RN(x * y) is within one ulp of x/6, the remainder r = x - 6q is exact in one FMA,
// and RN(q + r * y) is the correctly rounded quotient, provided nothing underflows or
// overflows: three instructions in place of the ~11 of a general FP64 division. Zeros return
// q (which keeps the sign); tiny, huge and non-finite x take the division.
CHQ_FN double div6(double x) {
  constexpr double y = 1.0 / 6.0;
  const double ax = fabs(x);
  const double q = x * y;
  if (ax >= 0x1p-900 && ax <= 0x1p1000) return fma(fma(-6.0, q, x), y, q);
  if (ax == 0.0) return q;
  return x / 6.0;
}
CHQ_FN float div6(float x) { return x / 6.0f; }

// ----------------------------------------------------------------- reconstruction
// s[0..5] = cells i-3..i+2 around face i-1/2; L = right value of cell i-1, R = left
// value of cell i (recon.hpp:3-7).
template <class T> struct Rec {
  // van Leer slope, recon.hpp:14-18
  static CHQ_FN T vl(T a, T b, T c) {
    const T l = b - a, r = c - b, p = l * r;
    return (p > T(0)) ? T(2) * p / (l + r) : T(0);
  }
  // MC slope, recon.hpp:21-27
  static CHQ_FN T mc(T a, T b, T c) {
    const T l = b - a, r = c - b;
    if (l * r <= T(0)) return T(0);
    const T m = T(0.5) * (c - a);
    const T cap = T(2) * vmin(vabs(l), vabs(r));
    return sgn(m) * vmin(vabs(m), cap);
  }
  // CW84 interface, recon.hpp:42-44
  static CHQ_FN T edge(T qa, T qb, T da, T db) { return T(0.5) * (qa + qb) + div6(da - db); }
  // CW84 monotonization, recon.hpp:47-57
  static CHQ_FN void mono(T q, T &lo, T &hi) {
    if ((hi - q) * (q - lo) <= T(0)) { lo = q; hi = q; return; }
    const T d = hi - lo, c = q - T(0.5) * (lo + hi), t = div6(d * d);
    if (d * c > t) lo = T(3) * q - T(2) * hi;
    else if (d * c < -t) hi = T(3) * q - T(2) * lo;
  }
  // PPMX interface, recon.hpp:85-100
  static CHQ_FN T xedge(T qm1, T q0, T qp1, T qp2) {
    const T C = T(1.25);
    T f = (T(7) / T(12)) * (q0 + qp1) - (T(1) / T(12)) * (qm1 + qp2);
    if ((f - q0) * (qp1 - f) < T(0)) {
      const T d2 = T(3) * ((q0 + qp1) - T(2) * f);
      const T dl = (qm1 + qp1) - T(2) * q0;
      const T dr = (q0 + qp2) - T(2) * qp1;
      T lim = T(0);
      if (d2 * dl > T(0) && d2 * dr > T(0)) lim = sgn(d2) * vmin(vmin(C * vabs(dl), C * vabs(dr)), vabs(d2));
      f = T(0.5) * (q0 + qp1) - div6(lim);
    }
    return f;
  }
  // PPMX limiter, recon.hpp:104-127
  static CHQ_FN void xmono(T qm2, T qm1, T q, T qp1, T qp2, T &lo, T &hi) {
    const T C = T(1.25);
    const T dm = q - lo, dp = hi - q;
    if ((dm * dp <= T(0)) || ((qp1 - q) * (q - qm1) <= T(0))) {
      const T d2 = T(6) * ((lo + hi) - T(2) * q);
      const T dc = (qm1 + qp1) - T(2) * q;
      const T dl = (qm2 + q) - T(2) * qm1;
      const T dr = (q + qp2) - T(2) * qp1;
      T lim = T(0);
      if (d2 * dc > T(0) && d2 * dl > T(0) && d2 * dr > T(0))
        lim = sgn(d2) * vmin(vmin(C * vabs(dl), C * vabs(dr)), vmin(C * vabs(dc), vabs(d2)));
      const T rho = (vabs(d2) > Lit<T>::rel * (vabs(q) + vabs(qm1) + vabs(qp1)) + Lit<T>::tiny) ? lim / d2 : T(0);
      if (rho < T(1) - Lit<T>::rel) { lo = q - rho * dm; hi = q + rho * dp; }
    } else {
      if (vabs(dm) >= T(2) * vabs(dp)) lo = q - T(2) * dp;
      if (vabs(dp) >= T(2) * vabs(dm)) hi = q + T(2) * dm;
    }
  }
  // WENO-Z right interface of the cell with 5-point stencil (a..e), recon.hpp:144-164
  static CHQ_FN T wz(T a, T b, T c, T d, T e) {
    const T eps = Lit<T>::weno_eps;
    const T k13 = T(13) / T(12);
    const T s0 = k13 * sq((a + c) - T(2) * b) + T(0.25) * sq((T(3) * c + a) - T(4) * b);
    const T s1 = k13 * sq((b + d) - T(2) * c) + T(0.25) * sq(b - d);
    const T s2 = k13 * sq((e + c) - T(2) * d) + T(0.25) * sq((T(3) * c + e) - T(4) * d);
    const T tau = vabs(s0 - s2);
    const T w0 = T(0.1) * (T(1) + tau / (s0 + eps));
    const T w1 = T(0.6) * (T(1) + tau / (s1 + eps));
    const T w2 = T(0.3) * (T(1) + tau / (s2 + eps));
    const T inv = T(1) / (w0 + w1 + w2);
    const T p0 = div6(T(2) * a - T(7) * b + T(11) * c);
    const T p1 = div6(-b + T(5) * c + T(2) * d);
    const T p2 = div6(T(2) * c + T(5) * d - e);
    return inv * (w0 * p0 + w1 * p1 + w2 * p2);
  }

  template <Recon R>
  static CHQ_FN void face(const T s[6], T &L, T &R_) {   // recon.hpp:189-196
    if constexpr (R == Recon::dc) { L = s[2]; R_ = s[3]; }
    else if constexpr (R == Recon::plm) {                 // recon.hpp:34-37
      L = s[2] + T(0.5) * vl(s[1], s[2], s[3]);
      R_ = s[3] - T(0.5) * vl(s[2], s[3], s[4]);
    } else if constexpr (R == Recon::ppm) {               // recon.hpp:60-78
      const T m1 = mc(s[0], s[1], s[2]), m2 = mc(s[1], s[2], s[3]);
      const T m3 = mc(s[2], s[3], s[4]), m4 = mc(s[3], s[4], s[5]);
      const T e1 = edge(s[2], s[3], m2, m3);
      T a = edge(s[1], s[2], m1, m2), b = e1;
      mono(s[2], a, b);
      T c = e1, d = edge(s[3], s[4], m3, m4);
      mono(s[3], c, d);
      L = b; R_ = c;
    } else if constexpr (R == Recon::ppmx) {              // recon.hpp:129-140
      const T e1 = xedge(s[1], s[2], s[3], s[4]);
      T a = xedge(s[0], s[1], s[2], s[3]), b = e1;
      xmono(s[0], s[1], s[2], s[3], s[4], a, b);
      T c = e1, d = xedge(s[2], s[3], s[4], s[5]);
      xmono(s[1], s[2], s[3], s[4], s[5], c, d);
      L = b; R_ = c;
    } else {                                              // recon.hpp:166-171
      L = wz(s[0], s[1], s[2], s[3], s[4]);
      R_ = wz(s[5], s[4], s[3], s[2], s[1]);
    }
  }

  // CW84 flattening coefficient of one cell, recon.hpp:179-187
  static CHQ_FN T flat(T p0, T p1, T p2, T p3, T p4, T um1, T up1) {
    (void)p2;
    const T dp = p3 - p1, dp2 = p4 - p0;
    const T pm = vmin(p1, p3);
    if (!((vabs(dp) > T(0.33) * pm) && (um1 > up1))) return T(0);
    const T r = (vabs(dp2) > Lit<T>::tiny) ? dp / dp2 : T(0);
    return vmin(T(1), vmax(T(0), T(10) * (r - T(0.75))));
  }
};

// ----------------------------------------------------------------- Riemann solvers
// Rotated frame (riemann.hpp:3-13): u normal; v, w transverse; by, bz transverse field.
template <class T> struct St { T d, u, v, w, p, by, bz; };
enum : int { FD = 0, FMU = 1, FMV = 2, FMW = 3, FEN = 4, FBY = 5, FBZ = 6, NFLX = 7 };

template <class T> struct Rsv {
  // riemann.hpp:19-25
  static CHQ_FN T cfast(const St<T> &s, T bn, T gamma) {
    const T b2 = bn * bn + s.by * s.by + s.bz * s.bz;
    const T gp = gamma * s.p;
    const T a = (gp + b2) / s.d;
    const T disc = a * a - T(4) * gp * bn * bn / (s.d * s.d);
    return vsqrt(T(0.5) * (a + vsqrt(vmax(disc, T(0)))));
  }
  // riemann.hpp:27-40
  static CHQ_FN void pflux(const St<T> &s, T bn, T gm1, T &e, T &pt, T f[NFLX]) {
    const T v2 = s.u * s.u + s.v * s.v + s.w * s.w;
    const T b2 = bn * bn + s.by * s.by + s.bz * s.bz;
    e = s.p / gm1 + T(0.5) * (s.d * v2 + b2);
    pt = s.p + T(0.5) * b2;
    const T vb = s.u * bn + s.v * s.by + s.w * s.bz;
    f[FD] = s.d * s.u;
    f[FMU] = s.d * s.u * s.u + pt - bn * bn;
    f[FMV] = s.d * s.u * s.v - bn * s.by;
    f[FMW] = s.d * s.u * s.w - bn * s.bz;
    f[FEN] = (e + pt) * s.u - bn * vb;
    f[FBY] = s.by * s.u - bn * s.v;
    f[FBZ] = s.bz * s.u - bn * s.w;
  }
  // riemann.hpp:44-51
  static CHQ_FN T cflow(const St<T> &s, T bn) {
    const T v2 = s.u * s.u + s.v * s.v + s.w * s.w;
    const T ca2 = (bn * bn + s.by * s.by + s.bz * s.bz) / s.d;
    const T cax2 = bn * bn / s.d;
    const T a = ca2 + v2;
    const T disc = a * a - T(4) * v2 * cax2;
    return vsqrt(T(0.5) * (a + vsqrt(vmax(disc, T(0)))));
  }
  // HLLE, riemann.hpp:55-68
  static CHQ_FN void hlle(const St<T> &L, const St<T> &R, T bn, T gamma, T gm1, T f[NFLX]) {
    T fl[NFLX], fr[NFLX], el, er, pl, pr;
    pflux(L, bn, gm1, el, pl, fl);
    pflux(R, bn, gm1, er, pr, fr);
    const T cl = cfast(L, bn, gamma), cr = cfast(R, bn, gamma);
    const T sl = vmin(L.u - cl, R.u - cr);
    const T sr = vmax(L.u + cl, R.u + cr);
    if (sl >= T(0)) { for (int n = 0; n < NFLX; ++n) f[n] = fl[n]; return; }
    if (sr <= T(0)) { for (int n = 0; n < NFLX; ++n) f[n] = fr[n]; return; }
    const T ul[NFLX] = {L.d, L.d * L.u, L.d * L.v, L.d * L.w, el, L.by, L.bz};
    const T ur[NFLX] = {R.d, R.d * R.u, R.d * R.v, R.d * R.w, er, R.by, R.bz};
    const T inv = T(1) / (sr - sl);
    for (int n = 0; n < NFLX; ++n) f[n] = (sr * fl[n] - sl * fr[n] + sl * sr * (ur[n] - ul[n])) * inv;
  }
  // flux of an intermediate state at total pressure pts, riemann.hpp:81-91
  static CHQ_FN void sflux(T d, T sm, T v, T w, T e, T by, T bz, T bn, T pts, T f[NFLX]) {
    const T vb = (sm * bn + v * by) + w * bz;
    f[FD] = d * sm;
    f[FMU] = (d * sm) * sm + (pts - bn * bn);
    f[FMV] = (d * sm) * v - bn * by;
    f[FMW] = (d * sm) * w - bn * bz;
    f[FEN] = (e + pts) * sm - bn * vb;
    f[FBY] = by * sm - bn * v;
    f[FBZ] = bz * sm - bn * w;
  }
  // one side's star state (riemann.hpp:127-147, its right-hand copy 149-169)
  static CHQ_FN void star(const St<T> &S, T a /*S-u*/, T bs /*S-S_M*/, T bn, T bn2, T sm, T pt, T e, T pts,
                          T &ds, T &vs, T &ws, T &bys, T &bzs, T &es, T &vbs) {
    ds = (S.d * a) / bs;
    const T den = (S.d * a) * bs - bn2;
    const T scale = vabs((S.d * a) * bs) + bn2;
    if (vabs(den) <= Lit<T>::rel * scale + Lit<T>::tiny) {
      vs = S.v; ws = S.w; bys = S.by; bzs = S.bz;
    } else {
      const T nv = (bn * (sm - S.u)) / den;
      vs = S.v - S.by * nv;
      ws = S.w - S.bz * nv;
      const T nb = ((S.d * a) * a - bn2) / den;
      bys = S.by * nb;
      bzs = S.bz * nb;
    }
    const T vb = (S.u * bn + S.v * S.by) + S.w * S.bz;
    vbs = (sm * bn + vs * bys) + ws * bzs;
    es = (((a * e - pt * S.u) + pts * sm) + bn * (vb - vbs)) / bs;
  }
  // the energy and total pressure of pflux (same expressions)
  static CHQ_FN void ept(const St<T> &s, T bn, T gm1, T &e, T &pt) {
    const T v2 = s.u * s.u + s.v * s.v + s.w * s.w;
    const T b2 = bn * bn + s.by * s.by + s.bz * s.bz;
    e = s.p / gm1 + T(0.5) * (s.d * v2 + b2);
    pt = s.p + T(0.5) * b2;
  }
  // fls = F(L) + S_L (U*_L - U_L) of one side (riemann.hpp:170-183), F(L) formed here
  static CHQ_FN void fstar_hll(const St<T> &S, T bn, T gm1, T sw, const T us[NFLX], T f[NFLX]) {
    T fs[NFLX], e, pt;
    pflux(S, bn, gm1, e, pt, fs);
    const T u[NFLX] = {S.d, S.d * S.u, S.d * S.v, S.d * S.w, e, S.by, S.bz};
    for (int n = 0; n < NFLX; ++n) f[n] = fs[n] + sw * (us[n] - u[n]);
  }
  // HLLD (Miyoshi & Kusano 2005) + LHLLD factors, riemann.hpp:96-234. Each flux is formed
  // only on the branch that returns it (same expressions, so the same values as forming all
  // of them first), which keeps fewer values live across the star-state arithmetic.
  template <bool LH>
  static CHQ_FN void hlld(const St<T> &L, const St<T> &R, T bn, T gamma, T gm1, T f[NFLX],
                          T theta = T(1), T phi = T(1)) {
    const T cl = cfast(L, bn, gamma), cr = cfast(R, bn, gamma);
    T sl, sr;
    if constexpr (LH) {
      const T cm = vmax(cl, cr);
      sl = vmin(T(0), vmin(L.u, R.u) - cm);
      sr = vmax(T(0), vmax(L.u, R.u) + cm);
    } else {
      sl = vmin(L.u - cl, R.u - cr);
      sr = vmax(L.u + cl, R.u + cr);
    }
    if (sl >= T(0)) { T e, pt; pflux(L, bn, gm1, e, pt, f); return; }
    if (sr <= T(0)) { T e, pt; pflux(R, bn, gm1, e, pt, f); return; }
    T el, er, ptl, ptr;
    ept(L, bn, gm1, el, ptl);
    ept(R, bn, gm1, er, ptr);
    const T bn2 = bn * bn;
    const T al = sl - L.u, ar = sr - R.u;
    const T ml = L.d * al, mr = R.d * ar;
    const T iden = T(1) / (mr - ml);
    const T sm = ((mr * R.u - ml * L.u) + theta * (ptl - ptr)) * iden;
    const T pts = ((mr * ptl - ml * ptr) + (phi * ((L.d * R.d) * (ar * al))) * (R.u - L.u)) * iden;
    T dls, vls, wls, byls, bzls, els, vbls;
    T drs, vrs, wrs, byrs, bzrs, ers, vbrs;
    star(L, al, sl - sm, bn, bn2, sm, ptl, el, pts, dls, vls, wls, byls, bzls, els, vbls);
    star(R, ar, sr - sm, bn, bn2, sm, ptr, er, pts, drs, vrs, wrs, byrs, bzrs, ers, vbrs);
    const T qdl = vsqrt(dls), qdr = vsqrt(drs);
    const T abn = vabs(bn);
    const T sls = sm - abn / qdl;
    const T srs = sm + abn / qdr;
    // The region of the fan that holds the face (riemann.hpp:193-233): the L* or R* state,
    // the degenerate case (either star state), or the L** or R** state; S_M = 0 (or NaN)
    // averages the two sides. Each side's flux is the same expression with L and R exchanged,
    // so a lane of one side selects its side's inputs and runs one flux evaluation: a wave whose
    // lanes lie on both sides of the contact evaluates one flux for them instead of two.
    const bool lreg = sls > T(0), rreg = !lreg && srs < T(0);
    const bool dd = !lreg && !rreg && !(bn2 <= Lit<T>::rel * (ptl + ptr));   // double-star state
    const bool two = !lreg && !rreg && !(sm > T(0)) && !(sm < T(0));
    const bool left = lreg || (!rreg && !(sm < T(0)));
    T vss = T(0), wss = T(0), byss = T(0), bzss = T(0), elss = T(0), erss = T(0);
    if (dd) {
      const T s = sgn(bn);                                         // riemann.hpp:208-233
      const T isum = T(1) / (qdl + qdr);
      vss = ((qdl * vls + qdr * vrs) + (byrs - byls) * s) * isum;
      wss = ((qdl * wls + qdr * wrs) + (bzrs - bzls) * s) * isum;
      byss = ((qdl * byrs + qdr * byls) + (qdl * qdr) * ((vrs - vls) * s)) * isum;
      bzss = ((qdl * bzrs + qdr * bzls) + (qdl * qdr) * ((wrs - wls) * s)) * isum;
      const T vbss = (sm * bn + vss * byss) + wss * bzss;
      elss = els - qdl * ((vbls - vbss) * s);
      erss = ers + qdr * ((vbrs - vbss) * s);
    }
    // one side's flux: F* = F(S) + S_S (U* - U) and, in the double-star region,
    // F** = F* + S*_S (U** - U*) (riemann.hpp:170-183, 219-229); LHLLD takes the flux of the
    // intermediate state itself (sflux)
    auto side = [&](const bool lf, T o[NFLX]) {
      const T ds = lf ? dls : drs, vs = lf ? vls : vrs, ws = lf ? wls : wrs;
      const T bys = lf ? byls : byrs, bzs = lf ? bzls : bzrs, es = lf ? els : ers;
      if constexpr (LH) {
        if (dd) sflux(ds, sm, vss, wss, lf ? elss : erss, byss, bzss, bn, pts, o);
        else sflux(ds, sm, vs, ws, es, bys, bzs, bn, pts, o);
      } else {
        const St<T> S{lf ? L.d : R.d, lf ? L.u : R.u, lf ? L.v : R.v, lf ? L.w : R.w,
                      lf ? L.p : R.p, lf ? L.by : R.by, lf ? L.bz : R.bz};
        const T us[NFLX] = {ds, ds * sm, ds * vs, ds * ws, es, bys, bzs};
        fstar_hll(S, bn, gm1, lf ? sl : sr, us, o);
        if (dd) {
          const T sw = lf ? sls : srs;
          const T uss[NFLX] = {ds, ds * sm, ds * vss, ds * wss, lf ? elss : erss, byss, bzss};
          for (int n = 0; n < NFLX; ++n) o[n] = o[n] + sw * (uss[n] - us[n]);
        }
      }
    };
    if (two) {   // both sides in one straight run, so their arithmetic interleaves
      T a[NFLX], b[NFLX];
      side(true, a);
      side(false, b);
      for (int n = 0; n < NFLX; ++n) f[n] = T(0.5) * (a[n] + b[n]);
    } else {
      side(left, f);
    }
  }
};

}  // namespace chq
