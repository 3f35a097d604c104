// Explicit diffusion and RKL2 super-time-stepping kernels (FP64), for gfx950.
//
// Reference: diffusion::FaceCtx::Add / Fluxes / ResistiveEMF / TimeStep (diffusion.hpp:25-174),
// the diffusive part of fofc::RepairFluxes (fofc.hpp:90-93) and sts::Operator / Combine
// (sts.hpp:43-107). Every expression keeps the reference's operation order (bitwise parity).
// All of this is memory-light stencil work on the active box: one thread per face, edge or
// cell, linearised over the box (i fastest) so wave64 loads run along rows.
#include "kernels.hpp"

namespace chq {
namespace {

constexpr int TPB = 256;

// linear index -> (k, j, i) of the box [k0..k1] x [j0..j1] x [i0..i1]
struct Box {
  int i0, j0, k0, ni, nj, nk;
  __host__ __device__ long long n() const { return (long long)ni * nj * nk; }
  __device__ __forceinline__ void at(long long t, int &k, int &j, int &i) const {
    const long long r = t / ni;
    i = i0 + int(t - r * ni);
    j = j0 + int(r % nj);
    k = k0 + int(r / nj);
  }
};
inline Box box(int i0, int i1, int j0, int j1, int k0, int k1) { return Box{i0, j0, k0, i1 - i0 + 1, j1 - j0 + 1, k1 - k0 + 1}; }
inline int blocks_for(long long n) { return int((n + TPB - 1) / TPB); }

// current density on edges from face B, diffusion.hpp:25-47
struct Curl {
  const double *b1, *b2, *b3;
  Geom g;
  __device__ __forceinline__ double J1(int k, int j, int i) const {
    double j1 = 0.0;
    const long long c = g.at(k, j, i);
    if (g.two_d) j1 += (b3[c] - b3[c - g.pitch]) * g.idx2;
    if (g.three_d) j1 -= (b2[c] - b2[c - g.plane]) * g.idx3;
    return j1;
  }
  __device__ __forceinline__ double J2(int k, int j, int i) const {
    const long long c = g.at(k, j, i);
    double j2 = -(b3[c] - b3[c - 1]) * g.idx1;
    if (g.three_d) j2 += (b1[c] - b1[c - g.plane]) * g.idx3;
    return j2;
  }
  __device__ __forceinline__ double J3(int k, int j, int i) const {
    const long long c = g.at(k, j, i);
    double j3 = (b2[c] - b2[c - 1]) * g.idx1;
    if (g.two_d) j3 -= (b1[c] - b1[c - g.pitch]) * g.idx2;
    return j3;
  }
};

// the diffusive fluxes of face (k, j, i) of direction DIR added to f[DIR], FaceCtx::Add
// (diffusion.hpp:73-128)
template <int DIR>
__device__ __forceinline__ void diff_face(const DiffArgs &a, int k, int j, int i) {
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  constexpr int t1 = (DIR + 1) % 3, t2 = (DIR + 2) % 3;
  const Geom &g = a.g;
  const long long cs = g.cstride;
  const long long sd = DIR == 0 ? 1 : DIR == 1 ? g.pitch : g.plane;
  const long long s1 = t1 == 0 ? 1 : t1 == 1 ? g.pitch : g.plane;
  const long long s2 = t2 == 0 ? 1 : t2 == 1 ? g.pitch : g.plane;
  const double idx[3] = {g.idx1, g.idx2, g.idx3};
  const double idn = idx[DIR], idt1 = idx[t1], idt2 = idx[t2];
  const bool has[3] = {true, g.two_d != 0, g.three_d != 0};
  const bool has_t1 = has[t1], has_t2 = has[t2];
  const long long c = g.at(k, j, i), cL = c - sd;
  const double *w = a.w;
  double *flx = a.f[DIR] + c;
  if (a.nu > 0.0) {
    const double *u = w + IU * cs, *v = w + IV * cs, *ww = w + IW * cs;
    const double mu = a.nu * 0.5 * (w[IDN * cs + cL] + w[IDN * cs + c]);
    const double dun = (u[c] - u[cL]) * idn;
    const double dvn = (v[c] - v[cL]) * idn;
    const double dwn = (ww[c] - ww[cL]) * idn;
    double dut1 = 0.0, dvt1 = 0.0, dut2 = 0.0, dwt2 = 0.0;
    if (has_t1) {
      dut1 = 0.25 * idt1 * ((u[c + s1] - u[c - s1]) + (u[cL + s1] - u[cL - s1]));
      dvt1 = 0.25 * idt1 * ((v[c + s1] - v[c - s1]) + (v[cL + s1] - v[cL - s1]));
    }
    if (has_t2) {
      dut2 = 0.25 * idt2 * ((u[c + s2] - u[c - s2]) + (u[cL + s2] - u[cL - s2]));
      dwt2 = 0.25 * idt2 * ((ww[c + s2] - ww[c - s2]) + (ww[cL + s2] - ww[cL - s2]));
    }
    const double divv = dun + dvt1 + dwt2;
    const double tnn = mu * (2.0 * dun - (2.0 / 3.0) * divv);
    const double tnt1 = mu * (dvn + dut1);
    const double tnt2 = mu * (dwn + dut2);
    const double uf = 0.5 * (u[cL] + u[c]);
    const double vf = 0.5 * (v[cL] + v[c]);
    const double wf = 0.5 * (ww[cL] + ww[c]);
    flx[IU * cs] -= tnn;
    flx[IV * cs] -= tnt1;
    flx[IW * cs] -= tnt2;
    flx[IEN * cs] -= tnn * uf + tnt1 * vf + tnt2 * wf;
  }
  if (a.kappa > 0.0) {
    const double tL = w[IPR * cs + cL] / w[IDN * cs + cL], tR = w[IPR * cs + c] / w[IDN * cs + c];
    flx[IEN * cs] -= a.kappa * (tR - tL) * idn;
  }
  if (a.eta > 0.0) {
    const Curl J{a.b1, a.b2, a.b3, g};
    double et1, et2;
    if (DIR == 0) {
      et1 = a.eta * 0.5 * (J.J2(k, j, i) + (has_t2 ? J.J2(k + 1, j, i) : J.J2(k, j, i)));
      et2 = a.eta * 0.5 * (J.J3(k, j, i) + (has_t1 ? J.J3(k, j + 1, i) : J.J3(k, j, i)));
    } else if (DIR == 1) {
      et1 = a.eta * 0.5 * (J.J3(k, j, i) + J.J3(k, j, i + 1));
      et2 = a.eta * 0.5 * (J.J1(k, j, i) + (has_t1 ? J.J1(k + 1, j, i) : J.J1(k, j, i)));
    } else {
      et1 = a.eta * 0.5 * (J.J1(k, j, i) + J.J1(k, j + 1, i));
      et2 = a.eta * 0.5 * (J.J2(k, j, i) + J.J2(k, j, i + 1));
    }
    const double bt1 = 0.5 * (a.bcc[t1 * cs + cL] + a.bcc[t1 * cs + c]);
    const double bt2 = 0.5 * (a.bcc[t2 * cs + cL] + a.bcc[t2 * cs + c]);
    flx[IEN * cs] += et1 * bt2 - et2 * bt1;
  }
}

__device__ __forceinline__ bool active_cell(const Geom &g, int k, int j, int i) {
  return i >= g.is && i <= g.ie && j >= g.js && j <= g.je && k >= g.ks && k <= g.ke;
}

// the active faces of DIR (diffusion.hpp:131-137). flagged_only: the faces next to a flagged
// active cell after the fofc repair rewrote them (fofc.hpp:71-75, 90-93)
template <int DIR>
__global__ void __launch_bounds__(TPB) k_diff_flux(const DiffArgs a, const Box b, const int flagged_only) {
  if (!a.ctrl->active) return;
  if (flagged_only && a.ctrl->nbad == 0) return;
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= b.n()) return;
  int k, j, i;
  b.at(t, k, j, i);
  if (flagged_only) {
    const Geom &g = a.g;
    const int kL = k - (DIR == 2), jL = j - (DIR == 1), iL = i - (DIR == 0);
    const bool fl = active_cell(g, k, j, i) && a.flag[g.at(k, j, i)];
    const bool fr = active_cell(g, kL, jL, iL) && a.flag[g.at(kL, jL, iL)];
    if (!(fl || fr)) return;
  }
  diff_face<DIR>(a, k, j, i);
}

// eta J added to the edge EMFs over [ks..ku] x [js..ju] x [is..ie+1], diffusion.hpp:141-156
__global__ void __launch_bounds__(TPB) k_resistive_emf(const DiffArgs a, const Box b) {
  if (!a.ctrl->active) return;
  if (a.only_if_bad && a.ctrl->nbad == 0) return;
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= b.n()) return;
  int k, j, i;
  b.at(t, k, j, i);
  const Geom &g = a.g;
  const Curl J{a.b1, a.b2, a.b3, g};
  const long long c = g.at(k, j, i);
  if (i <= g.ie) a.e1[c] += a.eta * J.J1(k, j, i);
  if (j <= g.je || !g.two_d) a.e2[c] += a.eta * J.J2(k, j, i);
  if (k <= g.ke || !g.three_d) a.e3[c] += a.eta * J.J3(k, j, i);
}

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// minimum density over the active cells (positive doubles order like their bit patterns)
__global__ void __launch_bounds__(TPB) k_rho_min(const double *w, const Geom g, const Box b, Ctrl *c) {
  double m = 1e300;
  for (long long t = (long long)blockIdx.x * TPB + threadIdx.x; t < b.n(); t += (long long)gridDim.x * TPB) {
    int k, j, i;
    b.at(t, k, j, i);
    m = fmin(m, w[g.at(k, j, i)]);
  }
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0 && m < 1e300) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
    atomicMin(&c->rho_bits, bits);
  }
}
__global__ void k_rho_reset(Ctrl *c) { c->rho_bits = (unsigned long long)__double_as_longlong(1e300); }

// parabolic limit folded into the CFL minimum, diffusion.hpp:159-174 and main.cpp:275
__global__ void k_diff_dt_fold(Ctrl *c, const DiffArgs a) {
  if (!c->active) return;
  const Geom &g = a.g;
  const int ndim = 1 + (g.two_d ? 1 : 0) + (g.three_d ? 1 : 0);
  double dxmin = g.dx1;
  if (g.two_d) dxmin = fmin(dxmin, g.dx2);
  if (g.three_d) dxmin = fmin(dxmin, g.dx3);
  const double rho_min = __longlong_as_double((long long)c->rho_bits);
  const double chi = (a.kappa > 0) ? a.kappa * a.gm1 / rho_min : 0.0;
  const double dmax = fmax(fmax(a.nu, a.eta), chi);
  c->min_dt = fmin(c->min_dt, dxmin * dxmin / (2.0 * ndim * dmax));
}

__global__ void k_take_min(Ctrl *c, int slot) {
  if (!c->active) return;
  c->min_dt = fmin(__longlong_as_double((long long)c->mdt_bits[slot]), 1e300);
}

__global__ void __launch_bounds__(TPB) k_fill_zero(double *p, long long n) {
  for (long long t = (long long)blockIdx.x * TPB + threadIdx.x; t < n; t += (long long)gridDim.x * TPB) p[t] = 0.0;
}

// sts::Operator's divergence and curl (sts.hpp:65-89) over [ks..ke+1] x [js..je+1] x [is..ie+1]
__global__ void __launch_bounds__(TPB) k_sts_div(const Geom g, const Box b, const double *f1, const double *f2,
                                                  const double *f3, const double *e1, const double *e2, const double *e3,
                                                  double *mu, double *mb1, double *mb2, double *mb3) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= b.n()) return;
  int k, j, i;
  b.at(t, k, j, i);
  const long long c = g.at(k, j, i), sj = g.pitch, sk = g.plane, cs = g.cstride;
  if (i <= g.ie && j <= g.je && k <= g.ke) {
    for (int n = 0; n < NVAR; ++n) {
      const long long o = n * cs + c;
      double d = (f1[o + 1] - f1[o]) * g.idx1;
      if (g.two_d) d += (f2[o + sj] - f2[o]) * g.idx2;
      if (g.three_d) d += (f3[o + sk] - f3[o]) * g.idx3;
      mu[o] = -d;
    }
  }
  if (j <= g.je && k <= g.ke) {
    double cr = 0.0;
    if (g.two_d) cr += (e3[c + sj] - e3[c]) * g.idx2;
    if (g.three_d) cr -= (e2[c + sk] - e2[c]) * g.idx3;
    mb1[c] = -cr;
  }
  if (i <= g.ie && k <= g.ke && (g.two_d || j == g.js)) {
    double cr = -(e3[c + 1] - e3[c]) * g.idx1;
    if (g.three_d) cr += (e1[c + sk] - e1[c]) * g.idx3;
    mb2[c] = -cr;
    if (!g.two_d) mb2[c + sj] = -cr;
  }
  if (i <= g.ie && j <= g.je && (g.three_d || k == g.ks)) {
    double cr = (e2[c + 1] - e2[c]) * g.idx1;
    if (g.two_d) cr -= (e1[c + sj] - e1[c]) * g.idx2;
    mb3[c] = -cr;
    if (!g.three_d) mb3[c + sk] = -cr;
  }
}

// out = ((a x + b y) + c z) + (d p + e q), sts.hpp:93-107
struct Terms { const double *x[5]; double a[5]; };
__global__ void __launch_bounds__(TPB) k_sts_combine(const Geom g, const Box b, double *out, const Terms T, const int ncomp) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= b.n()) return;
  int k, j, i;
  b.at(t, k, j, i);
  const long long c = g.at(k, j, i);
  for (int n = 0; n < ncomp; ++n) {
    const long long o = n * g.cstride + c;
    out[o] = ((T.a[0] * T.x[0][o] + T.a[1] * T.x[1][o]) + T.a[2] * T.x[2][o]) + (T.a[3] * T.x[3][o] + T.a[4] * T.x[4][o]);
  }
}

}  // namespace

hipError_t launch_diff_flux(const DiffArgs &a, int dir, bool flagged_only, hipStream_t s) {
  const Geom &g = a.g;
  const Box b = box(g.is, dir == 0 ? g.ie + 1 : g.ie, g.js, dir == 1 ? g.je + 1 : g.je, g.ks, dir == 2 ? g.ke + 1 : g.ke);
  const int nb = blocks_for(b.n()), fo = flagged_only ? 1 : 0;
  if (dir == 0) hipLaunchKernelGGL(k_diff_flux<0>, dim3(nb), dim3(TPB), 0, s, a, b, fo);
  else if (dir == 1) hipLaunchKernelGGL(k_diff_flux<1>, dim3(nb), dim3(TPB), 0, s, a, b, fo);
  else hipLaunchKernelGGL(k_diff_flux<2>, dim3(nb), dim3(TPB), 0, s, a, b, fo);
  return hipGetLastError();
}

hipError_t launch_resistive_emf(const DiffArgs &a, hipStream_t s) {
  const Geom &g = a.g;
  const Box b = box(g.is, g.ie + 1, g.js, g.two_d ? g.je + 1 : g.js + 1, g.ks, g.three_d ? g.ke + 1 : g.ks + 1);
  hipLaunchKernelGGL(k_resistive_emf, dim3(blocks_for(b.n())), dim3(TPB), 0, s, a, b);
  return hipGetLastError();
}

hipError_t launch_rho_min(const double *w, const Geom &g, Ctrl *c, bool reset, hipStream_t s) {
  if (reset) hipLaunchKernelGGL(k_rho_reset, dim3(1), dim3(1), 0, s, c);
  const Box b = box(g.is, g.ie, g.js, g.je, g.ks, g.ke);
  const int nb = std::min(blocks_for(b.n()), 4096);
  hipLaunchKernelGGL(k_rho_min, dim3(nb), dim3(TPB), 0, s, w, g, b, c);
  return hipGetLastError();
}

hipError_t launch_diff_dt_fold(Ctrl *c, const DiffArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_diff_dt_fold, dim3(1), dim3(1), 0, s, c, a);
  return hipGetLastError();
}

hipError_t launch_ctrl_take_min(Ctrl *c, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_take_min, dim3(1), dim3(1), 0, s, c, slot);
  return hipGetLastError();
}

hipError_t launch_fill_zero(double *p, long long n, hipStream_t s) {
  const int nb = std::min(blocks_for(n), 8192);
  hipLaunchKernelGGL(k_fill_zero, dim3(nb), dim3(TPB), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_sts_div(const Geom &g, const double *f1, const double *f2, const double *f3, const double *e1,
                          const double *e2, const double *e3, double *mu, double *mb1, double *mb2, double *mb3,
                          hipStream_t s) {
  const Box b = box(g.is, g.ie + 1, g.js, g.je + 1, g.ks, g.ke + 1);
  hipLaunchKernelGGL(k_sts_div, dim3(blocks_for(b.n())), dim3(TPB), 0, s, g, b, f1, f2, f3, e1, e2, e3, mu, mb1, mb2, mb3);
  return hipGetLastError();
}

hipError_t launch_sts_combine(const Geom &g, double *out, const StsTerm t[5], int ncomp, bool faces, hipStream_t s) {
  const Box b = faces ? box(g.is, g.ie + 1, g.js, g.je + 1, g.ks, g.ke + 1) : box(g.is, g.ie, g.js, g.je, g.ks, g.ke);
  Terms T;
  for (int n = 0; n < 5; ++n) { T.x[n] = t[n].x; T.a[n] = t[n].a; }
  hipLaunchKernelGGL(k_sts_combine, dim3(blocks_for(b.n())), dim3(TPB), 0, s, g, b, out, T, ncomp);
  return hipGetLastError();
}

}  // namespace chq
