// Positivity flux limiter and source terms (FP64), for gfx950.
//
// Reference: positivity::LimitFluxes<DIR> (positivity.hpp:24-89), srcterms::ApplyGravity and
// ApplyCooling (srcterms.hpp:21-71). Same expressions and operation order as the reference.
// One thread per active face (limiter) or active cell (sources), linearised with i fastest.
#include "kernels.hpp"
#include "numerics.hpp"

namespace chq {
namespace {

constexpr int TPB = 256;

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

__device__ __forceinline__ bool gated(const SrcArgs &a) {
  return a.ctrl->active && !(a.only_if_bad && a.ctrl->nbad == 0);
}
__device__ __forceinline__ double beta_dt(const SrcArgs &a) { return a.use_ctrl_dt ? a.beta * a.ctrl->dt : a.beta; }

// CellFlux (positivity.hpp:16-22): physical flux, conserved state and fast speed of a cell
__device__ __forceinline__ void cell_flux(const St<double> &s, double bn, double sent, double gm1, double gamma,
                                          double fc[6], double uc[6], double &cf) {
  double f[NFLX], e, pt;
  Rsv<double>::pflux(s, bn, gm1, e, pt, f);
  fc[0] = f[FD]; fc[1] = f[FMU]; fc[2] = f[FMV]; fc[3] = f[FMW]; fc[4] = f[FEN]; fc[5] = f[FD] * sent;
  uc[0] = s.d; uc[1] = s.d * s.u; uc[2] = s.d * s.v; uc[3] = s.d * s.w; uc[4] = e; uc[5] = s.d * sent;
  cf = Rsv<double>::cfast(s, bn, gamma);
}

template <int DIR>
__global__ void __launch_bounds__(TPB) k_limit_fluxes(const SrcArgs a, const Box b) {
  if (!gated(a)) return;
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  int limited = 0;
  if (t < b.n()) {
    int k, j, i;
    b.at(t, k, j, i);
    constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
    constexpr int B1 = (DIR + 1) % 3, B2 = (DIR + 2) % 3;
    const Geom &g = a.g;
    const Eos &eos = a.eos;
    const long long cs = g.cstride, sd = DIR == 0 ? 1 : DIR == 1 ? g.pitch : g.plane;
    const long long c = g.at(k, j, i), cL = c - sd;
    const int ndim = 1 + (g.two_d ? 1 : 0) + (g.three_d ? 1 : 0);
    const double dx = DIR == 0 ? g.dx1 : DIR == 1 ? g.dx2 : g.dx3;
    const double lam2 = 2.0 * ndim * beta_dt(a) / dx;
    const double *w = a.w, *bcc = a.bcc;
    const St<double> L{w[IDN * cs + cL], w[IU * cs + cL], w[IV * cs + cL], w[IW * cs + cL], w[IPR * cs + cL],
                       bcc[B1 * cs + cL], bcc[B2 * cs + cL]};
    const St<double> R{w[IDN * cs + c], w[IU * cs + c], w[IV * cs + c], w[IW * cs + c], w[IPR * cs + c],
                       bcc[B1 * cs + c], bcc[B2 * cs + c]};
    const double bnL = bcc[DIR * cs + cL], bnR = bcc[DIR * cs + c];
    double fL[6], fR[6], uL[6], uR[6], cfL, cfR;
    cell_flux(L, bnL, w[IENT * cs + cL], eos.gm1, eos.gamma, fL, uL, cfL);
    cell_flux(R, bnR, w[IENT * cs + c], eos.gm1, eos.gamma, fR, uR, cfR);
    const double alpha = fmax(fabs(L.u) + cfL, fabs(R.u) + cfR);
    double *flx = a.f[DIR] + c;
    const int nv[6] = {IDN, IU, IV, IW, IEN, IENT};
    double fho[6], flf[6];
    for (int n = 0; n < 6; ++n) {
      fho[n] = flx[nv[n] * cs];
      flf[n] = 0.5 * (fL[n] + fR[n]) - 0.5 * (alpha * (uR[n] - uL[n]));
    }
    const double emL = 0.5 * (bnL * bnL + L.by * L.by + L.bz * L.bz);
    const double emR = 0.5 * (bnR * bnR + R.by * R.by + R.bz * R.bz);
    auto ok = [&](double th) {
      double f[6];
      for (int n = 0; n < 6; ++n) f[n] = flf[n] + th * (fho[n] - flf[n]);
      double pa_[5], pb_[5];
      for (int n = 0; n < 5; ++n) { pa_[n] = uL[n] - lam2 * (f[n] - fL[n]); pb_[n] = uR[n] + lam2 * (f[n] - fR[n]); }
      if (pa_[0] < eos.dfloor || pb_[0] < eos.dfloor) return false;
      const double ea = pa_[4] - 0.5 * (pa_[1] * pa_[1] + pa_[2] * pa_[2] + pa_[3] * pa_[3]) / pa_[0] - emL;
      const double eb = pb_[4] - 0.5 * (pb_[1] * pb_[1] + pb_[2] * pb_[2] + pb_[3] * pb_[3]) / pb_[0] - emR;
      double pa = eos.gm1 * ea, pb = eos.gm1 * eb;
      if (eos.dual_energy) {
        const double sa = (uL[5] - lam2 * (f[5] - fL[5])) / pa_[0], sb = (uR[5] + lam2 * (f[5] - fR[5])) / pb_[0];
        if (ea < eos.de_eta * fabs(pa_[4])) pa = sa * pow(pa_[0], eos.gamma);
        if (eb < eos.de_eta * fabs(pb_[4])) pb = sb * pow(pb_[0], eos.gamma);
      }
      return pa >= eos.pfloor && pb >= eos.pfloor;
    };
    if (!ok(1.0)) {
      double lo = 0.0, hi = 1.0;
      for (int it = 0; it < 12; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (ok(mid)) lo = mid; else hi = mid;
      }
      const double th = lo;
      for (int n = 0; n < 6; ++n) flx[nv[n] * cs] = flf[n] + th * (fho[n] - flf[n]);
      limited = 1;
    }
  }
  int sum = limited;
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if ((threadIdx.x & 63) == 0 && sum) atomicAdd((unsigned long long *)&a.ctrl->limiter_total, (unsigned long long)sum);
}

// srcterms::ApplyGravity (srcterms.hpp:21-40) over the active cells
__global__ void __launch_bounds__(TPB) k_gravity(const SrcArgs a, const Box b) {
  if (!gated(a)) return;
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= b.n()) return;
  int k, j, i;
  b.at(t, k, j, i);
  const Geom &g = a.g;
  const double bdt = beta_dt(a);
  double a1 = a.g1, a2 = a.g2, a3 = a.g3;
  if (a.grav_point) {
    const double x = a.x1min + (i - g.is + 0.5) * g.dx1, y = a.x2min + (j - g.js + 0.5) * g.dx2;
    const double z = a.x3min + (k - g.ks + 0.5) * g.dx3;
    const double dx = x - a.gx0, dy = g.two_d ? y - a.gy0 : 0.0, dz = g.three_d ? z - a.gz0 : 0.0;
    const double r2 = dx * dx + dy * dy + dz * dz + a.rsoft * a.rsoft;
    const double f = -a.gm / (r2 * sqrt(r2));
    a1 += f * dx; a2 += f * dy; a3 += f * dz;
  }
  const long long c = g.at(k, j, i), cs = g.cstride;
  const double d = a.w[IDN * cs + c];
  double *u = a.u + c;
  u[IM1 * cs] += bdt * d * a1;
  u[IM2 * cs] += bdt * d * a2;
  u[IM3 * cs] += bdt * d * a3;
  u[IEN * cs] += bdt * d * (a.w[IVX * cs + c] * a1 + a.w[IVY * cs + c] * a2 + a.w[IVZ * cs + c] * a3);
}

// srcterms::ApplyCooling (srcterms.hpp:55-71) over the active cells
__global__ void __launch_bounds__(TPB) k_cooling(const SrcArgs a, const Box b) {
  if (!gated(a)) return;
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= b.n()) return;
  int k, j, i;
  b.at(t, k, j, i);
  const Geom &g = a.g;
  const long long c = g.at(k, j, i), cs = g.cstride;
  const double dt = beta_dt(a), gm1 = a.eos.gm1;   // the step's dt in the time loop (beta = 1)
  const double d = a.w[IDN * cs + c], p = a.w[IPR * cs + c];
  const double t0 = p / d;
  const double cc = gm1 * d * a.lambda0;
  double t1;
  if (fabs(a.cool_alpha - 1.0) < 1e-12) {
    t1 = t0 * exp(-cc * dt);
  } else {
    const double e = 1.0 - a.cool_alpha;
    const double base = pow(t0, e) + (a.cool_alpha - 1.0) * cc * dt;
    t1 = base > 0.0 ? pow(base, 1.0 / e) : a.tfloor;
  }
  t1 = fmax(t1, a.tfloor);
  a.u[IEN * cs + c] += d * (t1 - t0) / gm1;
}

}  // namespace

hipError_t launch_limit_fluxes(const SrcArgs &a, int dir, hipStream_t s) {
  const Geom &g = a.g;
  const Box b = box(g.is, dir == 0 ? g.ie + 1 : g.ie, g.js, dir == 1 ? g.je + 1 : g.je, g.ks, dir == 2 ? g.ke + 1 : g.ke);
  const int nb = blocks_for(b.n());
  if (dir == 0) hipLaunchKernelGGL(k_limit_fluxes<0>, dim3(nb), dim3(TPB), 0, s, a, b);
  else if (dir == 1) hipLaunchKernelGGL(k_limit_fluxes<1>, dim3(nb), dim3(TPB), 0, s, a, b);
  else hipLaunchKernelGGL(k_limit_fluxes<2>, dim3(nb), dim3(TPB), 0, s, a, b);
  return hipGetLastError();
}

hipError_t launch_gravity(const SrcArgs &a, hipStream_t s) {
  const Geom &g = a.g;
  const Box b = box(g.is, g.ie, g.js, g.je, g.ks, g.ke);
  hipLaunchKernelGGL(k_gravity, dim3(blocks_for(b.n())), dim3(TPB), 0, s, a, b);
  return hipGetLastError();
}

hipError_t launch_cooling(const SrcArgs &a, hipStream_t s) {
  const Geom &g = a.g;
  const Box b = box(g.is, g.ie, g.js, g.je, g.ks, g.ke);
  hipLaunchKernelGGL(k_cooling, dim3(blocks_for(b.n())), dim3(TPB), 0, s, a, b);
  return hipGetLastError();
}

}  // namespace chq
