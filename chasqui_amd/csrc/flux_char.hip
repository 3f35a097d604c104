// Face fluxes with reconstruction in characteristic variables (FP64), for gfx950.
//
// Reference: the characteristic branch of kernels::Fluxes (physics.hpp:108-150) with the
// eigensystem of eigen_mhd::Eigensystem (eigen_mhd.hpp:15-82). For each face the reference
// builds, for the cell on each side, the left/right eigenvectors of that cell's primitive
// state, projects the 6-point stencils of the 7 variables onto the left eigenvectors,
// reconstructs every characteristic variable, keeps that side's value and maps it back with
// the right eigenvectors. A side's value depends only on its own cell (eigenvectors and the
// 5-cell window c-2..c+2 its reconstruction reads), so, as for the primitive kernels, each
// cell is done once and hands its two half-states to its two faces:
//   * x faces: lane = cell, the window arrives through DPP whole-wave shifts;
//   * y and z faces: lane = i, each thread marches a.seg faces along the direction with the
//     raw window in registers.
// Sums keep the reference's order (ascending variable index for the projection, ascending
// wave index for the map back); the structurally zero matrix entries are left out. The
// accumulators start at +0 as in the reference and an added product with a zero entry can
// never change such a sum, so leaving them out is exact for finite states.
// Flattening, the donor-cell fallback and the Riemann solve are the primitive kernels' code.
#include "flux_impl.hpp"

namespace chq {
namespace fluxk {
namespace {

// The matrix entries the projection and the map back need, eigen_mhd.hpp:15-81 (same
// expressions, same association). L rows: waves u-cf, u-ca, u-cs, u, u+cs, u+ca, u+cf.
struct Eig {
  double l0[7], l1[7], l2[7], l3[7], l4[7], l5[7], l6[7];   // rows of L (zeros never read)
  double r0[7], r1[7], r2[7], r3[7], r4[7], r5[7], r6[7];   // rows of R (zeros never read)
};

__device__ __forceinline__ void eigen_of(double d, double p, double by, double bz, double bx, double gamma, Eig &E) {
  const double sqrtd = sqrt(d);
  const double isqrtd = 1.0 / sqrtd;
  const double asq = gamma * p / d;
  const double a = sqrt(asq);
  const double casq = bx * bx / d;
  const double bperpsq = by * by + bz * bz;
  const double ctsq = bperpsq / d;
  const double tsum = asq + casq + ctsq;
  const double tdif = asq + ctsq - casq;
  const double disc = sqrt(fmax(tdif * tdif + 4.0 * ctsq * casq, 0.0));
  const double cfsq = 0.5 * (tsum + disc);
  const double cssq = fmax(0.5 * (tsum - disc), 0.0);
  const double cf = sqrt(cfsq), cs = sqrt(cssq);
  double alf, als;
  if (cfsq - cssq <= 1e-12 * cfsq) {
    alf = 1.0;
    als = 0.0;
  } else if (ctsq <= 1e-12 * (asq + casq)) {
    const bool acoustic = asq >= casq;
    alf = acoustic ? 1.0 : 0.0;
    als = acoustic ? 0.0 : 1.0;
  } else if (asq - cssq <= 0.0) {
    alf = 0.0;
    als = 1.0;
  } else if (cfsq - asq <= 0.0) {
    alf = 1.0;
    als = 0.0;
  } else {
    alf = sqrt((asq - cssq) / (cfsq - cssq));
    als = sqrt((cfsq - asq) / (cfsq - cssq));
  }
  double bety, betz;
  if (bperpsq > 0.0) {
    const double ib = 1.0 / sqrt(bperpsq);
    bety = by * ib;
    betz = bz * ib;
  } else {
    bety = 1.0 / sqrt(2.0);
    betz = bety;
  }
  const double s = bx >= 0.0 ? 1.0 : -1.0;
  const double cff = cf * alf, css = cs * als;
  const double qf = cff * s, qs = css * s;
  const double af = a * alf * sqrtd, as = a * als * sqrtd;
  // right eigenvectors (eigen_mhd.hpp:63-70): row = variable, column = wave
  E.r0[0] = d * alf; E.r0[2] = d * als; E.r0[3] = 1.0; E.r0[4] = d * als; E.r0[6] = d * alf;
  E.r1[0] = -cff; E.r1[2] = -css; E.r1[4] = css; E.r1[6] = cff;
  E.r2[0] = qs * bety; E.r2[1] = -betz; E.r2[2] = -qf * bety; E.r2[4] = qf * bety; E.r2[5] = betz; E.r2[6] = -qs * bety;
  E.r3[0] = qs * betz; E.r3[1] = bety; E.r3[2] = -qf * betz; E.r3[4] = qf * betz; E.r3[5] = -bety; E.r3[6] = -qs * betz;
  E.r4[0] = d * alf * asq; E.r4[2] = d * als * asq; E.r4[4] = d * als * asq; E.r4[6] = d * alf * asq;
  E.r5[0] = as * bety; E.r5[1] = -betz * s * sqrtd; E.r5[2] = -af * bety; E.r5[4] = -af * bety;
  E.r5[5] = -betz * s * sqrtd; E.r5[6] = as * bety;
  E.r6[0] = as * betz; E.r6[1] = bety * s * sqrtd; E.r6[2] = -af * betz; E.r6[4] = -af * betz;
  E.r6[5] = bety * s * sqrtd; E.r6[6] = as * betz;
  // left eigenvectors (eigen_mhd.hpp:72-81): row = wave, column = variable
  const double norm = 0.5 / asq;
  const double ncff = norm * cff, ncss = norm * css, nqf = norm * qf, nqs = norm * qs;
  const double naf = norm * af / d, nas = norm * as / d;
  E.l0[1] = -ncff; E.l0[2] = nqs * bety; E.l0[3] = nqs * betz; E.l0[4] = norm * alf / d; E.l0[5] = nas * bety; E.l0[6] = nas * betz;
  E.l1[2] = -0.5 * betz; E.l1[3] = 0.5 * bety; E.l1[5] = -0.5 * betz * s * isqrtd; E.l1[6] = 0.5 * bety * s * isqrtd;
  E.l2[1] = -ncss; E.l2[2] = -nqf * bety; E.l2[3] = -nqf * betz; E.l2[4] = norm * als / d; E.l2[5] = -naf * bety; E.l2[6] = -naf * betz;
  E.l3[0] = 1.0; E.l3[4] = -1.0 / asq;
  E.l4[1] = ncss; E.l4[2] = nqf * bety; E.l4[3] = nqf * betz; E.l4[4] = norm * als / d; E.l4[5] = -naf * bety; E.l4[6] = -naf * betz;
  E.l5[2] = 0.5 * betz; E.l5[3] = -0.5 * bety; E.l5[5] = -0.5 * betz * s * isqrtd; E.l5[6] = 0.5 * bety * s * isqrtd;
  E.l6[1] = ncff; E.l6[2] = -nqs * bety; E.l6[3] = -nqs * betz; E.l6[4] = norm * alf / d; E.l6[5] = nas * bety; E.l6[6] = nas * betz;
}

// characteristic variable KK of one stencil point: sum_n L[KK][n] q[n], ascending n, from +0
template <int KK>
__device__ __forceinline__ double project(const Eig &E, const double *q) {
  double acc = 0.0;
  if constexpr (KK == 0) {
    acc += E.l0[1] * q[1]; acc += E.l0[2] * q[2]; acc += E.l0[3] * q[3];
    acc += E.l0[4] * q[4]; acc += E.l0[5] * q[5]; acc += E.l0[6] * q[6];
  } else if constexpr (KK == 1) {
    acc += E.l1[2] * q[2]; acc += E.l1[3] * q[3]; acc += E.l1[5] * q[5]; acc += E.l1[6] * q[6];
  } else if constexpr (KK == 2) {
    acc += E.l2[1] * q[1]; acc += E.l2[2] * q[2]; acc += E.l2[3] * q[3];
    acc += E.l2[4] * q[4]; acc += E.l2[5] * q[5]; acc += E.l2[6] * q[6];
  } else if constexpr (KK == 3) {
    acc += E.l3[0] * q[0]; acc += E.l3[4] * q[4];
  } else if constexpr (KK == 4) {
    acc += E.l4[1] * q[1]; acc += E.l4[2] * q[2]; acc += E.l4[3] * q[3];
    acc += E.l4[4] * q[4]; acc += E.l4[5] * q[5]; acc += E.l4[6] * q[6];
  } else if constexpr (KK == 5) {
    acc += E.l5[2] * q[2]; acc += E.l5[3] * q[3]; acc += E.l5[5] * q[5]; acc += E.l5[6] * q[6];
  } else {
    acc += E.l6[1] * q[1]; acc += E.l6[2] * q[2]; acc += E.l6[3] * q[3];
    acc += E.l6[4] * q[4]; acc += E.l6[5] * q[5]; acc += E.l6[6] * q[6];
  }
  return acc;
}

// map back, wave KK: out[n] += R[n][KK] * ch for the nonzero entries of column KK
template <int KK>
__device__ __forceinline__ void map_back(const Eig &E, double ch, double out[7]) {
  if constexpr (KK == 0 || KK == 6) {
    out[0] += E.r0[KK] * ch; out[1] += E.r1[KK] * ch; out[2] += E.r2[KK] * ch; out[3] += E.r3[KK] * ch;
    out[4] += E.r4[KK] * ch; out[5] += E.r5[KK] * ch; out[6] += E.r6[KK] * ch;
  } else if constexpr (KK == 1 || KK == 5) {
    out[2] += E.r2[KK] * ch; out[3] += E.r3[KK] * ch; out[5] += E.r5[KK] * ch; out[6] += E.r6[KK] * ch;
  } else if constexpr (KK == 2 || KK == 4) {
    out[0] += E.r0[KK] * ch; out[1] += E.r1[KK] * ch; out[2] += E.r2[KK] * ch; out[3] += E.r3[KK] * ch;
    out[4] += E.r4[KK] * ch; out[5] += E.r5[KK] * ch; out[6] += E.r6[KK] * ch;
  } else {   // KK == 3: entropy wave, R[n][3] = (1, 0, ...)
    out[0] += E.r0[3] * ch;
  }
}

template <Recon RC, int KK>
__device__ __forceinline__ void wave_step(const Eig &E, const double (&r)[5][8], double Lf[7], double Rf[7]) {
  constexpr bool W = wide_v<RC>;
  const double c1 = project<KK>(E, r[1]), c2 = project<KK>(E, r[2]), c3 = project<KK>(E, r[3]);
  const double c0 = W ? project<KK>(E, r[0]) : c1;
  const double c4 = W ? project<KK>(E, r[4]) : c3;
  double lo, hi;
  cell_halves<double, RC>(c0, c1, c2, c3, c4, lo, hi);
  map_back<KK>(E, hi, Lf);
  map_back<KK>(E, lo, Rf);
}

// Half-states of cell c (window r[0..4] = raw primitives of c-2..c+2, slots 0..6 in the
// rotated frame) in the characteristic variables of c; bxn = cell-centred normal field of c.
// Lf = right-edge value of c (left state of face c+1/2), Rf = left-edge value (right state of
// face c-1/2). Narrow methods read r[1..3] only.
template <Recon RC>
__device__ __forceinline__ void charac_cell(const double (&r)[5][8], double bxn, double gamma, double Lf[7],
                                            double Rf[7]) {
  Eig E;
  eigen_of(r[2][0], r[2][4], r[2][5], r[2][6], bxn, gamma, E);
#pragma unroll
  for (int n = 0; n < 7; ++n) {
    Lf[n] = 0.0;
    Rf[n] = 0.0;
  }
  wave_step<RC, 0>(E, r, Lf, Rf);
  wave_step<RC, 1>(E, r, Lf, Rf);
  wave_step<RC, 2>(E, r, Lf, Rf);
  wave_step<RC, 3>(E, r, Lf, Rf);
  wave_step<RC, 4>(E, r, Lf, Rf);
  wave_step<RC, 5>(E, r, Lf, Rf);
  wave_step<RC, 6>(E, r, Lf, Rf);
}

// ---------------------------------------------------------------- x faces: lane = cell
template <Recon RC, RSolver RS, bool ENT>
__global__ void __launch_bounds__(256) k_fluxc_x(const FluxArgsT<double> a, const int nseg, const int nwaves) {
  if (!a.ctrl->active) return;
  const int wave = int(blockIdx.x) * 4 + int(threadIdx.x >> 6);
  if (wave >= nwaves) return;
  constexpr int NV = ENT ? 8 : 7;
  constexpr bool W = wide_v<RC>;
  constexpr int H = W ? 3 : 2;
  constexpr int FPW = fpw_v<RC>;
  const int lane = int(threadIdx.x) & 63;
  const Geom &g = a.g;
  // the same lane -> cell map as the primitive x kernel (rows laid end to end)
  const int nrow_j = a.ju - a.jl + 1;
  const int ncell = nseg * nrow_j * (a.ku - a.kl + 1);   // below 2^30 (the mesh size limit)
  const int v = wave * FPW - H + lane;
  const unsigned vc = unsigned(v < 0 ? 0 : (v >= ncell ? ncell - 1 : v));
  const unsigned row = fdiv(vc, a.q1), kr = fdiv(row, a.q2);   // vc / nseg, row / nrow_j
  const int ci = int(vc - row * unsigned(nseg));
  const int j = a.jl + int(row - kr * unsigned(nrow_j)), k = a.kl + int(kr);
  const long long row0 = g.at(k, j, 0);
  const double *src[NV];
  var_ptrs<double, 0, NV>(a, src);
  double r[5][8];
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const double q = src[n][row0 + ci];
    r[2][n] = q;
    r[1][n] = lane_up(q);
    r[3][n] = lane_dn(q);
    r[0][n] = W ? lane_up<2>(q) : r[1][n];
    r[4][n] = W ? lane_dn<2>(q) : r[3][n];
  }
  const double bxn = a.bcc[row0 + ci];   // component 0 = the x1 (normal) field
  double Lf[NV], Rf[NV];
  charac_cell<RC>(r, bxn, a.eos.gamma, Lf, Rf);
  if constexpr (ENT) cell_halves<double, RC>(r[0][7], r[1][7], r[2][7], r[3][7], r[4][7], Rf[7], Lf[7]);
  double fco = 0.0;
  if constexpr (W) {
    if (a.flatten) {
      const double pr[5] = {r[0][4], r[1][4], r[2][4], r[3][4], r[4][4]};
      fco = Rec<double>::flat(pr[0], pr[1], pr[2], pr[3], pr[4], r[1][1], r[3][1]);
    }
  }
  double qc[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) qc[n] = r[2][n];
  finish_cell<double, NV>(Lf, Rf, qc, fco, W && a.flatten, a.eos.dfloor, a.eos.pfloor);
  double Lface[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) Lface[n] = lane_up(Lf[n]);
  if (lane < H || lane >= H + FPW || v != int(vc) || ci < a.il || ci > a.iu) return;
  solve_store<double, 0, RS, ENT, NV>(a, row0 + ci, 1, Lface, Rf);
}

// ---------------------------------------------------------------- y/z faces: register march
template <int DIR, Recon RC, RSolver RS, bool ENT>
__global__ void __launch_bounds__(BX * BY) k_fluxc_m(const FluxArgsT<double> a) {
  if (!a.ctrl->active) return;
  constexpr int NV = ENT ? 8 : 7;
  constexpr bool W = wide_v<RC>;
  const Geom &g = a.g;
  int i, t, sg;
  if (!m_item<double, DIR>(a, (long long)blockIdx.x * (BX * BY) + int(threadIdx.y) * BX + int(threadIdx.x), i, t, sg))
    return;
  const int lo_d = DIR == 1 ? a.jl : a.kl, hi_d = DIR == 1 ? a.ju : a.ku;
  const int F0 = lo_d + sg * a.seg;
  const int F1 = min(F0 + a.seg - 1, hi_d);
  int j, k;
  if constexpr (DIR == 1) { k = a.kl + t; j = 0; }
  else { j = a.jl + t; k = 0; }
  const long long sd = DIR == 1 ? g.pitch : g.plane;
  const long long base = DIR == 1 ? g.at(k, 0, i) : g.at(0, j, i);
  const double *src[NV];
  var_ptrs<double, DIR, NV>(a, src);
  const double *bnc = a.bcc + DIR * g.cstride;   // cell-centred normal field
  // raw window of the cell being finished, c: r[m] = q_{c-2+m} (narrow methods: m = 1..3)
  double r[5][8];
  int c = F0 - 1;
  {
    const long long cc = base + (long long)c * sd;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      r[1][n] = src[n][cc - sd];
      r[2][n] = src[n][cc];
      r[3][n] = src[n][cc + sd];
      r[0][n] = W ? src[n][cc - 2 * sd] : r[1][n];
      r[4][n] = W ? src[n][cc + 2 * sd] : r[3][n];
    }
  }
  double Lprev[NV];
  for (; c <= F1; ++c) {
    const long long cc = base + (long long)c * sd;
    double Lf[NV], Rf[NV];
    charac_cell<RC>(r, bnc[cc], a.eos.gamma, Lf, Rf);
    if constexpr (ENT) cell_halves<double, RC>(r[0][7], r[1][7], r[2][7], r[3][7], r[4][7], Rf[7], Lf[7]);
    double fco = 0.0;
    if constexpr (W) {
      if (a.flatten) fco = Rec<double>::flat(r[0][4], r[1][4], r[2][4], r[3][4], r[4][4], r[1][1], r[3][1]);
    }
    double qc[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) qc[n] = r[2][n];
    finish_cell<double, NV>(Lf, Rf, qc, fco, W && a.flatten, a.eos.dfloor, a.eos.pfloor);
    if (c >= F0) solve_store<double, DIR, RS, ENT, NV>(a, cc, sd, Lprev, Rf);   // face c
#pragma unroll
    for (int n = 0; n < NV; ++n) Lprev[n] = Lf[n];
    if (c < F1) {   // slide the window by one cell
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        if constexpr (W) {
          r[0][n] = r[1][n];
          r[1][n] = r[2][n];
          r[2][n] = r[3][n];
          r[3][n] = r[4][n];
          r[4][n] = src[n][cc + 3 * sd];
        } else {
          r[1][n] = r[2][n];
          r[2][n] = r[3][n];
          r[3][n] = src[n][cc + 2 * sd];
          r[0][n] = r[1][n];
          r[4][n] = r[3][n];
        }
      }
    }
  }
}

template <int DIR, Recon RC, RSolver RS, bool ENT>
hipError_t go_fluxc(const FluxArgsT<double> &a, hipStream_t s) {
  const int nx = a.iu - a.il + 1, ny = a.ju - a.jl + 1, nz = a.ku - a.kl + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
  if constexpr (DIR == 0) {
    int nseg, nwaves;
    x_shape(a, fpw_v<RC>, nseg, nwaves);
    hipLaunchKernelGGL((k_fluxc_x<RC, RS, ENT>), dim3(div_up(nwaves, 4)), dim3(256), 0, s, a, nseg, nwaves);
  } else {
    int gx, gy, gz;
    m_shape<double, DIR>(a, gx, gy, gz);
    hipLaunchKernelGGL((k_fluxc_m<DIR, RC, RS, ENT>), dim3(gx, gy, gz), dim3(BX, BY, 1), 0, s, a);
  }
  return hipGetLastError();
}

template <int DIR, Recon RC, bool ENT>
hipError_t go_fluxc_rs(const FluxArgsT<double> &a, RSolver rs, hipStream_t s) {
  switch (rs) {
    case RSolver::hlld: return go_fluxc<DIR, RC, RSolver::hlld, ENT>(a, s);
    case RSolver::lhlld: return go_fluxc<DIR, RC, RSolver::lhlld, ENT>(a, s);
    default: return go_fluxc<DIR, RC, RSolver::hlle, ENT>(a, s);
  }
}

template <int DIR, bool ENT>
hipError_t go_fluxc_rc(const FluxArgsT<double> &a, Recon rc, RSolver rs, hipStream_t s) {
  switch (rc) {
    case Recon::plm: return go_fluxc_rs<DIR, Recon::plm, ENT>(a, rs, s);
    case Recon::ppm: return go_fluxc_rs<DIR, Recon::ppm, ENT>(a, rs, s);
    case Recon::ppmx: return go_fluxc_rs<DIR, Recon::ppmx, ENT>(a, rs, s);
    case Recon::wenoz: return go_fluxc_rs<DIR, Recon::wenoz, ENT>(a, rs, s);
    default: return hipErrorInvalidValue;   // donor cell has no characteristic form (physics.hpp:71)
  }
}

}  // namespace
}  // namespace fluxk

hipError_t launch_flux_charac(const FluxArgsT<double> &a, int dir, Recon rc, RSolver rs, bool ent, hipStream_t s) {
  using namespace fluxk;
  if (dir == 0) return ent ? go_fluxc_rc<0, true>(a, rc, rs, s) : go_fluxc_rc<0, false>(a, rc, rs, s);
  if (dir == 1) return ent ? go_fluxc_rc<1, true>(a, rc, rs, s) : go_fluxc_rc<1, false>(a, rc, rs, s);
  return ent ? go_fluxc_rc<2, true>(a, rc, rs, s) : go_fluxc_rc<2, false>(a, rc, rs, s);
}

}  // namespace chq
