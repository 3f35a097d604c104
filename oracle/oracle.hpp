// CPU oracle — mesh, stage kernels, boundaries, con2prim, fofc, problem setups, driver.
//
// TEST INFRASTRUCTURE ONLY (parity oracle for the HIP stage pipeline; used by tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg; never by the product).
//
// A plain C++17/OpenMP restatement of the reference's per-RK-stage MHD pipeline, in the
// reference's own array layout (LayoutRight [var][k][j][i], faces/edges sized +1 in every
// dimension, mesh.hpp:37-63) so that device arrays downloaded in that layout compare 1:1.
// Every kernel keeps the reference's loop ranges and floating-point operation order and
// cites the reference file:line it follows. Min-reductions and integer counts are exact,
// so results do not depend on the OpenMP thread count.
//
// Every physics option of the reference's Stage() and time loop is restated.
// Parity unpinned (no golden vector, known answer or invariant of its own): the fofc flag count
// of a real run beyond its closed-form cases. The fofc repair (faces and values) and the aliased
// whole-domain re-update are pinned by numpy closed forms in tests/test_oracle_fofc_pins.py.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <sstream>
#include <stdexcept>
#include <fstream>
#include <string>
#include <vector>
#include <algorithm>
#include "numerics.hpp"

namespace orc {

enum { IDN = 0, IM1 = 1, IM2 = 2, IM3 = 3, IEN = 4, IENT = 5, NVAR = 6 };   // defs.hpp:16
enum { IVX = 1, IVY = 2, IVZ = 3, IPR = 4 };                               // defs.hpp:18
enum class Bc : int { periodic = 0, outflow = 1, reflect = 2 };            // defs.hpp:37

// LayoutRight arrays (defs.hpp:22-27)
struct A3 {
  int nk = 0, nj = 0, ni = 0;
  std::vector<Real> v;
  void alloc(int k, int j, int i) { nk = k; nj = j; ni = i; v.assign(size_t(k) * j * i, 0.0); }
  Real &operator()(int k, int j, int i) { return v[(size_t(k) * nj + j) * ni + i]; }
  Real operator()(int k, int j, int i) const { return v[(size_t(k) * nj + j) * ni + i]; }
};
struct A4 {
  int nn = 0, nk = 0, nj = 0, ni = 0;
  std::vector<Real> v;
  void alloc(int n, int k, int j, int i) { nn = n; nk = k; nj = j; ni = i; v.assign(size_t(n) * k * j * i, 0.0); }
  Real &operator()(int n, int k, int j, int i) { return v[((size_t(n) * nk + k) * nj + j) * ni + i]; }
  Real operator()(int n, int k, int j, int i) const { return v[((size_t(n) * nk + k) * nj + j) * ni + i]; }
};

// ---------------------------------------------------------------- parameters
// INI deck + block.key=value overrides with defaults recorded on first use (params.hpp:17-92).
class Deck {
 public:
  static std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
  }
  void parse(const std::string &text) {
    std::istringstream in(text);
    std::string line, block;
    while (std::getline(in, line)) {
      const auto h = line.find('#');
      if (h != std::string::npos) line = line.substr(0, h);
      line = trim(line);
      if (line.empty()) continue;
      if (line.front() == '[' && line.back() == ']') { block = trim(line.substr(1, line.size() - 2)); continue; }
      const auto eq = line.find('=');
      if (eq == std::string::npos) continue;
      kv_[block + "." + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
  }
  void override_kv(const std::string &arg) {
    const auto eq = arg.find('=');
    if (eq == std::string::npos) throw std::runtime_error("bad override: " + arg);
    kv_[trim(arg.substr(0, eq))] = trim(arg.substr(eq + 1));
  }
  bool has(const std::string &b, const std::string &k) const { return kv_.count(b + "." + k) > 0; }
  std::string str(const std::string &b, const std::string &k) const {
    auto it = kv_.find(b + "." + k);
    if (it == kv_.end()) throw std::runtime_error("missing parameter " + b + "." + k);
    return it->second;
  }
  std::string str(const std::string &b, const std::string &k, const std::string &d) {
    if (!has(b, k)) kv_[b + "." + k] = d;
    return kv_[b + "." + k];
  }
  double real(const std::string &b, const std::string &k, double d) {
    if (!has(b, k)) { std::ostringstream o; o.precision(17); o << d; kv_[b + "." + k] = o.str(); }
    return std::stod(kv_[b + "." + k]);
  }
  int integer(const std::string &b, const std::string &k) const { return std::stoi(str(b, k)); }
  int integer(const std::string &b, const std::string &k, int d) {
    if (!has(b, k)) kv_[b + "." + k] = std::to_string(d);
    return std::stoi(kv_[b + "." + k]);
  }
  bool boolean(const std::string &b, const std::string &k, bool d) {
    if (!has(b, k)) kv_[b + "." + k] = d ? "true" : "false";
    std::string v = kv_[b + "." + k];
    std::transform(v.begin(), v.end(), v.begin(), ::tolower);
    return v == "true" || v == "1" || v == "yes";
  }
 private:
  std::map<std::string, std::string> kv_;
};

struct Cfg {   // physics.hpp:22-33 + the driver switches of main.cpp:112-163 that the path uses
  Recon recon = Recon::ppm;
  RSolver rsolver = RSolver::hlld;
  int nstage = 3;
  Eos eos;
  Real cfl = 0.3;
  bool low_mach = false, flatten = true;
  Real lhlld_phi = -1.0;
  Bc bc[3] = {Bc::periodic, Bc::periodic, Bc::periodic};
  bool fofc = true;
  bool characteristic = false;   // physics.hpp:30
  Real nu = 0.0, kappa = 0.0, eta = 0.0;   // diffusion::Coeffs (diffusion.hpp:18-21)
  bool sts = false;                        // RKL2 super-time-stepping of the diffusive terms
  bool diffusive() const { return nu > 0.0 || kappa > 0.0 || eta > 0.0; }
  bool flux_limiter = false;               // Hu-Adams-Shu positivity flux limiter (main.cpp:141)
  // gravity (srcterms.hpp:11-17) and cooling (45-50), main.cpp:148-162
  bool grav_uniform = false, grav_point = false;
  Real g1 = 0.0, g2 = 0.0, g3 = 0.0, gm = 0.0, gx0 = 0.0, gy0 = 0.0, gz0 = 0.0, rsoft = 0.0;
  bool cooling = false;
  Real lambda0 = 0.0, cool_alpha = 1.0, tfloor = 0.0;
};

// ---------------------------------------------------------------- mesh (mesh.hpp:13-71)
struct Grid {
  int nx1 = 1, nx2 = 1, nx3 = 1, ng = 3;
  int ni = 0, nj = 0, nk = 0, is = 0, ie = 0, js = 0, je = 0, ks = 0, ke = 0;
  Real x1min = 0, x1max = 1, x2min = 0, x2max = 1, x3min = 0, x3max = 1, dx1 = 0, dx2 = 0, dx3 = 0;
  bool two_d = false, three_d = false;
  A4 u0, u1, w, bcc, ecc, f1, f2, f3;
  A3 b1f0, b2f0, b3f0, b1f1, b2f1, b3f1;
  A3 e2x1, e3x1, e1x2, e3x2, e1x3, e2x3, e1, e2, e3;
  std::vector<int> flag;

  void allocate() {
    ni = nx1 + 2 * ng;
    nj = (nx2 > 1) ? nx2 + 2 * ng : 1;
    nk = (nx3 > 1) ? nx3 + 2 * ng : 1;
    is = ng; ie = ng + nx1 - 1;
    js = (nx2 > 1) ? ng : 0; je = (nx2 > 1) ? ng + nx2 - 1 : 0;
    ks = (nx3 > 1) ? ng : 0; ke = (nx3 > 1) ? ng + nx3 - 1 : 0;
    two_d = nx2 > 1; three_d = nx3 > 1;
    dx1 = (x1max - x1min) / nx1; dx2 = (x2max - x2min) / nx2; dx3 = (x3max - x3min) / nx3;
    const int a = ni + 1, b = nj + 1, c = nk + 1;
    u0.alloc(NVAR, nk, nj, ni); u1.alloc(NVAR, nk, nj, ni); w.alloc(NVAR, nk, nj, ni);
    bcc.alloc(3, nk, nj, ni); ecc.alloc(3, nk, nj, ni);
    for (A3 *p : {&b1f0, &b2f0, &b3f0, &b1f1, &b2f1, &b3f1, &e2x1, &e3x1, &e1x2, &e3x2, &e1x3, &e2x3, &e1, &e2, &e3})
      p->alloc(c, b, a);
    f1.alloc(NVAR, c, b, a); f2.alloc(NVAR, c, b, a); f3.alloc(NVAR, c, b, a);
    flag.assign(size_t(nk) * nj * ni, 0);
  }
  int &flg(int k, int j, int i) { return flag[(size_t(k) * nj + j) * ni + i]; }
  Real x1v(int i) const { return x1min + (i - is + 0.5) * dx1; }
  Real x2v(int j) const { return x2min + (j - js + 0.5) * dx2; }
  Real x3v(int k) const { return x3min + (k - ks + 0.5) * dx3; }
  Real x1f(int i) const { return x1min + (i - is) * dx1; }
  Real x2f(int j) const { return x2min + (j - js) * dx2; }
  Real x3f(int k) const { return x3min + (k - ks) * dx3; }
  A4 &u(int s) { return s ? u1 : u0; }
  A3 &bf(int s, int d) { return s ? (d == 0 ? b1f1 : d == 1 ? b2f1 : b3f1) : (d == 0 ? b1f0 : d == 1 ? b2f0 : b3f0); }
};

// ---------------------------------------------------------------- fused face flux
// One direction of the reconstruction + Riemann kernel, physics.hpp:65-226, with the
// characteristic-variable branch of physics.hpp:110-150 (charac_side).
template <int DIR, Recon RC, RSolver RS>
void fluxes(Grid &m, const Cfg &cfg, const A3 &bn) {
  const Eos eos = cfg.eos;
  const bool flat = cfg.flatten && RC != Recon::dc && RC != Recon::plm;
  const bool charac = cfg.characteristic && RC != Recon::dc;      // physics.hpp:71
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  constexpr int B1 = (DIR + 1) % 3, B2 = (DIR + 2) % 3;
  constexpr int t1 = (DIR + 1) % 3, t2 = (DIR + 2) % 3;
  constexpr int di1 = t1 == 0, dj1 = t1 == 1, dk1 = t1 == 2, di2 = t2 == 0, dj2 = t2 == 1, dk2 = t2 == 2;
  const bool has1 = t1 == 0 || (t1 == 1 && m.two_d) || (t1 == 2 && m.three_d);
  const bool has2 = t2 == 0 || (t2 == 1 && m.two_d) || (t2 == 2 && m.three_d);
  // physics.hpp:85-88
  int il = m.is, iu = m.ie, jl = m.js, ju = m.je, kl = m.ks, ku = m.ke;
  if (DIR == 0) iu = m.ie + 1; else if (m.nx1 > 1) { il = m.is - 1; iu = m.ie + 1; }
  if (DIR == 1) ju = m.je + 1; else if (m.two_d) { jl = m.js - 1; ju = m.je + 1; }
  if (DIR == 2) ku = m.ke + 1; else if (m.three_d) { kl = m.ks - 1; ku = m.ke + 1; }
  const A4 &w = m.w, &bcc = m.bcc;
  A4 &fx = DIR == 0 ? m.f1 : DIR == 1 ? m.f2 : m.f3;
  A3 &et2 = DIR == 0 ? m.e3x1 : DIR == 1 ? m.e1x2 : m.e2x3;
  A3 &et1 = DIR == 0 ? m.e2x1 : DIR == 1 ? m.e3x2 : m.e1x3;
  const Real force_phi = cfg.lhlld_phi;
  const bool low_mach = cfg.low_mach;

#pragma omp parallel for collapse(2) schedule(static)
  for (int k = kl; k <= ku; ++k)
    for (int j = jl; j <= ju; ++j)
      for (int i = il; i <= iu; ++i) {
        // stencil loads, physics.hpp:40-57 / 99-106
        auto load = [&](const A4 &a, int n, Real q[6]) {
          if (RC == Recon::dc || RC == Recon::plm) {
            q[1] = a(n, k - 2 * dk, j - 2 * dj, i - 2 * di);
            q[2] = a(n, k - dk, j - dj, i - di);
            q[3] = a(n, k, j, i);
            q[4] = a(n, k + dk, j + dj, i + di);
            q[0] = q[1]; q[5] = q[4];
          } else {
            for (int s = 0; s < 6; ++s) q[s] = a(n, k + (s - 3) * dk, j + (s - 3) * dj, i + (s - 3) * di);
          }
        };
        Real st[7][6];
        load(w, IDN, st[0]); load(w, IU, st[1]); load(w, IV, st[2]); load(w, IW, st[3]);
        load(w, IPR, st[4]); load(bcc, B1, st[5]); load(bcc, B2, st[6]);
        Real wl[7], wr[7];
        if (!charac) {
          for (int n = 0; n < 7; ++n) reconstruct(RC, st[n], wl[n], wr[n]);   // physics.hpp:109
        } else {   // each side in the characteristic variables of its own cell (normal field bcc(DIR))
          charac_side(RC, st, 2, bcc(DIR, k - dk, j - dj, i - di), eos.gamma, 0, wl);
          charac_side(RC, st, 3, bcc(DIR, k, j, i), eos.gamma, 1, wr);
        }
        Real qe[6], sl, sr;                                                    // physics.hpp:152-154
        load(w, IENT, qe); reconstruct(RC, qe, sl, sr);
        const Real sl0 = qe[2], sr0 = qe[3];
        if (flat) {                                                            // physics.hpp:156-162
          const Real fL = flatten_coef(&st[4][0], st[1][1], st[1][3]);
          const Real fR = flatten_coef(&st[4][1], st[1][2], st[1][4]);
          if (fL > 0.0) { for (int n = 0; n < 7; ++n) wl[n] += fL * (st[n][2] - wl[n]); sl += fL * (sl0 - sl); }
          if (fR > 0.0) { for (int n = 0; n < 7; ++n) wr[n] += fR * (st[n][3] - wr[n]); sr += fR * (sr0 - sr); }
        }
        // donor-cell positivity fallback, physics.hpp:163-165
        if (wl[0] <= eos.dfloor || wl[4] <= eos.pfloor) { for (int n = 0; n < 7; ++n) wl[n] = st[n][2]; sl = sl0; }
        if (wr[0] <= eos.dfloor || wr[4] <= eos.pfloor) { for (int n = 0; n < 7; ++n) wr[n] = st[n][3]; sr = sr0; }
        Prim L{wl[0], wl[1], wl[2], wl[3], wl[4], wl[5], wl[6]};
        Prim R{wr[0], wr[1], wr[2], wr[3], wr[4], wr[5], wr[6]};
        const Real b_n = bn(k, j, i);
        if (low_mach) {                                                        // physics.hpp:170-178
          const Real cl = fast_speed(L, b_n, eos.gamma), cr = fast_speed(R, b_n, eos.gamma);
          const Real z = std::fmin(1.0, std::fmax(std::fabs(L.u) / cl, std::fabs(R.u) / cr));
          const Real um = 0.5 * (L.u + R.u), ud = 0.5 * (z * (L.u - R.u));
          L.u = um + ud; R.u = um - ud;
        }
        Real f[NF];
        if (RS == RSolver::hlld) {
          hlld(L, R, b_n, eos, f);
        } else if (RS == RSolver::lhlld) {                                     // physics.hpp:182-211
          const int kL = k - dk, jL = j - dj, iL = i - di;
          const Real cfm = std::fmax(fast_speed(L, b_n, eos.gamma), fast_speed(R, b_n, eos.gamma));
          const Real du = w(IU, k, j, i) - w(IU, kL, jL, iL);
          Real dvw = 0.0;
          if (has1) {
            const Real d1 = std::fmin(std::fmin(w(IV, kL, jL, iL) - w(IV, kL - dk1, jL - dj1, iL - di1),
                                                w(IV, kL + dk1, jL + dj1, iL + di1) - w(IV, kL, jL, iL)),
                                      std::fmin(w(IV, k, j, i) - w(IV, k - dk1, j - dj1, i - di1),
                                                w(IV, k + dk1, j + dj1, i + di1) - w(IV, k, j, i)));
            dvw = std::fmin(dvw, d1);
          }
          if (has2) {
            const Real d2 = std::fmin(std::fmin(w(IW, kL, jL, iL) - w(IW, kL - dk2, jL - dj2, iL - di2),
                                                w(IW, kL + dk2, jL + dj2, iL + di2) - w(IW, kL, jL, iL)),
                                      std::fmin(w(IW, k, j, i) - w(IW, k - dk2, j - dj2, i - di2),
                                                w(IW, k + dk2, j + dj2, i + di2) - w(IW, k, j, i)));
            dvw = std::fmin(dvw, d2);
          }
          const Real ratio = std::fmin(1.0, (-std::fmin(du, 0.0) + cfm) / (-dvw + cfm));
          const Real r2 = ratio * ratio;
          const Real theta = r2 * r2;
          const Real chi = std::fmin(1.0, std::fmax(flow_speed_u(L, b_n), flow_speed_u(R, b_n)) / cfm);
          const Real phi = (force_phi >= 0.0) ? force_phi : chi * (2.0 - chi);
          hlld(L, R, b_n, eos, f, theta, phi, true);
        } else {
          hlle(L, R, b_n, eos, f);
        }
        // flux / face-EMF store, physics.hpp:215-224
        fx(IDN, k, j, i) = f[F_D];
        fx(IU, k, j, i) = f[F_MU];
        fx(IV, k, j, i) = f[F_MV];
        fx(IW, k, j, i) = f[F_MW];
        fx(IEN, k, j, i) = f[F_E];
        fx(IENT, k, j, i) = (f[F_D] > 0.0) ? f[F_D] * sl : (f[F_D] < 0.0) ? f[F_D] * sr : 0.5 * (f[F_D] * (sl + sr));
        et2(k, j, i) = -f[F_BY];
        et1(k, j, i) = f[F_BZ];
      }
}

template <int DIR, Recon RC>
void fluxes_rs(Grid &m, const Cfg &c, const A3 &bn) {
  if (c.rsolver == RSolver::hlld) fluxes<DIR, RC, RSolver::hlld>(m, c, bn);
  else if (c.rsolver == RSolver::lhlld) fluxes<DIR, RC, RSolver::lhlld>(m, c, bn);
  else fluxes<DIR, RC, RSolver::hlle>(m, c, bn);
}
template <int DIR>
void fluxes_dir(Grid &m, const Cfg &c, const A3 &bn) {   // physics.hpp:498-514
  switch (c.recon) {
    case Recon::dc: fluxes_rs<DIR, Recon::dc>(m, c, bn); break;
    case Recon::plm: fluxes_rs<DIR, Recon::plm>(m, c, bn); break;
    case Recon::ppm: fluxes_rs<DIR, Recon::ppm>(m, c, bn); break;
    case Recon::ppmx: fluxes_rs<DIR, Recon::ppmx>(m, c, bn); break;
    default: fluxes_rs<DIR, Recon::wenoz>(m, c, bn);
  }
}
inline void fluxes_any(Grid &m, const Cfg &c, int dir, const A3 &bn) {
  if (dir == 0) fluxes_dir<0>(m, c, bn); else if (dir == 1) fluxes_dir<1>(m, c, bn); else fluxes_dir<2>(m, c, bn);
}

// ---------------------------------------------------------------- corner EMF
// upwind selection with the 1e-14 dead band, physics.hpp:233-236
inline Real upwind(Real flux, Real left, Real right) {
  const Real tol = 1e-14;
  return (flux > tol) ? left : (flux < -tol) ? right : 0.5 * (left + right);
}

// Gardiner-Stone corner EMFs, physics.hpp:242-297
inline void corner_emf(Grid &m) {
  const int is = m.is, ie = m.ie, js = m.js, je = m.je, ks = m.ks, ke = m.ke;
  A3 &e1 = m.e1, &e2 = m.e2, &e3 = m.e3;
  const A3 &e2x1 = m.e2x1, &e3x1 = m.e3x1, &e1x2 = m.e1x2, &e3x2 = m.e3x2, &e1x3 = m.e1x3, &e2x3 = m.e2x3;
  const A4 &ecc = m.ecc, &f1 = m.f1, &f2 = m.f2, &f3 = m.f3;
  if (!m.three_d) {
    for (int k = ks; k < ks + 1; ++k)
      for (int j = js; j < je + 2; ++j)
        for (int i = is; i < ie + 2; ++i) {
          e1(k, j, i) = e1x2(k, j, i); e1(k + 1, j, i) = e1x2(k, j, i);
          e2(k, j, i) = e2x1(k, j, i); e2(k + 1, j, i) = e2x1(k, j, i);
        }
    if (!m.two_d) {
      for (int i = is; i < ie + 2; ++i) { e3(ks, js, i) = e3x1(ks, js, i); e3(ks, js + 1, i) = e3x1(ks, js, i); }
      return;
    }
  } else {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = ks; k < ke + 2; ++k)
      for (int j = js; j < je + 2; ++j)
        for (int i = is; i < ie + 1; ++i) {
          const Real l3 = upwind(f2(IDN, k - 1, j, i), e1x3(k, j - 1, i) - ecc(0, k - 1, j - 1, i), e1x3(k, j, i) - ecc(0, k - 1, j, i));
          const Real r3 = upwind(f2(IDN, k, j, i), e1x3(k, j - 1, i) - ecc(0, k, j - 1, i), e1x3(k, j, i) - ecc(0, k, j, i));
          const Real l2 = upwind(f3(IDN, k, j - 1, i), e1x2(k - 1, j, i) - ecc(0, k - 1, j - 1, i), e1x2(k, j, i) - ecc(0, k, j - 1, i));
          const Real r2 = upwind(f3(IDN, k, j, i), e1x2(k - 1, j, i) - ecc(0, k - 1, j, i), e1x2(k, j, i) - ecc(0, k, j, i));
          e1(k, j, i) = 0.25 * (((l3 + r3) + (l2 + r2)) + ((e1x2(k - 1, j, i) + e1x2(k, j, i)) + (e1x3(k, j - 1, i) + e1x3(k, j, i))));
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = ks; k < ke + 2; ++k)
      for (int j = js; j < je + 1; ++j)
        for (int i = is; i < ie + 2; ++i) {
          const Real l3 = upwind(f1(IDN, k - 1, j, i), e2x3(k, j, i - 1) - ecc(1, k - 1, j, i - 1), e2x3(k, j, i) - ecc(1, k - 1, j, i));
          const Real r3 = upwind(f1(IDN, k, j, i), e2x3(k, j, i - 1) - ecc(1, k, j, i - 1), e2x3(k, j, i) - ecc(1, k, j, i));
          const Real l1 = upwind(f3(IDN, k, j, i - 1), e2x1(k - 1, j, i) - ecc(1, k - 1, j, i - 1), e2x1(k, j, i) - ecc(1, k, j, i - 1));
          const Real r1 = upwind(f3(IDN, k, j, i), e2x1(k - 1, j, i) - ecc(1, k - 1, j, i), e2x1(k, j, i) - ecc(1, k, j, i));
          e2(k, j, i) = 0.25 * (((l3 + r3) + (l1 + r1)) + ((e2x3(k, j, i - 1) + e2x3(k, j, i)) + (e2x1(k - 1, j, i) + e2x1(k, j, i))));
        }
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = ks; k < ke + 1; ++k)
    for (int j = js; j < je + 2; ++j)
      for (int i = is; i < ie + 2; ++i) {
        const Real l2 = upwind(f1(IDN, k, j - 1, i), e3x2(k, j, i - 1) - ecc(2, k, j - 1, i - 1), e3x2(k, j, i) - ecc(2, k, j - 1, i));
        const Real r2 = upwind(f1(IDN, k, j, i), e3x2(k, j, i - 1) - ecc(2, k, j, i - 1), e3x2(k, j, i) - ecc(2, k, j, i));
        const Real l1 = upwind(f2(IDN, k, j, i - 1), e3x1(k, j - 1, i) - ecc(2, k, j - 1, i - 1), e3x1(k, j, i) - ecc(2, k, j, i - 1));
        const Real r1 = upwind(f2(IDN, k, j, i), e3x1(k, j - 1, i) - ecc(2, k, j - 1, i), e3x1(k, j, i) - ecc(2, k, j, i));
        e3(k, j, i) = 0.25 * (((l2 + r2) + (l1 + r1)) + ((e3x2(k, j, i - 1) + e3x2(k, j, i)) + (e3x1(k, j - 1, i) + e3x1(k, j, i))));
      }
}

// ---------------------------------------------------------------- RK combine + CT
// physics.hpp:307-348. ua = u0 set, ub = u1 set; out may alias either (in-place is
// pointwise safe).
inline void update(Grid &m, int out, Real g0, Real g1, Real bdt) {
  const int is = m.is, ie = m.ie, js = m.js, je = m.je, ks = m.ks, ke = m.ke;
  const Real idx1 = 1.0 / m.dx1, idx2 = 1.0 / m.dx2, idx3 = 1.0 / m.dx3;
  const bool two_d = m.two_d, three_d = m.three_d;
  const A4 &f1 = m.f1, &f2 = m.f2, &f3 = m.f3;
  A4 &ua = m.u0, &ub = m.u1, &uo = m.u(out);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = ks; k <= ke; ++k)
    for (int j = js; j <= je; ++j)
      for (int i = is; i <= ie; ++i)
        for (int n = 0; n < NVAR; ++n) {
          Real dv = (f1(n, k, j, i + 1) - f1(n, k, j, i)) * idx1;
          if (two_d) dv += (f2(n, k, j + 1, i) - f2(n, k, j, i)) * idx2;
          if (three_d) dv += (f3(n, k + 1, j, i) - f3(n, k, j, i)) * idx3;
          uo(n, k, j, i) = g0 * ua(n, k, j, i) + g1 * ub(n, k, j, i) - bdt * dv;
        }
  const A3 &e1 = m.e1, &e2 = m.e2, &e3 = m.e3;
  A3 &b1a = m.b1f0, &b2a = m.b2f0, &b3a = m.b3f0, &b1b = m.b1f1, &b2b = m.b2f1, &b3b = m.b3f1;
  A3 &b1o = m.bf(out, 0), &b2o = m.bf(out, 1), &b3o = m.bf(out, 2);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = ks; k < ke + 2; ++k)
    for (int j = js; j < je + 2; ++j)
      for (int i = is; i < ie + 2; ++i) {
        if (j <= je && k <= ke) {
          Real c = 0.0;
          if (two_d) c += (e3(k, j + 1, i) - e3(k, j, i)) * idx2;
          if (three_d) c -= (e2(k + 1, j, i) - e2(k, j, i)) * idx3;
          b1o(k, j, i) = g0 * b1a(k, j, i) + g1 * b1b(k, j, i) - bdt * c;
        }
        if (i <= ie && k <= ke && (two_d || j == js)) {
          Real c = -(e3(k, j, i + 1) - e3(k, j, i)) * idx1;
          if (three_d) c += (e1(k + 1, j, i) - e1(k, j, i)) * idx3;
          b2o(k, j, i) = g0 * b2a(k, j, i) + g1 * b2b(k, j, i) - bdt * c;
          if (!two_d) b2o(k, j + 1, i) = b2o(k, j, i);
        }
        if (i <= ie && j <= je && (three_d || k == ks)) {
          Real c = (e2(k, j, i + 1) - e2(k, j, i)) * idx1;
          if (two_d) c -= (e1(k, j + 1, i) - e1(k, j, i)) * idx2;
          b3o(k, j, i) = g0 * b3a(k, j, i) + g1 * b3b(k, j, i) - bdt * c;
          if (!three_d) b3o(k + 1, j, i) = b3o(k, j, i);
        }
      }
}

// ---------------------------------------------------------------- boundaries
// physics.hpp:354-432: x1, then x2, then x3 (so edges and corners are filled).
inline void boundaries_on(Grid &m, const Cfg &cfg, A4 &u, A3 &b1, A3 &b2, A3 &b3) {
  const int ng = m.ng, nx1 = m.nx1, nx2 = m.nx2, nx3 = m.nx3;
  const int is = m.is, ie = m.ie, js = m.js, je = m.je, ks = m.ks, ke = m.ke, ni = m.ni, nj = m.nj, nk = m.nk;
  {
    const int bc = int(cfg.bc[0]);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nk + 1; ++k)
      for (int j = 0; j < nj + 1; ++j)
        for (int g = 0; g < ng; ++g) {
          const int il = g, ir = ie + 1 + g;
          const int sl = bc == 0 ? il + nx1 : bc == 1 ? is : 2 * is - 1 - il;
          const int sr = bc == 0 ? ir - nx1 : bc == 1 ? ie : 2 * ie + 1 - ir;
          const Real sg = bc == 2 ? -1.0 : 1.0;
          if (k < nk && j < nj) {
            for (int n = 0; n < NVAR; ++n) {
              const Real f = n == IM1 ? sg : 1.0;
              u(n, k, j, il) = f * u(n, k, j, sl); u(n, k, j, ir) = f * u(n, k, j, sr);
            }
            b2(k, j, il) = b2(k, j, sl); b2(k, j, ir) = b2(k, j, sr);
            b3(k, j, il) = b3(k, j, sl); b3(k, j, ir) = b3(k, j, sr);
          }
          const int fl = g, fr = ie + 2 + g;
          if (bc == 0) { b1(k, j, fl) = b1(k, j, fl + nx1); b1(k, j, fr) = b1(k, j, fr - nx1); }
          else if (bc == 1) { b1(k, j, fl) = b1(k, j, is); b1(k, j, fr) = b1(k, j, ie + 1); }
          else { b1(k, j, fl) = -b1(k, j, 2 * is - fl); b1(k, j, fr) = -b1(k, j, 2 * (ie + 1) - fr); }
          if (bc == 0 && g == 0) b1(k, j, ie + 1) = b1(k, j, is);
        }
  }
  if (m.two_d) {
    const int bc = int(cfg.bc[1]);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nk + 1; ++k)
      for (int g = 0; g < ng; ++g)
        for (int i = 0; i < ni + 1; ++i) {
          const int jl = g, jr = je + 1 + g;
          const int sl = bc == 0 ? jl + nx2 : bc == 1 ? js : 2 * js - 1 - jl;
          const int sr = bc == 0 ? jr - nx2 : bc == 1 ? je : 2 * je + 1 - jr;
          const Real sg = bc == 2 ? -1.0 : 1.0;
          if (k < nk && i < ni) {
            for (int n = 0; n < NVAR; ++n) {
              const Real f = n == IM2 ? sg : 1.0;
              u(n, k, jl, i) = f * u(n, k, sl, i); u(n, k, jr, i) = f * u(n, k, sr, i);
            }
            b3(k, jl, i) = b3(k, sl, i); b3(k, jr, i) = b3(k, sr, i);
          }
          if (k < nk) { b1(k, jl, i) = b1(k, sl, i); b1(k, jr, i) = b1(k, sr, i); }
          const int fl = g, fr = je + 2 + g;
          if (bc == 0) { b2(k, fl, i) = b2(k, fl + nx2, i); b2(k, fr, i) = b2(k, fr - nx2, i); }
          else if (bc == 1) { b2(k, fl, i) = b2(k, js, i); b2(k, fr, i) = b2(k, je + 1, i); }
          else { b2(k, fl, i) = -b2(k, 2 * js - fl, i); b2(k, fr, i) = -b2(k, 2 * (je + 1) - fr, i); }
          if (bc == 0 && g == 0) b2(k, je + 1, i) = b2(k, js, i);
        }
  }
  if (m.three_d) {
    const int bc = int(cfg.bc[2]);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nj + 1; ++j)
      for (int g = 0; g < ng; ++g)
        for (int i = 0; i < ni + 1; ++i) {
          const int kl = g, kr = ke + 1 + g;
          const int sl = bc == 0 ? kl + nx3 : bc == 1 ? ks : 2 * ks - 1 - kl;
          const int sr = bc == 0 ? kr - nx3 : bc == 1 ? ke : 2 * ke + 1 - kr;
          const Real sg = bc == 2 ? -1.0 : 1.0;
          if (j < nj && i < ni) {
            for (int n = 0; n < NVAR; ++n) {
              const Real f = n == IM3 ? sg : 1.0;
              u(n, kl, j, i) = f * u(n, sl, j, i); u(n, kr, j, i) = f * u(n, sr, j, i);
            }
          }
          b1(kl, j, i) = b1(sl, j, i); b1(kr, j, i) = b1(sr, j, i);
          b2(kl, j, i) = b2(sl, j, i); b2(kr, j, i) = b2(sr, j, i);
          const int fl = g, fr = ke + 2 + g;
          if (bc == 0) { b3(fl, j, i) = b3(fl + nx3, j, i); b3(fr, j, i) = b3(fr - nx3, j, i); }
          else if (bc == 1) { b3(fl, j, i) = b3(ks, j, i); b3(fr, j, i) = b3(ke + 1, j, i); }
          else { b3(fl, j, i) = -b3(2 * ks - fl, j, i); b3(fr, j, i) = -b3(2 * (ke + 1) - fr, j, i); }
          if (bc == 0 && g == 0) b3(ke + 1, j, i) = b3(ks, j, i);
        }
  }
}
inline void boundaries(Grid &m, const Cfg &cfg, int s) { boundaries_on(m, cfg, m.u(s), m.bf(s, 0), m.bf(s, 1), m.bf(s, 2)); }

// ---------------------------------------------------------------- con2prim + dt
// physics.hpp:437-494. Rewrites u[IENT] (and u[IEN] with dual energy) in place.
inline Real cons_to_prim_on(Grid &m, const Cfg &cfg, A4 &u, const A3 &b1, const A3 &b2, const A3 &b3,
                            bool compute_dt) {
  const Eos eos = cfg.eos;
  const int is = m.is, ie = m.ie, js = m.js, je = m.je, ks = m.ks, ke = m.ke;
  const bool two_d = m.two_d, three_d = m.three_d;
  const Real dx1 = m.dx1, dx2 = m.dx2, dx3 = m.dx3;
  A4 &w = m.w, &bcc = m.bcc, &ecc = m.ecc;
  Real mdt = 1e300;
#pragma omp parallel for collapse(2) schedule(static) reduction(min : mdt)
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        const Real bx = 0.5 * (b1(k, j, i) + b1(k, j, i + 1));
        const Real by = two_d ? 0.5 * (b2(k, j, i) + b2(k, j + 1, i)) : b2(k, j, i);
        const Real bz = three_d ? 0.5 * (b3(k, j, i) + b3(k + 1, j, i)) : b3(k, j, i);
        const Real d = std::fmax(u(IDN, k, j, i), eos.dfloor);
        const Real id = 1.0 / d;
        const Real vx = u(IM1, k, j, i) * id, vy = u(IM2, k, j, i) * id, vz = u(IM3, k, j, i) * id;
        const Real bsq = bx * bx + by * by + bz * bz;
        const Real kin = 0.5 * d * (vx * vx + vy * vy + vz * vz);
        const Real et = u(IEN, k, j, i);
        const Real ei = et - kin - 0.5 * bsq;
        Real p = eos.gm1 * ei;
        Real se = u(IENT, k, j, i) * id;
        if (eos.dual_energy && ei < eos.de_eta * std::fabs(et)) {
          p = se * std::pow(d, eos.gamma);
          p = std::fmax(p, eos.pfloor);
          u(IEN, k, j, i) = kin + 0.5 * bsq + p / eos.gm1;
        } else {
          p = std::fmax(p, eos.pfloor);
          se = p / std::pow(d, eos.gamma);
          u(IENT, k, j, i) = d * se;
        }
        w(IDN, k, j, i) = d; w(IVX, k, j, i) = vx; w(IVY, k, j, i) = vy; w(IVZ, k, j, i) = vz; w(IPR, k, j, i) = p;
        w(IENT, k, j, i) = se;
        bcc(0, k, j, i) = bx; bcc(1, k, j, i) = by; bcc(2, k, j, i) = bz;
        ecc(0, k, j, i) = vz * by - vy * bz;
        ecc(1, k, j, i) = vx * bz - vz * bx;
        ecc(2, k, j, i) = vy * bx - vx * by;
        if (compute_dt && i >= is && i <= ie && j >= js && j <= je && k >= ks && k <= ke) {
          const Real gp = eos.gamma * p;
          const Real a = (gp + bsq) * id;
          auto cf = [&](Real b) {
            const Real disc = a * a - 4.0 * gp * b * b * id * id;
            return std::sqrt(0.5 * (a + std::sqrt(std::fmax(disc, 0.0))));
          };
          Real dt = dx1 / (std::fabs(vx) + cf(bx));
          if (two_d) dt = std::fmin(dt, dx2 / (std::fabs(vy) + cf(by)));
          if (three_d) dt = std::fmin(dt, dx3 / (std::fabs(vz) + cf(bz)));
          mdt = std::fmin(mdt, dt);
        }
      }
  return mdt;
}
inline Real cons_to_prim(Grid &m, const Cfg &cfg, int s, bool compute_dt) {
  return cons_to_prim_on(m, cfg, m.u(s), m.bf(s, 0), m.bf(s, 1), m.bf(s, 2), compute_dt);
}

// ---------------------------------------------------------------- explicit diffusion
// Viscosity, conduction and Ohmic resistivity added to the fluxes / edge EMFs of the same
// stage (diffusion.hpp:1-176). Coefficients: nu, kappa, eta (a zero coefficient skips its term).
struct Diff {
  Real nu = 0.0, kappa = 0.0, eta = 0.0;
  bool any() const { return nu > 0.0 || kappa > 0.0 || eta > 0.0; }
};
inline Diff diff_of(const Cfg &c) { return Diff{c.nu, c.kappa, c.eta}; }

// current density on edges from the face fields, diffusion.hpp:25-47
struct Curl {
  const A3 &b1, &b2, &b3;
  Real idx1, idx2, idx3;
  bool two_d, three_d;
  Real J1(int k, int j, int i) const {
    Real j1 = 0.0;
    if (two_d) j1 += (b3(k, j, i) - b3(k, j - 1, i)) * idx2;
    if (three_d) j1 -= (b2(k, j, i) - b2(k - 1, j, i)) * idx3;
    return j1;
  }
  Real J2(int k, int j, int i) const {
    Real j2 = -(b3(k, j, i) - b3(k, j, i - 1)) * idx1;
    if (three_d) j2 += (b1(k, j, i) - b1(k - 1, j, i)) * idx3;
    return j2;
  }
  Real J3(int k, int j, int i) const {
    Real j3 = (b2(k, j, i) - b2(k, j, i - 1)) * idx1;
    if (two_d) j3 -= (b1(k, j, i) - b1(k, j - 1, i)) * idx2;
    return j3;
  }
};
inline Curl curl_of(const Grid &m, const A3 &b1, const A3 &b2, const A3 &b3) {
  return Curl{b1, b2, b3, 1.0 / m.dx1, 1.0 / m.dx2, 1.0 / m.dx3, m.two_d, m.three_d};
}

// diffusive fluxes of one face (k, j, i) of direction DIR added to f{DIR+1},
// FaceCtx::Add (diffusion.hpp:73-128)
template <int DIR>
void diff_face(Grid &m, const Diff &c, const Curl &J, int k, int j, int i) {
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  constexpr int t1 = (DIR + 1) % 3, t2 = (DIR + 2) % 3;
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  constexpr int di1 = t1 == 0, dj1 = t1 == 1, dk1 = t1 == 2, di2 = t2 == 0, dj2 = t2 == 1, dk2 = t2 == 2;
  const Real idx[3] = {1.0 / m.dx1, 1.0 / m.dx2, 1.0 / m.dx3};
  const Real idn = idx[DIR], idt1 = idx[t1], idt2 = idx[t2];
  const bool has[3] = {true, m.two_d, m.three_d};
  const bool has_t1 = has[t1], has_t2 = has[t2];
  const A4 &w = m.w, &bcc = m.bcc;
  A4 &flx = DIR == 0 ? m.f1 : DIR == 1 ? m.f2 : m.f3;
  const int kL = k - dk, jL = j - dj, iL = i - di;
  if (c.nu > 0.0) {
    const Real mu = c.nu * 0.5 * (w(IDN, kL, jL, iL) + w(IDN, k, j, i));
    const Real dun = (w(IU, k, j, i) - w(IU, kL, jL, iL)) * idn;
    const Real dvn = (w(IV, k, j, i) - w(IV, kL, jL, iL)) * idn;
    const Real dwn = (w(IW, k, j, i) - w(IW, kL, jL, iL)) * idn;
    Real dut1 = 0.0, dvt1 = 0.0, dut2 = 0.0, dwt2 = 0.0;
    if (has_t1) {
      dut1 = 0.25 * idt1 * ((w(IU, k + dk1, j + dj1, i + di1) - w(IU, k - dk1, j - dj1, i - di1))
                            + (w(IU, kL + dk1, jL + dj1, iL + di1) - w(IU, kL - dk1, jL - dj1, iL - di1)));
      dvt1 = 0.25 * idt1 * ((w(IV, k + dk1, j + dj1, i + di1) - w(IV, k - dk1, j - dj1, i - di1))
                            + (w(IV, kL + dk1, jL + dj1, iL + di1) - w(IV, kL - dk1, jL - dj1, iL - di1)));
    }
    if (has_t2) {
      dut2 = 0.25 * idt2 * ((w(IU, k + dk2, j + dj2, i + di2) - w(IU, k - dk2, j - dj2, i - di2))
                            + (w(IU, kL + dk2, jL + dj2, iL + di2) - w(IU, kL - dk2, jL - dj2, iL - di2)));
      dwt2 = 0.25 * idt2 * ((w(IW, k + dk2, j + dj2, i + di2) - w(IW, k - dk2, j - dj2, i - di2))
                            + (w(IW, kL + dk2, jL + dj2, iL + di2) - w(IW, kL - dk2, jL - dj2, iL - di2)));
    }
    const Real divv = dun + dvt1 + dwt2;
    const Real tnn = mu * (2.0 * dun - (2.0 / 3.0) * divv);
    const Real tnt1 = mu * (dvn + dut1);
    const Real tnt2 = mu * (dwn + dut2);
    const Real uf = 0.5 * (w(IU, kL, jL, iL) + w(IU, k, j, i));
    const Real vf = 0.5 * (w(IV, kL, jL, iL) + w(IV, k, j, i));
    const Real wf = 0.5 * (w(IW, kL, jL, iL) + w(IW, k, j, i));
    flx(IU, k, j, i) -= tnn;
    flx(IV, k, j, i) -= tnt1;
    flx(IW, k, j, i) -= tnt2;
    flx(IEN, k, j, i) -= tnn * uf + tnt1 * vf + tnt2 * wf;
  }
  if (c.kappa > 0.0) {
    const Real tL = w(IPR, kL, jL, iL) / w(IDN, kL, jL, iL), tR = w(IPR, k, j, i) / w(IDN, k, j, i);
    flx(IEN, k, j, i) -= c.kappa * (tR - tL) * idn;
  }
  if (c.eta > 0.0) {
    Real et1, et2;
    if (DIR == 0) {
      et1 = c.eta * 0.5 * (J.J2(k, j, i) + (has_t2 ? J.J2(k + 1, j, i) : J.J2(k, j, i)));
      et2 = c.eta * 0.5 * (J.J3(k, j, i) + (has_t1 ? J.J3(k, j + 1, i) : J.J3(k, j, i)));
    } else if (DIR == 1) {
      et1 = c.eta * 0.5 * (J.J3(k, j, i) + J.J3(k, j, i + 1));
      et2 = c.eta * 0.5 * (J.J1(k, j, i) + (has_t1 ? J.J1(k + 1, j, i) : J.J1(k, j, i)));
    } else {
      et1 = c.eta * 0.5 * (J.J1(k, j, i) + J.J1(k, j + 1, i));
      et2 = c.eta * 0.5 * (J.J2(k, j, i) + J.J2(k, j, i + 1));
    }
    const Real bt1 = 0.5 * (bcc(t1, kL, jL, iL) + bcc(t1, k, j, i));
    const Real bt2 = 0.5 * (bcc(t2, kL, jL, iL) + bcc(t2, k, j, i));
    flx(IEN, k, j, i) += et1 * bt2 - et2 * bt1;
  }
}

// the face range of the diffusive fluxes: active faces only (diffusion.hpp:131-137)
template <int DIR>
bool diff_face_active(const Grid &m, int k, int j, int i) {
  if (DIR == 0) return i >= m.is && i <= m.ie + 1 && j >= m.js && j <= m.je && k >= m.ks && k <= m.ke;
  if (DIR == 1) return j >= m.js && j <= m.je + 1 && i >= m.is && i <= m.ie && k >= m.ks && k <= m.ke;
  return k >= m.ks && k <= m.ke + 1 && i >= m.is && i <= m.ie && j >= m.js && j <= m.je;
}

template <int DIR>
void diffusion_fluxes(Grid &m, const Diff &c, const A3 &b1, const A3 &b2, const A3 &b3) {   // diffusion.hpp:131-138
  if (!c.any()) return;
  const Curl J = curl_of(m, b1, b2, b3);
  const int iu = DIR == 0 ? m.ie + 1 : m.ie, ju = DIR == 1 ? m.je + 1 : m.je, ku = DIR == 2 ? m.ke + 1 : m.ke;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = m.ks; k <= ku; ++k)
    for (int j = m.js; j <= ju; ++j)
      for (int i = m.is; i <= iu; ++i) diff_face<DIR>(m, c, J, k, j, i);
}
inline void diffusion_fluxes_all(Grid &m, const Diff &c, const A3 &b1, const A3 &b2, const A3 &b3) {
  diffusion_fluxes<0>(m, c, b1, b2, b3);
  if (m.two_d) diffusion_fluxes<1>(m, c, b1, b2, b3);
  if (m.three_d) diffusion_fluxes<2>(m, c, b1, b2, b3);
}

// eta J added to the edge EMFs, diffusion.hpp:141-156
inline void resistive_emf(Grid &m, const Diff &c, const A3 &b1, const A3 &b2, const A3 &b3) {
  if (c.eta <= 0.0) return;
  const Curl J = curl_of(m, b1, b2, b3);
  const int ku = m.three_d ? m.ke + 1 : m.ks + 1, ju = m.two_d ? m.je + 1 : m.js + 1;
  const int ie = m.ie, je = m.je, ke = m.ke;
  const bool two_d = m.two_d, three_d = m.three_d;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = m.ks; k <= ku; ++k)
    for (int j = m.js; j <= ju; ++j)
      for (int i = m.is; i <= ie + 1; ++i) {
        if (i <= ie) m.e1(k, j, i) += c.eta * J.J1(k, j, i);
        if (j <= je || !two_d) m.e2(k, j, i) += c.eta * J.J2(k, j, i);
        if (k <= ke || !three_d) m.e3(k, j, i) += c.eta * J.J3(k, j, i);
      }
}

// parabolic time-step limit over the active cells, diffusion.hpp:159-174
inline Real diffusion_dt(const Grid &m, const Cfg &cfg, const Diff &c) {
  if (!c.any()) return 1e300;
  const int ndim = 1 + (m.two_d ? 1 : 0) + (m.three_d ? 1 : 0);
  Real dxmin = m.dx1;
  if (m.two_d) dxmin = std::fmin(dxmin, m.dx2);
  if (m.three_d) dxmin = std::fmin(dxmin, m.dx3);
  Real rho_min = 1e300;
  for (int k = m.ks; k <= m.ke; ++k)
    for (int j = m.js; j <= m.je; ++j)
      for (int i = m.is; i <= m.ie; ++i) rho_min = std::fmin(rho_min, m.w(IDN, k, j, i));
  const Real chi = (c.kappa > 0) ? c.kappa * cfg.eos.gm1 / rho_min : 0.0;
  const Real dmax = std::fmax(std::fmax(c.nu, c.eta), chi);
  return dxmin * dxmin / (2.0 * ndim * dmax);
}

// ---------------------------------------------------------------- positivity flux limiter
// positivity::LimitFluxes<DIR> (positivity.hpp:24-89): each active face's high-order flux is
// blended toward the Lax-Friedrichs flux with the largest theta (12-step bisection) for which
// both one-sided half-updates keep rho >= dfloor and p >= pfloor (B held fixed).
inline void cell_flux(const Prim &s, Real bn, Real sent, Real gm1, Real gamma, Real fc[6], Real uc[6], Real &cf) {
  Real f[NF], e, pt;
  phys_flux(s, bn, gm1, e, pt, f);
  fc[0] = f[F_D]; fc[1] = f[F_MU]; fc[2] = f[F_MV]; fc[3] = f[F_MW]; fc[4] = f[F_E]; fc[5] = f[F_D] * sent;
  uc[0] = s.d; uc[1] = s.d * s.u; uc[2] = s.d * s.v; uc[3] = s.d * s.w; uc[4] = e; uc[5] = s.d * sent;
  cf = fast_speed(s, bn, gamma);
}

template <int DIR>
long limit_fluxes(Grid &m, const Cfg &cfg, Real beta_dt) {
  const Eos eos = cfg.eos;
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  constexpr int B1 = (DIR + 1) % 3, B2 = (DIR + 2) % 3;
  const int iu = DIR == 0 ? m.ie + 1 : m.ie, ju = DIR == 1 ? m.je + 1 : m.je, ku = DIR == 2 ? m.ke + 1 : m.ke;
  const int ndim = 1 + (m.two_d ? 1 : 0) + (m.three_d ? 1 : 0);
  const Real dx = DIR == 0 ? m.dx1 : DIR == 1 ? m.dx2 : m.dx3;
  const Real lam2 = 2.0 * ndim * beta_dt / dx;
  const A4 &w = m.w, &bcc = m.bcc;
  A4 &flx = DIR == 0 ? m.f1 : DIR == 1 ? m.f2 : m.f3;
  const int nv[6] = {IDN, IU, IV, IW, IEN, IENT};
  long count = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : count)
  for (int k = m.ks; k <= ku; ++k)
    for (int j = m.js; j <= ju; ++j)
      for (int i = m.is; i <= iu; ++i) {
        const int kL = k - dk, jL = j - dj, iL = i - di;
        const Prim L{w(IDN, kL, jL, iL), w(IU, kL, jL, iL), w(IV, kL, jL, iL), w(IW, kL, jL, iL), w(IPR, kL, jL, iL),
                     bcc(B1, kL, jL, iL), bcc(B2, kL, jL, iL)};
        const Prim R{w(IDN, k, j, i), w(IU, k, j, i), w(IV, k, j, i), w(IW, k, j, i), w(IPR, k, j, i),
                     bcc(B1, k, j, i), bcc(B2, k, j, i)};
        const Real bnL = bcc(DIR, kL, jL, iL), bnR = bcc(DIR, k, j, i);
        Real fL[6], fR[6], uL[6], uR[6], cfL, cfR;
        cell_flux(L, bnL, w(IENT, kL, jL, iL), eos.gm1, eos.gamma, fL, uL, cfL);
        cell_flux(R, bnR, w(IENT, k, j, i), eos.gm1, eos.gamma, fR, uR, cfR);
        const Real alpha = std::fmax(std::fabs(L.u) + cfL, std::fabs(R.u) + cfR);
        Real fho[6], flf[6];
        for (int n = 0; n < 6; ++n) {
          fho[n] = flx(nv[n], k, j, i);
          flf[n] = 0.5 * (fL[n] + fR[n]) - 0.5 * (alpha * (uR[n] - uL[n]));
        }
        const Real emL = 0.5 * (bnL * bnL + L.by * L.by + L.bz * L.bz);
        const Real emR = 0.5 * (bnR * bnR + R.by * R.by + R.bz * R.bz);
        auto ok = [&](Real th) {
          Real f[6];
          for (int n = 0; n < 6; ++n) f[n] = flf[n] + th * (fho[n] - flf[n]);
          Real a[5], b[5];
          for (int n = 0; n < 5; ++n) { a[n] = uL[n] - lam2 * (f[n] - fL[n]); b[n] = uR[n] + lam2 * (f[n] - fR[n]); }
          if (a[0] < eos.dfloor || b[0] < eos.dfloor) return false;
          const Real ea = a[4] - 0.5 * (a[1] * a[1] + a[2] * a[2] + a[3] * a[3]) / a[0] - emL;
          const Real eb = b[4] - 0.5 * (b[1] * b[1] + b[2] * b[2] + b[3] * b[3]) / b[0] - emR;
          Real pa = eos.gm1 * ea, pb = eos.gm1 * eb;
          if (eos.dual_energy) {
            const Real sa = (uL[5] - lam2 * (f[5] - fL[5])) / a[0], sb = (uR[5] + lam2 * (f[5] - fR[5])) / b[0];
            if (ea < eos.de_eta * std::fabs(a[4])) pa = sa * std::pow(a[0], eos.gamma);
            if (eb < eos.de_eta * std::fabs(b[4])) pb = sb * std::pow(b[0], eos.gamma);
          }
          return pa >= eos.pfloor && pb >= eos.pfloor;
        };
        if (ok(1.0)) continue;
        Real lo = 0.0, hi = 1.0;
        for (int it = 0; it < 12; ++it) {
          const Real mid = 0.5 * (lo + hi);
          if (ok(mid)) lo = mid; else hi = mid;
        }
        const Real th = lo;
        for (int n = 0; n < 6; ++n) flx(nv[n], k, j, i) = flf[n] + th * (fho[n] - flf[n]);
        ++count;
      }
  return count;
}

// ---------------------------------------------------------------- source terms
// srcterms::ApplyGravity (srcterms.hpp:21-40): momentum and energy sources with beta*dt from
// the stage primitives, uniform field and/or softened point mass
inline void apply_gravity(Grid &m, const Cfg &c, A4 &u, Real beta_dt) {
  if (!(c.grav_uniform || c.grav_point)) return;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = m.ks; k <= m.ke; ++k)
    for (int j = m.js; j <= m.je; ++j)
      for (int i = m.is; i <= m.ie; ++i) {
        Real a1 = c.g1, a2 = c.g2, a3 = c.g3;
        if (c.grav_point) {
          const Real dx = m.x1v(i) - c.gx0, dy = m.two_d ? m.x2v(j) - c.gy0 : 0.0, dz = m.three_d ? m.x3v(k) - c.gz0 : 0.0;
          const Real r2 = dx * dx + dy * dy + dz * dz + c.rsoft * c.rsoft;
          const Real f = -c.gm / (r2 * std::sqrt(r2));
          a1 += f * dx; a2 += f * dy; a3 += f * dz;
        }
        const Real d = m.w(IDN, k, j, i);
        u(IM1, k, j, i) += beta_dt * d * a1;
        u(IM2, k, j, i) += beta_dt * d * a2;
        u(IM3, k, j, i) += beta_dt * d * a3;
        u(IEN, k, j, i) += beta_dt * d * (m.w(IVX, k, j, i) * a1 + m.w(IVY, k, j, i) * a2 + m.w(IVZ, k, j, i) * a3);
      }
}

// srcterms::ApplyCooling (srcterms.hpp:55-71): exact power-law cooling of T = p / rho over dt
inline void apply_cooling(Grid &m, const Cfg &c, A4 &u, Real dt) {
  if (!c.cooling) return;
  const Real gm1 = c.eos.gm1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = m.ks; k <= m.ke; ++k)
    for (int j = m.js; j <= m.je; ++j)
      for (int i = m.is; i <= m.ie; ++i) {
        const Real d = m.w(IDN, k, j, i), p = m.w(IPR, k, j, i);
        const Real t0 = p / d;
        const Real cc = gm1 * d * c.lambda0;
        Real t1;
        if (std::fabs(c.cool_alpha - 1.0) < 1e-12) {
          t1 = t0 * std::exp(-cc * dt);
        } else {
          const Real e = 1.0 - c.cool_alpha;
          const Real base = std::pow(t0, e) + (c.cool_alpha - 1.0) * cc * dt;
          t1 = base > 0.0 ? std::pow(base, 1.0 / e) : c.tfloor;
        }
        t1 = std::fmax(t1, c.tfloor);
        u(IEN, k, j, i) += d * (t1 - t0) / gm1;
      }
}

// ---------------------------------------------------------------- fofc
// flag cells whose tentative rho or p is at/below the floors, fofc.hpp:17-44
inline int fofc_flag(Grid &m, const Cfg &cfg, int s) {
  const Eos eos = cfg.eos;
  const bool two_d = m.two_d, three_d = m.three_d;
  const A4 &u = m.u(s);
  const A3 &b1 = m.bf(s, 0), &b2 = m.bf(s, 1), &b3 = m.bf(s, 2);
  int nbad = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : nbad)
  for (int k = m.ks; k <= m.ke; ++k)
    for (int j = m.js; j <= m.je; ++j)
      for (int i = m.is; i <= m.ie; ++i) {
        const Real bx = 0.5 * (b1(k, j, i) + b1(k, j, i + 1));
        const Real by = two_d ? 0.5 * (b2(k, j, i) + b2(k, j + 1, i)) : b2(k, j, i);
        const Real bz = three_d ? 0.5 * (b3(k, j, i) + b3(k + 1, j, i)) : b3(k, j, i);
        const Real d = u(IDN, k, j, i);
        bool bad = d <= eos.dfloor;
        if (!bad) {
          const Real kin = 0.5 * (sqr(u(IM1, k, j, i)) + sqr(u(IM2, k, j, i)) + sqr(u(IM3, k, j, i))) / d;
          const Real ei = u(IEN, k, j, i) - kin - 0.5 * (bx * bx + by * by + bz * bz);
          Real p = eos.gm1 * ei;
          if (eos.dual_energy && ei < eos.de_eta * std::fabs(u(IEN, k, j, i)))
            p = (u(IENT, k, j, i) / d) * std::pow(d, eos.gamma);
          bad = p <= eos.pfloor;
        }
        m.flg(k, j, i) = bad ? 1 : 0;
        if (bad) ++nbad;
      }
  return nbad;
}

// donor-cell HLLE fluxes on faces touching a flagged active cell, fofc.hpp:46-96
// (diffusion is out of scope, so the diffusive re-add of fofc.hpp:89-93 never applies)
// fofc.hpp:46-96; with diffusion coefficients dc the repaired active faces get their
// diffusive fluxes again (dc is empty under super-time-stepping, main.cpp:82)
template <int DIR>
void fofc_repair(Grid &m, const Cfg &cfg, const A3 &bn, const Diff &dc, const A3 &b1c, const A3 &b2c, const A3 &b3c) {
  const Eos eos = cfg.eos;
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  constexpr int IU = IVX + DIR, IV = IVX + (DIR + 1) % 3, IW = IVX + (DIR + 2) % 3;
  constexpr int B1 = (DIR + 1) % 3, B2 = (DIR + 2) % 3;
  int il = m.is, iu = m.ie, jl = m.js, ju = m.je, kl = m.ks, ku = m.ke;
  if (DIR == 0) iu = m.ie + 1; else if (m.nx1 > 1) { il = m.is - 1; iu = m.ie + 1; }
  if (DIR == 1) ju = m.je + 1; else if (m.two_d) { jl = m.js - 1; ju = m.je + 1; }
  if (DIR == 2) ku = m.ke + 1; else if (m.three_d) { kl = m.ks - 1; ku = m.ke + 1; }
  const int is = m.is, ie = m.ie, js = m.js, je = m.je, ks = m.ks, ke = m.ke;
  const A4 &w = m.w, &bcc = m.bcc;
  A4 &fx = DIR == 0 ? m.f1 : DIR == 1 ? m.f2 : m.f3;
  A3 &et2 = DIR == 0 ? m.e3x1 : DIR == 1 ? m.e1x2 : m.e2x3;
  A3 &et1 = DIR == 0 ? m.e2x1 : DIR == 1 ? m.e3x2 : m.e1x3;
  const bool diff = dc.any();
  const Curl J = curl_of(m, b1c, b2c, b3c);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = kl; k <= ku; ++k)
    for (int j = jl; j <= ju; ++j)
      for (int i = il; i <= iu; ++i) {
        const int kL = k - dk, jL = j - dj, iL = i - di;
        auto inside = [&](int kk, int jj, int ii) { return ii >= is && ii <= ie && jj >= js && jj <= je && kk >= ks && kk <= ke; };
        const bool fl = inside(k, j, i) && m.flg(k, j, i);
        const bool fr = inside(kL, jL, iL) && m.flg(kL, jL, iL);
        if (!(fl || fr)) continue;
        Prim L{w(IDN, kL, jL, iL), w(IU, kL, jL, iL), w(IV, kL, jL, iL), w(IW, kL, jL, iL), w(IPR, kL, jL, iL),
               bcc(B1, kL, jL, iL), bcc(B2, kL, jL, iL)};
        Prim R{w(IDN, k, j, i), w(IU, k, j, i), w(IV, k, j, i), w(IW, k, j, i), w(IPR, k, j, i), bcc(B1, k, j, i),
               bcc(B2, k, j, i)};
        Real f[NF];
        hlle(L, R, bn(k, j, i), eos, f);
        fx(IDN, k, j, i) = f[F_D]; fx(IU, k, j, i) = f[F_MU]; fx(IV, k, j, i) = f[F_MV];
        fx(IW, k, j, i) = f[F_MW]; fx(IEN, k, j, i) = f[F_E];
        fx(IENT, k, j, i) = (f[F_D] > 0.0) ? f[F_D] * w(IENT, kL, jL, iL) : (f[F_D] < 0.0) ? f[F_D] * w(IENT, k, j, i)
                                           : 0.5 * (f[F_D] * (w(IENT, kL, jL, iL) + w(IENT, k, j, i)));
        et2(k, j, i) = -f[F_BY];
        et1(k, j, i) = f[F_BZ];
        if (diff && diff_face_active<DIR>(m, k, j, i)) diff_face<DIR>(m, dc, J, k, j, i);
      }
}
inline void fofc_repair_any(Grid &m, const Cfg &c, int dir, int cur, const Diff &dc) {
  const A3 &bn = m.bf(cur, dir), &b1 = m.bf(cur, 0), &b2 = m.bf(cur, 1), &b3 = m.bf(cur, 2);
  if (dir == 0) fofc_repair<0>(m, c, bn, dc, b1, b2, b3);
  else if (dir == 1) fofc_repair<1>(m, c, bn, dc, b1, b2, b3);
  else fofc_repair<2>(m, c, bn, dc, b1, b2, b3);
}

// ---------------------------------------------------------------- RK stage and step
// RK coefficient table (g0, g1, beta, cur, out), main.cpp:250-268
struct StageCoef { Real g0, g1, beta; int cur, out; };
inline StageCoef stage_coef(int nstage, int s) {
  if (nstage == 1) return {1, 0, 1, 0, 0};
  if (nstage == 2) return s == 1 ? StageCoef{1, 0, 1, 0, 1} : StageCoef{0.5, 0.5, 0.5, 1, 0};
  if (nstage == 3) {
    if (s == 1) return {1, 0, 1, 0, 1};
    if (s == 2) return {0.75, 0.25, 0.25, 1, 1};
    return {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 1, 0};
  }
  if (s == 1) return {1, 0, 0.5, 0, 1};
  if (s == 2) return {0, 1, 0.5, 1, 1};
  if (s == 3) return {2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 1, 1};
  return {0, 1, 0.5, 1, 0};
}

// Stage(), main.cpp:46-93 (explicit diffusion in the stage unless super-time-stepping);
// *limited (may be null) accumulates the flux-limiter activations
inline long stage(Grid &m, const Cfg &cfg, int cur, int out, Real g0, Real g1, Real bdt, long *limited = nullptr) {
  const A3 &b1c = m.bf(cur, 0), &b2c = m.bf(cur, 1), &b3c = m.bf(cur, 2);
  const Diff dc = cfg.sts ? Diff{} : diff_of(cfg);
  fluxes_any(m, cfg, 0, b1c);
  if (m.two_d) fluxes_any(m, cfg, 1, b2c);
  if (m.three_d) fluxes_any(m, cfg, 2, b3c);
  diffusion_fluxes_all(m, dc, b1c, b2c, b3c);                      // main.cpp:55-59
  if (cfg.flux_limiter) {                                          // main.cpp:60-64
    long n = limit_fluxes<0>(m, cfg, bdt);
    if (m.two_d) n += limit_fluxes<1>(m, cfg, bdt);
    if (m.three_d) n += limit_fluxes<2>(m, cfg, bdt);
    if (limited) *limited += n;
  }
  corner_emf(m);
  resistive_emf(m, dc, b1c, b2c, b3c);                             // main.cpp:66
  update(m, out, g0, g1, bdt);
  apply_gravity(m, cfg, m.u(out), bdt);                            // main.cpp:76
  long nbad = 0;
  if (cfg.fofc) {
    nbad = fofc_flag(m, cfg, out);
    if (nbad > 0) {
      fofc_repair_any(m, cfg, 0, cur, dc);
      if (m.two_d) fofc_repair_any(m, cfg, 1, cur, dc);
      if (m.three_d) fofc_repair_any(m, cfg, 2, cur, dc);
      corner_emf(m);
      resistive_emf(m, dc, b1c, b2c, b3c);
      update(m, out, g0, g1, bdt);
      apply_gravity(m, cfg, m.u(out), bdt);                        // main.cpp:89
    }
  }
  boundaries(m, cfg, out);
  return nbad;
}

// ---------------------------------------------------------------- RKL2 super-time-stepping
// sts.hpp:1-157 (Meyer, Balsara & Aslam 2014): the parabolic operator M(U) = -div F_diff,
// M(B) = -curl(eta J) advanced over the whole hyperbolic step tau with s Runge-Kutta-Legendre
// stages after the RK step (operator split).
struct StsWork {   // sts.hpp:16-32
  bool allocated = false;
  A4 y0, y1, y2, m0, mj;
  A3 b0[3], b1[3], b2[3], mb0[3], mbj[3];
  void allocate(const Grid &m) {
    if (allocated) return;
    for (A4 *p : {&y0, &y1, &y2, &m0, &mj}) p->alloc(NVAR, m.nk, m.nj, m.ni);
    for (int c = 0; c < 3; ++c)
      for (A3 *p : {&b0[c], &b1[c], &b2[c], &mb0[c], &mbj[c]}) p->alloc(m.nk + 1, m.nj + 1, m.ni + 1);
    allocated = true;
  }
};

inline int sts_stages(Real tau, Real dt_par) {   // sts.hpp:35-39
  const double r = tau / dt_par;
  const int s = static_cast<int>(std::ceil(0.5 * (-1.0 + std::sqrt(9.0 + 16.0 * r))));
  return s < 2 ? 2 : s;
}
inline double sts_b(int j) { return j <= 2 ? 1.0 / 3.0 : (double(j) * j + j - 2.0) / (2.0 * j * (j + 1.0)); }

// M(U, B) with the primitives already in m.w / m.bcc: -div F_diff into mu, -curl(eta J) into
// mb (sts.hpp:43-90)
inline void sts_operator(Grid &m, const Diff &dc, const A3 &b1c, const A3 &b2c, const A3 &b3c, A4 &mu, A3 &mb1,
                         A3 &mb2, A3 &mb3) {
  std::fill(m.f1.v.begin(), m.f1.v.end(), 0.0);
  std::fill(m.f2.v.begin(), m.f2.v.end(), 0.0);
  std::fill(m.f3.v.begin(), m.f3.v.end(), 0.0);
  std::fill(m.e1.v.begin(), m.e1.v.end(), 0.0);
  std::fill(m.e2.v.begin(), m.e2.v.end(), 0.0);
  std::fill(m.e3.v.begin(), m.e3.v.end(), 0.0);
  diffusion_fluxes_all(m, dc, b1c, b2c, b3c);
  resistive_emf(m, dc, b1c, b2c, b3c);
  const int is = m.is, ie = m.ie, js = m.js, je = m.je, ks = m.ks, ke = m.ke;
  const bool two_d = m.two_d, three_d = m.three_d;
  const Real idx1 = 1.0 / m.dx1, idx2 = 1.0 / m.dx2, idx3 = 1.0 / m.dx3;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = ks; k <= ke + 1; ++k)
    for (int j = js; j <= je + 1; ++j)
      for (int i = is; i <= ie + 1; ++i) {
        if (i <= ie && j <= je && k <= ke) {
          for (int n = 0; n < NVAR; ++n) {
            Real d = (m.f1(n, k, j, i + 1) - m.f1(n, k, j, i)) * idx1;
            if (two_d) d += (m.f2(n, k, j + 1, i) - m.f2(n, k, j, i)) * idx2;
            if (three_d) d += (m.f3(n, k + 1, j, i) - m.f3(n, k, j, i)) * idx3;
            mu(n, k, j, i) = -d;
          }
        }
        if (j <= je && k <= ke) {
          Real c = 0.0;
          if (two_d) c += (m.e3(k, j + 1, i) - m.e3(k, j, i)) * idx2;
          if (three_d) c -= (m.e2(k + 1, j, i) - m.e2(k, j, i)) * idx3;
          mb1(k, j, i) = -c;
        }
        if (i <= ie && k <= ke && (two_d || j == js)) {
          Real c = -(m.e3(k, j, i + 1) - m.e3(k, j, i)) * idx1;
          if (three_d) c += (m.e1(k + 1, j, i) - m.e1(k, j, i)) * idx3;
          mb2(k, j, i) = -c;
          if (!two_d) mb2(k, j + 1, i) = -c;
        }
        if (i <= ie && j <= je && (three_d || k == ks)) {
          Real c = (m.e2(k, j, i + 1) - m.e2(k, j, i)) * idx1;
          if (two_d) c -= (m.e1(k, j + 1, i) - m.e1(k, j, i)) * idx2;
          mb3(k, j, i) = -c;
          if (!three_d) mb3(k + 1, j, i) = -c;
        }
      }
}

// out = ((a x + b y) + c z) + (d p + e q) on the active cells (sts.hpp:93-100) and on the face
// box [ks..ke+1] x [js..je+1] x [is..ie+1] (sts.hpp:101-107)
inline void sts_combine(const Grid &m, A4 &out, Real a, const A4 &x, Real b, const A4 &y, Real c, const A4 &z, Real d,
                        const A4 &p, Real e, const A4 &q) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = m.ks; k <= m.ke; ++k)
    for (int j = m.js; j <= m.je; ++j)
      for (int i = m.is; i <= m.ie; ++i)
        for (int n = 0; n < NVAR; ++n)
          out(n, k, j, i) = ((a * x(n, k, j, i) + b * y(n, k, j, i)) + c * z(n, k, j, i)) +
                            (d * p(n, k, j, i) + e * q(n, k, j, i));
}
inline void sts_combine_b(const Grid &m, A3 &out, Real a, const A3 &x, Real b, const A3 &y, Real c, const A3 &z, Real d,
                          const A3 &p, Real e, const A3 &q) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = m.ks; k <= m.ke + 1; ++k)
    for (int j = m.js; j <= m.je + 1; ++j)
      for (int i = m.is; i <= m.ie + 1; ++i)
        out(k, j, i) = ((a * x(k, j, i) + b * y(k, j, i)) + c * z(k, j, i)) + (d * p(k, j, i) + e * q(k, j, i));
}

// sts::Advance (sts.hpp:111-157): the diffusive terms over tau from (u0, b?f0), result left
// there; u1 / b?f1 serve as the fourth rotating array. Returns the number of stages.
inline int sts_advance(Grid &m, const Cfg &cfg, const Diff &dc, StsWork &wk, Real tau, Real dt_par) {
  wk.allocate(m);
  const int s = sts_stages(tau, dt_par);
  const double w1 = 4.0 / (double(s) * s + s - 2.0);
  wk.y0 = m.u0;
  wk.b0[0] = m.b1f0; wk.b0[1] = m.b2f0; wk.b0[2] = m.b3f0;
  cons_to_prim_on(m, cfg, wk.y0, wk.b0[0], wk.b0[1], wk.b0[2], false);
  sts_operator(m, dc, wk.b0[0], wk.b0[1], wk.b0[2], wk.m0, wk.mb0[0], wk.mb0[1], wk.mb0[2]);
  const Real mt1 = Real(sts_b(1) * w1) * tau;
  sts_combine(m, wk.y1, 1, wk.y0, 0, wk.y0, 0, wk.y0, mt1, wk.m0, 0, wk.m0);
  for (int c = 0; c < 3; ++c) sts_combine_b(m, wk.b1[c], 1, wk.b0[c], 0, wk.b0[c], 0, wk.b0[c], mt1, wk.mb0[c], 0, wk.mb0[c]);
  boundaries_on(m, cfg, wk.y1, wk.b1[0], wk.b1[1], wk.b1[2]);
  A4 *yjm2 = &wk.y0, *yjm1 = &wk.y1, *yj = &wk.y2;
  A3 *bjm2[3] = {&wk.b0[0], &wk.b0[1], &wk.b0[2]}, *bjm1[3] = {&wk.b1[0], &wk.b1[1], &wk.b1[2]};
  A3 *bjn[3] = {&wk.b2[0], &wk.b2[1], &wk.b2[2]};
  A4 *spare = &m.u1;
  A3 *spareb[3] = {&m.b1f1, &m.b2f1, &m.b3f1};
  for (int j = 2; j <= s; ++j) {
    const double b_j = sts_b(j), b_jm1 = sts_b(j - 1), b_jm2 = sts_b(j - 2);
    const double mu = (2.0 * j - 1.0) / j * b_j / b_jm1;
    const double nu = -(j - 1.0) / j * b_j / b_jm2;
    const double mut = mu * w1;
    const double gt = -(1.0 - b_jm1) * mut;
    cons_to_prim_on(m, cfg, *yjm1, *bjm1[0], *bjm1[1], *bjm1[2], false);
    sts_operator(m, dc, *bjm1[0], *bjm1[1], *bjm1[2], wk.mj, wk.mbj[0], wk.mbj[1], wk.mbj[2]);
    sts_combine(m, *yj, Real(mu), *yjm1, Real(nu), *yjm2, Real(1.0 - mu - nu), wk.y0, Real(mut) * tau, wk.mj,
                Real(gt) * tau, wk.m0);
    for (int c = 0; c < 3; ++c)
      sts_combine_b(m, *bjn[c], Real(mu), *bjm1[c], Real(nu), *bjm2[c], Real(1.0 - mu - nu), wk.b0[c],
                    Real(mut) * tau, wk.mbj[c], Real(gt) * tau, wk.mb0[c]);
    boundaries_on(m, cfg, *yj, *bjn[0], *bjn[1], *bjn[2]);
    A4 *freed = j == 2 ? spare : yjm2;
    A3 *freedb[3] = {j == 2 ? spareb[0] : bjm2[0], j == 2 ? spareb[1] : bjm2[1], j == 2 ? spareb[2] : bjm2[2]};
    yjm2 = yjm1; yjm1 = yj; yj = freed;
    for (int c = 0; c < 3; ++c) { bjm2[c] = bjm1[c]; bjm1[c] = bjn[c]; bjn[c] = freedb[c]; }
  }
  m.u0 = *yjm1;
  m.b1f0 = *bjm1[0]; m.b2f0 = *bjm1[1]; m.b3f0 = *bjm1[2];
  return s;
}

// ---------------------------------------------------------------- problem setups
constexpr double kPi = 3.14159265358979323846;

struct Init { A4 w; A3 b1, b2, b3; };

// B = curl A on faces from a potential sampled on edges, pgen.hpp:30-50
template <class F>
void from_potential(const Grid &m, Init &h, F A) {
  const Real idx = 1.0 / m.dx1, idy = 1.0 / m.dx2, idz = 1.0 / m.dx3;
  for (int k = 0; k < m.nk + 1; ++k)
    for (int j = 0; j < m.nj + 1; ++j)
      for (int i = 0; i < m.ni + 1; ++i) {
        const Real xf = m.x1f(i), yf = m.x2f(j), zf = m.x3f(k);
        const Real xv = m.x1v(i), yv = m.x2v(j), zv = m.x3v(k);
        Real b1 = 0.0;
        if (m.two_d) b1 += (A(2, xf, m.x2f(j + 1), zv) - A(2, xf, yf, zv)) * idy;
        if (m.three_d) b1 -= (A(1, xf, yv, m.x3f(k + 1)) - A(1, xf, yv, zf)) * idz;
        Real b2 = -(A(2, m.x1f(i + 1), yf, zv) - A(2, xf, yf, zv)) * idx;
        if (m.three_d) b2 += (A(0, xv, yf, m.x3f(k + 1)) - A(0, xv, yf, zf)) * idz;
        Real b3 = (A(1, m.x1f(i + 1), yv, zf) - A(1, xf, yv, zf)) * idx;
        if (m.two_d) b3 -= (A(0, xv, m.x2f(j + 1), zf) - A(0, xv, yf, zf)) * idy;
        h.b1(k, j, i) = b1; h.b2(k, j, i) = b2; h.b3(k, j, i) = b3;
      }
}

// Circularly polarized Alfven wave, pgen.hpp:54-91. Extension (no reference counterpart):
// problem.angle2 = elevation of the propagation direction out of the x-y plane (degrees);
// with angle2 = 0 the reference's own expressions are used unchanged.
inline void pg_cpaw(Deck &p, const Grid &m, Init &h) {
  const Real dens = p.real("problem", "dens", 1.0);
  const Real pres = p.real("problem", "pres", 0.1);
  const Real b0 = p.real("problem", "b_par", 1.0);
  const Real amp = p.real("problem", "amp", 0.1);
  const Real lambda = p.real("problem", "lambda", 1.0);
  const Real th = p.real("problem", "angle", 0.0) * kPi / 180.0;
  const Real ph = p.real("problem", "angle2", 0.0) * kPi / 180.0;
  const Real vflow = p.real("problem", "vflow", 0.0);
  const Real dir = p.real("problem", "dir", 1.0);
  const Real kw = 2.0 * kPi / lambda;
  const Real cth = std::cos(th), sth = std::sin(th);
  const Real fac = -dir * amp / std::sqrt(dens);
  if (ph == 0.0) {
    from_potential(m, h, [=](int c, Real x, Real y, Real) -> Real {
      const Real xp = x * cth + y * sth, yp = -x * sth + y * cth;
      const Real ayp = (amp / kw) * std::sin(kw * xp);
      const Real az = b0 * yp + (amp / kw) * std::cos(kw * xp);
      if (c == 0) return -ayp * sth;
      if (c == 1) return ayp * cth;
      return az;
    });
    if (!m.two_d)
      for (int k = 0; k < m.nk + 1; ++k)
        for (int j = 0; j < m.nj + 1; ++j)
          for (int i = 0; i < m.ni + 1; ++i) h.b1(k, j, i) += b0 * cth;
    for (int k = 0; k < m.nk; ++k)
      for (int j = 0; j < m.nj; ++j)
        for (int i = 0; i < m.ni; ++i) {
          const Real xp = m.x1v(i) * cth + m.x2v(j) * sth;
          const Real vyp = fac * std::sin(kw * xp), vz = fac * std::cos(kw * xp);
          h.w(IDN, k, j, i) = dens;
          h.w(IVX, k, j, i) = vflow * cth - vyp * sth;
          h.w(IVY, k, j, i) = vflow * sth + vyp * cth;
          h.w(IVZ, k, j, i) = vz;
          h.w(IPR, k, j, i) = pres;
        }
    return;
  }
  if (!m.three_d) throw std::runtime_error("cpaw: problem.angle2 != 0 needs a 3D mesh");
  const Real cph = std::cos(ph), sph = std::sin(ph);
  // wave-frame basis: e1 along k, e2 in the x-y plane, e3 = e1 x e2
  const Real e1[3] = {cph * cth, cph * sth, sph};
  const Real e2[3] = {-sth, cth, 0.0};
  const Real e3[3] = {-sph * cth, -sph * sth, cph};
  from_potential(m, h, [=](int c, Real x, Real y, Real z) -> Real {
    const Real xp = (x * e1[0] + y * e1[1]) + z * e1[2];
    const Real yp = x * e2[0] + y * e2[1];
    const Real a2 = (amp / kw) * std::sin(kw * xp);
    const Real a3 = b0 * yp + (amp / kw) * std::cos(kw * xp);
    return a2 * e2[c] + a3 * e3[c];
  });
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        const Real xp = (m.x1v(i) * e1[0] + m.x2v(j) * e1[1]) + m.x3v(k) * e1[2];
        const Real v2 = fac * std::sin(kw * xp), v3 = fac * std::cos(kw * xp);
        h.w(IDN, k, j, i) = dens;
        h.w(IVX, k, j, i) = (vflow * e1[0] + v2 * e2[0]) + v3 * e3[0];
        h.w(IVY, k, j, i) = (vflow * e1[1] + v2 * e2[1]) + v3 * e3[1];
        h.w(IVZ, k, j, i) = (vflow * e1[2] + v2 * e2[2]) + v3 * e3[2];
        h.w(IPR, k, j, i) = pres;
      }
}

// Orszag-Tang, pgen.hpp:94-111. Extension: problem.pres and problem.b0 (defaults are the
// reference's constants, so an absent key reproduces it bitwise).
inline void pg_orszag_tang(Deck &p, const Grid &m, Init &h) {
  const Real b0 = p.real("problem", "b0", 1.0 / std::sqrt(4.0 * kPi));
  const Real pres = p.real("problem", "pres", 5.0 / (12.0 * kPi));
  from_potential(m, h, [=](int c, Real x, Real y, Real) -> Real {
    if (c != 2) return 0.0;
    return b0 / (2.0 * kPi) * std::cos(2.0 * kPi * y) + b0 / (4.0 * kPi) * std::cos(4.0 * kPi * x);
  });
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        h.w(IDN, k, j, i) = 25.0 / (36.0 * kPi);
        h.w(IVX, k, j, i) = -std::sin(2.0 * kPi * m.x2v(j));
        h.w(IVY, k, j, i) = std::sin(2.0 * kPi * m.x1v(i));
        h.w(IVZ, k, j, i) = 0.0;
        h.w(IPR, k, j, i) = pres;
      }
}

// counter-based uniform deviate in [-1, 1): splitmix64 of (seed, cell index)
inline double noise_pm1(uint64_t seed, uint64_t idx) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + (idx + 1) * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return double(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// MHD blast, pgen.hpp:114-135. Extension: problem.perturb (multiplicative density noise
// amplitude, default 0) and problem.seed (default 42); cell index = all-cell linear index.
inline void pg_blast(Deck &p, const Grid &m, Init &h) {
  const Real pout = p.real("problem", "pres_out", 0.1);
  const Real pin = p.real("problem", "pres_in", 10.0);
  const Real rad = p.real("problem", "radius", 0.1);
  const Real b0 = p.real("problem", "b0", 1.0);
  const Real ang = p.real("problem", "angle", 45.0) * kPi / 180.0;
  const Real pert = p.real("problem", "perturb", 0.0);
  const uint64_t seed = uint64_t(p.integer("problem", "seed", 42));
  Real bx = b0 * std::cos(ang), by = b0 * std::sin(ang);
  if (std::fabs(bx) < 1e-14 * b0) bx = 0.0;
  if (std::fabs(by) < 1e-14 * b0) by = 0.0;
  for (int k = 0; k < m.nk + 1; ++k)
    for (int j = 0; j < m.nj + 1; ++j)
      for (int i = 0; i < m.ni + 1; ++i) { h.b1(k, j, i) = bx; h.b2(k, j, i) = by; h.b3(k, j, i) = 0.0; }
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        const Real x = m.x1v(i), y = m.two_d ? m.x2v(j) : 0.0, z = m.three_d ? m.x3v(k) : 0.0;
        const Real r = std::sqrt(x * x + y * y + z * z);
        Real d = 1.0;
        if (pert != 0.0) d = 1.0 * (1.0 + pert * noise_pm1(seed, (uint64_t(k) * m.nj + j) * m.ni + i));
        h.w(IDN, k, j, i) = d;
        h.w(IVX, k, j, i) = 0.0; h.w(IVY, k, j, i) = 0.0; h.w(IVZ, k, j, i) = 0.0;
        h.w(IPR, k, j, i) = (r < rad) ? pin : pout;
      }
}

// field loop advection, pgen.hpp:138-155: A_z = max(amp (radius - r), 0), uniform flow
inline void pg_field_loop(Deck &p, const Grid &m, Init &h) {
  const Real a0 = p.real("problem", "amp", 1.0e-3);
  const Real rad = p.real("problem", "radius", 0.3);
  const Real vx = p.real("problem", "vx", 1.0), vy = p.real("problem", "vy", 0.5), vz = p.real("problem", "vz", 0.0);
  from_potential(m, h, [=](int c, Real x, Real y, Real) -> Real {
    if (c != 2) return 0.0;
    const Real r = std::sqrt(x * x + y * y);
    return std::fmax(a0 * (rad - r), 0.0);
  });
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        h.w(IDN, k, j, i) = 1.0; h.w(IVX, k, j, i) = vx; h.w(IVY, k, j, i) = vy; h.w(IVZ, k, j, i) = vz;
        h.w(IPR, k, j, i) = 1.0;
      }
}

// Brio-Wu shock tube, pgen.hpp:158-184. problem.perturb: density times (1 + perturb U) with U
// drawn from std::mt19937(42) through std::uniform_real_distribution<double>(-1, 1), one draw
// per cell in (k, j, i) order over all cells, ghosts included (the draw is made for every
// cell, also when perturb is 0).
inline void pg_shock_tube(Deck &p, const Grid &m, Init &h) {
  const Real dl = p.real("problem", "dl", 1.0), pl = p.real("problem", "pl", 1.0);
  const Real dr = p.real("problem", "dr", 0.125), pr = p.real("problem", "pr", 0.1);
  const Real byl = p.real("problem", "byl", 1.0), byr = p.real("problem", "byr", -1.0);
  const Real bx = p.real("problem", "bx", 0.75);
  const Real x0 = p.real("problem", "x0", 0.0);
  const Real ul = p.real("problem", "ul", 0.0), ur = p.real("problem", "ur", 0.0);
  const Real perturb = p.real("problem", "perturb", 0.0);
  std::mt19937 rng(42);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  for (int k = 0; k < m.nk + 1; ++k)
    for (int j = 0; j < m.nj + 1; ++j)
      for (int i = 0; i < m.ni + 1; ++i) {
        h.b1(k, j, i) = bx;
        h.b2(k, j, i) = (m.x1v(i) < x0) ? byl : byr;
        h.b3(k, j, i) = 0.0;
      }
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        const bool left = m.x1v(i) < x0;
        h.w(IDN, k, j, i) = (left ? dl : dr) * (1.0 + perturb * U(rng));
        h.w(IVX, k, j, i) = left ? ul : ur;
        h.w(IVY, k, j, i) = 0.0; h.w(IVZ, k, j, i) = 0.0;
        h.w(IPR, k, j, i) = left ? pl : pr;
      }
}

// uniform state (gravity and cooling tests), pgen.hpp:216-229
inline void pg_uniform(Deck &p, const Grid &m, Init &h) {
  const Real d = p.real("problem", "dens", 1.0), pr = p.real("problem", "pres", 1.0);
  const Real vx = p.real("problem", "vx", 0.0), vy = p.real("problem", "vy", 0.0), vz = p.real("problem", "vz", 0.0);
  const Real bx = p.real("problem", "bx", 0.0), by = p.real("problem", "by", 0.0), bz = p.real("problem", "bz", 0.0);
  for (int k = 0; k < m.nk + 1; ++k)
    for (int j = 0; j < m.nj + 1; ++j)
      for (int i = 0; i < m.ni + 1; ++i) { h.b1(k, j, i) = bx; h.b2(k, j, i) = by; h.b3(k, j, i) = bz; }
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        h.w(IDN, k, j, i) = d; h.w(IVX, k, j, i) = vx; h.w(IVY, k, j, i) = vy; h.w(IVZ, k, j, i) = vz;
        h.w(IPR, k, j, i) = pr;
      }
}

// Gresho vortex at a prescribed Mach number, pgen.hpp:188-214
inline void pg_gresho(Deck &p, const Grid &m, Init &h) {
  const Real mach = p.real("problem", "mach", 0.1);
  const Real gamma = p.real("mhd", "gamma", 5.0 / 3.0);
  const Real p0 = 1.0 / (gamma * mach * mach);
  for (int k = 0; k < m.nk + 1; ++k)
    for (int j = 0; j < m.nj + 1; ++j)
      for (int i = 0; i < m.ni + 1; ++i) { h.b1(k, j, i) = 0.0; h.b2(k, j, i) = 0.0; h.b3(k, j, i) = 0.0; }
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        const Real x = m.x1v(i), y = m.two_d ? m.x2v(j) : 0.0;
        const Real r = std::sqrt(x * x + y * y);
        Real vphi, pr;
        if (r < 0.2) { vphi = 5.0 * r; pr = p0 + 12.5 * r * r; }
        else if (r < 0.4) { vphi = 2.0 - 5.0 * r; pr = p0 + 12.5 * r * r - 20.0 * r + 4.0 * std::log(r / 0.2) + 4.0; }
        else { vphi = 0.0; pr = p0 - 2.0 + 4.0 * std::log(2.0); }
        const Real ir = r > 1e-14 ? 1.0 / r : 0.0;
        h.w(IDN, k, j, i) = 1.0;
        h.w(IVX, k, j, i) = -vphi * y * ir;
        h.w(IVY, k, j, i) = vphi * x * ir;
        h.w(IVZ, k, j, i) = 0.0;
        h.w(IPR, k, j, i) = pr;
      }
}

// sinusoidal decay test of one diffusive term along one axis, pgen.hpp:235-266: var = b
// (B_{axis+1} on faces, uniform total pressure), v (v_{axis+1}) or temp (T at uniform p)
inline void pg_diffusion_test(Deck &p, const Grid &m, Init &h) {
  const std::string var = p.str("problem", "var", "b");
  const int axis = p.integer("problem", "axis", 1);
  const Real amp = p.real("problem", "amp", 1.0e-3);
  const Real d0 = p.real("problem", "dens", 1.0), p0 = p.real("problem", "pres", 1.0);
  const Real L = axis == 1 ? (m.x1max - m.x1min) : axis == 2 ? (m.x2max - m.x2min) : (m.x3max - m.x3min);
  const Real kw = 2.0 * kPi / L;
  const int comp = axis % 3;
  auto coord = [&](int kk, int j, int i) { return axis == 1 ? m.x1v(i) : axis == 2 ? m.x2v(j) : m.x3v(kk); };
  for (int kk = 0; kk < m.nk + 1; ++kk)
    for (int j = 0; j < m.nj + 1; ++j)
      for (int i = 0; i < m.ni + 1; ++i) {
        h.b1(kk, j, i) = 0.0; h.b2(kk, j, i) = 0.0; h.b3(kk, j, i) = 0.0;
        if (var == "b") {
          const Real val = amp * std::sin(kw * coord(kk, j, i));
          if (comp == 0) h.b1(kk, j, i) = val; else if (comp == 1) h.b2(kk, j, i) = val; else h.b3(kk, j, i) = val;
        }
      }
  for (int kk = 0; kk < m.nk; ++kk)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        const Real x = coord(kk, j, i);
        Real d = d0, pr = p0, v[3] = {0.0, 0.0, 0.0};
        if (var == "v") v[comp] = amp * std::sin(kw * x);
        if (var == "temp") { const Real T = p0 / d0 + amp * std::sin(kw * x); d = p0 / T; }
        if (var == "b") pr = p0 - 0.5 * sqr(amp * std::sin(kw * x));
        h.w(IDN, kk, j, i) = d; h.w(IVX, kk, j, i) = v[0]; h.w(IVY, kk, j, i) = v[1]; h.w(IVZ, kk, j, i) = v[2];
        h.w(IPR, kk, j, i) = pr;
      }
}

// fill primitives and faces, convert to conserved (pgen.hpp:269-307)
inline void initialize(const std::string &name, Deck &p, Grid &m, const Cfg &cfg) {
  Init h;
  h.w.alloc(NVAR, m.nk, m.nj, m.ni);
  h.b1.alloc(m.nk + 1, m.nj + 1, m.ni + 1); h.b2.alloc(m.nk + 1, m.nj + 1, m.ni + 1); h.b3.alloc(m.nk + 1, m.nj + 1, m.ni + 1);
  if (name == "cpaw") pg_cpaw(p, m, h);
  else if (name == "orszag_tang") pg_orszag_tang(p, m, h);
  else if (name == "blast") pg_blast(p, m, h);
  else if (name == "field_loop") pg_field_loop(p, m, h);
  else if (name == "shock_tube") pg_shock_tube(p, m, h);
  else if (name == "diffusion_test") pg_diffusion_test(p, m, h);
  else if (name == "uniform") pg_uniform(p, m, h);
  else if (name == "gresho") pg_gresho(p, m, h);
  else throw std::runtime_error("unknown problem: " + name);
  for (int k = 0; k < m.nk; ++k)
    for (int j = 0; j < m.nj; ++j)
      for (int i = 0; i < m.ni; ++i) {
        const Real bx = 0.5 * (h.b1(k, j, i) + h.b1(k, j, i + 1));
        const Real by = m.two_d ? 0.5 * (h.b2(k, j, i) + h.b2(k, j + 1, i)) : h.b2(k, j, i);
        const Real bz = m.three_d ? 0.5 * (h.b3(k, j, i) + h.b3(k + 1, j, i)) : h.b3(k, j, i);
        const Real d = h.w(IDN, k, j, i), vx = h.w(IVX, k, j, i), vy = h.w(IVY, k, j, i), vz = h.w(IVZ, k, j, i);
        m.u0(IDN, k, j, i) = d;
        m.u0(IM1, k, j, i) = d * vx; m.u0(IM2, k, j, i) = d * vy; m.u0(IM3, k, j, i) = d * vz;
        m.u0(IEN, k, j, i) = h.w(IPR, k, j, i) / cfg.eos.gm1 + 0.5 * d * (vx * vx + vy * vy + vz * vz)
                             + 0.5 * (bx * bx + by * by + bz * bz);
        h.w(IENT, k, j, i) = h.w(IPR, k, j, i) / std::pow(d, cfg.eos.gamma);
        m.u0(IENT, k, j, i) = d * h.w(IENT, k, j, i);
      }
  m.b1f0.v = h.b1.v; m.b2f0.v = h.b2.v; m.b3f0.v = h.b3.v;
  m.w.v = h.w.v;
}

// ---------------------------------------------------------------- the simulation object
inline Bc parse_bc(const std::string &s) {
  if (s == "periodic") return Bc::periodic;
  if (s == "outflow") return Bc::outflow;
  if (s == "reflect") return Bc::reflect;
  throw std::runtime_error("unknown boundary condition: " + s);
}

struct Sim {
  Deck deck;
  Cfg cfg;
  Grid m;
  Real t = 0.0, tlim = 1.0, min_dt = 0.0, last_dt = 0.0;
  int cycle = 0, cycle0 = 0, nlim = -1;   // nlim counts from cycle0 (main.cpp:241-243)
  long fofc_count = 0, sts_count = 0, limiter_count = 0;
  StsWork sts_work;
  std::string problem;
  A4 u_init, bcc_init;

  // configuration + mesh + initial state (main.cpp:112-208); pgen_only stops after the
  // problem generator (no boundaries, no con2prim) so its output can be compared as is
  void setup(bool pgen_only = false) {
    Deck &p = deck;
    const std::string r = p.str("mhd", "reconstruct", "ppm");
    cfg.recon = r == "dc" ? Recon::dc : r == "plm" ? Recon::plm : r == "ppm" ? Recon::ppm : r == "ppmx" ? Recon::ppmx
              : r == "wenoz" ? Recon::wenoz : throw std::runtime_error("unknown reconstruct: " + r);
    const std::string s = p.str("mhd", "rsolver", "hlld");
    cfg.rsolver = s == "hlld" ? RSolver::hlld : s == "lhlld" ? RSolver::lhlld : s == "hlle" ? RSolver::hlle
                : throw std::runtime_error("unknown rsolver: " + s);
    cfg.characteristic = p.boolean("mhd", "characteristic", false);
    cfg.lhlld_phi = p.real("mhd", "lhlld_phi", -1.0);
    cfg.eos.dual_energy = p.boolean("mhd", "dual_energy", false);
    cfg.eos.de_eta = p.real("mhd", "de_eta", 1.0e-3);
    const std::string ti = p.str("time", "integrator", "rk3");
    cfg.nstage = ti == "rk1" ? 1 : ti == "rk2" ? 2 : ti == "rk3" ? 3 : (ti == "rk4" || ti == "ssprk43") ? 4
               : throw std::runtime_error("unknown integrator: " + ti);
    cfg.low_mach = p.boolean("mhd", "low_mach", false);
    cfg.flatten = p.boolean("mhd", "flatten", true);
    cfg.cfl = p.real("time", "cfl", 0.3);
    cfg.eos.gamma = p.real("mhd", "gamma", 5.0 / 3.0);
    cfg.eos.gm1 = cfg.eos.gamma - 1.0;
    cfg.eos.dfloor = p.real("mhd", "dfloor", 1.0e-10);
    cfg.eos.pfloor = p.real("mhd", "pfloor", 1.0e-12);
    cfg.bc[0] = parse_bc(p.str("mesh", "bc_x1", "periodic"));
    cfg.bc[1] = parse_bc(p.str("mesh", "bc_x2", "periodic"));
    cfg.bc[2] = parse_bc(p.str("mesh", "bc_x3", "periodic"));
    cfg.fofc = p.boolean("mhd", "fofc", true);
    cfg.flux_limiter = p.boolean("mhd", "flux_limiter", false);
    cfg.sts = p.boolean("diffusion", "sts", false);                  // main.cpp:142-146
    cfg.nu = p.real("diffusion", "nu", 0.0);
    cfg.kappa = p.real("diffusion", "kappa", 0.0);
    cfg.eta = p.real("diffusion", "eta", 0.0);
    cfg.grav_uniform = p.boolean("gravity", "uniform", false);          // main.cpp:148-162
    cfg.g1 = p.real("gravity", "g1", 0.0);
    cfg.g2 = p.real("gravity", "g2", 0.0);
    cfg.g3 = p.real("gravity", "g3", 0.0);
    cfg.grav_point = p.boolean("gravity", "point_mass", false);
    cfg.gm = p.real("gravity", "gm", 0.0);
    cfg.gx0 = p.real("gravity", "x0", 0.0);
    cfg.gy0 = p.real("gravity", "y0", 0.0);
    cfg.gz0 = p.real("gravity", "z0", 0.0);
    cfg.rsoft = p.real("gravity", "rsoft", 0.0);
    cfg.cooling = p.boolean("cooling", "enabled", false);
    cfg.lambda0 = p.real("cooling", "lambda0", 0.0);
    cfg.cool_alpha = p.real("cooling", "alpha", 1.0);
    cfg.tfloor = p.real("cooling", "tfloor", 0.0);
    m.nx1 = p.integer("mesh", "nx1");
    m.nx2 = p.integer("mesh", "nx2", 1);
    m.nx3 = p.integer("mesh", "nx3", 1);
    const int ngmin = (cfg.recon == Recon::ppm || cfg.recon == Recon::ppmx || cfg.recon == Recon::wenoz) ? 3 : 2;
    m.ng = p.integer("mesh", "nghost", ngmin);
    if (m.ng < ngmin) throw std::runtime_error("nghost too small for the reconstruction method");
    m.x1min = p.real("mesh", "x1min", 0.0); m.x1max = p.real("mesh", "x1max", 1.0);
    m.x2min = p.real("mesh", "x2min", 0.0); m.x2max = p.real("mesh", "x2max", 1.0);
    m.x3min = p.real("mesh", "x3min", 0.0); m.x3max = p.real("mesh", "x3max", 1.0);
    m.allocate();
    tlim = p.real("time", "tlim", 1.0);
    nlim = p.integer("time", "nlim", -1);
    problem = p.str("problem", "name");
    initialize(problem, p, m, cfg);
    if (pgen_only) return;
    boundaries(m, cfg, 0);
    min_dt = cons_to_prim(m, cfg, 0, true);
    if (!cfg.sts) min_dt = std::fmin(min_dt, diffusion_dt(m, cfg, diff_of(cfg)));   // main.cpp:202
    u_init = m.u0;
    bcc_init = m.bcc;
  }

  // one step of the time loop, main.cpp:247-292 (returns false when the loop would stop)
  bool step() {
    if (!(t < tlim && (nlim < 0 || cycle - cycle0 < nlim))) return false;
    Real dt = cfg.cfl * min_dt;
    if (t + dt > tlim) dt = tlim - t;
    for (int s = 1; s <= cfg.nstage; ++s) {
      const StageCoef c = stage_coef(cfg.nstage, s);
      const bool last = s == cfg.nstage;
      fofc_count += stage(m, cfg, c.cur, c.out, c.g0, c.g1, c.beta * dt, &limiter_count);
      const bool split = last && (cfg.sts || cfg.cooling);             // main.cpp:273-275
      const Real mdt = cons_to_prim(m, cfg, c.out, last && !split);
      if (last && !split) min_dt = std::fmin(mdt, diffusion_dt(m, cfg, diff_of(cfg)));
    }
    if (cfg.sts) {                                                     // main.cpp:278-281
      sts_count += sts_advance(m, cfg, diff_of(cfg), sts_work, dt, diffusion_dt(m, cfg, diff_of(cfg)));
      boundaries(m, cfg, 0);
    }
    if (cfg.cooling) {                                                 // main.cpp:282-286
      cons_to_prim(m, cfg, 0, false);
      apply_cooling(m, cfg, m.u0, dt);
      boundaries(m, cfg, 0);
    }
    if (cfg.sts || cfg.cooling) {                                      // main.cpp:287-290
      min_dt = cons_to_prim(m, cfg, 0, true);
      if (!cfg.sts) min_dt = std::fmin(min_dt, diffusion_dt(m, cfg, diff_of(cfg)));
    }
    t += dt;
    ++cycle;
    last_dt = dt;
    return true;
  }

  // binary restart dump (restart.hpp:11-34): header, then u0 and the three face fields of
  // set 0 in the reference layout
  struct RstHeader {
    int magic = 0x43485351;   // "CHSQ"
    int nx1 = 0, nx2 = 0, nx3 = 0, ng = 0;
    int cycle = 0;
    double time = 0.0;
    int real_size = 0;
  };
  void restart_write(const std::string &path) const {
    RstHeader h;
    h.nx1 = m.nx1; h.nx2 = m.nx2; h.nx3 = m.nx3; h.ng = m.ng; h.cycle = cycle; h.time = t;
    h.real_size = int(sizeof(Real));
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write(reinterpret_cast<const char *>(&h), sizeof(h));
    for (const auto *v : {&m.u0.v, &m.b1f0.v, &m.b2f0.v, &m.b3f0.v})
      f.write(reinterpret_cast<const char *>(v->data()), std::streamsize(v->size() * sizeof(Real)));
  }
  // restart.hpp:37-58, then the driver's boundaries + con2prim (main.cpp:196-201)
  void restart_read(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open restart file " + path);
    RstHeader h;
    f.read(reinterpret_cast<char *>(&h), sizeof(h));
    if (h.magic != 0x43485351) throw std::runtime_error("not a Chasqui restart file: " + path);
    if (h.nx1 != m.nx1 || h.nx2 != m.nx2 || h.nx3 != m.nx3 || h.ng != m.ng || h.real_size != int(sizeof(Real)))
      throw std::runtime_error("restart file does not match the mesh");
    for (auto *v : {&m.u0.v, &m.b1f0.v, &m.b2f0.v, &m.b3f0.v})
      f.read(reinterpret_cast<char *>(v->data()), std::streamsize(v->size() * sizeof(Real)));
    if (!f) throw std::runtime_error("restart file is truncated: " + path);
    t = h.time;
    cycle = h.cycle;
    cycle0 = h.cycle;
    boundaries(m, cfg, 0);
    min_dt = cons_to_prim(m, cfg, 0, true);
  }

  // L1 error vs the initial state over the 9 fields (main.cpp:310-327)
  void l1_error(double err[10]) const {
    for (int n = 0; n < 10; ++n) err[n] = 0.0;
    for (int k = m.ks; k <= m.ke; ++k)
      for (int j = m.js; j <= m.je; ++j)
        for (int i = m.is; i <= m.ie; ++i) {
          for (int n = 0; n < NVAR; ++n) err[n] += std::fabs(m.u0(n, k, j, i) - u_init(n, k, j, i));
          for (int n = 0; n < 3; ++n) err[NVAR + n] += std::fabs(m.bcc(n, k, j, i) - bcc_init(n, k, j, i));
        }
    const double zones = double(m.nx1) * m.nx2 * m.nx3;
    double tot = 0.0;
    for (int n = 0; n < 9; ++n) { err[n] /= zones; tot += err[n]; }
    err[9] = tot;
  }
};

}  // namespace orc
