// Run configuration and problem generators (host side), and the chq_deck_* C ABI.
//
// Problems: cpaw, orszag_tang, blast, shock_tube (pgen.hpp:54-184), with the keys the
// benchmark decks need and the reference lacks: cpaw problem.angle2 (propagation direction
// out of the x-y plane), orszag_tang problem.pres / problem.b0, blast problem.perturb /
// problem.seed (counter-based density noise). Absent keys reproduce the reference bitwise.
#include <cmath>
#include <random>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "deck.hpp"

namespace chq {

constexpr double kPi = 3.14159265358979323846;

void HostGrid::build() {
  ni = nx1 + 2 * ng;
  nj = (nx2 > 1) ? nx2 + 2 * ng : 1;
  nk = (nx3 > 1) ? nx3 + 2 * ng : 1;
  is = ng;
  ie = ng + nx1 - 1;
  js = (nx2 > 1) ? ng : 0;
  je = (nx2 > 1) ? ng + nx2 - 1 : 0;
  ks = (nx3 > 1) ? ng : 0;
  ke = (nx3 > 1) ? ng + nx3 - 1 : 0;
  two_d = nx2 > 1;
  three_d = nx3 > 1;
  dx1 = (x1max - x1min) / nx1;
  dx2 = (x2max - x2min) / nx2;
  dx3 = (x3max - x3min) / nx3;
}

namespace {

// fn(k, j) over the nk x nj rows, split across host threads for large meshes (every row
// writes only its own elements, so the result does not depend on the split)
void for_rows(int nk, int nj, int ni, const std::function<void(int, int)> &fn) {
  const long long rows = (long long)nk * nj;
  unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (rows * ni < (1 << 17) || nt == 1) {
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nj; ++j) fn(k, j);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {   // a contiguous block of rows per thread
      for (long long r = rows * t / nt; r < rows * (t + 1) / nt; ++r) fn(int(r / nj), int(r % nj));
    });
  for (auto &x : th) x.join();
}

int parse_bc(const std::string &s) {
  if (s == "periodic") return 0;
  if (s == "outflow") return 1;
  if (s == "reflect") return 2;
  throw std::runtime_error("unknown boundary condition: " + s);
}

struct Fields {
  const HostGrid &g;
  HostArray &w, &b1, &b2, &b3;
  double &W(int n, int k, int j, int i) { return w[size_t(n) * g.ncell() + g.c(k, j, i)]; }
};

// B = curl A from an edge-sampled potential (pgen.hpp:30-50)
void curl_potential(Fields &h, const std::function<double(int, double, double, double)> &A) {
  const HostGrid &m = h.g;
  const double idx = 1.0 / m.dx1, idy = 1.0 / m.dx2, idz = 1.0 / m.dx3;
  for_rows(m.nk + 1, m.nj + 1, m.ni + 1, [&](int k, int j) {
      for (int i = 0; i < m.ni + 1; ++i) {
        const double xf = m.x1f(i), yf = m.x2f(j), zf = m.x3f(k);
        const double xv = m.x1v(i), yv = m.x2v(j), zv = m.x3v(k);
        double bx = 0.0;
        if (m.two_d) bx += (A(2, xf, m.x2f(j + 1), zv) - A(2, xf, yf, zv)) * idy;
        if (m.three_d) bx -= (A(1, xf, yv, m.x3f(k + 1)) - A(1, xf, yv, zf)) * idz;
        double by = -(A(2, m.x1f(i + 1), yf, zv) - A(2, xf, yf, zv)) * idx;
        if (m.three_d) by += (A(0, xv, yf, m.x3f(k + 1)) - A(0, xv, yf, zf)) * idz;
        double bz = (A(1, m.x1f(i + 1), yv, zf) - A(1, xf, yv, zf)) * idx;
        if (m.two_d) bz -= (A(0, xv, m.x2f(j + 1), zf) - A(0, xv, yf, zf)) * idy;
        const size_t f = m.f(k, j, i);
        h.b1[f] = bx;
        h.b2[f] = by;
        h.b3[f] = bz;
      }
  });
}

// circularly polarized Alfven wave, pgen.hpp:54-91 (+ angle2, 3D propagation direction)
void cpaw(Params &p, Fields &h) {
  const HostGrid &m = h.g;
  const double dens = p.get_real("problem", "dens", 1.0);
  const double pres = p.get_real("problem", "pres", 0.1);
  const double b0 = p.get_real("problem", "b_par", 1.0);
  const double amp = p.get_real("problem", "amp", 0.1);
  const double lambda = p.get_real("problem", "lambda", 1.0);
  const double th = p.get_real("problem", "angle", 0.0) * kPi / 180.0;
  const double ph = p.get_real("problem", "angle2", 0.0) * kPi / 180.0;
  const double vflow = p.get_real("problem", "vflow", 0.0);
  const double dir = p.get_real("problem", "dir", 1.0);
  const double kw = 2.0 * kPi / lambda;
  const double cth = std::cos(th), sth = std::sin(th);
  const double fac = -dir * amp / std::sqrt(dens);
  if (ph == 0.0) {
    curl_potential(h, [=](int c, double x, double y, double) {
      const double xp = x * cth + y * sth, yp = -x * sth + y * cth;
      const double a_y = (amp / kw) * std::sin(kw * xp);
      const double a_z = b0 * yp + (amp / kw) * std::cos(kw * xp);
      return c == 0 ? -a_y * sth : c == 1 ? a_y * cth : a_z;
    });
    if (!m.two_d)   // the uniform B1 of a 1D run cannot come from the potential
      for (size_t f = 0; f < m.nface(); ++f) h.b1[f] += b0 * cth;
    for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
        for (int i = 0; i < m.ni; ++i) {
          const double xp = m.x1v(i) * cth + m.x2v(j) * sth;
          const double vyp = fac * std::sin(kw * xp), vz = fac * std::cos(kw * xp);
          h.W(0, k, j, i) = dens;
          h.W(1, k, j, i) = vflow * cth - vyp * sth;
          h.W(2, k, j, i) = vflow * sth + vyp * cth;
          h.W(3, k, j, i) = vz;
          h.W(4, k, j, i) = pres;
        }
    });
    return;
  }
  if (!m.three_d) throw std::runtime_error("cpaw: problem.angle2 != 0 needs a 3D mesh");
  const double cph = std::cos(ph), sph = std::sin(ph);
  const double e1[3] = {cph * cth, cph * sth, sph};      // along the wave vector
  const double e2[3] = {-sth, cth, 0.0};                 // in the x-y plane
  const double e3[3] = {-sph * cth, -sph * sth, cph};    // e1 x e2
  curl_potential(h, [=](int c, double x, double y, double z) {
    const double xp = (x * e1[0] + y * e1[1]) + z * e1[2];
    const double yp = x * e2[0] + y * e2[1];
    const double a2 = (amp / kw) * std::sin(kw * xp);
    const double a3 = b0 * yp + (amp / kw) * std::cos(kw * xp);
    return a2 * e2[c] + a3 * e3[c];
  });
  for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
      for (int i = 0; i < m.ni; ++i) {
        const double xp = (m.x1v(i) * e1[0] + m.x2v(j) * e1[1]) + m.x3v(k) * e1[2];
        const double v2 = fac * std::sin(kw * xp), v3 = fac * std::cos(kw * xp);
        h.W(0, k, j, i) = dens;
        for (int c = 0; c < 3; ++c) h.W(1 + c, k, j, i) = (vflow * e1[c] + v2 * e2[c]) + v3 * e3[c];
        h.W(4, k, j, i) = pres;
      }
  });
}

// Orszag-Tang vortex, pgen.hpp:94-111 (+ problem.pres, problem.b0)
void orszag_tang(Params &p, Fields &h) {
  const HostGrid &m = h.g;
  const double b0 = p.get_real("problem", "b0", 1.0 / std::sqrt(4.0 * kPi));
  const double pres = p.get_real("problem", "pres", 5.0 / (12.0 * kPi));
  curl_potential(h, [=](int c, double x, double y, double) {
    if (c != 2) return 0.0;
    return b0 / (2.0 * kPi) * std::cos(2.0 * kPi * y) + b0 / (4.0 * kPi) * std::cos(4.0 * kPi * x);
  });
  for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
      for (int i = 0; i < m.ni; ++i) {
        h.W(0, k, j, i) = 25.0 / (36.0 * kPi);
        h.W(1, k, j, i) = -std::sin(2.0 * kPi * m.x2v(j));
        h.W(2, k, j, i) = std::sin(2.0 * kPi * m.x1v(i));
        h.W(3, k, j, i) = 0.0;
        h.W(4, k, j, i) = pres;
      }
  });
}

// splitmix64 of (seed, all-cell index) mapped to [-1, 1): a counter-based generator, so the
// noise does not depend on traversal order or thread count
double cell_noise(uint64_t seed, uint64_t idx) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + (idx + 1) * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return double(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// MHD blast, pgen.hpp:114-135 (+ problem.perturb, problem.seed)
void blast(Params &p, Fields &h) {
  const HostGrid &m = h.g;
  const double pout = p.get_real("problem", "pres_out", 0.1);
  const double pin = p.get_real("problem", "pres_in", 10.0);
  const double rad = p.get_real("problem", "radius", 0.1);
  const double b0 = p.get_real("problem", "b0", 1.0);
  const double ang = p.get_real("problem", "angle", 45.0) * kPi / 180.0;
  const double pert = p.get_real("problem", "perturb", 0.0);
  const uint64_t seed = uint64_t(p.get_int("problem", "seed", 42));
  double bx = b0 * std::cos(ang), by = b0 * std::sin(ang);
  if (std::fabs(bx) < 1e-14 * b0) bx = 0.0;
  if (std::fabs(by) < 1e-14 * b0) by = 0.0;
  for_rows(m.nk + 1, m.nj + 1, m.ni + 1, [&](int k, int j) {
    for (int i = 0; i < m.ni + 1; ++i) {
      const size_t f = m.f(k, j, i);
      h.b1[f] = bx;
      h.b2[f] = by;
      h.b3[f] = 0.0;
    }
  });
  for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
      for (int i = 0; i < m.ni; ++i) {
        const double x = m.x1v(i), y = m.two_d ? m.x2v(j) : 0.0, z = m.three_d ? m.x3v(k) : 0.0;
        const double r = std::sqrt(x * x + y * y + z * z);
        double d = 1.0;
        if (pert != 0.0) d = 1.0 * (1.0 + pert * cell_noise(seed, m.c(k, j, i)));
        h.W(0, k, j, i) = d;
        h.W(1, k, j, i) = 0.0;
        h.W(2, k, j, i) = 0.0;
        h.W(3, k, j, i) = 0.0;
        h.W(4, k, j, i) = (r < rad) ? pin : pout;
      }
  });
}

// field loop advection, pgen.hpp:138-155: A_z = max(amp (radius - r), 0), uniform flow
void field_loop(Params &p, Fields &h) {
  const HostGrid &m = h.g;
  const double a0 = p.get_real("problem", "amp", 1.0e-3);
  const double rad = p.get_real("problem", "radius", 0.3);
  const double vx = p.get_real("problem", "vx", 1.0), vy = p.get_real("problem", "vy", 0.5);
  const double vz = p.get_real("problem", "vz", 0.0);
  curl_potential(h, [=](int c, double x, double y, double) {
    if (c != 2) return 0.0;
    const double r = std::sqrt(x * x + y * y);
    return std::fmax(a0 * (rad - r), 0.0);
  });
  for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
      for (int i = 0; i < m.ni; ++i) {
        h.W(0, k, j, i) = 1.0;
        h.W(1, k, j, i) = vx;
        h.W(2, k, j, i) = vy;
        h.W(3, k, j, i) = vz;
        h.W(4, k, j, i) = 1.0;
      }
  });
}

// Brio-Wu shock tube, pgen.hpp:158-184. problem.perturb: density times (1 + perturb U), U from
// std::mt19937(42) through std::uniform_real_distribution<double>(-1, 1), one draw per cell in
// (k, j, i) order over all cells, ghosts included (so that pass runs on one thread).
void shock_tube(Params &p, Fields &h) {
  const HostGrid &m = h.g;
  const double dl = p.get_real("problem", "dl", 1.0), pl = p.get_real("problem", "pl", 1.0);
  const double dr = p.get_real("problem", "dr", 0.125), pr = p.get_real("problem", "pr", 0.1);
  const double byl = p.get_real("problem", "byl", 1.0), byr = p.get_real("problem", "byr", -1.0);
  const double bx = p.get_real("problem", "bx", 0.75);
  const double x0 = p.get_real("problem", "x0", 0.0);
  const double ul = p.get_real("problem", "ul", 0.0), ur = p.get_real("problem", "ur", 0.0);
  const double perturb = p.get_real("problem", "perturb", 0.0);
  for_rows(m.nk + 1, m.nj + 1, m.ni + 1, [&](int k, int j) {
      for (int i = 0; i < m.ni + 1; ++i) {
        const size_t f = m.f(k, j, i);
        h.b1[f] = bx;
        h.b2[f] = (m.x1v(i) < x0) ? byl : byr;
        h.b3[f] = 0.0;
      }
  });
  for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
      for (int i = 0; i < m.ni; ++i) {
        const bool left = m.x1v(i) < x0;
        h.W(0, k, j, i) = (left ? dl : dr) * (1.0 + 0.0);
        h.W(1, k, j, i) = left ? ul : ur;
        h.W(2, k, j, i) = 0.0;
        h.W(3, k, j, i) = 0.0;
        h.W(4, k, j, i) = left ? pl : pr;
      }
  });
  if (perturb != 0.0) {   // (with perturb = 0 every factor is 1 + 0 U = 1, as above)
    std::mt19937 rng(42);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int k = 0; k < m.nk; ++k)
      for (int j = 0; j < m.nj; ++j)
        for (int i = 0; i < m.ni; ++i) {
          const bool left = m.x1v(i) < x0;
          h.W(0, k, j, i) = (left ? dl : dr) * (1.0 + perturb * U(rng));
        }
  }
}

// uniform state (gravity and cooling tests), pgen.hpp:216-229
void uniform(Params &p, Fields &h) {
  const HostGrid &m = h.g;
  const double d = p.get_real("problem", "dens", 1.0), pr = p.get_real("problem", "pres", 1.0);
  const double vx = p.get_real("problem", "vx", 0.0), vy = p.get_real("problem", "vy", 0.0);
  const double vz = p.get_real("problem", "vz", 0.0);
  const double bx = p.get_real("problem", "bx", 0.0), by = p.get_real("problem", "by", 0.0);
  const double bz = p.get_real("problem", "bz", 0.0);
  for_rows(m.nk + 1, m.nj + 1, m.ni + 1, [&](int k, int j) {
      for (int i = 0; i < m.ni + 1; ++i) {
        const size_t f = m.f(k, j, i);
        h.b1[f] = bx;
        h.b2[f] = by;
        h.b3[f] = bz;
      }
  });
  for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
      for (int i = 0; i < m.ni; ++i) {
        h.W(0, k, j, i) = d;
        h.W(1, k, j, i) = vx;
        h.W(2, k, j, i) = vy;
        h.W(3, k, j, i) = vz;
        h.W(4, k, j, i) = pr;
      }
  });
}

// Gresho vortex at Mach `mach` (pressure equilibrium of a rotating ring), pgen.hpp:188-214
void gresho(Params &p, Fields &h) {
  const HostGrid &m = h.g;
  const double mach = p.get_real("problem", "mach", 0.1);
  const double gamma = p.get_real("mhd", "gamma", 5.0 / 3.0);
  const double p0 = 1.0 / (gamma * mach * mach);
  for_rows(m.nk + 1, m.nj + 1, m.ni + 1, [&](int k, int j) {
      for (int i = 0; i < m.ni + 1; ++i) {
        const size_t f = m.f(k, j, i);
        h.b1[f] = 0.0;
        h.b2[f] = 0.0;
        h.b3[f] = 0.0;
      }
  });
  for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
      for (int i = 0; i < m.ni; ++i) {
        const double x = m.x1v(i), y = m.two_d ? m.x2v(j) : 0.0;
        const double r = std::sqrt(x * x + y * y);
        double vphi, pr;
        if (r < 0.2) {
          vphi = 5.0 * r;
          pr = p0 + 12.5 * r * r;
        } else if (r < 0.4) {
          vphi = 2.0 - 5.0 * r;
          pr = p0 + 12.5 * r * r - 20.0 * r + 4.0 * std::log(r / 0.2) + 4.0;
        } else {
          vphi = 0.0;
          pr = p0 - 2.0 + 4.0 * std::log(2.0);
        }
        const double ir = r > 1e-14 ? 1.0 / r : 0.0;
        h.W(0, k, j, i) = 1.0;
        h.W(1, k, j, i) = -vphi * y * ir;
        h.W(2, k, j, i) = vphi * x * ir;
        h.W(3, k, j, i) = 0.0;
        h.W(4, k, j, i) = pr;
      }
  });
}

// decay of a sinusoid in B, v or T along one axis under one diffusive term, pgen.hpp:235-266
void diffusion_test(Params &p, Fields &h) {
  const HostGrid &m = h.g;
  const std::string var = p.get_string("problem", "var", "b");
  const int axis = p.get_int("problem", "axis", 1);
  const double amp = p.get_real("problem", "amp", 1.0e-3);
  const double d0 = p.get_real("problem", "dens", 1.0), p0 = p.get_real("problem", "pres", 1.0);
  const double L = axis == 1 ? (m.x1max - m.x1min) : axis == 2 ? (m.x2max - m.x2min) : (m.x3max - m.x3min);
  const double kw = 2.0 * kPi / L;
  const int comp = axis % 3;
  auto coord = [&](int k, int j, int i) { return axis == 1 ? m.x1v(i) : axis == 2 ? m.x2v(j) : m.x3v(k); };
  for_rows(m.nk + 1, m.nj + 1, m.ni + 1, [&](int k, int j) {
      for (int i = 0; i < m.ni + 1; ++i) {
        const size_t f = m.f(k, j, i);
        h.b1[f] = 0.0;
        h.b2[f] = 0.0;
        h.b3[f] = 0.0;
        if (var == "b") {
          const double val = amp * std::sin(kw * coord(k, j, i));
          (comp == 0 ? h.b1 : comp == 1 ? h.b2 : h.b3)[f] = val;
        }
      }
  });
  for_rows(m.nk, m.nj, m.ni, [&](int k, int j) {
      for (int i = 0; i < m.ni; ++i) {
        const double x = coord(k, j, i);
        double d = d0, pr = p0, v[3] = {0.0, 0.0, 0.0};
        if (var == "v") v[comp] = amp * std::sin(kw * x);
        if (var == "temp") {
          const double T = p0 / d0 + amp * std::sin(kw * x);
          d = p0 / T;
        }
        if (var == "b") {
          const double s = amp * std::sin(kw * x);
          pr = p0 - 0.5 * (s * s);
        }
        h.W(0, k, j, i) = d;
        h.W(1, k, j, i) = v[0];
        h.W(2, k, j, i) = v[1];
        h.W(3, k, j, i) = v[2];
        h.W(4, k, j, i) = pr;
      }
  });
}

}  // namespace

void Deck::configure() {
  const std::string r = p.get_string("mhd", "reconstruct", "ppm");
  cfg.recon = r == "dc" ? 0 : r == "plm" ? 1 : r == "ppm" ? 2 : r == "wenoz" ? 3 : r == "ppmx" ? 4 : -1;
  if (cfg.recon < 0) throw std::runtime_error("unknown reconstruct: " + r);
  const std::string s = p.get_string("mhd", "rsolver", "hlld");
  cfg.rsolver = s == "hlle" ? 0 : s == "hlld" ? 1 : s == "lhlld" ? 2 : -1;
  if (cfg.rsolver < 0) throw std::runtime_error("unknown rsolver: " + s);
  cfg.characteristic = p.get_bool("mhd", "characteristic", false) ? 1 : 0;   // main.cpp:123
  cfg.lhlld_phi = p.get_real("mhd", "lhlld_phi", -1.0);
  cfg.dual_energy = p.get_bool("mhd", "dual_energy", false);
  cfg.de_eta = p.get_real("mhd", "de_eta", 1.0e-3);
  const std::string t = p.get_string("time", "integrator", "rk3");
  cfg.nstage = t == "rk1" ? 1 : t == "rk2" ? 2 : t == "rk3" ? 3 : (t == "rk4" || t == "ssprk43") ? 4 : -1;
  if (cfg.nstage < 0) throw std::runtime_error("unknown integrator: " + t);
  cfg.low_mach = p.get_bool("mhd", "low_mach", false);
  cfg.flatten = p.get_bool("mhd", "flatten", true);
  cfg.cfl = p.get_real("time", "cfl", 0.3);
  cfg.gamma = p.get_real("mhd", "gamma", 5.0 / 3.0);
  cfg.dfloor = p.get_real("mhd", "dfloor", 1.0e-10);
  cfg.pfloor = p.get_real("mhd", "pfloor", 1.0e-12);
  cfg.bc[0] = parse_bc(p.get_string("mesh", "bc_x1", "periodic"));
  cfg.bc[1] = parse_bc(p.get_string("mesh", "bc_x2", "periodic"));
  cfg.bc[2] = parse_bc(p.get_string("mesh", "bc_x3", "periodic"));
  cfg.fofc = p.get_bool("mhd", "fofc", true);
  const std::string pr = p.get_string("mhd", "precision", "fp64");
  cfg.precision = pr == "fp64" ? 0 : pr == "mixed" ? 1 : -1;
  if (cfg.precision < 0) throw std::runtime_error("unknown precision: " + pr + " (fp64 | mixed)");
  cfg.flux_limiter = p.get_bool("mhd", "flux_limiter", false) ? 1 : 0;   // main.cpp:141
  cfg.sts = p.get_bool("diffusion", "sts", false) ? 1 : 0;   // main.cpp:142-146
  cfg.nu = p.get_real("diffusion", "nu", 0.0);
  cfg.kappa = p.get_real("diffusion", "kappa", 0.0);
  cfg.eta = p.get_real("diffusion", "eta", 0.0);
  cfg.grav_uniform = p.get_bool("gravity", "uniform", false) ? 1 : 0;   // main.cpp:148-162
  cfg.g1 = p.get_real("gravity", "g1", 0.0);
  cfg.g2 = p.get_real("gravity", "g2", 0.0);
  cfg.g3 = p.get_real("gravity", "g3", 0.0);
  cfg.grav_point = p.get_bool("gravity", "point_mass", false) ? 1 : 0;
  cfg.gm = p.get_real("gravity", "gm", 0.0);
  cfg.gx0 = p.get_real("gravity", "x0", 0.0);
  cfg.gy0 = p.get_real("gravity", "y0", 0.0);
  cfg.gz0 = p.get_real("gravity", "z0", 0.0);
  cfg.rsoft = p.get_real("gravity", "rsoft", 0.0);
  cfg.cooling = p.get_bool("cooling", "enabled", false) ? 1 : 0;
  cfg.lambda0 = p.get_real("cooling", "lambda0", 0.0);
  cfg.cool_alpha = p.get_real("cooling", "alpha", 1.0);
  cfg.tfloor = p.get_real("cooling", "tfloor", 0.0);
  g.nx1 = p.get_int("mesh", "nx1");
  g.nx2 = p.get_int("mesh", "nx2", 1);
  g.nx3 = p.get_int("mesh", "nx3", 1);
  const int ngmin = (cfg.recon == 2 || cfg.recon == 3 || cfg.recon == 4) ? 3 : 2;
  g.ng = p.get_int("mesh", "nghost", ngmin);
  if (g.ng < ngmin) throw std::runtime_error("nghost too small for the reconstruction method");
  g.x1min = p.get_real("mesh", "x1min", 0.0); g.x1max = p.get_real("mesh", "x1max", 1.0);
  g.x2min = p.get_real("mesh", "x2min", 0.0); g.x2max = p.get_real("mesh", "x2max", 1.0);
  g.x3min = p.get_real("mesh", "x3min", 0.0); g.x3max = p.get_real("mesh", "x3max", 1.0);
  g.build();
  mesh.nx1 = g.nx1; mesh.nx2 = g.nx2; mesh.nx3 = g.nx3; mesh.ng = g.ng;
  mesh.x1min = g.x1min; mesh.x1max = g.x1max; mesh.x2min = g.x2min; mesh.x2max = g.x2max;
  mesh.x3min = g.x3min; mesh.x3max = g.x3max;
  tlim = p.get_real("time", "tlim", 1.0);
  nlim = p.get_int("time", "nlim", -1);
  ndiag = p.get_int("time", "ndiag", 100);
  problem = p.get_string("problem", "name");
  basename = p.get_string("output", "basename", problem);
  vtk_dt = p.get_real("output", "vtk_dt", -1.0);
  hst_dt = p.get_real("output", "hst_dt", -1.0);
  rst_dt = p.get_real("output", "rst_dt", -1.0);
  compute_error = p.get_bool("output", "compute_error", false);
}

void Deck::initialize() {
  const size_t nc = g.ncell(), nf = g.nface();
  // every element is written below (no zero fill)
  w.resize(6 * nc);
  u0.resize(6 * nc);
  b1.resize(nf);
  b2.resize(nf);
  b3.resize(nf);
  Fields h{g, w, b1, b2, b3};
  if (problem == "cpaw") cpaw(p, h);
  else if (problem == "orszag_tang") orszag_tang(p, h);
  else if (problem == "blast") blast(p, h);
  else if (problem == "field_loop") field_loop(p, h);
  else if (problem == "shock_tube") shock_tube(p, h);
  else if (problem == "diffusion_test") diffusion_test(p, h);
  else if (problem == "uniform") uniform(p, h);
  else if (problem == "gresho") gresho(p, h);
  else throw std::runtime_error("unknown problem: " + problem);
  // conserved variables from primitives and face fields, pgen.hpp:286-301
  const double gm1 = cfg.gamma - 1.0;
  for_rows(g.nk, g.nj, g.ni, [&](int k, int j) {
      for (int i = 0; i < g.ni; ++i) {
        const size_t c = g.c(k, j, i);
        const double bx = 0.5 * (b1[g.f(k, j, i)] + b1[g.f(k, j, i + 1)]);
        const double by = g.two_d ? 0.5 * (b2[g.f(k, j, i)] + b2[g.f(k, j + 1, i)]) : b2[g.f(k, j, i)];
        const double bz = g.three_d ? 0.5 * (b3[g.f(k, j, i)] + b3[g.f(k + 1, j, i)]) : b3[g.f(k, j, i)];
        const double d = w[c], vx = w[nc + c], vy = w[2 * nc + c], vz = w[3 * nc + c], pr = w[4 * nc + c];
        u0[c] = d;
        u0[nc + c] = d * vx;
        u0[2 * nc + c] = d * vy;
        u0[3 * nc + c] = d * vz;
        u0[4 * nc + c] = pr / gm1 + 0.5 * d * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz);
        w[5 * nc + c] = pr / std::pow(d, cfg.gamma);
        u0[5 * nc + c] = d * w[5 * nc + c];
      }
  });
}

}  // namespace chq

// ---------------------------------------------------------------- C ABI
struct chq_deck {
  chq::Deck d;
};

namespace {
thread_local std::string g_deck_err;
}

extern "C" {

const char *chq_deck_last_error(void) { return g_deck_err.c_str(); }

namespace {
int open_deck(const char *text, const char *overrides, bool lazy, chq_deck **out) {
  if (!out || !text) { g_deck_err = "null argument"; return CHQ_EINVAL; }
  *out = nullptr;
  chq_deck *k = new chq_deck();
  try {
    k->d.p.read_text(text);
    if (overrides) {
      std::istringstream in(overrides);
      std::string line;
      while (std::getline(in, line)) {
        const auto a = line.find_first_not_of(" \t\r");
        if (a == std::string::npos) continue;
        k->d.p.override_kv(line.substr(a));
      }
    }
    k->d.configure();
    if (!lazy) {
      k->d.initialize();
      k->d.ready = true;
    }
  } catch (const std::exception &e) {
    g_deck_err = e.what();
    delete k;
    return CHQ_EINVAL;
  }
  *out = k;
  return CHQ_OK;
}

// the initial state of a lazily opened deck is built on first use
int ensure_state(const chq_deck *k) {
  if (k->d.ready) return CHQ_OK;
  try {
    auto &d = const_cast<chq::Deck &>(k->d);
    d.initialize();
    d.ready = true;
  } catch (const std::exception &e) {
    g_deck_err = e.what();
    return CHQ_EINVAL;
  }
  return CHQ_OK;
}
}  // namespace

int chq_deck_open(const char *text, const char *overrides, chq_deck **out) {
  return open_deck(text, overrides, false, out);
}

int chq_deck_open_lazy(const char *text, const char *overrides, chq_deck **out) {
  return open_deck(text, overrides, true, out);
}

int chq_deck_prepare(chq_deck *k) {
  if (!k) return CHQ_EINVAL;
  return ensure_state(k);
}

int chq_deck_close(chq_deck *k) {
  delete k;
  return CHQ_OK;
}

int chq_deck_mesh(const chq_deck *k, chq_mesh_desc *m) {
  if (!k || !m) return CHQ_EINVAL;
  *m = k->d.mesh;
  return CHQ_OK;
}

int chq_deck_config(const chq_deck *k, chq_config *c) {
  if (!k || !c) return CHQ_EINVAL;
  *c = k->d.cfg;
  return CHQ_OK;
}

int chq_deck_run(const chq_deck *k, double *tlim, int *nlim, int *ndiag) {
  if (!k) return CHQ_EINVAL;
  if (tlim) *tlim = k->d.tlim;
  if (nlim) *nlim = k->d.nlim;
  if (ndiag) *ndiag = k->d.ndiag;
  return CHQ_OK;
}

int chq_deck_initial(const chq_deck *k, int array, void *host, size_t count) {
  if (!k || !host) return CHQ_EINVAL;
  if (int e = ensure_state(k)) return e;
  const chq::HostArray *v = nullptr;
  switch (array) {
    case CHQ_U0: v = &k->d.u0; break;
    case CHQ_W: v = &k->d.w; break;
    case CHQ_B1F0: v = &k->d.b1; break;
    case CHQ_B2F0: v = &k->d.b2; break;
    case CHQ_B3F0: v = &k->d.b3; break;
    default: g_deck_err = "the deck holds u0, w and b1f0..b3f0 only"; return CHQ_EINVAL;
  }
  if (count != v->size()) { g_deck_err = "count mismatch"; return CHQ_EINVAL; }
  std::memcpy(host, v->data(), count * sizeof(double));
  return CHQ_OK;
}

const char *chq_deck_get(const chq_deck *k, const char *block, const char *key) {
  static thread_local std::string v;
  if (!k || !block || !key || !k->d.p.has(block, key)) return "";
  v = k->d.p.get_string(block, key);
  return v.c_str();
}

size_t chq_deck_print(const chq_deck *k, char *buf, size_t len) {
  if (!k) return 0;
  std::ostringstream o;
  k->d.p.print(o);
  const std::string s = o.str();
  if (buf && len > 0) {
    const size_t n = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return s.size() + 1;
}

int chq_deck_upload(const chq_deck *k, chq_ctx *ctx) {
  if (!k || !ctx) return CHQ_EINVAL;
  if (int e = ensure_state(k)) return e;
  // the primitives are not uploaded: chq_init derives them from u0 and the face fields with
  // con2prim over every cell, as the reference's driver does before its loop (main.cpp:200-201)
  int rc = chq_upload(ctx, CHQ_U0, k->d.u0.data(), k->d.u0.size());
  if (!rc) rc = chq_upload(ctx, CHQ_B1F0, k->d.b1.data(), k->d.b1.size());
  if (!rc) rc = chq_upload(ctx, CHQ_B2F0, k->d.b2.data(), k->d.b2.size());
  if (!rc) rc = chq_upload(ctx, CHQ_B3F0, k->d.b3.data(), k->d.b3.size());
  return rc;
}

}  // extern "C"
