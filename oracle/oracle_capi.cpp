// CPU oracle — C entry points for ctypes (tests/, smoke(), bench.py's CPU-baseline leg).
// TEST INFRASTRUCTURE ONLY: the product library never links or calls this.
#include <omp.h>
#include <cstring>
#include <exception>
#include <string>
#include "oracle.hpp"

using namespace orc;

namespace {
thread_local std::string g_err;
void set_err(char *err, int len, const std::string &msg) {
  g_err = msg;
  if (err && len > 0) { std::strncpy(err, msg.c_str(), size_t(len) - 1); err[len - 1] = 0; }
}
Sim *S(void *p) { return static_cast<Sim *>(p); }
}  // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int orc_get_threads() { return omp_get_max_threads(); }

// deck_text: INI text; overrides: newline-separated block.key=value lines (may be null)
void *orc_create_ex(const char *deck_text, const char *overrides, int pgen_only, char *err, int errlen) {
  try {
    Sim *s = new Sim();
    s->deck.parse(deck_text ? deck_text : "");
    if (overrides) {
      std::string o(overrides), line;
      std::istringstream in(o);
      while (std::getline(in, line)) {
        line = Deck::trim(line);
        if (!line.empty()) s->deck.override_kv(line);
      }
    }
    s->setup(pgen_only != 0);
    return s;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

void *orc_create(const char *deck_text, const char *overrides, char *err, int errlen) {
  return orc_create_ex(deck_text, overrides, 0, err, errlen);
}

void orc_destroy(void *p) { delete S(p); }

// iv[16] = nx1 nx2 nx3 ng ni nj nk is ie js je ks ke two_d three_d cycle
// rv[12] = dx1 dx2 dx3 t tlim min_dt last_dt x1min x2min x3min cfl gamma
void orc_geometry(void *p, int *iv, double *rv) {
  const Sim &s = *S(p);
  const Grid &m = s.m;
  const int a[16] = {m.nx1, m.nx2, m.nx3, m.ng, m.ni, m.nj, m.nk, m.is, m.ie, m.js, m.je, m.ks, m.ke,
                     m.two_d, m.three_d, s.cycle};
  std::memcpy(iv, a, sizeof(a));
  const double b[12] = {m.dx1, m.dx2, m.dx3, s.t, s.tlim, s.min_dt, s.last_dt, m.x1min, m.x2min, m.x3min,
                        s.cfg.cfl, s.cfg.eos.gamma};
  std::memcpy(rv, b, sizeof(b));
}

double *orc_array(void *p, const char *name, long *count) {
  Grid &m = S(p)->m;
  const std::string n(name);
  std::vector<Real> *v = nullptr;
  if (n == "u0") v = &m.u0.v; else if (n == "u1") v = &m.u1.v; else if (n == "w") v = &m.w.v;
  else if (n == "bcc") v = &m.bcc.v; else if (n == "ecc") v = &m.ecc.v;
  else if (n == "b1f0") v = &m.b1f0.v; else if (n == "b2f0") v = &m.b2f0.v; else if (n == "b3f0") v = &m.b3f0.v;
  else if (n == "b1f1") v = &m.b1f1.v; else if (n == "b2f1") v = &m.b2f1.v; else if (n == "b3f1") v = &m.b3f1.v;
  else if (n == "f1") v = &m.f1.v; else if (n == "f2") v = &m.f2.v; else if (n == "f3") v = &m.f3.v;
  else if (n == "e2x1") v = &m.e2x1.v; else if (n == "e3x1") v = &m.e3x1.v; else if (n == "e1x2") v = &m.e1x2.v;
  else if (n == "e3x2") v = &m.e3x2.v; else if (n == "e1x3") v = &m.e1x3.v; else if (n == "e2x3") v = &m.e2x3.v;
  else if (n == "e1") v = &m.e1.v; else if (n == "e2") v = &m.e2.v; else if (n == "e3") v = &m.e3.v;
  else if (n == "u_init") v = &S(p)->u_init.v; else if (n == "bcc_init") v = &S(p)->bcc_init.v;
  if (!v) { g_err = "unknown array " + n; if (count) *count = 0; return nullptr; }
  if (count) *count = long(v->size());
  return v->data();
}

int *orc_flag(void *p, long *count) {
  Grid &m = S(p)->m;
  if (count) *count = long(m.flag.size());
  return m.flag.data();
}

// runs up to n steps of the time loop; returns the number taken
int orc_step(void *p, int n) {
  int k = 0;
  while (k < n && S(p)->step()) ++k;
  return k;
}

void orc_l1_error(void *p, double *err10) { S(p)->l1_error(err10); }

int orc_restart_write(void *p, const char *path) {
  try { S(p)->restart_write(path); } catch (const std::exception &e) { g_err = e.what(); return -1; }
  return 0;
}
int orc_restart_read(void *p, const char *path) {
  try { S(p)->restart_read(path); } catch (const std::exception &e) { g_err = e.what(); return -1; }
  return 0;
}
long orc_fofc_count(void *p) { return S(p)->fofc_count; }

// ---- single kernels, acting on the mesh state exactly as the reference kernels do
int orc_fluxes(void *p, int dir, int cur) {
  Sim &s = *S(p);
  fluxes_any(s.m, s.cfg, dir, s.m.bf(cur, dir));
  return 0;
}
int orc_corner_emf(void *p) { corner_emf(S(p)->m); return 0; }
int orc_update(void *p, int out, double g0, double g1, double bdt) { update(S(p)->m, out, g0, g1, bdt); return 0; }
int orc_boundaries(void *p, int s) { boundaries(S(p)->m, S(p)->cfg, s); return 0; }
double orc_cons_to_prim(void *p, int s, int compute_dt) { return cons_to_prim(S(p)->m, S(p)->cfg, s, compute_dt != 0); }
int orc_fofc_flag(void *p, int s) { return fofc_flag(S(p)->m, S(p)->cfg, s); }
void orc_diffusion_fluxes(void *p, int dir, int cur) {
  Sim &s = *S(p);
  const Diff dc = diff_of(s.cfg);
  const A3 &b1 = s.m.bf(cur, 0), &b2 = s.m.bf(cur, 1), &b3 = s.m.bf(cur, 2);
  if (dir == 0) diffusion_fluxes<0>(s.m, dc, b1, b2, b3);
  else if (dir == 1) diffusion_fluxes<1>(s.m, dc, b1, b2, b3);
  else diffusion_fluxes<2>(s.m, dc, b1, b2, b3);
}
void orc_resistive_emf(void *p, int cur) {
  Sim &s = *S(p);
  resistive_emf(s.m, diff_of(s.cfg), s.m.bf(cur, 0), s.m.bf(cur, 1), s.m.bf(cur, 2));
}
double orc_diffusion_dt(void *p) {
  Sim &s = *S(p);
  return diffusion_dt(s.m, s.cfg, diff_of(s.cfg));
}
long orc_sts_count(void *p) { return S(p)->sts_count; }
int orc_fofc_repair(void *p, int dir, int cur) {
  Sim &s = *S(p);
  fofc_repair_any(s.m, s.cfg, dir, cur, s.cfg.sts ? Diff{} : diff_of(s.cfg));
  return 0;
}
long orc_limiter_count(void *p) { return S(p)->limiter_count; }
long orc_limit_fluxes(void *p, int dir, double bdt) {
  Sim &s = *S(p);
  if (dir == 0) return limit_fluxes<0>(s.m, s.cfg, bdt);
  if (dir == 1) return limit_fluxes<1>(s.m, s.cfg, bdt);
  return limit_fluxes<2>(s.m, s.cfg, bdt);
}
void orc_apply_gravity(void *p, int out, double bdt) {
  Sim &s = *S(p);
  apply_gravity(s.m, s.cfg, s.m.u(out), bdt);
}
void orc_apply_cooling(void *p, int buf, double dt) {
  Sim &s = *S(p);
  apply_cooling(s.m, s.cfg, s.m.u(buf), dt);
}
long orc_stage(void *p, int cur, int out, double g0, double g1, double bdt) {
  Sim &s = *S(p);
  return stage(s.m, s.cfg, cur, out, g0, g1, bdt);
}
// overwrite the floors / switches after setup (used to force fofc in tests)
void orc_set_floors(void *p, double dfloor, double pfloor) { S(p)->cfg.eos.dfloor = dfloor; S(p)->cfg.eos.pfloor = pfloor; }

// ---- pointwise numerics, vectorised over n samples (unit tests)
void orc_reconstruct(int recon, const double *q, long n, double *L, double *R) {
  for (long s = 0; s < n; ++s) reconstruct(Recon(recon), q + 6 * s, L[s], R[s]);
}
void orc_flatten(const double *p5, const double *um1, const double *up1, long n, double *out) {
  for (long s = 0; s < n; ++s) out[s] = flatten_coef(p5 + 5 * s, um1[s], up1[s]);
}
// solver: 0 hlle, 1 hlld, 2 hlld with explicit (theta, phi) and LHLLD speeds
void orc_riemann(int solver, const double *L7, const double *R7, const double *bn, long n, double gamma,
                 const double *theta, const double *phi, double *f7) {
  Eos eos;
  eos.gamma = gamma; eos.gm1 = gamma - 1.0;
  for (long s = 0; s < n; ++s) {
    const double *a = L7 + 7 * s, *b = R7 + 7 * s;
    Prim L{a[0], a[1], a[2], a[3], a[4], a[5], a[6]}, R{b[0], b[1], b[2], b[3], b[4], b[5], b[6]};
    if (solver == 0) hlle(L, R, bn[s], eos, f7 + 7 * s);
    else if (solver == 1) hlld(L, R, bn[s], eos, f7 + 7 * s);
    else hlld(L, R, bn[s], eos, f7 + 7 * s, theta[s], phi[s], true);
  }
}
void orc_phys_flux(const double *S7, const double *bn, long n, double gamma, double *f7) {
  for (long s = 0; s < n; ++s) {
    const double *a = S7 + 7 * s;
    Prim P{a[0], a[1], a[2], a[3], a[4], a[5], a[6]};
    double e, pt;
    phys_flux(P, bn[s], gamma - 1.0, e, pt, f7 + 7 * s);
  }
}
double orc_fast_speed(const double *S7, double bn, double gamma) {
  Prim P{S7[0], S7[1], S7[2], S7[3], S7[4], S7[5], S7[6]};
  return fast_speed(P, bn, gamma);
}
double orc_noise(unsigned long long seed, unsigned long long idx) { return noise_pm1(seed, idx); }
// eigensystem of n states st7 = (d, p, by, bz, bx) x n: L and R as n row-major 7x7 blocks
void orc_eigensystem(const double *st5, long n, double gamma, double *L, double *R) {
  for (long s = 0; s < n; ++s) {
    const double *a = st5 + 5 * s;
    eigensystem(a[0], a[1], a[2], a[3], a[4], gamma, reinterpret_cast<double(*)[7]>(L + 49 * s),
                reinterpret_cast<double(*)[7]>(R + 49 * s));
  }
}
// one side of a face in characteristic variables: st = 7x6 stencils per sample, bx per sample
void orc_charac_side(int recon, const double *st, const double *bx, long n, double gamma, int cell, int side,
                     double *out7) {
  for (long s = 0; s < n; ++s)
    charac_side(Recon(recon), reinterpret_cast<const double(*)[6]>(st + 42 * s), cell, bx[s], gamma, side,
                out7 + 7 * s);
}

}  // extern "C"
