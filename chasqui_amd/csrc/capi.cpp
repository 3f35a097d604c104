// C ABI of the stage pipeline (include/chasqui_amd.h): context, device arena, transfers,
// the per-kernel entry points with the reference's semantics, and the device-resident
// time loop.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace chq {
hipError_t launch_relayout(void *ref, void *dev, const Geom &g, int nc, int nk, int nj, int ni, int to_dev,
                           int elem_bytes, hipStream_t s);
}

namespace {
thread_local std::string g_err;
}  // namespace

int chq::fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}


namespace {

template <class T>
void field_desc(const FieldSet<T> &f, int wcur, int id, void **dev, int *nc, int *face) {
  switch (id) {
    case CHQ_W: *dev = f.w[wcur]; *nc = 6; *face = 0; break;
    case CHQ_BCC: *dev = f.bcc[wcur]; *nc = 3; *face = 0; break;
    case CHQ_ECC: *dev = f.ecc[wcur]; *nc = 3; *face = 0; break;
    case CHQ_F1: case CHQ_F2: case CHQ_F3: *dev = f.flx[id - CHQ_F1]; *nc = 6; break;
    case CHQ_E1: case CHQ_E2: case CHQ_E3: *dev = f.ee[id - CHQ_E1]; break;
    default: *dev = f.ef[id - CHQ_E2X1]; break;
  }
}

// Device pointer, component count, face (+1) extent and device element size of an array.
// `wide` = 1 when the device holds float and the host side is double (mixed precision).
bool array_desc(const chq_ctx *c, int id, void **dev, int *nc, int *face, int *ebytes, int *wide) {
  *face = 1; *ebytes = 8; *nc = 1; *dev = nullptr; *wide = 0;
  switch (id) {
    case CHQ_U0: *dev = c->u[0]; *nc = 6; *face = 0; break;
    case CHQ_U1: *dev = c->u[1]; *nc = 6; *face = 0; break;
    case CHQ_W: case CHQ_BCC: case CHQ_ECC: case CHQ_F1: case CHQ_F2: case CHQ_F3:
    case CHQ_E2X1: case CHQ_E3X1: case CHQ_E1X2: case CHQ_E3X2: case CHQ_E1X3: case CHQ_E2X3:
    case CHQ_E1: case CHQ_E2: case CHQ_E3:
      if (c->mixed) { field_desc(c->f32, c->wcur, id, dev, nc, face); *ebytes = 4; *wide = 1; }
      else field_desc(c->f64, c->wcur, id, dev, nc, face);
      break;
    case CHQ_B1F0: *dev = c->bf[0][0]; break;
    case CHQ_B2F0: *dev = c->bf[0][1]; break;
    case CHQ_B3F0: *dev = c->bf[0][2]; break;
    case CHQ_B1F1: *dev = c->bf[1][0]; break;
    case CHQ_B2F1: *dev = c->bf[1][1]; break;
    case CHQ_B3F1: *dev = c->bf[1][2]; break;
    case CHQ_FLAG: *dev = c->flag; *face = 0; *ebytes = 4; break;
    default: return false;
  }
  return true;
}

// The time loop flips the current primitive set on the device, and only in stages that ran
// (a run that reached tlim stops flipping): read it back before the host uses it.
int sync_wcur(chq_ctx *c) {
  if (!c->wcur_stale) return CHQ_OK;
  int v = 0;
  HIPCHK(hipMemcpyAsync(&v, &c->ctrl->wcur, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->wcur = v;
  c->wcur_stale = false;
  return CHQ_OK;
}

// The fast path's last RK stage leaves u[IENT] of the final state (buffer 0) to be
// resynchronised here, before the host reads the state or a per-kernel entry point runs on it.
int sync_ient(chq_ctx *c) {
  if (!c->ient_stale) return CHQ_OK;
  if (c->batch_stream) HIPCHK(hipStreamSynchronize(c->batch_stream));
  HIPCHK(launch_ient_sync(c->c2p_args(0, c->wcur, c->ctrl_api), c->stream));
  c->ient_stale = false;
  return CHQ_OK;
}
int sync_state(chq_ctx *c) {
  if (int rc = sync_wcur(c)) return rc;
  return sync_ient(c);
}

int relayout(chq_ctx *c, int id, void *host, size_t count, bool to_dev) {
  if (c->batch_stream) HIPCHK(hipStreamSynchronize(c->batch_stream));
  if (int rc = sync_wcur(c)) return rc;
  // the deferred u[IENT] of buffer 0 is stored before its state is read or its face B is
  // replaced (it is computed from both); an upload of u0 replaces every component
  if (id == CHQ_U0 && to_dev) c->ient_stale = false;
  else if (id == CHQ_U0 || (to_dev && (id == CHQ_B1F0 || id == CHQ_B2F0 || id == CHQ_B3F0)))
    if (int rc = sync_ient(c)) return rc;
  void *dev;
  int nc, face, eb, wide;
  if (!array_desc(c, id, &dev, &nc, &face, &eb, &wide)) return fail(CHQ_EINVAL, "unknown array id");
  const Geom &g = c->g;
  const int nk = g.nk + face, nj = g.nj + face, ni = g.ni + face;
  const size_t n = size_t(nc) * nk * nj * ni;
  if (count != n) return fail(CHQ_EINVAL, "host buffer has " + std::to_string(count) + " elements, expected " + std::to_string(n));
  if (!host) return fail(CHQ_EINVAL, "null host buffer");
  const size_t bytes = n * (wide ? 8 : eb);   // host element size
  if (bytes > c->staging_bytes) {
    if (c->staging) HIPCHK(hipFree(c->staging));
    c->staging = nullptr;
    c->staging_bytes = 0;
    HIPCHK(hipMalloc(&c->staging, bytes));
    c->staging_bytes = bytes;
  }
  if (to_dev) {
    HIPCHK(hipMemcpyAsync(c->staging, host, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_relayout(c->staging, dev, g, nc, nk, nj, ni, 1, wide ? -4 : eb, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    HIPCHK(launch_relayout(c->staging, dev, g, nc, nk, nj, ni, 0, wide ? -4 : eb, c->stream));
    HIPCHK(hipMemcpyAsync(host, c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return CHQ_OK;
}

}  // namespace

int chq::ctx_ready(chq_ctx *c) {
  if (c->batch_stream) HIPCHK(hipStreamSynchronize(c->batch_stream));
  return sync_wcur(c);
}

namespace {

// Mesh::Allocate, mesh.hpp:37-48, plus the padded device layout
void fill_geom(const chq_mesh_desc *m, Geom &g) {
  g = Geom{};
  g.nx1 = m->nx1; g.nx2 = m->nx2; g.nx3 = m->nx3; g.ng = m->ng;
  g.ni = g.nx1 + 2 * g.ng;
  g.nj = (g.nx2 > 1) ? g.nx2 + 2 * g.ng : 1;
  g.nk = (g.nx3 > 1) ? g.nx3 + 2 * g.ng : 1;
  g.is = g.ng; g.ie = g.ng + g.nx1 - 1;
  g.js = (g.nx2 > 1) ? g.ng : 0; g.je = (g.nx2 > 1) ? g.ng + g.nx2 - 1 : 0;
  g.ks = (g.nx3 > 1) ? g.ng : 0; g.ke = (g.nx3 > 1) ? g.ng + g.nx3 - 1 : 0;
  g.two_d = g.nx2 > 1; g.three_d = g.nx3 > 1;
  g.dx1 = (m->x1max - m->x1min) / g.nx1;
  g.dx2 = (m->x2max - m->x2min) / g.nx2;
  g.dx3 = (m->x3max - m->x3min) / g.nx3;
  g.idx1 = 1.0 / g.dx1; g.idx2 = 1.0 / g.dx2; g.idx3 = 1.0 / g.dx3;   // physics.hpp:309
  g.ioff = (16 - g.is % 16) % 16;
  g.pitch = ((g.ioff + g.ni + 1 + 15) / 16) * 16;
  g.plane = (long long)g.pitch * (g.nj + 1);
  g.cstride = g.plane * (g.nk + 1);
}

// device arena of a context: conserved state (12), face B (6), flags (2) as double-sized
// components, then the field set (w 12, bcc 6, ecc 6, fluxes 18, face EMFs 6, edge EMFs 3) in
// double (FP64) or float (mixed), padded to 256 B, then two 256-B control blocks
size_t arena_bytes_for(const Geom &g, bool mixed) {
  const size_t comp = size_t(g.cstride) * sizeof(double);
  const size_t nfield = 12 + 6 + 6 + 18 + 6 + 3;
  const size_t field_bytes = nfield * size_t(g.cstride) * (mixed ? sizeof(float) : sizeof(double));
  return (12 + 6 + 2) * comp + ((field_bytes + 255) / 256) * 256 + 2 * 256;
}

int check_cfg(const chq_mesh_desc *m, const chq_config *c) {
  if (!m || !c) return fail(CHQ_EINVAL, "null mesh or config");
  if (m->nx1 < 1 || m->nx2 < 1 || m->nx3 < 1) return fail(CHQ_EINVAL, "mesh sizes must be >= 1");
  if (m->nx3 > 1 && m->nx2 <= 1) return fail(CHQ_EINVAL, "a 3D mesh needs nx2 > 1");
  if (c->recon < 0 || c->recon > 4) return fail(CHQ_EINVAL, "unknown reconstruction");
  if (c->precision < 0 || c->precision > 1) return fail(CHQ_EINVAL, "precision must be 0 (fp64) or 1 (mixed)");
  if (c->precision == 1 && c->dual_energy) return fail(CHQ_EUNSUP, "mixed precision with dual energy is not supported");
  if (c->precision == 1 && c->characteristic && c->recon != 0)
    return fail(CHQ_EUNSUP, "mixed precision with characteristic reconstruction is not supported");
  if (c->precision == 1 && (c->nu > 0.0 || c->kappa > 0.0 || c->eta > 0.0 || c->sts))
    return fail(CHQ_EUNSUP, "mixed precision with diffusion is not supported");
  if (c->precision == 1 && (c->flux_limiter || c->grav_uniform || c->grav_point || c->cooling))
    return fail(CHQ_EUNSUP, "mixed precision with the flux limiter or source terms is not supported");
  if (c->rsolver < 0 || c->rsolver > 2) return fail(CHQ_EINVAL, "unknown Riemann solver");
  if (c->nstage < 1 || c->nstage > 4) return fail(CHQ_EINVAL, "nstage must be 1..4");
  const int ngmin = (c->recon == 2 || c->recon == 3 || c->recon == 4) ? 3 : 2;   // main.cpp:169-171
  if (m->ng < ngmin) return fail(CHQ_EINVAL, "nghost too small for the reconstruction method");
  for (int d = 0; d < 3; ++d)
    if (c->bc[d] < 0 || c->bc[d] > 2) return fail(CHQ_EINVAL, "unknown boundary condition");
  // cell and thread indices of the kernels are 32-bit (far above what fits in one GPU's HBM)
  const long long padded = (long long)(m->nx1 + 2 * m->ng) * (m->nx2 > 1 ? m->nx2 + 2 * m->ng : 1) *
                           (m->nx3 > 1 ? m->nx3 + 2 * m->ng : 1);
  if (padded >= (1LL << 30)) return fail(CHQ_EINVAL, "mesh too large: 2^30 cells or more, ghosts included");
  // a periodic ghost fill reads nx cells back from the ghost: it needs nx >= ng
  if (m->nx1 < m->ng || (m->nx2 > 1 && m->nx2 < m->ng) || (m->nx3 > 1 && m->nx3 < m->ng))
    return fail(CHQ_EINVAL, "every multi-cell dimension needs at least nghost active cells");
  return CHQ_OK;
}

}  // namespace

extern "C" {

const char *chq_last_error(void) { return g_err.c_str(); }
const char *chq_version(void) { return "chasqui_amd 0.1 (gfx950)"; }

int chq_create(const chq_mesh_desc *m, const chq_config *cfg, chq_ctx **out) {
  if (!out) return fail(CHQ_EINVAL, "null output pointer");
  *out = nullptr;
  int rc = check_cfg(m, cfg);
  if (rc) return rc;
  chq_ctx *c = new (std::nothrow) chq_ctx();
  if (!c) return fail(CHQ_ENOMEM, "host allocation failed");
  c->cfg = *cfg;
  Geom &g = c->g;
  fill_geom(m, g);
  c->x1min = m->x1min; c->x2min = m->x2min; c->x3min = m->x3min;
  c->eos.gamma = cfg->gamma;
  c->eos.gm1 = cfg->gamma - 1.0;
  c->eos.dfloor = cfg->dfloor;
  c->eos.pfloor = cfg->pfloor;
  c->eos.dual_energy = cfg->dual_energy;
  c->eos.de_eta = cfg->de_eta;
  c->entropy = cfg->dual_energy != 0;
  c->mixed = cfg->precision == 1;

  auto cleanup = [&](int code) { chq_destroy(c); return code; };
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) return cleanup(fail(CHQ_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e)));
  if (int rs = history_scratch(c->device, &c->hst_scratch)) return cleanup(rs);
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess)
    return cleanup(fail(CHQ_EHIP, std::string("stream/event: ") + hipGetErrorString(e)));
  const size_t comp = size_t(g.cstride) * sizeof(double);
  c->arena_bytes = arena_bytes_for(g, c->mixed);
  e = hipMalloc((void **)&c->arena, c->arena_bytes);
  if (e != hipSuccess) { c->arena = nullptr; return cleanup(fail(CHQ_ENOMEM, "device arena of " + std::to_string(c->arena_bytes) + " bytes: " + hipGetErrorString(e))); }
  e = hipMemsetAsync(c->arena, 0, c->arena_bytes, c->stream);
  if (e != hipSuccess) return cleanup(fail(CHQ_EHIP, "arena clear"));
  char *p = c->arena;
  auto take = [&](int n) { double *r = (double *)p; p += size_t(n) * comp; return r; };
  for (int s = 0; s < 2; ++s) c->u[s] = take(6);
  for (int s = 0; s < 2; ++s) for (int d = 0; d < 3; ++d) c->bf[s][d] = take(1);
  c->flag = (int *)take(1);
  c->badlist = (int *)take(1);
  auto carve = [&](auto &f) {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(f.flx[0])>>;
    auto t = [&](int n) { T *r = (T *)p; p += size_t(n) * size_t(g.cstride) * sizeof(T); return r; };
    for (int s = 0; s < 2; ++s) f.w[s] = t(6);
    for (int s = 0; s < 2; ++s) f.bcc[s] = t(3);
    for (int s = 0; s < 2; ++s) f.ecc[s] = t(3);
    for (int d = 0; d < 3; ++d) f.flx[d] = t(6);
    for (int q = 0; q < 6; ++q) f.ef[q] = t(1);
    for (int q = 0; q < 3; ++q) f.ee[q] = t(1);
  };
  if (c->mixed) carve(c->f32); else carve(c->f64);
  // float arenas end on a 4-byte multiple of cstride elements; keep the control blocks aligned
  p = c->arena + (((p - c->arena) + 255) / 256) * 256;
  c->ctrl = (Ctrl *)p;
  c->ctrl_api = (Ctrl *)(p + 256);
  Ctrl h{};
  h.active = 1;
  h.tlim = 1e300;
  h.nlim = -1;
  h.cfl = cfg->cfl;
  e = hipMemcpyAsync(c->ctrl_api, &h, sizeof(h), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(c->ctrl, &h, sizeof(h), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return cleanup(fail(CHQ_EHIP, std::string("control block: ") + hipGetErrorString(e)));
  *out = c;
  return CHQ_OK;
}

int chq_destroy(chq_ctx *c) {
  if (!c) return CHQ_OK;
  if (c->batch) (void)chq_batch_remove(static_cast<chq_batch *>(c->batch), c);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->arena) (void)hipFree(c->arena);
  if (c->staging) (void)hipFree(c->staging);
  if (c->hst_work) (void)hipFree(c->hst_work);
  if (c->sts_buf) (void)hipFree(c->sts_buf);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CHQ_OK;
}

int chq_geometry(const chq_ctx *c, int *iv, double *dx) {
  if (!c) return fail(CHQ_EINVAL, "null context");
  const Geom &g = c->g;
  if (iv) {
    const int a[16] = {g.nx1, g.nx2, g.nx3, g.ng, g.ni, g.nj, g.nk, g.is, g.ie, g.js, g.je, g.ks, g.ke,
                       g.two_d, g.three_d, g.pitch};
    std::memcpy(iv, a, sizeof(a));
  }
  if (dx) { dx[0] = g.dx1; dx[1] = g.dx2; dx[2] = g.dx3; }
  return CHQ_OK;
}

size_t chq_device_bytes(const chq_ctx *c) { return c ? c->arena_bytes + c->staging_bytes + c->hst_bytes : 0; }

size_t chq_bytes_for(const chq_mesh_desc *m, const chq_config *cfg) {
  if (check_cfg(m, cfg)) return 0;
  Geom g;
  fill_geom(m, g);
  return arena_bytes_for(g, cfg->precision == 1);
}

size_t chq_array_count(const chq_ctx *c, int id) {
  void *dev;
  int nc, face, eb, wide;
  if (!c || !array_desc(c, id, &dev, &nc, &face, &eb, &wide)) return 0;
  return size_t(nc) * (c->g.nk + face) * (c->g.nj + face) * (c->g.ni + face);
}

int chq_upload(chq_ctx *c, int id, const void *host, size_t count) {
  if (!c) return fail(CHQ_EINVAL, "null context");
  return relayout(c, id, const_cast<void *>(host), count, true);
}

int chq_download(chq_ctx *c, int id, void *host, size_t count) {
  if (!c) return fail(CHQ_EINVAL, "null context");
  return relayout(c, id, host, count, false);
}

// ---------------------------------------------------------------- per-kernel entry points
int chq_fluxes(chq_ctx *c, int dir, int cur) {
  if (!c || dir < 0 || dir > 2 || (cur & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if ((dir == 1 && !c->g.two_d) || (dir == 2 && !c->g.three_d)) return fail(CHQ_EINVAL, "direction not present in this mesh");
  if (c->mixed) {   // FP32 fluxes; the entropy component is not computed on this path
    FluxArgsT<float> a = c->flux_args_t(c->f32, dir, cur, c->wcur, c->ctrl_api);
    a.entropy = 0;
    HIPCHK(launch_flux(a, dir, Recon(c->cfg.recon), RSolver(c->cfg.rsolver), false, c->stream));
    return CHQ_OK;
  }
  FluxArgs a = c->flux_args(dir, cur, c->wcur, c->ctrl_api);
  HIPCHK(launch_flux(a, dir, Recon(c->cfg.recon), RSolver(c->cfg.rsolver), true, c->stream));
  return CHQ_OK;
}

int chq_corner_emf(chq_ctx *c) {
  if (!c) return fail(CHQ_EINVAL, "null context");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) HIPCHK(launch_corner<float>(c->corner_args_t(c->f32, c->wcur, c->ctrl_api, 0), c->stream));
  else HIPCHK(launch_corner<double>(c->corner_args(c->wcur, c->ctrl_api, 0), c->stream));
  return CHQ_OK;
}

int chq_update(chq_ctx *c, int out, double g0, double g1, double beta_dt) {
  if (!c || (out & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (c->mixed) {
    UpdateArgsT<float> a = c->update_args_t(c->f32, out, g0, g1, 5, c->ctrl_api, 0);
    a.use_ctrl_dt = 0;
    a.beta_dt = beta_dt;
    HIPCHK(launch_update<float>(a, c->stream));
    return CHQ_OK;
  }
  UpdateArgs a = c->update_args(out, g0, g1, 6, c->ctrl_api, 0);
  a.use_ctrl_dt = 0;
  a.beta_dt = beta_dt;
  HIPCHK(launch_update<double>(a, c->stream));
  return CHQ_OK;
}

int chq_boundaries(chq_ctx *c, int buf) {
  if (!c || (buf & ~1)) return fail(CHQ_EINVAL, "bad argument");
  HIPCHK(c->boundaries(buf, 6, c->ctrl_api, 0));
  return CHQ_OK;
}

int chq_cons_to_prim(chq_ctx *c, int buf, int compute_dt, double *min_dt) {
  if (!c || (buf & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  HIPCHK(launch_ctrl_begin_stage(c->ctrl_api, 1, c->stream));
  if (c->mixed) {
    C2pArgsT<float> a = c->c2p_args_t(c->f32, buf, c->wcur, c->ctrl_api);
    a.compute_dt = compute_dt ? 1 : 0;
    a.write_went = 0;
    HIPCHK(launch_c2p<float>(a, c->stream));
  } else {
    C2pArgs a = c->c2p_args(buf, c->wcur, c->ctrl_api);
    a.compute_dt = compute_dt ? 1 : 0;
    a.slot = 0;
    HIPCHK(launch_c2p<double>(a, c->stream));
  }
  if (min_dt) {
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, &c->ctrl_api->mdt_bits[0], sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *min_dt = compute_dt ? v : 1e300;
  }
  return CHQ_OK;
}

int chq_fofc_flag(chq_ctx *c, int buf, int *count) {
  if (!c || (buf & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_fofc_flag: FP64 contexts only");
  C2pArgs a = c->c2p_args(buf, c->wcur, c->ctrl_api);
  HIPCHK(launch_ctrl_begin_stage(c->ctrl_api, 0, c->stream));
  HIPCHK(launch_flag(a, c->stream));
  if (count) {
    HIPCHK(hipMemcpyAsync(count, &c->ctrl_api->nbad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return CHQ_OK;
}

int chq_fofc_repair(chq_ctx *c, int dir, int cur) {
  if (!c || dir < 0 || dir > 2 || (cur & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_fofc_repair: FP64 contexts only");
  if ((dir == 1 && !c->g.two_d) || (dir == 2 && !c->g.three_d)) return fail(CHQ_EINVAL, "direction not present in this mesh");
  FluxArgs a = c->flux_args(dir, cur, c->wcur, c->ctrl_api);
  HIPCHK(launch_repair(a, dir, false, c->stream));
  return CHQ_OK;
}

// ---------------------------------------------------------------- diffusion
namespace {
// diffusion::TimeStep's limit (diffusion.hpp:159-174) from the minimum active density
double parabolic_dt(const chq_ctx *c, double rho_min) {
  if (!c->diffusive()) return 1e300;
  const Geom &g = c->g;
  const int ndim = 1 + (g.two_d ? 1 : 0) + (g.three_d ? 1 : 0);
  double dxmin = g.dx1;
  if (g.two_d) dxmin = std::fmin(dxmin, g.dx2);
  if (g.three_d) dxmin = std::fmin(dxmin, g.dx3);
  const double chi = (c->cfg.kappa > 0) ? c->cfg.kappa * c->eos.gm1 / rho_min : 0.0;
  const double dmax = std::fmax(std::fmax(c->cfg.nu, c->cfg.eta), chi);
  return dxmin * dxmin / (2.0 * ndim * dmax);
}

// minimum active density of primitive set `wset` (synchronises)
int rho_min_now(chq_ctx *c, Ctrl *k, int wset, double *rho) {
  HIPCHK(launch_rho_min(c->f64.w[wset], c->g, k, true, c->stream));
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, &k->rho_bits, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::memcpy(rho, &bits, sizeof(bits));
  return CHQ_OK;
}

// the diffusive fluxes of all present directions (diffusion::Fluxes, main.cpp:55-59)
int enqueue_diff_fluxes(chq_ctx *c, const DiffArgs &a, bool flagged_only) {
  if (!c->diff_any(a)) return CHQ_OK;
  for (int d = 0; d < c->ndir(); ++d) HIPCHK(launch_diff_flux(a, d, flagged_only, c->stream));
  return CHQ_OK;
}

int sts_alloc(chq_ctx *c) {
  if (c->sts_buf) return CHQ_OK;
  const size_t cs = size_t(c->g.cstride);
  HIPCHK(hipMalloc(&c->sts_buf, (5 * NVAR + 15) * cs * sizeof(double)));
  // the face combination covers [is..ie+1] x [js..je+1] x [ks..ke+1] on every component, beyond
  // where the operator writes mb?: those entries stay zero for the life of the context
  HIPCHK(launch_fill_zero(c->sts_buf, (long long)((5 * NVAR + 15) * cs), c->stream));
  double *p = c->sts_buf;
  for (int n = 0; n < 3; ++n, p += NVAR * cs) c->sy[n] = p;
  c->sm0 = p; p += NVAR * cs;
  c->smj = p; p += NVAR * cs;
  for (int n = 0; n < 3; ++n)
    for (int d = 0; d < 3; ++d, p += cs) c->syb[n][d] = p;
  for (int d = 0; d < 3; ++d, p += cs) c->smb0[d] = p;
  for (int d = 0; d < 3; ++d, p += cs) c->smbj[d] = p;
  return CHQ_OK;
}

// sts::Operator (sts.hpp:43-90) on (u, b) whose primitives are in set wset: -div F_diff into
// mu, -curl(eta J) into mb
int sts_operator(chq_ctx *c, Ctrl *k, int wset, double *const b[3], double *mu, double *const mb[3]) {
  const size_t cs = size_t(c->g.cstride);
  for (int d = 0; d < 3; ++d) HIPCHK(launch_fill_zero(c->f64.flx[d], (long long)(NVAR * cs), c->stream));
  for (int d = 0; d < 3; ++d) HIPCHK(launch_fill_zero(c->f64.ee[d], (long long)cs, c->stream));
  const DiffArgs a = c->diff_args(b[0], b[1], b[2], wset, k, false);
  if (int rc = enqueue_diff_fluxes(c, a, false)) return rc;
  if (a.eta > 0.0) HIPCHK(launch_resistive_emf(a, c->stream));
  HIPCHK(launch_sts_div(c->g, c->f64.flx[0], c->f64.flx[1], c->f64.flx[2], c->f64.ee[0], c->f64.ee[1], c->f64.ee[2],
                        mu, mb[0], mb[1], mb[2], c->stream));
  return CHQ_OK;
}

int sts_c2p(chq_ctx *c, Ctrl *k, int wset, double *u, double *const b[3]) {
  C2pArgs a = c->c2p_args(0, wset, k);
  a.u = u;
  a.b1 = b[0]; a.b2 = b[1]; a.b3 = b[2];
  a.compute_dt = 0;
  a.flag_mode = 0;
  HIPCHK(launch_c2p<double>(a, c->stream));
  return CHQ_OK;
}

int sts_bc(chq_ctx *c, Ctrl *k, double *u, double *const b[3]) {
  BcArgs a{};
  a.g = c->g;
  a.nvar = NVAR;
  a.u = u;
  a.b1 = b[0]; a.b2 = b[1]; a.b3 = b[2];
  a.ctrl = k;
  a.only_if_bad = 0;
  a.bc = c->cfg.bc[0];
  HIPCHK(launch_bc(a, 0, c->stream));
  if (c->g.two_d) { a.bc = c->cfg.bc[1]; HIPCHK(launch_bc(a, 1, c->stream)); }
  if (c->g.three_d) { a.bc = c->cfg.bc[2]; HIPCHK(launch_bc(a, 2, c->stream)); }
  return CHQ_OK;
}

double sts_b(int j) { return j <= 2 ? 1.0 / 3.0 : (double(j) * j + j - 2.0) / (2.0 * j * (j + 1.0)); }

// sts::Advance (sts.hpp:111-157) over tau from (u0, b?f0), result left there; u1 / b?f1 are
// the fourth rotating set. Same stage count, coefficients and array rotation as the reference.
int sts_advance(chq_ctx *c, Ctrl *k, int wset, double tau, double dt_par, int *stages) {
  if (int rc = sts_alloc(c)) return rc;
  const size_t cs = size_t(c->g.cstride);
  const double r = tau / dt_par;
  int s = static_cast<int>(std::ceil(0.5 * (-1.0 + std::sqrt(9.0 + 16.0 * r))));
  if (s < 2) s = 2;
  const double w1 = 4.0 / (double(s) * s + s - 2.0);
  HIPCHK(hipMemcpyAsync(c->sy[0], c->u[0], NVAR * cs * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  for (int d = 0; d < 3; ++d)
    HIPCHK(hipMemcpyAsync(c->syb[0][d], c->bf[0][d], cs * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (int rc = sts_c2p(c, k, wset, c->sy[0], c->syb[0])) return rc;
  if (int rc = sts_operator(c, k, wset, c->syb[0], c->sm0, c->smb0)) return rc;
  const double mt1 = double(sts_b(1) * w1) * tau;
  {
    const StsTerm t[5] = {{c->sy[0], 1}, {c->sy[0], 0}, {c->sy[0], 0}, {c->sm0, mt1}, {c->sm0, 0}};
    HIPCHK(launch_sts_combine(c->g, c->sy[1], t, NVAR, false, c->stream));
    for (int d = 0; d < 3; ++d) {
      const StsTerm tb[5] = {{c->syb[0][d], 1}, {c->syb[0][d], 0}, {c->syb[0][d], 0}, {c->smb0[d], mt1}, {c->smb0[d], 0}};
      HIPCHK(launch_sts_combine(c->g, c->syb[1][d], tb, 1, true, c->stream));
    }
  }
  if (int rc = sts_bc(c, k, c->sy[1], c->syb[1])) return rc;
  double *yjm2 = c->sy[0], *yjm1 = c->sy[1], *yj = c->sy[2];
  double *bjm2[3], *bjm1[3], *bjn[3], *spareb[3];
  for (int d = 0; d < 3; ++d) {
    bjm2[d] = c->syb[0][d];
    bjm1[d] = c->syb[1][d];
    bjn[d] = c->syb[2][d];
    spareb[d] = c->bf[1][d];
  }
  double *spare = c->u[1];
  for (int j = 2; j <= s; ++j) {
    const double b_j = sts_b(j), b_jm1 = sts_b(j - 1), b_jm2 = sts_b(j - 2);
    const double mu = (2.0 * j - 1.0) / j * b_j / b_jm1;
    const double nu = -(j - 1.0) / j * b_j / b_jm2;
    const double mut = mu * w1;
    const double gt = -(1.0 - b_jm1) * mut;
    if (int rc = sts_c2p(c, k, wset, yjm1, bjm1)) return rc;
    if (int rc = sts_operator(c, k, wset, bjm1, c->smj, c->smbj)) return rc;
    const StsTerm t[5] = {{yjm1, mu}, {yjm2, nu}, {c->sy[0], 1.0 - mu - nu}, {c->smj, mut * tau}, {c->sm0, gt * tau}};
    HIPCHK(launch_sts_combine(c->g, yj, t, NVAR, false, c->stream));
    for (int d = 0; d < 3; ++d) {
      const StsTerm tb[5] = {{bjm1[d], mu}, {bjm2[d], nu}, {c->syb[0][d], 1.0 - mu - nu}, {c->smbj[d], mut * tau},
                             {c->smb0[d], gt * tau}};
      HIPCHK(launch_sts_combine(c->g, bjn[d], tb, 1, true, c->stream));
    }
    if (int rc = sts_bc(c, k, yj, bjn)) return rc;
    double *freed = j == 2 ? spare : yjm2;
    double *freedb[3];
    for (int d = 0; d < 3; ++d) freedb[d] = j == 2 ? spareb[d] : bjm2[d];
    yjm2 = yjm1; yjm1 = yj; yj = freed;
    for (int d = 0; d < 3; ++d) { bjm2[d] = bjm1[d]; bjm1[d] = bjn[d]; bjn[d] = freedb[d]; }
  }
  HIPCHK(hipMemcpyAsync(c->u[0], yjm1, NVAR * cs * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  for (int d = 0; d < 3; ++d)
    HIPCHK(hipMemcpyAsync(c->bf[0][d], bjm1[d], cs * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  *stages = s;
  return CHQ_OK;
}
}  // namespace

int chq_diffusion_fluxes(chq_ctx *c, int dir, int cur) {
  if (!c || dir < 0 || dir > 2 || (cur & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_diffusion_fluxes: FP64 contexts only");
  if ((dir == 1 && !c->g.two_d) || (dir == 2 && !c->g.three_d)) return fail(CHQ_EINVAL, "direction not present in this mesh");
  const DiffArgs a = c->diff_args(cur, c->wcur, c->ctrl_api, false);
  if (c->diff_any(a)) HIPCHK(launch_diff_flux(a, dir, false, c->stream));
  return CHQ_OK;
}

int chq_resistive_emf(chq_ctx *c, int cur) {
  if (!c || (cur & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_resistive_emf: FP64 contexts only");
  const DiffArgs a = c->diff_args(cur, c->wcur, c->ctrl_api, false);
  if (a.eta > 0.0) HIPCHK(launch_resistive_emf(a, c->stream));
  return CHQ_OK;
}

int chq_diffusion_dt(chq_ctx *c, double *dt) {
  if (!c || !dt) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_diffusion_dt: FP64 contexts only");
  double rho = 1e300;
  if (c->diffusive())
    if (int rc = rho_min_now(c, c->ctrl_api, c->wcur, &rho)) return rc;
  *dt = parabolic_dt(c, rho);
  return CHQ_OK;
}

int chq_limit_fluxes(chq_ctx *c, int dir, double beta_dt, long *limited) {
  if (!c || dir < 0 || dir > 2) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_limit_fluxes: FP64 contexts only");
  if ((dir == 1 && !c->g.two_d) || (dir == 2 && !c->g.three_d)) return fail(CHQ_EINVAL, "direction not present in this mesh");
  long long before = 0, after = 0;
  if (limited) {
    HIPCHK(hipMemcpyAsync(&before, &c->ctrl_api->limiter_total, sizeof(before), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  HIPCHK(launch_limit_fluxes(c->src_args(0, c->wcur, c->ctrl_api, beta_dt, false), dir, c->stream));
  if (limited) {
    HIPCHK(hipMemcpyAsync(&after, &c->ctrl_api->limiter_total, sizeof(after), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *limited = long(after - before);
  }
  return CHQ_OK;
}

int chq_apply_gravity(chq_ctx *c, int out, double beta_dt) {
  if (!c || (out & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_apply_gravity: FP64 contexts only");
  if (c->gravity()) HIPCHK(launch_gravity(c->src_args(out, c->wcur, c->ctrl_api, beta_dt, false), c->stream));
  return CHQ_OK;
}

int chq_apply_cooling(chq_ctx *c, int buf, double dt) {
  if (!c || (buf & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_apply_cooling: FP64 contexts only");
  if (c->cfg.cooling) HIPCHK(launch_cooling(c->src_args(buf, c->wcur, c->ctrl_api, dt, false), c->stream));
  return CHQ_OK;
}

// Stage(), main.cpp:46-93, reference order (flag before boundaries; fofc branch taken on the
// device when the flag kernel counted cells)
int chq_stage(chq_ctx *c, int cur, int out, double g0, double g1, double beta_dt, long *nbad) {
  if (!c || (cur & ~1) || (out & ~1)) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_state(c)) return rc_;
  if (c->mixed) return fail(CHQ_EUNSUP, "chq_stage: FP64 contexts only");
  const Ctrl *k = c->ctrl_api;
  const Recon rc = Recon(c->cfg.recon);
  const RSolver rs = RSolver(c->cfg.rsolver);
  const DiffArgs da = c->diff_args(cur, c->wcur, k, true);
  const SrcArgs sa = c->src_args(out, c->wcur, k, beta_dt, false);
  for (int d = 0; d < c->ndir(); ++d) HIPCHK(launch_flux(c->flux_args(d, cur, c->wcur, k), d, rc, rs, true, c->stream));
  if (int rc_ = enqueue_diff_fluxes(c, da, false)) return rc_;
  if (c->cfg.flux_limiter)
    for (int d = 0; d < c->ndir(); ++d) HIPCHK(launch_limit_fluxes(sa, d, c->stream));
  HIPCHK(launch_corner<double>(c->corner_args(c->wcur, k, 0), c->stream));
  if (da.eta > 0.0) HIPCHK(launch_resistive_emf(da, c->stream));
  UpdateArgs ua = c->update_args(out, g0, g1, 6, k, 0);
  ua.beta_dt = beta_dt;
  HIPCHK(launch_update<double>(ua, c->stream));
  if (c->gravity()) HIPCHK(launch_gravity(sa, c->stream));
  HIPCHK(launch_ctrl_begin_stage(c->ctrl_api, 0, c->stream));
  if (c->cfg.fofc) {
    HIPCHK(launch_flag(c->c2p_args(out, c->wcur, c->ctrl_api), c->stream));
    for (int d = 0; d < c->ndir(); ++d) HIPCHK(launch_repair(c->flux_args(d, cur, c->wcur, k), d, true, c->stream));
    if (int rc_ = enqueue_diff_fluxes(c, da, true)) return rc_;
    HIPCHK(launch_corner<double>(c->corner_args(c->wcur, k, 1), c->stream));
    if (da.eta > 0.0) {
      DiffArgs db = da;
      db.only_if_bad = 1;
      HIPCHK(launch_resistive_emf(db, c->stream));
    }
    UpdateArgs ub = c->update_args(out, g0, g1, 6, k, 1);
    ub.beta_dt = beta_dt;
    HIPCHK(launch_update<double>(ub, c->stream));
    if (c->gravity()) {
      SrcArgs sb = sa;
      sb.only_if_bad = 1;
      HIPCHK(launch_gravity(sb, c->stream));
    }
  }
  HIPCHK(c->boundaries(out, 6, k, 0));
  if (nbad) {
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, &c->ctrl_api->nbad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *nbad = n;
  }
  return CHQ_OK;
}

// ---------------------------------------------------------------- time loop
int chq_init(chq_ctx *c, double t, double tlim, int nlim) { return chq_init_resume(c, t, 0, tlim, nlim); }

int chq_init_resume(chq_ctx *c, double t, int cycle, double tlim, int nlim) {
  if (!c) return fail(CHQ_EINVAL, "null context");
  if (int rc_ = sync_state(c)) return rc_;
  HIPCHK(c->boundaries(0, 6, c->ctrl_api, 0));
  double mdt = 0.0;
  int rc = chq_cons_to_prim(c, 0, 1, &mdt);
  if (rc) return rc;
  Ctrl h{};
  h.t = t;
  h.tlim = tlim;
  h.nlim = nlim;
  h.cfl = c->cfg.cfl;
  h.min_dt = std::fmin(mdt, 1e300);   // main.cpp:201-202
  if (!c->cfg.sts && c->diffusive()) {
    double rho = 1e300;
    if (int rc2 = rho_min_now(c, c->ctrl_api, c->wcur, &rho)) return rc2;
    h.min_dt = std::fmin(h.min_dt, parabolic_dt(c, rho));
  }
  h.active = 0;
  h.cycle = cycle;
  h.cycle0 = cycle;   // time.nlim counts from here (main.cpp:241-243)
  // the epoch restarts at 0: clear the near-marks of the local fofc re-update it is compared with
  HIPCHK(launch_fill_zero(reinterpret_cast<double *>(c->near_marks()), c->g.cstride / 2, c->stream));
  h.wcur = c->wcur;
  HIPCHK(hipMemcpyAsync(c->ctrl, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CHQ_OK;
}

extern "C++" {
// One fused RK stage (dual energy off) over field set `fs` (FP64, or FP32 in mixed precision).
template <class T>
static int enqueue_fused_stage(chq_ctx *c, const FieldSet<T> &fs, const Coef &cf, int last) {
  const Recon rc = Recon(c->cfg.recon);
  const RSolver rs = RSolver(c->cfg.rsolver);
  Ctrl *k = c->ctrl;
  const int nvar = c->entropy ? 6 : 5;
  const int wc = c->wcur, wn = c->wcur ^ 1;
  for (int d = 0; d < c->ndir(); ++d) {
    FluxArgsT<T> fa = c->flux_args_t(fs, d, cf.cur, wc, k);
    fa.entropy = c->entropy;
    c->tmark(CHQ_K_FLUX, true);
    HIPCHK(launch_flux(fa, d, rc, rs, c->entropy, c->stream));
    c->tmark(CHQ_K_FLUX, false);
  }
  // flux limiter and gravity (FP64 contexts only) read the stage-input primitives, set wc
  constexpr bool fp64 = std::is_same<T, double>::value;
  const bool lim = fp64 && c->cfg.flux_limiter;
  if constexpr (fp64) {
    if (lim) {   // main.cpp:60-64
      const SrcArgs sa = c->src_args(cf.out, wc, k, cf.beta, true);
      for (int d = 0; d < c->ndir(); ++d) HIPCHK(launch_limit_fluxes(sa, d, c->stream));
    }
  }
  c->tmark(CHQ_K_CORNER, true);
  HIPCHK(launch_corner<T>(c->corner_args_t(fs, wc, k, 0), c->stream));
  c->tmark(CHQ_K_CORNER, false);
  const bool strict = fp64 && c->cfg.fofc && (lim || c->eos.dual_energy);
  // The first stage of RK2 and RK3 (u0 -> u1 with g1 = 0) re-updates only the points around the
  // flagged cells when it fires (k_mark_near / k_update_near), unless its first pass counted a
  // point where that could differ from the whole-domain re-update (FP64; the FP32 field set's
  // update does not read u1 when g1 = 0). Needs the fast fofc order and no source term after
  // the update.
  const bool local = c->cfg.fofc && !strict && !(fp64 && c->gravity()) && c->cfg.nstage <= 3 && cf.out == 1 &&
                     cf.g1 == 0.0;
  UpdateArgsT<T> ua = c->update_args_t(fs, cf.out, cf.g0, cf.g1, nvar, k, 0);
  ua.use_ctrl_dt = 1;
  ua.beta = cf.beta;
  if (local && fp64) {
    ua.zwatch = 1;
    ua.zcount = &k->zsens;
  }
  c->tmark(CHQ_K_UPDATE, true);
  HIPCHK(launch_update<T>(ua, c->stream));
  c->tmark(CHQ_K_UPDATE, false);
  if (fp64 && c->gravity()) HIPCHK(launch_gravity(c->src_args(cf.out, wc, k, cf.beta, true), c->stream));   // main.cpp:76
  // Strict order (flux limiter or dual energy on): the cells are flagged here, before the
  // boundary fill, as the reference does (fofc.hpp:17-44, main.cpp:78-79), and the first fill
  // and con2prim run only when no cell was flagged, because the fofc re-update must read the
  // tentative state as the update left it (main.cpp:78-91): the limiter can leave the two
  // copies of a periodic shared face unequal (the fill would replace one), and dual-energy
  // con2prim rewrites u[IEN] of switched cells (physics.hpp:463-468).
  if constexpr (fp64) {
    if (strict) {
      C2pArgs fl = c->c2p_args(cf.out, wc, k);
      fl.flag_mode = 2;
      HIPCHK(launch_flag(fl, c->stream));
    }
  }
  // Otherwise the order adapts on the device (ctrl_begin_stage): a stage whose counterpart in
  // the previous step flagged cells flags here, before the fill (gate 3), and skips its first
  // fill and con2prim when it flags again (gate 4); any other stage flags inside con2prim
  // (flag mode 3).
  const bool adaptive = c->cfg.fofc && !strict;
  if (adaptive) {
    C2pArgs fl = c->c2p_args(cf.out, wc, k);   // reads the conserved state and face B only
    fl.flag_mode = 2;
    fl.only_if_bad = 3;
    HIPCHK(launch_flag(fl, c->stream));
  }
  const int first = strict ? 2 : adaptive ? 4 : 0;
  c->tmark(CHQ_K_BC, true);
  HIPCHK(c->boundaries(cf.out, nvar, k, first));
  c->tmark(CHQ_K_BC, false);
  C2pArgsT<T> ca = c->c2p_args_t(fs, cf.out, wn, k);
  ca.only_if_bad = first;
  ca.write_went = c->entropy;
  // u[IENT] is read by nothing in this path unless entropy is transported: the intermediate RK
  // states never resynchronise it, and the last stage leaves it to sync_ient (the first host
  // read of the state or per-kernel entry point), which stores what this con2prim would have
  ca.entropy_out = c->entropy;
  if (last && !c->entropy) c->ient_stale = true;   // the last stage writes buffer 0
  ca.compute_dt = last && !(c->cfg.sts || c->cfg.cooling);   // split steps take dt after the splits
  ca.slot = 0;
  ca.flag_mode = adaptive ? 3 : 0;
  c->tmark(CHQ_K_C2P, true);
  HIPCHK(launch_c2p<T>(ca, c->stream));
  c->tmark(CHQ_K_C2P, false);
  if (c->cfg.fofc) {
    c->tmark(CHQ_K_REPAIR, true);
    FluxArgsT<T> fr[3];
    for (int d = 0; d < c->ndir(); ++d) {
      fr[d] = c->flux_args_t(fs, d, cf.cur, wc, k);
      fr[d].entropy = c->entropy;
    }
    HIPCHK(launch_repair_list(fr, c->ndir(), c->badlist, c->stream));
    HIPCHK(launch_corner_list<T>(c->corner_args_t(fs, wc, k, 1), c->badlist, c->stream));
    UpdateArgsT<T> ub = c->update_args_t(fs, cf.out, cf.g0, cf.g1, nvar, k, 1);
    ub.use_ctrl_dt = 1;
    ub.beta = cf.beta;
    if (local) {
      HIPCHK(launch_mark_near(c->g, k, c->badlist, c->near_marks(), c->stream));   // gate 5
      HIPCHK(launch_update_near<T>(ub, c->near_marks(), c->stream));               // gate 5
      ub.only_if_bad = 6;   // the whole domain when the first pass counted a point
    }
    HIPCHK(launch_update<T>(ub, c->stream));
    if (fp64 && c->gravity()) {   // main.cpp:89
      SrcArgs sb = c->src_args(cf.out, wc, k, cf.beta, true);
      sb.only_if_bad = 1;
      HIPCHK(launch_gravity(sb, c->stream));
    }
    HIPCHK(c->boundaries(cf.out, nvar, k, 1));
    C2pArgsT<T> cb = ca;
    cb.flag_mode = 0;
    cb.slot = 1;
    cb.only_if_bad = 1;   // (after ca's gate 2: exactly one of the two runs)
    HIPCHK(launch_c2p<T>(cb, c->stream));
    c->tmark(CHQ_K_REPAIR, false);
  }
  return CHQ_OK;
}
}  // extern "C++"

// One RK step on the device. Fast order (dual energy off): the fofc flag is evaluated inside
// con2prim of the tentative state, whose primitives go to the second (w, bcc, ecc) set, so
// the stage-input primitives the repair needs stay intact; the repair branch (list-driven
// repair and corner EMF, update, boundaries, con2prim) always launches and exits at once on
// the device when nothing was flagged. Entropy (IENT) is not transported unless dual energy is
// on: without it con2prim recomputes u[IENT] from rho and p in every stage, so its transport
// cannot change any result. With dual energy or the flux limiter on, the flag runs before the
// boundary fill (see enqueue_fused_stage); explicit diffusion keeps the reference kernel order.
static int enqueue_step(chq_ctx *c) {
  const Recon rc = Recon(c->cfg.recon);
  const RSolver rs = RSolver(c->cfg.rsolver);
  Ctrl *k = c->ctrl;
  const int nst = c->cfg.nstage;
  // explicit diffusion runs in the reference kernel order; dual energy, the flux limiter,
  // gravity and the operator-split steps (super-time-stepping, cooling: the stages carry no
  // diffusion then) also fit the fused stage
  const bool split = c->cfg.sts || c->cfg.cooling;   // main.cpp:273-290: dt taken after the splits
  const bool fused = !(c->diffusive() && !c->cfg.sts);
  const int nvar = c->entropy ? 6 : 5;
  HIPCHK(launch_ctrl_begin_step(k, c->stream));
  for (int s = 1; s <= nst; ++s) {
    const Coef cf = rk_coef(nst, s);
    const int last = s == nst;
    HIPCHK(launch_ctrl_begin_stage(k, last, c->stream, s - 1));
    if (fused) {
      const int rc2 = c->mixed ? enqueue_fused_stage<float>(c, c->f32, cf, last)
                               : enqueue_fused_stage<double>(c, c->f64, cf, last);
      if (rc2) return rc2;
      c->wcur ^= 1;
    } else {
      const int wc = c->wcur;
      const DiffArgs da = c->diff_args(cf.cur, wc, k, true);
      const SrcArgs sa = c->src_args(cf.out, wc, k, cf.beta, true);
      for (int d = 0; d < c->ndir(); ++d) {
        FluxArgs fa = c->flux_args(d, cf.cur, wc, k);
        fa.entropy = c->entropy;
        HIPCHK(launch_flux(fa, d, rc, rs, c->entropy, c->stream));
      }
      if (int rc_ = enqueue_diff_fluxes(c, da, false)) return rc_;        // main.cpp:55-59
      if (c->cfg.flux_limiter)                                             // main.cpp:60-64
        for (int d = 0; d < c->ndir(); ++d) HIPCHK(launch_limit_fluxes(sa, d, c->stream));
      HIPCHK(launch_corner<double>(c->corner_args(wc, k, 0), c->stream));
      if (da.eta > 0.0) HIPCHK(launch_resistive_emf(da, c->stream));     // main.cpp:66
      UpdateArgs ua = c->update_args(cf.out, cf.g0, cf.g1, nvar, k, 0);
      ua.use_ctrl_dt = 1;
      ua.beta = cf.beta;
      HIPCHK(launch_update<double>(ua, c->stream));
      if (c->gravity()) HIPCHK(launch_gravity(sa, c->stream));           // main.cpp:76
      if (c->cfg.fofc) {
        HIPCHK(launch_flag(c->c2p_args(cf.out, wc, k), c->stream));
        for (int d = 0; d < c->ndir(); ++d) HIPCHK(launch_repair(c->flux_args(d, cf.cur, wc, k), d, true, c->stream));
        if (int rc_ = enqueue_diff_fluxes(c, da, true)) return rc_;       // fofc.hpp:90-93
        HIPCHK(launch_corner<double>(c->corner_args(wc, k, 1), c->stream));
        if (da.eta > 0.0) {
          DiffArgs db = da;
          db.only_if_bad = 1;
          HIPCHK(launch_resistive_emf(db, c->stream));
        }
        UpdateArgs ub = c->update_args(cf.out, cf.g0, cf.g1, nvar, k, 1);
        ub.use_ctrl_dt = 1;
        ub.beta = cf.beta;
        HIPCHK(launch_update<double>(ub, c->stream));
        if (c->gravity()) {                                                // main.cpp:89
          SrcArgs sb = sa;
          sb.only_if_bad = 1;
          HIPCHK(launch_gravity(sb, c->stream));
        }
      }
      HIPCHK(c->boundaries(cf.out, nvar, k, 0));
      C2pArgs ca = c->c2p_args(cf.out, wc, k);
      ca.compute_dt = last && !split;   // main.cpp:273-274: split steps take dt after the splits
      ca.slot = 0;
      HIPCHK(launch_c2p<double>(ca, c->stream));
    }
    HIPCHK(launch_ctrl_end_stage(k, last, fused ? 1 : 0, c->stream));
    if (last && !fused && !split && c->diffusive()) {   // main.cpp:275
      HIPCHK(launch_rho_min(c->f64.w[c->wcur], c->g, k, true, c->stream));
      HIPCHK(launch_diff_dt_fold(k, c->diff_args(0, c->wcur, k, false), c->stream));
    }
  }
  if (c->cfg.sts) {   // operator-split diffusion on the end-of-step state, main.cpp:278-290
    double rho = 1e300;
    if (c->diffusive())
      if (int rc_ = rho_min_now(c, k, c->wcur, &rho)) return rc_;   // synchronises
    Ctrl h{};
    HIPCHK(hipMemcpyAsync(&h, k, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h.active) {
      int s = 0;
      if (int rc_ = sts_advance(c, k, c->wcur, h.dt, parabolic_dt(c, rho), &s)) return rc_;
      c->sts_stages += s;
      HIPCHK(c->boundaries(0, NVAR, k, 0));
    }
  }
  if (c->cfg.cooling) {   // main.cpp:282-286, with the step's dt (the kernels exit when the loop has ended)
    C2pArgs ca = c->c2p_args(0, c->wcur, k);
    ca.compute_dt = 0;
    ca.slot = 0;
    HIPCHK(launch_c2p<double>(ca, c->stream));
    HIPCHK(launch_cooling(c->src_args(0, c->wcur, k, 1.0, true), c->stream));
    HIPCHK(c->boundaries(0, NVAR, k, 0));
  }
  if (split) {            // main.cpp:287-290
    HIPCHK(launch_ctrl_begin_stage(k, 1, c->stream));
    C2pArgs ca = c->c2p_args(0, c->wcur, k);
    ca.compute_dt = 1;
    ca.slot = 0;
    HIPCHK(launch_c2p<double>(ca, c->stream));
    HIPCHK(launch_ctrl_take_min(k, 0, c->stream));
    if (!c->cfg.sts && c->diffusive()) {
      HIPCHK(launch_rho_min(c->f64.w[c->wcur], c->g, k, true, c->stream));
      HIPCHK(launch_diff_dt_fold(k, c->diff_args(0, c->wcur, k, false), c->stream));
    }
  }
  HIPCHK(launch_ctrl_end_step(k, c->stream));
  c->wcur_stale = fused;
  return CHQ_OK;
}

int chq_step(chq_ctx *c, int nsteps) {
  if (!c || nsteps < 0) return fail(CHQ_EINVAL, "bad argument");
  if (c->batch) return fail(CHQ_EINVAL, "this context is advanced by a batch");
  if (int rc_ = sync_wcur(c)) return rc_;
  for (int n = 0; n < nsteps; ++n) {
    const int rc = enqueue_step(c);
    if (rc) return rc;
  }
  return CHQ_OK;
}

int chq_step_timed(chq_ctx *c, int nsteps, double *ms) {
  if (!c || nsteps < 0 || !ms) return fail(CHQ_EINVAL, "bad argument");
  if (int rc_ = sync_wcur(c)) return rc_;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  const int rc = chq_step(c, nsteps);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float f = 0.0f;
  HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
  *ms = f;
  return CHQ_OK;
}

int chq_kernel_timing(chq_ctx *c, int kernel_class) {
  if (!c || kernel_class < -1 || (kernel_class >= CHQ_K_NCLASS && kernel_class != CHQ_K_ALL))
    return fail(CHQ_EINVAL, "bad kernel class");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->timing_class = kernel_class;
  c->tev_used = 0;
  return CHQ_OK;
}

int chq_kernel_timing_read(chq_ctx *c, double *total_ms, int *launches) {
  if (!c || !total_ms || !launches) return fail(CHQ_EINVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  int n = 0;
  for (size_t q = 0; q + 1 < c->tev_used; q += 2) {
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, c->tev[q], c->tev[q + 1]));
    tot += f;
    ++n;
  }
  *total_ms = tot;
  *launches = n;
  c->tev_used = 0;
  return CHQ_OK;
}

int chq_kernel_timing_breakdown(chq_ctx *c, double *ms, int *launches) {
  if (!c || !ms || !launches) return fail(CHQ_EINVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int q = 0; q < CHQ_K_NCLASS; ++q) { ms[q] = 0.0; launches[q] = 0; }
  for (size_t q = 0; q + 1 < c->tev_used; q += 2) {
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, c->tev[q], c->tev[q + 1]));
    const int cls = c->tcls[q / 2];
    if (cls >= 0 && cls < CHQ_K_NCLASS) { ms[cls] += f; launches[cls] += 1; }
  }
  c->tev_used = 0;
  return CHQ_OK;
}

int chq_get_status(chq_ctx *c, chq_status *st) {
  if (!c || !st) return fail(CHQ_EINVAL, "bad argument");
  if (c->batch_stream) HIPCHK(hipStreamSynchronize(c->batch_stream));
  Ctrl h{};
  HIPCHK(hipMemcpyAsync(&h, c->ctrl, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  st->t = h.t;
  st->dt = h.dt;
  st->tlim = h.tlim;
  st->min_dt = h.min_dt;
  st->cycle = h.cycle;
  st->nlim = h.nlim;
  st->active = h.active;
  st->nbad_last = h.nbad_last;
  st->fofc_total = h.fofc_total;
  st->sts_stages = c->sts_stages;
  st->limiter_total = h.limiter_total;
  return CHQ_OK;
}

int chq_sync(chq_ctx *c) {
  if (!c) return fail(CHQ_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return CHQ_OK;
}

void *chq_stream(chq_ctx *c) { return c ? (void *)c->stream : nullptr; }

}  // extern "C"
