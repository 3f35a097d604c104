// Batches (include/chasqui_amd.h, chq_batch_*): many independent meshes of one scheme advanced
// by shared launches. Every run keeps its own context (arena, control block, boundary
// conditions, CFL, gamma, floors, size and dimensionality); the batch owns the device table of
// their launch descriptions (RunArgs) and the (run, tile) work lists of each kernel class.
//
// One RK step is the same kernel sequence as a single context's fused step, each launch
// covering every run: flux x1..x3, corner EMF, update, boundaries, con2prim (+ fofc list),
// then the fofc branch over the runs that flagged cells (compacted on the device), then the
// control-block updates (per-run dt, t, cycle, primitive-set flip). No reduction spans runs,
// so every run's result is bitwise the one it gets alone.
//
// The step is captured once in a hipGraph and replayed. Lists and the run count live in device
// memory and every kernel walks its list with a grid-stride loop, so runs joining or leaving
// only rewrite list contents: the graph is re-captured only when a list outgrows its buffer or
// the autotuner changes a grid size.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace chq;

namespace {

// kernel classes with a host-built work list
enum : int { K_FLUX1 = 0, K_FLUX2, K_FLUX3, K_CORNER, K_UPDATE, K_BC1, K_BC2, K_BC3, K_C2P, K_NLIST };
// grid-size slots: the listed classes, then the fofc branch (repair, conditional re-runs)
enum : int { G_REPAIR = K_NLIST, G_COND, G_N };

const char *const kClassName[G_N] = {"flux_x1", "flux_x2", "flux_x3", "corner", "update",
                                     "bc_x1",   "bc_x2",   "bc_x3",   "c2p",    "repair", "fofc_rerun"};

struct BatchBase {
  virtual ~BatchBase() = default;
  virtual int rebuild() = 0;
  virtual int enqueue_step() = 0;
  virtual int tune(int reps, double *best_ms) = 0;
  virtual int live(std::vector<int> &out) = 0;   // per run: not finished
};

}  // namespace

struct chq_batch {
  int capacity = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<chq_ctx *> runs;
  // scheme fixed by the first run
  bool keyed = false;
  int recon = 0, rsolver = 0, nstage = 0, precision = 0;
  int grid[G_N];
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  bool graph_ok = false;
  bool dirty = false;   // membership changed since the device tables were written
  std::unique_ptr<BatchBase> impl;
  chq_batch() {
    for (int q = 0; q < G_N; ++q) grid[q] = q < K_NLIST ? 4096 : 1024;
    grid[G_REPAIR] = 256;
  }
};

namespace {

template <class T>
FieldSet<T> &fields(chq_ctx *c);
template <>
FieldSet<double> &fields<double>(chq_ctx *c) { return c->f64; }
template <>
FieldSet<float> &fields<float>(chq_ctx *c) { return c->f32; }

template <class T>
struct Batch final : BatchBase {
  chq_batch *b;
  RunArgs<T> *d_runs = nullptr;        // capacity entries
  int *d_count = nullptr;              // [0] runs, [1 + k] items of list k
  Item *d_items[K_NLIST]{};
  size_t item_cap[K_NLIST]{};
  int *d_cond = nullptr;               // CondList storage
  int *d_live = nullptr;               // per run: not finished (written at the end of every step)
  std::vector<RunArgs<T>> h_runs;
  std::vector<Item> h_items[K_NLIST];
  int h_count[1 + K_NLIST]{};

  explicit Batch(chq_batch *bb) : b(bb) {}
  ~Batch() override {
    if (d_runs) (void)hipFree(d_runs);
    if (d_count) (void)hipFree(d_count);
    for (auto *p : d_items) if (p) (void)hipFree(p);
    if (d_cond) (void)hipFree(d_cond);
    if (d_live) (void)hipFree(d_live);
  }

  int alloc() {
    HIPCHK(hipMalloc(&d_runs, sizeof(RunArgs<T>) * size_t(b->capacity)));
    HIPCHK(hipMalloc(&d_count, sizeof(int) * (1 + K_NLIST)));
    HIPCHK(hipMalloc(&d_cond, sizeof(int) * CondList::ints(b->capacity)));
    HIPCHK(hipMalloc(&d_live, sizeof(int) * size_t(b->capacity)));
    // the scan kernel writes every field the conditional kernels read; start from zero runs
    std::vector<int> z(CondList::ints(b->capacity), 0);
    HIPCHK(hipMemcpy(d_cond, z.data(), z.size() * sizeof(int), hipMemcpyHostToDevice));
    std::memset(h_count, 0, sizeof(h_count));
    HIPCHK(hipMemcpy(d_count, h_count, sizeof(h_count), hipMemcpyHostToDevice));
    return CHQ_OK;
  }

  // launch descriptions of one run (the fused step's, entropy not transported)
  RunArgs<T> describe(chq_ctx *c) {
    const FieldSet<T> &fs = fields<T>(c);
    RunArgs<T> r{};
    Ctrl *k = c->ctrl;
    for (int cur = 0; cur < 2; ++cur)
      for (int ws = 0; ws < 2; ++ws)
        for (int d = 0; d < 3; ++d) {
          FluxArgsT<T> a = c->flux_args_t(fs, d, cur, ws, k);
          a.entropy = 0;
          r.flux[cur][ws][d] = a;
        }
    for (int ws = 0; ws < 2; ++ws) r.corner[ws] = c->corner_args_t(fs, ws, k, 0);
    for (int out = 0; out < 2; ++out) {
      UpdateArgsT<T> u = c->update_args_t(fs, out, 0.0, 0.0, 5, k, 0);
      u.use_ctrl_dt = 1;
      r.update[out] = u;
    }
    for (int buf = 0; buf < 2; ++buf)
      for (int ws = 0; ws < 2; ++ws) {
        C2pArgsT<T> a = c->c2p_args_t(fs, buf, ws, k);
        a.write_went = 0;
        a.flag_mode = c->cfg.fofc ? 2 : 0;
        r.c2p[buf][ws] = a;
      }
    for (int buf = 0; buf < 2; ++buf)
      for (int d = 0; d < 3; ++d) {
        BcArgs a{};
        a.g = c->g;
        a.nvar = 5;
        a.u = c->u[buf];
        a.b1 = c->bf[buf][0]; a.b2 = c->bf[buf][1]; a.b3 = c->bf[buf][2];
        a.ctrl = k;
        a.only_if_bad = 0;
        a.bc = c->cfg.bc[d];
        r.bc[buf][d] = a;
      }
    r.ctrl = k;
    r.badlist = c->badlist;
    r.ndir = c->ndir();
    const Geom &g = c->g;
    r.cond_tiles[0] = update_vb(g).n;
    r.cond_tiles[1] = int((bc_points(g, 0) + 255) / 256);
    r.cond_tiles[2] = g.two_d ? int((bc_points(g, 1) + 255) / 256) : 0;
    r.cond_tiles[3] = g.three_d ? int((bc_points(g, 2) + 255) / 256) : 0;
    r.cond_tiles[4] = c2p_vb(g).n;
    return r;
  }

  // tiles of run r in list k
  int tiles_of(const RunArgs<T> &r, int k) const {
    const Geom &g = r.flux[0][0][0].g;
    if (k <= K_FLUX3) {
      const int d = k - K_FLUX1;
      if (d >= r.ndir) return 0;
      const FluxArgsT<T> &a = r.flux[0][0][d];
      if (d == 0) {
        int nseg, nwaves;
        fluxk::x_shape(a, fluxk::fpw_of(Recon(b->recon)), nseg, nwaves);
        return div_up(nwaves, 4);
      }
      int gx, gy, gz;
      if (d == 1) fluxk::m_shape<T, 1>(a, gx, gy, gz);
      else fluxk::m_shape<T, 2>(a, gx, gy, gz);
      return gx * gy * gz;
    }
    switch (k) {
      case K_CORNER: return corner_vb(g).n;
      case K_UPDATE: return update_vb(g).n;
      case K_BC1: return int((bc_points(g, 0) + 255) / 256);
      case K_BC2: return g.two_d ? int((bc_points(g, 1) + 255) / 256) : 0;
      case K_BC3: return g.three_d ? int((bc_points(g, 2) + 255) / 256) : 0;
      default: return c2p_vb(g).n;
    }
  }

  // rewrite the run table and the work lists after a membership change
  int rebuild() override {
    HIPCHK(hipStreamSynchronize(b->stream));   // previous uploads and steps are done
    const int n = int(b->runs.size());
    h_runs.resize(size_t(n));
    for (int r = 0; r < n; ++r) h_runs[r] = describe(b->runs[r]);
    // the y/z march segment follows the march threads of the whole batch per direction
    for (int d = 1; d < 3; ++d) {
      long long threads = 0;
      for (int r = 0; r < n; ++r) {
        const FluxArgsT<T> &a = h_runs[r].flux[0][0][d];
        if (d < h_runs[r].ndir) threads += fluxk::march_threads(a.il, a.iu, a.jl, a.ju, a.kl, a.ku, d, fluxk::SEG);
      }
      const int seg = fluxk::march_seg(threads);
      for (int r = 0; r < n; ++r)
        for (auto &row : h_runs[r].flux)
          for (auto &ws : row) ws[d].seg = seg;
    }
    for (int k = 0; k < K_NLIST; ++k) h_items[k].clear();
    for (int r = 0; r < n; ++r) {
      for (int k = 0; k < K_NLIST; ++k) {
        const int t = tiles_of(h_runs[r], k);
        for (int q = 0; q < t; ++q) h_items[k].push_back(Item{r, q});
      }
    }
    h_count[0] = n;
    for (int k = 0; k < K_NLIST; ++k) {
      h_count[1 + k] = int(h_items[k].size());
      if (h_items[k].size() > item_cap[k]) {   // grow: the captured graph holds the old pointer
        size_t cap = std::max<size_t>(h_items[k].size() + h_items[k].size() / 2, 1024);
        if (d_items[k]) HIPCHK(hipFree(d_items[k]));
        d_items[k] = nullptr;
        HIPCHK(hipMalloc(&d_items[k], cap * sizeof(Item)));
        item_cap[k] = cap;
        b->graph_ok = false;
      }
      if (!h_items[k].empty())
        HIPCHK(hipMemcpyAsync(d_items[k], h_items[k].data(), h_items[k].size() * sizeof(Item), hipMemcpyHostToDevice,
                              b->stream));
    }
    if (n) HIPCHK(hipMemcpyAsync(d_runs, h_runs.data(), sizeof(RunArgs<T>) * size_t(n), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(d_count, h_count, sizeof(h_count), hipMemcpyHostToDevice, b->stream));
    if (present_mask() != captured_present) b->graph_ok = false;   // a class appeared or emptied
    // live flags in the new table order, from the runs' control blocks (after the run count)
    if (int e = ctrl(CTRL_LIVE, 0)) return e;
    HIPCHK(hipStreamSynchronize(b->stream));
    return CHQ_OK;
  }

  // live flags; runs never stepped in this batch get theirs from their control block
  int live(std::vector<int> &out) override {
    const int n = int(b->runs.size());
    out.assign(size_t(n), 0);
    if (!n) return CHQ_OK;
    HIPCHK(hipMemcpyAsync(out.data(), d_live, sizeof(int) * size_t(n), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return CHQ_OK;
  }

  ItemList list(int k) const { return ItemList{d_items[k], d_count + 1 + k}; }
  CondList cond() const { return CondList{b->capacity, d_cond}; }
  int ctrl(int op, int arg) {
    HIPCHK(launchb_ctrl<T>(d_runs, d_count, b->capacity, op, arg, d_live, b->stream));
    return CHQ_OK;
  }

  // The launches of a stage; flux and boundary classes that no run of the batch has (the y and z
  // directions of 1D runs, z of 2D) are left out (present[], fixed when the step is enqueued or
  // captured; a change re-captures the graph). The control launches between stages are merged:
  // begin step + begin stage 1, end stage + begin next, end last stage + end step.
  int enqueue_stage(int s) {
    const Coef cf = rk_coef(b->nstage, s);
    const int last = s == b->nstage;
    const Recon rc = Recon(b->recon);
    const RSolver rs = RSolver(b->rsolver);
    hipStream_t st = b->stream;
    const int *g = b->grid;
    for (int d = 0; d < 3; ++d)
      if (present[K_FLUX1 + d]) HIPCHK(launchb_flux(d_runs, list(K_FLUX1 + d), d, cf.cur, rc, rs, g[K_FLUX1 + d], st));
    HIPCHK(launchb_corner<T>(d_runs, list(K_CORNER), g[K_CORNER], st));
    HIPCHK(launchb_update<T>(d_runs, list(K_UPDATE), cf.out, cf.g0, cf.g1, cf.beta, g[K_UPDATE], st));
    for (int d = 0; d < 3; ++d)
      if (present[K_BC1 + d]) HIPCHK(launchb_bc<T>(d_runs, list(K_BC1 + d), d, cf.out, g[K_BC1 + d], st));
    HIPCHK(launchb_c2p<T>(d_runs, list(K_C2P), cf.out, last, g[K_C2P], st));
    // fofc branch of the runs that flagged cells: for batches of small runs, one block per run
    // in two launches; otherwise compacted over the flagged runs, one launch per class
    if (small_branch) {
      HIPCHK(launchb_repair_runs(d_runs, d_count, b->capacity, cf.cur, st));
      HIPCHK(launchb_fofc_runs<T>(d_runs, d_count, b->capacity, cf.out, cf.g0, cf.g1, cf.beta, last, st));
    } else {
      HIPCHK(launchb_fofc_scan<T>(d_runs, d_count, cond(), st));
      HIPCHK(launchb_repair(d_runs, cond(), cf.cur, g[G_REPAIR], st));
      HIPCHK(launchb_corner_cond<T>(d_runs, cond(), g[G_REPAIR], st));
      HIPCHK(launchb_update_cond<T>(d_runs, cond(), cf.out, cf.g0, cf.g1, cf.beta, g[G_COND], st));
      for (int d = 0; d < 3; ++d)
        if (present[K_BC1 + d]) HIPCHK(launchb_bc_cond<T>(d_runs, cond(), d, cf.out, g[G_COND], st));
      HIPCHK(launchb_c2p_cond<T>(d_runs, cond(), cf.out, last, g[G_COND], st));
    }
    if (last) return ctrl(CTRL_END_LAST, 0);
    return ctrl(CTRL_TURN, s + 1 == b->nstage);
  }

  int enqueue_step() override {
    for (int k = 0; k < K_NLIST; ++k) present[k] = h_count[1 + k] > 0;
    small_branch = small_runs();
    captured_present = present_mask();
    if (int e = ctrl(CTRL_STEP_FIRST, b->nstage == 1)) return e;
    for (int s = 1; s <= b->nstage; ++s)
      if (int e = enqueue_stage(s)) return e;
    return CHQ_OK;
  }

  bool present[K_NLIST]{};
  bool small_branch = false;
  unsigned captured_present = ~0u;
  // the runs' fofc branch fits one block each: at most SMALL_BRANCH_TILES tiles of re-update,
  // boundaries and con2prim per run (a few thousand cells), where launches cost more than the
  // block's serial walk over its tiles
  static constexpr int SMALL_BRANCH_TILES = 48;
  bool small_runs() const {
    if (h_runs.empty()) return false;
    for (const RunArgs<T> &r : h_runs) {
      int t = 0;
      for (int q = 0; q < 5; ++q) t += r.cond_tiles[q];
      if (t > SMALL_BRANCH_TILES) return false;
    }
    return true;
  }
  unsigned present_mask() const {
    unsigned m = 0;
    for (int k = 0; k < K_NLIST; ++k) m |= (h_count[1 + k] > 0 ? 1u : 0u) << k;
    return m | (small_runs() ? 1u << K_NLIST : 0u);
  }

  // one timed launch of class q with grid gsz, in the stage-1 configuration (input buffer 0,
  // output buffer 1, the tentative primitive set): it writes only arrays a step rewrites
  // before reading (fluxes, EMFs, u1, face B 1, the tentative primitives)
  int launch_class(int q, int gsz) {
    const Recon rc = Recon(b->recon);
    const RSolver rs = RSolver(b->rsolver);
    hipStream_t st = b->stream;
    switch (q) {
      case K_FLUX1: case K_FLUX2: case K_FLUX3:
        HIPCHK(launchb_flux(d_runs, list(q), q - K_FLUX1, 0, rc, rs, gsz, st));
        break;
      case K_CORNER: HIPCHK(launchb_corner<T>(d_runs, list(q), gsz, st)); break;
      case K_UPDATE: HIPCHK(launchb_update<T>(d_runs, list(q), 1, 1.0, 0.0, 1.0, gsz, st)); break;
      case K_BC1: case K_BC2: case K_BC3: HIPCHK(launchb_bc<T>(d_runs, list(q), q - K_BC1, 1, gsz, st)); break;
      case K_C2P: HIPCHK(launchb_c2p<T>(d_runs, list(q), 1, 0, gsz, st)); break;
      default: break;
    }
    return CHQ_OK;
  }

  // Autotuner: for every listed kernel class, time each candidate grid size (median of
  // `reps` launches) and keep the fastest. Con2prim is timed with the fofc flags of the
  // tentative state disabled, so repeated launches do not grow the flag lists.
  int tune(int reps, double *best_ms) override {
    static const int cand[] = {256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072, 262144};
    if (int e = ctrl(CTRL_BEGIN_STEP, 0)) return e;
    if (int e = ctrl(CTRL_BEGIN_STAGE, 0)) return e;
    // con2prim without flags while tuning
    std::vector<RunArgs<T>> saved = h_runs;
    for (auto &r : h_runs)
      for (auto &row : r.c2p)
        for (auto &a : row) a.flag_mode = 0;
    if (!h_runs.empty())
      HIPCHK(hipMemcpyAsync(d_runs, h_runs.data(), sizeof(RunArgs<T>) * h_runs.size(), hipMemcpyHostToDevice, b->stream));
    double total = 0.0;
    for (int q = 0; q < K_NLIST; ++q) {
      if (h_count[1 + q] == 0) continue;
      double best = 1e30;
      int best_g = b->grid[q];
      for (int gsz : cand) {
        if (gsz > 4 * h_count[1 + q] && gsz > 256) break;
        if (int e = launch_class(q, gsz)) return e;   // warm
        std::vector<float> t;
        for (int rep = 0; rep < reps; ++rep) {
          HIPCHK(hipEventRecord(b->ev0, b->stream));
          if (int e = launch_class(q, gsz)) return e;
          HIPCHK(hipEventRecord(b->ev1, b->stream));
          HIPCHK(hipEventSynchronize(b->ev1));
          float ms = 0.0f;
          HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
          t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        const double med = t[t.size() / 2];
        if (med < best) { best = med; best_g = gsz; }
      }
      b->grid[q] = best_g;
      total += best;
    }
    h_runs = saved;
    if (!h_runs.empty())
      HIPCHK(hipMemcpyAsync(d_runs, h_runs.data(), sizeof(RunArgs<T>) * h_runs.size(), hipMemcpyHostToDevice, b->stream));
    if (int e = ctrl(CTRL_BEGIN_STAGE, 1)) return e;   // clear the flag counts of the tuning launches
    HIPCHK(hipStreamSynchronize(b->stream));
    b->graph_ok = false;
    if (best_ms) *best_ms = total;
    return CHQ_OK;
  }
};

int capture(chq_batch *b) {
  if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
  b->graph_ok = false;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = b->impl->enqueue_step();
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(CHQ_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { b->graph = nullptr; return fail(CHQ_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)); }
  b->graph_ok = true;
  return CHQ_OK;
}

}  // namespace

extern "C" {

int chq_batch_create(int capacity, chq_batch **out) {
  if (!out || capacity < 1 || capacity > (1 << 20)) return fail(CHQ_EINVAL, "capacity must be 1 .. 2^20");
  *out = nullptr;
  auto *b = new (std::nothrow) chq_batch();
  if (!b) return fail(CHQ_ENOMEM, "batch");
  b->capacity = capacity;
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    chq_batch_destroy(b);
    return fail(CHQ_EHIP, std::string("batch stream: ") + hipGetErrorString(e));
  }
  *out = b;
  return CHQ_OK;
}

int chq_batch_destroy(chq_batch *b) {
  if (!b) return CHQ_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (chq_ctx *c : b->runs) { c->batch = nullptr; c->batch_stream = nullptr; c->wcur_stale = true; }
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->impl.reset();
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CHQ_OK;
}

int chq_batch_add(chq_batch *b, chq_ctx *c) {
  if (!b || !c) return fail(CHQ_EINVAL, "null batch or context");
  if (c->batch) return fail(CHQ_EINVAL, "context already belongs to a batch");
  if (int(b->runs.size()) >= b->capacity) return fail(CHQ_EINVAL, "batch is full");
  if (c->eos.dual_energy) return fail(CHQ_EUNSUP, "batched runs need dual energy off");
  if (c->cfg.characteristic && c->cfg.recon != 0)
    return fail(CHQ_EUNSUP, "batched runs need primitive-variable reconstruction");
  if (c->diffusive() || c->cfg.sts) return fail(CHQ_EUNSUP, "batched runs need diffusion off");
  if (c->cfg.flux_limiter || c->gravity() || c->cfg.cooling)
    return fail(CHQ_EUNSUP, "batched runs need the flux limiter and source terms off");
  if (c->device != b->device) return fail(CHQ_EINVAL, "context and batch are on different devices");
  const int prec = c->mixed ? 1 : 0;
  if (!b->keyed) {
    b->recon = c->cfg.recon; b->rsolver = c->cfg.rsolver; b->nstage = c->cfg.nstage; b->precision = prec;
    if (prec) { auto p = std::make_unique<Batch<float>>(b); if (int e = p->alloc()) return e; b->impl = std::move(p); }
    else { auto p = std::make_unique<Batch<double>>(b); if (int e = p->alloc()) return e; b->impl = std::move(p); }
    b->keyed = true;
  } else if (c->cfg.recon != b->recon || c->cfg.rsolver != b->rsolver || c->cfg.nstage != b->nstage || prec != b->precision) {
    return fail(CHQ_EINVAL, "a batch holds runs of one scheme (reconstruction, Riemann solver, integrator, precision)");
  }
  // the run's device state must be complete and its current primitive set known on the device
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->wcur_stale) {
    int v = 0;
    HIPCHK(hipMemcpy(&v, &c->ctrl->wcur, sizeof(int), hipMemcpyDeviceToHost));
    c->wcur = v;
    c->wcur_stale = false;
  }
  HIPCHK(hipMemcpy(&c->ctrl->wcur, &c->wcur, sizeof(int), hipMemcpyHostToDevice));
  b->runs.push_back(c);
  c->batch = b;
  c->batch_stream = b->stream;
  b->dirty = true;
  return CHQ_OK;
}

int chq_batch_remove(chq_batch *b, chq_ctx *c) {
  if (!b || !c) return fail(CHQ_EINVAL, "null batch or context");
  auto it = std::find(b->runs.begin(), b->runs.end(), c);
  if (it == b->runs.end()) return fail(CHQ_EINVAL, "context is not in this batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  *it = b->runs.back();   // keep the table dense
  b->runs.pop_back();
  c->batch = nullptr;
  c->batch_stream = nullptr;
  c->wcur_stale = true;
  b->dirty = true;
  return CHQ_OK;
}

int chq_batch_size(const chq_batch *b) { return b ? int(b->runs.size()) : 0; }

namespace {
// device tables follow the membership lazily: joins and leaves between two steps cost one rewrite
int refresh(chq_batch *b) {
  if (!b->dirty || !b->impl) return CHQ_OK;
  if (int e = b->impl->rebuild()) return e;
  b->dirty = false;
  return CHQ_OK;
}
}  // namespace

int chq_batch_step(chq_batch *b, int nsteps) {
  if (!b || nsteps < 0) return fail(CHQ_EINVAL, "bad argument");
  if (int e = refresh(b)) return e;
  if (b->runs.empty() || nsteps == 0) return CHQ_OK;
  if (b->use_graph && !b->graph_ok)
    if (int e = capture(b)) return e;
  for (int n = 0; n < nsteps; ++n) {
    if (b->use_graph) HIPCHK(hipGraphLaunch(b->graph, b->stream));
    else if (int e = b->impl->enqueue_step()) return e;
  }
  for (chq_ctx *c : b->runs) {
    c->wcur_stale = true;
    c->ient_stale = true;   // batches transport no entropy; the last stage leaves u[IENT] to sync_ient
  }
  return CHQ_OK;
}

int chq_batch_step_timed(chq_batch *b, int nsteps, double *ms) {
  if (!b || !ms) return fail(CHQ_EINVAL, "bad argument");
  if (int e = refresh(b)) return e;
  if (b->use_graph && !b->graph_ok && !b->runs.empty())
    if (int e = capture(b)) return e;
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  if (int e = chq_batch_step(b, nsteps)) return e;
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  float f = 0.0f;
  HIPCHK(hipEventElapsedTime(&f, b->ev0, b->ev1));
  *ms = f;
  return CHQ_OK;
}

int chq_batch_sync(chq_batch *b) {
  if (!b) return fail(CHQ_EINVAL, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  return CHQ_OK;
}

int chq_batch_unfinished(chq_batch *b, int *count, int *flags) {
  if (!b || !count) return fail(CHQ_EINVAL, "bad argument");
  *count = 0;
  if (!b->impl) return CHQ_OK;
  if (int e = refresh(b)) return e;
  std::vector<int> live;
  if (int e = b->impl->live(live)) return e;
  int n = 0;
  for (size_t r = 0; r < live.size(); ++r) {
    if (flags) flags[r] = live[r];
    n += live[r];
  }
  *count = n;
  return CHQ_OK;
}

chq_ctx *chq_batch_run(chq_batch *b, int index) {
  if (!b || index < 0 || index >= int(b->runs.size())) return nullptr;
  return b->runs[size_t(index)];
}

int chq_batch_use_graph(chq_batch *b, int on) {
  if (!b) return fail(CHQ_EINVAL, "null batch");
  b->use_graph = on != 0;
  return CHQ_OK;
}

int chq_batch_tune(chq_batch *b, int reps, double *best_ms) {
  if (!b || reps < 1) return fail(CHQ_EINVAL, "bad argument");
  if (b->runs.empty()) return fail(CHQ_EINVAL, "empty batch");
  if (int e = refresh(b)) return e;
  return b->impl->tune(reps, best_ms);
}

int chq_batch_grid(chq_batch *b, int cls, int *grid, const char **name) {
  if (!b || cls < 0 || cls >= G_N) return fail(CHQ_EINVAL, "bad argument");
  if (grid) *grid = b->grid[cls];
  if (name) *name = kClassName[cls];
  return CHQ_OK;
}

int chq_batch_set_grid(chq_batch *b, int cls, int grid) {
  if (!b || cls < 0 || cls >= G_N || grid < 1 || grid > (1 << 22)) return fail(CHQ_EINVAL, "bad argument");
  b->grid[cls] = grid;
  b->graph_ok = false;
  return CHQ_OK;
}

}  // extern "C"
