// History sums of the .hst file reduced on the device (output::History, output.hpp:91-119):
// mass, mom1..3, etot, ekin, emag, eint as volume integrals over the active cells, and the
// maximum and mean of |div B|, from u and face B of one buffer set and the current primitives.
//
// Two kernels on the context's stream, in a fixed summation order (the result does not depend
// on scheduling, so it is the same on every call and every device): each block sums a slab of
// HST_ROWS active (k, j) rows, every thread walking i with stride 256 and accumulating in the
// reference's expression order (output.hpp:101-112), then wave64 xor-shuffles and one pass over
// the block's four waves; a one-block kernel folds the per-block partials the same way. Only
// the ten results (80 bytes) cross PCIe; the host applies the cell volume and the cell count.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.hpp"

namespace chq {

// The partials and results of chq_history_noalloc: a file-scope device array, one copy per GPU
// (320 KiB, held until the process ends), shared by every call on that GPU. Host code reaches
// it only through the address hipGetSymbolAddress returns on each device (history_scratch);
// its name in host code is a host-side stand-in, never a device address.
static __device__ __align__(256) double scratch[(4096 + 1) * HST_NV];

namespace {

constexpr int HB = 256;        // threads per block
constexpr int HST_ROWS = 32;   // active (k, j) rows per block of the first pass
// chq_history_noalloc's limit: 4096 first-pass blocks, i.e. 131072 active (k, j) rows (a 256^3
// mesh has 65536, a 1024^2 one 1024)
constexpr int HST_MAX_BLOCKS = 4096;
static_assert(sizeof(scratch) == (HST_MAX_BLOCKS + 1) * HST_NV * sizeof(double), "scratch size");

std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};   // device address of `scratch`, per device ordinal

template <class T>
struct HstArgs {
  Geom g;
  const double *u, *b1, *b2, *b3;
  const T *w, *bcc;
  int nrows;      // active (k, j) rows
  int rows_j;     // active rows per k-plane
  double *part;   // nblocks x HST_NV partials
};

// block-wide reduction of v[HST_NV] (sums, except the maximum in slot HST_MAX) in a fixed order;
// the result is valid in thread 0
__device__ __forceinline__ void block_fold(double v[HST_NV], double *lds) {
  for (int q = 0; q < HST_NV; ++q)
    for (int o = 32; o > 0; o >>= 1) {
      const double x = __shfl_xor(v[q], o, 64);
      v[q] = q == HST_MAX ? fmax(v[q], x) : v[q] + x;
    }
  const int lane = int(threadIdx.x) & 63, wid = int(threadIdx.x) >> 6;
  if (lane == 0)
    for (int q = 0; q < HST_NV; ++q) lds[wid * HST_NV + q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < HST_NV; ++q) {
      const double a = lds[q], b = lds[HST_NV + q], c = lds[2 * HST_NV + q], d = lds[3 * HST_NV + q];
      v[q] = q == HST_MAX ? fmax(fmax(a, b), fmax(c, d)) : (a + b) + (c + d);
    }
}

template <class T>
__global__ void __launch_bounds__(HB) k_history(const HstArgs<T> a) {
  __shared__ double lds[4 * HST_NV];
  const Geom &g = a.g;
  const long long cs = g.cstride, sj = g.pitch, sk = g.plane;
  double s[HST_NV];
  for (int q = 0; q < HST_NV; ++q) s[q] = 0.0;
  const int r0 = int(blockIdx.x) * HST_ROWS;
  const int r1 = min(r0 + HST_ROWS, a.nrows);
  for (int r = r0; r < r1; ++r) {
    const int k = g.ks + r / a.rows_j, j = g.js + r % a.rows_j;
    for (int i = g.is + int(threadIdx.x); i <= g.ie; i += HB) {
      const long long c = g.at(k, j, i);
      const double *u = a.u + c;
      s[0] += u[IDN * cs];
      s[1] += u[IM1 * cs];
      s[2] += u[IM2 * cs];
      s[3] += u[IM3 * cs];
      s[4] += u[IEN * cs];
      const double d = double(a.w[c]);
      const double vx = double(a.w[c + IVX * cs]), vy = double(a.w[c + IVY * cs]), vz = double(a.w[c + IVZ * cs]);
      const double bx = double(a.bcc[c]), by = double(a.bcc[c + cs]), bz = double(a.bcc[c + 2 * cs]);
      const double v2 = vx * vx + vy * vy + vz * vz;
      const double bb = bx * bx + by * by + bz * bz;
      s[5] += 0.5 * d * v2;
      s[6] += 0.5 * bb;
      s[7] += u[IEN * cs] - 0.5 * d * v2 - 0.5 * bb;
      double db = (a.b1[c + 1] - a.b1[c]) * g.idx1;
      if (g.two_d) db += (a.b2[c + sj] - a.b2[c]) * g.idx2;
      if (g.three_d) db += (a.b3[c + sk] - a.b3[c]) * g.idx3;
      s[HST_MAX] = fmax(s[HST_MAX], fabs(db));
      s[9] += fabs(db);
    }
  }
  block_fold(s, lds);
  if (threadIdx.x == 0)
    for (int q = 0; q < HST_NV; ++q) a.part[size_t(blockIdx.x) * HST_NV + q] = s[q];
}

// one block: thread t folds partials t, t + 256, ... in order, then the block fold
__global__ void __launch_bounds__(HB) k_history_final(const double *part, int nblocks, double *res) {
  __shared__ double lds[4 * HST_NV];
  double s[HST_NV];
  for (int q = 0; q < HST_NV; ++q) s[q] = 0.0;
  for (int b = int(threadIdx.x); b < nblocks; b += HB)
    for (int q = 0; q < HST_NV; ++q) {
      const double x = part[size_t(b) * HST_NV + q];
      s[q] = q == HST_MAX ? fmax(s[q], x) : s[q] + x;
    }
  block_fold(s, lds);
  if (threadIdx.x == 0)
    for (int q = 0; q < HST_NV; ++q) res[q] = s[q];
}

}  // namespace

// Device address of the static scratch on device `dev`, looked up once per device with that
// device current (chq_create calls it).
int history_scratch(int dev, double **ptr) {
  *ptr = nullptr;
  if (dev < 0 || dev >= 64) return fail(CHQ_EINVAL, "device ordinal out of range for the history scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void *q = nullptr;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess)
      return fail(CHQ_EHIP, std::string("hipGetSymbolAddress(history scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = q;
  }
  *ptr = static_cast<double *>(g_scratch_ptr[dev]);
  return CHQ_OK;
}

int history_blocks(const Geom &g) {
  const int rows = (g.two_d ? g.nx2 : 1) * (g.three_d ? g.nx3 : 1);
  return (rows + HST_ROWS - 1) / HST_ROWS;
}

// The two passes over set `buf` of context c with the primitives of set c->wcur; `work` holds
// history_blocks(g) partials followed by the HST_NV results; out[] as chq_history.
int history_on(chq_ctx *c, int buf, double *work, double out[HST_NV]) {
  const Geom &g = c->g;
  const int nb = history_blocks(g);
  double *res = work + size_t(nb) * HST_NV;
  auto launch = [&](auto w, auto bcc) -> hipError_t {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(w)>>;
    HstArgs<T> a{};
    a.g = g;
    a.u = c->u[buf];
    a.b1 = c->bf[buf][0]; a.b2 = c->bf[buf][1]; a.b3 = c->bf[buf][2];
    a.w = w;
    a.bcc = bcc;
    a.rows_j = g.two_d ? g.nx2 : 1;
    a.nrows = a.rows_j * (g.three_d ? g.nx3 : 1);
    a.part = work;
    k_history<T><<<nb, HB, 0, c->stream>>>(a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_history_final<<<1, HB, 0, c->stream>>>(work, nb, res);
    return hipGetLastError();
  };
  if (c->mixed) HIPCHK(launch(c->f32.w[c->wcur], c->f32.bcc[c->wcur]));
  else HIPCHK(launch(c->f64.w[c->wcur], c->f64.bcc[c->wcur]));
  double r[HST_NV];
  HIPCHK(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // output.hpp:115-118
  const double vol = g.dx1 * (g.two_d ? g.dx2 : 1.0) * (g.three_d ? g.dx3 : 1.0);
  for (int q = 0; q < 8; ++q) out[q] = r[q] * vol;
  out[HST_MAX] = r[HST_MAX];
  out[9] = r[9] / (double(g.nx1) * g.nx2 * g.nx3);
  return CHQ_OK;
}

}  // namespace chq

extern "C" {

// output::History (output.hpp:91-119) reduced on the device. The partials live in a buffer the
// library allocates at the first call (and grows for a larger mesh); `mesh` is the mesh
// description the context was created with (kept for the reference's signature).
int chq_history(chq_ctx *c, const chq_mesh_desc *mesh, int buf, double out[10]) {
  if (!c || !mesh || !out || (buf & ~1)) return chq::fail(CHQ_EINVAL, "chq_history: null argument or bad buffer index");
  if (int rc = chq::ctx_ready(c)) return rc;
  const size_t need = (size_t(chq::history_blocks(c->g)) + 1) * chq::HST_NV * sizeof(double);
  if (need > c->hst_bytes) {
    if (c->hst_work) HIPCHK(hipFree(c->hst_work));
    c->hst_work = nullptr;
    c->hst_bytes = 0;
    HIPCHK(hipMalloc(&c->hst_work, need));
    c->hst_bytes = need;
  }
  return chq::history_on(c, buf, c->hst_work, out);
}

// chq_history without any allocation: the partials go to the static per-device scratch.
int chq_history_noalloc(chq_ctx *c, const chq_mesh_desc *mesh, int buf, double out[10]) {
  if (!c || !mesh || !out || (buf & ~1))
    return chq::fail(CHQ_EINVAL, "chq_history_noalloc: null argument or bad buffer index");
  const int nb = chq::history_blocks(c->g);
  if (nb > chq::HST_MAX_BLOCKS)
    return chq::fail(CHQ_EINVAL, "chq_history_noalloc: the mesh has " + std::to_string(nb * chq::HST_ROWS) +
                                     " or more active (k, j) rows; the static scratch holds " +
                                     std::to_string(chq::HST_MAX_BLOCKS * chq::HST_ROWS) + " (use chq_history)");
  if (!c->hst_scratch) return chq::fail(CHQ_EHIP, "chq_history_noalloc: no history scratch on this device");
  if (int rc = chq::ctx_ready(c)) return rc;
  return chq::history_on(c, buf, c->hst_scratch, out);
}

}  // extern "C"
