// chq_stream_probe: HBM streaming kernels with the stage kernels' access pattern (one double
// per lane, a wave64 touching 512 contiguous bytes, blocks of 256 threads, one element per
// thread), timed with HIP events. Two uses:
//  * the practical streaming rate of the device (GB/s of copy, read, write and triad over
//    2 GiB arrays), the yardstick for the memory-bound stage kernels;
//  * calibration of the profiler's memory counters: run under rocprofv3 --pmc, each kernel's
//    per-dispatch reading is compared with its known byte count (below).
//
// Output: one JSON line {"bytes_per_array", "reps", kernel: {"ms", "read_bytes",
// "write_bytes", "gb_per_s"}}.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_stream_copy(const double *__restrict__ src, double *__restrict__ dst) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  dst[i] = src[i];
}

__global__ void __launch_bounds__(kBlock) k_stream_write(double *__restrict__ dst, double v) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  dst[i] = v;
}

// read-only: each block folds its 256 values into one double written to `part` (n / 256
// doubles, 1/256 of the read volume)
__global__ void __launch_bounds__(kBlock) k_stream_read(const double *__restrict__ src, double *__restrict__ part) {
  __shared__ double s[kBlock / 64];
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  double v = src[i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

__global__ void __launch_bounds__(kBlock) k_stream_triad(const double *__restrict__ a, const double *__restrict__ b,
                                                           double *__restrict__ dst, double q) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  dst[i] = a[i] + q * b[i];
}

bool ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "chq_stream_probe: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main(int argc, char **argv) {
  // array length: 2^28 doubles (2 GiB) by default, a multiple of the block size
  size_t n = size_t(1) << 28;
  int reps = 10;
  if (argc > 1) n = size_t(std::strtoull(argv[1], nullptr, 10));
  if (argc > 2) reps = std::atoi(argv[2]);
  n = (n / kBlock) * kBlock;
  if (n == 0 || reps < 1) {
    std::fprintf(stderr, "usage: chq_stream_probe [doubles per array (multiple of 256)] [reps]\n");
    return 2;
  }
  const unsigned blocks = unsigned(n / kBlock);
  double *a = nullptr, *b = nullptr, *c = nullptr, *part = nullptr;
  if (!ok(hipMalloc(&a, n * sizeof(double)), "hipMalloc") || !ok(hipMalloc(&b, n * sizeof(double)), "hipMalloc") ||
      !ok(hipMalloc(&c, n * sizeof(double)), "hipMalloc") || !ok(hipMalloc(&part, blocks * sizeof(double)), "hipMalloc"))
    return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return 1;
  // initialise every element of the three arrays (write kernel), then warm each kernel once
  k_stream_write<<<blocks, kBlock>>>(a, 1.0);
  k_stream_write<<<blocks, kBlock>>>(b, 2.0);
  k_stream_write<<<blocks, kBlock>>>(c, 0.0);
  if (!ok(hipGetLastError(), "launch") || !ok(hipDeviceSynchronize(), "init")) return 1;
  const double bytes = double(n) * sizeof(double);
  std::string out = "{\"bytes_per_array\": " + std::to_string(size_t(bytes)) + ", \"reps\": " + std::to_string(reps);
  auto run = [&](const char *name, double rd, double wr, auto launch) -> bool {
    launch();
    if (!ok(hipGetLastError(), name) || !ok(hipDeviceSynchronize(), name)) return false;
    if (!ok(hipEventRecord(e0, nullptr), "record")) return false;
    for (int r = 0; r < reps; ++r) launch();
    if (!ok(hipEventRecord(e1, nullptr), "record") || !ok(hipEventSynchronize(e1), "sync")) return false;
    float ms = 0.0f;
    if (!ok(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return false;
    const double per = double(ms) / reps;
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  ", \"%s\": {\"ms\": %.4f, \"read_bytes\": %.0f, \"write_bytes\": %.0f, \"gb_per_s\": %.1f}", name,
                  per, rd, wr, (rd + wr) / (per * 1e-3) / 1e9);
    out += buf;
    return true;
  };
  if (!run("copy", bytes, bytes, [&] { k_stream_copy<<<blocks, kBlock>>>(a, c); })) return 1;
  if (!run("read", bytes, bytes / kBlock, [&] { k_stream_read<<<blocks, kBlock>>>(a, part); })) return 1;
  if (!run("write", 0.0, bytes, [&] { k_stream_write<<<blocks, kBlock>>>(c, 3.0); })) return 1;
  if (!run("triad", 2.0 * bytes, bytes, [&] { k_stream_triad<<<blocks, kBlock>>>(a, b, c, 0.5); })) return 1;
  std::printf("%s}\n", out.c_str());
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(c);
  (void)hipFree(part);
  return 0;
}
