// Conversion between the reference layout (dense LayoutRight [n][k][j][i], mesh.hpp:37-63)
// and the padded device layout of Geom::at(). Used by chq_upload / chq_download only.
#include "kernels.hpp"

namespace chq {

namespace {

// H: element type of the reference-layout (host-side) buffer, D: of the device array.
// float device arrays (mixed precision) are widened on download and rounded on upload.
template <class H, class D>
__global__ void __launch_bounds__(256) k_relayout(H *ref, D *dev, Geom g, int nc, int nk, int nj, int ni, int to_dev) {
  const long long n_all = (long long)nc * nk * nj * ni;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n_all;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = int(t % ni);
    long long r = t / ni;
    const int j = int(r % nj);
    r /= nj;
    const int k = int(r % nk);
    const int n = int(r / nk);
    const long long d = (long long)n * g.cstride + g.at(k, j, i);
    if (to_dev) dev[d] = D(ref[t]);
    else ref[t] = H(dev[d]);
  }
}

}  // namespace

hipError_t launch_relayout(void *ref, void *dev, const Geom &g, int nc, int nk, int nj, int ni, int to_dev,
                           int elem_bytes, hipStream_t s) {
  const long long n_all = (long long)nc * nk * nj * ni;
  const int nb = int(std::min<long long>((n_all + 255) / 256, 65536));
  if (elem_bytes == 8)
    hipLaunchKernelGGL((k_relayout<double, double>), dim3(nb), dim3(256), 0, s, (double *)ref, (double *)dev, g, nc, nk, nj, ni, to_dev);
  else if (elem_bytes == -4)   // double on the host side, float on the device
    hipLaunchKernelGGL((k_relayout<double, float>), dim3(nb), dim3(256), 0, s, (double *)ref, (float *)dev, g, nc, nk, nj, ni, to_dev);
  else
    hipLaunchKernelGGL((k_relayout<int, int>), dim3(nb), dim3(256), 0, s, (int *)ref, (int *)dev, g, nc, nk, nj, ni, to_dev);
  return hipGetLastError();
}

}  // namespace chq
