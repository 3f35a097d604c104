// FP64 face-flux kernels (the drop-in path; strict IEEE arithmetic, -ffp-contract=off) and the
// flag-array form of the fofc repair. The kernel bodies are in flux_impl.hpp.
#include "flux_impl.hpp"

namespace chq {

namespace {
using fluxk::BX;
using fluxk::BY;

// RepairFluxes<DIR> over the whole face range with the flag array (fofc.hpp:66-94).
template <int DIR>
__global__ void __launch_bounds__(BX * BY) k_repair(const FluxArgs a, const int only_if_bad) {
  if (!a.ctrl->active) return;
  if (only_if_bad && a.ctrl->nbad == 0) return;
  const int i = a.il + int(blockIdx.x) * BX + int(threadIdx.x);
  const int j = a.jl + int(blockIdx.y) * BY + int(threadIdx.y);
  const int k = a.kl + int(blockIdx.z);
  if (i > a.iu || j > a.ju) return;
  const Geom &g = a.g;
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  const int kL = k - dk, jL = j - dj, iL = i - di;
  auto inside = [&](int kk, int jj, int ii) {
    return ii >= g.is && ii <= g.ie && jj >= g.js && jj <= g.je && kk >= g.ks && kk <= g.ke;
  };
  const long long c = g.at(k, j, i), cL = g.at(kL, jL, iL);
  const bool fl = inside(k, j, i) && a.flag[c];
  const bool fr = inside(kL, jL, iL) && a.flag[cL];
  if (!(fl || fr)) return;
  fluxk::repair_face<double, DIR>(a, c, c - cL);
}
}  // namespace

hipError_t launch_flux(const FluxArgsT<double> &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s) {
  if (a.charac) return launch_flux_charac(a, dir, rc, rs, entropy, s);
  return fluxk::launch_flux_t<double>(a, dir, rc, rs, entropy, s);
}

hipError_t launch_repair_list(const FluxArgsT<double> a[3], int ndir, const int *list, hipStream_t s) {
  return fluxk::launch_repair_list_t<double>(a, ndir, list, s);
}

hipError_t launch_repair(const FluxArgs &a, int dir, bool only_if_bad, hipStream_t s) {
  const int nx = a.iu - a.il + 1, ny = a.ju - a.jl + 1, nz = a.ku - a.kl + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
  dim3 grid(div_up(nx, BX), div_up(ny, BY), nz), block(BX, BY, 1);
  const int oib = only_if_bad ? 1 : 0;
  if (dir == 0) hipLaunchKernelGGL(k_repair<0>, grid, block, 0, s, a, oib);
  else if (dir == 1) hipLaunchKernelGGL(k_repair<1>, grid, block, 0, s, a, oib);
  else hipLaunchKernelGGL(k_repair<2>, grid, block, 0, s, a, oib);
  return hipGetLastError();
}

hipError_t launchb_flux(const RunArgs<double> *R, ItemList L, int dir, int cur, Recon rc, RSolver rs, int grid, hipStream_t s) {
  return fluxk::launchb_flux_t<double>(R, L, dir, cur, rc, rs, grid, s);
}

hipError_t launchb_repair(const RunArgs<double> *R, CondList C, int cur, int grid, hipStream_t s) {
  return fluxk::launchb_repair_t<double>(R, C, cur, grid, s);
}
hipError_t launchb_repair_runs(const RunArgs<double> *R, const int *nruns, int cap, int cur, hipStream_t s) {
  return fluxk::launchb_repair_runs_t<double>(R, nruns, cap, cur, s);
}

}  // namespace chq
