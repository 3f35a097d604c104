// FP32 face-flux kernels of the mixed-precision path: reconstruction, flattening, Riemann
// solve and face EMFs in FP32 registers, from FP32 primitives (written by the FP64
// con2prim) to FP32 fluxes and EMFs (widened to FP64 by the update). This file is compiled
// with fast FP32 division/square root, FMA contraction, flushed denormals, finite-math and
// no-signed-zero semantics and reciprocal division; the kernel bodies are the same
// templates as the FP64 path (flux_impl.hpp) with the reference's single-precision literals
// (recon.hpp:145-149 WENO epsilon; Real(1e-300) -> 0).
#include "flux_impl.hpp"

namespace chq {

// The mixed path never transports the entropy variable (it requires dual energy off, where
// con2prim rebuilds u[IENT] from rho and p), so only the ENT = false kernels are built.
hipError_t launch_flux(const FluxArgsT<float> &a, int dir, Recon rc, RSolver rs, bool entropy, hipStream_t s) {
  if (entropy || a.charac) return hipErrorInvalidValue;
  if (dir == 0) return fluxk::go_flux_rc<float, 0, false>(a, rc, rs, s);
  if (dir == 1) return fluxk::go_flux_rc<float, 1, false>(a, rc, rs, s);
  return fluxk::go_flux_rc<float, 2, false>(a, rc, rs, s);
}

hipError_t launch_repair_list(const FluxArgsT<float> a[3], int ndir, const int *list, hipStream_t s) {
  return fluxk::launch_repair_list_t<float>(a, ndir, list, s);
}

hipError_t launchb_flux(const RunArgs<float> *R, ItemList L, int dir, int cur, Recon rc, RSolver rs, int grid, hipStream_t s) {
  return fluxk::launchb_flux_t<float>(R, L, dir, cur, rc, rs, grid, s);
}

hipError_t launchb_repair(const RunArgs<float> *R, CondList C, int cur, int grid, hipStream_t s) {
  return fluxk::launchb_repair_t<float>(R, C, cur, grid, s);
}
hipError_t launchb_repair_runs(const RunArgs<float> *R, const int *nruns, int cap, int cur, hipStream_t s) {
  return fluxk::launchb_repair_runs_t<float>(R, nruns, cap, cur, s);
}

}  // namespace chq
