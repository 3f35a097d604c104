// chasqui_amd — command-line driver of the MI355X stage pipeline (restates main.cpp:97-329
// of the reference over the C ABI):
//
//   chasqui_amd -i input.in [-r restart.rst] [block.key=value ...]
//
// Same deck syntax and overrides, same stdout summary (parameter echo, cycle lines every
// ndiag, zone-cycles/second, fofc count, L1 errors), same .hst / .vtk / .rst / _errors.dat
// files, restart from a .rst dump.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "chasqui_amd.h"

namespace {
int die(const std::string &msg) {
  std::cerr << "chasqui_amd: " << msg << "\n";
  return 1;
}
}  // namespace

int main(int argc, char *argv[]) {
  std::string input, overrides, rstfile;
  for (int a = 1; a < argc; ++a) {
    if (std::strcmp(argv[a], "-i") == 0 && a + 1 < argc) { input = argv[++a]; continue; }
    if (std::strcmp(argv[a], "-r") == 0 && a + 1 < argc) { rstfile = argv[++a]; continue; }
    if (std::strchr(argv[a], '=')) overrides += std::string(argv[a]) + "\n";
  }
  if (input.empty()) {
    std::cerr << "usage: chasqui_amd -i input.in [-r restart.rst] [block.key=value ...]\n";
    return 1;
  }
  std::ifstream fin(input);
  if (!fin) return die("cannot open input file: " + input);
  std::stringstream ss;
  ss << fin.rdbuf();

  chq_deck *deck = nullptr;
  if (chq_deck_open(ss.str().c_str(), overrides.c_str(), &deck)) return die(chq_deck_last_error());
  chq_mesh_desc mesh;
  chq_config cfg;
  double tlim = 1.0;
  int nlim = -1, ndiag = 100;
  chq_deck_mesh(deck, &mesh);
  chq_deck_config(deck, &cfg);
  chq_deck_run(deck, &tlim, &nlim, &ndiag);
  const std::string basename = chq_deck_get(deck, "output", "basename");
  const double dt_vtk = std::stod(chq_deck_get(deck, "output", "vtk_dt"));
  const double dt_hst = std::stod(chq_deck_get(deck, "output", "hst_dt"));
  const double dt_rst = std::stod(chq_deck_get(deck, "output", "rst_dt"));
  // GetOrAddBool semantics (params.hpp:67-72): the value is compared in lower case
  std::string ce = chq_deck_get(deck, "output", "compute_error");
  std::transform(ce.begin(), ce.end(), ce.begin(), [](unsigned char ch) { return char(std::tolower(ch)); });
  const bool compute_error = ce == "true" || ce == "1" || ce == "yes";

  std::cout << "# Chasqui MHD\n";
  std::vector<char> pbuf(chq_deck_print(deck, nullptr, 0));
  chq_deck_print(deck, pbuf.data(), pbuf.size());
  std::cout << pbuf.data();
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return die("no HIP device");
  std::cout << "# execution space: HIP " << prop.gcnArchName << " (" << prop.name << ")\n";

  // initial conditions or restart, main.cpp:191-201
  double t = 0.0;
  int cycle = 0;
  chq_ctx *ctx = nullptr;
  if (chq_create(&mesh, &cfg, &ctx)) return die(chq_last_error());
  if (rstfile.empty()) {
    if (chq_deck_upload(deck, ctx)) return die(chq_last_error());
  } else {
    if (chq_restart_read(ctx, rstfile.c_str(), &t, &cycle)) return die(chq_last_error());
    std::cout << "# restarted from " << rstfile << " at t = " << t << ", cycle = " << cycle << "\n";
  }
  if (chq_init_resume(ctx, t, cycle, tlim, nlim)) return die(chq_last_error());
  chq_deck_close(deck);

  std::vector<double> u_init, b_init;
  if (compute_error) {
    u_init.resize(chq_array_count(ctx, CHQ_U0));
    b_init.resize(chq_array_count(ctx, CHQ_BCC));
    chq_download(ctx, CHQ_U0, u_init.data(), u_init.size());
    chq_download(ctx, CHQ_BCC, b_init.data(), b_init.size());
  }

  // output scheduling, main.cpp:211-239
  int nvtk = 0, nrst = 0;
  double next_vtk = t, next_hst = t, next_rst = t + dt_rst;
  std::ofstream hst;
  if (dt_hst > 0) {
    hst.open(basename + ".hst", rstfile.empty() ? std::ios::out : std::ios::app);
    if (rstfile.empty()) hst << "# time dt mass mom1 mom2 mom3 etot ekin emag eint divB_max divB_L1\n";
  }
  auto do_outputs = [&](double dt, bool force, bool final = false) {
    const double eps = 1e-10 * (tlim > 0 ? tlim : 1.0);
    if (dt_vtk > 0 && (t >= next_vtk - eps || force)) {
      char fn[512];
      std::snprintf(fn, sizeof(fn), "%s.%04d.vtk", basename.c_str(), nvtk++);
      chq_write_vtk(ctx, &mesh, fn, t, cycle);
      next_vtk += dt_vtk;
    }
    if (dt_hst > 0 && (t >= next_hst - eps || force)) {
      double h[10];
      chq_history(ctx, &mesh, 0, h);
      hst << std::setprecision(12) << t << " " << dt;
      for (double v : h) hst << " " << v;
      hst << "\n";
      hst.flush();
      next_hst += dt_hst;
    }
    if (dt_rst > 0 && (t >= next_rst - eps || final)) {
      char fn[512];
      std::snprintf(fn, sizeof(fn), "%s.%04d.rst", basename.c_str(), nrst++);
      if (chq_restart_write(ctx, fn, t, cycle)) std::cerr << "chasqui_amd: " << chq_last_error() << "\n";
      next_rst += dt_rst;
    }
  };
  if (rstfile.empty()) do_outputs(0.0, true);

  // time loop, main.cpp:241-305. Without per-step outputs or diagnostics the steps are
  // enqueued in chunks and the device decides when t reaches tlim.
  const int cycle0 = cycle;
  const bool per_step = dt_vtk > 0 || dt_hst > 0 || dt_rst > 0 || ndiag > 0;
  chq_sync(ctx);
  const auto t0 = std::chrono::steady_clock::now();
  chq_status st{};
  for (;;) {
    const int chunk = per_step ? 1 : 64;
    if (chq_step(ctx, chunk) || chq_get_status(ctx, &st)) return die(chq_last_error());
    const int done = st.cycle - cycle;
    cycle = st.cycle;
    t = st.t;
    if (per_step && done > 0) {
      if (ndiag > 0 && cycle % ndiag == 0)
        std::cout << "cycle=" << cycle << " time=" << std::setprecision(8) << t << " dt=" << st.dt << "\n";
      do_outputs(st.dt, false);
    }
    if (done < chunk) break;
  }
  chq_sync(ctx);
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (dt_vtk > 0 || dt_hst > 0 || dt_rst > 0) do_outputs(0.0, true, true);

  const double zones = double(mesh.nx1) * mesh.nx2 * mesh.nx3;
  std::cout << "# cycles = " << cycle - cycle0 << ", time = " << t << ", wall = " << wall << " s\n";
  std::cout << "# zone-cycles/second = " << std::scientific << zones * (cycle - cycle0) / wall << "\n";
  if (cfg.fofc) std::cout << "# first-order flux corrections applied: " << st.fofc_total << "\n";
  if (cfg.sts) std::cout << "# super-time-stepping stages (total): " << st.sts_stages << "\n";
  if (cfg.flux_limiter) std::cout << "# positivity flux limiter activations: " << st.limiter_total << "\n";

  if (compute_error) {   // main.cpp:310-327
    std::vector<double> u(u_init.size()), b(b_init.size());
    chq_download(ctx, CHQ_U0, u.data(), u.size());
    chq_download(ctx, CHQ_BCC, b.data(), b.size());
    int iv[16];
    chq_geometry(ctx, iv, nullptr);
    const int ni = iv[4], nj = iv[5], nk = iv[6], is = iv[7], ie = iv[8], js = iv[9], je = iv[10], ks = iv[11], ke = iv[12];
    const size_t nc = size_t(ni) * nj * nk;
    double err[9] = {0}, tot = 0;
    for (int k = ks; k <= ke; ++k)
      for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie; ++i) {
          const size_t c = (size_t(k) * nj + j) * ni + i;
          for (int n = 0; n < 6; ++n) err[n] += std::fabs(u[n * nc + c] - u_init[n * nc + c]);
          for (int n = 0; n < 3; ++n) err[6 + n] += std::fabs(b[n * nc + c] - b_init[n * nc + c]);
        }
    std::cout << "# L1 errors vs initial state (rho m1 m2 m3 E rho*s B1 B2 B3 | total):\n# ";
    for (int n = 0; n < 9; ++n) { err[n] /= zones; tot += err[n]; std::cout << err[n] << " "; }
    std::cout << "| " << tot << "\n";
    std::ofstream ef(basename + "_errors.dat", std::ios::app);
    ef << mesh.nx1 << " " << mesh.nx2 << " " << mesh.nx3 << " " << std::scientific << std::setprecision(8);
    for (int n = 0; n < 9; ++n) ef << err[n] << " ";
    ef << tot << "\n";
  }
  chq_destroy(ctx);
  return 0;
}
