// Outputs: legacy-VTK dumps in the PLUTO layout (output.hpp:14-85) and the restart files
// (restart.hpp), from state downloaded off the device. The history sums of the .hst file are
// reduced on the device (history.hip).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/chasqui_amd.h"

namespace chq {
int fail(int code, const std::string &msg);   // chq_last_error() message (capi.cpp)
}

namespace {

void put_be_floats(std::ofstream &f, const std::vector<float> &v) {
  std::vector<unsigned char> b(v.size() * 4);
  for (size_t n = 0; n < v.size(); ++n) {
    uint32_t u;
    std::memcpy(&u, &v[n], 4);
    b[4 * n] = (u >> 24) & 0xff;
    b[4 * n + 1] = (u >> 16) & 0xff;
    b[4 * n + 2] = (u >> 8) & 0xff;
    b[4 * n + 3] = u & 0xff;
  }
  f.write(reinterpret_cast<const char *>(b.data()), std::streamsize(b.size()));
}

void put_be_int(std::ofstream &f, int32_t v) {
  const unsigned char b[4] = {static_cast<unsigned char>((v >> 24) & 0xff), static_cast<unsigned char>((v >> 16) & 0xff),
                              static_cast<unsigned char>((v >> 8) & 0xff), static_cast<unsigned char>(v & 0xff)};
  f.write(reinterpret_cast<const char *>(b), 4);
}

struct Geo {
  int nx1, nx2, nx3, ni, nj, nk, is, js, ks, two_d, three_d;
};

Geo geo_of(chq_ctx *ctx) {
  int iv[16];
  chq_geometry(ctx, iv, nullptr);
  return Geo{iv[0], iv[1], iv[2], iv[4], iv[5], iv[6], iv[7], iv[9], iv[11], iv[13], iv[14]};
}

// the restart header of restart.hpp:11-17, field for field (a file written by either code is
// read by the other)
struct RstHeader {
  int32_t magic;
  int32_t nx1, nx2, nx3, ng;
  int32_t cycle;
  double time;
  int32_t real_size;
};
constexpr int32_t kRstMagic = 0x43485351;   // "CHSQ"
constexpr int kRstArrays[4] = {CHQ_U0, CHQ_B1F0, CHQ_B2F0, CHQ_B3F0};

}  // namespace

extern "C" {

// WriteVTK (output.hpp:32-85): BINARY RECTILINEAR_GRID, big-endian float32 cell data
// rho prs vx1 vx2 vx3 Bx1 Bx2 Bx3 over the active cells, i fastest.
int chq_write_vtk(chq_ctx *ctx, const chq_mesh_desc *mesh, const char *path, double time, int cycle) {
  if (!ctx || !mesh || !path) return chq::fail(CHQ_EINVAL, "chq_write_vtk: null context, mesh or path");
  const Geo g = geo_of(ctx);
  std::vector<double> w(chq_array_count(ctx, CHQ_W)), b(chq_array_count(ctx, CHQ_BCC));
  int rc = chq_download(ctx, CHQ_W, w.data(), w.size());
  if (!rc) rc = chq_download(ctx, CHQ_BCC, b.data(), b.size());
  if (rc) return rc;
  std::ofstream f(path, std::ios::binary);
  if (!f) return chq::fail(CHQ_EINVAL, std::string("cannot write ") + path);
  const double dx1 = (mesh->x1max - mesh->x1min) / g.nx1, dx2 = (mesh->x2max - mesh->x2min) / g.nx2,
               dx3 = (mesh->x3max - mesh->x3min) / g.nx3;
  f << "# vtk DataFile Version 2.0\n" << "Chasqui MHD VTK Data\n" << "BINARY\n" << "DATASET RECTILINEAR_GRID\n";
  f << "FIELD FieldData 2\n" << "TIME 1 1 float\n";
  put_be_floats(f, {static_cast<float>(time)});
  f << "\nCYCLE 1 1 int\n";
  put_be_int(f, cycle);
  f << "\n";
  const int n2 = g.two_d ? g.nx2 + 1 : 1, n3 = g.three_d ? g.nx3 + 1 : 1;
  f << "DIMENSIONS " << g.nx1 + 1 << " " << n2 << " " << n3 << "\n";
  std::vector<float> x(g.nx1 + 1), y(n2), z(n3);
  for (int i = 0; i <= g.nx1; ++i) x[i] = static_cast<float>(mesh->x1min + i * dx1);
  for (int j = 0; j < n2; ++j) y[j] = static_cast<float>(g.two_d ? mesh->x2min + j * dx2 : mesh->x2min);
  for (int k = 0; k < n3; ++k) z[k] = static_cast<float>(g.three_d ? mesh->x3min + k * dx3 : mesh->x3min);
  f << "X_COORDINATES " << g.nx1 + 1 << " float\n";
  put_be_floats(f, x);
  f << "\nY_COORDINATES " << n2 << " float\n";
  put_be_floats(f, y);
  f << "\nZ_COORDINATES " << n3 << " float\n";
  put_be_floats(f, z);
  f << "\n";
  const size_t ncells = size_t(g.nx1) * g.nx2 * g.nx3, nc = size_t(g.ni) * g.nj * g.nk;
  f << "CELL_DATA " << ncells << "\n";
  std::vector<float> buf(ncells);
  auto dump = [&](const char *name, const std::vector<double> &a, int comp) {
    size_t n = 0;
    for (int k = 0; k < (g.three_d ? g.nx3 : 1); ++k)
      for (int j = 0; j < (g.two_d ? g.nx2 : 1); ++j)
        for (int i = 0; i < g.nx1; ++i)
          buf[n++] = static_cast<float>(a[comp * nc + (size_t(g.ks + k) * g.nj + (g.js + j)) * g.ni + (g.is + i)]);
    f << "SCALARS " << name << " float\nLOOKUP_TABLE default\n";
    put_be_floats(f, buf);
    f << "\n";
  };
  dump("rho", w, 0);
  dump("prs", w, 4);
  dump("vx1", w, 1);
  dump("vx2", w, 2);
  dump("vx3", w, 3);
  dump("Bx1", b, 0);
  dump("Bx2", b, 1);
  dump("Bx3", b, 2);
  return f ? CHQ_OK : chq::fail(CHQ_EINVAL, std::string("write failed: ") + path);
}

// restart::Write (restart.hpp:19-34): header, then u0 and b1f0..b3f0 in the reference layout
int chq_restart_write(chq_ctx *ctx, const char *path, double time, int cycle) {
  if (!ctx || !path) return chq::fail(CHQ_EINVAL, "chq_restart_write: null context or path");
  int iv[16];
  if (int e = chq_geometry(ctx, iv, nullptr)) return e;
  RstHeader h;
  std::memset(&h, 0, sizeof(h));
  h.magic = kRstMagic;
  h.nx1 = iv[0]; h.nx2 = iv[1]; h.nx3 = iv[2]; h.ng = iv[3];
  h.cycle = cycle;
  h.time = time;
  h.real_size = int32_t(sizeof(double));
  std::ofstream f(path, std::ios::binary);
  if (!f) return chq::fail(CHQ_EINVAL, std::string("cannot write restart file ") + path);
  f.write(reinterpret_cast<const char *>(&h), sizeof(h));
  std::vector<double> buf;
  for (int a : kRstArrays) {
    buf.resize(chq_array_count(ctx, a));
    if (int e = chq_download(ctx, a, buf.data(), buf.size())) return e;
    f.write(reinterpret_cast<const char *>(buf.data()), std::streamsize(buf.size() * sizeof(double)));
  }
  return f ? CHQ_OK : chq::fail(CHQ_EINVAL, std::string("write failed: ") + path);
}

// restart::Read (restart.hpp:37-58) into u0 and b1f0..b3f0; the caller then runs chq_init_resume
// (boundaries + con2prim, main.cpp:196-201)
int chq_restart_read(chq_ctx *ctx, const char *path, double *time, int *cycle) {
  if (!ctx || !path) return chq::fail(CHQ_EINVAL, "chq_restart_read: null context or path");
  std::ifstream f(path, std::ios::binary);
  if (!f) return chq::fail(CHQ_EINVAL, std::string("cannot open restart file ") + path);
  RstHeader h;
  f.read(reinterpret_cast<char *>(&h), sizeof(h));
  int iv[16];
  if (int e = chq_geometry(ctx, iv, nullptr)) return e;
  if (!f || h.magic != kRstMagic) return chq::fail(CHQ_EINVAL, std::string("not a Chasqui restart file: ") + path);
  if (h.nx1 != iv[0] || h.nx2 != iv[1] || h.nx3 != iv[2] || h.ng != iv[3] || h.real_size != int32_t(sizeof(double)))
    return chq::fail(CHQ_EINVAL, "restart file does not match the mesh (nx, nghost or precision differ)");
  std::vector<double> buf;
  for (int a : kRstArrays) {
    buf.resize(chq_array_count(ctx, a));
    f.read(reinterpret_cast<char *>(buf.data()), std::streamsize(buf.size() * sizeof(double)));
    if (!f) return chq::fail(CHQ_EINVAL, std::string("restart file is truncated: ") + path);
    if (int e = chq_upload(ctx, a, buf.data(), buf.size())) return e;
  }
  if (time) *time = h.time;
  if (cycle) *cycle = h.cycle;
  return CHQ_OK;
}

}  // extern "C"
