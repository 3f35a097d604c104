// A parsed input deck: run configuration (main.cpp:112-185) and the problem's initial state
// in the reference layout (pgen.hpp:269-307), built on the host before upload.
#pragma once
#include <string>
#include <sys/mman.h>
#include <cstdlib>
#include <memory>
#include <new>
#include <utility>
#include <vector>

#include "../../../include/chasqui_amd.h"
#include "params.hpp"

namespace chq {

struct HostGrid {
  int nx1 = 1, nx2 = 1, nx3 = 1, ng = 3;
  int ni = 0, nj = 0, nk = 0, is = 0, ie = 0, js = 0, je = 0, ks = 0, ke = 0;
  double x1min = 0, x1max = 1, x2min = 0, x2max = 1, x3min = 0, x3max = 1, dx1 = 0, dx2 = 0, dx3 = 0;
  bool two_d = false, three_d = false;
  void build();   // Mesh::Allocate index arithmetic (mesh.hpp:37-48)
  double x1v(int i) const { return x1min + (i - is + 0.5) * dx1; }   // mesh.hpp:66-71
  double x2v(int j) const { return x2min + (j - js + 0.5) * dx2; }
  double x3v(int k) const { return x3min + (k - ks + 0.5) * dx3; }
  double x1f(int i) const { return x1min + (i - is) * dx1; }
  double x2f(int j) const { return x2min + (j - js) * dx2; }
  double x3f(int k) const { return x3min + (k - ks) * dx3; }
  size_t ncell() const { return size_t(nk) * nj * ni; }
  size_t nface() const { return size_t(nk + 1) * (nj + 1) * (ni + 1); }
  size_t c(int k, int j, int i) const { return (size_t(k) * nj + j) * ni + i; }
  size_t f(int k, int j, int i) const { return (size_t(k) * (nj + 1) + j) * (ni + 1) + i; }
};

// Allocator of the host-side initial state: new elements are not zero-filled (the problem
// generators write every element, from several threads, so pages are first touched in
// parallel), and large arrays are 2-MB aligned and marked for transparent huge pages (512x
// fewer page faults on a 200^3 mesh).
template <class T>
struct NoInitAlloc {
  using value_type = T;
  static constexpr size_t kHuge = size_t(2) << 20;
  NoInitAlloc() = default;
  template <class U> NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  template <class U> struct rebind { using other = NoInitAlloc<U>; };
  T *allocate(size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes < 4 * kHuge) return static_cast<T *>(::operator new(bytes));
    const size_t len = (bytes + kHuge - 1) / kHuge * kHuge;
    void *p = std::aligned_alloc(kHuge, len);
    if (!p) throw std::bad_alloc();
    madvise(p, len, MADV_HUGEPAGE);
    return static_cast<T *>(p);
  }
  void deallocate(T *p, size_t n) noexcept {
    if (n * sizeof(T) < 4 * kHuge) ::operator delete(p);
    else std::free(p);
  }
  template <class U> void construct(U *p) noexcept { ::new (static_cast<void *>(p)) U; }
  template <class U, class... A> void construct(U *p, A &&...a) { ::new (static_cast<void *>(p)) U(std::forward<A>(a)...); }
  template <class U> bool operator==(const NoInitAlloc<U> &) const noexcept { return true; }
  template <class U> bool operator!=(const NoInitAlloc<U> &) const noexcept { return false; }
};
using HostArray = std::vector<double, NoInitAlloc<double>>;

struct Deck {
  Params p;
  chq_mesh_desc mesh{};
  chq_config cfg{};
  HostGrid g;
  double tlim = 1.0;
  int nlim = -1, ndiag = 100;
  std::string problem, basename;
  double vtk_dt = -1, hst_dt = -1, rst_dt = -1;
  bool compute_error = false;
  // initial state, reference layout: u0 [6][nk][nj][ni], w [6][..], b? [(nk+1)(nj+1)(ni+1)]
  HostArray u0, w, b1, b2, b3;
  bool ready = false;  // the initial state has been built

  void configure();    // main.cpp:112-185
  void initialize();   // pgen.hpp:269-307
};

}  // namespace chq
