// Per-cell reconstruction for the face-flux kernels.
//
// The reference reconstructs both sides of every face from a 6-cell stencil, so each cell's
// limited parabola (and its flattening coefficient and donor fallback) is computed twice,
// once for each of its two faces (physics.hpp:99-165). The result for a cell depends only on
// that cell's stencil, so computing it once per cell and handing the two half-states to the
// two faces gives bitwise the same face states with about half the arithmetic:
//   Rf(c) = state on the right side of face c-1/2 (left edge of cell c),
//   Lf(c) = state on the left side of face c+1/2 (right edge of cell c).
// The identities used (each is the reference's own call, re-indexed to the cell):
//   PPM  (recon.hpp:60-78):  slope d_c = MC(q_{c-1}, q_c, q_{c+1});
//        e_{c+1/2} = Interface(q_c, q_{c+1}, d_c, d_{c+1});
//        (Rf, Lf) = Limit(q_c, e_{c-1/2}, e_{c+1/2})
//   PPMX (recon.hpp:129-140): e_{c+1/2} = XInterface(q_{c-1}, q_c, q_{c+1}, q_{c+2});
//        (Rf, Lf) = XLimit(q_{c-2..c+2}, e_{c-1/2}, e_{c+1/2})
//   WENO-Z (recon.hpp:166-171): Lf = WZ(q_{c-2..c+2}), Rf = WZ(q_{c+2..c-2})
//   PLM  (recon.hpp:34-37):  s = VL(q_{c-1}, q_c, q_{c+1}); Lf = q_c + s/2; Rf = q_c - s/2
//   DC:  Lf = Rf = q_c
//   flattening of cell c (recon.hpp:179-187, physics.hpp:156-162) pulls both toward q_c;
//   the donor fallback (physics.hpp:163-165) replaces a whole half-state by the cell values.
#pragma once
#include "common.hpp"
#include "numerics.hpp"

namespace chq {

// Finish one cell: flattening + donor fallback on the NV half-state pairs.
// v0 = density slot, v4 = pressure slot (the fallback tests them).
template <class T, int NV>
__device__ __forceinline__ void finish_cell(T Lf[NV], T Rf[NV], const T qc[NV], T fcoef, bool flat, T dfloor,
                                            T pfloor) {
  if (flat && fcoef > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      Lf[n] += fcoef * (qc[n] - Lf[n]);
      Rf[n] += fcoef * (qc[n] - Rf[n]);
    }
  }
  if (Lf[0] <= dfloor || Lf[4] <= pfloor) {
#pragma unroll
    for (int n = 0; n < NV; ++n) Lf[n] = qc[n];
  }
  if (Rf[0] <= dfloor || Rf[4] <= pfloor) {
#pragma unroll
    for (int n = 0; n < NV; ++n) Rf[n] = qc[n];
  }
}

// Half-states of one cell from its 5-cell window q[-2..2] (all methods); PPM recomputes the
// three slopes and two edges it needs (the march kernels use the incremental form instead).
template <class T, Recon RC>
__device__ __forceinline__ void cell_halves(T qm2, T qm1, T q0, T qp1, T qp2, T &Rf, T &Lf) {
  using R = Rec<T>;
  if constexpr (RC == Recon::dc) {
    Lf = q0;
    Rf = q0;
  } else if constexpr (RC == Recon::plm) {
    const T s = R::vl(qm1, q0, qp1);
    Lf = q0 + T(0.5) * s;
    Rf = q0 - T(0.5) * s;
  } else if constexpr (RC == Recon::ppm) {
    const T dm = R::mc(qm2, qm1, q0), d0 = R::mc(qm1, q0, qp1), dp = R::mc(q0, qp1, qp2);
    T lo = R::edge(qm1, q0, dm, d0), hi = R::edge(q0, qp1, d0, dp);
    R::mono(q0, lo, hi);
    Rf = lo;
    Lf = hi;
  } else if constexpr (RC == Recon::ppmx) {
    T lo = R::xedge(qm2, qm1, q0, qp1), hi = R::xedge(qm1, q0, qp1, qp2);
    R::xmono(qm2, qm1, q0, qp1, qp2, lo, hi);
    Rf = lo;
    Lf = hi;
  } else {
    Lf = R::wz(qm2, qm1, q0, qp1, qp2);
    Rf = R::wz(qp2, qp1, q0, qm1, qm2);
  }
}

}  // namespace chq
