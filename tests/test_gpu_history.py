"""GPU: the .hst columns reduced on the device (chq_history) against numpy sums of the same
arrays downloaded in the reference layout, written out here from output.hpp:98-118: the volume
integrals of mass, momentum, total, kinetic, magnetic and internal energy over the active
cells, the maximum of |div B| (bitwise: the per-cell expression is evaluated in the same order)
and its mean. The sums' order differs (blocks of rows on the device, numpy's pairwise sum), so
they are held to 1e-12 of the sum of the absolute terms."""
import numpy as np
import pytest

from helpers import deck

pytestmark = pytest.mark.gpu

CASES = [
    ("blast_3d.in", ["mesh.nx1=20", "mesh.nx2=18", "mesh.nx3=16"], "fp64", 3),
    ("blast_3d.in", ["mesh.nx1=20", "mesh.nx2=18", "mesh.nx3=16"], "mixed", 3),
    ("orszag_tang.in", ["mesh.nx1=70", "mesh.nx2=36"], "fp64", 6),
    ("orszag_tang.in", ["mesh.nx1=300", "mesh.nx2=40"], "mixed", 2),     # rows longer than a block
    ("brio_wu.in", ["mesh.nx1=600"], "fp64", 20),
    ("cpaw_3d_oblique.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=40"], "fp64", 2),  # > 1 block of rows
]


def expected(m, buf):
    u, w, bcc = m.download(f"u{buf}"), m.download("w"), m.download("bcc")
    b1, b2, b3 = (m.download(f"b{d}f{buf}") for d in (1, 2, 3))
    k, j, i = slice(m.ks, m.ke + 1), slice(m.js, m.je + 1), slice(m.is_, m.ie + 1)
    dx1, dx2, dx3 = m.dx
    d = w[0][k, j, i]
    v2 = w[1][k, j, i] * w[1][k, j, i] + w[2][k, j, i] * w[2][k, j, i] + w[3][k, j, i] * w[3][k, j, i]
    bb = bcc[0][k, j, i] * bcc[0][k, j, i] + bcc[1][k, j, i] * bcc[1][k, j, i] + bcc[2][k, j, i] * bcc[2][k, j, i]
    terms = [u[n][k, j, i] for n in range(5)] + [0.5 * d * v2, 0.5 * bb, u[4][k, j, i] - 0.5 * d * v2 - 0.5 * bb]
    db = (b1[k, j, m.is_ + 1:m.ie + 2] - b1[k, j, i]) * (1.0 / dx1)
    if m.two_d:
        db = db + (b2[k, m.js + 1:m.je + 2, i] - b2[k, j, i]) * (1.0 / dx2)
    if m.three_d:
        db = db + (b3[m.ks + 1:m.ke + 2, j, i] - b3[k, j, i]) * (1.0 / dx3)
    vol = dx1 * (dx2 if m.two_d else 1.0) * (dx3 if m.three_d else 1.0)
    ncell = float(m.nx1) * m.nx2 * m.nx3
    sums = [float(np.sum(t)) * vol for t in terms] + [float(np.max(np.abs(db))), float(np.sum(np.abs(db))) / ncell]
    scale = [float(np.sum(np.abs(t))) * vol for t in terms] + [0.0, float(np.sum(np.abs(db))) / ncell]
    return sums, scale


@pytest.mark.parametrize("name,ov,prec,steps", CASES, ids=[f"{c[0]}-{c[2]}-{i}" for i, c in enumerate(CASES)])
def test_history_columns_match_numpy_sums(name, ov, prec, steps):
    from chasqui_amd.deck import Deck
    print(f"case history {name} {ov} {prec}", flush=True)
    d = Deck(deck(name), ov + [f"mhd.precision={prec}", "time.tlim=1e9", "time.nlim=-1"])
    m = d.make_mesh()
    d.close()
    m.step(steps)
    m.sync()
    for buf in (0, 1):
        got = m.history(buf)
        exp, scale = expected(m, buf)
        for q in range(10):
            if q == 8:
                assert got[q] == exp[q], (buf, q, got[q], exp[q])
            else:
                assert abs(got[q] - exp[q]) <= 1e-12 * max(scale[q], 1e-300), (buf, q, got[q], exp[q])
        again = m.history(buf)
        assert again == got       # fixed order: the same bits on every call
    assert m.history(0)[0] > 0
    m.close()
