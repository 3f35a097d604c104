"""GPU at the headline size: the 3D MHD blast at 256^3 (PPM + HLLD + RK3 + CT + fofc +
flattening, periodic, seeded density noise; the benchmark workload of bench.py).

* FP64: the configuration's twenty RK3 steps bitwise against the oracle on the full mesh (u0
  except the entropy slot, the face fields, t, the CFL minimum and the fofc count, which is
  nonzero; u0[IENT] goes through pow() on both sides, 1e-14).
* FP64 and mixed precision, twenty steps (the configuration's step count): properties that do not
  depend on the size - div B unchanged to round-off (<= 5e-13 relative to max|B| / dx), mass,
  momentum and energy conserved by the periodic flux form (1e-12 of the sum of |q|; the fofc repair
  replaces fluxes, it does not break the flux form), density and pressure positive.
"""
import numpy as np
import pytest

import bench
from helpers import face_region, gpu_mesh_for, max_rel, n_diff, upload_oracle_state

pytestmark = pytest.mark.gpu

po = pytest.importorskip("pyoracle")

NG = 3
ACT = (slice(NG, NG + 256),) * 3


def _divb(m):
    b1, b2, b3 = (m.download(f"b{d}f0") for d in (1, 2, 3))
    k, j, i = ACT
    sh = slice(NG + 1, NG + 257)
    div = (b1[k, j, sh] - b1[k, j, i]) + (b2[k, sh, i] - b2[k, j, i]) + (b3[sh, j, i] - b3[k, j, i])
    bmax = max(float(np.max(np.abs(b))) for b in (b1, b2, b3))
    return div, bmax   # div * dx (dx is the same on the three axes)


def _totals(m):
    u = m.download("u0")
    k, j, i = ACT
    return np.array([np.sum(u[n][k, j, i]) for n in range(5)]), np.array([np.sum(np.abs(u[n][k, j, i])) for n in range(5)])


def test_blast256_fp64_twenty_steps_bitwise():
    """The configuration's 20 steps on the full 256^3 mesh, the oracle on the host's cores
    (about 1.5 s per step on 16 threads) against the device; the fofc branch fires on the way."""
    print("case blast 256^3 fp64 vs oracle, 20 steps", flush=True)
    o = po.OracleSim(bench.DECK, [])
    gm, _ = gpu_mesh_for(bench.DECK)
    upload_oracle_state(gm, o)
    gm.init(0.0, 1.0e9, -1)
    assert gm.status()["min_dt"] == o.min_dt
    for n in range(20):   # one step at a time, so the run reports progress
        o.step(1)
        gm.step(1)
        if n % 5 == 4:
            print(f"  step {n + 1}: fofc {o.fofc_count()}", flush=True)
    st = gm.status()
    assert st["cycle"] == o.cycle == 20
    assert st["t"] == o.time and st["min_dt"] == o.min_dt
    assert st["fofc_total"] == o.fofc_count() > 0
    u, ou = gm.download("u0"), o.get("u0")
    for n in range(5):
        assert n_diff(u[n], ou[n]) == 0, f"u0[{n}]: {n_diff(u[n], ou[n])} entries differ"
    assert max_rel(u[5], ou[5]) <= 1e-14
    for d in range(3):
        r = face_region(o, d)
        b = f"b{d + 1}f0"
        assert n_diff(gm.download(b)[r], o.get(b)[r]) == 0, b
    gm.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_blast256_twenty_steps_invariants(prec):
    from chasqui_amd.deck import Deck
    print(f"case blast 256^3 {prec}, 20 steps", flush=True)
    d = Deck(bench.DECK, [f"mhd.precision={prec}"])
    m = d.make_mesh()
    d.close()
    m.init(0.0, 1.0e9, -1)
    div0, _ = _divb(m)
    tot0, mag0 = _totals(m)
    m.step(20)
    st = m.status()
    assert st["cycle"] == 20 and st["t"] > 0
    div1, bmax = _divb(m)
    assert float(np.max(np.abs(div1 - div0))) / bmax <= 5e-13
    tot1, mag1 = _totals(m)
    for n in range(5):   # the momenta start at zero: their scale is the sum of |M| at the end
        assert abs(tot1[n] - tot0[n]) <= 1e-12 * max(mag0[n], mag1[n]), (n, tot0[n], tot1[n])
    w = m.download("w")
    k, j, i = ACT
    assert np.min(w[0][k, j, i]) > 0 and np.min(w[4][k, j, i]) > 0
    m.close()


def test_blast_mirror_symmetry_on_the_device():
    """README.md:150-152 on the device time loop: the noise-free blast (64^3, 20 steps, the
    fused stage with PPM + HLLD + flattening + fofc) keeps its mirror symmetries bitwise: the
    z-mirror (B lies in the x-y plane) and the combined (x, y) -> (-x, -y) flip."""
    from chasqui_amd.deck import Deck
    print("case blast 64^3 mirror symmetry, 20 steps", flush=True)
    d = Deck(bench.DECK, ["mesh.nx1=64", "mesh.nx2=64", "mesh.nx3=64", "problem.perturb=0"])
    m = d.make_mesh()
    d.close()
    m.init(0.0, 1.0e9, -1)
    m.step(20)
    u = m.download("u0")
    a = (slice(None), slice(NG, NG + 64), slice(NG, NG + 64), slice(NG, NG + 64))
    u = u[a]
    assert np.array_equal(u[0], u[0][::-1, :, :]) and np.array_equal(u[4], u[4][::-1, :, :])
    assert np.array_equal(u[3], -u[3][::-1, :, :]) and np.array_equal(u[1], u[1][::-1, :, :])
    assert np.array_equal(u[0], u[0][:, ::-1, ::-1]) and np.array_equal(u[4], u[4][:, ::-1, ::-1])
    assert np.array_equal(u[1], -u[1][:, ::-1, ::-1]) and np.array_equal(u[2], -u[2][:, ::-1, ::-1])
    m.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_blast256_runs_are_deterministic(prec):
    """Two independent contexts advanced the same 20 steps at 256^3 end bitwise equal: the
    atomically built fofc list (its order varies from run to run), the atomic CFL minimum and
    the grid-wide launches leave no trace of scheduling in the results."""
    from chasqui_amd.deck import Deck
    print(f"case blast 256^3 {prec} determinism, 2 x 20 steps", flush=True)
    out = []
    for _ in range(2):
        d = Deck(bench.DECK, [f"mhd.precision={prec}", "mhd.pfloor=0.05"])   # fofc fires every stage
        m = d.make_mesh()
        d.close()
        m.init(0.0, 1.0e9, -1)
        m.step(20)
        st = m.status()
        out.append((m.download("u0"), [m.download(f"b{q}f0") for q in (1, 2, 3)], st["t"], st["fofc_total"]))
        m.close()
    (u_a, b_a, t_a, n_a), (u_b, b_b, t_b, n_b) = out
    assert n_a == n_b and n_a > 0 and t_a == t_b
    assert np.array_equal(u_a, u_b)
    for x, y in zip(b_a, b_b):
        assert np.array_equal(x, y)
