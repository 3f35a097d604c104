"""GPU: seeded combinations of the options against the CPU oracle, bitwise.

Each case draws a problem deck, a mesh (1D, 2D or 3D, ragged sizes), reconstruction, Riemann
solver, integrator, boundary conditions per axis and the switches (fofc, flattening, low-Mach
blend, characteristic variables, dual energy, flux limiter, gravity, super-time-stepped
resistivity) from a fixed seed, runs a few steps through the deck pipeline (host problem setup,
device time loop) and through the oracle, and compares the whole state: u0 bitwise (u0[IENT]
to 1e-14: pow), the face fields bitwise, t, the CFL minimum and the fofc / limiter / RKL2 counts.
Combinations that the single-kernel and time-loop tests do not list one by one are the point.
"""
import random

import numpy as np
import pytest

from helpers import deck, face_region

pytestmark = pytest.mark.gpu

po = pytest.importorskip("pyoracle")

DECKS = ["blast_3d.in", "orszag_tang.in", "brio_wu.in", "cpaw_1d.in", "field_loop.in", "cpaw_3d_oblique.in"]


def _cases(seed=2026, n=36):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        name = rng.choice(DECKS)
        dims = rng.choice([1, 2, 2, 3, 3])
        if name == "blast_3d.in" and dims == 1:
            dims = 2
        if name == "cpaw_3d_oblique.in":   # its wave vector leaves the x-y plane
            dims = 3
        recon = rng.choice(["dc", "plm", "ppm", "ppmx", "wenoz"])
        ng = 2 if recon in ("dc", "plm") else 3
        nx = [rng.randint(max(ng, 4), 20) for _ in range(3)]
        ov = [f"mesh.nx1={nx[0]}", f"mesh.nx2={nx[1] if dims > 1 else 1}", f"mesh.nx3={nx[2] if dims > 2 else 1}",
              f"mhd.reconstruct={recon}", f"mhd.rsolver={rng.choice(['hlle', 'hlld', 'lhlld'])}",
              f"time.integrator={rng.choice(['rk1', 'rk2', 'rk3', 'rk4'])}", "time.nlim=-1", "time.tlim=1e9",
              f"time.cfl={rng.choice([0.2, 0.3, 0.4])}"]
        for ax in range(dims):
            ov.append(f"mesh.bc_x{ax + 1}={rng.choice(['periodic', 'periodic', 'outflow', 'reflect'])}")
        ov.append(f"mhd.fofc={rng.choice(['true', 'true', 'false'])}")
        ov.append(f"mhd.flatten={rng.choice(['true', 'false'])}")
        if rng.random() < 0.25:
            ov.append("mhd.low_mach=true")
        if rng.random() < 0.25 and recon != "dc":
            ov.append("mhd.characteristic=true")
        if rng.random() < 0.3:
            ov.append("mhd.dual_energy=true")
        if rng.random() < 0.3:
            ov.append("mhd.flux_limiter=true")
        if rng.random() < 0.25:
            ov += ["gravity.uniform=true", "gravity.g1=0.05", "gravity.g2=-0.03", "gravity.g3=0.02"]
        if rng.random() < 0.15:
            ov += ["diffusion.eta=0.002", "diffusion.sts=true"]
        if rng.random() < 0.35:
            ov.append("mhd.pfloor=0.11")   # above the problems' low pressures: fofc (and the limiter) work
        out.append((name, ov, rng.randint(2, 6)))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,ov,steps", CASES, ids=[f"{i:02d}-{c[0].split('.')[0]}" for i, c in enumerate(CASES)])
def test_option_combination_bitwise(name, ov, steps):
    from chasqui_amd.deck import Deck
    print(f"case {name} {ov} steps={steps}", flush=True)
    text = deck(name)
    d = Deck(text, ov)
    m = d.make_mesh()
    d.close()
    m.step(steps)
    o = po.OracleSim(text, ov)
    o.step(steps)
    st = m.status()
    assert st["cycle"] == o.cycle == steps
    assert st["t"] == o.time and st["min_dt"] == o.min_dt
    assert st["fofc_total"] == o.fofc_count() and st["limiter_total"] == o.limiter_count
    assert st["sts_stages"] == o.sts_count
    u, ou = m.download("u0"), o.get("u0")
    for q in range(5):
        assert np.array_equal(u[q], ou[q], equal_nan=True), f"u0[{q}]: {np.count_nonzero(u[q] != ou[q])} differ"
    scale = max(float(np.max(np.abs(ou[5]))), 1e-300)
    assert float(np.max(np.abs(u[5] - ou[5]))) <= 1e-14 * scale
    for dd in range(3):
        b = f"b{dd + 1}f0"
        r = face_region(o, dd)
        assert np.array_equal(m.download(b)[r], o.get(b)[r], equal_nan=True), b
    m.close()
