"""Whole-pipeline invariants of the oracle (README.md:100, 150-152; physics.sh:57, 91-92):
div B at round-off, conservation in periodic and closed boxes, bitwise mirror symmetry of a
blast, fofc forced from the host."""
import numpy as np
import pytest

from helpers import deck

po = pytest.importorskip("pyoracle")


def div_b(o, s=0):
    b1, b2, b3 = o.get(f"b1f{s}"), o.get(f"b2f{s}"), o.get(f"b3f{s}")
    ks, ke, js, je, is_, ie = o.ks, o.ke, o.js, o.je, o.is_, o.ie
    g, r = o.geometry()
    dx1, dx2, dx3 = r[0], r[1], r[2]
    k, j, i = slice(ks, ke + 1), slice(js, je + 1), slice(is_, ie + 1)
    d = (b1[k, j, is_ + 1:ie + 2] - b1[k, j, i]) / dx1
    if o.two_d:
        d = d + (b2[k, js + 1:je + 2, i] - b2[k, j, i]) / dx2
    if o.three_d:
        d = d + (b3[ks + 1:ke + 2, j, i] - b3[k, j, i]) / dx3
    return float(np.max(np.abs(d)))


def totals(o):
    u = o.get("u0")[:, o.ks:o.ke + 1, o.js:o.je + 1, o.is_:o.ie + 1]
    return u.reshape(6, -1).sum(axis=1)


@pytest.mark.parametrize("name,ov,steps", [
    ("blast_3d.in", ["mesh.nx1=24", "mesh.nx2=24", "mesh.nx3=24"], 15),
    ("orszag_tang.in", ["mesh.nx1=48", "mesh.nx2=48"], 60),
    ("cpaw_3d_oblique.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12"], 10),
    # the reference's recommended preset (inputs/recommended_mhd.in): characteristic PPMX,
    # LHLLD, flattening, dual energy, fofc
    ("orszag_tang.in", ["mesh.nx1=48", "mesh.nx2=48", "mhd.characteristic=true", "mhd.reconstruct=ppmx",
                        "mhd.rsolver=lhlld", "mhd.dual_energy=true", "time.cfl=0.4"], 40),
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16", "mhd.characteristic=true"], 8),
])
def test_divb_and_conservation(name, ov, steps):
    o = po.OracleSim(deck(name), ov)
    t0 = totals(o)
    assert div_b(o) <= 5e-13
    o.step(steps)
    assert o.cycle == steps
    assert div_b(o) <= 5e-13
    t1 = totals(o)
    for n in (0, 4):    # mass and total energy in a periodic domain
        assert abs(t1[n] - t0[n]) <= 1e-12 * abs(t0[n]), (n, t0[n], t1[n])
    for n in (1, 2, 3):
        assert abs(t1[n] - t0[n]) <= 1e-12 * max(abs(t0[0]), 1.0)


def test_reflecting_box_conserves_mass():
    """physics.sh:57, 91-92: blast in a box with reflecting x walls (angle 90: B normal to
    the walls vanishes) keeps its mass to 1e-9."""
    o = po.OracleSim(deck("blast_3d.in"), ["mesh.nx1=32", "mesh.nx2=32", "mesh.nx3=1", "mesh.bc_x1=reflect",
                                           "problem.angle=90", "problem.perturb=0"])
    m0 = totals(o)[0]
    o.step(40)
    assert abs(totals(o)[0] - m0) <= 1e-9 * m0


def test_blast_mirror_symmetry_bitwise():
    """README.md:150-152: a blast keeps its mirror symmetry to 0.0. With B in the x-y plane
    the z-mirror and the combined (x, y) flip are symmetries of ideal MHD (B -> -B)."""
    o = po.OracleSim(deck("blast_3d.in"), ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16", "problem.perturb=0"])
    o.step(8)
    a = (slice(None), slice(o.ks, o.ke + 1), slice(o.js, o.je + 1), slice(o.is_, o.ie + 1))
    u = o.get("u0")[a]
    assert np.array_equal(u[0], u[0][::-1, :, :])                    # density, z-mirror
    assert np.array_equal(u[3], -u[3][::-1, :, :])                   # z-momentum flips
    assert np.array_equal(u[4], u[4][::-1, :, :])
    assert np.array_equal(u[0], u[0][:, ::-1, ::-1])                 # density, (x, y) flip
    assert np.array_equal(u[1], -u[1][:, ::-1, ::-1])


def test_fofc_forced_by_floor():
    """Raising the pressure floor above the ambient state forces fofc (main.cpp:135-136, 78-91)."""
    o = po.OracleSim(deck("blast_3d.in"), ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12", "mhd.pfloor=0.1"])
    o.step(2)
    assert o.fofc_count() > 0
    o2 = po.OracleSim(deck("blast_3d.in"), ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12"])
    o2.step(2)
    assert o2.fofc_count() == 0


def test_bad_deck_key_raises():
    with pytest.raises(RuntimeError):
        po.OracleSim(deck("blast_3d.in"), ["mhd.reconstruct=nonsense"])
    with pytest.raises(RuntimeError):
        po.OracleSim(deck("blast_3d.in"), ["mesh.bc_x1=bogus"])


def test_dual_energy_without_switch_is_the_total_energy_path():
    """With dual energy on but e_int >= de_eta * E everywhere, the advected entropy is never
    used and con2prim resynchronises u[IENT] from rho and p in every cell (physics.hpp:463-472):
    the whole state is bitwise the plain path's."""
    ov = ["mesh.nx1=32", "mesh.nx2=32", "mhd.reconstruct=ppmx", "mhd.rsolver=lhlld"]
    a = po.OracleSim(deck("orszag_tang.in"), ov + ["mhd.dual_energy=true"])
    b = po.OracleSim(deck("orszag_tang.in"), ov)
    a.step(10)
    b.step(10)
    for f in ("u0", "b1f0", "b2f0", "b3f0", "w", "bcc"):
        assert np.array_equal(a.get(f), b.get(f)), f


def test_dual_energy_switch_uses_the_entropy_pressure():
    """Where the switch fires (de_eta raised), con2prim's pressure is s * rho^gamma with
    s = u[IENT] / rho, and the total energy is rebuilt from it (physics.hpp:463-469)."""
    o = po.OracleSim(deck("brio_wu.in"), ["mesh.nx1=200", "mhd.dual_energy=true", "mhd.de_eta=0.5"])
    o.step(30)
    gamma = 2.0
    ks, js, i = o.ks, o.js, slice(o.is_, o.ie + 1)
    u, w, bc = o.get("u0")[:, ks, js, i], o.get("w")[:, ks, js, i], o.get("bcc")[:, ks, js, i]
    d = u[0]
    kin = 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / d
    ei = u[4] - kin - 0.5 * (bc ** 2).sum(0)
    sw = ei < 0.5 * np.abs(u[4])
    assert np.count_nonzero(sw) > 0.3 * d.size
    p_ent = (u[5][sw] / d[sw]) * d[sw] ** gamma
    assert np.max(np.abs(w[4][sw] - p_ent) / p_ent) < 1e-13
    # E = kin + B^2/2 + p/(gamma-1) in switched cells
    e_re = kin[sw] + 0.5 * (bc[:, sw] ** 2).sum(0) + w[4][sw] / (gamma - 1.0)
    assert np.max(np.abs(u[4][sw] - e_re) / np.abs(u[4][sw])) < 1e-13


def test_fofc_flag_count_closed_forms():
    """fofc.hpp:24-40, known counts: a pressure floor above every state flags every active cell
    of the tentative state; Brio-Wu with dfloor = 0.5 flags exactly the right half (its uniform
    cells keep rho = 0.125 through a stage, the cells next to the jump stay far from 0.5), and a
    negligible floor flags nothing."""
    o = po.OracleSim(deck("blast_3d.in"), ["mesh.nx1=8", "mesh.nx2=6", "mesh.nx3=4", "mhd.pfloor=1e6"])
    assert o.stage(0, 1, 1.0, 0.0, 0.1 * o.min_dt) == 8 * 6 * 4
    o = po.OracleSim(deck("brio_wu.in"), ["mesh.nx1=64", "mhd.dfloor=0.5"])
    assert o.stage(0, 1, 1.0, 0.0, 0.5 * o.min_dt) == 32
    o = po.OracleSim(deck("brio_wu.in"), ["mesh.nx1=64"])
    assert o.stage(0, 1, 1.0, 0.0, 0.5 * o.min_dt) == 0
