"""Pins the CPU oracle against the reference's published CPAW L1 table (README.md:90-96).

Each entry is a full run (RK3, T = 1) of the whole stage pipeline — reconstruction, HLLD,
flattening, corner EMF, CT update, boundaries, con2prim, the dt logic — so a dropped term
or a wrong index anywhere on the path moves the 3-digit value.
"""
import json
import math
import os

import numpy as np
import pytest

from helpers import deck

po = pytest.importorskip("pyoracle")

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "readme_cpaw_l1.json")))


def _l1_total(method, n):
    s = po.OracleSim(deck("cpaw_1d.in"), [f"mesh.nx1={n}", f"mhd.reconstruct={method}"])
    s.run()
    assert s.time == 1.0
    e = s.l1_error()
    return e[9], e[9] - e[5]


@pytest.mark.parametrize("method", ["plm", "ppm", "ppmx", "wenoz"])
def test_cpaw_1d_table(method):
    errs = {}
    tol_units = GOLD["tolerance_units_last_digit"]
    for n_str, expected in GOLD["table"][method].items():
        _, got = _l1_total(method, int(n_str))
        errs[int(n_str)] = got
        unit = 10.0 ** (math.floor(math.log10(expected)) - 2)   # last quoted digit
        assert abs(got - expected) <= tol_units * unit, f"{method} N={n_str}: {got:.4e} vs README {expected:.2e}"
    # the README's order row is a one-decimal summary (e.g. PLM 2.06 is quoted as 2.0)
    order = math.log2(errs[128] / errs[256])
    assert abs(order - GOLD["orders"][method]) <= 0.1, order


def test_cpaw_config0_value():
    """Config 0 (PPM, N = 128): 6.65e-4, the README's headline accuracy number."""
    full, no_ent = _l1_total("ppm", 128)
    assert f"{no_ent:.2e}" == "6.65e-04"
    assert 6.6e-4 < full < 6.7e-4


@pytest.mark.parametrize("method", ["plm", "wenoz"])
def test_cpaw_2d_oblique_order(method):
    """README.md:98 (tests/convergence.sh:14-17 decks): 45-degree CPAW on a sqrt(2) box,
    second order for PLM and WENO-Z (limited by the face-to-centre B average and the corner EMF)."""
    text = deck("cpaw_1d.in")
    errs = []
    for n in (64, 128):
        ov = [f"mesh.nx1={n}", f"mesh.nx2={n}", "mesh.x1max=1.4142135623730951", "mesh.x2max=1.4142135623730951",
              "problem.angle=45.0", f"mhd.reconstruct={method}"]
        s = po.OracleSim(text, ov)
        s.run()
        e = s.l1_error()
        errs.append(e[9] - e[5])
    order = math.log2(errs[0] / errs[1])
    assert abs(order - 2.0) < 0.1, order


# ---------------------------------------------------------------- explicit diffusion, RKL2 STS
from diffusion_cases import DECAY_CASES, DIFF_DECK, decay_amplitude, exact_amplitude  # noqa: E402


def _active(o):
    return (slice(o.ks, o.ke + 1), slice(o.js, o.je + 1), slice(o.is_, o.ie + 1))


@pytest.mark.parametrize("sts", [False, True], ids=["explicit", "sts"])
@pytest.mark.parametrize("case", DECAY_CASES, ids=[c[0] for c in DECAY_CASES])
def test_diffusion_decay_exact_solution(case, sts):
    """physics.sh:30-36 and 72-81: viscous, resistive and conductive decay of a sinusoid along
    each axis against exp(-D k^2 t), in the stage (explicit) and operator-split (RKL2 STS)."""
    label, ov, var, axis, D, t, tol = case
    o = po.OracleSim(DIFF_DECK, ov + (["diffusion.sts=true"] if sts else []))
    o.run()
    assert abs(o.time - t) < 1e-12
    A = decay_amplitude(o.get("w"), o.get("bcc"), _active(o), var, axis)
    assert abs(A / exact_amplitude(D, t) - 1) < tol, (label, A, exact_amplitude(D, t))


def test_sts_takes_hyperbolic_steps():
    """RKL2 (sts.hpp:35-39): the loop keeps the hyperbolic step, the parabolic limit is covered by
    s stages with tau <= dt_par (s^2 + s - 2) / 4, so far fewer steps are taken."""
    ov = DECAY_CASES[0][1]
    a = po.OracleSim(DIFF_DECK, ov)
    a.run()
    b = po.OracleSim(DIFF_DECK, ov + ["diffusion.sts=true"])
    b.run()
    assert b.cycle < a.cycle / 1.5


def test_resistive_mhd_keeps_div_b_and_conserves():
    """Ohmic resistivity and viscosity on the 3D blast: CT keeps div B at round-off (the resistive
    EMF goes through the same curl, diffusion.hpp:1-7) and the periodic box conserves mass,
    momentum and energy."""
    from test_oracle_invariants import div_b, totals
    ov = ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16", "diffusion.eta=0.01", "diffusion.nu=0.01",
          "diffusion.kappa=0.01"]
    for extra in ([], ["diffusion.sts=true"]):
        o = po.OracleSim(deck("blast_3d.in"), ov + extra)
        t0 = totals(o)
        o.step(6)
        assert div_b(o) < 5e-13
        t1 = totals(o)
        assert np.all(np.abs(t1[:5] - t0[:5]) <= 1e-12 * np.maximum(np.abs(t0[:5]), 1.0))
