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
