"""The C++ host layer on a CPU-only host: deck parsing and problem setups (bitwise equal to
the oracle's restatement of pgen.hpp), error handling of bad decks, parameter echo."""
import ctypes as C

import numpy as np
import pytest

from helpers import deck

po = pytest.importorskip("pyoracle")
from chasqui_amd.deck import Deck, DeckError  # noqa: E402

CASES = [
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8"]),
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=1", "problem.perturb=0"]),
    ("cpaw_1d.in", ["mesh.nx1=64"]),
    ("cpaw_1d.in", ["mesh.nx1=16", "mesh.nx2=16", "mesh.x1max=1.4142135623730951",
                    "mesh.x2max=1.4142135623730951", "problem.angle=45"]),
    ("cpaw_3d_oblique.in", ["mesh.nx1=8", "mesh.nx2=8", "mesh.nx3=8"]),
    ("orszag_tang.in", ["mesh.nx1=20", "mesh.nx2=24"]),
    ("orszag_tang.in", ["mesh.nx1=16", "mesh.nx2=16", "problem.pres=0.14", "problem.b0=0.3"]),
    ("brio_wu.in", ["mesh.nx1=100"]),
    # large enough for the host layer to split the setup across threads
    ("blast_3d.in", ["mesh.nx1=64", "mesh.nx2=64", "mesh.nx3=64"]),
    ("orszag_tang.in", ["mesh.nx1=384", "mesh.nx2=384"]),
    ("cpaw_3d_oblique.in", ["mesh.nx1=48", "mesh.nx2=48", "mesh.nx3=48"]),
    ("field_loop.in", []),
    ("field_loop.in", ["mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8", "mesh.x3min=-0.5", "mesh.x3max=0.5",
                       "problem.vz=0.25"]),
    ("brio_wu.in", ["mesh.nx1=100", "problem.perturb=0.05"]),
    ("brio_wu.in", ["mesh.nx1=40", "mesh.nx2=6", "problem.perturb=0.2"]),
]


DIFF_CASES = [(None, c[1]) for c in __import__("diffusion_cases").DECAY_CASES]


@pytest.mark.parametrize("name,ov", CASES + DIFF_CASES)
def test_initial_state_bitwise_vs_oracle(name, ov):
    text = deck(name) if name else __import__("diffusion_cases").DIFF_DECK
    d = Deck(text, ov)
    o = po.OracleSim(text, ov, pgen_only=True)
    for host, ora in (("u0", "u0"), ("w", "w"), ("b1f0", "b1f0"), ("b2f0", "b2f0"), ("b3f0", "b3f0")):
        a, b = d.initial(host), o.get(ora)
        assert a.shape == b.shape
        assert np.array_equal(a, b), f"{host}: {np.count_nonzero(a != b)} entries differ"


def test_blast_noise_statistics():
    d = Deck(deck("blast_3d.in"), ["mesh.nx1=32", "mesh.nx2=32", "mesh.nx3=32", "problem.perturb=1e-3"])
    rho = d.initial("w")[0]
    dev = (rho - 1.0) / 1e-3
    assert dev.min() >= -1.0 and dev.max() < 1.0
    assert abs(dev.mean()) < 0.02 and abs(dev.std() - 1 / np.sqrt(3)) < 0.01   # U[-1, 1)


def test_defaults_reproduce_reference_orszag_tang():
    """problem.pres / problem.b0 absent: the reference constants 5/(12 pi), 1/sqrt(4 pi)."""
    d = Deck(deck("orszag_tang.in"), ["mesh.nx1=8", "mesh.nx2=8"])
    assert float(d.get("problem", "pres")) == 5.0 / (12.0 * np.pi)
    assert float(d.get("problem", "b0")) == 1.0 / np.sqrt(4.0 * np.pi)


def test_deck_defaults_recorded_and_printed():
    d = Deck(deck("cpaw_1d.in"))
    assert d.get("mhd", "fofc") == "true"
    assert d.get("time", "nlim") == "-1"
    assert d.get("mesh", "nghost") == "3"
    txt = d.text()
    assert "[mhd]" in txt and "  reconstruct = ppm" in txt
    assert d.tlim == 1.0 and d.nlim == -1


def test_overrides_win_over_file():
    d = Deck(deck("cpaw_1d.in"), ["mesh.nx1=32", "mhd.reconstruct=wenoz", "time.cfl=0.25"])
    assert d.mesh_desc.nx1 == 32 and d.config.recon == 3 and d.config.cfl == 0.25


@pytest.mark.parametrize("ov,msg", [
    (["mhd.reconstruct=nonsense"], "unknown reconstruct"),
    (["mhd.rsolver=roe"], "unknown rsolver"),
    (["mesh.bc_x1=bogus"], "unknown boundary"),
    (["time.integrator=rk7"], "unknown integrator"),
    (["mesh.nghost=2"], "nghost too small"),
    (["problem.name=nothere"], "unknown problem"),
])
def test_bad_decks_raise(ov, msg):
    with pytest.raises(DeckError, match=msg):
        Deck(deck("cpaw_1d.in"), ov)


def test_missing_required_key():
    with pytest.raises(DeckError, match="missing parameter mesh.nx1"):
        Deck("[problem]\nname = cpaw\n")


def test_characteristic_key_and_its_limits():
    """mhd.characteristic (main.cpp:123) reaches the configuration; the device accepts it in FP64
    and refuses it in mixed precision (checked without a GPU through chq_bytes_for)."""
    from chasqui_amd import _capi
    d = Deck(deck("orszag_tang.in"), ["mhd.characteristic=true", "mhd.reconstruct=ppmx"])
    assert d.config.characteristic == 1
    L = _capi.lib()
    md, cfg = d.mesh_desc, d.config
    assert L.chq_bytes_for(C.byref(md), C.byref(cfg)) > 0
    cfg.precision = 1
    assert L.chq_bytes_for(C.byref(md), C.byref(cfg)) == 0
    assert "characteristic" in L.chq_last_error().decode()
    cfg.recon = 0   # donor cell ignores the switch (physics.hpp:71)
    assert L.chq_bytes_for(C.byref(md), C.byref(cfg)) > 0



def test_limiter_and_source_keys_and_their_limits():
    """mhd.flux_limiter, [gravity] and [cooling] (main.cpp:141, 148-162) reach the configuration;
    the device accepts them in FP64 and refuses them in mixed precision."""
    from chasqui_amd import _capi
    d = Deck(deck("blast_3d.in"), ["mhd.flux_limiter=true", "gravity.uniform=true", "gravity.g3=-0.5",
                                   "gravity.point_mass=true", "gravity.gm=2", "gravity.x0=0.1", "gravity.rsoft=0.05",
                                   "cooling.enabled=true", "cooling.lambda0=3", "cooling.alpha=0.5",
                                   "cooling.tfloor=0.01"])
    c = d.config
    assert (c.flux_limiter, c.grav_uniform, c.grav_point, c.cooling) == (1, 1, 1, 1)
    assert (c.g1, c.g2, c.g3, c.gm, c.gx0, c.rsoft) == (0.0, 0.0, -0.5, 2.0, 0.1, 0.05)
    assert (c.lambda0, c.cool_alpha, c.tfloor) == (3.0, 0.5, 0.01)
    L = _capi.lib()
    md = d.mesh_desc
    assert L.chq_bytes_for(C.byref(md), C.byref(c)) > 0
    for field in ("flux_limiter", "grav_uniform", "grav_point", "cooling"):
        one = type(c).from_buffer_copy(c)
        for f in ("flux_limiter", "grav_uniform", "grav_point", "cooling"):
            setattr(one, f, 1 if f == field else 0)
        one.precision = 1
        assert L.chq_bytes_for(C.byref(md), C.byref(one)) == 0, field
        assert "source terms" in L.chq_last_error().decode()
