"""GPU: the positivity flux limiter, gravity and cooling against the CPU oracle; and the
reference's own gravity and cooling checks (tests/physics.sh:53-54, 84-86) run on the device.

Stage: positivity::LimitFluxes<DIR> after the diffusion fluxes (main.cpp:60-64,
positivity.hpp:24-89), srcterms::ApplyGravity after the update and after the fofc re-update
(main.cpp:76, 89; srcterms.hpp:21-40). Step: srcterms::ApplyCooling on the end-of-step state with
the step's dt, then con2prim and the new dt (main.cpp:273-290; srcterms.hpp:55-71).

The limiter and gravity are bitwise the oracle's (sqrt and division are correctly rounded on both
sides). Cooling calls exp or pow, whose device versions may differ from the host's in the last
place: cooling runs are held to COOL_RTOL relative to the oracle.
"""
import math

import numpy as np
import pytest

from helpers import deck, face_region, gpu_mesh_for, max_rel, n_diff, upload_oracle_state

pytestmark = pytest.mark.gpu

po = pytest.importorskip("pyoracle")

ENT_RTOL = 1e-14
COOL_RTOL = 1e-12

UNI_DECK = """[mesh]
nx1 = 16
nx2 = 16
nx3 = 16
[time]
tlim = 1.0
[mhd]
reconstruct = plm
[problem]
name = uniform
dens = 2.0
pres = 1.0
"""

LIMITED_BLAST = ["mesh.nx1=24", "mesh.nx2=24", "mesh.nx3=1", "problem.pres_in=1e4", "problem.pres_out=1e-4",
                 "mhd.fofc=false", "mhd.flux_limiter=true", "time.cfl=0.4"]
POINT = ["gravity.point_mass=true", "gravity.gm=0.05", "gravity.rsoft=0.05", "gravity.x0=0.1", "gravity.y0=-0.05",
         "gravity.z0=0.02"]
UNIFORM = ["gravity.uniform=true", "gravity.g1=0.3", "gravity.g2=-0.2", "gravity.g3=0.1"]


def _pair(text, ov):
    o = po.OracleSim(text, list(ov))
    gm, kv = gpu_mesh_for(text, ov)
    upload_oracle_state(gm, o)
    return o, gm, kv


def _same(gpu, ora, label):
    assert gpu.shape == ora.shape, label
    assert n_diff(gpu, ora) == 0, f"{label}: {n_diff(gpu, ora)} entries differ, max rel {max_rel(gpu, ora):.3e}"


def _close(gpu, ora, rtol, label):
    assert gpu.shape == ora.shape, label
    assert max_rel(gpu, ora) <= rtol, f"{label}: max rel {max_rel(gpu, ora):.3e} above {rtol}"


STAGE_CASES = [
    ("blast_3d.in", LIMITED_BLAST),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8", "mhd.flux_limiter=true", "mhd.pfloor=0.1"] + UNIFORM),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8"] + POINT),
    ("orszag_tang.in", ["mesh.nx1=24", "mesh.nx2=20", "mhd.flux_limiter=true", "mhd.dual_energy=true"] + POINT),
]


@pytest.mark.parametrize("name,ov", STAGE_CASES, ids=[f"{c[0]}-{len(c[1])}" for c in STAGE_CASES])
def test_stage_entry_with_limiter_and_gravity_bitwise(name, ov):
    """chq_stage over a full RK3 step: limited fluxes, gravity after each (re-)update. The entropy
    flux (component 5) goes through pow(rho, gamma), held to ENT_RTOL like the entropy itself."""
    print(f"case sources stage {name} {ov}", flush=True)
    o, gm, _ = _pair(deck(name), ov)
    dt = 0.3 * o.min_dt
    table = [(1.0, 0.0, 1.0, 0, 1), (0.75, 0.25, 0.25, 1, 1), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 1, 0)]
    for g0, g1, beta, cur, out in table:
        assert gm.stage(cur, out, g0, g1, beta * dt) == o.stage(cur, out, g0, g1, beta * dt)
        for d in range(1 + o.two_d + o.three_d):
            f = f"f{d + 1}"
            a, b = gm.download(f), o.get(f)
            _same(np.delete(a, 5, axis=0), np.delete(b, 5, axis=0), f"{f} after stage out={out}")
            assert max_rel(a[5], b[5]) <= ENT_RTOL, f"{f}[IENT] (entropy flux: pow on both sides)"
        o.cons_to_prim(out, False)
        gm.cons_to_prim(out, False)
    u, ou = gm.download("u0"), o.get("u0")
    for n in range(5):
        _same(u[n], ou[n], f"u0[{n}]")
    for b in ("b1f0", "b2f0", "b3f0"):
        r = face_region(o, int(b[1]) - 1)
        _same(gm.download(b)[r], o.get(b)[r], b)
    gm.close()


STEP_CASES = [
    ("blast_3d.in", LIMITED_BLAST, 8, False),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12", "mhd.flux_limiter=true", "mhd.pfloor=0.1"], 4, False),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12", "mhd.pfloor=0.1"] + UNIFORM, 4, False),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8", "mhd.reconstruct=ppmx"] + UNIFORM + POINT, 4, False),
    ("orszag_tang.in", ["mesh.nx1=24", "mesh.nx2=24", "mhd.flux_limiter=true", "mhd.dual_energy=true"] + POINT, 6,
     False),
    ("orszag_tang.in", ["mesh.nx1=24", "mesh.nx2=24", "diffusion.nu=0.002"] + UNIFORM, 5, False),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12", "cooling.enabled=true", "cooling.lambda0=0.5",
                     "cooling.alpha=1.0"], 4, True),
    ("orszag_tang.in", ["mesh.nx1=24", "mesh.nx2=24", "cooling.enabled=true", "cooling.lambda0=2.0",
                        "cooling.alpha=0.5", "cooling.tfloor=0.05"] + UNIFORM, 5, True),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12", "cooling.enabled=true", "cooling.lambda0=0.5",
                     "cooling.alpha=1.5", "diffusion.kappa=0.02", "diffusion.sts=true"], 3, True),
    ("cpaw_1d.in", ["mesh.nx1=64", "cooling.enabled=true", "cooling.lambda0=0.2", "cooling.alpha=2.0",
                    "diffusion.nu=0.01"], 10, True),
]


@pytest.mark.parametrize("name,ov,nsteps,cool", STEP_CASES,
                         ids=[f"{c[0]}-{c[1][-1]}-{len(c[1])}" for c in STEP_CASES])
def test_time_loop_with_sources(name, ov, nsteps, cool):
    print(f"case sources loop {name} {ov} steps={nsteps}", flush=True)
    o, gm, kv = _pair(deck(name), ov)
    gm.init(0.0, float(kv.get("time.tlim", 1.0)), -1)
    assert gm.status()["min_dt"] == o.min_dt
    o.step(nsteps)
    gm.step(nsteps)
    st = gm.status()
    assert st["cycle"] == o.cycle == nsteps
    assert st["fofc_total"] == o.fofc_count()
    assert st["limiter_total"] == o.limiter_count
    if ov is LIMITED_BLAST:
        assert o.limiter_count > 0
    u, ou = gm.download("u0"), o.get("u0")
    if not cool:
        assert st["t"] == o.time and st["min_dt"] == o.min_dt
        for n in range(5):
            _same(u[n], ou[n], f"u0[{n}]")
        assert max_rel(u[5], ou[5]) <= ENT_RTOL
        for b in ("b1f0", "b2f0", "b3f0"):
            r = face_region(o, int(b[1]) - 1)
            _same(gm.download(b)[r], o.get(b)[r], b)
    else:
        # vectors are compared on the scale of the whole vector: a component that stays near
        # zero (v_z, B_z of a 2D run) carries only the rounding of the others
        assert abs(st["t"] / o.time - 1) <= COOL_RTOL and abs(st["min_dt"] / o.min_dt - 1) <= COOL_RTOL
        for n in (0, 4):
            _close(u[n], ou[n], COOL_RTOL, f"u0[{n}]")
        _close(u[1:4], ou[1:4], COOL_RTOL, "u0[1:4]")
        gb, ob = [], []
        for b in ("b1f0", "b2f0", "b3f0"):
            r = face_region(o, int(b[1]) - 1)
            gb.append(gm.download(b)[r].ravel())
            ob.append(o.get(b)[r].ravel())
        _close(np.concatenate(gb), np.concatenate(ob), COOL_RTOL, "face B")
    gm.close()


def _device_run(ov, nsteps=None):
    gm, kv = gpu_mesh_for(UNI_DECK, ov)
    o = po.OracleSim(UNI_DECK, list(ov))
    upload_oracle_state(gm, o)
    if nsteps is None:
        gm.init(0.0, 1.0, -1)
        gm.step(100000)   # steps past tlim are no-ops on the device
        assert gm.status()["t"] == 1.0
    else:
        gm.init(0.0, 1.0, nsteps)
        gm.step(nsteps)
    return gm, o


def test_uniform_gravity_momentum_on_the_device():
    """physics.sh:53, 84: a uniform box under uniform gravity gains momentum rho g t V exactly
    (history mom1 = 1, mom3 = 2 to 1e-9)."""
    gm, o = _device_run(["gravity.uniform=true", "gravity.g1=0.5", "gravity.g2=-0.25", "gravity.g3=1.0"])
    _, r = o.geometry()
    u = gm.download("u0")[:, o.ks:o.ke + 1, o.js:o.je + 1, o.is_:o.ie + 1]
    mom = u[1:4].reshape(3, -1).sum(axis=1) * (r[0] * r[1] * r[2])
    assert abs(mom[0] - 1.0) < 1e-9 and abs(mom[1] + 0.5) < 1e-9 and abs(mom[2] - 2.0) < 1e-9, mom
    gm.close()


@pytest.mark.parametrize("alpha", [0.5, 1.0])
def test_power_law_cooling_exact_on_the_device(alpha):
    """physics.sh:54, 86: T(1) = (sqrt(T0) - c/2)^2 for alpha = 1/2 with c = (gamma-1) rho lambda0,
    to 1e-6; alpha = 1 decays like exp(-c t)."""
    gm, o = _device_run(["cooling.enabled=true", "cooling.lambda0=0.1", f"cooling.alpha={alpha}"])
    w = gm.download("w")[:, o.ks:o.ke + 1, o.js:o.je + 1, o.is_:o.ie + 1]
    c = (5.0 / 3.0 - 1) * 2.0 * 0.1
    T = (w[4] / w[0]).mean()
    if alpha == 0.5:
        assert abs(T / (math.sqrt(0.5) - 0.5 * c) ** 2 - 1) < 1e-6
    else:
        assert abs(T / (0.5 * math.exp(-c)) - 1) < 1e-12
    gm.close()


def test_point_mass_pull_is_symmetric_on_the_device():
    """A softened point mass at the box centre pulls a uniform box symmetrically: the total
    momentum stays zero while the gas moves."""
    gm, _ = _device_run(["gravity.point_mass=true", "gravity.gm=0.05", "gravity.rsoft=0.05", "gravity.x0=0.5",
                         "gravity.y0=0.5", "gravity.z0=0.5"], nsteps=10)
    g = gm.download("u0")
    o = po.OracleSim(UNI_DECK, [])
    u = g[:, o.ks:o.ke + 1, o.js:o.je + 1, o.is_:o.ie + 1]
    assert np.max(np.abs(u[1:4].reshape(3, -1).sum(axis=1))) < 1e-12
    assert np.max(np.abs(u[1])) > 1e-4
    gm.close()


def test_sources_need_fp64_single_mesh():
    from chasqui_amd import Batch
    from chasqui_amd.deck import Deck
    with pytest.raises(RuntimeError, match="source terms"):
        Deck(deck("cpaw_1d.in"), ["mesh.nx1=32", "cooling.enabled=true", "mhd.precision=mixed"]).make_mesh()
    for ov in (["mhd.flux_limiter=true"], UNIFORM, ["cooling.enabled=true", "cooling.lambda0=0.1"]):
        m = Deck(deck("cpaw_1d.in"), ["mesh.nx1=32"] + ov).make_mesh()
        b = Batch(2)
        with pytest.raises(RuntimeError, match="source terms"):
            b.add(m)
        b.close()
        m.close()


KERNEL_CASES = [
    ("blast_3d.in", LIMITED_BLAST + UNIFORM),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8", "mhd.flux_limiter=true", "mhd.pfloor=1e-3",
                     "problem.pres_out=1e-3"] + POINT),
    ("orszag_tang.in", ["mesh.nx1=24", "mesh.nx2=20", "mhd.flux_limiter=true", "mhd.dual_energy=true",
                        "cooling.enabled=true", "cooling.lambda0=2.0", "cooling.alpha=0.5"] + UNIFORM),
]


@pytest.mark.parametrize("name,ov", KERNEL_CASES, ids=[f"{c[0]}-{len(c[1])}" for c in KERNEL_CASES])
def test_source_entry_points(name, ov):
    """chq_limit_fluxes / chq_apply_gravity / chq_apply_cooling (positivity::LimitFluxes<DIR>,
    srcterms::ApplyGravity, ApplyCooling) against the oracle's functions on the same state: the
    limited fluxes and the gravity update bitwise with equal activation counts, cooling within
    COOL_RTOL (exp/pow)."""
    print(f"case source entry points {name} {ov}", flush=True)
    o, gm, _ = _pair(deck(name), ov)
    gm.upload("u1", o.get("u1"))
    bdt = 2.0 * o.min_dt
    limited = 0
    for d in range(1 + o.two_d + o.three_d):
        o.fluxes(d, 0)
        gm.fluxes(d, 0)
        n = gm.limit_fluxes(d, bdt)
        assert n == o.limit_fluxes(d, bdt)
        limited += n
        f = f"f{d + 1}"
        a, b = gm.download(f), o.get(f)
        _same(np.delete(a, 5, axis=0), np.delete(b, 5, axis=0), f + " limited")
    if ov[:len(LIMITED_BLAST)] == LIMITED_BLAST:
        assert limited > 0
    o.apply_gravity(1, bdt)
    gm.apply_gravity(1, bdt)
    u, ou = gm.download("u1"), o.get("u1")
    for n in range(5):
        _same(u[n], ou[n], f"u1[{n}] after gravity")
    o.apply_cooling(0, bdt)
    gm.apply_cooling(0, bdt)
    u, ou = gm.download("u0"), o.get("u0")
    for n in range(5):
        _close(u[n], ou[n], COOL_RTOL, f"u0[{n}] after cooling")
    gm.close()
