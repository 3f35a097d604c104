"""GPU: explicit diffusion (viscosity, conduction, Ohmic resistivity) and RKL2 super-time-stepping
against the CPU oracle, bitwise; and the reference's own decay tests (tests/physics.sh:30-36,
72-81) run on the device.

Kernel level: diffusion::Fluxes (diffusion.hpp:131-138) added to the hyperbolic fluxes,
ResistiveEMF (141-156) added to the corner EMFs, TimeStep (159-174). Time loop: the stage with
diffusion and its fofc repair (main.cpp:55-91, fofc.hpp:90-93), the parabolic dt
(main.cpp:202, 275), and the operator-split STS step (sts.hpp:111-157, main.cpp:278-290).
"""
import numpy as np
import pytest

from diffusion_cases import DECAY_CASES, DIFF_DECK, decay_amplitude, exact_amplitude
from helpers import deck, face_region, gpu_mesh_for, max_rel, n_diff, upload_oracle_state

pytestmark = pytest.mark.gpu

po = pytest.importorskip("pyoracle")

ENT_RTOL = 1e-14
ALL3 = ["diffusion.nu=0.01", "diffusion.kappa=0.02", "diffusion.eta=0.015"]


def _pair(text, ov):
    o = po.OracleSim(text, list(ov))
    gm, kv = gpu_mesh_for(text, ov)
    upload_oracle_state(gm, o)
    return o, gm, kv


def _same(gpu, ora, label):
    assert gpu.shape == ora.shape, label
    assert n_diff(gpu, ora) == 0, f"{label}: {n_diff(gpu, ora)} entries differ, max rel {max_rel(gpu, ora):.3e}"


KERNEL_CASES = [
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8"] + ALL3),
    ("orszag_tang.in", ["mesh.nx1=24", "mesh.nx2=20"] + ALL3),
    ("brio_wu.in", ["mesh.nx1=64"] + ALL3),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8", "diffusion.eta=0.03", "mesh.bc_x1=reflect",
                     "mesh.bc_x3=outflow"]),
]


@pytest.mark.parametrize("name,ov", KERNEL_CASES, ids=[f"{c[0]}-{len(c[1])}" for c in KERNEL_CASES])
def test_diffusion_kernels_bitwise(name, ov):
    print(f"case diffusion kernels {name} {ov}", flush=True)
    o, gm, _ = _pair(deck(name), ov)
    assert gm.diffusion_dt() == o.diffusion_dt()
    ndir = 1 + o.two_d + o.three_d
    for d in range(ndir):
        o.fluxes(d, 0)
        gm.fluxes(d, 0)
        o.diffusion_fluxes(d, 0)
        gm.diffusion_fluxes(d, 0)
        f = f"f{d + 1}"
        _same(gm.download(f), o.get(f), f + " with diffusion")
    o.corner_emf()
    gm.corner_emf()
    o.resistive_emf(0)
    gm.resistive_emf(0)
    for e in ("e1", "e2", "e3"):
        _same(gm.download(e), o.get(e), e + " with eta J")
    gm.close()


STEP_CASES = [
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12"] + ALL3, 4),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12", "mhd.pfloor=0.1"] + ALL3, 3),      # fofc fires
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12", "diffusion.sts=true"] + ALL3, 4),
    ("orszag_tang.in", ["mesh.nx1=32", "mesh.nx2=24", "diffusion.eta=0.05", "diffusion.sts=true"], 6),
    ("orszag_tang.in", ["mesh.nx1=24", "mesh.nx2=24", "diffusion.nu=0.002", "mhd.dual_energy=true"], 8),
    ("cpaw_1d.in", ["mesh.nx1=64", "diffusion.nu=0.01", "diffusion.eta=0.01"], 10),
]


@pytest.mark.parametrize("name,ov,nsteps", STEP_CASES, ids=[f"{c[0]}-{c[1][-1]}-{len(c[1])}" for c in STEP_CASES])
def test_diffusion_time_loop_bitwise(name, ov, nsteps):
    print(f"case diffusion loop {name} {ov} steps={nsteps}", flush=True)
    o, gm, kv = _pair(deck(name), ov)
    gm.init(0.0, float(kv.get("time.tlim", 1.0)), -1)
    assert gm.status()["min_dt"] == o.min_dt
    o.step(nsteps)
    gm.step(nsteps)
    st = gm.status()
    assert st["cycle"] == o.cycle == nsteps
    assert st["t"] == o.time and st["min_dt"] == o.min_dt
    assert st["fofc_total"] == o.fofc_count()
    u, ou = gm.download("u0"), o.get("u0")
    for n in range(5):
        _same(u[n], ou[n], f"u0[{n}]")
    assert max_rel(u[5], ou[5]) <= ENT_RTOL
    for b in ("b1f0", "b2f0", "b3f0"):
        r = face_region(o, int(b[1]) - 1)
        _same(gm.download(b)[r], o.get(b)[r], b)
    _same(gm.download("bcc"), o.get("bcc"), "bcc")
    gm.close()


@pytest.mark.parametrize("sts", [False, True], ids=["explicit", "sts"])
@pytest.mark.parametrize("case", DECAY_CASES, ids=[c[0] for c in DECAY_CASES])
def test_decay_on_the_device(case, sts):
    """physics.sh:30-36, 72-81 on the device time loop to tlim; and the final state is bitwise
    the oracle's."""
    label, ov, var, axis, D, t, tol = case
    ov = ov + (["diffusion.sts=true"] if sts else [])
    print(f"case decay {label} sts={sts}", flush=True)
    o, gm, kv = _pair(DIFF_DECK, ov)
    gm.init(0.0, t, -1)
    o.run()
    gm.step(o.cycle + 5)   # steps past tlim are no-ops on the device
    st = gm.status()
    assert st["cycle"] == o.cycle and st["t"] == o.time
    w, bcc = gm.download("w"), gm.download("bcc")
    region = (slice(o.ks, o.ke + 1), slice(o.js, o.je + 1), slice(o.is_, o.ie + 1))
    A = decay_amplitude(w, bcc, region, var, axis)
    assert abs(A / exact_amplitude(D, t) - 1) < tol, (label, A)
    _same(gm.download("bcc"), o.get("bcc"), "bcc")
    u, ou = gm.download("u0"), o.get("u0")
    for n in range(5):
        _same(u[n], ou[n], f"u0[{n}]")
    gm.close()


def test_diffusion_needs_fp64_single_mesh():
    from chasqui_amd import Batch
    from chasqui_amd.deck import Deck
    with pytest.raises(RuntimeError, match="diffusion"):
        Deck(deck("cpaw_1d.in"), ["mesh.nx1=32", "diffusion.nu=0.01", "mhd.precision=mixed"]).make_mesh()
    d = Deck(deck("cpaw_1d.in"), ["mesh.nx1=32", "diffusion.eta=0.01"])
    m = d.make_mesh()
    b = Batch(2)
    with pytest.raises(RuntimeError, match="diffusion"):
        b.add(m)
    b.close()
    m.close()


def test_stiff_diffusion_step_counts_on_the_device():
    """README.md:126-127 on the device loop: eta = 0.5 on the 64-cell deck to t = 0.05 takes 683
    explicit steps, or 14 RKL2 super-time-steps with 111 stages in total."""
    ov = ["mesh.nx1=64", "problem.var=b", "problem.axis=1", "diffusion.eta=0.5"]
    for sts, steps, stages in ((False, 683, 0), (True, 14, 111)):
        o, gm, _ = _pair(DIFF_DECK, ov + (["diffusion.sts=true"] if sts else []))
        gm.init(0.0, 0.05, -1)
        gm.step(1000)
        st = gm.status()
        assert st["cycle"] == steps and st["t"] == 0.05 and st["sts_stages"] == stages, st
        gm.close()


def test_sts_scratch_starts_clear_after_other_contexts():
    """Regression: the RKL2 face combination covers [is..ie+1] x [js..je+1] x [ks..ke+1] on every
    face component, beyond the entries where the resistive operator writes its curl terms. A 3D
    RKL2 run fills and frees its scratch; a 1D run created after it (whose scratch may reuse that
    memory) must still match the oracle bitwise, b2 at i = ie+1 included."""
    from chasqui_amd.deck import Deck
    sts = ["diffusion.eta=0.002", "diffusion.sts=true", "time.nlim=-1", "time.tlim=1e9"]
    print("case RKL2 3D blast 12^3, then RKL2 1D Brio-Wu 8 cells", flush=True)
    d = Deck(deck("blast_3d.in"), sts + ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12"])
    m = d.make_mesh()
    d.close()
    m.step(3)
    assert m.status()["sts_stages"] > 0
    m.close()
    ov = sts + ["mesh.nx1=8", "mesh.nx2=1", "mesh.nx3=1", "mhd.reconstruct=dc", "mhd.rsolver=hlld",
                "time.integrator=rk3", "time.cfl=0.2", "mesh.bc_x1=outflow", "mhd.pfloor=0.11"]
    d = Deck(deck("brio_wu.in"), ov)
    m = d.make_mesh()
    d.close()
    m.step(5)
    o = po.OracleSim(deck("brio_wu.in"), ov)
    o.step(5)
    assert m.status()["sts_stages"] == o.sts_count > 0
    u, ou = m.download("u0"), o.get("u0")
    for q in range(5):
        assert n_diff(u[q], ou[q]) == 0, f"u0[{q}]"
    for dd in range(3):
        b = f"b{dd + 1}f0"
        r = face_region(o, dd)
        assert n_diff(m.download(b)[r], o.get(b)[r]) == 0, b
    m.close()
