"""GPU: the mixed-precision path (FP32 primitives, reconstruction, Riemann solve, fluxes and
EMFs; FP64 con2prim, RK combine, flux divergence, CT and conserved state).

There is no FP32 form of the reference method to be bitwise equal to, so the bar is accuracy
against the FP64 path (itself bitwise equal to the oracle):
  * CPAW L1 error (vs the initial state after one period) within 2x of FP64's;
  * div B at round-off: max |div B| <= 5e-13 where the initial state is itself below that
    (1D, 32^3), and everywhere a change of max |div B| * dx / max |B| <= 5e-13 (the
    vector-potential initialisation of the 128^3 oblique wave already holds 2.4e-12 in
    absolute terms, so the dimensionless change is the round-off measure there);
  * total energy conserved to 1e-12 relative, since the
    CT update and the flux divergence stay FP64 and every face flux/EMF enters both
    neighbours with the same value;
  * single-kernel FP32 fluxes within 2e-4 of the FP64 oracle's (relative to each
    component's largest value) on a 3-D blast state with strong shocks.
"""
import numpy as np
import pytest

from helpers import deck, gpu_mesh_for, upload_oracle_state

pytestmark = pytest.mark.gpu

po = pytest.importorskip("pyoracle")

MIXED = ["mhd.precision=mixed"]


def _open(name, ov):
    from chasqui_amd.deck import Deck
    return Deck(deck(name), ov)


def _active(md):
    ng = md.ng
    k = slice(ng, ng + md.nx3) if md.nx3 > 1 else slice(0, 1)
    j = slice(ng, ng + md.nx2) if md.nx2 > 1 else slice(0, 1)
    i = slice(ng, ng + md.nx1)
    return k, j, i


def _cell_b(b1, b2, b3, md):
    """Cell-centred B as the face average (the con2prim formula), active cells only."""
    k, j, i = _active(md)
    i1 = slice(i.start + 1, i.stop + 1)
    j1 = slice(j.start + 1, j.stop + 1) if md.nx2 > 1 else j
    k1 = slice(k.start + 1, k.stop + 1) if md.nx3 > 1 else k
    bx = 0.5 * (b1[k, j, i] + b1[k, j, i1])
    by = 0.5 * (b2[k, j, i] + b2[k, j1, i]) if md.nx2 > 1 else b2[k, j, i]
    bz = 0.5 * (b3[k, j, i] + b3[k1, j, i]) if md.nx3 > 1 else b3[k, j, i]
    return bx, by, bz


def _faces(m):
    return [m.download(f"b{d}f0") for d in (1, 2, 3)]


def _l1_total(m, d):
    """Sum over rho, M1..3, E and B1..3 of the mean |q - q_init| (the README's CPAW error,
    entropy excluded)."""
    md = d.mesh_desc
    k, j, i = _active(md)
    u, ui = m.download("u0"), d.initial("u0")
    err = [np.mean(np.abs(u[n][k, j, i] - ui[n][k, j, i])) for n in range(5)]
    b = _cell_b(*_faces(m), md)
    bi = _cell_b(*[d.initial(f"b{q}f0") for q in (1, 2, 3)], md)
    err += [np.mean(np.abs(x - y)) for x, y in zip(b, bi)]
    return float(np.sum(err))


def _divb(m, d):
    md = d.mesh_desc
    b1, b2, b3 = _faces(m)
    k, j, i = _active(md)
    dx1 = (md.x1max - md.x1min) / md.nx1
    div = (b1[k, j, i.start + 1:i.stop + 1] - b1[k, j, i]) / dx1
    if md.nx2 > 1:
        dx2 = (md.x2max - md.x2min) / md.nx2
        div = div + (b2[k, j.start + 1:j.stop + 1, i] - b2[k, j, i]) / dx2
    if md.nx3 > 1:
        dx3 = (md.x3max - md.x3min) / md.nx3
        div = div + (b3[k.start + 1:k.stop + 1, j, i] - b3[k, j, i]) / dx3
    return div


def _energy(m, d):
    k, j, i = _active(d.mesh_desc)
    return float(np.sum(m.download("u0")[4][k, j, i]))


@pytest.mark.parametrize("name,ov", [
    ("cpaw_1d.in", []),                                                    # N = 128, PPM, T = 1
    ("cpaw_3d_oblique.in", []),                                            # 32^3, WENO-Z, T = 1
    ("cpaw_3d_oblique.in", ["mesh.nx1=128", "mesh.nx2=128", "mesh.nx3=128"]),   # 128^3
])
def test_mixed_cpaw_accuracy(name, ov):
    """L1 error of the mixed path within 2x of FP64 after one wave period, div B at
    round-off, total energy conserved to round-off."""
    print(f"case {name} {ov}", flush=True)
    res = {}
    for prec in ("fp64", "mixed"):
        d = _open(name, ov + [f"mhd.precision={prec}"])
        m = d.make_mesh()
        e0 = _energy(m, d)
        div0 = _divb(m, d)
        st = m.run(chunk=256)
        assert st["t"] == 1.0
        div1 = _divb(m, d)
        md = d.mesh_desc
        dx = (md.x1max - md.x1min) / md.nx1
        bmax = max(float(np.max(np.abs(b))) for b in _faces(m))
        res[prec] = dict(l1=_l1_total(m, d), div_end=float(np.max(np.abs(div1))),
                         div_change_rel=float(np.max(np.abs(div1 - div0))) * dx / bmax,
                         div0=float(np.max(np.abs(div0))), drift=abs(_energy(m, d) - e0) / abs(e0),
                         cycle=st["cycle"])
        m.close()
        d.close()
    print(res, flush=True)
    assert res["mixed"]["l1"] <= 2.0 * res["fp64"]["l1"], res
    if name == "cpaw_1d.in":
        assert f"{res['fp64']['l1']:.2e}" == "6.65e-04"
    for prec in ("fp64", "mixed"):
        assert res[prec]["div_change_rel"] <= 5e-13, (prec, res[prec])
        if res[prec]["div0"] <= 5e-14:
            assert res[prec]["div_end"] <= 5e-13, (prec, res[prec])
        assert res[prec]["drift"] <= 1e-12, (prec, res[prec])


@pytest.mark.parametrize("name,ov,steps", [
    ("blast_3d.in", ["mesh.nx1=32", "mesh.nx2=32", "mesh.nx3=32"], 10),
    ("orszag_tang.in", ["mesh.nx1=64", "mesh.nx2=64"], 40),
    ("brio_wu.in", ["mesh.nx1=400"], 80),
])
def test_mixed_close_to_fp64(name, ov, steps):
    """Shocked problems: the mixed state stays within 1e-3 (relative L1) of FP64's, with
    positive density and pressure and div B unchanged to round-off."""
    print(f"case {name} {ov} steps={steps}", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        d = _open(name, ov + [f"mhd.precision={prec}"])
        m = d.make_mesh()
        div0 = _divb(m, d)
        m.step(steps)
        md = d.mesh_desc
        bmax = max(float(np.max(np.abs(b))) for b in _faces(m))
        div_change = float(np.max(np.abs(_divb(m, d) - div0))) * (md.x1max - md.x1min) / md.nx1 / bmax
        out[prec] = (m.download("u0"), m.download("w"), m.status(), div_change, _active(d.mesh_desc))
        m.close()
        d.close()
    u64, w64, s64, _, (k, j, i) = out["fp64"]
    u32, w32, s32, div32, _ = out["mixed"]
    assert s32["cycle"] == s64["cycle"] == steps
    assert abs(s32["t"] - s64["t"]) <= 1e-4 * s64["t"]
    for n in range(5):
        a, b = u32[n][k, j, i], u64[n][k, j, i]
        rel = np.sum(np.abs(a - b)) / max(np.sum(np.abs(b)), 1e-300)
        assert rel <= 1e-3, (n, rel)
    assert np.min(w32[0][k, j, i]) > 0 and np.min(w32[4][k, j, i]) > 0
    assert div32 <= 5e-13


def test_mixed_single_kernel_fluxes():
    """FP32 fluxes and face EMFs from the oracle's (FP64) primitives, against the oracle."""
    ov = ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16"]
    text = deck("blast_3d.in")
    o = po.OracleSim(text, ov)
    o.step(2)                      # a state with developed shocks
    gm, _ = gpu_mesh_for(text, ov + MIXED)
    upload_oracle_state(gm, o)
    worst = 0.0
    for d in range(3):
        o.fluxes(d, 0)
        gm.fluxes(d, 0)
        f = f"f{d + 1}"
        a, b = gm.download(f), o.get(f)
        for n in range(5):
            scale = np.max(np.abs(b[n])) + 1e-30
            worst = max(worst, np.max(np.abs(a[n] - b[n])) / scale)
    for e in ("e2x1", "e3x1", "e1x2", "e3x2", "e1x3", "e2x3"):
        a, b = gm.download(e), o.get(e)
        worst = max(worst, np.max(np.abs(a - b)) / (np.max(np.abs(b)) + 1e-30))
    print("worst relative flux/EMF difference", worst, flush=True)
    assert worst <= 2e-4
    gm.close()


def test_mixed_rejects_reference_order_entries():
    """chq_stage / fofc_flag / fofc_repair are FP64-only: they refuse a mixed context."""
    gm, _ = gpu_mesh_for(deck("cpaw_1d.in"), ["mesh.nx1=32"] + MIXED)
    with pytest.raises(RuntimeError, match="FP64 contexts only"):
        gm.stage(0, 1, 1.0, 0.0, 1e-3)
    with pytest.raises(RuntimeError, match="FP64 contexts only"):
        gm.flag_bad_cells(0)
    gm.close()


def _mixed_cases(seed=31, n=16):
    import random
    rng = random.Random(seed)
    decks = ["blast_3d.in", "orszag_tang.in", "brio_wu.in", "cpaw_1d.in", "field_loop.in", "cpaw_3d_oblique.in"]
    out = []
    for _ in range(n):
        name = rng.choice(decks)
        dims = 3 if name in ("blast_3d.in", "cpaw_3d_oblique.in") else rng.choice([1, 2])
        recon = rng.choice(["dc", "plm", "ppm", "ppmx", "wenoz"])
        nx = [rng.randint(6, 24) for _ in range(3)]
        ov = [f"mesh.nx1={nx[0]}", f"mesh.nx2={nx[1] if dims > 1 else 1}", f"mesh.nx3={nx[2] if dims > 2 else 1}",
              f"mhd.reconstruct={recon}", f"mhd.rsolver={rng.choice(['hlle', 'hlld', 'lhlld'])}",
              f"time.integrator={rng.choice(['rk1', 'rk2', 'rk3', 'rk4'])}", "time.nlim=-1", "time.tlim=1e9",
              f"time.cfl={rng.choice([0.2, 0.3, 0.4])}", f"mhd.fofc={rng.choice(['true', 'false'])}",
              f"mhd.flatten={rng.choice(['true', 'false'])}"]
        for ax in range(dims):
            ov.append(f"mesh.bc_x{ax + 1}={rng.choice(['periodic', 'outflow', 'reflect'])}")
        if rng.random() < 0.3:
            ov.append("mhd.low_mach=true")
        out.append((name, ov, rng.randint(3, 8)))
    return out


MIXED_CASES = _mixed_cases()


@pytest.mark.parametrize("name,ov,steps", MIXED_CASES,
                         ids=[f"{i:02d}-{c[0].split('.')[0]}" for i, c in enumerate(MIXED_CASES)])
def test_mixed_seeded_options_close_to_fp64(name, ov, steps):
    """Seeded problems, meshes, schemes, boundaries and switches (fofc, flattening, low-Mach
    blend; the mixed path takes neither dual energy nor characteristic variables) in mixed
    precision against the same run in FP64:
    the same number of cycles, t within 1e-4, each conserved field within 1e-3 relative L1
    (the scale of a field that stays at zero by symmetry is 1e-6 of the total energy), positive
    density and pressure, div B unchanged to round-off."""
    print(f"case mixed {name} {ov} steps={steps}", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        d = _open(name, ov + [f"mhd.precision={prec}"])
        m = d.make_mesh()
        div0 = _divb(m, d)
        m.step(steps)
        md = d.mesh_desc
        bmax = max(float(np.max(np.abs(b))) for b in _faces(m))
        div_change = float(np.max(np.abs(_divb(m, d) - div0))) * (md.x1max - md.x1min) / md.nx1 / max(bmax, 1e-300)
        out[prec] = (m.download("u0"), m.download("w"), m.status(), div_change, _active(md))
        m.close()
        d.close()
    u64, w64, s64, _, (k, j, i) = out["fp64"]
    u32, w32, s32, div32, _ = out["mixed"]
    assert s32["cycle"] == s64["cycle"] == steps
    assert abs(s32["t"] - s64["t"]) <= 1e-4 * s64["t"]
    e_scale = float(np.sum(np.abs(u64[4][k, j, i])))
    for n in range(5):
        a, b = u32[n][k, j, i], u64[n][k, j, i]
        rel = np.sum(np.abs(a - b)) / max(float(np.sum(np.abs(b))), 1e-6 * e_scale)
        assert rel <= 1e-3, (n, rel)
    assert np.min(w32[0][k, j, i]) > 0 and np.min(w32[4][k, j, i]) > 0
    assert div32 <= 5e-13
