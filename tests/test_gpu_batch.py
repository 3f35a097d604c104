"""GPU: batches of uneven runs against the same runs stepped alone.

A batch advances every run with shared launches (device work lists, one launch per kernel
class and stage, hipGraph replay); no reduction spans runs, so each run must be bitwise what
`Mesh.step` gives it alone (and the solo path is pinned bitwise to the oracle elsewhere).
Runs differ in dimensionality, size, boundaries, CFL, tlim/nlim (so they finish at different
steps), and one of them has fofc firing every stage.
"""
import numpy as np
import pytest

from helpers import deck

pytestmark = pytest.mark.gpu

SCHEME = ["mhd.reconstruct=ppm", "mhd.rsolver=hlld", "time.integrator=rk3"]

RUNS = [
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16"]),
    ("blast_3d.in", ["mesh.nx1=20", "mesh.nx2=12", "mesh.nx3=8", "mesh.bc_x1=reflect", "mesh.bc_x2=outflow",
                     "time.nlim=5"]),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=12", "mesh.nx3=12", "mhd.pfloor=0.1"]),          # fofc fires
    ("orszag_tang.in", ["mesh.nx1=40", "mesh.nx2=24", "time.cfl=0.3"]),
    ("orszag_tang.in", ["mesh.nx1=32", "mesh.nx2=32", "time.tlim=0.01"]),                     # finishes early
    ("brio_wu.in", ["mesh.nx1=300"]),
    ("cpaw_1d.in", ["mesh.nx1=64", "time.tlim=0.05"]),
]


def _mesh(name, ov, prec):
    from chasqui_amd.deck import Deck
    d = Deck(deck(name), SCHEME + ov + [f"mhd.precision={prec}"])
    m = d.make_mesh()
    d.close()
    return m


def _state(m):
    out = {n: m.download(n) for n in ("u0", "b1f0", "b2f0", "b3f0", "w", "bcc", "ecc")}
    st = m.status()
    out["status"] = (st["t"], st["cycle"], st["min_dt"], st["fofc_total"], st["dt"])
    return out


def _same(a, b, label):
    for k in a:
        if k == "status":
            assert a[k] == b[k], f"{label}: status {a[k]} != {b[k]}"
        elif k == "u0":   # the entropy slot is rebuilt by pow() from rho and p in both
            assert np.array_equal(a[k][:5], b[k][:5]), f"{label}: {k}"
            assert np.array_equal(a[k][5], b[k][5]), f"{label}: u0[IENT]"
        else:
            assert np.array_equal(a[k], b[k]), f"{label}: {k}"


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_solo(prec, graph):
    from chasqui_amd import Batch
    steps = 12
    print(f"case batch prec={prec} graph={graph}", flush=True)
    solo = []
    for name, ov in RUNS:
        m = _mesh(name, ov, prec)
        m.step(steps)
        solo.append(_state(m))
        m.close()
    b = Batch(16)
    b.use_graph(graph)
    meshes = [_mesh(name, ov, prec) for name, ov in RUNS]
    for m in meshes:
        b.add(m)
    b.step(steps)
    b.sync()
    for i, m in enumerate(meshes):
        _same(_state(m), solo[i], f"run {i} {RUNS[i]}")
    assert solo[2]["status"][3] > 0, "the fofc case must flag cells"
    cnt, flags = b.unfinished()
    assert flags[4] == 0 and flags[6] == 0 and flags[1] == 0
    b.close()
    for m in meshes:
        m.close()


def test_batch_membership_changes():
    """Runs join and leave between replays of the captured step; the others are unaffected."""
    from chasqui_amd import Batch
    print("case membership", flush=True)
    A, B, Cc = RUNS[0], RUNS[3], RUNS[5]
    ref = {}
    for key, (name, ov), n in (("a", A, 9), ("b", B, 5), ("c", Cc, 6)):
        m = _mesh(name, ov, "fp64")
        m.step(n)
        ref[key] = _state(m)
        m.close()
    b = Batch(4)
    ma, mb, mc = _mesh(*A, "fp64"), _mesh(*B, "fp64"), _mesh(*Cc, "fp64")
    b.add(ma)
    b.add(mb)
    b.step(3)
    b.add(mc)        # joins after 3 steps
    b.step(2)
    b.remove(mb)     # leaves after 5 steps
    b.step(4)        # a: 9 steps, c: 6 steps
    b.sync()
    _same(_state(ma), ref["a"], "a")
    _same(_state(mb), ref["b"], "b")
    _same(_state(mc), ref["c"], "c")
    # a context that left a batch steps on its own again
    mb.step(1)
    b.close()
    for m in (ma, mb, mc):
        m.close()


def test_batch_direction_classes_appear_and_vanish():
    """The batched step leaves out the flux and boundary launches of directions no run has
    (1D and 2D batches); a run that brings a new direction re-captures the graph with them, and
    one that takes the last run of a direction away captures it without them."""
    from chasqui_amd import Batch
    print("case direction classes", flush=True)
    one, two, three = RUNS[5], RUNS[3], RUNS[0]
    ref = {}
    for key, (name, ov), n in (("1", one, 10), ("2", two, 6), ("3", three, 3)):
        m = _mesh(name, ov, "fp64")
        m.step(n)
        ref[key] = _state(m)
        m.close()
    b = Batch(4)
    m1, m2, m3 = _mesh(*one, "fp64"), _mesh(*two, "fp64"), _mesh(*three, "fp64")
    b.add(m1)
    b.add(m2)
    b.step(2)        # x and y classes only
    b.add(m3)        # z appears: re-captured
    b.step(3)
    b.remove(m3)     # 3 steps for the 3D run
    b.step(1)
    b.remove(m2)     # 6 steps for the 2D run; y disappears
    b.step(4)        # 10 steps for the 1D run
    b.sync()
    _same(_state(m1), ref["1"], "1D")
    _same(_state(m2), ref["2"], "2D")
    _same(_state(m3), ref["3"], "3D")
    b.close()
    for m in (m1, m2, m3):
        m.close()


SMALL_RUNS = [
    ("blast_3d.in", ["mesh.nx1=10", "mesh.nx2=10", "mesh.nx3=10", "mhd.pfloor=0.0999"]),   # fofc fires
    ("blast_3d.in", ["mesh.nx1=8", "mesh.nx2=10", "mesh.nx3=6"]),
    ("orszag_tang.in", ["mesh.nx1=24", "mesh.nx2=20"]),
    ("brio_wu.in", ["mesh.nx1=200", "mhd.dfloor=0.2"]),                                      # fofc fires
    ("cpaw_1d.in", ["mesh.nx1=64"]),
]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_small_runs_fofc_equals_solo(prec):
    """A batch whose runs are all small takes the fofc branch with one block per run (two
    launches per stage instead of one per kernel class); two of its runs flag cells in every
    stage. Each run is bitwise its solo run."""
    from chasqui_amd import Batch
    print(f"case small runs {prec}", flush=True)
    steps = 6
    ref = []
    for name, ov in SMALL_RUNS:
        m = _mesh(name, ov, prec)
        m.step(steps)
        ref.append(_state(m))
        m.close()
    assert ref[0]["status"][3] > 0 and ref[3]["status"][3] > 0   # fofc fired in the solo runs
    b = Batch(len(SMALL_RUNS))
    meshes = [_mesh(name, ov, prec) for name, ov in SMALL_RUNS]
    for m in meshes:
        b.add(m)
    b.step(steps)
    b.sync()
    for q, m in enumerate(meshes):
        _same(_state(m), ref[q], f"small run {q}")
    b.close()
    for m in meshes:
        m.close()


def test_batch_autotune_keeps_results():
    """The autotuner times kernel classes on scratch arrays only: tuning between steps must not
    change any run's trajectory."""
    from chasqui_amd import Batch
    print("case autotune", flush=True)
    runs = RUNS[:4]
    solo = []
    for name, ov in runs:
        m = _mesh(name, ov, "fp64")
        m.step(6)
        solo.append(_state(m))
        m.close()
    b = Batch(8)
    meshes = [_mesh(name, ov, "fp64") for name, ov in runs]
    for m in meshes:
        b.add(m)
    b.step(2)
    best = b.tune(reps=2)
    assert best > 0
    grids = b.grids()
    print("tuned grids", grids, flush=True)
    b.step(4)
    b.sync()
    for i, m in enumerate(meshes):
        _same(_state(m), solo[i], f"run {i}")
    b.close()
    for m in meshes:
        m.close()


def test_batch_rejects_mixed_schemes():
    from chasqui_amd import Batch
    b = Batch(2)
    m1 = _mesh("cpaw_1d.in", ["mesh.nx1=32"], "fp64")
    from chasqui_amd.deck import Deck
    d = Deck(deck("cpaw_1d.in"), ["mesh.nx1=32", "mhd.reconstruct=wenoz"])
    m2 = d.make_mesh()
    b.add(m1)
    with pytest.raises(RuntimeError, match="one scheme"):
        b.add(m2)
    with pytest.raises(RuntimeError, match="advanced by a batch"):
        m1.step(1)
    b.close()
    m1.close()
    m2.close()


def _batch_cases(seed=7, n=8):
    """Seeded schemes (shared by the batch) and per-run options (kept per run)."""
    import random
    rng = random.Random(seed)
    decks = ["blast_3d.in", "orszag_tang.in", "brio_wu.in", "cpaw_1d.in", "field_loop.in"]
    out = []
    for _ in range(n):
        recon = rng.choice(["dc", "plm", "ppm", "ppmx", "wenoz"])
        scheme = [f"mhd.reconstruct={recon}", f"mhd.rsolver={rng.choice(['hlle', 'hlld', 'lhlld'])}",
                  f"time.integrator={rng.choice(['rk1', 'rk2', 'rk3', 'rk4'])}"]
        prec = rng.choice(["fp64", "fp64", "mixed"])
        runs = []
        for _ in range(rng.randint(2, 5)):
            name = rng.choice(decks)
            dims = 3 if name == "blast_3d.in" else rng.choice([1, 2])
            nx = [rng.randint(4, 18) for _ in range(3)]
            ov = [f"mesh.nx1={nx[0]}", f"mesh.nx2={nx[1] if dims > 1 else 1}", f"mesh.nx3={nx[2] if dims > 2 else 1}",
                  f"time.cfl={rng.choice([0.2, 0.3, 0.4])}", f"mhd.fofc={rng.choice(['true', 'false'])}",
                  f"mhd.flatten={rng.choice(['true', 'false'])}"]
            for ax in range(dims):
                ov.append(f"mesh.bc_x{ax + 1}={rng.choice(['periodic', 'outflow', 'reflect'])}")
            if rng.random() < 0.3:
                ov.append("mhd.low_mach=true")
            if rng.random() < 0.3:
                ov.append("mhd.pfloor=0.11")   # fofc fires where it is on
            if rng.random() < 0.3:
                ov.append(f"time.nlim={rng.randint(1, 5)}")   # finishes before the others
            runs.append((name, ov))
        out.append((scheme, prec, runs, rng.randint(3, 8)))
    return out


BATCH_CASES = _batch_cases()


@pytest.mark.parametrize("scheme,prec,runs,steps", BATCH_CASES,
                         ids=[f"{i}-{c[1]}-{c[0][0].split('=')[1]}" for i, c in enumerate(BATCH_CASES)])
def test_batch_seeded_schemes_equal_solo(scheme, prec, runs, steps):
    """Every scheme the batch accepts, with runs that differ in problem, dimensionality, size,
    boundaries, CFL, fofc, flattening, the low-Mach blend, a pressure floor and their step limit:
    each batched run ends bitwise where it ends stepped alone."""
    from chasqui_amd import Batch
    from chasqui_amd.deck import Deck
    print(f"case batch {scheme} {prec} {runs} steps={steps}", flush=True)

    def mk(name, ov):
        d = Deck(deck(name), scheme + ["time.tlim=1e9", "time.nlim=-1"] + ov + [f"mhd.precision={prec}"])
        m = d.make_mesh()
        d.close()
        return m

    solo = []
    for name, ov in runs:
        m = mk(name, ov)
        m.step(steps)
        solo.append(_state(m))
        m.close()
    b = Batch(8)
    meshes = [mk(name, ov) for name, ov in runs]
    for m in meshes:
        b.add(m)
    b.step(steps)
    b.sync()
    for i, m in enumerate(meshes):
        nlim = [int(o.split("=")[1]) for o in runs[i][1] if o.startswith("time.nlim=")]
        assert solo[i]["status"][1] == min([steps] + nlim), f"run {i}: cycle {solo[i]['status'][1]}"
        _same(_state(m), solo[i], f"run {i} {runs[i]}")
    b.close()
    for m in meshes:
        m.close()


def _ensemble_runs(n, seed, sizes):
    """Orszag-Tang ensemble members: size from `sizes`, per-run seeded +-5% perturbation of the
    pressure and of the field amplitude (problem.pres / problem.b0 around pgen.hpp:94-111's
    5/(12 pi) and 1/sqrt(4 pi))."""
    rng = np.random.default_rng(seed)
    p0, b0 = 5.0 / (12.0 * np.pi), 1.0 / np.sqrt(4.0 * np.pi)
    out = []
    for _ in range(n):
        s = int(rng.choice(sizes))
        fp, fb = 1.0 + 0.05 * rng.uniform(-1, 1), 1.0 + 0.05 * rng.uniform(-1, 1)
        out.append([f"mesh.nx1={s}", f"mesh.nx2={s}", f"problem.pres={float(p0 * fp)!r}",
                    f"problem.b0={float(b0 * fb)!r}"])
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_orszag_tang_ensemble_config(prec):
    """The batched-ensemble workload's own scheme, PPMX + HLLD + RK3 on Orszag-Tang with sizes
    from its set {64, 96, 128}^2 (the smaller end, so the solo runs finish in seconds) and the
    per-run +-5% pressure / field amplitude: every member bitwise its solo run, and the
    perturbations really differ between members."""
    from chasqui_amd import Batch
    from chasqui_amd.deck import Deck
    scheme = ["mhd.reconstruct=ppmx", "mhd.rsolver=hlld", "time.integrator=rk3", "time.tlim=1e9", "time.nlim=-1"]
    runs = _ensemble_runs(6, 3, [64, 96, 128])
    steps = 8
    print(f"case OT ensemble {prec} {runs}", flush=True)

    def mk(ov):
        d = Deck(deck("orszag_tang.in"), scheme + ov + [f"mhd.precision={prec}"])
        m = d.make_mesh()
        d.close()
        return m

    solo = []
    for ov in runs:
        m = mk(ov)
        m.step(steps)
        solo.append(_state(m))
        m.close()
    assert len({r[2] for r in runs}) == len(runs) and len({r[3] for r in runs}) == len(runs)
    assert len({int(r[0].split("=")[1]) for r in runs}) >= 2
    b = Batch(8)
    meshes = [mk(ov) for ov in runs]
    for m in meshes:
        b.add(m)
    b.step(2)
    b.tune(reps=1)
    b.step(steps - 2)
    b.sync()
    for i, m in enumerate(meshes):
        assert solo[i]["status"][1] == steps
        _same(_state(m), solo[i], f"member {i} {runs[i]}")
    b.close()
    for m in meshes:
        m.close()
