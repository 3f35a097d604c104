"""GPU: the continuous-batching server. Runs of mixed dimension, size, scheme and precision
arrive while others are on the device, are admitted against an HBM budget, share batched
launches, and leave when done. Each result must be bitwise the one the same deck gives when run
alone (no reduction spans runs; the solo path is pinned to the oracle elsewhere), and the
budget must hold at every iteration."""
import numpy as np
import pytest

from helpers import deck

pytestmark = pytest.mark.gpu

JOBS = [   # (deck, overrides, precision)
    ("brio_wu.in", ["mesh.nx1=256", "time.tlim=0.02"], "mixed"),
    ("orszag_tang.in", ["mesh.nx1=48", "mesh.nx2=48", "time.nlim=12", "mhd.reconstruct=ppm"], "mixed"),
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16", "time.nlim=6"], "mixed"),
    ("brio_wu.in", ["mesh.nx1=400", "time.tlim=0.015", "mhd.reconstruct=wenoz"], "mixed"),
    ("orszag_tang.in", ["mesh.nx1=32", "mesh.nx2=64", "time.nlim=9", "mhd.reconstruct=wenoz"], "fp64"),
    ("blast_3d.in", ["mesh.nx1=12", "mesh.nx2=20", "mesh.nx3=16", "time.nlim=7", "mhd.pfloor=0.1"], "fp64"),
    ("orszag_tang.in", ["mesh.nx1=40", "mesh.nx2=40", "time.nlim=15", "mhd.reconstruct=ppm"], "mixed"),
    ("brio_wu.in", ["mesh.nx1=300", "time.tlim=0.02"], "fp64"),
    ("blast_3d.in", ["mesh.nx1=20", "mesh.nx2=20", "mesh.nx3=20", "time.nlim=5", "mhd.reconstruct=wenoz"], "mixed"),
    ("orszag_tang.in", ["mesh.nx1=64", "mesh.nx2=64", "time.nlim=10", "mhd.reconstruct=ppm"], "mixed"),
]

ARRAYS = ("u0", "b1f0", "b2f0", "b3f0", "w", "bcc")


def _result(m):
    out = {n: m.download(n) for n in ARRAYS}
    st = m.status()
    out["status"] = (st["t"], st["cycle"], st["min_dt"], st["fofc_total"])
    return out


def _solo(name, ov, prec):
    from chasqui_amd.deck import Deck
    d = Deck(deck(name), ov + [f"mhd.precision={prec}"])
    m = d.make_mesh()
    m.run(chunk=8)
    r = _result(m)
    m.close()
    d.close()
    return r


def test_served_runs_equal_solo_runs():
    from chasqui_amd.serve import Server
    budget = 16 * 2 ** 20           # 0.4-12 MB runs: a few fit at a time, so the queue must build
    srv = Server(budget, slots=3, chunk=3)
    results, ids, order = {}, {}, list(range(len(JOBS)))
    saw_overlap = False
    nxt = 0
    while nxt < len(order) or len(results) < len(JOBS):
        # two arrivals per iteration while the earlier ones are running
        for _ in range(2):
            if nxt < len(order):
                name, ov, prec = JOBS[nxt]
                ids[srv.submit(deck(name), ov, prec)] = nxt
                nxt += 1
        done = srv.pump()
        st = srv.stats()
        assert st["bytes_in_use"] <= budget
        if st["queued"] and st["running"]:
            saw_overlap = True
        for jid in done:
            info = srv.job(jid)
            assert info["state"] == "done", info
            results[ids[jid]] = _result(srv.mesh(jid))
            srv.release(jid)
    assert saw_overlap, "runs must arrive and wait while others are on the device"
    srv.close()
    for i, (name, ov, prec) in enumerate(JOBS):
        print(f"case job {i} {name} {ov} {prec}", flush=True)
        ref = _solo(name, ov, prec)
        got = results[i]
        assert got["status"] == ref["status"], (i, got["status"], ref["status"])
        for a in ARRAYS:
            assert np.array_equal(got[a], ref[a]), (i, a)


def test_server_rejects_oversized_and_bad_decks():
    from chasqui_amd.serve import Server
    srv = Server(8 * 2 ** 20, slots=2, chunk=1)
    with pytest.raises(RuntimeError, match="larger than the HBM budget"):
        srv.submit(deck("blast_3d.in"), ["mesh.nx1=64", "mesh.nx2=64", "mesh.nx3=64"])
    with pytest.raises(RuntimeError, match="deck"):
        srv.submit("[mesh]\nnx1 = -4\n", [])
    srv.close()
