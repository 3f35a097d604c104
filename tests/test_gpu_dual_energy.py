"""GPU: the dual-energy path (mhd.dual_energy = true) against the CPU oracle.

With dual energy on, the entropy IENT is advected with its own upwind flux
(physics.hpp:218-223), con2prim takes the pressure from it where e_int < de_eta * E
(physics.hpp:463-472), and the fofc flag does the same (fofc.hpp:33-36); the device runs the
reference's kernel order for it. Where the switch fires, p = s * pow(rho, gamma) and glibc's pow
and the device library's may differ by an ulp, so the bar is the FP64 drop-in tolerance:
1e-12 relative L-infinity per field (relative to the field's largest magnitude). Cases where
the switch fires in most cells (de_eta raised), where fofc repairs every stage, and the
reference's recommended preset (ppmx + lhlld + flattening + fofc + dual energy) on
Orszag-Tang.
"""
import pytest

from helpers import deck, face_region, gpu_mesh_for, max_rel, upload_oracle_state

pytestmark = pytest.mark.gpu

po = pytest.importorskip("pyoracle")

RTOL = 1e-12

CASES = [
    ("orszag_tang.in", ["mesh.nx1=32", "mesh.nx2=32", "mhd.dual_energy=true", "mhd.reconstruct=ppmx",
                        "mhd.rsolver=lhlld", "mhd.flatten=true"], 10),
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=16", "mesh.nx3=16", "mhd.dual_energy=true", "mhd.pfloor=0.1"], 3),
    ("brio_wu.in", ["mesh.nx1=200", "mhd.dual_energy=true", "mhd.de_eta=0.5"], 30),
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=12", "mesh.nx3=10", "mhd.dual_energy=true", "mhd.de_eta=0.3"], 4),
]


@pytest.mark.parametrize("name,ov,nsteps", CASES, ids=[f"{c[0]}-{c[1][-1]}" for c in CASES])
def test_dual_energy_time_loop(name, ov, nsteps):
    print(f"case dual energy {name} {ov} steps={nsteps}", flush=True)
    text = deck(name)
    o = po.OracleSim(text, ov)
    gm, kv = gpu_mesh_for(text, ov)
    upload_oracle_state(gm, o)
    tlim = float(kv.get("time.tlim", 1.0))
    gm.init(0.0, tlim, -1)
    o.step(nsteps)
    gm.step(nsteps)
    st = gm.status()
    assert st["cycle"] == o.cycle == nsteps
    assert abs(st["t"] - o.time) <= RTOL * o.time
    assert abs(st["min_dt"] - o.min_dt) <= RTOL * o.min_dt
    assert st["fofc_total"] == o.fofc_count()
    worst = {}
    u, ou = gm.download("u0"), o.get("u0")
    for n in range(ou.shape[0]):
        worst[f"u0[{n}]"] = max_rel(u[n], ou[n])
    for b in ("b1f0", "b2f0", "b3f0"):
        r = face_region(o, int(b[1]) - 1)
        worst[b] = max_rel(gm.download(b)[r], o.get(b)[r])
    w, ow = gm.download("w"), o.get("w")
    for n in range(ow.shape[0]):
        worst[f"w[{n}]"] = max_rel(w[n], ow[n])
    worst["bcc"] = max_rel(gm.download("bcc"), o.get("bcc"))
    print("worst relative differences", {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
    bad = {k: v for k, v in worst.items() if v > RTOL}
    assert not bad, bad
    gm.close()
