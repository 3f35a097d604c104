"""Shared test helpers: decks, oracle/GPU pairing, array comparison regions."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "decks")


def deck(name):
    with open(os.path.join(DECKS, name)) as f:
        return f.read()


def parse_deck(text, overrides=()):
    """Minimal reader of the block.key=value deck (for building GPU configs in tests)."""
    kv, block = {}, ""
    for line in text.splitlines():
        line = line.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("[") and line.endswith("]"):
            block = line[1:-1].strip()
            continue
        if "=" in line:
            k, v = line.split("=", 1)
            kv[block + "." + k.strip()] = v.strip()
    for o in overrides:
        k, v = o.split("=", 1)
        kv[k.strip()] = v.strip()
    return kv


def _b(v):
    return v.lower() in ("true", "1", "yes")


def gpu_mesh_for(text, overrides=()):
    """chasqui_amd.Mesh configured like the deck (the state itself is uploaded separately)."""
    from chasqui_amd import Mesh, make_config, make_mesh_desc
    kv = parse_deck(text, overrides)
    g = lambda k, d: kv.get(k, d)  # noqa: E731
    recon = g("mhd.reconstruct", "ppm")
    ng = int(g("mesh.nghost", 3 if recon in ("ppm", "ppmx", "wenoz") else 2))
    md = make_mesh_desc(int(kv["mesh.nx1"]), int(g("mesh.nx2", 1)), int(g("mesh.nx3", 1)), ng,
                        (float(g("mesh.x1min", 0.0)), float(g("mesh.x1max", 1.0))),
                        (float(g("mesh.x2min", 0.0)), float(g("mesh.x2max", 1.0))),
                        (float(g("mesh.x3min", 0.0)), float(g("mesh.x3max", 1.0))))
    cfg = make_config(recon=recon, rsolver=g("mhd.rsolver", "hlld"), integrator=g("time.integrator", "rk3"),
                      gamma=float(g("mhd.gamma", 5.0 / 3.0)), dfloor=float(g("mhd.dfloor", 1e-10)),
                      pfloor=float(g("mhd.pfloor", 1e-12)), cfl=float(g("time.cfl", 0.3)),
                      flatten=_b(g("mhd.flatten", "true")), low_mach=_b(g("mhd.low_mach", "false")),
                      fofc=_b(g("mhd.fofc", "true")), dual_energy=_b(g("mhd.dual_energy", "false")),
                      de_eta=float(g("mhd.de_eta", 1e-3)), lhlld_phi=float(g("mhd.lhlld_phi", -1.0)),
                      bc=(g("mesh.bc_x1", "periodic"), g("mesh.bc_x2", "periodic"), g("mesh.bc_x3", "periodic")),
                      precision=g("mhd.precision", "fp64"),
                      characteristic=_b(g("mhd.characteristic", "false")),
                      nu=float(g("diffusion.nu", 0.0)), kappa=float(g("diffusion.kappa", 0.0)),
                      eta=float(g("diffusion.eta", 0.0)), sts=_b(g("diffusion.sts", "false")),
                      flux_limiter=int(_b(g("mhd.flux_limiter", "false"))),
                      grav_uniform=int(_b(g("gravity.uniform", "false"))), grav_point=int(_b(g("gravity.point_mass", "false"))),
                      g1=float(g("gravity.g1", 0.0)), g2=float(g("gravity.g2", 0.0)), g3=float(g("gravity.g3", 0.0)),
                      gm=float(g("gravity.gm", 0.0)), gx0=float(g("gravity.x0", 0.0)), gy0=float(g("gravity.y0", 0.0)),
                      gz0=float(g("gravity.z0", 0.0)), rsoft=float(g("gravity.rsoft", 0.0)),
                      cooling=int(_b(g("cooling.enabled", "false"))), lambda0=float(g("cooling.lambda0", 0.0)),
                      cool_alpha=float(g("cooling.alpha", 1.0)), tfloor=float(g("cooling.tfloor", 0.0)))
    return Mesh(md, cfg), kv


STATE = ["u0", "u1", "w", "bcc", "ecc", "b1f0", "b2f0", "b3f0", "b1f1", "b2f1", "b3f1"]


def upload_oracle_state(gm, osim, names=STATE):
    for n in names:
        gm.upload(n, osim.get(n))


def cell_region(o):
    return (slice(None), slice(0, o.nk), slice(0, o.nj), slice(0, o.ni))


def face_region(o, d):
    """Entries of face array b{d+1} that the pipeline reads or writes meaningfully."""
    nk, nj, ni = o.nk, o.nj, o.ni
    if d == 0:
        return (slice(0, nk), slice(0, nj), slice(0, ni + 1))
    if d == 1:
        return (slice(0, nk), slice(0, nj + 1), slice(0, ni))
    return (slice(0, nk + (1 if o.three_d else 0)), slice(0, nj), slice(0, ni))


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


def n_diff(a, b):
    return int(np.count_nonzero(np.asarray(a) != np.asarray(b)))
