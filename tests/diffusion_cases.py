"""The reference's diffusion decay tests (tests/physics.sh:13-36, 60-83) as data: a sinusoid of
amplitude 1e-3 in B, v or T along one axis decays like exp(-D k^2 t), k = 2 pi; the fitted
amplitude must be within 2e-3 relative (3e-2 for the temperature) of the exact one."""
import math

import numpy as np

DIFF_DECK = """[mesh]
nx1 = 64
[time]
integrator = rk3
cfl = 0.3
tlim = 0.5
[mhd]
reconstruct = plm
[problem]
name = diffusion_test
amp = 1.0e-3
"""
GAMMA = 5.0 / 3.0
AMP = 1e-3
K = 2 * math.pi

# label, overrides, field, axis, diffusivity, t, tolerance (physics.sh:30-36, 78-81)
DECAY_CASES = [
    ("eta_x", ["problem.var=b", "problem.axis=1", "mesh.nx1=64", "mesh.nx2=1", "mesh.nx3=1", "diffusion.eta=0.02"], "b", 1, 0.02, 0.5, 2e-3),
    ("eta_y", ["problem.var=b", "problem.axis=2", "mesh.nx1=8", "mesh.nx2=64", "mesh.nx3=1", "diffusion.eta=0.02"], "b", 2, 0.02, 0.5, 2e-3),
    ("eta_z", ["problem.var=b", "problem.axis=3", "mesh.nx1=8", "mesh.nx2=8", "mesh.nx3=64", "diffusion.eta=0.02"], "b", 3, 0.02, 0.5, 2e-3),
    ("nu_x", ["problem.var=v", "problem.axis=1", "mesh.nx1=64", "mesh.nx2=1", "mesh.nx3=1", "diffusion.nu=0.02"], "v", 1, 0.02, 0.5, 2e-3),
    ("nu_y", ["problem.var=v", "problem.axis=2", "mesh.nx1=8", "mesh.nx2=64", "mesh.nx3=1", "diffusion.nu=0.02"], "v", 2, 0.02, 0.5, 2e-3),
    ("nu_z", ["problem.var=v", "problem.axis=3", "mesh.nx1=8", "mesh.nx2=8", "mesh.nx3=64", "diffusion.nu=0.02"], "v", 3, 0.02, 0.5, 2e-3),
    ("kap_x", ["problem.var=temp", "problem.axis=1", "mesh.nx1=64", "mesh.nx2=1", "mesh.nx3=1", "diffusion.kappa=0.003",
               "time.tlim=2.0"], "T", 1, 0.003 * (GAMMA - 1) / GAMMA, 2.0, 3e-2),
]


def decay_amplitude(w, bcc, region, var, axis):
    """The fitted sin(k x) amplitude of the perturbed field over the active cells (the mean over
    the other two axes first), as physics.sh:72-77 does on the VTK output."""
    comp = axis % 3
    if var == "b":
        f = bcc[comp]
    elif var == "v":
        f = w[1 + comp]
    else:
        f = w[4] / w[0] - 1.0
    f = f[region]
    ax = {1: 2, 2: 1, 3: 0}[axis]
    prof = f.mean(axis=tuple(a for a in range(3) if a != ax))
    x = (np.arange(prof.size) + 0.5) / prof.size
    return float(np.sum(prof * np.sin(K * x)) / np.sum(np.sin(K * x) ** 2))


def exact_amplitude(D, t):
    return AMP * math.exp(-D * K * K * t)
