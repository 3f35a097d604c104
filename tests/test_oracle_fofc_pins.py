"""Pins of the oracle's fofc branch by what the reference's definitions fix, independently of
the oracle's own restatement of RepairFluxes and Stage() (fofc.hpp:46-96, main.cpp:69-91).

* Face selection and values (fofc.hpp:66-94): after a repair, a face carries the donor-cell
  HLLE flux of its two neighbouring cells' stage-input primitives exactly when one of those
  cells is an active cell with its flag set; every other face (and every face outside the
  repair ranges) keeps its value. The expected values are assembled here in numpy from the
  primitive arrays, the face field of the stage-input set and the pointwise HLLE (itself pinned
  by consistency and supersonic upwinding in test_oracle_numerics.py); flags planted in ghost
  cells must be ignored. On a uniform state the repaired flux is the physical flux of that
  state (HLLE consistency), and on a supersonic state it is the upwind side's physical flux,
  both written out here from the ideal-MHD equations.
* Re-update aliasing (main.cpp:69-75, 86-88, 258-260): when a stage fires, the whole domain is
  updated again on the same buffers, so in RK3 stage 2 (out = ub = u1) and stage 3
  (out = ua = u0) the second update reads the first update's output:
  out = g0*a' + g1*b' - beta_dt*div F_repaired with a', b' the buffers after the first pass.
  The identity is evaluated in numpy (same expression order as physics.hpp:317-320, 331-345)
  from arrays taken before and after the oracle's composed stage; a re-update that read the
  stage inputs, or one limited to the flagged cells, fails it.
"""
import numpy as np
import pytest

from helpers import deck

po = pytest.importorskip("pyoracle")

IDN, IVX, IVY, IVZ, IPR, IENT = 0, 1, 2, 3, 4, 5
ET2 = ("e3x1", "e1x2", "e2x3")   # -F_By of direction 0, 1, 2 (physics.hpp:93-94)
ET1 = ("e2x1", "e3x2", "e1x3")   # +F_Bz


def _ranges(o, d):
    """Face ranges of a direction's flux launch (physics.hpp:85-88), inclusive."""
    il, iu, jl, ju, kl, ku = o.is_, o.ie, o.js, o.je, o.ks, o.ke
    if d == 0:
        iu = o.ie + 1
    elif o.nx1 > 1:
        il, iu = o.is_ - 1, o.ie + 1
    if d == 1:
        ju = o.je + 1
    elif o.two_d:
        jl, ju = o.js - 1, o.je + 1
    if d == 2:
        ku = o.ke + 1
    elif o.three_d:
        kl, ku = o.ks - 1, o.ke + 1
    return il, iu, jl, ju, kl, ku


def _active(o):
    a = np.zeros((o.nk, o.nj, o.ni), dtype=bool)
    a[o.ks:o.ke + 1, o.js:o.je + 1, o.is_:o.ie + 1] = True
    return a


def _expected_repair(o, d, flag, before, gamma):
    """Face arrays after RepairFluxes<d> (fofc.hpp:66-94), built in numpy."""
    il, iu, jl, ju, kl, ku = _ranges(o, d)
    w, bcc = o.get("w"), o.get("bcc")
    bn = o.get(f"b{d + 1}f0")
    act = _active(o) & (flag != 0)
    k, j, i = np.meshgrid(np.arange(kl, ku + 1), np.arange(jl, ju + 1), np.arange(il, iu + 1), indexing="ij")
    k, j, i = k.ravel(), j.ravel(), i.ravel()
    kL, jL, iL = k - (d == 2), j - (d == 1), i - (d == 0)
    ok = (kL >= 0) & (jL >= 0) & (iL >= 0) & (k < o.nk) & (j < o.nj) & (i < o.ni)
    touched = np.zeros(k.shape, dtype=bool)
    touched[ok] = act[k[ok], j[ok], i[ok]] | act[kL[ok], jL[ok], iL[ok]]
    k, j, i, kL, jL, iL = (x[touched] for x in (k, j, i, kL, jL, iL))
    iu_, iv_, iw_ = IVX + d, IVX + (d + 1) % 3, IVX + (d + 2) % 3
    b1_, b2_ = (d + 1) % 3, (d + 2) % 3

    def state(kk, jj, ii):
        return np.stack([w[IDN, kk, jj, ii], w[iu_, kk, jj, ii], w[iv_, kk, jj, ii], w[iw_, kk, jj, ii],
                         w[IPR, kk, jj, ii], bcc[b1_, kk, jj, ii], bcc[b2_, kk, jj, ii]], axis=-1)

    f = po.riemann("hlle", state(kL, jL, iL), state(k, j, i), bn[k, j, i], gamma)
    fx = before[f"f{d + 1}"].copy()
    et2, et1 = before[ET2[d]].copy(), before[ET1[d]].copy()
    fd = f[:, 0]
    for n, col in ((IDN, 0), (iu_, 1), (iv_, 2), (iw_, 3), (IPR, 4)):
        fx[n, k, j, i] = f[:, col]
    sl, sr = w[IENT, kL, jL, iL], w[IENT, k, j, i]
    fx[IENT, k, j, i] = np.where(fd > 0, fd * sl, np.where(fd < 0, fd * sr, 0.5 * (fd * (sl + sr))))
    et2[k, j, i] = -f[:, 5]
    et1[k, j, i] = f[:, 6]
    return {f"f{d + 1}": fx, ET2[d]: et2, ET1[d]: et1}, int(touched.sum())


def _flag_some(o, buf, target=0.5):
    """Set the pressure floor so that the flag of set `buf` marks some but not all of the active
    cells, as close to the fraction `target` as bisection on the floor finds (the count grows
    with the floor; uniform regions make it jump); returns (count, floor)."""
    nact = o.nx1 * o.nx2 * o.nx3
    a, b, best = 0.0, 1e3, None
    for _ in range(120):
        m = 0.5 * (a + b)
        o.set_floors(1e-10, m)
        n = o.fofc_flag(buf)
        if 0 < n < nact and (best is None or abs(n - target * nact) < abs(best[0] - target * nact)):
            best = (n, m)
        if n < target * nact:
            a = m
        else:
            b = m
    assert best is not None, "no pressure floor flags some but not all cells"
    o.set_floors(1e-10, best[1])
    assert o.fofc_flag(buf) == best[0]
    return best


def _face_arrays(o):
    names = ["f1", "f2", "f3"] + list(ET2) + list(ET1)
    return {n: o.get(n) for n in names}


CASES = [
    ("blast_3d.in", ["mesh.nx1=16", "mesh.nx2=14", "mesh.nx3=12"]),
    ("orszag_tang.in", ["mesh.nx1=20", "mesh.nx2=14"]),
    ("brio_wu.in", ["mesh.nx1=40"]),
    ("orszag_tang.in", ["mesh.nx1=20", "mesh.nx2=14", "mhd.reconstruct=ppmx", "mesh.bc_x1=outflow",
                        "mesh.bc_x2=reflect"]),
]


@pytest.mark.parametrize("name,ov", CASES)
def test_repair_rewrites_exactly_the_faces_of_flagged_active_cells(name, ov):
    o = po.OracleSim(deck(name), ov)
    o.step(2)
    gamma = o.geometry()[1][11]
    dirs = [0] + ([1] if o.two_d else []) + ([2] if o.three_d else [])
    for d in dirs:
        o.fluxes(d, 0)
    o.corner_emf()
    o.update(1, 1.0, 0.0, 0.3 * o.min_dt)
    nbad, _ = _flag_some(o, 1)
    fv = o.flag_view()
    flag = fv.copy()
    assert flag[_active(o)].sum() == nbad
    fv[~_active(o)] = 1            # planted in the ghosts: the repair trusts active flags only
    before = _face_arrays(o)
    for d in dirs:
        o.fofc_repair(d, 0)
    after = _face_arrays(o)
    touched_total = 0
    for d in dirs:
        exp, touched = _expected_repair(o, d, flag, before, gamma)
        touched_total += touched
        for n, a in exp.items():
            assert np.array_equal(after[n], a), (name, d, n)
    assert touched_total > nbad
    for d in set(range(3)) - set(dirs):
        for n in (f"f{d + 1}", ET2[d], ET1[d]):
            assert np.array_equal(after[n], before[n])


def _phys_flux(s, bn, gamma):
    """Ideal-MHD flux along the face normal of the primitive state s = (d, u, v, w, p, by, bz)."""
    d, u, v, ww, p, by, bz = (s[..., q] for q in range(7))
    bsq = bn * bn + by * by + bz * bz
    e = p / (gamma - 1.0) + 0.5 * (d * (u * u + v * v + ww * ww) + bsq)
    pt = p + 0.5 * bsq
    vb = u * bn + v * by + ww * bz
    return np.stack([d * u, d * u * u + pt - bn * bn, d * u * v - bn * by, d * u * ww - bn * bz,
                     (e + pt) * u - bn * vb, by * u - bn * v, bz * u - bn * ww], axis=-1)


def _fast_speed(s, bn, gamma):
    d, p, by, bz = s[..., 0], s[..., 4], s[..., 5], s[..., 6]
    a2, b2 = gamma * p / d, (bn * bn + by * by + bz * bz) / d
    return np.sqrt(0.5 * (a2 + b2 + np.sqrt((a2 + b2) ** 2 - 4.0 * a2 * bn * bn / d)))


def _repaired_closed_form(o, d, kind, gamma):
    """Compare the repaired faces of direction d with the physical flux of the upwind state
    (kind 'L' or 'R') or of the common state ('U'), on the faces of the active cells (the faces
    of the transverse ghost ring touch no active cell and keep their values)."""
    il, iu, jl, ju, kl, ku = o.is_, o.ie + (d == 0), o.js, o.je + (d == 1), o.ks, o.ke + (d == 2)
    w, bcc, bn = o.get("w"), o.get("bcc"), o.get(f"b{d + 1}f0")
    iu_, iv_, iw_ = IVX + d, IVX + (d + 1) % 3, IVX + (d + 2) % 3
    b1_, b2_ = (d + 1) % 3, (d + 2) % 3
    fx = o.get(f"f{d + 1}")
    et2, et1 = o.get(ET2[d]), o.get(ET1[d])
    sl_ = (slice(kl - (d == 2), ku + 1 - (d == 2)), slice(jl - (d == 1), ju + 1 - (d == 1)),
           slice(il - (d == 0), iu + 1 - (d == 0)))
    sr_ = (slice(kl, ku + 1), slice(jl, ju + 1), slice(il, iu + 1))
    pick = sl_ if kind == "L" else sr_

    def st(sl):
        return np.stack([w[IDN][sl], w[iu_][sl], w[iv_][sl], w[iw_][sl], w[IPR][sl], bcc[b1_][sl], bcc[b2_][sl]],
                        axis=-1)

    if kind != "U":   # the wave fan lies on one side of the face
        sL, sR, b = st(sl_), st(sr_), bn[sr_]
        smin = np.minimum(sL[..., 1] - _fast_speed(sL, b, gamma), sR[..., 1] - _fast_speed(sR, b, gamma))
        smax = np.maximum(sL[..., 1] + _fast_speed(sL, b, gamma), sR[..., 1] + _fast_speed(sR, b, gamma))
        assert (smin > 0).all() if kind == "L" else (smax < 0).all()
    F = _phys_flux(st(pick), bn[sr_], gamma)
    got = np.stack([fx[IDN][sr_], fx[iu_][sr_], fx[iv_][sr_], fx[iw_][sr_], fx[IPR][sr_], -et2[sr_], et1[sr_]],
                   axis=-1)
    scale = np.max(np.abs(F), axis=tuple(range(F.ndim - 1))) + 1e-300
    assert np.all(np.abs(got - F) <= 1e-13 * scale), (d, kind, np.max(np.abs(got - F) / scale))


def _all_flagged(o):
    fv = o.flag_view()
    fv[...] = 0
    fv[_active(o)] = 1


UNIFORM = ["problem.name=uniform", "mesh.nx1=8", "mesh.nx2=6", "mesh.nx3=5", "problem.dens=1.3", "problem.pres=0.7",
           "problem.bx=0.4", "problem.by=-0.3", "problem.bz=0.2"]


def test_repaired_uniform_state_carries_its_physical_flux():
    """HLLE consistency F(U, U) = F(U) on every repaired face of every direction."""
    o = po.OracleSim(deck("blast_3d.in"), UNIFORM + ["problem.vx=0.2", "problem.vy=-0.1", "problem.vz=0.05"])
    gamma = o.geometry()[1][11]
    _all_flagged(o)
    for d in range(3):
        o.fofc_repair(d, 0)
        _repaired_closed_form(o, d, "U", gamma)


def test_repaired_supersonic_faces_carry_the_upwind_flux():
    """Supersonic flow (+x, -y, +z) with unequal neighbours (a seeded 1% perturbation of every
    primitive and cell field): each repaired face carries the upwind cell's physical flux."""
    o = po.OracleSim(deck("blast_3d.in"), UNIFORM + ["problem.vx=4.0", "problem.vy=-4.5", "problem.vz=5.0"])
    gamma = o.geometry()[1][11]
    rng = np.random.default_rng(7)
    w, bcc = o.view("w"), o.view("bcc")
    w[:5] *= 1.0 + 0.01 * rng.uniform(-1, 1, size=w[:5].shape)
    bcc *= 1.0 + 0.01 * rng.uniform(-1, 1, size=bcc.shape)
    _all_flagged(o)
    for d, kind in ((0, "L"), (1, "R"), (2, "L")):
        o.fofc_repair(d, 0)
        _repaired_closed_form(o, d, kind, gamma)


def _divf(o, F, n):
    """Flux divergence over the active cells, physics.hpp:317-319 order."""
    g, r = o.geometry()
    i1, i2, i3 = 1.0 / r[0], 1.0 / r[1], 1.0 / r[2]
    k, j, i = slice(o.ks, o.ke + 1), slice(o.js, o.je + 1), slice(o.is_, o.ie + 1)
    f1, f2, f3 = F
    dv = (f1[n, k, j, o.is_ + 1:o.ie + 2] - f1[n, k, j, i]) * i1
    if o.two_d:
        dv = dv + (f2[n, k, o.js + 1:o.je + 2, i] - f2[n, k, j, i]) * i2
    if o.three_d:
        dv = dv + (f3[n, o.ks + 1:o.ke + 2, j, i] - f3[n, k, j, i]) * i3
    return dv


def _ct_b1(o, E, b_a, b_b, g0, g1, bdt):
    """B1 on x1-faces (physics.hpp:329-334) over k <= ke, j <= je, i in is..ie+1."""
    g, r = o.geometry()
    i2, i3 = 1.0 / r[1], 1.0 / r[2]
    e1, e2, e3 = E
    k, j, i = slice(o.ks, o.ke + 1), slice(o.js, o.je + 1), slice(o.is_, o.ie + 2)
    curl = np.zeros_like(b_a[k, j, i])
    if o.two_d:
        curl = curl + (e3[k, o.js + 1:o.je + 2, i] - e3[k, j, i]) * i2
    if o.three_d:
        curl = curl - (e2[o.ks + 1:o.ke + 2, j, i] - e2[k, j, i]) * i3
    return g0 * b_a[k, j, i] + g1 * b_b[k, j, i] - bdt * curl


@pytest.mark.parametrize("stage", [1, 2, 3])
@pytest.mark.parametrize("name,ov", [CASES[0], CASES[1]])
def test_fofc_reupdate_reads_the_aliased_first_pass(name, ov, stage):
    """RK3 (main.cpp:258-260): stage 1 u0 -> u1 (1, 0, 1), stage 2 u1 -> u1 (3/4, 1/4, 1/4),
    stage 3 u1 -> u0 (1/3, 2/3, 2/3). When fofc fires, the re-update reads the buffers the first
    update wrote (main.cpp:86-88 with the StageArrays of 69-75)."""
    g0, g1, beta, cur, out = {1: (1.0, 0.0, 1.0, 0, 1), 2: (0.75, 0.25, 0.25, 1, 1),
                              3: (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 1, 0)}[stage]
    o = po.OracleSim(deck(name), ov)
    o.step(2)
    dt = 0.3 * o.min_dt
    if stage > 1:   # bring the state to the input of this stage (stage 1, then con2prim of u1)
        o.stage(0, 1, 1.0, 0.0, dt)
        o.cons_to_prim(1, False)
        if stage == 3:
            o.stage(1, 1, 0.75, 0.25, 0.25 * dt)
            o.cons_to_prim(1, False)
    bdt = beta * dt
    dirs = [0] + ([1] if o.two_d else []) + ([2] if o.three_d else [])
    # a pressure floor that flags part of the first pass's output (found on a copy of the mesh;
    # the floor also enters the fluxes through the donor-cell fallback, so it is set first)
    probe = po.OracleSim(deck(name), ov)
    for nm in ("u0", "u1", "b1f0", "b2f0", "b3f0", "b1f1", "b2f1", "b3f1", "w", "bcc", "ecc"):
        probe.view(nm)[...] = o.view(nm)
    for d in dirs:
        probe.fluxes(d, cur)
    probe.corner_emf()
    probe.update(out, g0, g1, bdt)
    _, pfloor = _flag_some(probe, out, 0.2)
    probe.close()
    o.set_floors(1e-10, pfloor)
    for d in dirs:                   # first-pass fluxes and edge EMFs (stage() recomputes the same)
        o.fluxes(d, cur)
    o.corner_emf()
    F1 = [o.get(f"f{d}") for d in (1, 2, 3)]
    E1 = [o.get(f"e{d}") for d in (1, 2, 3)]
    ua, ub = o.get("u0"), o.get("u1")
    b1a, b1b = o.get("b1f0"), o.get("b1f1")
    nbad = o.stage(cur, out, g0, g1, bdt)
    assert nbad > 0
    F2 = [o.get(f"f{d}") for d in (1, 2, 3)]
    E2 = [o.get(f"e{d}") for d in (1, 2, 3)]
    assert any(not np.array_equal(a, b) for a, b in zip(F1, F2))
    uo = o.get(f"u{out}")
    b1o = o.get(f"b1f{out}")
    act = (slice(o.ks, o.ke + 1), slice(o.js, o.je + 1), slice(o.is_, o.ie + 1))
    # first pass: out = g0*a + g1*b - bdt*div F1, written over the aliased buffer
    first = {}
    for n in range(6):
        first[n] = g0 * ua[n][act] + g1 * ub[n][act] - bdt * _divf(o, F1, n)
    a_after = {n: (first[n] if out == 0 else ua[n][act]) for n in range(6)}
    b_after = {n: (first[n] if out == 1 else ub[n][act]) for n in range(6)}
    stale = 0
    for n in range(6):
        exp = g0 * a_after[n] + g1 * b_after[n] - bdt * _divf(o, F2, n)
        assert np.array_equal(uo[n][act], exp), (name, stage, n)
        # the same expression from the stage inputs (a non-aliased re-update) is another state
        naive = g0 * ua[n][act] + g1 * ub[n][act] - bdt * _divf(o, F2, n)
        stale += int(np.count_nonzero(naive != exp))
    if stage > 1:
        assert stale > 0.5 * uo[0][act].size
    # CT of the x1 faces follows the same aliasing
    b1_first = _ct_b1(o, E1, b1a, b1b, g0, g1, bdt)
    k, j, i = act[0], act[1], slice(o.is_, o.ie + 2)
    ba = b1a.copy()
    bb = b1b.copy()
    (ba if out == 0 else bb)[k, j, i] = b1_first
    assert np.array_equal(b1o[k, j, i], _ct_b1(o, E2, ba, bb, g0, g1, bdt)), (name, stage)


@pytest.mark.parametrize("target", [0.03, 0.3])
@pytest.mark.parametrize("name,ov", CASES)
def test_first_stage_reupdate_changes_only_points_near_flagged_cells(name, ov, target):
    """The premise of a re-update limited to the flagged cells' neighbourhoods: in the first RK
    stage (u0 -> u1 with g1 = 0) the reference's whole-domain re-update after the repair
    (main.cpp:78-91: RepairFluxes, CornerEMF, Update on the aliased buffers) reproduces the first
    pass bit for bit, zero signs included, at every point of the update box
    [is..ie+1] x [js..je+1] x [ks..ke+1] farther than one cell from a flagged active cell. The
    repair changes the fluxes of the flagged cells' faces and the EMFs of their edges; a point
    farther away reads neither, and its aliased input u1 enters as 0 * u1."""
    o = po.OracleSim(deck(name), ov)
    o.step(2)
    bdt = 0.3 * o.min_dt
    dirs = [0] + ([1] if o.two_d else []) + ([2] if o.three_d else [])
    probe = po.OracleSim(deck(name), ov)
    for nm in ("u0", "u1", "b1f0", "b2f0", "b3f0", "b1f1", "b2f1", "b3f1", "w", "bcc", "ecc"):
        probe.view(nm)[...] = o.view(nm)
    for d in dirs:
        probe.fluxes(d, 0)
    probe.corner_emf()
    probe.update(1, 1.0, 0.0, bdt)
    _, pfloor = _flag_some(probe, 1, target)
    probe.close()
    o.set_floors(1e-10, pfloor)
    for d in dirs:
        o.fluxes(d, 0)
    o.corner_emf()
    o.update(1, 1.0, 0.0, bdt)
    names = ["u1", "b1f1", "b2f1", "b3f1"]
    first = {nm: o.get(nm) for nm in names}
    assert o.fofc_flag(1) > 0
    flagged = _active(o) & (o.flag() != 0)
    for d in dirs:
        o.fofc_repair(d, 0)
    o.corner_emf()
    o.update(1, 1.0, 0.0, bdt)
    whole = {nm: o.get(nm) for nm in names}
    # points within one cell of a flagged active cell (a 3x3x3 dilation, clipped to the mesh)
    near = np.zeros_like(flagged)
    kk, jj, ii = np.nonzero(flagged)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                k, j, i = kk + dk, jj + dj, ii + di
                ok = (k >= 0) & (k < o.nk) & (j >= 0) & (j < o.nj) & (i >= 0) & (i < o.ni)
                near[k[ok], j[ok], i[ok]] = True
    ku, ju = (o.ke + 1 if o.three_d else o.ks), (o.je + 1 if o.two_d else o.js)
    box = np.zeros_like(flagged)
    box[o.ks:ku + 1, o.js:ju + 1, o.is_:o.ie + 2] = True
    changed_near = 0
    for nm in names:
        a, b = first[nm], whole[nm]
        a = a.reshape((-1,) + a.shape[-3:])
        b = b.reshape((-1,) + b.shape[-3:])
        diff = np.any(a.view(np.int64) != b.view(np.int64), axis=0)   # bit patterns: -0 != +0
        pts = np.zeros(diff.shape, dtype=bool)    # face arrays: one more entry per dimension
        pts[:o.nk, :o.nj, :o.ni] = near & box
        # a point's update also writes the copy of B2 one row up in 1D, of B3 one plane up
        # below 3D (physics.hpp:339, 345)
        if nm == "b2f1" and not o.two_d:
            pts[:, o.js + 1] |= pts[:, o.js]
        if nm == "b3f1" and not o.three_d:
            pts[o.ks + 1] |= pts[o.ks]
        assert not np.any(diff & ~pts), (nm, int(np.count_nonzero(diff & ~pts)))
        changed_near += int(np.count_nonzero(diff))
    assert changed_near > 0
