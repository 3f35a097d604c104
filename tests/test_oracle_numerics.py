"""Pins the oracle's pointwise numerics: the reference's bitwise mirror-symmetry invariant
(tests/symmetry_test.cpp:20-70, same sample classes, extended to DC, LHLLD and the
flattening coefficient) and closed forms that a plausible slip would break."""
import numpy as np
import pytest

po = pytest.importorskip("pyoracle")

N = 200_000


def _recon_samples(rng):
    q = rng.uniform(-2.0, 2.0, size=(N, 6))
    idx = np.arange(N)
    q[idx % 3 == 0] = np.abs(q[idx % 3 == 0])                     # positive data
    ev = idx % 5 == 0                                               # even-symmetric stencil
    q[ev, 3], q[ev, 4], q[ev, 5] = q[ev, 2], q[ev, 1], q[ev, 0]
    return q


@pytest.mark.parametrize("method", ["dc", "plm", "ppm", "ppmx", "wenoz"])
def test_recon_mirror_bitwise(method):
    rng = np.random.default_rng(12345)
    q = _recon_samples(rng)
    L, R = po.reconstruct(method, q)
    Lm, Rm = po.reconstruct(method, q[:, ::-1].copy())
    La, Ra = po.reconstruct(method, -q[:, ::-1].copy())
    assert np.array_equal(Lm, R) and np.array_equal(Rm, L)
    assert np.array_equal(La, -R) and np.array_equal(Ra, -L)


def _riemann_samples(rng):
    P = lambda: rng.uniform(0.05, 2.0, N)  # noqa: E731
    U = lambda: rng.uniform(-1.0, 1.0, N)  # noqa: E731
    L = np.stack([P(), U(), U(), U(), P(), U(), U()], axis=1)
    R = np.stack([P(), U(), U(), U(), P(), U(), U()], axis=1)
    bn = U()
    idx = np.arange(N)
    bn[idx % 4 == 0] = 0.0                                          # degenerate branch
    m = idx % 7 == 0                                                # exact mirror pair
    R[m] = L[m]
    R[m, 1] = -L[m, 1]
    return L, R, bn


def _mirror(L, R):
    Lm = R.copy()
    Lm[:, 1] = -R[:, 1]
    Rm = L.copy()
    Rm[:, 1] = -L[:, 1]
    return Lm, Rm


SIGNS = np.array([-1, 1, -1, -1, -1, -1, -1], dtype=float)   # FD FMU FMV FMW FEN FBY FBZ


@pytest.mark.parametrize("solver", ["hlle", "hlld", "lhlld"])
def test_riemann_mirror_bitwise(solver):
    rng = np.random.default_rng(777)
    L, R, bn = _riemann_samples(rng)
    theta = rng.uniform(0.0, 1.0, N) if solver == "lhlld" else None
    phi = rng.uniform(0.0, 1.0, N) if solver == "lhlld" else None
    f = po.riemann(solver, L, R, bn, theta=theta, phi=phi)
    Lm, Rm = _mirror(L, R)
    g = po.riemann(solver, Lm, Rm, -bn, theta=theta, phi=phi)
    assert np.array_equal(g, f * SIGNS), int(np.count_nonzero(g != f * SIGNS))


def test_flatten_mirror_bitwise():
    rng = np.random.default_rng(99)
    p = rng.uniform(0.01, 5.0, size=(N, 5))
    um, up = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    f = po.flatten(p, um, up)
    fm = po.flatten(p[:, ::-1].copy(), -up, -um)
    assert np.array_equal(f, fm)


@pytest.mark.parametrize("method", ["dc", "plm", "ppm", "ppmx", "wenoz"])
def test_recon_constant_data(method):
    q = np.full((5, 6), 0.0) + np.array([0.3, -1.7, 2.5, 1e-8, 7.0])[:, None]
    L, R = po.reconstruct(method, q)
    assert np.array_equal(L, q[:, 0]) and np.array_equal(R, q[:, 0])


@pytest.mark.parametrize("method", ["plm", "ppm", "ppmx", "wenoz"])
def test_recon_linear_data_exact(method):
    # cell averages of a linear profile: every method is exact at the face (x = 0)
    a, b = 0.7, -1.3
    x = np.arange(-3, 3) + 0.5
    q = (a + b * x)[None, :]
    L, R = po.reconstruct(method, q)
    assert abs(L[0] - a) < 1e-14 and abs(R[0] - a) < 1e-14


def test_wenoz_smooth_quadratic_is_fifth_order_linear_scheme():
    # For cell averages of a quadratic the three smoothness indicators are equal, tau = 0,
    # so WENO-Z uses the ideal weights and reduces to the 5th-order upwind formula
    # (2a - 13b + 47c + 27d - 3e) / 60 of Jiang & Shu 1996 — exact for polynomials of
    # degree <= 4, hence equal to the true face value of the quadratic.
    c0, c1, c2 = 0.4, -0.9, 0.35
    avg = lambda x0: c0 + c1 * x0 + c2 * (x0 ** 2 + 1.0 / 12.0)  # noqa: E731 cell average on [x0-1/2, x0+1/2]
    x = np.arange(-3, 3) + 0.5
    q = avg(x)[None, :]
    L, R = po.reconstruct("wenoz", q)
    a, b, c, d, e = q[0, :5]
    lin = (2 * a - 13 * b + 47 * c + 27 * d - 3 * e) / 60.0
    assert abs(L[0] - lin) < 1e-14 and abs(L[0] - c0) < 1e-14 and abs(R[0] - c0) < 1e-14


def test_ppm_extremum_gives_donor_cell():
    # cell i-1 (q[2]) is a local maximum: the CW84 limiter flattens its parabola
    q = np.array([[0.0, 1.0, 3.0, 1.0, 0.5, 0.2]])
    L, _ = po.reconstruct("ppm", q)
    assert L[0] == 3.0


def test_plm_van_leer_slope_closed_form():
    q = np.array([[0.0, 1.0, 2.0, 4.0, 5.0, 6.0]])
    L, R = po.reconstruct("plm", q)
    vl = lambda a, b, c: 2 * (b - a) * (c - b) / ((b - a) + (c - b)) if (b - a) * (c - b) > 0 else 0.0  # noqa: E731
    assert L[0] == 2.0 + 0.5 * vl(1.0, 2.0, 4.0)
    assert R[0] == 4.0 - 0.5 * vl(2.0, 4.0, 5.0)


def test_flattening_step_closed_form():
    p = np.array([[1.0, 1.0, 1.0, 10.0, 10.0], [1.0, 1.0, 5.0, 10.0, 10.0], [1.0, 1.0, 1.0, 1.2, 1.2]])
    f_comp = po.flatten(p, np.array([1.0, 1.0, 1.0]), np.array([0.0, 0.0, 0.0]))
    f_exp = po.flatten(p, np.array([0.0, 0.0, 0.0]), np.array([1.0, 1.0, 1.0]))
    # dp / dp2 = 1 -> 10 * (1 - 0.75) clipped to 1; expanding flow or weak jump -> 0
    assert list(f_comp) == [1.0, 1.0, 0.0]
    assert list(f_exp) == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("solver", ["hlle", "hlld", "lhlld"])
def test_riemann_consistency(solver):
    rng = np.random.default_rng(5)
    S = np.stack([rng.uniform(0.1, 2, 500), rng.uniform(-1, 1, 500), rng.uniform(-1, 1, 500),
                  rng.uniform(-1, 1, 500), rng.uniform(0.1, 2, 500), rng.uniform(-1, 1, 500),
                  rng.uniform(-1, 1, 500)], axis=1)
    bn = rng.uniform(-1, 1, 500)
    f = po.riemann(solver, S, S, bn)
    F = po.phys_flux(S, bn)
    assert np.allclose(f, F, rtol=1e-12, atol=1e-13)


def _phys_flux_closed_form(s, bn, gamma):
    d, u, v, w, p, by, bz = s
    b2 = bn * bn + by * by + bz * bz
    e = p / (gamma - 1) + 0.5 * (d * (u * u + v * v + w * w) + b2)
    pt = p + 0.5 * b2
    vb = u * bn + v * by + w * bz
    return np.array([d * u, d * u * u + pt - bn * bn, d * u * v - bn * by, d * u * w - bn * bz,
                     (e + pt) * u - bn * vb, by * u - bn * v, bz * u - bn * w])


def test_phys_flux_closed_form():
    s = np.array([1.3, 0.4, -0.2, 0.7, 0.9, 0.3, -0.6])
    assert np.allclose(po.phys_flux(s, 0.8)[0], _phys_flux_closed_form(s, 0.8, 5.0 / 3.0), rtol=1e-14)


@pytest.mark.parametrize("solver", ["hlle", "hlld"])
def test_supersonic_upwind(solver):
    L = np.array([[1.0, 20.0, 0.1, -0.2, 0.5, 0.3, 0.2]])
    R = np.array([[0.5, 19.0, 0.0, 0.1, 0.4, -0.2, 0.1]])
    f = po.riemann(solver, L, R, 0.4)
    assert np.array_equal(f, po.phys_flux(L, 0.4))
    f = po.riemann(solver, -L * np.array([-1, 1, -1, -1, -1, -1, -1]), R * np.array([1, -1, 1, 1, 1, 1, 1]), 0.4)
    assert np.isfinite(f).all()


def test_hlld_resolves_stationary_contact():
    """Miyoshi & Kusano 2005: HLLD is exact for an isolated contact (density jump only)."""
    bn = 0.7
    L = np.array([[1.0, 0.0, 0.0, 0.0, 1.0, 0.4, -0.3]])
    R = np.array([[0.2, 0.0, 0.0, 0.0, 1.0, 0.4, -0.3]])
    f = po.riemann("hlld", L, R, bn)[0]
    pt = 1.0 + 0.5 * (bn * bn + 0.4 ** 2 + 0.3 ** 2)
    exact = np.array([0.0, pt - bn * bn, -bn * 0.4, bn * 0.3, 0.0, 0.0, 0.0])
    assert np.allclose(f, exact, atol=1e-15)
    fe = po.riemann("hlle", L, R, bn)[0]
    assert abs(fe[0]) > 1e-3   # HLLE smears the contact (sanity of the test itself)


def test_fast_speed_closed_form():
    # c_f^2 = 1/2 (a^2 + b^2 + sqrt((a^2 + b^2)^2 - 4 a^2 b_n^2)) with a^2 = gamma p / rho
    s = np.array([2.0, 0.0, 0.0, 0.0, 0.6, 0.5, -0.2])
    bn, gamma = 0.8, 5.0 / 3.0
    a2 = gamma * 0.6 / 2.0
    b2 = (bn ** 2 + 0.5 ** 2 + 0.2 ** 2) / 2.0
    bx2 = bn ** 2 / 2.0
    cf = np.sqrt(0.5 * ((a2 + b2) + np.sqrt((a2 + b2) ** 2 - 4 * a2 * bx2)))
    assert abs(po.fast_speed(s, bn, gamma) - cf) < 1e-14


# ---------------------------------------------------------------- characteristic variables
def _eigen_samples(n=20000, seed=7, gamma=5.0 / 3.0):
    """Random primitive states with the degenerate families of tests/eigen_test.cpp:16-22:
    no transverse field, no normal field, the hydro limit, and cf = cs = a."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.05, 3.0, n)
    p = rng.uniform(0.05, 3.0, n)
    u, v, w = (rng.uniform(-1.5, 1.5, n) for _ in range(3))
    by, bz, bx = (rng.uniform(-1.5, 1.5, n) for _ in range(3))
    case = np.arange(n) % 5
    by[case == 1] = 0.0
    bz[case == 1] = 0.0
    bx[case == 2] = 0.0
    bx[case == 3] = 0.0
    by[case == 3] = 0.0
    bz[case == 3] = 0.0
    by[case == 4] = 0.0
    bz[case == 4] = 0.0
    bx[case == 4] = np.sqrt(gamma * p[case == 4])
    return d, u, v, w, p, by, bz, bx


def test_eigensystem_inverse_and_jacobian():
    """The reference's own check (tests/eigen_test.cpp:24-51): L R = I and A R = R Lambda for
    the primitive Jacobian A, to 1e-9, over 20000 states including the degenerate ones."""
    g = 5.0 / 3.0
    d, u, v, w, p, by, bz, bx = _eigen_samples(gamma=g)
    L, R = po.eigensystem(d, p, by, bz, bx, g)
    n = d.size
    A = np.zeros((n, 7, 7))
    A[:, 0, 0] = u; A[:, 0, 1] = d
    A[:, 1, 1] = u; A[:, 1, 4] = 1.0 / d; A[:, 1, 5] = by / d; A[:, 1, 6] = bz / d
    A[:, 2, 2] = u; A[:, 2, 5] = -bx / d
    A[:, 3, 3] = u; A[:, 3, 6] = -bx / d
    A[:, 4, 1] = g * p; A[:, 4, 4] = u
    A[:, 5, 1] = by; A[:, 5, 2] = -bx; A[:, 5, 5] = u
    A[:, 6, 1] = bz; A[:, 6, 3] = -bx; A[:, 6, 6] = u
    asq, casq, ctsq = g * p / d, bx * bx / d, (by * by + bz * bz) / d
    tsum, tdif = asq + casq + ctsq, asq + ctsq - casq
    disc = np.sqrt(np.maximum(tdif * tdif + 4 * ctsq * casq, 0.0))
    cf = np.sqrt(0.5 * (tsum + disc))
    cs = np.sqrt(np.maximum(0.5 * (tsum - disc), 0.0))
    ca = np.sqrt(casq)
    lam = np.stack([u - cf, u - ca, u - cs, u, u + cs, u + ca, u + cf], axis=1)
    elr = np.abs(L @ R - np.eye(7)).max(axis=(1, 2))
    ear = np.abs(A @ R - R * lam[:, None, :]).max(axis=(1, 2))
    assert elr.max() < 1e-9, elr.max()
    assert ear.max() < 1e-9, ear.max()


def test_eigensystem_closed_forms():
    """Hydro limit (B = 0): the acoustic eigenvectors of Euler in primitive variables, and the
    entropy wave row L[3] = (1, 0, 0, 0, -1/a^2, 0, 0)."""
    g = 1.4
    d, p = 1.3, 0.7
    L, R = po.eigensystem(d, p, 0.0, 0.0, 0.0, g)
    L, R = L[0], R[0]
    a2 = g * p / d
    a = np.sqrt(a2)
    assert np.allclose(R[:, 6], [d, a, 0, 0, d * a2, 0, 0], rtol=1e-15, atol=0)
    assert np.allclose(R[:, 0], [d, -a, 0, 0, d * a2, 0, 0], rtol=1e-15, atol=0)
    assert np.array_equal(L[3], [1.0, 0, 0, 0, -1.0 / a2, 0, 0])
    assert np.allclose(L @ R, np.eye(7), atol=1e-15)


def test_charac_reconstruction_of_constant_and_linear_data():
    """R L = I, so a constant stencil reconstructs to the state itself and, for PLM and PPM, a
    linear profile (a single slope, not limited) to the primitive reconstruction's values, up
    to round-off; a transposed L or R fails this."""
    rng = np.random.default_rng(3)
    n = 200
    base = np.stack([rng.uniform(0.5, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                     rng.uniform(0.3, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)], axis=1)
    bx = rng.uniform(-1, 1, n)
    const = np.repeat(base[:, :, None], 6, axis=2)
    for m in ("plm", "ppm", "ppmx", "wenoz"):
        for side in (0, 1):
            out = po.charac_side(m, const, bx, side)
            assert np.max(np.abs(out - base)) < 1e-13, m
    slope = rng.uniform(-1e-3, 1e-3, (n, 7))
    lin = base[:, :, None] + slope[:, :, None] * (np.arange(6) - 2.5)[None, None, :]
    for m in ("plm", "ppm"):
        for side in (0, 1):
            out = po.charac_side(m, lin, bx, side)
            prim = np.stack([np.array(po.reconstruct(m, lin[:, v, :])[side]) for v in range(7)], axis=1)
            assert np.max(np.abs(out - prim)) < 1e-12, (m, side, np.max(np.abs(out - prim)))


def test_charac_limits_a_single_wave_like_the_primitive_scheme():
    """A jump carried by one characteristic field (a right eigenvector times a step) is, in
    characteristic variables, a scalar step in that field alone: PPM then gives the cell
    values exactly (no overshoot in any primitive), where the primitive scheme mixes."""
    d, p, by, bz, bx, g = 1.0, 0.6, 0.8, -0.3, 0.5, 5.0 / 3.0
    L, R = po.eigensystem(d, p, by, bz, bx, g)
    q0 = np.array([d, 0.1, -0.2, 0.05, p, by, bz])
    step = np.array([0, 0, 0, 1.0, 1.0, 1.0]) * 1e-6    # a tiny jump: the eigenvectors barely change
    st = q0[:, None] + R[0][:, 6][:, None] * step[None, :]
    for side, cell in ((0, 2), (1, 3)):
        out = po.charac_side("ppm", st[None], np.array([bx]), side)[0]
        assert np.max(np.abs(out - st[:, cell])) < 1e-13


# ---------------------------------------------------------------- LHLLD and low-Mach pins
def test_lhlld_supersonic_upwind():
    """riemann.hpp:104-113: the LHLLD speeds are clamped at zero, so a supersonic face takes
    S_L = 0 and returns the left physical flux exactly."""
    L = np.array([[1.0, 20.0, 0.1, -0.2, 0.5, 0.3, 0.2]])
    R = np.array([[0.5, 19.0, 0.0, 0.1, 0.4, -0.2, 0.1]])
    assert np.array_equal(po.riemann("lhlld", L, R, 0.4, theta=0.3, phi=0.7), po.phys_flux(L, 0.4))


def test_lhlld_consistent_for_any_theta_phi():
    """theta and phi scale jump terms only (riemann.hpp:123-124): F(U, U) = F(U) for any
    theta, phi in [0, 1]."""
    rng = np.random.default_rng(3)
    n = 500
    S = np.stack([rng.uniform(0.1, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                  rng.uniform(0.1, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)], axis=1)
    bn = rng.uniform(-1, 1, n)
    f = po.riemann("lhlld", S, S, bn, theta=rng.uniform(0, 1, n), phi=rng.uniform(0, 1, n))
    assert np.allclose(f, po.phys_flux(S, bn), rtol=1e-12, atol=1e-13)


def test_lhlld_theta_scales_the_pressure_term_of_the_contact_speed():
    """Gas at rest with a pressure jump (B = 0): with theta = 0 the contact speed
    ((dsr uR - dsl uL) + theta (ptL - ptR)) / (dsr - dsl) is exactly 0 and so is the mass
    flux; theta = 1 gives the usual pressure-driven mass flux toward the low side."""
    L = np.array([[1.0, 0.0, 0.0, 0.0, 2.0, 0.0, 0.0]])
    R = np.array([[1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0]])
    f0 = po.riemann("lhlld", L, R, 0.0, theta=0.0, phi=1.0)[0]
    f1 = po.riemann("lhlld", L, R, 0.0, theta=1.0, phi=1.0)[0]
    assert f0[0] == 0.0
    assert f1[0] > 1e-3


def test_lhlld_phi_scales_the_velocity_term_of_the_star_pressure():
    """A symmetric collision at equal pressure (B = 0): the contact is at rest, so the normal
    momentum flux is the star pressure pts = P + phi rho^2 (sR - uR)(sL - uL)(uR - uL) /
    (rho (sR - uR) - rho (sL - uL)) (riemann.hpp:124), with the LHLLD speeds
    sL = min(0, min(uL, uR) - cmax), sR = max(0, max(uL, uR) + cmax)."""
    rho, P, a, g = 1.0, 1.0, 0.3, 5.0 / 3.0
    L = np.array([[rho, a, 0.0, 0.0, P, 0.0, 0.0]])
    R = np.array([[rho, -a, 0.0, 0.0, P, 0.0, 0.0]])
    cmax = np.sqrt(g * P / rho)
    sl, sr = min(0.0, -a - cmax), max(0.0, a + cmax)
    for phi in (0.0, 0.4, 1.0):
        f = po.riemann("lhlld", L, R, 0.0, theta=1.0, phi=phi)[0]
        pts = P + phi * rho * rho * (sr + a) * (sl - a) * (-2 * a) / (rho * (sr + a) - rho * (sl - a))
        assert abs(f[1] - pts) <= 1e-14 * pts, (phi, f[1], pts)
        assert abs(f[0]) <= 1e-15


def _flux_grid(text, setup):
    o = po.OracleSim(text, [])
    w, bcc = o.view("w"), o.view("bcc")
    setup(o, w, bcc)
    o.fluxes(0, 0)
    return o, o.get("f1")


def _state7(w, bcc, k, j, i):
    return np.array([[w[0, k, j, i], w[1, k, j, i], w[2, k, j, i], w[3, k, j, i], w[4, k, j, i],
                      bcc[1, k, j, i], bcc[2, k, j, i]]])


def test_lhlld_shock_detector_wiring():
    """physics.hpp:182-211 on a grid: an x-face with a pressure jump whose two cells sit on a
    compressive jump of v_y along y gets theta = (cf / (2A + cf))^4 (the transverse one-sided
    differences of the face's two cells, normal jump du = 0); the same face one row away gets
    theta = 1. Donor-cell states, phi forced to 1: the stored flux is the LHLLD flux of the
    two cell states with that theta, bitwise."""
    text = ("[mesh]\nnx1 = 12\nnx2 = 12\n[mhd]\nreconstruct = dc\nrsolver = lhlld\nlhlld_phi = 1.0\n"
            "flatten = false\n[problem]\nname = uniform\ndens = 1.0\npres = 1.0\n")
    A = 0.75

    def setup(o, w, bcc):
        w[0] = 1.0
        w[1] = 0.0
        w[3] = 0.0
        w[2] = np.where(np.arange(o.nj)[None, :, None] >= o.js + 6, -A, A)   # v_y jumps down across row js+6
        w[4] = np.where(np.arange(o.ni)[None, None, :] >= o.is_ + 5, 2.0, 1.0)   # p jumps across x-face is+5
        bcc[:] = 0.0

    o, f1 = _flux_grid(text, setup)
    w, bcc = o.get("w"), o.get("bcc")
    i = o.is_ + 5
    for j, sheared in ((o.js + 5, True), (o.js + 6, True), (o.js + 2, False), (o.js + 9, False)):
        L, R = _state7(w, bcc, 0, j, i - 1), _state7(w, bcc, 0, j, i)
        cf = max(po.fast_speed(L[0], 0.0, 5.0 / 3.0), po.fast_speed(R[0], 0.0, 5.0 / 3.0))
        ratio = min(1.0, cf / (2 * A + cf)) if sheared else 1.0
        theta = (ratio * ratio) * (ratio * ratio)
        f = po.riemann("lhlld", L, R, 0.0, theta=theta, phi=1.0)[0]
        assert np.array_equal(f1[:5, 0, j, i], f[:5]), (j, f1[:5, 0, j, i], f[:5])
        if sheared:
            assert not np.array_equal(f[:5], po.riemann("lhlld", L, R, 0.0, theta=1.0, phi=1.0)[0][:5])


def test_low_mach_blend_wiring():
    """physics.hpp:170-178 on a grid: the face's normal velocities are blended with
    z = min(1, max(|uL| / cL, |uR| / cR)) before the HLLD solve (um + ud, um - ud with
    um = (uL + uR) / 2, ud = z (uL - uR) / 2); a supersonic pair (z = 1) and a subsonic one."""
    for uL, uR in ((0.05, -0.02), (3.0, 2.5)):
        text = ("[mesh]\nnx1 = 12\n[mhd]\nreconstruct = dc\nrsolver = hlld\nlow_mach = true\n"
                "[problem]\nname = uniform\ndens = 1.0\npres = 1.0\nbx = 0.3\nby = 0.2\n")

        def setup(o, w, bcc, uL=uL, uR=uR):
            w[1] = np.where(np.arange(o.ni)[None, None, :] >= o.is_ + 6, uR, uL)
            w[4] = np.where(np.arange(o.ni)[None, None, :] >= o.is_ + 6, 0.8, 1.0)

        o, f1 = _flux_grid(text, setup)
        w, bcc = o.get("w"), o.get("bcc")
        i = o.is_ + 6
        bn = o.get("b1f0")[0, 0, i]
        L, R = _state7(w, bcc, 0, 0, i - 1), _state7(w, bcc, 0, 0, i)
        g = 5.0 / 3.0
        cl, cr = po.fast_speed(L[0], bn, g), po.fast_speed(R[0], bn, g)
        z = min(1.0, max(abs(L[0, 1]) / cl, abs(R[0, 1]) / cr))
        um, ud = 0.5 * (L[0, 1] + R[0, 1]), 0.5 * (z * (L[0, 1] - R[0, 1]))
        Lb, Rb = L.copy(), R.copy()
        Lb[0, 1], Rb[0, 1] = um + ud, um - ud
        f = po.riemann("hlld", Lb, Rb, bn)[0]
        assert np.array_equal(f1[:5, 0, 0, i], f[:5]), (uL, uR)
        if z < 1.0:
            assert not np.array_equal(f[:5], po.riemann("hlld", L, R, bn)[0][:5])
