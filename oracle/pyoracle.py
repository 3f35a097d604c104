"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg, never by the chasqui_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

ARRAYS_CELL4 = {"u0": 6, "u1": 6, "w": 6, "bcc": 3, "ecc": 3, "u_init": 6, "bcc_init": 3}
ARRAYS_FACE4 = {"f1": 6, "f2": 6, "f3": 6}
ARRAYS_FACE3 = ["b1f0", "b2f0", "b3f0", "b1f1", "b2f1", "b3f1",
                "e2x1", "e3x1", "e1x2", "e3x2", "e1x3", "e2x3", "e1", "e2", "e3"]

RECON = {"dc": 0, "plm": 1, "ppm": 2, "wenoz": 3, "ppmx": 4}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
    L.orc_create.restype = vp
    L.orc_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    L.orc_create_ex.restype = vp
    L.orc_create_ex.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    L.orc_destroy.argtypes = [vp]
    L.orc_geometry.argtypes = [vp, ip, dp]
    L.orc_array.restype = dp
    L.orc_array.argtypes = [vp, C.c_char_p, lp]
    L.orc_flag.restype = ip
    L.orc_flag.argtypes = [vp, lp]
    L.orc_step.argtypes = [vp, C.c_int]
    L.orc_l1_error.argtypes = [vp, dp]
    L.orc_restart_write.argtypes = [vp, C.c_char_p]
    L.orc_restart_read.argtypes = [vp, C.c_char_p]
    L.orc_fofc_count.restype = C.c_long
    L.orc_fofc_count.argtypes = [vp]
    L.orc_fluxes.argtypes = [vp, C.c_int, C.c_int]
    L.orc_corner_emf.argtypes = [vp]
    L.orc_update.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double]
    L.orc_boundaries.argtypes = [vp, C.c_int]
    L.orc_cons_to_prim.restype = C.c_double
    L.orc_cons_to_prim.argtypes = [vp, C.c_int, C.c_int]
    L.orc_fofc_flag.argtypes = [vp, C.c_int]
    L.orc_fofc_repair.argtypes = [vp, C.c_int, C.c_int]
    L.orc_diffusion_fluxes.argtypes = [vp, C.c_int, C.c_int]
    L.orc_resistive_emf.argtypes = [vp, C.c_int]
    L.orc_diffusion_dt.restype = C.c_double
    L.orc_diffusion_dt.argtypes = [vp]
    L.orc_sts_count.restype = C.c_long
    L.orc_sts_count.argtypes = [vp]
    L.orc_limiter_count.restype = C.c_long
    L.orc_limiter_count.argtypes = [vp]
    L.orc_limit_fluxes.restype = C.c_long
    L.orc_limit_fluxes.argtypes = [vp, C.c_int, C.c_double]
    L.orc_apply_gravity.argtypes = [vp, C.c_int, C.c_double]
    L.orc_apply_cooling.argtypes = [vp, C.c_int, C.c_double]
    L.orc_stage.restype = C.c_long
    L.orc_stage.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
    L.orc_set_floors.argtypes = [vp, C.c_double, C.c_double]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.restype = C.c_int
    L.orc_last_error.restype = C.c_char_p
    L.orc_reconstruct.argtypes = [C.c_int, dp, C.c_long, dp, dp]
    L.orc_flatten.argtypes = [dp, dp, dp, C.c_long, dp]
    L.orc_riemann.argtypes = [C.c_int, dp, dp, dp, C.c_long, C.c_double, dp, dp, dp]
    L.orc_phys_flux.argtypes = [dp, dp, C.c_long, C.c_double, dp]
    L.orc_fast_speed.restype = C.c_double
    L.orc_fast_speed.argtypes = [dp, C.c_double, C.c_double]
    L.orc_noise.restype = C.c_double
    L.orc_noise.argtypes = [C.c_ulonglong, C.c_ulonglong]
    L.orc_eigensystem.argtypes = [dp, C.c_long, C.c_double, dp, dp]
    L.orc_charac_side.argtypes = [C.c_int, dp, dp, C.c_long, C.c_double, C.c_int, C.c_int, dp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n):
    lib().orc_set_threads(int(n))


class OracleSim:
    """One mesh of the CPU oracle, initialised from a deck (text) plus overrides."""

    def __init__(self, deck_text, overrides=(), pgen_only=False):
        L = lib()
        err = C.create_string_buffer(1024)
        ov = "\n".join(overrides).encode()
        self.h = L.orc_create_ex(deck_text.encode(), ov, int(pgen_only), err, 1024)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.L = L
        iv = (C.c_int * 16)()
        rv = (C.c_double * 12)()
        L.orc_geometry(self.h, iv, rv)
        (self.nx1, self.nx2, self.nx3, self.ng, self.ni, self.nj, self.nk, self.is_, self.ie,
         self.js, self.je, self.ks, self.ke, td, thd, _) = list(iv)
        self.two_d, self.three_d = bool(td), bool(thd)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self):
        iv = (C.c_int * 16)()
        rv = (C.c_double * 12)()
        self.L.orc_geometry(self.h, iv, rv)
        return list(iv), list(rv)

    @property
    def time(self):
        return self.geometry()[1][3]

    @property
    def cycle(self):
        return self.geometry()[0][15]

    @property
    def min_dt(self):
        return self.geometry()[1][5]

    def shape_of(self, name):
        ni, nj, nk = self.ni, self.nj, self.nk
        if name in ARRAYS_CELL4:
            return (ARRAYS_CELL4[name], nk, nj, ni)
        if name in ARRAYS_FACE4:
            return (6, nk + 1, nj + 1, ni + 1)
        if name in ARRAYS_FACE3:
            return (nk + 1, nj + 1, ni + 1)
        raise KeyError(name)

    def view(self, name):
        """A live numpy view of an oracle array (reference layout)."""
        n = C.c_long()
        p = self.L.orc_array(self.h, name.encode(), C.byref(n))
        if not p:
            raise KeyError(name)
        a = np.ctypeslib.as_array(p, shape=(n.value,))
        return a.reshape(self.shape_of(name))

    def get(self, name):
        return self.view(name).copy()

    def flag(self):
        n = C.c_long()
        p = self.L.orc_flag(self.h, C.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).reshape(self.nk, self.nj, self.ni).copy()

    def flag_view(self):
        """A live numpy view of the oracle's fofc flag array (cells, reference layout)."""
        n = C.c_long()
        p = self.L.orc_flag(self.h, C.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).reshape(self.nk, self.nj, self.ni)

    def step(self, n=1):
        return self.L.orc_step(self.h, int(n))

    def run(self):
        total = 0
        while True:
            k = self.step(1000)
            total += k
            if k < 1000:
                return total

    def l1_error(self):
        e = (C.c_double * 10)()
        self.L.orc_l1_error(self.h, e)
        return np.array(list(e))

    def fofc_count(self):
        return self.L.orc_fofc_count(self.h)

    def restart_write(self, path):
        if self.L.orc_restart_write(self.h, str(path).encode()):
            raise RuntimeError(self.L.orc_last_error().decode())

    def restart_read(self, path):
        """Restart dump into this (deck-configured) simulation, then boundaries + con2prim."""
        if self.L.orc_restart_read(self.h, str(path).encode()):
            raise RuntimeError(self.L.orc_last_error().decode())

    # single kernels
    def fluxes(self, d, cur):
        self.L.orc_fluxes(self.h, d, cur)

    def corner_emf(self):
        self.L.orc_corner_emf(self.h)

    def update(self, out, g0, g1, bdt):
        self.L.orc_update(self.h, out, g0, g1, bdt)

    def boundaries(self, s):
        self.L.orc_boundaries(self.h, s)

    def cons_to_prim(self, s, compute_dt):
        return self.L.orc_cons_to_prim(self.h, s, int(compute_dt))

    def fofc_flag(self, s):
        return self.L.orc_fofc_flag(self.h, s)

    def fofc_repair(self, d, cur):
        self.L.orc_fofc_repair(self.h, d, cur)

    def diffusion_fluxes(self, d, cur):
        self.L.orc_diffusion_fluxes(self.h, d, cur)

    def resistive_emf(self, cur):
        self.L.orc_resistive_emf(self.h, cur)

    def diffusion_dt(self):
        return self.L.orc_diffusion_dt(self.h)

    def limit_fluxes(self, d, bdt):
        return self.L.orc_limit_fluxes(self.h, d, bdt)

    def apply_gravity(self, out, bdt):
        self.L.orc_apply_gravity(self.h, out, bdt)

    def apply_cooling(self, buf, dt):
        self.L.orc_apply_cooling(self.h, buf, dt)

    @property
    def sts_count(self):
        return self.L.orc_sts_count(self.h)

    @property
    def limiter_count(self):
        return self.L.orc_limiter_count(self.h)

    def stage(self, cur, out, g0, g1, bdt):
        return self.L.orc_stage(self.h, cur, out, g0, g1, bdt)

    def set_floors(self, dfloor, pfloor):
        self.L.orc_set_floors(self.h, dfloor, pfloor)


def reconstruct(method, q):
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 6)
    n = q.shape[0]
    L = np.empty(n)
    R = np.empty(n)
    lib().orc_reconstruct(RECON[method], _dp(q), n, _dp(L), _dp(R))
    return L, R


def flatten(p5, um1, up1):
    p5 = np.ascontiguousarray(p5, dtype=np.float64).reshape(-1, 5)
    um1 = np.ascontiguousarray(um1, dtype=np.float64)
    up1 = np.ascontiguousarray(up1, dtype=np.float64)
    out = np.empty(p5.shape[0])
    lib().orc_flatten(_dp(p5), _dp(um1), _dp(up1), p5.shape[0], _dp(out))
    return out


def riemann(solver, Ls, Rs, bn, gamma=5.0 / 3.0, theta=None, phi=None):
    """solver: 'hlle' | 'hlld' | 'lhlld' (explicit theta/phi). States: (n,7) d,u,v,w,p,by,bz."""
    Ls = np.ascontiguousarray(Ls, dtype=np.float64).reshape(-1, 7)
    Rs = np.ascontiguousarray(Rs, dtype=np.float64).reshape(-1, 7)
    n = Ls.shape[0]
    bn = np.ascontiguousarray(np.broadcast_to(bn, (n,)), dtype=np.float64)
    th = np.ascontiguousarray(np.broadcast_to(1.0 if theta is None else theta, (n,)), dtype=np.float64)
    ph = np.ascontiguousarray(np.broadcast_to(1.0 if phi is None else phi, (n,)), dtype=np.float64)
    f = np.empty((n, 7))
    code = {"hlle": 0, "hlld": 1, "lhlld": 2}[solver]
    lib().orc_riemann(code, _dp(Ls), _dp(Rs), _dp(bn), n, gamma, _dp(th), _dp(ph), _dp(f))
    return f


def phys_flux(S, bn, gamma=5.0 / 3.0):
    S = np.ascontiguousarray(S, dtype=np.float64).reshape(-1, 7)
    n = S.shape[0]
    bn = np.ascontiguousarray(np.broadcast_to(bn, (n,)), dtype=np.float64)
    f = np.empty((n, 7))
    lib().orc_phys_flux(_dp(S), _dp(bn), n, gamma, _dp(f))
    return f


def fast_speed(S, bn, gamma=5.0 / 3.0):
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().orc_fast_speed(_dp(S), bn, gamma)


def eigensystem(d, p, by, bz, bx, gamma=5.0 / 3.0):
    """Left (rows) and right (columns) eigenvectors of the primitive MHD system (n, 7, 7)."""
    st = np.ascontiguousarray(np.stack(np.broadcast_arrays(d, p, by, bz, bx), axis=-1), dtype=np.float64).reshape(-1, 5)
    n = st.shape[0]
    L = np.empty((n, 7, 7))
    R = np.empty((n, 7, 7))
    lib().orc_eigensystem(_dp(st), n, gamma, _dp(L), _dp(R))
    return L, R


def charac_side(method, st, bx, side, gamma=5.0 / 3.0):
    """One face side in characteristic variables: st (n, 7, 6) stencils, bx (n,) normal field of
    the side's cell; side 0 = left state (cell = stencil point 2), 1 = right (point 3)."""
    st = np.ascontiguousarray(st, dtype=np.float64).reshape(-1, 7, 6)
    n = st.shape[0]
    bx = np.ascontiguousarray(np.broadcast_to(bx, (n,)), dtype=np.float64)
    out = np.empty((n, 7))
    lib().orc_charac_side(RECON[method], _dp(st), _dp(bx), n, gamma, 2 + side, side, _dp(out))
    return out
