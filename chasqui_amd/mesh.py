"""Python face of the C ABI: one device-resident MHD mesh per `Mesh` object.

Method names follow the reference's kernel interface (physics.hpp / fofc.hpp):
fluxes = FluxesDispatch<DIR>, corner_emf = CornerEMF, update = Update,
boundaries = Boundaries, cons_to_prim = ConsToPrim, flag_bad_cells = FlagBadCells,
repair_fluxes = RepairFluxes<DIR>, stage = Stage(). Arrays move in the reference layout.
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import ARRAY_IDS, CELL_ARRAYS, FACE_MULTI, check, lib


def make_config(recon="ppm", rsolver="hlld", integrator="rk3", gamma=5.0 / 3.0, dfloor=1e-10,
                pfloor=1e-12, cfl=0.3, flatten=True, low_mach=False, fofc=True, dual_energy=False,
                de_eta=1e-3, lhlld_phi=-1.0, bc=("periodic", "periodic", "periodic"), precision="fp64",
                characteristic=False, nu=0.0, kappa=0.0, eta=0.0, sts=False, **extra):
    c = _capi.Config()
    c.recon = _capi.RECON[recon]
    c.rsolver = _capi.RSOLVER[rsolver]
    c.nstage = _capi.INTEGRATOR[integrator]
    c.gamma, c.dfloor, c.pfloor, c.cfl = gamma, dfloor, pfloor, cfl
    c.flatten, c.low_mach, c.fofc, c.dual_energy = int(flatten), int(low_mach), int(fofc), int(dual_energy)
    c.de_eta, c.lhlld_phi = de_eta, lhlld_phi
    for d in range(3):
        c.bc[d] = _capi.BC[bc[d]]
    c.precision = _capi.PRECISION[precision]
    c.characteristic = int(characteristic)
    c.nu, c.kappa, c.eta, c.sts = nu, kappa, eta, int(sts)
    c.cool_alpha = 1.0
    for k, v in extra.items():   # flux limiter, gravity and cooling fields of chq_config
        if k not in dict(_capi.Config._fields_):
            raise TypeError(f"unknown configuration field {k}")
        setattr(c, k, v)
    return c


def make_mesh_desc(nx1, nx2=1, nx3=1, ng=3, x1=(0.0, 1.0), x2=(0.0, 1.0), x3=(0.0, 1.0)):
    m = _capi.MeshDesc()
    m.nx1, m.nx2, m.nx3, m.ng = nx1, nx2, nx3, ng
    m.x1min, m.x1max = x1
    m.x2min, m.x2max = x2
    m.x3min, m.x3max = x3
    return m


class Mesh:
    def __init__(self, mesh_desc, config):
        L = lib()
        h = C.c_void_p()
        check(L.chq_create(C.byref(mesh_desc), C.byref(config), C.byref(h)), "chq_create")
        self._init_handle(L, h, owned=True)

    @classmethod
    def borrow(cls, handle):
        """A view of a context owned elsewhere (a served run): closing it does not destroy it."""
        m = cls.__new__(cls)
        m._init_handle(lib(), C.c_void_p(handle), owned=False)
        return m

    def _init_handle(self, L, h, owned):
        self.h = h
        self.L = L
        self.owned = owned
        iv = (C.c_int * 16)()
        dx = (C.c_double * 3)()
        check(L.chq_geometry(self.h, iv, dx), "chq_geometry")
        (self.nx1, self.nx2, self.nx3, self.ng, self.ni, self.nj, self.nk, self.is_, self.ie,
         self.js, self.je, self.ks, self.ke, td, thd, self.pitch) = list(iv)
        self.two_d, self.three_d = bool(td), bool(thd)
        self.dx = tuple(dx)

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                self.L.chq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- arrays
    def shape_of(self, name):
        if name in CELL_ARRAYS:
            n = CELL_ARRAYS[name]
            s = (self.nk, self.nj, self.ni)
            return s if name == "flag" else (n,) + s
        s = (self.nk + 1, self.nj + 1, self.ni + 1)
        return (FACE_MULTI[name],) + s if name in FACE_MULTI else s

    def upload(self, name, arr):
        dtype = np.int32 if name == "flag" else np.float64
        a = np.ascontiguousarray(arr, dtype=dtype).reshape(self.shape_of(name))
        check(self.L.chq_upload(self.h, ARRAY_IDS[name], a.ctypes.data_as(C.c_void_p), a.size), f"upload {name}")

    def download(self, name):
        dtype = np.int32 if name == "flag" else np.float64
        a = np.empty(self.shape_of(name), dtype=dtype)
        check(self.L.chq_download(self.h, ARRAY_IDS[name], a.ctypes.data_as(C.c_void_p), a.size),
              f"download {name}")
        return a

    def device_bytes(self):
        return self.L.chq_device_bytes(self.h)

    # ---- kernels with the reference semantics
    def fluxes(self, d, cur):
        check(self.L.chq_fluxes(self.h, d, cur), "chq_fluxes")

    def corner_emf(self):
        check(self.L.chq_corner_emf(self.h), "chq_corner_emf")

    def update(self, out, g0, g1, beta_dt):
        check(self.L.chq_update(self.h, out, g0, g1, beta_dt), "chq_update")

    def boundaries(self, buf):
        check(self.L.chq_boundaries(self.h, buf), "chq_boundaries")

    def cons_to_prim(self, buf, compute_dt):
        v = C.c_double(0.0)
        check(self.L.chq_cons_to_prim(self.h, buf, int(compute_dt), C.byref(v)), "chq_cons_to_prim")
        return v.value

    def flag_bad_cells(self, buf):
        n = C.c_int(0)
        check(self.L.chq_fofc_flag(self.h, buf, C.byref(n)), "chq_fofc_flag")
        return n.value

    def repair_fluxes(self, d, cur):
        check(self.L.chq_fofc_repair(self.h, d, cur), "chq_fofc_repair")

    def diffusion_fluxes(self, d, cur):
        check(self.L.chq_diffusion_fluxes(self.h, d, cur), "chq_diffusion_fluxes")

    def resistive_emf(self, cur):
        check(self.L.chq_resistive_emf(self.h, cur), "chq_resistive_emf")

    def diffusion_dt(self):
        v = C.c_double(0.0)
        check(self.L.chq_diffusion_dt(self.h, C.byref(v)), "chq_diffusion_dt")
        return v.value

    def limit_fluxes(self, d, beta_dt):
        n = C.c_long(0)
        check(self.L.chq_limit_fluxes(self.h, d, beta_dt, C.byref(n)), "chq_limit_fluxes")
        return n.value

    def apply_gravity(self, out, beta_dt):
        check(self.L.chq_apply_gravity(self.h, out, beta_dt), "chq_apply_gravity")

    def apply_cooling(self, buf, dt):
        check(self.L.chq_apply_cooling(self.h, buf, dt), "chq_apply_cooling")

    def stage(self, cur, out, g0, g1, beta_dt):
        n = C.c_long(0)
        check(self.L.chq_stage(self.h, cur, out, g0, g1, beta_dt, C.byref(n)), "chq_stage")
        return n.value

    # ---- device-resident time loop
    def init(self, t=0.0, tlim=1.0, nlim=-1, cycle=0):
        check(self.L.chq_init_resume(self.h, t, int(cycle), tlim, nlim), "chq_init")

    def restart_write(self, path, t, cycle):
        """The reference's binary restart dump (restart.hpp) of u0 and the face fields."""
        check(self.L.chq_restart_write(self.h, str(path).encode(), t, int(cycle)), "chq_restart_write")

    def restart_read(self, path):
        """Load a restart dump into u0 / b?f0; returns (t, cycle) for init(t, tlim, nlim, cycle)."""
        t, cyc = C.c_double(), C.c_int()
        check(self.L.chq_restart_read(self.h, str(path).encode(), C.byref(t), C.byref(cyc)), "chq_restart_read")
        return t.value, cyc.value

    def step(self, n=1):
        check(self.L.chq_step(self.h, int(n)), "chq_step")

    def step_timed(self, n):
        ms = C.c_double(0.0)
        check(self.L.chq_step_timed(self.h, int(n), C.byref(ms)), "chq_step_timed")
        return ms.value

    def kernel_timing(self, kernel_class):
        """Select a kernel class ('flux', 'corner', 'update', 'bc', 'c2p', 'repair' (the whole
        fofc branch), 'all', or None) whose launches inside step() are bracketed by HIP events."""
        code = -1 if kernel_class is None else _capi.KERNEL_CLASSES[kernel_class]
        check(self.L.chq_kernel_timing(self.h, code), "chq_kernel_timing")

    def kernel_timing_read(self):
        ms, n = C.c_double(0.0), C.c_int(0)
        check(self.L.chq_kernel_timing_read(self.h, C.byref(ms), C.byref(n)), "chq_kernel_timing_read")
        return ms.value, n.value

    def history(self, buf=0):
        """The .hst columns of set `buf` (output.hpp:91-119), reduced on the device: mass, mom1,
        mom2, mom3, etot, ekin, emag, eint, divB_max, divB_L1."""
        out = (C.c_double * 10)()
        desc = _capi.MeshDesc()   # the geometry comes from the context
        check(self.L.chq_history(self.h, C.byref(desc), int(buf), out), "chq_history")
        return list(out)

    def kernel_timing_breakdown(self):
        """{class: (ms, launches)} since kernel_timing('all') (then starts over)."""
        n = len(_capi.CLASS_NAMES)
        ms, k = (C.c_double * n)(), (C.c_int * n)()
        check(self.L.chq_kernel_timing_breakdown(self.h, ms, k), "chq_kernel_timing_breakdown")
        return {c: (ms[q], k[q]) for q, c in enumerate(_capi.CLASS_NAMES)}

    def status(self):
        s = _capi.Status()
        check(self.L.chq_get_status(self.h, C.byref(s)), "chq_get_status")
        return {"t": s.t, "dt": s.dt, "tlim": s.tlim, "min_dt": s.min_dt, "cycle": s.cycle,
                "nlim": s.nlim, "active": s.active, "nbad_last": s.nbad_last,
                "fofc_total": s.fofc_total, "sts_stages": s.sts_stages, "limiter_total": s.limiter_total}

    def sync(self):
        check(self.L.chq_sync(self.h), "chq_sync")

    def run(self, chunk=64):
        """Step until t reaches tlim (or nlim), checking the device clock every `chunk` steps."""
        while True:
            before = self.status()["cycle"]
            self.step(chunk)
            st = self.status()
            if st["cycle"] - before < chunk:
                return st
