"""Input decks through the C++ host layer: parse `block.key=value` decks (with overrides),
build the problem's initial state on the host, and start a device-resident simulation."""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import lib
from .mesh import Mesh


class DeckError(RuntimeError):
    pass


class Deck:
    def __init__(self, text, overrides=()):
        L = lib()
        h = C.c_void_p()
        rc = L.chq_deck_open(text.encode(), "\n".join(overrides).encode(), C.byref(h))
        if rc != 0:
            raise DeckError(L.chq_deck_last_error().decode())
        self.h, self.L = h, L
        self.mesh_desc = _capi.MeshDesc()
        self.config = _capi.Config()
        L.chq_deck_mesh(h, C.byref(self.mesh_desc))
        L.chq_deck_config(h, C.byref(self.config))
        tl, nl, nd = C.c_double(), C.c_int(), C.c_int()
        L.chq_deck_run(h, C.byref(tl), C.byref(nl), C.byref(nd))
        self.tlim, self.nlim, self.ndiag = tl.value, nl.value, nd.value

    @classmethod
    def from_file(cls, path, overrides=()):
        with open(path) as f:
            return cls(f.read(), overrides)

    def close(self):
        if getattr(self, "h", None):
            self.L.chq_deck_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, block, key):
        return self.L.chq_deck_get(self.h, block.encode(), key.encode()).decode()

    def text(self):
        n = self.L.chq_deck_print(self.h, None, 0)
        buf = C.create_string_buffer(n)
        self.L.chq_deck_print(self.h, buf, n)
        return buf.value.decode()

    def _shape(self, name):
        m = self.mesh_desc
        ng = m.ng
        ni = m.nx1 + 2 * ng
        nj = m.nx2 + 2 * ng if m.nx2 > 1 else 1
        nk = m.nx3 + 2 * ng if m.nx3 > 1 else 1
        if name in ("u0", "w"):
            return (6, nk, nj, ni)
        return (nk + 1, nj + 1, ni + 1)

    def initial(self, name):
        a = np.empty(self._shape(name))
        rc = self.L.chq_deck_initial(self.h, _capi.ARRAY_IDS[name], a.ctypes.data_as(C.c_void_p), a.size)
        if rc != 0:
            raise DeckError(self.L.chq_deck_last_error().decode())
        return a

    def make_mesh(self, t=0.0):
        """Device mesh with this deck's initial state, boundaries and primitives (ready to step)."""
        m = Mesh(self.mesh_desc, self.config)
        _capi.check(self.L.chq_deck_upload(self.h, m.h), "chq_deck_upload")
        m.init(t, self.tlim, self.nlim)
        return m
