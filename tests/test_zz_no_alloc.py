"""GPU: chq_history_noalloc (partials in the static per-device scratch, no allocation) against
chq_history (partials in a buffer the library allocates), through the C ABI: the same kernels in
the same order, so every column is bitwise equal; on two calls in a row with the second mesh
smaller (the scratch keeps the first call's larger contents), then the first again, in FP64 and
mixed precision; a mesh beyond the scratch's 131072 active rows is refused with a message."""
import ctypes as C

import pytest

from helpers import deck

pytestmark = pytest.mark.gpu


def _mesh(name, ov, prec="fp64", steps=3):
    from chasqui_amd.deck import Deck
    d = Deck(deck(name), ov + [f"mhd.precision={prec}", "time.tlim=1e9", "time.nlim=-1"])
    m = d.make_mesh()
    d.close()
    m.step(steps)
    m.sync()
    return m


def _both(m, buf):
    from chasqui_amd import _capi
    L = _capi.lib()
    desc = _capi.MeshDesc()
    a, b = (C.c_double * 10)(), (C.c_double * 10)()
    assert L.chq_history(m.h, C.byref(desc), buf, a) == 0, L.chq_last_error()
    assert L.chq_history_noalloc(m.h, C.byref(desc), buf, b) == 0, L.chq_last_error()
    return list(a), list(b)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_history_equals_history(prec):
    print(f"case noalloc history {prec}", flush=True)
    big = _mesh("blast_3d.in", ["mesh.nx1=24", "mesh.nx2=40", "mesh.nx3=36"], prec)        # 1440 rows, 45 blocks
    small = _mesh("orszag_tang.in", ["mesh.nx1=48", "mesh.nx2=20"], prec, steps=5)         # 20 rows, 1 block
    for m in (big, small, big):
        for buf in (0, 1):
            a, b = _both(m, buf)
            assert a == b, (buf, a, b)
            assert a[0] > 0
    big.close()
    small.close()


def test_noalloc_history_refuses_meshes_beyond_the_scratch():
    """131769 active rows (2 x 363 x 363, donor cell) exceed the 131072 rows of the scratch."""
    from chasqui_amd import Mesh, _capi, make_config, make_mesh_desc
    print("case noalloc limit", flush=True)
    L = _capi.lib()
    m = Mesh(make_mesh_desc(2, 363, 363, 2), make_config(recon="dc", rsolver="hlle"))
    desc = _capi.MeshDesc()
    out = (C.c_double * 10)()
    assert L.chq_history_noalloc(m.h, C.byref(desc), 0, out) == -1
    assert b"131072" in L.chq_last_error()
    m.close()
