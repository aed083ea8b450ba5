"""The boundary in the reference's own View layouts (src/types.h:88-100):
x LayoutRight, v and f LayoutLeft with a leading dimension (extent(0)) above
the atom count, Binning's bincount/binoffsets LayoutLeft View<int***>
(src/binning.h:54-59).  A system imported from those layouts gives bitwise
the forces, velocities and neighbour rows of the same system uploaded in
this library's AoS layout, and exports back bit for bit."""
import ctypes as C

import numpy as np
import pytest

import examinimd_amd as emd
from examinimd_amd.abi import (EMD_FP64, EMD_LAYOUT_LEFT, EMD_LAYOUT_RIGHT, EmdBins, EmdNeigh,
                               EmdVec3View)
from examinimd_amd.pipeline import Pipeline, lattice

pytestmark = pytest.mark.gpu


class Dev:
    """device buffers of one context, freed together"""

    def __init__(self, pl):
        self.pl, self.bufs = pl, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.pl.chk(self.pl.L.emd_malloc(self.pl.ctx.h, C.byref(p), max(a.nbytes, 8)), "malloc")
        self.pl.chk(self.pl.L.emd_memcpy_h2d(self.pl.ctx.h, p, a.ctypes.data_as(C.c_void_p),
                                             a.nbytes), "h2d")
        self.bufs.append(p)
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        self.pl.chk(self.pl.L.emd_memcpy_d2h(self.pl.ctx.h, out.ctypes.data_as(C.c_void_p), p,
                                             out.nbytes), "d2h")
        return out

    def free(self):
        for p in self.bufs:
            self.pl.L.emd_free(self.pl.ctx.h, p)


def view(p, layout, ld=0):
    return EmdVec3View(C.cast(p, C.POINTER(C.c_double)) if p else None, layout, ld)


def left(a, ld):
    """a [n][3] array stored LayoutLeft with leading dimension ld"""
    out = np.full(3 * ld, np.nan)
    for d in range(3):
        out[d * ld:d * ld + len(a)] = a[:, d]
    return out


def test_layout_left_import_export_bitwise():
    print("case: LayoutLeft v/f import and export, FCC 15^3 displaced", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    nc = 15
    rng = np.random.default_rng(21)
    x = lattice("fcc", nc, a) + rng.normal(0, 0.05, (4 * nc ** 3, 3))
    v = rng.normal(0, 1.0, x.shape)
    n, box = len(x), (nc * a,) * 3
    ref = Pipeline()
    ref.load(x, box, v=v)
    ref.build(2.8)
    f_ref = ref.lj(EMD_FP64)

    pl = Pipeline(ref.ctx)
    dev = Dev(pl)
    ld = n + 77                       # extent(0) of the reference's Views: N_max > n
    xd = dev.put(x)                   # LayoutRight
    vd = dev.put(left(v, ld))
    fd = dev.put(np.full(3 * ld, np.nan))
    types, ids, q = np.zeros(n, np.int32), np.arange(1, n + 1, dtype=np.int32), np.zeros(n)
    td, idd, qd = dev.put(types), dev.put(ids), dev.put(q)
    pl.chk(pl.L.emd_system_import(pl.ctx.h, C.byref(pl.s), n, view(xd, EMD_LAYOUT_RIGHT),
                                  view(vd, EMD_LAYOUT_LEFT, ld), view(None, EMD_LAYOUT_LEFT),
                                  td, idd, qd), "import")
    m = np.ones(1)
    pl.chk(pl.L.emd_system_set_mass(pl.ctx.h, C.byref(pl.s), 1, m.ctypes.data_as(
        C.POINTER(C.c_double))), "mass")
    pl.s.n_local = pl.s.n_global = n
    pl.s.n_ghost = 0
    for d in range(3):
        pl.s.domain[d] = pl.s.hi[d] = box[d]
        pl.s.lo[d] = 0.0
    pl.build(2.8)
    f_imp = pl.lj(EMD_FP64)
    assert np.array_equal(f_imp, f_ref)
    # export the (sorted, wrapped) state into the Views and read it back
    xo = dev.put(np.full(3 * n, np.nan))
    pl.chk(pl.L.emd_system_export(pl.ctx.h, C.byref(pl.s), n, view(xo, EMD_LAYOUT_RIGHT),
                                  view(vd, EMD_LAYOUT_LEFT, ld), view(fd, EMD_LAYOUT_LEFT, ld)),
           "export")
    xs, vs, fs = (np.zeros((n, 3)) for _ in range(3))
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    pl.chk(pl.L.emd_system_download(pl.ctx.h, C.byref(pl.s), n, P(xs), P(vs), P(fs), None, None,
                                    None), "download")
    assert np.array_equal(dev.get(xo, (n, 3), np.float64), xs)
    vl, fl = dev.get(vd, 3 * ld, np.float64), dev.get(fd, 3 * ld, np.float64)
    assert np.array_equal(np.stack([vl[d * ld:d * ld + n] for d in range(3)], 1), vs)
    assert np.array_equal(np.stack([fl[d * ld:d * ld + n] for d in range(3)], 1), fs)
    assert np.all(np.isnan(fl[n:ld]))          # padding rows untouched
    dev.free()
    pl.close()
    ref.close()


def test_bins_import_layout_left_gives_same_rows():
    print("case: Binning fields imported LayoutLeft, same neighbour rows", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    nc = 12
    x = lattice("fcc", nc, a) + np.random.default_rng(4).normal(0, 0.05, (4 * nc ** 3, 3))
    pl = Pipeline()
    pl.load(x, (nc * a,) * 3)
    pl.build(2.8)
    b, L, c = pl.b, pl.L, pl.ctx.h
    dev = Dev(pl)
    nb = (b.nbinx, b.nbiny, b.nbinz)
    cnt = dev.get(b.bincount, nb, np.int32)          # x-major = C order of [nbx][nby][nbz]
    off = dev.get(b.binoffsets, nb, np.int32)
    perm = dev.get(b.permute, b.n_range, np.int32)
    cl, ol = dev.put(cnt.ravel(order="F")), dev.put(off.ravel(order="F"))   # LayoutLeft
    pd = dev.put(perm)
    lo = (C.c_double * 3)(b.minx, b.miny, b.minz)
    hi = (C.c_double * 3)(b.maxx, b.maxy, b.maxz)
    b2, nl2 = EmdBins(), EmdNeigh()
    pl.chk(L.emd_bins_import(c, nb[0], nb[1], nb[2], b.nhalo, lo, hi, cl, ol, EMD_LAYOUT_LEFT, pd,
                             b.n_range, C.byref(b2)), "bins_import")
    assert np.array_equal(dev.get(b2.bincount, nb, np.int32), cnt)
    assert b2.max_count == cnt.max()
    pl.chk(L.emd_neigh_build(c, C.byref(pl.s), C.byref(b2), 2.8, 0, 0, emd.EMD_NEIGH_CSR,
                             C.byref(nl2)), "neigh")
    r1 = emd.download_neigh(c, pl.nl)
    r2 = emd.download_neigh(c, nl2)
    assert all(np.array_equal(u, w) for u, w in zip(r1, r2))
    L.emd_neigh_free(c, C.byref(nl2))
    L.emd_bins_free(c, C.byref(b2))
    dev.free()
    pl.close()
