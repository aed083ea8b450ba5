"""Kernel-level driver: one force evaluation of an arbitrary
configuration through the C ABI (wrap, bin + sort, halo, bin, neighbour
list, force, reverse fold), results returned in input order."""
import ctypes as C

import numpy as np

import examinimd_amd as emd
from examinimd_amd.abi import (EMD_FP64, EMD_MIXED, EMD_NEIGH_CSR, EmdBins, EmdHalo,
                               EmdLJIDialParams, EmdLJParams, EmdNeigh, EmdSnapParams, EmdSystem)

P = C.POINTER


def _d(a):
    return a.ctypes.data_as(P(C.c_double))


def _i(a):
    return a.ctypes.data_as(P(C.c_int))


class Pipeline:
    def __init__(self, ctx=None):
        self.ctx = ctx or emd.Context()
        self.L = emd.hip_lib()
        self.s = EmdSystem()
        self.h = EmdHalo()
        self.b = EmdBins()
        self.nl = EmdNeigh()
        self.snap = None

    def chk(self, st, what):
        self.ctx.check(st, what)

    def load(self, x, box, mass=1.0, v=None, types=None):
        """types: 0-based type per atom (default all 0); every type gets
        the same mass"""
        self._ids = None
        n = len(x)
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float64)
        ids = np.arange(1, n + 1, dtype=np.int32)
        types = (np.zeros(n, np.int32) if types is None
                 else np.ascontiguousarray(types, dtype=np.int32))
        ntypes = int(types.max()) + 1 if n else 1
        if types.shape != (n,) or (n and types.min() < 0) or ntypes > 1024:
            raise ValueError("types: one value in [0, 1024) per atom")
        q = np.zeros(n)
        s = self.s
        self.chk(self.L.emd_system_upload(self.ctx.h, C.byref(s), n, _d(x), _d(v), _i(types),
                                          _i(ids), _d(q)), "upload")
        m = np.full(ntypes, mass, dtype=np.float64)
        self.chk(self.L.emd_system_set_mass(self.ctx.h, C.byref(s), ntypes, _d(m)), "mass")
        s.n_local = n
        s.n_ghost = 0
        s.n_global = n
        for d in range(3):
            s.domain[d] = box[d]
            s.lo[d] = 0.0
            s.hi[d] = box[d]

    def build(self, cut, half=0, newton=0, layout=EMD_NEIGH_CSR):
        """exchange, binning sort, halo, binning and the neighbour list;
        layout: EMD_NEIGH_CSR or EMD_NEIGH_2D"""
        L, c, s = self.L, self.ctx.h, C.byref(self.s)
        self._ids = None
        self.chk(L.emd_comm_exchange(c, s), "exchange")
        self.chk(L.emd_binning(c, s, cut, cut, cut, 1, 1, 0, 1, C.byref(self.b)), "bin sort")
        self.chk(L.emd_comm_exchange_halo(c, s, cut, C.byref(self.h)), "halo")
        self.chk(L.emd_binning(c, s, cut, cut, cut, 1, 1, 1, 0, C.byref(self.b)), "bin")
        self.chk(L.emd_neigh_build(c, s, C.byref(self.b), cut, half, newton, layout,
                                   C.byref(self.nl)), "neigh")

    def lj(self, precision=EMD_FP64, eps=1.0, sigma=1.0, rc=2.5):
        p = EmdLJParams()
        p.ntypes = 1
        p.lj1[0][0] = 48.0 * eps * sigma ** 12
        p.lj2[0][0] = 24.0 * eps * sigma ** 6
        p.cutsq[0][0] = rc * rc
        L, c, s = self.L, self.ctx.h, C.byref(self.s)
        self.chk(L.emd_force_zero(c, s, precision), "zero")
        self.chk(L.emd_lj_force(c, s, C.byref(self.nl), C.byref(p), precision, 0), "lj")
        return self.forces(precision)

    def lj_idial(self, intensity, precision=EMD_FP64, eps=1.0, sigma=1.0, rc=2.5, half=0):
        """pair_style lj/cut/idial forces of the local atoms (input order) on
        the list of the last build()"""
        p = EmdLJIDialParams()
        p.lj.ntypes = 1
        p.lj.lj1[0][0] = 48.0 * eps * sigma ** 12
        p.lj.lj2[0][0] = 24.0 * eps * sigma ** 6
        p.lj.cutsq[0][0] = rc * rc
        p.intensity[0][0] = intensity
        L, c, s = self.L, self.ctx.h, C.byref(self.s)
        self.chk(L.emd_force_zero(c, s, precision), "zero")
        self.chk(L.emd_lj_idial_force(c, s, C.byref(self.nl), C.byref(p), precision, half),
                 "lj_idial")
        return self.forces(precision)

    def snap_setup(self, beta, twojmax=8, rcutfac=4.73442, rfac0=0.99363, rmin0=0.0,
                   switchflag=1, bzeroflag=0, radelem=0.5, wjelem=1.0):
        self._beta = np.ascontiguousarray(beta, dtype=np.float64)
        p = EmdSnapParams()
        p.twojmax, p.rfac0, p.rmin0, p.rcutfac = twojmax, rfac0, rmin0, rcutfac
        p.switchflag, p.bzeroflag, p.radelem, p.wjelem = switchflag, bzeroflag, radelem, wjelem
        p.ncoeff = len(beta) - 1
        p.coeff = _d(self._beta)
        h = C.c_void_p()
        self.chk(self.L.emd_snap_create(self.ctx.h, C.byref(p), C.byref(h)), "snap_create")
        self.snap = h
        self._snap_owned = True

    def snap_use(self, other):
        """use another Pipeline's SNAP object (same context): its tables and
        coefficients; the owner destroys it"""
        self.snap = other.snap
        self._beta = other._beta
        self._snap_owned = False

    def snap_force(self, precision=EMD_FP64):
        L, c, s = self.L, self.ctx.h, C.byref(self.s)
        self.chk(L.emd_force_zero(c, s, precision), "zero")
        self.chk(L.emd_snap_force(c, s, C.byref(self.nl), self.snap, precision), "snap")
        self.chk(L.emd_comm_update_force(c, s, C.byref(self.h), precision), "fold")
        return self.forces(precision)

    def snap_energy(self):
        e = C.c_double()
        self.chk(self.L.emd_snap_energy(self.ctx.h, C.byref(self.s), C.byref(self.nl), self.snap,
                                        C.byref(e), None), "snap_energy")
        return e.value

    def snap_set_coeff(self, beta):
        self._beta = np.ascontiguousarray(beta, dtype=np.float64)
        self.chk(self.L.emd_snap_set_coeff(self.ctx.h, self.snap, _d(self._beta)), "set_coeff")

    def _cotangent(self, gF):
        """upload dL/dF given in input order into the device atom order"""
        ids = self._device_ids()
        g = np.ascontiguousarray(np.asarray(gF, dtype=np.float64)[ids - 1])
        if getattr(self, "_gdev", None) is None or self._gcap < g.nbytes:
            if getattr(self, "_gdev", None) is not None:
                self.L.emd_free(self.ctx.h, self._gdev)
            self._gdev = C.c_void_p()
            self.chk(self.L.emd_malloc(self.ctx.h, C.byref(self._gdev), g.nbytes), "malloc")
            self._gcap = g.nbytes
        self.chk(self.L.emd_memcpy_h2d(self.ctx.h, self._gdev, g.ctypes.data_as(C.c_void_p),
                                       g.nbytes), "h2d")
        return C.cast(self._gdev, P(C.c_double))

    def snap_adjoint(self, gF, gE, precision=EMD_FP64):
        """sum_i gF_i . dF_i/dbeta + gE dE/dbeta (host array, ncoeff+1)"""
        out = np.zeros(len(self._beta))
        g = self._cotangent(gF)
        self.chk(self.L.emd_snap_adjoint_beta(self.ctx.h, C.byref(self.s), C.byref(self.nl),
                                              C.byref(self.h), self.snap, g, float(gE),
                                              precision, _d(out)), "snap_adjoint")
        return out

    def lj_adjoint(self, gF, gE, eps=1.0, sigma=1.0, rc=2.5, precision=EMD_FP64):
        """(dL/deps, dL/dsigma) for cotangents gF (input order), gE"""
        out = np.zeros(2)
        g = self._cotangent(gF)
        self.chk(self.L.emd_lj_adjoint(self.ctx.h, C.byref(self.s), C.byref(self.nl), eps, sigma,
                                       rc, g, float(gE), precision, _d(out)), "lj_adjoint")
        return out

    def lj_energy(self, eps=1.0, sigma=1.0, rc=2.5):
        p = EmdLJParams()
        p.ntypes = 1
        p.lj1[0][0] = 48.0 * eps * sigma ** 12
        p.lj2[0][0] = 24.0 * eps * sigma ** 6
        p.cutsq[0][0] = rc * rc
        e = C.c_double()
        self.chk(self.L.emd_lj_energy(self.ctx.h, C.byref(self.s), C.byref(self.nl), C.byref(p), 0,
                                      C.byref(e)), "lj_energy")
        return e.value

    def _device_ids(self):
        """atom ids in device order (cached until the next load or build:
        only those reorder the atoms)"""
        if getattr(self, "_ids", None) is None:
            n = self.s.n_local
            ids = np.zeros(n, np.int32)
            self.chk(self.L.emd_system_download(self.ctx.h, C.byref(self.s), n, None, None, None,
                                                None, _i(ids), None), "ids")
            self._ids = ids
        return self._ids

    def forces(self, precision):
        n = self.s.n_local
        f = np.zeros((n, 3))
        L, c = self.L, self.ctx.h
        if precision == EMD_MIXED:
            f32 = np.zeros((n, 3), np.float32)
            self.chk(L.emd_memcpy_d2h(c, f32.ctypes.data_as(C.c_void_p),
                                      C.cast(self.s.f32, C.c_void_p), f32.nbytes), "f32")
            f[:] = f32
        else:
            self.chk(L.emd_system_download(c, C.byref(self.s), n, None, None, _d(f), None, None,
                                           None), "f")
        ids = self._device_ids()
        out = np.zeros_like(f)
        out[ids - 1] = f
        return out

    def close(self):
        L, c = self.L, self.ctx.h
        if self.snap:
            if getattr(self, "_snap_owned", True):
                L.emd_snap_destroy(c, self.snap)
            self.snap = None
        if getattr(self, "_gdev", None) is not None:
            L.emd_free(c, self._gdev)
            self._gdev = None
        L.emd_neigh_free(c, C.byref(self.nl))
        L.emd_bins_free(c, C.byref(self.b))
        L.emd_halo_free(c, C.byref(self.h))
        L.emd_system_free(c, C.byref(self.s))


def lattice(kind, n, a):
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(float)
    if kind == "fcc":
        basis = [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)]
    elif kind == "bcc":
        basis = [(0, 0, 0), (0.5, 0.5, 0.5)]
    else:
        basis = [(0, 0, 0)]
    return np.concatenate([(g + np.array(b)) * a for b in basis])
