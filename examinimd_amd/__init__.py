"""examinimd_amd -- MI355X (gfx950) implementation of ExaMiniMD's per-step hot path.

The compute path is native: libemd_hip.so (HIP kernels behind the C ABI in
include/emd_hip.h) and libexaminimd.so (the C++ module classes and driver
behind include/emd_app.h).  This module only loads them with ctypes; it has no
fallback.  If a library is missing every entry point raises.
"""
import ctypes as C
import os

from . import abi
from .abi import (EMD_FP64, EMD_MIXED, EMD_NEIGH_CSR, EMD_NEIGH_2D, EmdSystem, EmdHalo, EmdBins,
                  EmdNeigh, EmdLJParams, EmdSnapParams)

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")
ROOT = os.path.dirname(PKG)

_hip = None
_app = None


class NativeLibraryMissing(RuntimeError):
    pass


def hip_lib():
    """libemd_hip.so with prototypes attached (raises if it is not built)."""
    global _hip
    if _hip is None:
        path = os.path.join(LIBDIR, "libemd_hip.so")
        if not os.path.exists(path):
            raise NativeLibraryMissing(
                f"{path} is not built; run `python -m examinimd_amd.build` first")
        _hip = C.CDLL(path, mode=C.RTLD_GLOBAL)
        abi.attach_hip(_hip)
    return _hip


def app_lib():
    """libexaminimd.so (driver + module classes) with prototypes attached."""
    global _app
    if _app is None:
        hip_lib()
        path = os.path.join(LIBDIR, "libexaminimd.so")
        if not os.path.exists(path):
            raise NativeLibraryMissing(
                f"{path} is not built; run `python -m examinimd_amd.build` first")
        _app = C.CDLL(path, mode=C.RTLD_GLOBAL)
        abi.attach_app(_app)
    return _app


class EmdError(RuntimeError):
    pass


def check(ctx, status, what=""):
    if status != 0:
        msg = hip_lib().emd_last_error(ctx).decode() if ctx else ""
        raise EmdError(f"{what} failed with status {status}: {msg}")


class Context:
    """An emd_ctx (one HIP stream on one GPU)."""

    def __init__(self, device=0):
        L = hip_lib()
        self.L = L
        h = C.c_void_p()
        st = L.emd_create(device, C.byref(h))
        if st != 0:
            raise EmdError(f"emd_create({device}) failed with status {st}")
        self.h = h

    def check(self, st, what=""):
        check(self.h, st, what)

    def sync(self):
        self.check(self.L.emd_synchronize(self.h), "emd_synchronize")

    def close(self):
        if self.h:
            self.L.emd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class App:
    """The ExaMiniMD driver (C++ class ExaMiniMD) behind include/emd_app.h."""

    def __init__(self, argv, quiet=True):
        L = app_lib()
        self.L = L
        args = ["ExaMiniMD"] + [str(a) for a in argv]
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        h = C.c_void_p()
        st = L.emd_app_create(len(args), arr, 1 if quiet else 0, C.byref(h))
        if st != 0:
            raise EmdError(f"emd_app_create failed with status {st}")
        self.h = h

    def init(self):
        self.L.emd_app_init(self.h)
        return self

    def run(self, nsteps):
        self.L.emd_app_run(self.h, int(nsteps))

    def set_phase_sync(self, on):
        self.L.emd_app_set_phase_sync(self.h, 1 if on else 0)

    def set_fused_integrate(self, on):
        self.L.emd_app_set_fused_integrate(self.h, 1 if on else 0)

    @property
    def nsteps(self):
        return self.L.emd_app_nsteps(self.h)

    @property
    def natoms(self):
        return self.L.emd_app_natoms(self.h)

    @property
    def nlocal(self):
        return self.L.emd_app_nlocal(self.h)

    @property
    def nghost(self):
        return self.L.emd_app_nghost(self.h)

    def thermo(self):
        out = (C.c_double * 3)()
        self.L.emd_app_thermo(self.h, out)
        return tuple(out)

    def time(self, which=0):
        return self.L.emd_app_time(self.h, which)

    def arrays(self):
        import numpy as np
        n = self.nlocal
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        t = np.zeros(n, np.int32); i = np.zeros(n, np.int32); q = np.zeros(n)
        P = lambda a, ct: a.ctypes.data_as(C.POINTER(ct))
        st = self.L.emd_app_download(self.h, P(x, C.c_double), P(v, C.c_double), P(f, C.c_double),
                                     P(t, C.c_int), P(i, C.c_int), P(q, C.c_double))
        if st != 0:
            raise EmdError(f"emd_app_download failed with status {st}")
        return {"x": x, "v": v, "f": f, "type": t, "id": i, "q": q}

    def dump(self, path):
        return self.L.emd_app_dump(self.h, path.encode())

    def dump_neigh(self, path):
        """canonical neighbour-list file of the current list (--dumpneigh format)"""
        return self.L.emd_app_dump_neigh(self.h, path.encode())

    @property
    def ctx(self):
        return self.L.emd_app_ctx(self.h)

    def neigh(self):
        """(row_offsets, entries) of the current neighbour list as numpy arrays"""
        import numpy as np
        p = self.L.emd_app_neigh(self.h)
        if not p:
            return None
        return download_neigh(self.ctx, p.contents)

    def close(self):
        if self.h:
            self.L.emd_app_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_neigh_dump(path):
    """{id: sorted (id_j, image_j) pairs as an int array [count, 2]} from a
    --dumpneigh / emd_app_dump_neigh file"""
    import numpy as np
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    total = int(np.frombuffer(raw, np.int64, 1, 4)[0])
    body = np.frombuffer(raw, np.int32, offset=12)
    rows, k = {}, 0
    for _ in range(n):
        aid, cnt = int(body[k]), int(body[k + 1])
        rows[aid] = body[k + 2:k + 2 + 2 * cnt].reshape(cnt, 2)
        k += 2 + 2 * cnt
    assert k == len(body) and sum(len(r) for r in rows.values()) == total
    return rows


def download_ints(ctx, ptr, n):
    import numpy as np
    out = np.zeros(max(int(n), 0), np.int32)
    if n > 0:
        check(ctx, hip_lib().emd_memcpy_d2h(ctx, out.ctypes.data_as(C.c_void_p),
                                            C.cast(ptr, C.c_void_p), out.nbytes), "d2h")
    return out


def download_neigh(ctx, nl):
    """(rows, lengths, entries_per_row as a list of arrays) from an emd_neigh"""
    import numpy as np
    n = nl.n
    num = download_ints(ctx, nl.num, n)
    if nl.layout == EMD_NEIGH_CSR:
        row = download_ints(ctx, nl.row, n + 1)
        ent = download_ints(ctx, nl.entries, int(row[-1]) if n > 0 else 0)
        return row, num, ent
    pitch = nl.maxneighs
    ent2 = download_ints(ctx, nl.entries, n * pitch).reshape(n, pitch) if n else np.zeros((0, 0))
    row = np.zeros(n + 1, np.int64)
    row[1:] = np.cumsum(num)
    ent = ent2[np.arange(pitch)[None, :] < num[:, None]] if n else np.zeros(0, np.int32)
    return row, num, ent
