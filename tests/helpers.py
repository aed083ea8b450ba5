"""shared helpers for the parity tests"""
import os

import numpy as np

from conftest import ROOT


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0)) if a.size else 0.0


def sorted_rows(row, ent, key=None):
    """per-row sorted neighbour lists (optionally mapped through key)"""
    out = []
    for i in range(len(row) - 1):
        r = ent[row[i]:row[i + 1]]
        if key is not None:
            r = key[r]
        out.append(np.sort(r))
    return out


def oracle_lj_md(oracle, x, v, box, nsteps, dt=0.005, eps=1.0, sig=1.0, rc=2.5, skin=0.3,
                 every=20, mass=1.0):
    """the reference's serial-comm LJ loop (src/examinimd.cpp:176-294) on the
    CPU oracle for one periodic box, from given positions and velocities:
    set-up force, then per step NVE initial, rebuild (step % every == 0) or
    halo update, full-list force, NVE final.  Returns x, v, f of the local
    atoms in input order."""
    import ctypes as C
    L = oracle.lib()
    s = oracle.System()
    L.or_system_init(C.byref(s))
    n = len(x)
    L.or_system_grow(C.byref(s), n)
    s.N_local = s.N = n
    s.ntypes = 1
    s.mass[0] = mass
    s.dt, s.mvv2e, s.boltz = dt, 1.0, 1.0
    for d in range(3):
        s.domain[d] = s.hi[d] = box[d]
        s.lo[d] = 0.0
    np.ctypeslib.as_array(s.x, shape=(3 * n,))[:] = np.asarray(x, dtype=np.float64).ravel()
    np.ctypeslib.as_array(s.v, shape=(3 * n,))[:] = np.asarray(v, dtype=np.float64).ravel()
    np.ctypeslib.as_array(s.id, shape=(n,))[:] = np.arange(1, n + 1)
    np.ctypeslib.as_array(s.type, shape=(n,))[:] = 0
    cut = rc + skin

    def build():
        h, b, c = oracle.Halo(), oracle.Bins(), oracle.CSR()
        L.or_comm_exchange(C.byref(s))
        L.or_binning(C.byref(s), C.byref(b), cut, cut, cut, 1, 1, 0, 1)
        L.or_comm_exchange_halo(C.byref(s), C.byref(h), cut)
        L.or_binning(C.byref(s), C.byref(b), cut, cut, cut, 1, 1, 1, 0)
        L.or_neigh_csr(C.byref(s), C.byref(b), cut, 0, 0, C.byref(c))
        return h, b, c

    def force(c):
        np.ctypeslib.as_array(s.f, shape=(3 * s.N_max,))[:] = 0
        L.or_lj_force(C.byref(s), C.byref(c), eps, sig, rc, 0)

    h, b, c = build()
    force(c)
    for step in range(1, nsteps + 1):
        L.or_nve_initial(C.byref(s))
        if step % every == 0:
            L.or_halo_free(C.byref(h))
            L.or_bins_free(C.byref(b))
            L.or_csr_free(C.byref(c))
            h, b, c = build()
        else:
            L.or_comm_update_halo(C.byref(s), C.byref(h))
        force(c)
        L.or_nve_final(C.byref(s))
    ids = oracle.arr(s.id, n, np.int32) - 1
    out = []
    for p in (s.x, s.v, s.f):
        a = np.zeros((n, 3))
        a[ids] = oracle.arr(p, 3 * n).reshape(n, 3)
        out.append(a)
    L.or_system_free(C.byref(s))
    L.or_halo_free(C.byref(h))
    L.or_bins_free(C.byref(b))
    L.or_csr_free(C.byref(c))
    return tuple(out)


def oracle_csr_from(oracle, x, ids, box, cut=2.8):
    """the oracle's full CSR list (exchange, binning sort, halo, binning,
    NeighborCSR) for local positions given in the caller's order"""
    import ctypes as C
    L = oracle.lib()
    s = oracle.System()
    L.or_system_init(C.byref(s))
    n = len(x)
    L.or_system_grow(C.byref(s), n)
    s.N_local = s.N = n
    s.ntypes = 1
    s.mass[0] = 1.0
    for d in range(3):
        s.domain[d] = s.hi[d] = box[d]
        s.lo[d] = 0.0
    np.ctypeslib.as_array(s.x, shape=(3 * n,))[:] = np.asarray(x, dtype=np.float64).ravel()
    np.ctypeslib.as_array(s.id, shape=(n,))[:] = ids
    h, b, c = oracle.Halo(), oracle.Bins(), oracle.CSR()
    L.or_comm_exchange(C.byref(s))
    L.or_binning(C.byref(s), C.byref(b), cut, cut, cut, 1, 1, 0, 1)
    L.or_comm_exchange_halo(C.byref(s), C.byref(h), cut)
    L.or_binning(C.byref(s), C.byref(b), cut, cut, cut, 1, 1, 1, 0)
    L.or_neigh_csr(C.byref(s), C.byref(b), cut, 0, 0, C.byref(c))
    row, ent = oracle.csr_arrays(c)
    out_ids = oracle.arr(s.id, n, np.int32)
    for f, o in ((L.or_system_free, s), (L.or_halo_free, h), (L.or_bins_free, b),
                 (L.or_csr_free, c)):
        f(C.byref(o))
    return row, ent, out_ids


def pair_forces_typed(x, box, types, lj1, lj2, cutsq):
    """brute-force minimum-image LJ forces with per-type-pair coefficients
    and the strict cutoff test rsq < cutsq[ti][tj]
    (force_lj_neigh_impl.h:161-206)"""
    d = x[:, None, :] - x[None, :, :]
    d -= box * np.round(d / box)
    rsq = np.sum(d * d, axis=2)
    ti, tj = types[:, None], types[None, :]
    m = (rsq < cutsq[ti, tj]) & ~np.eye(len(x), dtype=bool)
    r2inv = np.where(m, 1.0 / np.where(m, rsq, 1.0), 0.0)
    r6inv = r2inv ** 3
    fpair = r6inv * (lj1[ti, tj] * r6inv - lj2[ti, tj]) * r2inv
    return np.sum(d * fpair[:, :, None], axis=1)

