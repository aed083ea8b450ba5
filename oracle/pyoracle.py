"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the examinimd_amd package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class System(C.Structure):
    _fields_ = [("N", C.c_int), ("N_local", C.c_int), ("N_ghost", C.c_int),
                ("N_max", C.c_int), ("ntypes", C.c_int),
                ("x", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
                ("f", C.POINTER(C.c_double)),
                ("type", C.POINTER(C.c_int)), ("id", C.POINTER(C.c_int)),
                ("q", C.POINTER(C.c_double)),
                ("mass", C.c_double * 16), ("domain", C.c_double * 3),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("boltz", C.c_double), ("mvv2e", C.c_double), ("dt", C.c_double)]


class Halo(C.Structure):
    _fields_ = [("num_ghost", C.c_int * 6), ("pack", C.POINTER(C.c_int) * 6),
                ("cap", C.c_int * 6)]


class Bins(C.Structure):
    _fields_ = [("nbin", C.c_int * 3), ("nhalo", C.c_int),
                ("min", C.c_double * 3), ("max", C.c_double * 3),
                ("nbins_total", C.c_int), ("n_range", C.c_int),
                ("count", C.POINTER(C.c_int)), ("offset", C.POINTER(C.c_int)),
                ("permute", C.POINTER(C.c_int))]


class CSR(C.Structure):
    _fields_ = [("n", C.c_int), ("total", C.c_int),
                ("row", C.POINTER(C.c_int)), ("entries", C.POINTER(C.c_int))]


class List2D(C.Structure):
    _fields_ = [("n", C.c_int), ("maxneighs", C.c_int),
                ("num", C.POINTER(C.c_int)), ("neighs", C.POINTER(C.c_int))]


class SnapParams(C.Structure):
    _fields_ = [("twojmax", C.c_int), ("rfac0", C.c_double), ("rmin0", C.c_double),
                ("rcutfac", C.c_double), ("switchflag", C.c_int), ("bzeroflag", C.c_int),
                ("radelem", C.c_double), ("wjelem", C.c_double), ("ncoeff", C.c_int),
                ("coeff", C.c_double * 1024)]


class Config(C.Structure):
    _fields_ = [("units", C.c_int), ("lattice_style", C.c_int),
                ("lattice_arg", C.c_double), ("lattice_offset", C.c_double * 3),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("ntypes", C.c_int), ("mass", C.c_double * 16),
                ("temperature", C.c_double), ("seed", C.c_int),
                ("vel_gaussian", C.c_int), ("force_type", C.c_int),
                ("force_cutoff", C.c_double), ("lj_eps", C.c_double),
                ("lj_sigma", C.c_double), ("lj_cut", C.c_double),
                ("skin", C.c_double), ("exchange_rate", C.c_int),
                ("neigh_type", C.c_int), ("half_neigh", C.c_int), ("newton", C.c_int),
                ("dt", C.c_double), ("thermo_rate", C.c_int), ("nthreads", C.c_int),
                ("lj_intensity", C.c_double)]


UNITS = {"real": 0, "lj": 1, "metal": 2}
LATTICE = {"sc": 0, "fcc": 1, "bcc": 2}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        P = C.POINTER
        sig = {
            "or_rng_uniform": (C.c_double, [P(C.c_int)]),
            "or_rng_gaussian": (C.c_double, [P(C.c_int)]),
            "or_rng_reset": (C.c_int, [C.c_int, P(C.c_double)]),
            "or_hash_bytes": (C.c_uint, [C.c_char_p, C.c_int, C.c_uint]),
            "or_units": (None, [P(System), C.c_int]),
            "or_create_lattice": (C.c_int, [P(System), C.c_int, C.c_double, P(C.c_double),
                                            C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                            C.c_int]),
            "or_temperature": (C.c_double, [P(System)]),
            "or_kinetic_energy": (C.c_double, [P(System)]),
            "or_system_init": (None, [P(System)]),
            "or_system_free": (None, [P(System)]),
            "or_system_grow": (None, [P(System), C.c_int]),
            "or_comm_exchange": (None, [P(System)]),
            "or_comm_exchange_halo": (None, [P(System), P(Halo), C.c_double]),
            "or_comm_update_halo": (None, [P(System), P(Halo)]),
            "or_comm_update_force": (None, [P(System), P(Halo)]),
            "or_halo_free": (None, [P(Halo)]),
            "or_binning": (None, [P(System), P(Bins), C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_int, C.c_int, C.c_int]),
            "or_bins_free": (None, [P(Bins)]),
            "or_neigh_csr": (None, [P(System), P(Bins), C.c_double, C.c_int, C.c_int, P(CSR)]),
            "or_neigh_2d": (None, [P(System), P(Bins), C.c_double, C.c_int, C.c_int,
                                   P(List2D)]),
            "or_csr_free": (None, [P(CSR)]),
            "or_list2d_free": (None, [P(List2D)]),
            "or_brute_neigh_count": (None, [P(System), C.c_double, P(C.c_int)]),
            "or_lj_force": (None, [P(System), P(CSR), C.c_double, C.c_double, C.c_double,
                                   C.c_int]),
            "or_lj_energy": (C.c_double, [P(System), P(CSR), C.c_double, C.c_double,
                                          C.c_double, C.c_int]),
            "or_lj_idial_force": (None, [P(System), P(CSR), C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_int]),
            "or_nve_initial": (None, [P(System)]),
            "or_nve_final": (None, [P(System)]),
            "or_snap_ncoeff": (C.c_int, [C.c_int]),
            "or_snap_idxj": (C.c_int, [C.c_int, P(C.c_int), P(C.c_int), P(C.c_int)]),
            "or_snap_cg": (C.c_double, [C.c_int] * 6),
            "or_snap_cg_all": (None, [C.c_int, P(C.c_double)]),
            "or_snap_force": (None, [P(System), P(CSR), P(SnapParams)]),
            "or_snap_energy": (C.c_double, [P(System), P(CSR), P(SnapParams), P(C.c_double)]),
            "or_snap_bispectrum": (None, [P(SnapParams), C.c_int, P(C.c_double),
                                          P(C.c_double)]),
            "or_snap_ui": (None, [P(SnapParams), C.c_int, P(C.c_double), P(C.c_double),
                                  P(C.c_double)]),
            "or_snap_dbidrj": (None, [P(SnapParams), C.c_int, P(C.c_double), C.c_int,
                                      P(C.c_double)]),
            "or_dump_write": (C.c_int, [P(System), C.c_char_p]),
            "or_dump_compare": (C.c_int, [P(System), C.c_char_p, P(C.c_double)]),
            "or_md_create": (C.c_void_p, [P(Config), P(SnapParams)]),
            "or_md_destroy": (None, [C.c_void_p]),
            "or_md_init": (C.c_int, [C.c_void_p]),
            "or_md_run": (C.c_int, [C.c_void_p, C.c_int]),
            "or_md_system": (P(System), [C.c_void_p]),
            "or_md_csr": (P(CSR), [C.c_void_p]),
            "or_md_halo": (P(Halo), [C.c_void_p]),
            "or_md_thermo": (None, [C.c_void_p, P(C.c_double)]),
            "or_md_potential": (C.c_double, [C.c_void_p]),
            "or_md_time": (C.c_double, [C.c_void_p, C.c_int]),
            "or_md_dump": (C.c_int, [C.c_void_p, C.c_char_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def arr(ptr, n, dtype=np.float64):
    """copy n items from a ctypes pointer"""
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def system_arrays(s, nall=None):
    """numpy copies of an or_system's per-atom arrays (locals + ghosts)"""
    n = s.N_local + s.N_ghost if nall is None else nall
    return {
        "x": arr(s.x, 3 * n).reshape(n, 3), "v": arr(s.v, 3 * n).reshape(n, 3),
        "f": arr(s.f, 3 * n).reshape(n, 3), "type": arr(s.type, n, np.int32),
        "id": arr(s.id, n, np.int32), "q": arr(s.q, n),
    }


def csr_arrays(c):
    row = arr(c.row, c.n + 1, np.int32)
    ent = arr(c.entries, int(row[-1]), np.int32)
    return row, ent


def make_config(units="lj", lattice="fcc", lattice_arg=0.8442, n=(20, 20, 20), mass=1.0,
                temperature=1.44, seed=87287, gaussian=False, force="lj", cutoff=2.5,
                eps=1.0, sigma=1.0, skin=0.3, every=20, half=False, newton=False, dt=0.0,
                nthreads=0, intensity=None):
    """force: "lj", "snap", or "lj/cut/idial" (with intensity = pair_coeff's
    repeat count)"""
    cfg = Config()
    cfg.units = UNITS[units]
    cfg.lattice_style = LATTICE[lattice]
    cfg.lattice_arg = lattice_arg
    cfg.nx, cfg.ny, cfg.nz = n
    cfg.ntypes = 1
    cfg.mass[0] = mass
    cfg.temperature = temperature
    cfg.seed = seed
    cfg.vel_gaussian = 1 if gaussian else 0
    cfg.force_type = {"lj": 0, "snap": 1, "lj/cut/idial": 2}[force]
    cfg.force_cutoff = 4.73442 if force == "snap" else cutoff
    cfg.lj_intensity = 1.0 if intensity is None else float(intensity)
    cfg.lj_eps, cfg.lj_sigma, cfg.lj_cut = eps, sigma, cutoff
    cfg.skin = skin
    cfg.exchange_rate = every
    cfg.neigh_type = 0
    cfg.half_neigh = 1 if half else 0
    cfg.newton = 1 if newton else 0
    cfg.dt = dt
    cfg.thermo_rate = 0
    cfg.nthreads = nthreads
    return cfg


def snap_params(coeff, twojmax=8, rcutfac=4.73442, rfac0=0.99363, rmin0=0.0, switchflag=1,
                bzeroflag=0, radelem=0.5, wjelem=1.0):
    p = SnapParams()
    p.twojmax = twojmax
    p.rfac0, p.rmin0, p.rcutfac = rfac0, rmin0, rcutfac
    p.switchflag, p.bzeroflag = switchflag, bzeroflag
    p.radelem, p.wjelem = radelem, wjelem
    p.ncoeff = len(coeff) - 1
    for k, c in enumerate(coeff):
        p.coeff[k] = c
    return p


class MD:
    """full ExaMiniMD init + run on the CPU oracle"""

    def __init__(self, cfg, snap=None):
        L = lib()
        self.L = L
        self.cfg = cfg
        self.snap = snap
        self.h = L.or_md_create(C.byref(cfg), C.byref(snap) if snap is not None else None)
        if not self.h:
            raise RuntimeError("or_md_create failed")

    def init(self):
        return self.L.or_md_init(self.h)

    def run(self, n):
        return self.L.or_md_run(self.h, n)

    @property
    def system(self):
        return self.L.or_md_system(self.h).contents

    def arrays(self):
        return system_arrays(self.system)

    def csr(self):
        return csr_arrays(self.L.or_md_csr(self.h).contents)

    def halo(self):
        h = self.L.or_md_halo(self.h).contents
        return [arr(h.pack[p], h.num_ghost[p], np.int32) for p in range(6)]

    def thermo(self):
        out = (C.c_double * 3)()
        self.L.or_md_thermo(self.h, out)
        return tuple(out)

    def potential(self):
        return self.L.or_md_potential(self.h)

    def time(self, which=0):
        return self.L.or_md_time(self.h, which)

    def dump(self, path):
        return self.L.or_md_dump(self.h, path.encode())

    def close(self):
        if self.h:
            self.L.or_md_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
