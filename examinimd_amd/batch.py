"""Batched periodic LJ replicas through the C ABI (emd_batch_* in
include/emd_hip.h), plus the seeded uneven-size FCC ensemble they are
benchmarked on.

Per replica the step is the reference's serial-comm LJ step
(src/examinimd.cpp:176-294, src/integrator_nve.cpp:65-121,
src/force_types/force_lj_neigh_impl.h:158-206); the batch advances all
replicas with one launch per kernel and replays one captured HIP graph per
rebuild period.
"""
import ctypes as C

import numpy as np

import examinimd_amd as emd
from examinimd_amd.abi import EMD_FP64, EMD_MIXED, EmdBatchDesc

P = C.POINTER
CODE_SHIFT = 26
JMASK = (1 << CODE_SHIFT) - 1

_FCC_BASIS = np.array([(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)])


def _d(a):
    return a.ctypes.data_as(P(C.c_double))


def fcc(ncell, rho):
    """FCC lattice of ncell^3 cells at reduced density rho (a = (4/rho)^(1/3),
    src/input.cpp:319) -> (positions [4 ncell^3][3], box length)"""
    a = (4.0 / rho) ** (1.0 / 3.0)
    g = np.stack(np.meshgrid(*[np.arange(ncell)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = ((g[:, None, :] + _FCC_BASIS[None]) * a).reshape(-1, 3)
    return x, ncell * a


def velocities(n, temp, rng, mass=1.0):
    """seeded Gaussian velocities, zero net momentum, scaled so that
    sum m v^2 / (3n - 3) = temp (Temperature, src/property_temperature.cpp:43-62,
    k_B = mvv2e = 1)"""
    v = rng.normal(size=(n, 3)) / np.sqrt(mass)
    v -= v.mean(0)
    t = mass * (v * v).sum() / (3 * n - 3)
    return v * np.sqrt(temp / t)


def ensemble(nrep=256, seed=4, nmin=256, nmax=8192, rho=(0.70, 0.95), temp=(0.8, 2.0)):
    """nrep replicas: N drawn log-uniform in [nmin, nmax], rounded to a cubic
    FCC box of n^3 cells (n clamped to [4, 12]), density and temperature
    uniform in the given ranges.  Returns a dict of host arrays."""
    rng = np.random.default_rng(seed)
    target = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nrep))
    ncell = np.clip(np.rint((target / 4.0) ** (1.0 / 3.0)), 4, 12).astype(int)
    rhos = rng.uniform(rho[0], rho[1], nrep)
    temps = rng.uniform(temp[0], temp[1], nrep)
    xs, vs, boxes = [], [], []
    for r in range(nrep):
        x, L = fcc(int(ncell[r]), rhos[r])
        xs.append(x)
        vs.append(velocities(len(x), temps[r], rng))
        boxes.append((L, L, L))
    natoms = np.array([len(x) for x in xs], dtype=np.int32)
    return {"natoms": natoms, "box": np.array(boxes), "x": np.concatenate(xs),
            "v": np.concatenate(vs), "rho": rhos, "temp": temps, "ncell": ncell}


def _desc(natoms, box, mass, precision, eps, sigma, cut, skin, dt, rebuild_every):
    """emd_batch_desc over host arrays (returned with the arrays it points to)"""
    natoms = np.ascontiguousarray(natoms, dtype=np.int32)
    box = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(-1, 3))
    mass = np.ascontiguousarray(np.ones(len(natoms)) if mass is None else mass,
                                dtype=np.float64)
    d = EmdBatchDesc()
    d.nrep = len(natoms)
    d.natoms = natoms.ctypes.data_as(P(C.c_int))
    d.box = _d(box)
    d.mass = _d(mass)
    d.eps, d.sigma, d.cut, d.skin, d.dt = eps, sigma, cut, skin, dt
    d.rebuild_every, d.precision = rebuild_every, precision
    return d, (natoms, box, mass)


def replica_loop(natoms, box, x, v, nsteps, precision=EMD_MIXED, mass=None, eps=1.0,
                 sigma=1.0, cut=2.5, skin=0.3, dt=0.005, rebuild_every=20, ctx=None):
    """the same ensemble run replica after replica through the per-system C
    ABI (emd_replica_loop_run); returns (seconds of the nsteps steps, x, v)"""
    ctx = ctx or emd.Context()
    d, keep = _desc(natoms, box, mass, precision, eps, sigma, cut, skin, dt, rebuild_every)
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    xo, vo = np.zeros_like(x), np.zeros_like(v)
    sec = C.c_double()
    ctx.check(emd.app_lib().emd_replica_loop_run(ctx.h, C.byref(d), _d(x), _d(v), int(nsteps),
                                                 C.byref(sec), _d(xo), _d(vo)),
              "replica_loop_run")
    return sec.value, xo, vo


class Batch:
    """one emd_batch: replicas with natoms[r] atoms in boxes box[r] (3 lengths)"""

    def __init__(self, natoms, box, mass=None, precision=EMD_MIXED, eps=1.0, sigma=1.0,
                 cut=2.5, skin=0.3, dt=0.005, rebuild_every=20, ctx=None):
        self.ctx = ctx or emd.Context()
        self.L = emd.hip_lib()
        d, (self.natoms, self.box, self.mass) = _desc(natoms, box, mass, precision, eps, sigma,
                                                      cut, skin, dt, rebuild_every)
        self.total = int(self.natoms.sum())
        self.start = np.concatenate([[0], np.cumsum(self.natoms)[:-1]]).astype(np.int64)
        self.precision = precision
        h = C.c_void_p()
        self.chk(self.L.emd_batch_create(self.ctx.h, C.byref(d), C.byref(h)), "batch_create")
        self.h = h

    def chk(self, st, what):
        self.ctx.check(st, what)

    def upload(self, x, v):
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert x.shape == (self.total, 3) and v.shape == (self.total, 3)
        self.chk(self.L.emd_batch_upload(self.ctx.h, self.h, _d(x), _d(v)), "batch_upload")

    def setup(self):
        self.chk(self.L.emd_batch_setup(self.ctx.h, self.h), "batch_setup")

    def run(self, nsteps, graph=True):
        self.chk(self.L.emd_batch_run(self.ctx.h, self.h, int(nsteps), int(bool(graph))),
                 "batch_run")

    def thermo(self):
        """per replica (ke, pe): 0.5 sum m v^2 and the shifted LJ energy"""
        n = len(self.natoms)
        ke, pe = np.zeros(n), np.zeros(n)
        self.chk(self.L.emd_batch_thermo(self.ctx.h, self.h, _d(ke), _d(pe)), "batch_thermo")
        return ke, pe

    def download(self):
        x, v, f = (np.zeros((self.total, 3)) for _ in range(3))
        self.chk(self.L.emd_batch_download(self.ctx.h, self.h, _d(x), _d(v), _d(f)),
                 "batch_download")
        return x, v, f

    def set_tile(self, lanes, block):
        self.chk(self.L.emd_batch_set_tile(self.ctx.h, self.h, lanes, block), "batch_set_tile")

    def autotune(self, trial_steps=10):
        lanes, block, ms = C.c_int(), C.c_int(), C.c_double()
        self.chk(self.L.emd_batch_autotune(self.ctx.h, self.h, trial_steps, C.byref(lanes),
                                           C.byref(block), C.byref(ms)), "batch_autotune")
        return lanes.value, block.value, ms.value

    def info(self):
        step, mrow, cap = C.c_long(), C.c_int(), C.c_int()
        self.chk(self.L.emd_batch_info(self.ctx.h, self.h, C.byref(step), C.byref(mrow),
                                       C.byref(cap)), "batch_info")
        return step.value, mrow.value, cap.value

    def neigh(self):
        """(num [total], entries [total][cap]) with entries j | code << 26"""
        _, _, cap = self.info()
        num = np.zeros(self.total, np.int32)
        ent = np.zeros((self.total, cap), np.uint32)
        self.chk(self.L.emd_batch_neigh(self.ctx.h, self.h, num.ctypes.data_as(P(C.c_int)),
                                        ent.ctypes.data_as(P(C.c_uint))), "batch_neigh")
        return num, ent

    def close(self):
        if self.h:
            self.L.emd_batch_destroy(self.ctx.h, self.h)
            self.h = None
