"""The large LJ configuration (4,000,000 atoms, FCC 100^3 at density
0.8442, rc 2.5, skin 0.3, rebuild every 20 steps) on the GPU against the CPU
oracle:
  * FP64 (16 lanes per atom, the default): positions bitwise equal and
    forces within 1e-12 relative at set-up; positions, velocities and forces
    within 1e-12 after 20 steps (one rebuild);
  * neighbour rows entry for entry equal to the oracle's list at set-up, and
    after the rebuild equal to the oracle's list built from the GPU's own
    positions (src/examinimd.cpp:199-221 rebuild schedule);
  * mixed precision (FP32 pair math, FP64 state): after 100 steps of melting
    the FP32 forces are within 1e-5 RMS (reduced units) of FP64 forces
    evaluated at the same positions, and the rows built at step 100 equal the
    oracle's from the same positions.
"""
import ctypes as C
import os

import numpy as np
import pytest

import examinimd_amd as emd
from helpers import oracle_csr_from
from test_gpu_lj import compare, gpu_app

pytestmark = pytest.mark.gpu

N_CELLS = 100
BOX = (N_CELLS * (4 / 0.8442) ** (1 / 3),) * 3


def _threads():
    return max(1, min(os.cpu_count() or 1, 16))


def _oracle(oracle, steps):
    cfg = oracle.make_config(n=(N_CELLS,) * 3, gaussian=True, nthreads=_threads())
    m = oracle.MD(cfg)
    m.init()
    if steps:
        m.run(steps)
    return m


def _rows_equal(a, b):
    (row, _, ent), (orow, oent) = a, b
    return np.array_equal(row, orow) and np.array_equal(ent, oent)


def test_lj4m_fp64_setup_rows_and_first_rebuild(tmp_path, oracle):
    print("case: LJ 4M FP64 CSR, set-up and 20 steps", flush=True)
    app = gpu_app(tmp_path, 20, n=N_CELLS)
    m = _oracle(oracle, 0)
    g, o = app.arrays(), m.arrays()
    n = app.nlocal
    assert n == 4_000_000 and np.array_equal(g["id"], o["id"][:n])
    assert np.array_equal(g["x"], o["x"][:n])
    # on the perfect lattice every force is rounding noise (~1e-13): the
    # error is taken relative to a force scale of 1 (the melt's RMS is 24)
    e = float(np.max(np.abs(g["f"] - o["f"][:n]))) / max(float(np.max(np.abs(o["f"][:n]))), 1.0)
    print("set-up force error relative to max(|f|, 1)", e, flush=True)
    assert e <= 1e-12
    assert _rows_equal(app.neigh(), m.csr()), "set-up rows differ"
    app.run(20)
    a = app.arrays()
    orow, oent, oid = oracle_csr_from(oracle, a["x"], a["id"], BOX)
    assert np.array_equal(oid, a["id"])
    assert _rows_equal(app.neigh(), (orow, oent)), "rows after the step-20 rebuild differ"
    m.run(20)
    compare(app, m, 1e-12)
    app.close()
    m.close()


def test_lj4m_mixed_melt_force_error_and_rows(tmp_path, oracle):
    from examinimd_amd.abi import EMD_FP64, EmdLJParams
    print("case: LJ 4M mixed 2D, 100 steps, force error on the melt", flush=True)
    app = gpu_app(tmp_path, 100, n=N_CELLS, neigh="2D", precision="mixed")
    app.run(100)
    a = app.arrays()
    orow, oent, oid = oracle_csr_from(oracle, a["x"], a["id"], BOX)
    assert np.array_equal(oid, a["id"])
    assert _rows_equal(app.neigh(), (orow, oent)), "rows of the step-100 rebuild differ"
    # FP64 forces at the same positions with the same list
    L, ctx = emd.hip_lib(), app.ctx
    s = app.L.emd_app_system(app.h)
    nl = app.L.emd_app_neigh(app.h)
    p = EmdLJParams()
    p.ntypes = 1
    p.lj1[0][0], p.lj2[0][0], p.cutsq[0][0] = 48.0, 24.0, 6.25
    emd.check(ctx, L.emd_force_zero(ctx, s, EMD_FP64), "force_zero")
    emd.check(ctx, L.emd_lj_force(ctx, s, nl, C.byref(p), EMD_FP64, 0), "lj_force")
    n = app.nlocal
    f64 = np.zeros((n, 3))
    emd.check(ctx, L.emd_system_download(ctx, s, n, None, None,
                                         f64.ctypes.data_as(C.POINTER(C.c_double)), None, None,
                                         None), "download")
    df = a["f"] - f64
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    print("mixed force RMS error on the 4M melt", rms, "force RMS",
          float(np.sqrt(np.mean(np.sum(f64 * f64, axis=1)))), flush=True)
    assert rms <= 1e-5
    app.close()
