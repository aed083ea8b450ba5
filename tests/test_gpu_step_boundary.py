"""emd_nve_step_boundary: the driver's step boundary (ghost-force fold of
step s, final_integrate of s, initial_integrate of s+1, halo position update
of s+1; src/examinimd.cpp:243, 249, 196, 226) as one workgroup launch for
small systems.  From the same device state (seeded forces on locals and
ghosts, seeded velocities) it must give the bits of emd_comm_update_force,
emd_nve_final_initial and emd_comm_update_halo called in turn: positions of
locals and ghosts, velocities and forces.  BCC 10^3 W with a 5.73 A halo has
~3 000 ghosts (one launch); 30^3 has ~22 000 (the separate launches)."""
import ctypes as C

import numpy as np
import pytest

import examinimd_amd as emd
from examinimd_amd.pipeline import Pipeline, lattice

pytestmark = pytest.mark.gpu

A_W, CUT = 3.1803, 5.73442
DTV = 0.0005
DTF = 0.5 * DTV / 1.0364269e-4


def _get(pl, ptr, n, dt):
    a = np.zeros(n, dt)
    if n:
        pl.chk(pl.L.emd_memcpy_d2h(pl.ctx.h, a.ctypes.data_as(C.c_void_p),
                                   C.cast(ptr, C.c_void_p), a.nbytes), "d2h")
    return a


def _put(pl, ptr, a):
    pl.chk(pl.L.emd_memcpy_h2d(pl.ctx.h, C.cast(ptr, C.c_void_p), a.ctypes.data_as(C.c_void_p),
                               a.nbytes), "h2d")


@pytest.mark.parametrize("nc", [10, 30])
@pytest.mark.parametrize("precision", [emd.EMD_FP64, emd.EMD_MIXED])
@pytest.mark.parametrize("fold,halo", [(1, 1), (1, 0), (0, 1)])
def test_step_boundary_equals_separate_launches(nc, precision, fold, halo):
    print(f"case: step boundary, BCC {nc}^3, precision {precision}, fold {fold}, halo {halo}",
          flush=True)
    rng = np.random.default_rng(nc + 7 * fold + 3 * halo)
    x = lattice("bcc", nc, A_W) + rng.normal(0, 0.05, (2 * nc ** 3, 3))
    v = rng.normal(0, 0.01, x.shape)
    pl = Pipeline()
    try:
        pl.load(x, (nc * A_W,) * 3, mass=183.84, v=v)
        pl.build(CUT, newton=1)
        L, c, s, h = pl.L, pl.ctx.h, C.byref(pl.s), C.byref(pl.h)
        n, ng = pl.s.n_local, pl.s.n_ghost
        mixed = precision == emd.EMD_MIXED
        fdt = np.float32 if mixed else np.float64
        fptr = pl.s.f32 if mixed else pl.s.f
        f0 = rng.normal(0, 1, 3 * (n + ng)).astype(fdt)
        x0 = _get(pl, pl.s.x, 3 * (n + ng), np.float64)
        v0 = _get(pl, pl.s.v, 3 * n, np.float64)
        out = []
        for fused in (False, True):
            _put(pl, pl.s.x, x0)
            _put(pl, pl.s.v, v0)
            _put(pl, fptr, f0)
            if fused:
                pl.chk(L.emd_nve_step_boundary(c, s, h, fold, halo, DTF, DTV, precision),
                       "step boundary")
            else:
                if fold:
                    pl.chk(L.emd_comm_update_force(c, s, h, precision), "fold")
                pl.chk(L.emd_nve_final_initial(c, s, DTF, DTV, precision), "nve")
                if halo:
                    pl.chk(L.emd_comm_update_halo(c, s, h), "halo")
            out.append({"x": _get(pl, pl.s.x, 3 * (n + ng), np.float64),
                        "v": _get(pl, pl.s.v, 3 * n, np.float64),
                        "f": _get(pl, fptr, 3 * (n + ng), fdt)})
    finally:
        pl.close()
    print("locals", n, "ghosts", ng)
    assert (ng <= 16384) == (nc == 10)
    for k in ("x", "v", "f"):
        assert np.array_equal(out[0][k], out[1][k]), k
    # the step moved every atom and (with halo) every ghost
    assert not np.array_equal(out[1]["x"][:3 * n], x0[:3 * n])
    if halo:
        assert not np.array_equal(out[1]["x"][3 * n:], x0[3 * n:])
    else:
        assert np.array_equal(out[1]["x"][3 * n:], x0[3 * n:])
