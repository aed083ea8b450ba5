"""emd_lj_energy_noalloc (all scratch in a static device array, no
allocation) against emd_lj_energy through the C ABI: two calls in a row, the
second on a smaller system, each within 1e-12 relative (the two add the same
block sums in a different order); sizes beyond the array are refused."""
import ctypes as C

import numpy as np
import pytest

import examinimd_amd as emd
from examinimd_amd.abi import EmdLJParams
from examinimd_amd.pipeline import Pipeline, lattice

pytestmark = pytest.mark.gpu


def _params():
    p = EmdLJParams()
    p.ntypes = 1
    p.lj1[0][0], p.lj2[0][0], p.cutsq[0][0] = 48.0, 24.0, 6.25
    return p


def _energies(pl, half=0):
    L, c = pl.L, pl.ctx.h
    p = _params()
    e_ref, e_na = C.c_double(), C.c_double()
    pl.chk(L.emd_lj_energy(c, C.byref(pl.s), C.byref(pl.nl), C.byref(p), half, C.byref(e_ref)),
           "lj_energy")
    pl.chk(L.emd_lj_energy_noalloc(c, C.byref(pl.s), C.byref(pl.nl), C.byref(p), half,
                                   C.byref(e_na)), "lj_energy_noalloc")
    return e_ref.value, e_na.value


def test_noalloc_energy_two_calls_second_smaller():
    a = (4 / 0.8442) ** (1 / 3)
    pl = Pipeline()
    for nc, seed in ((20, 1), (7, 2)):   # 32 000 atoms, then 1 372
        print(f"case: LJ energy noalloc, FCC {nc}^3 displaced", flush=True)
        x = lattice("fcc", nc, a) + np.random.default_rng(seed).normal(0, 0.05, (4 * nc ** 3, 3))
        pl.load(x, (nc * a,) * 3)
        pl.build(2.8)
        ref, na = _energies(pl)
        print("energy", ref, "noalloc", na, flush=True)
        assert np.isfinite(na) and abs(na - ref) <= 1e-12 * abs(ref)
    pl.close()


def test_noalloc_energy_half_list_and_size_limit():
    a = (4 / 0.8442) ** (1 / 3)
    x = lattice("fcc", 6, a) + np.random.default_rng(3).normal(0, 0.05, (864, 3))
    pl = Pipeline()
    pl.load(x, (6 * a,) * 3)
    pl.build(2.8, half=1, newton=0)
    ref, na = _energies(pl, half=1)
    assert abs(na - ref) <= 1e-12 * abs(ref)
    # a system beyond the array's 4,194,304 atoms is refused before any launch
    n0 = pl.s.n_local
    pl.s.n_local = (1 << 22) + 1
    e = C.c_double()
    st = pl.L.emd_lj_energy_noalloc(pl.ctx.h, C.byref(pl.s), C.byref(pl.nl), C.byref(_params()),
                                    0, C.byref(e))
    pl.s.n_local = n0
    assert st == 1  # EMD_ERR_ARG
    assert b"limit" in pl.L.emd_last_error(pl.ctx.h)
    pl.close()
