"""SNAP at twojmax other than 8.  The reference sizes its tables from twojmax
at run time (sna_impl.hpp:25-53, build_indexlist 86-132); this library has
two size classes, twojmax <= 8 and twojmax <= 12 (larger LDS tables, one
neighbour per wavefront in the force kernel).  Forces and energies against
the oracle (whose tables at twojmax 12 are pinned in tests/test_oracle.py),
descriptors against the oracle's bispectrum, the adjoint against the
forward pass (the force is affine in beta), and a short MD run through the
host driver against the oracle's driver."""
import ctypes as C
import os

import numpy as np
import pytest

import examinimd_amd as emd
from examinimd_amd import decks
from helpers import rel_err

pytestmark = pytest.mark.gpu

A_W, RC = 3.1803, 4.73442


def shaken_bcc(nc, seed, sigma=0.1):
    from examinimd_amd.pipeline import lattice
    x = lattice("bcc", nc, A_W) + np.random.default_rng(seed).normal(0, sigma, (2 * nc ** 3, 3))
    return x, nc * A_W


def ncoeff(oracle, tj):
    return oracle.lib().or_snap_ncoeff(tj)


@pytest.mark.parametrize("twojmax", [2, 4, 6, 10, 12])
def test_snap_twojmax_force_energy_against_oracle(oracle, twojmax):
    from examinimd_amd.pipeline import Pipeline
    from test_oracle import _energy_force
    print(f"case: SNAP twojmax {twojmax}, BCC W 4^3 shaken, kernel level", flush=True)
    beta = list(decks.snap_beta(seed=77 + twojmax, sigma=0.05, ncoeff=ncoeff(oracle, twojmax)))
    x, box = shaken_bcc(4, twojmax)
    pl = Pipeline()
    pl.load(x, (box,) * 3, mass=183.84)
    pl.build(RC + 1.0, newton=1)
    pl.snap_setup(beta, twojmax=twojmax)
    f64 = pl.snap_force(emd.EMD_FP64)
    e64 = pl.snap_energy()
    f32 = pl.snap_force(emd.EMD_MIXED)
    pl.close()
    E, F = _energy_force(oracle, x % box, box, "snap",
                         snap=oracle.snap_params(beta, twojmax=twojmax), cut=RC + 1.0)
    print("force rel err", rel_err(f64, F), "energy", e64, E)
    assert np.abs(F).max() > 0.1
    assert rel_err(f64, F) <= 1e-12
    assert abs(e64 - E) <= 1e-12 * abs(E)
    df = f32 - f64
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    frms = float(np.sqrt(np.mean(np.sum(f64 * f64, axis=1))))
    print("mixed force RMS error", rms, "force RMS", frms)
    assert rms <= 1e-5 * max(frms, 1.0) * 10


def test_snap_twojmax12_descriptors_against_oracle_bispectrum(oracle):
    from examinimd_amd.pipeline import Pipeline
    print("case: SNAP twojmax 12 descriptors, BCC W 3^3 shaken", flush=True)
    tj = 12
    nb = ncoeff(oracle, tj)
    beta = np.array(decks.snap_beta(seed=5, sigma=0.05, ncoeff=nb))
    x, box = shaken_bcc(3, 8)
    pl = Pipeline()
    pl.load(x, (box,) * 3, mass=183.84)
    pl.build(RC + 1.0, newton=1)
    pl.snap_setup(list(beta), twojmax=tj)
    L, c = pl.L, pl.ctx.h
    n = pl.s.n_local
    dev = C.c_void_p()
    pl.chk(L.emd_malloc(c, C.byref(dev), n * nb * 8), "malloc")
    pl.chk(L.emd_snap_descriptors(c, C.byref(pl.s), C.byref(pl.nl), pl.snap,
                                  C.cast(dev, C.POINTER(C.c_double))), "desc")
    B = np.zeros((n, nb))
    pl.chk(L.emd_memcpy_d2h(c, B.ctypes.data_as(C.c_void_p), dev, B.nbytes), "d2h")
    ids = pl._device_ids()
    L.emd_free(c, dev)
    pl.close()
    sp = oracle.snap_params(list(beta), twojmax=tj)
    xw = x % box
    worst = 0.0
    for k in range(0, n, 3):
        i = ids[k] - 1
        d = xw - xw[i]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(1)
        sel = (r2 < RC * RC) & (np.arange(len(x)) != i)
        rij = np.ascontiguousarray(d[sel])
        Bo = np.zeros(nb)
        oracle.lib().or_snap_bispectrum(C.byref(sp), len(rij), oracle.dptr(rij), oracle.dptr(Bo))
        worst = max(worst, np.abs(B[k] - Bo).max() / np.abs(Bo).max())
    print("descriptor relative error", worst)
    assert worst <= 1e-12


@pytest.mark.parametrize("twojmax", [10, 12])
def test_snap_twojmax_adjoint_is_the_transpose_of_the_force(oracle, twojmax):
    """sum_i g_i . dF_i/dbeta_k + gE dE/dbeta_k from emd_snap_adjoint_beta
    equals the same contraction of forward passes with unit beta vectors"""
    from examinimd_amd.pipeline import Pipeline
    print(f"case: SNAP twojmax {twojmax} adjoint vs forward, BCC W 3^3", flush=True)
    nb = ncoeff(oracle, twojmax)
    x, box = shaken_bcc(3, 30 + twojmax, sigma=0.08)
    pl = Pipeline()
    pl.load(x, (box,) * 3, mass=183.84)
    pl.build(RC + 1.0, newton=1)
    pl.snap_setup([0.0] * (nb + 1), twojmax=twojmax)
    rng = np.random.default_rng(twojmax)
    gF = rng.normal(size=(len(x), 3))
    gE = 0.37
    F0 = pl.snap_force(emd.EMD_FP64)
    E0 = pl.snap_energy()
    want = np.zeros(nb + 1)
    for k in range(nb + 1):
        e = np.zeros(nb + 1)
        e[k] = 1.0
        pl.snap_set_coeff(e)
        want[k] = np.sum(gF * (pl.snap_force(emd.EMD_FP64) - F0)) + gE * (pl.snap_energy() - E0)
    got = {p: pl.snap_adjoint(gF, gE, p) for p in (emd.EMD_FP64, emd.EMD_MIXED)}
    pl.close()
    scale = np.abs(want).max()
    e64 = np.abs(got[emd.EMD_FP64] - want).max() / scale
    e32 = np.abs(got[emd.EMD_MIXED] - want).max() / scale
    print("adjoint rel err FP64", e64, "mixed", e32)
    assert e64 <= 1e-9
    assert e32 <= 1e-4


def test_snap_twojmax10_md_against_oracle(tmp_path, oracle):
    """the host driver with a twojmax-10 parameter file: 20 steps of BCC W
    4^3 (rebuild every 10) against the oracle's driver, 1e-12 relative"""
    from test_gpu_snap import compare
    print("case: SNAP twojmax 10, driver, BCC 4^3, 20 steps", flush=True)
    tj = 10
    beta = list(decks.snap_beta(seed=4, sigma=0.05, ncoeff=ncoeff(oracle, tj)))
    coeff = str(tmp_path / "W10.snapcoeff")
    decks.write(coeff, decks.snap_coeff_text(beta))
    param = str(tmp_path / "W10.snapparam")
    decks.write(param, decks.SNAP_PARAM.replace("twojmax 8", f"twojmax {tj}"))
    deck = decks.snap_deck(n=4, steps=20, every=10, coeff=coeff, param=param, gaussian=False,
                           seed=91, dt=0.0005)
    path = decks.write(str(tmp_path / "in.snap10"), deck)
    app = emd.App(["-il", path, "--neigh-type", "CSR", "--precision", "fp64"]).init()
    app.run(20)
    cfg = oracle.make_config(units="metal", lattice="bcc", lattice_arg=A_W, n=(4, 4, 4),
                             mass=183.84, temperature=300.0, seed=91, gaussian=False,
                             force="snap", skin=1.0, every=10, newton=True, dt=0.0005)
    m = oracle.MD(cfg, oracle.snap_params(beta, twojmax=tj))
    m.init()
    m.run(20)
    compare(app, m, 1e-12)
    app.close()


@pytest.mark.parametrize("twojmax", [8, 6])
def test_snap_dense_neighbourhoods_against_oracle(oracle, twojmax):
    """compressed BCC W (a = 2.6 A, 0.05 A disorder): about 58 neighbours
    inside the cutoff, so the force kernel runs three block rows of 28
    neighbours per atom and the U kernel nine rounds of seven; FP64 forces
    and energy against the oracle, the mixed path's force RMS error"""
    from examinimd_amd.pipeline import Pipeline, lattice
    from test_oracle import _energy_force
    print(f"case: SNAP dense BCC W a = 2.6, twojmax {twojmax}", flush=True)
    a, nc = 2.6, 5
    beta = list(decks.snap_beta(seed=11 + twojmax, sigma=0.05, ncoeff=ncoeff(oracle, twojmax)))
    x = lattice("bcc", nc, a) + np.random.default_rng(7).normal(0, 0.05, (2 * nc ** 3, 3))
    box = nc * a
    pl = Pipeline()
    pl.load(x, (box,) * 3, mass=183.84)
    pl.build(RC + 1.0, newton=1)
    pl.snap_setup(beta, twojmax=twojmax)
    f64 = pl.snap_force(emd.EMD_FP64)
    e64 = pl.snap_energy()
    f32 = pl.snap_force(emd.EMD_MIXED)
    pl.close()
    xw = x % box
    d = xw[:, None, :] - xw[None, :, :]
    d -= box * np.round(d / box)
    nin = ((d ** 2).sum(-1) < RC * RC).sum(1) - 1
    print("in-cutoff neighbours min / max", nin.min(), nin.max())
    assert nin.min() > 28
    E, F = _energy_force(oracle, xw, box, "snap",
                         snap=oracle.snap_params(beta, twojmax=twojmax), cut=RC + 1.0)
    print("force rel err", rel_err(f64, F), "energy", e64, E)
    assert rel_err(f64, F) <= 1e-12
    assert abs(e64 - E) <= 1e-12 * abs(E)
    df = f32 - f64
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    frms = float(np.sqrt(np.mean(np.sum(f64 * f64, axis=1))))
    print("mixed force RMS error", rms, "force RMS", frms)
    assert rms <= 1e-4 * frms


@pytest.mark.parametrize("natoms", [1, 127])
def test_snap_odd_atom_counts_against_oracle(oracle, natoms):
    """an odd number of atoms (the Y kernel takes two atoms per block, so the
    last block has one) and a single atom with no neighbours: FP64 forces
    and energy against the oracle in both precisions' code paths"""
    from examinimd_amd.pipeline import Pipeline
    from test_oracle import _energy_force
    print(f"case: SNAP {natoms} atoms (odd count)", flush=True)
    beta = list(decks.snap_beta(seed=9, sigma=0.05))
    x, box = shaken_bcc(4, 21)
    x = x[:natoms]
    pl = Pipeline()
    pl.load(x, (box,) * 3, mass=183.84)
    pl.build(RC + 1.0, newton=1)
    pl.snap_setup(beta)
    f64 = pl.snap_force(emd.EMD_FP64)
    e64 = pl.snap_energy()
    f32 = pl.snap_force(emd.EMD_MIXED)
    pl.close()
    E, F = _energy_force(oracle, x % box, box, "snap", snap=oracle.snap_params(beta), cut=RC + 1.0)
    scale = max(np.abs(F).max(), 1.0)
    print("force max abs err", np.abs(f64 - F).max(), "energy", e64, E)
    assert np.abs(f64 - F).max() <= 1e-12 * scale
    assert abs(e64 - E) <= 1e-12 * max(abs(E), 1.0)
    assert np.abs(f32 - f64).max() <= 1e-4 * scale
