"""GPU parity of SNAP (adjoint form on the GPU vs the reference Z/dB form in
the CPU oracle).  The two evaluate the same force with a different order of
floating-point operations, so agreement is to rounding: 1e-12 relative on
forces, energies and on positions, velocities and forces after 100 steps."""
import os

import numpy as np
import pytest

import examinimd_amd as emd
from examinimd_amd import decks
from conftest import REF_INPUT, read_snapcoeff
from helpers import rel_err

pytestmark = pytest.mark.gpu

W_COEFF = os.path.join(REF_INPUT, "W.snapcoeff")
W_PARAM = os.path.join(REF_INPUT, "W.snapparam")


def oracle_md(oracle, beta, lattice, n, steps, every, gaussian, dt, seed, T=300.0):
    cfg = oracle.make_config(units="metal", lattice=lattice, lattice_arg=3.1803, n=n,
                             mass=183.84, temperature=T, seed=seed, gaussian=gaussian,
                             force="snap", skin=1.0, every=every, newton=True, dt=dt)
    sp = oracle.snap_params(beta)
    m = oracle.MD(cfg, sp)
    m.init()
    if steps:
        m.run(steps)
    return m


def gpu_md(tmp_path, coeff, lattice, n, steps, every, gaussian, dt, seed, precision="fp64",
           neigh="CSR", T=300.0):
    deck = decks.snap_deck(n=n[0], ny=n[1], nz=n[2], steps=steps, every=every, lattice=lattice,
                           coeff=coeff, param=W_PARAM, gaussian=gaussian, dt=dt, seed=seed,
                           temperature=T)
    path = decks.write(str(tmp_path / f"in.snap.{precision}"), deck)
    return emd.App(["-il", path, "--neigh-type", neigh, "--precision", precision]).init()


def compare(app, m, tol):
    """max |gpu - oracle| over max(|oracle|, floor); the force floor of
    1 eV/A keeps perfect-lattice forces (pure rounding noise, ~1e-13) from
    being divided by themselves"""
    g, o = app.arrays(), m.arrays()
    n = app.nlocal
    assert np.array_equal(g["id"], o["id"][:n])
    floor = {"x": 0.0, "v": 0.0, "f": 1.0}
    errs = {k: float(np.max(np.abs(g[k] - o[k][:n])) / max(np.max(np.abs(o[k][:n])), floor[k]))
            for k in ("x", "v", "f")}
    print("relative errors", errs)
    for k, e in errs.items():
        assert e <= tol, f"{k}: {e} > {tol}"


def bcc_coeff(tmp_path, oracle):
    beta = decks.snap_beta(seed=2024, sigma=0.05)
    path = str(tmp_path / "W_bcc.snapcoeff")
    decks.write(path, decks.snap_coeff_text(beta))
    return list(beta), path


def test_snap_reference_deck_init(tmp_path, oracle):
    """the reference's in.snap.W: SC W 4x8x8, newton on, rebuild every step"""
    print("case: SNAP SC 256 atoms init", flush=True)
    beta = read_snapcoeff(W_COEFF)
    app = gpu_md(tmp_path, W_COEFF, "sc", (4, 8, 8), 0, 1, False, 0.001, 4928459)
    m = oracle_md(oracle, beta, "sc", (4, 8, 8), 0, 1, False, 0.001, 4928459)
    compare(app, m, 1e-12)


def test_snap_reference_deck_100_steps(tmp_path, oracle):
    print("case: SNAP SC 256 atoms, 100 steps", flush=True)
    beta = read_snapcoeff(W_COEFF)
    app = gpu_md(tmp_path, W_COEFF, "sc", (4, 8, 8), 100, 1, False, 0.001, 4928459)
    app.run(100)
    m = oracle_md(oracle, beta, "sc", (4, 8, 8), 100, 1, False, 0.001, 4928459)
    compare(app, m, 1e-12)


def test_snap_bcc_2000_init_and_energy(tmp_path, oracle):
    print("case: SNAP BCC 2000 atoms init + energy", flush=True)
    beta, path = bcc_coeff(tmp_path, oracle)
    app = gpu_md(tmp_path, path, "bcc", (10, 10, 10), 0, 10, True, 0.0005, 4928459)
    m = oracle_md(oracle, beta, "bcc", (10, 10, 10), 0, 10, True, 0.0005, 4928459)
    compare(app, m, 1e-12)
    tg, to = app.thermo(), m.thermo()
    print("thermo gpu", tg, "oracle", to)
    assert abs(tg[1] - to[1]) <= 1e-12 * abs(to[1])


def test_snap_bcc_2000_100_steps(tmp_path, oracle):
    print("case: SNAP BCC 2000 atoms, 100 steps", flush=True)
    beta, path = bcc_coeff(tmp_path, oracle)
    app = gpu_md(tmp_path, path, "bcc", (10, 10, 10), 100, 10, True, 0.0005, 4928459)
    app.run(100)
    m = oracle_md(oracle, beta, "bcc", (10, 10, 10), 100, 10, True, 0.0005, 4928459)
    compare(app, m, 1e-12)


def test_snap_2d_list(tmp_path, oracle):
    print("case: SNAP SC 256 atoms, 2D list, 20 steps", flush=True)
    beta = read_snapcoeff(W_COEFF)
    app = gpu_md(tmp_path, W_COEFF, "sc", (4, 8, 8), 20, 1, False, 0.001, 4928459, neigh="2D")
    app.run(20)
    m = oracle_md(oracle, beta, "sc", (4, 8, 8), 20, 1, False, 0.001, 4928459)
    compare(app, m, 1e-12)


def test_snap_kernel_level_disordered(oracle):
    """thermally displaced BCC W (0.1 A): GPU FP64 force and energy against
    the oracle, and the FP32 path's force RMS error against FP64"""
    from examinimd_amd.pipeline import Pipeline, lattice
    from test_oracle import _energy_force
    print("case: SNAP BCC 2000 disordered, kernel level", flush=True)
    beta = list(decks.snap_beta(seed=2024, sigma=0.05))
    a = 3.1803
    x = lattice("bcc", 10, a) + np.random.default_rng(3).normal(0, 0.1, (2000, 3))
    box = (10 * a,) * 3
    pl = Pipeline()
    pl.load(x, box, mass=183.84)
    pl.build(4.73442 + 1.0, newton=1)
    pl.snap_setup(beta)
    f64 = pl.snap_force(emd.EMD_FP64)
    e64 = pl.snap_energy()
    f32 = pl.snap_force(emd.EMD_MIXED)
    pl.close()
    E, F = _energy_force(oracle, x % (10 * a), 10 * a, "snap", snap=oracle.snap_params(beta),
                         cut=5.73442)
    print("force rel err", rel_err(f64, F), "energy", e64, E)
    assert rel_err(f64, F) <= 1e-12
    assert abs(e64 - E) <= 1e-12 * abs(E)
    df = f32 - f64
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    print("mixed SNAP force RMS error (eV/A)", rms, "force RMS", np.sqrt(np.mean(np.sum(f64 ** 2, 1))))
    assert rms <= 1e-4


@pytest.mark.parametrize("nc,cut", [(10, 5.73442), (30, 5.73442)])
@pytest.mark.parametrize("mixed", [False, True])
def test_ghost_force_fold_against_numpy(monkeypatch, nc, cut, mixed):
    """CommSerial::update_force on seeded random forces: the one-workgroup
    six-phase fold (small halos) and the per-phase launches both equal the
    sequential fold f[pack_p] += f[ghosts_p], p = 5..0, done in numpy.
    (Whole SNAP runs cannot be compared bit for bit: the force scatter onto
    neighbours uses atomics.)  BCC 10^3 gives ~3 000 ghosts (single launch),
    30^3 ~ 22 000 (per-phase launches)."""
    import ctypes as C
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: ghost force fold, BCC {nc}^3, cut {cut}, mixed={mixed}", flush=True)
    a = 3.1803
    x = lattice("bcc", nc, a) + np.random.default_rng(41).normal(0, 0.05, (2 * nc ** 3, 3))
    prec = emd.EMD_MIXED if mixed else emd.EMD_FP64
    dt = np.float32 if mixed else np.float64
    results = []
    for one in ("1", "0"):
        monkeypatch.setenv("EMD_FOLD_ONE", one)
        pl = Pipeline()
        pl.load(x, (nc * a,) * 3, mass=183.84)
        pl.build(cut, newton=1)
        L, c = pl.L, pl.ctx.h
        n, ng = pl.s.n_local, pl.s.n_ghost
        counts = [pl.h.num_ghost[p] for p in range(6)]
        assert sum(counts) == ng
        packs = []
        for p in range(6):
            arr = np.zeros(counts[p], np.int32)
            if counts[p]:
                pl.chk(L.emd_memcpy_d2h(c, arr.ctypes.data_as(C.c_void_p),
                                        C.cast(pl.h.pack[p], C.c_void_p), arr.nbytes), "pack")
            packs.append(arr)
        f0 = np.random.default_rng(7).normal(0, 1, 3 * (n + ng)).astype(dt)
        dst = pl.s.f32 if mixed else pl.s.f
        pl.chk(L.emd_memcpy_h2d(c, C.cast(dst, C.c_void_p), f0.ctypes.data_as(C.c_void_p),
                                f0.nbytes), "upload f")
        pl.chk(L.emd_comm_update_force(c, C.byref(pl.s), C.byref(pl.h), prec), "fold")
        got = np.zeros(3 * n, dt)
        pl.chk(L.emd_memcpy_d2h(c, got.ctypes.data_as(C.c_void_p), C.cast(dst, C.c_void_p),
                                got.nbytes), "download f")
        want = f0.reshape(-1, 3).copy()
        off = [n]
        for p in range(5):
            off.append(off[-1] + counts[p])
        for p in range(5, -1, -1):
            if counts[p]:
                assert len(np.unique(packs[p])) == counts[p]
                want[packs[p]] += want[off[p]:off[p] + counts[p]]
        assert np.array_equal(got.reshape(-1, 3), want[:n]), f"fold one={one}"
        results.append((ng, got))
        pl.close()
    print("ghosts", results[0][0])
    assert (results[0][0] <= 16384) == (nc == 10)
    assert np.array_equal(results[0][1], results[1][1])


@pytest.mark.parametrize("nc,cut", [(10, 5.73442), (6, 5.73442), (8, 20.0)])
def test_halo_exchange_one_launch_equals_scans(monkeypatch, nc, cut):
    """exchange_halo of a small system in one workgroup launch gives the
    ghosts of the six per-phase scans: counts, pack lists, positions,
    velocities, ids, types, charges, roots and image codes, bit for bit (including the first
    call, which starts with no capacity and grows).  BCC 8^3 with a 20 A
    halo depth has ~27 000 local + ghost atoms, more than one workgroup
    handles: that call falls back to the scans mid-way."""
    import ctypes as C
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: exchange_halo one launch vs scans, BCC {nc}^3, depth {cut}", flush=True)
    a = 3.1803
    x = lattice("bcc", nc, a) + np.random.default_rng(43).normal(0, 0.08, (2 * nc ** 3, 3))
    v = np.random.default_rng(44).normal(0, 1, x.shape)
    out = []
    for one in ("1", "0"):
        monkeypatch.setenv("EMD_HALO_ONE", one)
        pl = Pipeline()
        pl.load(x, (nc * a,) * 3, mass=183.84, v=v)
        L, c, s = pl.L, pl.ctx.h, C.byref(pl.s)
        pl.chk(L.emd_comm_exchange(c, s), "exchange")
        pl.chk(L.emd_binning(c, s, cut, cut, cut, 1, 1, 0, 1, C.byref(pl.b)), "bin")
        pl.chk(L.emd_comm_exchange_halo(c, s, cut, C.byref(pl.h)), "halo")
        n, ng = pl.s.n_local, pl.s.n_ghost
        tot = n + ng
        rec = {"counts": [pl.h.num_ghost[p] for p in range(6)], "n_ghost": ng}
        xs = np.zeros(3 * tot)
        vs = np.zeros(3 * tot)
        ids = np.zeros(tot, np.int32)
        types = np.full(tot, -1, np.int32)
        qs = np.full(tot, np.nan)
        pl.chk(L.emd_system_download(c, s, tot, xs.ctypes.data_as(C.POINTER(C.c_double)),
                                     vs.ctypes.data_as(C.POINTER(C.c_double)), None,
                                     types.ctypes.data_as(C.POINTER(C.c_int)),
                                     ids.ctypes.data_as(C.POINTER(C.c_int)),
                                     qs.ctypes.data_as(C.POINTER(C.c_double))), "download")
        rec["x"], rec["v"], rec["id"], rec["type"], rec["q"] = xs, vs, ids, types, qs
        for name, ptr, cnt in [("root", pl.h.root, ng), ("image", pl.h.image, ng)] + \
                [(f"pack{p}", pl.h.pack[p], pl.h.num_ghost[p]) for p in range(6)]:
            arr = np.zeros(cnt, np.int32)
            if cnt:
                pl.chk(L.emd_memcpy_d2h(c, arr.ctypes.data_as(C.c_void_p), C.cast(ptr, C.c_void_p),
                                        arr.nbytes), name)
            rec[name] = arr
        # a second exchange on the grown capacities gives the same ghosts
        pl.chk(L.emd_comm_exchange_halo(c, s, cut, C.byref(pl.h)), "halo again")
        assert pl.s.n_ghost == ng
        out.append(rec)
        pl.close()
    assert out[0]["n_ghost"] > 0
    for k in out[0]:
        assert np.array_equal(np.asarray(out[0][k]), np.asarray(out[1][k])), k



@pytest.mark.parametrize("prec", [emd.EMD_FP64, emd.EMD_MIXED])
def test_snap_force_zeroed_equals_zero_then_force(prec):
    """emd_snap_force_zeroed zeroes every local and ghost force row in its
    first kernel: with f poisoned (NaN) beforehand, the folded forces match
    force_zero + emd_snap_force (the force kernel's atomic additions make the
    order, not the values, vary: 1e-12 relative)"""
    import ctypes as C
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: SNAP zeroing fused into the first kernel, precision {prec}", flush=True)
    a = 3.1803
    x = lattice("bcc", 8, a) + np.random.default_rng(4).normal(0, 0.1, (1024, 3))
    pl = Pipeline()
    pl.load(x, (8 * a,) * 3, mass=183.84)
    pl.build(4.73442 + 1.0, newton=1)
    pl.snap_setup(list(decks.snap_beta(seed=2024, sigma=0.05)))
    want = pl.snap_force(prec)
    L, c, s = pl.L, pl.ctx.h, C.byref(pl.s)
    ntot = pl.s.n_local + pl.s.n_ghost
    if prec == emd.EMD_MIXED:
        nan, dst = np.full(3 * ntot, np.nan, np.float32), pl.s.f32
    else:
        nan, dst = np.full(3 * ntot, np.nan, np.float64), pl.s.f
    pl.chk(L.emd_memcpy_h2d(c, C.cast(dst, C.c_void_p), nan.ctypes.data_as(C.c_void_p),
                            nan.nbytes), "poison")
    pl.chk(L.emd_snap_force_zeroed(c, s, C.byref(pl.nl), pl.snap, prec), "snap zeroed")
    pl.chk(L.emd_comm_update_force(c, s, C.byref(pl.h), prec), "fold")
    got = pl.forces(prec)
    pl.close()
    assert np.isfinite(got).all()
    assert rel_err(got, want) <= 1e-12 if prec == emd.EMD_FP64 else rel_err(got, want) <= 1e-5


def test_snap_ragged_rows_patch_and_isolated_atoms(oracle):
    """a BCC W patch (432 atoms) in one corner of a 70 A box, 64 isolated
    atoms on a 10 A grid and 8 dimers (2.5 A): atoms with no neighbour, one
    neighbour, surface and bulk rows.  FP64 forces and energy match the
    oracle; isolated atoms get exactly zero force in both precisions"""
    from examinimd_amd.pipeline import Pipeline, lattice
    from test_oracle import _energy_force
    print("case: SNAP ragged rows", flush=True)
    beta = list(decks.snap_beta(seed=2024, sigma=0.05))
    patch = lattice("bcc", 6, 3.1803)                 # spans [0, 19.1)
    g = np.array([28.0, 38.0, 48.0, 58.0])
    iso = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(12)
    iso = iso + rng.uniform(-0.3, 0.3, iso.shape)
    d = rng.normal(0, 1, (8, 3))
    dimer = iso[:8] + 2.5 * d / np.linalg.norm(d, axis=1, keepdims=True)
    x = np.concatenate([patch, iso, dimer])
    box = 70.0
    pl = Pipeline()
    pl.load(x, (box,) * 3, mass=183.84)
    pl.build(4.73442 + 1.0, newton=1)
    pl.snap_setup(beta)
    f64 = pl.snap_force(emd.EMD_FP64)
    e64 = pl.snap_energy()
    f32 = pl.snap_force(emd.EMD_MIXED)
    pl.close()
    E, F = _energy_force(oracle, x, box, "snap", snap=oracle.snap_params(beta), cut=5.73442)
    print("force rel err", rel_err(f64, F), "energy", e64, E)
    assert rel_err(f64, F) <= 1e-12
    assert abs(e64 - E) <= 1e-12 * abs(E)
    np_, ni = len(patch), len(iso)
    assert not f64[np_ + 8:np_ + ni].any() and not f32[np_ + 8:np_ + ni].any()
    fd = f64[np_:np_ + 8]
    assert np.abs(fd).min() > 0
    assert np.abs(fd + f64[np_ + ni:]).max() <= 1e-12 * np.abs(fd).max()
    df = f32 - f64
    assert float(np.sqrt(np.mean(np.sum(df * df, axis=1)))) <= 1e-4


def test_snap_descriptors_against_oracle_bispectrum(oracle):
    """emd_snap_descriptors: per-atom B_k on a disordered BCC W patch equal
    the oracle's bispectrum of each atom's neighbour shell (rij = x_j - x_i
    within rc, minimum image) to 1e-12 relative, and beta_0 + beta . B plus
    the repulsion reproduces the per-atom energy"""
    import ctypes as C
    from examinimd_amd.pipeline import Pipeline, lattice
    print("case: SNAP descriptors, BCC W 4^3 displaced", flush=True)
    a, nc, rc = 3.1803, 4, 4.73442
    rng = np.random.default_rng(12)
    x = lattice("bcc", nc, a) + rng.normal(0, 0.1, (2 * nc ** 3, 3))
    box = nc * a
    beta = np.array(decks.snap_beta(seed=2024, sigma=0.05))
    pl = Pipeline()
    pl.load(x, (box,) * 3, mass=183.84)
    pl.build(rc + 1.0, newton=1)
    pl.snap_setup(list(beta))
    L, c = pl.L, pl.ctx.h
    n, ncoef = pl.s.n_local, len(beta) - 1
    dev, e_dev = C.c_void_p(), C.c_void_p()
    pl.chk(L.emd_malloc(c, C.byref(dev), n * ncoef * 8), "malloc")
    pl.chk(L.emd_malloc(c, C.byref(e_dev), n * 8), "malloc")
    dptr = lambda p: C.cast(p, C.POINTER(C.c_double))
    pl.chk(L.emd_snap_descriptors(c, C.byref(pl.s), C.byref(pl.nl), pl.snap, dptr(dev)), "desc")
    B = np.zeros((n, ncoef))
    pl.chk(L.emd_memcpy_d2h(c, B.ctypes.data_as(C.c_void_p), dev, B.nbytes), "d2h")
    e = C.c_double()
    pl.chk(L.emd_snap_energy(c, C.byref(pl.s), C.byref(pl.nl), pl.snap, C.byref(e), dptr(e_dev)),
           "energy")
    ea = np.zeros(n)
    pl.chk(L.emd_memcpy_d2h(c, ea.ctypes.data_as(C.c_void_p), e_dev, ea.nbytes), "d2h")
    pl.snap_set_coeff(np.zeros_like(beta))
    pl.chk(L.emd_snap_energy(c, C.byref(pl.s), C.byref(pl.nl), pl.snap, C.byref(e), dptr(e_dev)),
           "energy")
    rep = np.zeros(n)
    pl.chk(L.emd_memcpy_d2h(c, rep.ctypes.data_as(C.c_void_p), e_dev, rep.nbytes), "d2h")
    ids = pl._device_ids()
    L.emd_free(c, dev)
    L.emd_free(c, e_dev)
    pl.close()
    assert np.abs(ea - (beta[0] + B @ beta[1:] + rep)).max() <= 1e-12 * np.abs(ea).max()
    sp = oracle.snap_params(list(beta))
    xw = x % box
    worst = 0.0
    for k in range(0, n, 5):
        i = ids[k] - 1
        d = xw - xw[i]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(1)
        sel = (r2 < rc * rc) & (np.arange(len(x)) != i)
        rij = np.ascontiguousarray(d[sel])
        Bo = np.zeros(ncoef)
        oracle.lib().or_snap_bispectrum(C.byref(sp), len(rij), oracle.dptr(rij), oracle.dptr(Bo))
        worst = max(worst, np.abs(B[k] - Bo).max() / np.abs(Bo).max())
    print("descriptor relative error", worst)
    assert worst <= 1e-12
