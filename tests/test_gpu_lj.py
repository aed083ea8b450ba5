"""GPU parity of the LJ path (wrap, halo, binning, neighbour list, LJ force,
NVE) against the CPU oracle, end to end through the C ABI driver.

The oracle runs the reference algorithm in its serial order; the GPU path is
deterministic and, with one lane per atom in the LJ kernel, performs the same
FP64 operations in the same order, so positions, velocities and forces are
expected to be bit-identical.  The default multi-lane kernel changes only the
summation order of each atom's force: bound 1e-12 relative over 100 steps.
"""
import ctypes as C
import os

import numpy as np
import pytest

import examinimd_amd as emd
from examinimd_amd import decks
from helpers import pair_forces_typed, rel_err

pytestmark = pytest.mark.gpu

N_CELLS = 20      # 32 000 atoms: the CPU-runnable LJ configuration
STEPS = 100


def oracle_run(oracle, steps, n=N_CELLS, half=False, newton=False, gaussian=True,
               intensity=None):
    cfg = oracle.make_config(n=(n, n, n), gaussian=gaussian, half=half, newton=newton,
                             force="lj" if intensity is None else "lj/cut/idial",
                             intensity=intensity)
    m = oracle.MD(cfg)
    m.init()
    if steps:
        m.run(steps)
    return m


def gpu_app(tmp_path, steps, n=N_CELLS, neigh="CSR", precision="fp64", half=False,
            newton=False, gaussian=True, lanes=None, intensity=None):
    deck = decks.write(str(tmp_path / "in.lj"),
                       decks.lj_deck(n=n, steps=steps, newton=newton, gaussian=gaussian,
                                     intensity=intensity))
    if lanes is not None:
        os.environ["EMD_LJ_LANES"] = str(lanes)
    try:
        args = ["-il", deck, "--neigh-type", neigh, "--comm-type", "SERIAL",
                "--force-iteration", "NEIGH_HALF" if half else "NEIGH_FULL",
                "--precision", precision]
        app = emd.App(args).init()
    finally:
        os.environ.pop("EMD_LJ_LANES", None)
    return app


def compare(app, m, tol, exact=False):
    g = app.arrays()
    o = m.arrays()
    n = app.nlocal
    assert n == m.system.N_local
    assert np.array_equal(g["id"], o["id"][:n]), "atom order differs"
    for k in ("x", "v", "f"):
        if exact:
            assert np.array_equal(g[k], o[k][:n]), f"{k} not bit-identical: max diff " \
                f"{np.max(np.abs(g[k] - o[k][:n]))}"
        else:
            e = rel_err(g[k], o[k][:n])
            assert e <= tol, f"{k}: relative error {e} > {tol}"


def test_lj32k_init_bitexact(tmp_path, oracle):
    print("case: LJ 32k init, CSR, 1 lane", flush=True)
    app = gpu_app(tmp_path, 0, lanes=1)
    m = oracle_run(oracle, 0)
    compare(app, m, 0.0, exact=True)
    # neighbour list: identical rows in identical order
    row, num, ent = app.neigh()
    orow, oent = m.csr()
    assert np.array_equal(row, orow)
    assert np.array_equal(ent, oent)
    assert app.nghost == m.system.N_ghost


def test_lj32k_100_steps_bitexact(tmp_path, oracle):
    print("case: LJ 32k, 100 steps, CSR, 1 lane", flush=True)
    app = gpu_app(tmp_path, STEPS, lanes=1)
    app.run(STEPS)
    m = oracle_run(oracle, STEPS)
    compare(app, m, 0.0, exact=True)
    th_g, th_o = app.thermo(), m.thermo()
    assert abs(th_g[0] - th_o[0]) <= 1e-12 * abs(th_o[0])
    assert abs(th_g[1] - th_o[1]) <= 1e-12 * abs(th_o[1])
    assert abs(th_g[2] - th_o[2]) <= 1e-12 * abs(th_o[2])


@pytest.mark.parametrize("lanes", [2, 4, 8, 16])
def test_lj32k_100_steps_lanes(tmp_path, oracle, lanes):
    print(f"case: LJ 32k, 100 steps, CSR, {lanes} lanes", flush=True)
    app = gpu_app(tmp_path, STEPS, lanes=lanes)
    app.run(STEPS)
    m = oracle_run(oracle, STEPS)
    compare(app, m, 1e-12)


def test_lj32k_2d_list(tmp_path, oracle):
    print("case: LJ 32k, 100 steps, 2D list", flush=True)
    app = gpu_app(tmp_path, STEPS, neigh="2D", lanes=1)
    app.run(STEPS)
    m = oracle_run(oracle, STEPS)
    compare(app, m, 0.0, exact=True)


@pytest.mark.parametrize("newton", [False, True])
def test_lj_half_list(tmp_path, oracle, newton):
    print(f"case: LJ 4000 atoms half list newton={newton}", flush=True)
    app = gpu_app(tmp_path, 40, n=10, half=True, newton=newton)
    m = oracle_run(oracle, 0, n=10, half=True, newton=newton)
    # the half-list rows hold the same pairs as the oracle's
    row, num, ent = app.neigh()
    orow, oent = m.csr()
    assert np.array_equal(row, orow)
    assert np.array_equal(ent, oent)
    assert np.array_equal(app.arrays()["x"], m.arrays()["x"][:app.nlocal])
    # on the perfect lattice every force is rounding noise (~1e-14) whose
    # value depends on the order of the atomic updates: absolute bound
    assert np.abs(app.arrays()["f"] - m.arrays()["f"][:app.nlocal]).max() < 1e-12
    app.run(40)
    m.run(40)
    compare(app, m, 1e-12)
    # PE/N with the reference's half-list accounting (a pair with a ghost j
    # counts half, force_lj_neigh_impl.h:329-339, with newton on as well)
    pe, pe_o = app.thermo()[1], m.thermo()[1]
    print("PE/N", pe, pe_o)
    assert abs(pe - pe_o) <= 1e-12 * abs(pe_o)


def test_lj_uniform_velocity_reference_deck(tmp_path, oracle):
    """the reference's own generator (uniform 'loop geom' velocities)"""
    print("case: LJ 4000 atoms uniform velocities", flush=True)
    app = gpu_app(tmp_path, 30, n=10, gaussian=False, lanes=1)
    m = oracle_run(oracle, 0, n=10, gaussian=False)
    compare(app, m, 0.0, exact=True)
    app.run(30)
    m.run(30)
    compare(app, m, 0.0, exact=True)


def test_lj_mixed_force_error(tmp_path, oracle):
    """FP32 pair arithmetic on the FP64 positions: force error at step 0"""
    print("case: LJ 32k mixed precision force error", flush=True)
    app = gpu_app(tmp_path, 0, precision="mixed")
    m = oracle_run(oracle, 0)
    g = app.arrays()
    o = m.arrays()
    n = app.nlocal
    assert np.array_equal(g["x"], o["x"][:n])
    df = g["f"] - o["f"][:n]
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    print("mixed force RMS error", rms)
    assert rms <= 1e-5
    # identical pair set: the neighbour list is the FP64 one
    row, num, ent = app.neigh()
    orow, oent = m.csr()
    assert np.array_equal(ent, oent)


def test_lj_kernel_level_disordered(oracle):
    """displaced FCC (0.05 sigma): FP64 force against the oracle through the
    kernel-level C ABI; FP32 pair arithmetic error against FP64"""
    from examinimd_amd.pipeline import Pipeline, lattice
    from test_oracle import _energy_force
    print("case: LJ 32k disordered, kernel level", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    x = lattice("fcc", 20, a) + np.random.default_rng(5).normal(0, 0.05, (32000, 3))
    box = (20 * a,) * 3
    pl = Pipeline()
    pl.load(x, box)
    pl.build(2.8)
    f64 = pl.lj(emd.EMD_FP64)
    f32 = pl.lj(emd.EMD_MIXED)
    pl.close()
    E, F = _energy_force(oracle, x % (20 * a), 20 * a, "lj")
    assert rel_err(f64, F) <= 1e-13
    df = f32 - f64
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    print("mixed LJ force RMS error", rms, "force RMS", np.sqrt(np.mean(np.sum(f64 ** 2, 1))))
    assert rms <= 1e-5


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16])
def test_lj_mixed_force_error_every_lane_count(tmp_path, lanes):
    """emd_lj_set_lanes accepts 1, 2, 4, 8 and 16 lanes per atom; with fewer
    lanes each lane sums more pair terms in FP32.  On the LJ-32k melt (100
    mixed steps) the FP32 forces stay within 1e-5 RMS of the FP64 forces at
    the same positions with the same list, at every lane count"""
    import ctypes as C
    from examinimd_amd.abi import EMD_FP64, EMD_MIXED, EmdLJParams
    print(f"case: LJ 32k melt, mixed force error at {lanes} lanes", flush=True)
    app = gpu_app(tmp_path, 100, precision="mixed")
    app.run(100)
    L, ctx = emd.hip_lib(), app.ctx
    s = app.L.emd_app_system(app.h)
    nl = app.L.emd_app_neigh(app.h)
    p = EmdLJParams()
    p.ntypes = 1
    p.lj1[0][0], p.lj2[0][0], p.cutsq[0][0] = 48.0, 24.0, 6.25
    n = app.nlocal
    out = {}
    for prec in (EMD_FP64, EMD_MIXED):
        emd.check(ctx, L.emd_lj_set_lanes(ctx, prec, lanes), "set_lanes")
        emd.check(ctx, L.emd_force_zero(ctx, s, prec), "force_zero")
        emd.check(ctx, L.emd_lj_force(ctx, s, nl, C.byref(p), prec, 0), "lj_force")
        if prec == EMD_FP64:
            f = np.zeros((n, 3))
            emd.check(ctx, L.emd_system_download(ctx, s, n, None, None,
                                                 f.ctypes.data_as(C.POINTER(C.c_double)),
                                                 None, None, None), "download")
        else:
            f32 = np.zeros((n, 3), np.float32)
            emd.check(ctx, L.emd_memcpy_d2h(ctx, f32.ctypes.data_as(C.c_void_p),
                                            C.cast(s.contents.f32, C.c_void_p), f32.nbytes), "d2h")
            f = f32.astype(np.float64)
        out[prec] = f
    df = out[EMD_MIXED] - out[EMD_FP64]
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    print("mixed LJ force RMS error on the melt,", lanes, "lanes:", rms)
    assert rms <= 1e-5
    app.close()


def test_lost_atom_reported_by_binning():
    """an atom displaced by more than one box length (exchange wraps by one
    length only, as CommSerial::exchange does, src/comm_types/
    comm_serial.h:97-107) lies outside the bin grid: the binning reports
    EMD_ERR_OVERFLOW instead of binning it (the kernel clamps the bin index
    it flags, so nothing is written out of bounds)"""
    from examinimd_amd.pipeline import Pipeline, lattice
    print("case: lost atom, FCC 5^3 with one atom 3.5 box lengths away", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    box = 5 * a
    x = lattice("fcc", 5, a)
    x[17, 0] += 3.5 * box       # 2.5 box lengths out after the one wrap
    pl = Pipeline()
    pl.load(x, (box,) * 3)
    with pytest.raises(emd.EmdError) as err:
        pl.build(2.8)
    print(err.value)
    assert "status 4" in str(err.value) and "outside the bin grid" in str(err.value)
    pl.close()


def _dev(ctx, ptr, n, dtype):
    import ctypes as C
    out = np.zeros(n, dtype)
    if n:
        ctx.check(emd.hip_lib().emd_memcpy_d2h(ctx.h, out.ctypes.data_as(C.c_void_p),
                                                C.cast(ptr, C.c_void_p), out.nbytes), "d2h")
    return out


@pytest.mark.parametrize("layout", [emd.EMD_NEIGH_CSR, emd.EMD_NEIGH_2D])
def test_neighbour_rebuild_capacity_fallback(layout):
    """a rebuild whose rows outgrow the previous build's capacity (one-pass
    fill into padded rows overflows) gives the same list as a fresh build"""
    import ctypes as C
    from examinimd_amd.abi import EmdNeigh
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: neighbour capacity fallback, layout {layout}", flush=True)
    rng = np.random.default_rng(12)

    def system(rho):
        a = (4 / rho) ** (1 / 3)
        return lattice("fcc", 10, a) + rng.normal(0, 0.03, (4000, 3)), (10 * a,) * 3

    def rows(pl, nl):
        n = pl.s.n_local
        num = _dev(pl.ctx, nl.num, n, np.int32)
        if layout == emd.EMD_NEIGH_CSR:
            row = _dev(pl.ctx, nl.row, n + 1, np.int32)
            ent = _dev(pl.ctx, nl.entries, int(row[-1]), np.int32)
            return [ent[row[i]:row[i] + num[i]] for i in range(n)]
        ent = _dev(pl.ctx, nl.entries, n * nl.maxneighs, np.int32).reshape(n, nl.maxneighs)
        return [ent[i, :num[i]] for i in range(n)]

    def build(pl, nl, x, box):
        pl.load(x, box)
        L, c, s = pl.L, pl.ctx.h, C.byref(pl.s)
        pl.chk(L.emd_comm_exchange(c, s), "exchange")
        pl.chk(L.emd_binning(c, s, 2.8, 2.8, 2.8, 1, 1, 0, 1, C.byref(pl.b)), "bin")
        pl.chk(L.emd_comm_exchange_halo(c, s, 2.8, C.byref(pl.h)), "halo")
        pl.chk(L.emd_binning(c, s, 2.8, 2.8, 2.8, 1, 1, 1, 0, C.byref(pl.b)), "bin")
        pl.chk(L.emd_neigh_build(c, s, C.byref(pl.b), 2.8, 0, 0, layout, C.byref(nl)), "neigh")

    x1, b1 = system(0.45)
    x2, b2 = system(0.95)
    pl = Pipeline()
    nl = EmdNeigh()
    build(pl, nl, x1, b1)          # sparse: short rows set the capacity
    build(pl, nl, x2, b2)          # dense: rows ~2x longer than that capacity
    got = rows(pl, nl)
    fresh = EmdNeigh()
    build(pl, fresh, x2, b2)       # first build of a new list (exact passes)
    want = rows(pl, fresh)
    assert len(got) == len(want)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    for l in (nl, fresh):
        pl.L.emd_neigh_free(pl.ctx.h, C.byref(l))
    pl.close()


def test_lj_mixed_neighbour_identity_at_every_rebuild(tmp_path, oracle):
    """mixed-precision trajectory: at each of 5 rebuilds the GPU list equals
    the oracle's list built from the same (GPU) positions, entry for entry"""
    from helpers import oracle_csr_from
    print("case: LJ 32k mixed, neighbour identity at 5 rebuilds", flush=True)
    app = gpu_app(tmp_path, 100, precision="mixed")
    box = (20 * (4 / 0.8442) ** (1 / 3),) * 3
    for r in range(5):
        app.run(20)                       # step 20 of each chunk rebuilds
        a = app.arrays()
        row, num, ent = app.neigh()
        orow, oent, oid = oracle_csr_from(oracle, a["x"], a["id"], box)
        assert np.array_equal(oid, a["id"]), r       # same order after the oracle's sort
        assert np.array_equal(row, orow) and np.array_equal(ent, oent), r


@pytest.mark.parametrize("cut", [2.8, 6.0])
def test_counting_sort_binning_against_oracle(oracle, cut):
    """the counting-sort binning (13,500 atoms, beyond the one-launch
    path) against the oracle's BinSort restatement: counts, offsets and the
    permutation equal, sorted positions equal; bins of 6 sigma hold ~180
    atoms, so the ordering pass takes its insertion-sort branch"""
    import ctypes as C
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: counting-sort binning vs oracle, FCC 15^3, bins {cut}", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    nc = 15
    box = nc * a
    x = (lattice("fcc", nc, a) + np.random.default_rng(61).normal(0, 0.1, (4 * nc ** 3, 3))) % box
    pl = Pipeline()
    pl.load(x, (box,) * 3)
    L, c, s = pl.L, pl.ctx.h, C.byref(pl.s)
    pl.chk(L.emd_binning(c, s, cut, cut, cut, 1, 1, 0, 1, C.byref(pl.b)), "bin")
    nb, nr = pl.b.nbins, pl.b.n_range
    g = {"count": _dev(pl.ctx, pl.b.bincount, nb, np.int32),
         "offset": _dev(pl.ctx, pl.b.binoffsets, nb, np.int32),
         "permute": _dev(pl.ctx, pl.b.permute, nr, np.int32),
         "x": _dev(pl.ctx, pl.s.x, 3 * nr, np.float64)}
    maxc = pl.b.max_count
    pl.close()
    Lo = oracle.lib()
    so, bo = oracle.System(), oracle.Bins()
    Lo.or_system_init(C.byref(so))
    Lo.or_system_grow(C.byref(so), len(x))
    so.N_local = so.N = len(x)
    for d in range(3):
        so.domain[d] = so.hi[d] = box
    np.ctypeslib.as_array(so.x, shape=(3 * len(x),))[:] = x.ravel()
    Lo.or_binning(C.byref(so), C.byref(bo), cut, cut, cut, 1, 1, 0, 1)
    assert bo.nbins_total == nb
    o = {"count": oracle.arr(bo.count, nb, np.int32), "offset": oracle.arr(bo.offset, nb, np.int32),
         "permute": oracle.arr(bo.permute, nr, np.int32), "x": oracle.arr(so.x, 3 * nr)}
    Lo.or_bins_free(C.byref(bo))
    Lo.or_system_free(C.byref(so))
    print("largest bin", maxc)
    assert maxc == o["count"].max() and (cut < 6 or maxc > 64)
    for k in g:
        assert np.array_equal(g[k], o[k]), k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_step_boundary_integration_bitwise(tmp_path, precision):
    """final_integrate + the next initial_integrate in one launch gives the
    same bits as the two launches (60 steps: three rebuilds)"""
    print(f"case: LJ 32k, 60 steps, fused vs split integration, {precision}", flush=True)
    out = []
    for fused in (True, False):
        app = gpu_app(tmp_path, 60, precision=precision)
        app.set_fused_integrate(fused)
        app.run(60)
        out.append(app.arrays())
        app.close()
    for k in ("x", "v", "f"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_neighbour_list_candidates_beyond_one_lds_stage(oracle):
    """a cutoff of 4.0 sigma at rho 0.8442 gives ~1700 candidates per bin,
    more than one LDS staging round of the neighbour kernel holds: rows that
    continue across rounds equal the oracle's entry for entry"""
    from examinimd_amd.abi import EmdNeigh  # noqa: F401
    from examinimd_amd.pipeline import Pipeline, lattice
    from helpers import oracle_csr_from
    print("case: neighbour list, cut 4.0, 4000 atoms, multi-round staging", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    nc, cut = 10, 4.0
    box = (nc * a,) * 3
    x = lattice("fcc", nc, a) + np.random.default_rng(31).normal(0, 0.05, (4 * nc ** 3, 3))
    pl = Pipeline()
    pl.load(x, box)
    pl.build(cut)
    n = pl.s.n_local
    xs = _dev(pl.ctx, pl.s.x, 3 * n, np.float64).reshape(n, 3)
    ids = _dev(pl.ctx, pl.s.id, n, np.int32)
    num = _dev(pl.ctx, pl.nl.num, n, np.int32)
    row = _dev(pl.ctx, pl.nl.row, n + 1, np.int32)
    ent = _dev(pl.ctx, pl.nl.entries, int(row[-1]), np.int32)
    pl.close()
    orow, oent, oid = oracle_csr_from(oracle, xs, ids, box, cut=cut)
    assert np.array_equal(oid, ids)
    assert int(num.max()) > 200
    assert np.array_equal(row, orow) and np.array_equal(ent, oent)


@pytest.mark.parametrize("layout", [emd.EMD_NEIGH_CSR, emd.EMD_NEIGH_2D])
@pytest.mark.parametrize("lanes", [1, 16])
def test_lj_force_store_equals_zero_then_accumulate(oracle, layout, lanes):
    """emd_lj_force_store writes each local force without reading f: with f
    poisoned (NaN) beforehand it gives the bits of force_zero + lj_force"""
    import ctypes as C
    from examinimd_amd.abi import EmdLJParams
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: LJ store vs zero+accumulate, layout {layout}, {lanes} lanes", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    nc = 12
    x = lattice("fcc", nc, a) + np.random.default_rng(5).normal(0, 0.05, (4 * nc ** 3, 3))
    os.environ["EMD_LJ_LANES"] = str(lanes)
    try:
        pl = Pipeline()
    finally:
        os.environ.pop("EMD_LJ_LANES", None)
    pl.load(x, (nc * a,) * 3)
    L, c, s = pl.L, pl.ctx.h, C.byref(pl.s)
    pl.chk(L.emd_comm_exchange(c, s), "exchange")
    pl.chk(L.emd_binning(c, s, 2.8, 2.8, 2.8, 1, 1, 0, 1, C.byref(pl.b)), "bin")
    pl.chk(L.emd_comm_exchange_halo(c, s, 2.8, C.byref(pl.h)), "halo")
    pl.chk(L.emd_binning(c, s, 2.8, 2.8, 2.8, 1, 1, 1, 0, C.byref(pl.b)), "bin")
    pl.chk(L.emd_neigh_build(c, s, C.byref(pl.b), 2.8, 0, 0, layout, C.byref(pl.nl)), "neigh")
    p = EmdLJParams()
    p.ntypes = 1
    p.lj1[0][0], p.lj2[0][0], p.cutsq[0][0] = 48.0, 24.0, 6.25
    n = pl.s.n_local
    for prec in (emd.EMD_FP64, emd.EMD_MIXED):
        pl.chk(L.emd_force_zero(c, s, prec), "zero")
        pl.chk(L.emd_lj_force(c, s, C.byref(pl.nl), C.byref(p), prec, 0), "lj")
        want = pl.forces(prec)
        ntot = n + pl.s.n_ghost
        if prec == emd.EMD_MIXED:
            nan = np.full(3 * ntot, np.nan, np.float32)
            dst = pl.s.f32
        else:
            nan = np.full(3 * ntot, np.nan, np.float64)
            dst = pl.s.f
        pl.chk(L.emd_memcpy_h2d(c, C.cast(dst, C.c_void_p), nan.ctypes.data_as(C.c_void_p),
                                nan.nbytes), "poison")
        pl.chk(L.emd_lj_force_store(c, s, C.byref(pl.nl), C.byref(p), prec), "lj store")
        got = pl.forces(prec)
        assert np.isfinite(got).all()
        assert np.array_equal(got, want), prec
    pl.close()


@pytest.mark.parametrize("nc", [6, 9])
def test_binning_one_launch_equals_counting_sort(monkeypatch, nc):
    """binning of a small range in one workgroup launch gives the counts,
    offsets, permutation and largest count of the counting-sort path, for
    the sorting local binning and the local + ghost binning after the halo"""
    import ctypes as C
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: binning one launch vs counting sort, FCC {nc}^3", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    x = lattice("fcc", nc, a) + np.random.default_rng(47).normal(0, 0.1, (4 * nc ** 3, 3))
    out = []
    for one in ("1", "0"):
        monkeypatch.setenv("EMD_BIN_ONE", one)
        pl = Pipeline()
        pl.load(x, (nc * a,) * 3)
        L, c, s = pl.L, pl.ctx.h, C.byref(pl.s)
        pl.chk(L.emd_comm_exchange(c, s), "exchange")
        rec = {}
        for tag, ghost, sort in (("local", 0, 1), ("all", 1, 0)):
            if ghost:
                pl.chk(L.emd_comm_exchange_halo(c, s, 2.8, C.byref(pl.h)), "halo")
            pl.chk(L.emd_binning(c, s, 2.8, 2.8, 2.8, 1, 1, ghost, sort, C.byref(pl.b)), "bin")
            nb, nr = pl.b.nbins, pl.b.n_range
            rec[tag + "_count"] = _dev(pl.ctx, pl.b.bincount, nb, np.int32)
            rec[tag + "_offset"] = _dev(pl.ctx, pl.b.binoffsets, nb, np.int32)
            rec[tag + "_permute"] = _dev(pl.ctx, pl.b.permute, nr, np.int32)
            rec[tag + "_max"] = np.array([pl.b.max_count])
        n = pl.s.n_local
        rec["x"] = _dev(pl.ctx, pl.s.x, 3 * n, np.float64)
        rec["id"] = _dev(pl.ctx, pl.s.id, n, np.int32)
        out.append(rec)
        pl.close()
    assert out[0]["all_count"].sum() == len(out[0]["all_permute"])
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj32k_2d_equals_csr_default_lanes(tmp_path, precision):
    """the 2D list's first round loads row slots before the row length
    (clamped to the pitch and to existing atoms, masked by length): 100 steps
    give the bits of the CSR run at the default 16 lanes"""
    print(f"case: LJ 32k, 100 steps, 2D vs CSR, 16 lanes, {precision}", flush=True)
    out = []
    for neigh in ("2D", "CSR"):
        app = gpu_app(tmp_path, STEPS, neigh=neigh, precision=precision)
        app.run(STEPS)
        out.append(app.arrays())
        app.close()
    for k in ("x", "v", "f"):
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restart_from_binary_dump_continues_bitwise(tmp_path, precision):
    """--restart from a --dumpbinary file written after step 40 (a rebuild
    step) and 40 more steps give the bits of an uninterrupted 80-step run:
    the dump holds the sorted, wrapped positions the rebuild used, so the
    restarted run rebuilds the same list"""
    print(f"case: LJ 32k restart at step 40 vs 80 steps, {precision}", flush=True)
    deck = decks.write(str(tmp_path / "in.lj"), decks.lj_deck(n=N_CELLS, steps=80))
    dump_dir = tmp_path / "dump"
    dump_dir.mkdir()
    base = ["-il", deck, "--neigh-type", "CSR", "--comm-type", "SERIAL", "--precision", precision]
    a = emd.App(base + ["--dumpbinary", "40", str(dump_dir)]).init()
    a.run(40)
    a.close()
    dump = dump_dir / "output.0000000040.000"
    assert dump.exists()
    b = emd.App(base).init()
    b.run(80)
    want = b.arrays()
    b.close()
    c = emd.App(base + ["--restart", str(dump)]).init()
    c.run(40)
    got = c.arrays()
    c.close()
    for k in ("id", "x", "v", "f"):
        assert np.array_equal(got[k], want[k]), k


# ------------------------------------------------------- pair_style lj/cut/idial
@pytest.mark.parametrize("lanes,intensity", [(1, 3), (1, 16), (16, 3), (16, 16)])
def test_lj_idial_32k_100_steps(tmp_path, oracle, lanes, intensity):
    """lj/cut/idial through the driver against the oracle's restatement of
    ForceLJIDialNeigh: one lane per atom repeats the reference's operations
    in its order (bit-identical trajectory), 16 lanes change only each
    atom's summation order; no energy term (PE = 0)"""
    print(f"case: LJ idial 32k, dial {intensity}, {lanes} lanes", flush=True)
    app = gpu_app(tmp_path, STEPS, lanes=lanes, intensity=intensity)
    m = oracle_run(oracle, 0, intensity=intensity)
    if lanes == 1:
        compare(app, m, 0.0, exact=True)
    else:
        # the perfect lattice's forces are rounding noise (~1e-14) whose
        # value depends on the summation order: absolute bound
        assert np.abs(app.arrays()["f"] - m.arrays()["f"][:app.nlocal]).max() < 1e-12
    app.run(STEPS)
    m.run(STEPS)
    compare(app, m, 1e-12, exact=lanes == 1)
    assert app.thermo()[1] == 0.0 and m.thermo()[1] == 0.0


@pytest.mark.parametrize("newton", [False, True])
def test_lj_idial_half_list(tmp_path, oracle, newton):
    """half list: the reaction goes to local j only (ghost rows stay zero,
    so with newton on the folded ghost forces are zero as in the
    reference); atomic updates make the order, not the values, vary"""
    print(f"case: LJ idial 4000 atoms half list newton={newton}", flush=True)
    app = gpu_app(tmp_path, 40, n=10, half=True, newton=newton, intensity=4)
    m = oracle_run(oracle, 0, n=10, half=True, newton=newton, intensity=4)
    assert np.array_equal(app.arrays()["x"], m.arrays()["x"][:app.nlocal])
    assert np.abs(app.arrays()["f"] - m.arrays()["f"][:app.nlocal]).max() < 1e-12
    app.run(40)
    m.run(40)
    compare(app, m, 1e-12)


def test_lj_idial_kernel_level_mixed(oracle):
    """displaced FCC (0.05 sigma), dial 8: FP64 forces against the oracle,
    FP32 pair arithmetic and FP32 repeat sums against FP64 (RMS 1e-5)"""
    from examinimd_amd.pipeline import Pipeline, lattice
    from test_oracle import _pair_forces
    print("case: LJ idial 32k disordered, kernel level, dial 8", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    x = lattice("fcc", 20, a) + np.random.default_rng(9).normal(0, 0.05, (32000, 3))
    box = 20 * a
    pl = Pipeline()
    pl.load(x, (box,) * 3)
    pl.build(2.8)
    f64 = pl.lj_idial(8, emd.EMD_FP64)
    f32 = pl.lj_idial(8, emd.EMD_MIXED)
    lj = pl.lj(emd.EMD_FP64)
    pl.close()
    F = _pair_forces(oracle, x % box, box, "idial", 8)[0]
    assert rel_err(f64, F) <= 1e-13
    assert rel_err(f64, lj) <= 1e-13 and not np.array_equal(f64, lj)
    df = f32 - f64
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    print("mixed idial force RMS error", rms)
    assert rms <= 1e-5


@pytest.mark.parametrize("layout", [emd.EMD_NEIGH_CSR, emd.EMD_NEIGH_2D])
@pytest.mark.parametrize("jitter", [0.0, 1e-9])
def test_neighbour_list_pairs_at_the_cutoff(oracle, layout, jitter):
    """SC lattice with spacing cut/2: the second neighbours along each axis
    sit at the cutoff to within rounding, where the inclusive FP64 test
    (rsq <= cutsq, neighbor_csr.h:206) decides pair by pair; rows and
    entries equal the oracle's"""
    from examinimd_amd.pipeline import Pipeline, lattice
    from test_oracle import oracle_csr
    print(f"case: neighbour list at the cutoff, layout={layout} jitter={jitter}", flush=True)
    n, a, cut = 12, 1.4, 2.8
    x = lattice("sc", n, a)
    if jitter:
        x = x + np.random.default_rng(2).normal(0, jitter, x.shape)
    box = n * a
    orow, oent, oghost = oracle_csr(oracle, x % box, box, cut)
    pl = Pipeline()
    pl.load(x, (box,) * 3)
    pl.build(cut, layout=layout)
    row, num, ent = emd.download_neigh(pl.ctx.h, pl.nl)
    ghosts = pl.s.n_ghost
    pl.close()
    assert ghosts == oghost
    assert np.array_equal(np.asarray(row, np.int64), orow.astype(np.int64))
    assert np.array_equal(ent, oent)
    # 26 neighbours are inside; of the 6 axis pairs at 2a some land on each
    # side of the FP64 test (30 per atom unjittered, 29.05 jittered)
    assert 26 * len(x) < len(oent) < 32 * len(x)


@pytest.mark.parametrize("layout", [emd.EMD_NEIGH_CSR, emd.EMD_NEIGH_2D])
def test_ragged_rows_dense_patch_and_isolated_atoms(oracle, layout):
    """a dense FCC patch (2048 atoms, 8^3 cells at rho 0.8442) in one corner
    of a 40-sigma box, 64 isolated atoms on a 5-sigma grid far from it and 8
    dimers: rows of 0, 1, surface and bulk length, most bins empty.  Lists
    equal the oracle's; FP64 forces match it, isolated atoms get exactly
    zero force and each dimer equal and opposite forces"""
    from examinimd_amd.pipeline import Pipeline, lattice
    from test_oracle import _energy_force, oracle_csr
    print(f"case: ragged rows, layout={layout}", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    patch = lattice("fcc", 8, a)                     # spans [0, 13.44)
    g = np.array([18.0, 23.0, 28.0, 33.0])         # >= 4.5 from the patch, also across the wrap
    iso = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(11)
    iso = iso + rng.uniform(-0.3, 0.3, iso.shape)
    d = rng.normal(0, 1, (8, 3))
    dimer = iso[:8] + 1.1 * d / np.linalg.norm(d, axis=1, keepdims=True)
    x = np.concatenate([patch, iso, dimer])
    box, cut = 40.0, 2.8
    orow, oent, _ = oracle_csr(oracle, x, box, cut)
    pl = Pipeline()
    pl.load(x, (box,) * 3)
    pl.build(cut, layout=layout)
    row, num, ent = emd.download_neigh(pl.ctx.h, pl.nl)
    f64 = pl.lj(emd.EMD_FP64)
    f32 = pl.lj(emd.EMD_MIXED)
    pl.close()
    assert np.array_equal(np.asarray(row, np.int64), orow.astype(np.int64))
    assert np.array_equal(ent, oent)
    lens = np.diff(orow)
    assert lens.min() == 0 and (lens == 1).sum() == 16 and lens.max() > 70
    E, F = _energy_force(oracle, x, box, "lj")
    assert rel_err(f64, F) <= 1e-13
    np_, ni = len(patch), len(iso)
    assert not f64[np_ + 8:np_ + ni].any() and not f32[np_ + 8:np_ + ni].any()
    assert np.array_equal(f64[np_:np_ + 8], -f64[np_ + ni:])
    df = f32 - f64
    assert float(np.sqrt(np.mean(np.sum(df * df, axis=1)))) <= 1e-5


@pytest.mark.parametrize("layout,half", [(emd.EMD_NEIGH_CSR, 0), (emd.EMD_NEIGH_2D, 0),
                                         (emd.EMD_NEIGH_CSR, 1)])
def test_lj_two_types_against_brute_force(layout, half):
    """two atom types with different eps, sigma and cutoff per type pair
    (the per-pair table path of the full- and half-list LJ kernels, the half
    list with newton off): FP64 forces against a brute-force pair sum, FP32
    pair arithmetic within the mixed bound"""
    import ctypes as C
    from examinimd_amd.abi import EmdLJParams
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: LJ two types, layout {layout}, half {half}", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    nc = 6
    rng = np.random.default_rng(21)
    x = lattice("fcc", nc, a) + rng.normal(0, 0.05, (4 * nc ** 3, 3))
    box = nc * a
    x %= box
    types = rng.integers(0, 2, len(x)).astype(np.int32)
    eps = np.array([[1.0, 0.7], [0.7, 1.3]])
    sig = np.array([[1.0, 1.1], [1.1, 0.9]])
    cut = np.array([[2.5, 2.7], [2.7, 2.3]])
    lj1, lj2, cutsq = 48.0 * eps * sig ** 12, 24.0 * eps * sig ** 6, cut * cut
    F = pair_forces_typed(x, box, types, lj1, lj2, cutsq)
    pl = Pipeline()
    pl.load(x, (box,) * 3, types=types)
    pl.build(2.8, half=half, layout=layout)
    p = EmdLJParams()
    p.ntypes = 2
    for t1 in range(2):
        for t2 in range(2):
            p.lj1[t1][t2], p.lj2[t1][t2], p.cutsq[t1][t2] = lj1[t1, t2], lj2[t1, t2], cutsq[t1, t2]
    L, c, s = pl.L, pl.ctx.h, C.byref(pl.s)
    out = {}
    for prec in (emd.EMD_FP64, emd.EMD_MIXED):
        pl.chk(L.emd_force_zero(c, s, prec), "zero")
        pl.chk(L.emd_lj_force(c, s, C.byref(pl.nl), C.byref(p), prec, half), "lj")
        out[prec] = pl.forces(prec)
    pl.close()
    assert rel_err(out[emd.EMD_FP64], F) <= 1e-12
    df = out[emd.EMD_MIXED] - out[emd.EMD_FP64]
    assert float(np.sqrt(np.mean(np.sum(df * df, axis=1)))) <= 1e-5


def _many_type_coeffs(rng, nt):
    eps = rng.uniform(0.6, 1.4, (nt, nt))
    sig = rng.uniform(0.85, 1.1, (nt, nt))
    cut = rng.uniform(2.0, 2.7, (nt, nt))
    eps, sig, cut = (np.triu(m) + np.triu(m, 1).T for m in (eps, sig, cut))
    return 48.0 * eps * sig ** 12, 24.0 * eps * sig ** 6, cut * cut


def _pair_energy_typed(x, box, types, lj1, lj2, cutsq):
    """the reference's shifted pair energy per ordered pair of a full list,
    0.5 (r6 (0.5 lj1 r6 - lj2) / 6 - the same at rc) (force_lj_neigh_impl.h:
    287-293), brute force over minimum images"""
    d = x[:, None, :] - x[None, :, :]
    d -= box * np.round(d / box)
    rsq = np.sum(d * d, axis=2)
    ti, tj = types[:, None], types[None, :]
    c, a, b = cutsq[ti, tj], lj1[ti, tj], lj2[ti, tj]
    m = (rsq < c) & ~np.eye(len(x), dtype=bool)
    r6 = np.where(m, 1.0 / np.where(m, rsq, 1.0), 0.0) ** 3
    rc6 = (1.0 / c) ** 3
    e = 0.5 * (r6 * (0.5 * a * r6 - b) / 6.0 - rc6 * (0.5 * a * rc6 - b) / 6.0)
    return float(np.sum(np.where(m, e, 0.0)))


@pytest.mark.parametrize("layout,half", [(emd.EMD_NEIGH_CSR, 0), (emd.EMD_NEIGH_2D, 0),
                                         (emd.EMD_NEIGH_CSR, 1)])
def test_lj_twenty_types_table_against_brute_force(layout, half):
    """20 atom types, more than the fixed 16 x 16 arrays hold: the
    coefficients come through emd_lj_params.table (the reference's View path
    for many types, force_lj_neigh_impl.h:100-126).  FP64 forces and (full
    lists) the energy against brute-force pair sums, FP32 pair arithmetic
    within the mixed bound; without the table 20 types are refused"""
    from examinimd_amd.abi import EmdLJParams
    from examinimd_amd.pipeline import Pipeline, lattice
    print(f"case: LJ 20 types (table), layout {layout}, half {half}", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    nc, nt = 6, 20
    rng = np.random.default_rng(44)
    x = lattice("fcc", nc, a) + rng.normal(0, 0.05, (4 * nc ** 3, 3))
    box = nc * a
    x %= box
    types = rng.integers(0, nt, len(x)).astype(np.int32)
    lj1, lj2, cutsq = _many_type_coeffs(rng, nt)
    F = pair_forces_typed(x, box, types, lj1, lj2, cutsq)
    pl = Pipeline()
    pl.load(x, (box,) * 3, types=types)
    pl.build(2.8, half=half, layout=layout)
    table = np.ascontiguousarray(np.concatenate([lj1.ravel(), lj2.ravel(), cutsq.ravel()]))
    p = EmdLJParams()
    p.ntypes = nt
    L, c, s = pl.L, pl.ctx.h, C.byref(pl.s)
    assert L.emd_lj_force(c, s, C.byref(pl.nl), C.byref(p), emd.EMD_FP64, half) == 1
    p.table = table.ctypes.data_as(C.POINTER(C.c_double))
    out = {}
    for prec in (emd.EMD_FP64, emd.EMD_MIXED):
        pl.chk(L.emd_force_zero(c, s, prec), "zero")
        pl.chk(L.emd_lj_force(c, s, C.byref(pl.nl), C.byref(p), prec, half), "lj")
        out[prec] = pl.forces(prec)
    if not half:
        e = C.c_double()
        pl.chk(L.emd_lj_energy(c, s, C.byref(pl.nl), C.byref(p), 0, C.byref(e)), "energy")
        E = _pair_energy_typed(x, box, types, lj1, lj2, cutsq)
        print("energy", e.value, E)
        assert abs(e.value - E) <= 1e-12 * abs(E)
    pl.close()
    print("force rel err", rel_err(out[emd.EMD_FP64], F))
    assert rel_err(out[emd.EMD_FP64], F) <= 1e-12
    df = out[emd.EMD_MIXED] - out[emd.EMD_FP64]
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    frms = float(np.sqrt(np.mean(np.sum(F * F, axis=1))))
    print("mixed force RMS error", rms, "force RMS", frms)
    assert rms <= 1e-6 * frms


def test_lj_twenty_types_through_the_driver(tmp_path):
    """a deck with create_box 20 and one pair_coeff line per type pair: the
    driver's LJ module builds the coefficient table, and its set-up forces
    equal a brute-force pair sum over its own positions and types"""
    print("case: LJ 20 types, driver deck, set-up forces", flush=True)
    nt, n = 20, 5
    rng = np.random.default_rng(7)
    eps = rng.uniform(0.6, 1.4, (nt, nt))
    sig = rng.uniform(0.85, 1.1, (nt, nt))
    cut = rng.uniform(2.0, 2.5, (nt, nt))
    lines = [f"mass            {t + 1} 1.0" for t in range(nt)]
    for i in range(nt):
        for j in range(i, nt):
            lines.append(f"pair_coeff      {i + 1} {j + 1} {eps[i, j]:.17g} {sig[i, j]:.17g} "
                         f"{cut[i, j]:.17g}")
    deck = "\n".join(["units           lj", "atom_style      atomic", "newton          off",
                      "lattice         fcc 0.8442", f"region          box block 0 {n} 0 {n} 0 {n}",
                      f"create_box      {nt} box", "create_atoms    1 box"] + lines[:nt] +
                     ["velocity        all create 1.0 87287 loop geom",
                      "pair_style      lj/cut 2.5"] + lines[nt:] +
                     ["neighbor        0.3 bin", "neigh_modify    delay 0 every 20 check no",
                      "fix             1 all nve", "timestep        0.005", "thermo          0",
                      "run             0"]) + "\n"
    path = decks.write(str(tmp_path / "in.lj20"), deck)
    app = emd.App(["-il", path, "--neigh-type", "CSR", "--precision", "fp64"]).init()
    g = app.arrays()
    app.close()
    up = lambda m: np.triu(m) + np.triu(m, 1).T
    eps, sig, cut = up(eps), up(sig), up(cut)
    box = n * (4 / 0.8442) ** (1 / 3)
    t = g["type"]
    assert t.min() >= 0 and t.max() < nt and len(np.unique(t)) > 16
    F = pair_forces_typed(g["x"], box, t, 48.0 * eps * sig ** 12, 24.0 * eps * sig ** 6,
                          cut * cut)
    print("force rel err", rel_err(g["f"], F))
    assert rel_err(g["f"], F) <= 1e-12
