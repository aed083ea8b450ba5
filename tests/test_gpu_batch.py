"""Batched periodic LJ replicas (emd_batch_*) against the CPU oracle.

Each replica is an independent periodic box; per replica the batch must
give what the reference's serial-comm LJ loop gives for that box alone:
  * neighbour sets: exactly the (j, periodic image) pairs within rc + skin
    found by brute force over all 27 images (bit-exact set identity);
  * FP64: positions, velocities and forces within 1e-12 relative of the
    oracle loop over 60 steps (3 rebuilds); the rows are in a different
    order than the oracle's ghost rows, so sums agree to rounding only;
  * mixed (FP32 pair math and forces, FP64 state): force RMS error against
    the FP64 batch at the same positions <= 1e-5 (reduced units);
  * graph replay and eager launches give bit-identical trajectories;
  * autotuning leaves the state untouched.
"""
import numpy as np
import pytest

from examinimd_amd import batch
from examinimd_amd.abi import EMD_FP64, EMD_MIXED
from helpers import oracle_lj_md, rel_err

pytestmark = pytest.mark.gpu


def small_ensemble(seed=11):
    """uneven boxes incl. one with only 2 bins per axis (images repeat bins)"""
    rng = np.random.default_rng(seed)
    spec = [(4, 0.95, 1.5), (5, 0.80, 0.9), (6, 0.70, 2.0), (4, 0.72, 1.2), (7, 0.88, 1.1)]
    xs, vs, boxes = [], [], []
    for nc, rho, t in spec:
        x, L = batch.fcc(nc, rho)
        x = (x + rng.normal(0, 0.03, x.shape)) % L
        xs.append(x)
        vs.append(batch.velocities(len(x), t, rng))
        boxes.append((L, L, L))
    n = np.array([len(x) for x in xs], np.int32)
    return n, np.array(boxes), np.concatenate(xs), np.concatenate(vs)


def make(n, box, x, v, precision, lanes=4, block=256):
    b = batch.Batch(n, box, precision=precision)
    b.set_tile(lanes, block)
    b.upload(x, v)
    b.setup()
    return b


def split(a, n):
    return np.split(a, np.cumsum(n)[:-1])


def test_batch_neighbour_sets_brute_force():
    print("case: batch neighbour sets, 5 replicas", flush=True)
    n, box, x, v = small_ensemble()
    b = make(n, box, x, v, EMD_FP64)
    xw, _, _ = b.download()
    num, ent = b.neigh()
    start = np.concatenate([[0], np.cumsum(n)])
    cut2 = 2.8 ** 2
    shifts = np.array([(a, c, d) for a in (-1, 0, 1) for c in (-1, 0, 1) for d in (-1, 0, 1)])
    for r in range(len(n)):
        L = box[r]
        xr = xw[start[r]:start[r + 1]]
        assert np.all(xr >= 0) and np.all(xr <= L)
        for i in range(0, len(xr), 7):
            gi = start[r] + i
            # brute force: every j and every image j + s L
            d = xr[i][None, None, :] - (xr[None, :, :] + shifts[:, None, :] * L)
            r2 = (d * d).sum(-1)
            ks, js = np.nonzero(r2 <= cut2)
            want = {(start[r] + j, tuple(shifts[k])) for k, j in zip(ks, js)
                    if not (j == i and not shifts[k].any())}
            row = ent[gi, :num[gi]]
            got = set()
            for e in row:
                code = int(e) >> batch.CODE_SHIFT
                s = tuple(0 if c == 0 else (-1 if c == 1 else 1)
                          for c in ((code >> 2 * a) & 3 for a in range(3)))
                got.add((int(e) & batch.JMASK, s))
            assert len(got) == len(row) == num[gi]
            assert got == want, (r, i)
    b.close()


@pytest.mark.parametrize("lanes,block", [(1, 256), (8, 128)])
def test_batch_fp64_matches_oracle(oracle, lanes, block):
    print(f"case: batch FP64 vs oracle, lanes {lanes} block {block}", flush=True)
    n, box, x, v = small_ensemble()
    b = make(n, box, x, v, EMD_FP64, lanes, block)
    _, _, f0 = b.download()
    b.run(60)
    xg, vg, fg = b.download()
    step, mrow, cap = b.info()
    assert step == 60 and mrow <= cap
    for r, (xr, vr, x0, v0, f0r) in enumerate(zip(split(xg, n), split(vg, n), split(x, n),
                                                  split(v, n), split(f0, n))):
        _, _, fo0 = oracle_lj_md(oracle, x0, v0, box[r], 0)
        assert rel_err(f0r, fo0) < 1e-13, r
        xo, vo, fo = oracle_lj_md(oracle, x0, v0, box[r], 60)
        # oracle positions are wrapped at rebuilds, batch positions too
        errs = (rel_err(xr, xo), rel_err(vr, vo), rel_err(split(fg, n)[r], fo))
        print("replica", r, "x v f relative errors", errs)
        assert max(errs) <= 1e-12, r
    b.close()


def test_batch_graph_equals_eager():
    print("case: batch graph replay vs eager launches", flush=True)
    n, box, x, v = small_ensemble(seed=5)
    out = []
    for graph in (True, False):
        b = make(n, box, x, v, EMD_MIXED)
        b.run(7, graph=graph)    # ends before the first rebuild step
        b.run(75, graph=graph)   # crosses rebuilds, several full periods
        out.append(b.download() + b.thermo())
        b.close()
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)


def test_batch_mixed_force_error_and_energy():
    print("case: batch mixed force error and energy drift", flush=True)
    n, box, x, v = small_ensemble(seed=8)
    b64 = make(n, box, x, v, EMD_FP64)
    b32 = make(n, box, x, v, EMD_MIXED)
    _, _, f64 = b64.download()
    _, _, f32 = b32.download()
    df = f32 - f64
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    print("batch mixed force RMS error", rms)
    assert rms <= 1e-5
    e0 = sum(b32.thermo())
    b32.run(200)
    e1 = sum(b32.thermo())
    b64.run(200)
    e1d = sum(b64.thermo())
    drift = np.abs(e1 - e0) / n
    drift64 = np.abs(e1d - e0) / n
    print("energy change per atom mixed", drift.max(), "fp64", drift64.max())
    # the change is the Verlet error of the hot, dilute boxes (dt 0.005),
    # which FP32 pair arithmetic must not alter
    assert np.abs(e1 - e1d).max() / n.min() < 1e-5
    assert drift64.max() < 5e-3
    b64.close()
    b32.close()


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16])
def test_batch_mixed_force_error_every_lane_count(lanes):
    """the autotuner may pick any of these lane counts in mixed precision:
    the FP32 force error bound (1e-5 RMS, reduced units) holds at each, on
    boxes melted for 100 FP64 steps (the melt is where forces are largest)"""
    print(f"case: batch mixed force error at {lanes} lanes", flush=True)
    n, box, x, v = small_ensemble(seed=8)
    b = make(n, box, x, v, EMD_FP64, 4, 256)
    b.run(100)
    xm, vm, f64 = b.download()
    b.close()
    b32 = make(n, box, xm, vm, EMD_MIXED, lanes, 128)
    _, _, f32 = b32.download()
    b32.close()
    df = f32 - f64
    rms = float(np.sqrt(np.mean(np.sum(df * df, axis=1))))
    print("batch mixed force RMS error", lanes, "lanes:", rms)
    assert rms <= 1e-5


def test_batch_autotune_keeps_state():
    print("case: batch autotune", flush=True)
    n, box, x, v = small_ensemble(seed=9)
    b = make(n, box, x, v, EMD_MIXED)
    b.run(5)
    before = b.download()
    lanes, block, ms = b.autotune(trial_steps=3)
    assert lanes in (1, 2, 4, 8, 16) and block in (64, 128, 256, 512) and ms > 0
    after = b.download()
    for a, c in zip(before, after):
        assert np.array_equal(a, c)
    assert b.info()[0] == 5
    b.run(30)
    b.close()


def test_batch_uneven_ensemble_runs():
    print("case: 256-replica uneven ensemble, 45 steps", flush=True)
    e = batch.ensemble(nrep=256, seed=4)
    b = batch.Batch(e["natoms"], e["box"], precision=EMD_MIXED)
    b.upload(e["x"], e["v"])
    b.setup()
    ke0, pe0 = b.thermo()
    b.run(45)
    ke1, pe1 = b.thermo()
    step, mrow, cap = b.info()
    assert step == 45 and mrow <= cap
    drift = np.abs((ke1 + pe1) - (ke0 + pe0)) / e["natoms"]
    print("max energy change per atom", drift.max())
    assert np.all(np.isfinite(ke1)) and drift.max() < 5e-3
    b.close()


def test_replica_loop_matches_batch():
    """the unbatched reference point (one replica after another through the
    per-system C ABI) integrates the same trajectories as the batch"""
    print("case: per-replica launch loop vs batch, FP64, 45 steps", flush=True)
    n, box, x, v = small_ensemble(seed=3)
    sec, xl, vl = batch.replica_loop(n, box, x, v, 45, precision=EMD_FP64)
    b = make(n, box, x, v, EMD_FP64)
    b.run(45)
    xb, vb, _ = b.download()
    assert sec > 0
    for r in range(len(n)):
        errs = (rel_err(split(xl, n)[r], split(xb, n)[r]), rel_err(split(vl, n)[r], split(vb, n)[r]))
        print("replica", r, "loop vs batch x v relative errors", errs)
        assert max(errs) <= 1e-12, r
    b.close()


def test_batch_row_overflow_grows_and_rebuilds():
    """one replica implodes (velocities toward its centre, 14 sigma/tau at the
    faces): its rows outgrow the capacity set up on the lattice, and the
    rebuilds at steps 20 (before a graph-replayed period) and 40 (the run's
    last step, eager) grow the rows and rebuild; the rows of step 40 are then
    exactly the brute-force neighbour sets over all 27 images"""
    print("case: batch rebuild with a row overflow, grow and rebuild", flush=True)
    n, box, x, v = small_ensemble(seed=13)
    start = np.concatenate([[0], np.cumsum(n)])
    r = 2
    L = box[r][0]
    v[start[r]:start[r + 1]] = -3.5 * (x[start[r]:start[r + 1]] - L / 2)
    b = make(n, box, x, v, EMD_FP64)
    cap0 = b.info()[2]
    b.run(40)                     # ends with the step-40 rebuild and its force
    step, mrow, cap = b.info()
    print("capacity", cap0, "->", cap, "longest row", mrow)
    assert step == 40 and cap > cap0 and mrow <= cap
    xw, _, _ = b.download()
    num, ent = b.neigh()
    b.close()
    xr = xw[start[r]:start[r + 1]]
    shifts = np.array([(p, q, t) for p in (-1, 0, 1) for q in (-1, 0, 1) for t in (-1, 0, 1)])
    for i in range(len(xr)):
        d = xr[i][None, None, :] - (xr[None, :, :] + shifts[:, None, :] * L)
        ks, js = np.nonzero((d * d).sum(-1) <= 2.8 ** 2)
        want = {(start[r] + j, tuple(shifts[k])) for k, j in zip(ks, js)
                if not (j == i and not shifts[k].any())}
        row = ent[start[r] + i, :num[start[r] + i]]
        got = set()
        for e in row:
            code = int(e) >> batch.CODE_SHIFT
            sh = tuple(0 if c == 0 else (-1 if c == 1 else 1)
                       for c in ((code >> 2 * a) & 3 for a in range(3)))
            got.add((int(e) & batch.JMASK, sh))
        assert got == want and len(row) == len(want), i
