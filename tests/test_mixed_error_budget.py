"""Error budget of the mixed-precision LJ force (CPU model, no GPU).

The bound is a force RMS error <= 1e-5 (reduced units) against FP64 at the
same positions.  On an LJ melt (32,000 atoms, density 0.8442, 200 steps from
T* = 1.44 on the CPU oracle, T* ~ 0.7, force RMS ~ 24) the FP32 operations of
the pair force are modelled one precision choice at a time, with FP32
rounding after every operation, FP32 fused multiply-add accumulation in row
order (one lane per atom: the longest FP32 sums), against the FP64 sum:

  kernel   FP64 distance, 1/r^2 and r^-6; r^-6, 1/r^2 and dx rounded to FP32;
           FP32 pair factor and accumulation        -> fits (4.3e-6; the GPU
           kernel at one lane per atom measures 4.1e-6)
  rsq32    r^2 rounded to FP32, then all FP32        -> 1.06e-5, exceeds 1e-5
  dx32     dx rounded to FP32, r^2 and on in FP32    -> 1.6e-5
  x32      positions in FP32, all FP32               -> exceeds 30-fold

fpair ~ r^-14 turns the relative rounding of r^2 (2^-24) into ~7 times
that in the pair force, and pair forces up to ~100 cancel to a net force of
~24, so only the first form fits.  The distance work therefore stays FP64 in
the mixed kernel, which bounds what FP32 arithmetic can save there
(force_lj_neigh_impl.h:161-206 is the pair force modelled)."""
import numpy as np
import pytest

LJ1, LJ2, CUTSQ = 48.0, 24.0, 6.25


@pytest.fixture(scope="module")
def melt(oracle):
    cfg = oracle.make_config(n=(20, 20, 20), gaussian=True)
    m = oracle.MD(cfg)
    m.init()
    m.run(200)
    a = m.arrays()
    row, ent = m.csr()
    n = m.system.N_local
    m.close()
    x = a["x"]
    cnt = np.diff(row)
    P = int(cnt.max())
    J = np.full((n, P), -1, np.int64)
    col = np.arange(P)[None, :]
    mask = col < cnt[:, None]
    J[mask] = ent
    return x, n, J, mask


def _f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32)


def forces(x, n, J, mask, mode):
    """per-atom forces [n][3] with the given precision choices"""
    xi = x[:n, None, :]
    xj = x[np.where(mask, J, 0)]
    if mode == "fp64":
        d = xi - xj
        rsq = np.sum(d * d, axis=2)
        inside = mask & (rsq < CUTSQ)
        r2 = 1.0 / np.where(inside, rsq, 1.0)
        r6 = r2 * r2 * r2
        fpair = np.where(inside, r6 * (LJ1 * r6 - LJ2) * r2, 0.0)
        return np.sum(d * fpair[:, :, None], axis=1)
    if mode == "x32":
        d = (_f32(xi) - _f32(xj)).astype(np.float32)
    else:
        d = xi - xj
    if mode in ("dx32", "x32"):
        d = _f32(d)
        rsq = _f32(_f32(_f32(d[..., 0] * d[..., 0]) + _f32(d[..., 1] * d[..., 1]))
                   + _f32(d[..., 2] * d[..., 2]))
    else:
        rsq = np.sum(d * d, axis=2)
    inside = mask & (rsq < CUTSQ)                    # the pair set: FP64 test unless x32/dx32
    safe = np.where(inside, rsq, 1.0)
    if mode == "kernel":
        r2 = 1.0 / safe
        r6 = r2 * r2 * r2
        r6f, r2f = _f32(r6), _f32(r2)
        dxf = _f32(d)
    else:                                            # rsq32, dx32, x32: all FP32 from r^2 on
        rsqf = _f32(safe)
        r2f = _f32(np.float32(1.0) / rsqf)
        r6f = _f32(_f32(r2f * r2f) * r2f)
        dxf = _f32(d)
    lj1, lj2 = np.float32(LJ1), np.float32(LJ2)
    fpair = _f32(_f32(r6f * _f32(_f32(lj1 * r6f) - lj2)) * r2f)
    fpair = np.where(inside, fpair, np.float32(0.0))
    acc = np.zeros((n, 3), np.float32)
    for k in range(J.shape[1]):                      # FP32 FMA in row order
        acc = _f32(dxf[:, k, :].astype(np.float64) * fpair[:, k, None].astype(np.float64)
                   + acc.astype(np.float64))
    return acc.astype(np.float64)


def _rms(a, b):
    d = a - b
    return float(np.sqrt(np.mean(np.sum(d * d, axis=1))))


def test_fp64_model_matches_oracle_forces(melt, oracle):
    """the FP64 model reproduces the oracle's forces (pins the model)"""
    x, n, J, mask = melt
    f = forces(x, n, J, mask, "fp64")
    cfg = oracle.make_config(n=(20, 20, 20), gaussian=True)
    m = oracle.MD(cfg)
    m.init()
    m.run(200)
    fo = m.arrays()["f"][:n]
    m.close()
    assert np.max(np.abs(f - fo)) <= 1e-12 * np.max(np.abs(fo))
    assert 15 < np.sqrt(np.mean(np.sum(fo * fo, axis=1))) < 40      # a melt, not a lattice


def test_error_budget(melt):
    x, n, J, mask = melt
    ref = forces(x, n, J, mask, "fp64")
    err = {m: _rms(forces(x, n, J, mask, m), ref) for m in ("kernel", "rsq32", "dx32", "x32")}
    print("force RMS error by precision choice:", err)
    assert err["kernel"] <= 6e-6          # the mixed kernel's choices fit with margin
    assert err["rsq32"] > 1e-5            # rounding r^2 to FP32 alone breaks the bound
    assert err["dx32"] > 1e-5
    assert err["x32"] > 1e-4
