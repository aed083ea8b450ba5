"""Host-side pieces of the batched-replica path (no GPU): the seeded
uneven-size ensemble and the per-replica oracle loop that the GPU batch is
checked against."""
import numpy as np

from examinimd_amd import batch
from helpers import oracle_lj_md, rel_err


def test_ensemble_sizes_densities_temperatures():
    e = batch.ensemble(nrep=256, seed=4)
    n, nc, L = e["natoms"], e["ncell"], e["box"][:, 0]
    assert len(n) == 256
    assert nc.min() >= 4 and nc.max() <= 12 and np.all(n == 4 * nc ** 3)
    assert np.allclose(n / L ** 3, e["rho"], rtol=1e-12)
    assert e["rho"].min() >= 0.70 and e["rho"].max() <= 0.95
    assert e["temp"].min() >= 0.8 and e["temp"].max() <= 2.0
    # every box holds the +-1 image stencil of the neighbour cutoff 2.8
    assert L.min() >= 2 * 2.8
    # uneven sizes: log-uniform targets give many small and some large boxes
    assert len(np.unique(n)) >= 6 and n.min() <= 500 and n.max() >= 4000
    start = np.concatenate([[0], np.cumsum(n)])
    for r in range(256):
        x = e["x"][start[r]:start[r + 1]]
        v = e["v"][start[r]:start[r + 1]]
        assert np.all(x >= 0) and np.all(x < L[r])
        assert np.abs(v.sum(0)).max() < 1e-10
        t = (v * v).sum() / (3 * len(v) - 3)
        assert abs(t - e["temp"][r]) < 1e-12 * e["temp"][r]


def test_oracle_loop_helper_matches_driver(oracle):
    """the per-replica oracle loop used for the batch parity tests reproduces
    the oracle's own config-driven MD driver from the same initial state"""
    cfg = oracle.make_config(n=(6, 6, 6), every=20)
    m = oracle.MD(cfg)
    m.init()
    s = m.system
    n = s.N_local
    a = m.arrays()
    order = np.argsort(a["id"][:n])
    x0, v0 = a["x"][:n][order], a["v"][:n][order]
    box = tuple(s.domain[d] for d in range(3))
    m.run(50)
    b = m.arrays()
    o = np.argsort(b["id"][:n])
    x, v, f = oracle_lj_md(oracle, x0, v0, box, 50)
    assert rel_err(x, b["x"][:n][o]) < 1e-12
    assert rel_err(v, b["v"][:n][o]) < 1e-12
    assert rel_err(f, b["f"][:n][o]) < 1e-12
