"""The reference's oracle files across implementations (src/examinimd.cpp:
296-482): the GPU driver's --dumpbinary files read by the CPU oracle's
dump comparison, the GPU driver's --correctness check against dumps the
oracle wrote, and the canonical neighbour-list dump (--dumpneigh) against
the oracle's list.  One lane per atom: the FP64 GPU path then reproduces the
oracle's arithmetic, so every difference is exactly zero."""
import ctypes as C
import os

import numpy as np
import pytest

import examinimd_amd as emd
from examinimd_amd import decks
from test_gpu_lj import oracle_run

pytestmark = pytest.mark.gpu


def _app(tmp_path, extra, steps, precision="fp64", neigh="CSR", lanes=1):
    deck = decks.write(str(tmp_path / f"in.lj.{precision}.{neigh}"), decks.lj_deck(n=20, steps=steps))
    os.environ["EMD_LJ_LANES"] = str(lanes)
    try:
        return emd.App(["-il", deck, "--neigh-type", neigh, "--precision", precision] +
                       [str(a) for a in extra]).init()
    finally:
        os.environ.pop("EMD_LJ_LANES", None)


def _dump_name(d, step, stem="output"):
    return os.path.join(str(d), f"{stem}.{step:010d}.000")


def test_gpu_dumpbinary_read_by_oracle(tmp_path, oracle):
    print("case: LJ 32k, GPU --dumpbinary every 10 steps, oracle reads", flush=True)
    d = tmp_path / "gpu"
    d.mkdir()
    app = _app(tmp_path, ["--dumpbinary", 10, d], 20)
    app.run(20)
    app.close()
    L = oracle.lib()
    out = (C.c_double * 6)()
    for step in (0, 10, 20):
        # run(k) counts steps 1..k (src/examinimd.cpp:192), so the state after
        # step s of the 20-step run is an oracle run of s steps from set-up
        m = oracle_run(oracle, step)
        rc = L.or_dump_compare(C.byref(m.system), _dump_name(d, step).encode(), out)
        print("step", step, "dr maxdr dv maxdv df maxdf", list(out), flush=True)
        assert rc == 0 and all(v == 0.0 for v in out), step
        m.close()


def _read_correctness(path):
    lines = open(path).read().splitlines()
    assert lines[0].startswith("# timestep deltarnorm maxdelr deltavnorm maxdelv deltafnorm maxdelf")
    return np.array([[float(w) for w in l.split()] for l in lines[1:]])


def test_gpu_correctness_against_oracle_dumps(tmp_path, oracle):
    print("case: LJ 32k, GPU --correctness against oracle dumps", flush=True)
    ref = tmp_path / "ref"
    ref.mkdir()
    for step in (0, 10, 20):
        m = oracle_run(oracle, step)
        assert m.dump(_dump_name(ref, step)) == 0
        m.close()
    # FP64, one lane: every column exactly zero
    out = str(tmp_path / "correctness.fp64")
    app = _app(tmp_path, ["--correctness", 10, ref, out], 20)
    app.run(20)
    app.close()
    c = _read_correctness(out)
    print(c, flush=True)
    assert np.array_equal(c[:, 0], [0, 10, 20]) and np.all(c[:, 1:] == 0.0)
    # mixed precision at set-up (same positions): force RMS error within 1e-5
    out = str(tmp_path / "correctness.mixed")
    app = _app(tmp_path, ["--correctness", 10, ref, out], 0, precision="mixed", lanes=16)
    n = app.nlocal
    app.close()
    c = _read_correctness(out)
    print(c, flush=True)
    assert c.shape[0] == 1 and c[0, 1] == 0.0 and c[0, 3] == 0.0
    assert c[0, 5] / np.sqrt(n) <= 1e-5


def test_gpu_neighbour_dump_matches_oracle(tmp_path, oracle):
    print("case: LJ 32k, --dumpneigh at set-up and step 20, CSR and 2D", flush=True)
    files = {}
    for neigh in ("CSR", "2D"):
        d = tmp_path / f"neigh_{neigh}"
        d.mkdir()
        app = _app(tmp_path, ["--dumpneigh", 20, d], 20, neigh=neigh, lanes=16)
        app.run(20)
        app.close()
        files[neigh] = [open(_dump_name(d, s, "neigh"), "rb").read() for s in (0, 20)]
    # the two layouts hold the same list
    assert files["CSR"] == files["2D"]
    # the set-up list equals the oracle's, mapped to (id_j, image_j)
    m = oracle_run(oracle, 0)
    s = m.system
    n, g = s.N_local, s.N_ghost
    a = oracle.system_arrays(s)
    box = np.array(list(s.domain))
    owner = np.zeros(int(a["id"].max()) + 1, np.int64)
    owner[a["id"][:n]] = np.arange(n)
    shift = np.rint((a["x"][n:] - a["x"][owner[a["id"][n:]]]) / box).astype(np.int64)
    code = np.zeros(n + g, np.int64)
    code[n:] = sum(np.where(shift[:, k] < 0, 1, np.where(shift[:, k] > 0, 2, 0)) << (2 * k)
                   for k in range(3))
    row, ent = m.csr()
    rows = emd.read_neigh_dump(_dump_name(tmp_path / "neigh_CSR", 0, "neigh"))
    assert sorted(rows) == sorted(a["id"][:n].tolist())
    for i in range(n):
        e = ent[row[i]:row[i + 1]]
        want = np.stack([a["id"][e], code[e]], 1)
        want = want[np.lexsort((want[:, 1], want[:, 0]))]
        assert np.array_equal(rows[int(a["id"][i])], want), i
    m.close()
