"""Force-matching gradients through the hand-written adjoint kernels,
wrapped in torch.autograd.Function.

Loss = sum_i |F_i(theta) - F*_i|^2 + 100 ((E(theta) - E*) / N)^2 with targets
from a 'true' parameter set.  The adjoint gradient (FP64, and FP32
arithmetic with FP64 accumulation) is checked against

  * central finite differences of the FP64 loss (h = 1e-6),
  * the exact gradient from the CPU oracle: SNAP forces and energies are
    affine in beta, so dF/dbeta_k = F(e_k) - F(0) and dE/dbeta_k likewise,
    which the oracle evaluates in the reference's Z/dB form;
    for LJ (eps, sigma) the oracle's central differences.
"""
import numpy as np
import pytest
import torch

import examinimd_amd as emd
from examinimd_amd import decks
from examinimd_amd.pipeline import Pipeline, lattice

pytestmark = pytest.mark.gpu

A_W = 3.1803


class SnapForceEnergy(torch.autograd.Function):
    """(F, E) of one configuration as a function of beta"""

    @staticmethod
    def forward(ctx, beta, pl, precision):
        pl.snap_set_coeff(beta.detach().numpy())
        F = pl.snap_force(precision)
        E = pl.snap_energy()
        ctx.pl, ctx.precision = pl, precision
        return torch.from_numpy(F), torch.tensor(E, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gF, gE):
        g = ctx.pl.snap_adjoint(gF.detach().numpy(), float(gE), ctx.precision)
        return torch.from_numpy(g), None, None


class LJForceEnergy(torch.autograd.Function):
    """(F, E) as a function of theta = (eps, sigma), cutoff 2.5"""

    @staticmethod
    def forward(ctx, theta, pl, precision):
        eps, sig = (float(v) for v in theta.detach().numpy())
        F = pl.lj(precision, eps=eps, sigma=sig)
        E = pl.lj_energy(eps=eps, sigma=sig)
        ctx.pl, ctx.precision, ctx.theta = pl, precision, (eps, sig)
        return torch.from_numpy(F), torch.tensor(E, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gF, gE):
        eps, sig = ctx.theta
        g = ctx.pl.lj_adjoint(gF.detach().numpy(), float(gE), eps, sig, 2.5, ctx.precision)
        return torch.from_numpy(g), None, None


def loss_fn(F, E, Ft, Et, n):
    return ((F - Ft) ** 2).sum() + 100.0 * ((E - Et) / n) ** 2


def snap_config(ncell, seed):
    x = lattice("bcc", ncell, A_W) + np.random.default_rng(seed).normal(0, 0.05,
                                                                          (2 * ncell ** 3, 3))
    pl = Pipeline()
    pl.load(x, (ncell * A_W,) * 3, mass=183.84)
    pl.build(4.73442 + 1.0, newton=1)
    return x, pl


@pytest.mark.parametrize("ncell,seed", [(4, 1), (5, 2)])
def test_snap_beta_gradient(oracle, ncell, seed):
    from test_oracle import _energy_force
    print(f"case: SNAP dL/dbeta, BCC {ncell}^3 seed {seed}", flush=True)
    rng = np.random.default_rng(100 + seed)
    beta_true = decks.snap_beta(seed=2024, sigma=0.05)
    beta = beta_true + np.concatenate([[0.0], rng.normal(0, 0.01, 55)])
    x, pl = snap_config(ncell, seed)
    n = len(x)
    pl.snap_setup(list(beta_true))
    Ft = torch.from_numpy(pl.snap_force(emd.EMD_FP64))
    Et = torch.tensor(pl.snap_energy(), dtype=torch.float64)

    grads = {}
    for prec in (emd.EMD_FP64, emd.EMD_MIXED):
        b = torch.tensor(beta, dtype=torch.float64, requires_grad=True)
        F, E = SnapForceEnergy.apply(b, pl, prec)
        L = loss_fn(F, E, Ft, Et, n)
        L.backward()
        grads[prec] = b.grad.numpy().copy()

    # exact gradient from the oracle (affine in beta)
    b = torch.tensor(beta, dtype=torch.float64)
    F, E = SnapForceEnergy.apply(b, pl, emd.EMD_FP64)
    gF = (2 * (F - Ft)).numpy()
    gE = float(200.0 * (E - Et) / n / n)
    box = ncell * A_W
    xw = x % box
    zero = [0.0] * 56
    E0, F0 = _energy_force(oracle, xw, box, "snap", snap=oracle.snap_params(zero), cut=5.73442)
    exact = np.zeros(56)
    exact[0] = gE * n
    for k in range(1, 56):
        ek = [0.0] * 56
        ek[k] = 1.0
        Ek, Fk = _energy_force(oracle, xw, box, "snap", snap=oracle.snap_params(ek), cut=5.73442)
        exact[k] = np.sum(gF * (Fk - F0)) + gE * (Ek - E0)
    scale = np.abs(exact).max()
    e64 = np.abs(grads[emd.EMD_FP64] - exact).max() / scale
    e32 = np.abs(grads[emd.EMD_MIXED] - exact).max() / scale
    print("gradient rel err FP64", e64, "mixed", e32)
    assert e64 < 1e-9
    assert e32 < 1e-4

    # central finite differences of the FP64 loss on a few components
    h = 1e-6
    for k in (0, 1, 17, 40, 55):
        bp, bm = beta.copy(), beta.copy()
        bp[k] += h
        bm[k] -= h
        Lp = loss_fn(*SnapForceEnergy.apply(torch.tensor(bp), pl, emd.EMD_FP64), Ft, Et, n)
        Lm = loss_fn(*SnapForceEnergy.apply(torch.tensor(bm), pl, emd.EMD_FP64), Ft, Et, n)
        fd = float(Lp - Lm) / (2 * h)
        assert abs(fd - grads[emd.EMD_FP64][k]) <= 1e-5 * scale + 1e-6 * abs(fd), (k, fd)
    pl.close()


def test_lj_eps_sigma_gradient(oracle):
    from test_oracle import _energy_force
    print("case: LJ dL/d(eps, sigma), FCC 500 atoms", flush=True)
    a = (4 / 0.8442) ** (1 / 3)
    x = lattice("fcc", 5, a) + np.random.default_rng(7).normal(0, 0.05, (500, 3))
    box = 5 * a
    n = len(x)
    pl = Pipeline()
    pl.load(x, (box,) * 3)
    pl.build(2.8)
    true = (1.0, 1.0)
    Ft = torch.from_numpy(pl.lj(emd.EMD_FP64, eps=true[0], sigma=true[1]))
    Et = torch.tensor(pl.lj_energy(*true), dtype=torch.float64)
    theta0 = np.array([1.07, 0.97])
    grads = {}
    for prec in (emd.EMD_FP64, emd.EMD_MIXED):
        th = torch.tensor(theta0, requires_grad=True)
        F, E = LJForceEnergy.apply(th, pl, prec)
        loss_fn(F, E, Ft, Et, n).backward()
        grads[prec] = th.grad.numpy().copy()
    # central differences of the FP64 loss: GPU forward and oracle forward,
    # force and energy parts separately
    h = 1e-6
    xw = x % box
    fd = {"gpu": np.zeros((2, 2)), "oracle": np.zeros((2, 2))}
    for k in range(2):
        parts = {"gpu": [], "oracle": []}
        for sgn in (1, -1):
            th = theta0.copy()
            th[k] += sgn * h
            Fg = torch.from_numpy(pl.lj(emd.EMD_FP64, eps=th[0], sigma=th[1]))
            Eg = torch.tensor(pl.lj_energy(th[0], th[1]), dtype=torch.float64)
            Eo, Fo = _energy_force(oracle, xw, box, "lj", eps=th[0], sig=th[1], rc=2.5)
            Eo = torch.tensor(Eo, dtype=torch.float64)
            for tag, F, E in (("gpu", Fg, Eg), ("oracle", torch.from_numpy(Fo), Eo)):
                parts[tag].append((float(((F - Ft) ** 2).sum()),
                                   float(100.0 * ((E - Et) / n) ** 2)))
        for tag in parts:
            (fp, ep), (fm, em) = parts[tag]
            fd[tag][k] = [(fp - fm) / (2 * h), (ep - em) / (2 * h)]
    g_fd = fd["gpu"].sum(1)
    o_fd = fd["oracle"].sum(1)
    scale = np.abs(g_fd).max()
    print("LJ grads fp64", grads[emd.EMD_FP64], "mixed", grads[emd.EMD_MIXED])
    print("FD gpu (force, energy parts)", fd["gpu"].tolist(), "FD oracle", fd["oracle"].tolist())
    assert np.abs(o_fd - g_fd).max() <= 1e-8 * scale
    assert np.abs(grads[emd.EMD_FP64] - o_fd).max() <= 1e-8 * scale
    assert np.abs(grads[emd.EMD_MIXED] - o_fd).max() <= 1e-5 * scale
    pl.close()


@pytest.mark.parametrize("precision", [emd.EMD_FP64, emd.EMD_MIXED])
def test_snap_static_geometry_reuse_matches_recomputation(precision):
    """emd_snap_static_geometry: forces and beta adjoints of a fixed
    configuration that reuse the kept compaction, neighbour parameters and
    U_tot equal those recomputed from scratch (1e-12 relative in FP64, the
    atomic force scatter aside), for a second beta too, and an energy call
    in between (which rewrites U_tot in FP64) does not leave stale values"""
    from helpers import rel_err
    print(f"case: SNAP static geometry reuse, precision {precision}", flush=True)
    a = 3.1803
    x = lattice("bcc", 6, a) + np.random.default_rng(12).normal(0, 0.08, (432, 3))
    beta1 = np.array(decks.snap_beta(seed=3, sigma=0.05))
    beta2 = np.array(decks.snap_beta(seed=4, sigma=0.05))
    gF = np.random.default_rng(5).normal(size=(432, 3))
    tol = 1e-12 if precision == emd.EMD_FP64 else 1e-5
    pl = Pipeline()
    try:
        pl.load(x, (6 * a,) * 3, mass=183.84)
        pl.build(4.73442 + 1.0, newton=1)
        pl.snap_setup(list(beta1))
        L, c = pl.L, pl.ctx.h
        want = {}
        for b in (1, 2):
            pl.snap_set_coeff(beta1 if b == 1 else beta2)
            want[b] = (pl.snap_force(precision), pl.snap_adjoint(gF, 0.0, precision))
        pl.chk(L.emd_snap_static_geometry(c, pl.snap, 1), "static on")
        got = {}
        for b in (1, 2, 1):
            pl.snap_set_coeff(beta1 if b == 1 else beta2)
            got[b] = (pl.snap_force(precision), pl.snap_adjoint(gF, 0.0, precision))
        pl.snap_energy()
        pl.snap_set_coeff(beta2)
        after_energy = pl.snap_force(precision)
        pl.chk(L.emd_snap_static_geometry(c, pl.snap, 0), "static off")
    finally:
        pl.close()
    for b in (1, 2):
        ef = rel_err(got[b][0], want[b][0])
        ea = np.abs(got[b][1] - want[b][1]).max() / np.abs(want[b][1]).max()
        print("beta", b, "force rel err", ef, "adjoint rel err", ea)
        assert ef <= tol and ea <= tol
    assert rel_err(after_energy, want[2][0]) <= tol
