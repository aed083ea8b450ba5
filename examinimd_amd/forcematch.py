"""Force-matching loss and gradient over sets of configurations through the
hand-written adjoint kernels (emd_snap_adjoint_beta, emd_lj_adjoint).

Per configuration c with N_c atoms:
    L_c = sum_i |F_i(theta) - F*_i|^2 + 100 ((E(theta) - E*) / N_c)^2
with targets F*, E* from the plain FP64 path at a 'true' parameter set.
The gradient is sum_c [sum_i g_i . dF_i/dtheta + g_E dE/dtheta] with
g = 2 (F - F*) and g_E = 200 (E - E*) / N_c^2, evaluated by the adjoint
kernels (FP32 arithmetic with FP64 accumulation in mixed precision).

The workloads are the seeded sets of the force-matching benchmark:
  snap_set(): 64 BCC tungsten configurations, a = 3.1803 A, 2 n^3 atoms with
              n in {4..8} (128..1024 atoms), Gaussian displacements 0.05 A;
  lj_set():   32 FCC LJ boxes at density 0.8442, 4 n^3 atoms with n in
              {5..10} (500..4000 atoms), displacements 0.05 sigma.
SnapForceMatch and LJForceMatch keep one device state per configuration
(an evaluation is a launch sequence and a host round trip per
configuration); SnapForceMatchMerged merges the configurations into one
system so an evaluation is one launch sequence for the whole set.
"""
import ctypes as C

import numpy as np

import examinimd_amd as emd
from examinimd_amd import decks
from examinimd_amd.abi import EMD_FP64, EMD_MIXED, EMD_NEIGH_CSR, EmdHalo, EmdNeigh, EmdSystem
from examinimd_amd.pipeline import Pipeline, lattice

A_W = 3.1803
W_MASS = 183.84
SNAP_CUT = 4.73442           # rcutfac * 2 * radelem (0.5)
LJ_A = (4 / 0.8442) ** (1 / 3)


def snap_set(count=64, seed=17, sizes=(4, 5, 6, 7, 8), disp=0.05):
    """[(positions, box)] of uneven BCC W configurations"""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        n = int(sizes[c % len(sizes)])
        x = lattice("bcc", n, A_W) + rng.normal(0, disp, (2 * n ** 3, 3))
        out.append((x, (n * A_W,) * 3))
    return out


def lj_set(count=32, seed=23, sizes=(5, 6, 7, 8, 9, 10), disp=0.05):
    """[(positions, box)] of uneven FCC LJ boxes"""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        n = int(sizes[c % len(sizes)])
        x = lattice("fcc", n, LJ_A) + rng.normal(0, disp, (4 * n ** 3, 3))
        out.append((x, (n * LJ_A,) * 3))
    return out


class SnapForceMatch:
    """loss(beta) and dloss/dbeta over SNAP configurations (one SNAP object,
    one coefficient upload per evaluation, shared by all configurations)"""

    def __init__(self, configs, beta_true, ctx=None):
        self.pls = []
        self.targets = []
        for x, box in configs:
            pl = Pipeline(ctx)
            ctx = pl.ctx
            pl.load(x, box, mass=W_MASS)
            # static configurations: the list needs no skin; newton on for
            # the ghost force fold
            pl.build(SNAP_CUT, newton=1)
            if self.pls:
                pl.snap_use(self.pls[0])
            else:
                pl.snap_setup(list(beta_true))
            F = pl.snap_force(EMD_FP64)
            E = pl.snap_energy()
            self.pls.append(pl)
            self.targets.append((F, E, len(x)))
        self.ctx = ctx

    def evaluate(self, beta, precision, grad=True):
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        loss = 0.0
        g = np.zeros(len(beta))
        self.pls[0].snap_set_coeff(beta)
        for pl, (Ft, Et, n) in zip(self.pls, self.targets):
            pl._beta = self.pls[0]._beta
            F = pl.snap_force(precision)
            E = pl.snap_energy()
            dF = F - Ft
            loss += float((dF * dF).sum()) + 100.0 * ((E - Et) / n) ** 2
            if grad:
                g += pl.snap_adjoint(2.0 * dF, 200.0 * (E - Et) / n ** 2, precision)
        return loss, g

    def close(self):
        # the first pipeline owns the shared SNAP object: close it last
        for pl in reversed(self.pls):
            pl.close()
        self.pls = []


class MergedSystem:
    """Configurations merged into one system on the device: the locals of
    every configuration, then their ghosts phase by phase (phase p of every
    configuration before phase p + 1), one CSR list and one halo whose pack
    lists name merged indices.  Each configuration is first built on its
    own (exchange, binning sort, halo, neighbour list, in `pls`); the merge
    only renumbers, so every kernel sees the same neighbourhoods.  Atom ids
    are offset per configuration to stay unique.  Nothing is rebuilt
    afterwards: the configurations are static."""

    def __init__(self, pls, mass):
        self.L = L = emd.hip_lib()
        self.ctx = pls[0].ctx
        c = self.ctx.h
        n_c = [pl.s.n_local for pl in pls]
        g_c = [pl.s.n_ghost for pl in pls]
        ph_c = [list(pl.h.num_ghost) for pl in pls]
        N, G = sum(n_c), sum(g_c)
        Gp = [sum(ph[p] for ph in ph_c) for p in range(6)]
        off_p = np.concatenate([[N], N + np.cumsum(Gp)[:-1]]).astype(np.int64)
        self.n_atoms, self.starts = np.array(n_c), np.concatenate([[0], np.cumsum(n_c)[:-1]])
        x = np.zeros((N + G, 3))
        ids = np.zeros(N + G, np.int32)
        packs = [[] for _ in range(6)]
        root = np.zeros(G, np.int32)
        image = np.zeros(G, np.int32)
        rows, ents = [np.zeros(1, np.int64)], []
        seen = [0] * 6                      # ghosts of phase p placed so far
        for k, pl in enumerate(pls):
            n, g = n_c[k], g_c[k]
            xs = self._dl(pl.s.x, 3 * (n + g)).reshape(-1, 3)
            idk = emd.download_ints(c, pl.s.id, n + g)
            mp = np.zeros(n + g, np.int64)
            mp[:n] = self.starts[k] + np.arange(n)
            gl = n
            for p in range(6):
                m = ph_c[k][p]
                mp[gl:gl + m] = off_p[p] + seen[p] + np.arange(m)
                gl += m
            for p in range(6):
                m = ph_c[k][p]
                if m:
                    packs[p].append(mp[emd.download_ints(c, pl.h.pack[p], m)])
                seen[p] += m
            x[mp] = xs
            ids[mp] = idk + self.starts[k]
            if g:
                gi = mp[n:] - N
                root[gi] = mp[emd.download_ints(c, pl.h.root, g)]
                image[gi] = emd.download_ints(c, pl.h.image, g)
            row, num, ent = emd.download_neigh(c, pl.nl)
            rows.append(rows[-1][-1] + row[1:].astype(np.int64))
            ents.append(mp[ent])
        row = np.concatenate(rows).astype(np.int32)
        ent = np.concatenate(ents).astype(np.int32)
        self._bufs = []
        s = self.s = EmdSystem()
        zeros = np.zeros_like(x)
        self.chk(L.emd_system_upload(c, C.byref(s), N + G, _dp(x), _dp(zeros),
                                     _ip(np.zeros(N + G, np.int32)), _ip(ids),
                                     _dp(np.zeros(N + G))), "upload")
        m = np.array([mass])
        self.chk(L.emd_system_set_mass(c, C.byref(s), 1, _dp(m)), "mass")
        s.n_local, s.n_ghost, s.n_global = N, G, N
        h = self.h = EmdHalo()
        for p in range(6):
            h.num_ghost[p] = Gp[p]
            h.cap[p] = Gp[p]
            h.pack[p] = self._put(np.concatenate(packs[p]).astype(np.int32) if packs[p]
                                  else np.zeros(1, np.int32))
        h.root, h.image = self._put(root), self._put(image)
        h.root_cap = h.n_ghost = G
        nl = self.nl = EmdNeigh()
        nl.layout, nl.n, nl.total = EMD_NEIGH_CSR, N, int(row[-1])
        nl.maxneighs = int(np.diff(row).max()) if N else 0
        nl.row, nl.num = self._put(row), self._put(np.diff(row).astype(np.int32))
        nl.entries = self._put(ent)
        nl.cap_entries, nl.cap_rows = int(row[-1]), N + 1

    def chk(self, st, what):
        self.ctx.check(st, what)

    def _dl(self, ptr, n):
        out = np.zeros(n)
        if n:
            self.chk(self.L.emd_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p),
                                           C.cast(ptr, C.c_void_p), out.nbytes), "d2h")
        return out

    def _put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.chk(self.L.emd_malloc(self.ctx.h, C.byref(p), max(a.nbytes, 4)), "malloc")
        self.chk(self.L.emd_memcpy_h2d(self.ctx.h, p, a.ctypes.data_as(C.c_void_p), a.nbytes),
                 "h2d")
        self._bufs.append(p)
        return C.cast(p, C.POINTER(C.c_int if a.dtype == np.int32 else C.c_double))

    def forces(self, precision):
        """forces of the merged locals (library order)"""
        n = self.s.n_local
        if precision == EMD_MIXED:
            f32 = np.zeros((n, 3), np.float32)
            self.chk(self.L.emd_memcpy_d2h(self.ctx.h, f32.ctypes.data_as(C.c_void_p),
                                           C.cast(self.s.f32, C.c_void_p), f32.nbytes), "f32")
            return f32.astype(np.float64)
        return self._dl(self.s.f, 3 * n).reshape(n, 3)

    def close(self):
        for p in self._bufs:
            self.L.emd_free(self.ctx.h, p)
        self._bufs = []
        self.L.emd_system_free(self.ctx.h, C.byref(self.s))


def _dp(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(C.POINTER(C.c_int))


class SnapForceMatchMerged:
    """loss(beta) and dloss/dbeta over SNAP configurations merged into one
    system: per evaluation one coefficient update, one force launch
    sequence (zeroing fused), one ghost fold and one adjoint for the whole
    set.  The configurations are static, so their energies are affine in
    beta through per-atom descriptors computed once at set-up
    (emd_snap_descriptors): E_c = N_c beta_0 + S_c . beta_1.. + R_c with S_c
    the sum of the configuration's descriptor rows and R_c its energy at
    beta = 0 (the pair repulsion); the energy part of the gradient is
    sum_c g_E,c (N_c, S_c), and the adjoint supplies the force part.  Same
    loss and gradient as SnapForceMatch."""

    def __init__(self, configs, beta_true, ctx=None):
        pls = []
        for x, box in configs:
            pl = Pipeline(ctx)
            ctx = pl.ctx
            pl.load(x, box, mass=W_MASS)
            pl.build(SNAP_CUT, newton=1)
            pls.append(pl)
        self.m = MergedSystem(pls, W_MASS)
        for pl in pls:
            pl.close()
        self.ctx = ctx
        self.L = emd.hip_lib()
        m, L, c = self.m, self.L, ctx.h
        self.pl = Pipeline(ctx)          # owner of the SNAP object
        self.pl.snap_setup(list(beta_true))
        n = m.s.n_local
        nc = len(beta_true) - 1
        # descriptors and the repulsion energy per configuration
        bdev = m._put(np.zeros(n * nc))
        m.chk(L.emd_snap_descriptors(c, C.byref(m.s), C.byref(m.nl), self.pl.snap, bdev), "desc")
        B = m._dl(bdev, n * nc).reshape(n, nc)
        self.S = np.add.reduceat(B, m.starts, axis=0)
        eatom = m._put(np.zeros(n))
        self.pl.snap_set_coeff(np.zeros(nc + 1))
        e = C.c_double()
        m.chk(L.emd_snap_energy(c, C.byref(m.s), C.byref(m.nl), self.pl.snap, C.byref(e), eatom),
              "energy")
        self.R = np.add.reduceat(m._dl(eatom, n), m.starts)
        self._g = m._put(np.zeros(3 * n))
        self.pl.snap_set_coeff(beta_true)
        # the configurations never move: the compaction, neighbour parameters
        # and U_tot depend on them only, so force and adjoint calls reuse them
        m.chk(L.emd_snap_static_geometry(c, self.pl.snap, 1), "static geometry")
        self.Ft = self._forces(EMD_FP64)
        self.Et = self.energies(beta_true)

    def energies(self, beta):
        """per-configuration energies, E_c = N_c beta_0 + S_c . beta_1.. + R_c"""
        beta = np.asarray(beta, dtype=np.float64)
        return self.m.n_atoms * beta[0] + self.S @ beta[1:] + self.R

    def _forces(self, precision):
        L, c, m = self.L, self.ctx.h, self.m
        s = C.byref(m.s)
        m.chk(L.emd_snap_force_zeroed(c, s, C.byref(m.nl), self.pl.snap, precision), "snap")
        m.chk(L.emd_comm_update_force(c, s, C.byref(m.h), precision), "fold")
        return m.forces(precision)

    def evaluate(self, beta, precision, grad=True):
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        self.pl.snap_set_coeff(beta)
        dF = self._forces(precision) - self.Ft
        n = self.m.n_atoms
        dE = self.energies(beta) - self.Et
        loss = float((dF * dF).sum()) + 100.0 * float(np.sum((dE / n) ** 2))
        g = np.zeros(len(beta))
        if grad:
            L, c, m = self.L, self.ctx.h, self.m
            gF = np.ascontiguousarray(2.0 * dF)
            m.chk(L.emd_memcpy_h2d(c, self._g, gF.ctypes.data_as(C.c_void_p), gF.nbytes), "h2d")
            m.chk(L.emd_snap_adjoint_beta(c, C.byref(m.s), C.byref(m.nl), C.byref(m.h),
                                          self.pl.snap, self._g, 0.0, precision,
                                          g.ctypes.data_as(C.POINTER(C.c_double))), "adjoint")
            gE = 200.0 * dE / n ** 2
            g[0] += float(np.sum(gE * n))
            g[1:] += gE @ self.S
        return loss, g

    def close(self):
        self.m.close()
        self.pl.close()


class LJForceMatchMerged:
    """loss(eps, sigma) and its gradient over LJ boxes merged into one system:
    per evaluation one force launch (stored forces), one per-atom energy
    launch and one adjoint launch for the whole set, the energy cotangent of
    each box entering the adjoint as a weight on its atoms
    (emd_lj_adjoint_atoms).  Same loss and gradient as LJForceMatch."""

    def __init__(self, configs, true=(1.0, 1.0), ctx=None, rc=2.5):
        self.rc = rc
        pls = []
        for x, box in configs:
            pl = Pipeline(ctx)
            ctx = pl.ctx
            pl.load(x, box)
            pl.build(rc)
            pls.append(pl)
        self.m = MergedSystem(pls, 1.0)
        for pl in pls:
            pl.close()
        self.ctx = ctx
        self.L = emd.hip_lib()
        n = self.m.s.n_local
        self._eatom = self.m._put(np.zeros(n))
        self._g = self.m._put(np.zeros(3 * n))
        self._w = self.m._put(np.zeros(n))
        self.Ft, self.Et = self._force_energy(EMD_FP64, *true)

    def _params(self, eps, sig):
        from examinimd_amd.abi import EmdLJParams
        p = EmdLJParams()
        p.ntypes = 1
        p.lj1[0][0] = 48.0 * eps * sig ** 12
        p.lj2[0][0] = 24.0 * eps * sig ** 6
        p.cutsq[0][0] = self.rc * self.rc
        return p

    def _force_energy(self, precision, eps, sig):
        L, c, m = self.L, self.ctx.h, self.m
        p = self._params(eps, sig)
        m.chk(L.emd_lj_force_store(c, C.byref(m.s), C.byref(m.nl), C.byref(p), precision), "lj")
        F = m.forces(precision)
        m.chk(L.emd_lj_energy_atoms(c, C.byref(m.s), C.byref(m.nl), C.byref(p), 0, self._eatom),
              "energy")
        return F, np.add.reduceat(m._dl(self._eatom, m.s.n_local), m.starts)

    def evaluate(self, theta, precision, grad=True):
        eps, sig = (float(t) for t in theta)
        F, E = self._force_energy(precision, eps, sig)
        dF = F - self.Ft
        n = self.m.n_atoms
        dE = E - self.Et
        loss = float((dF * dF).sum()) + 100.0 * float(np.sum((dE / n) ** 2))
        g = np.zeros(2)
        if grad:
            L, c, m = self.L, self.ctx.h, self.m
            w = np.repeat(200.0 * dE / n ** 2, n)
            gF = np.ascontiguousarray(2.0 * dF)
            m.chk(L.emd_memcpy_h2d(c, self._g, gF.ctypes.data_as(C.c_void_p), gF.nbytes), "h2d")
            m.chk(L.emd_memcpy_h2d(c, self._w, w.ctypes.data_as(C.c_void_p), w.nbytes), "h2d")
            m.chk(L.emd_lj_adjoint_atoms(c, C.byref(m.s), C.byref(m.nl), eps, sig, self.rc,
                                         self._g, self._w, precision,
                                         g.ctypes.data_as(C.POINTER(C.c_double))), "adjoint")
        return loss, g

    def close(self):
        self.m.close()


class LJForceMatch:
    """loss(eps, sigma) and its gradient over LJ boxes (cutoff 2.5)"""

    def __init__(self, configs, true=(1.0, 1.0), ctx=None, rc=2.5):
        self.rc = rc
        self.pls = []
        self.targets = []
        for x, box in configs:
            pl = Pipeline(ctx)
            ctx = pl.ctx
            pl.load(x, box)
            pl.build(rc)
            F = pl.lj(EMD_FP64, eps=true[0], sigma=true[1], rc=rc)
            E = pl.lj_energy(true[0], true[1], rc)
            self.pls.append(pl)
            self.targets.append((F, E, len(x)))
        self.ctx = ctx

    def evaluate(self, theta, precision, grad=True):
        eps, sig = (float(t) for t in theta)
        loss = 0.0
        g = np.zeros(2)
        for pl, (Ft, Et, n) in zip(self.pls, self.targets):
            F = pl.lj(precision, eps=eps, sigma=sig, rc=self.rc)
            E = pl.lj_energy(eps, sig, self.rc)
            dF = F - Ft
            loss += float((dF * dF).sum()) + 100.0 * ((E - Et) / n) ** 2
            if grad:
                g += pl.lj_adjoint(2.0 * dF, 200.0 * (E - Et) / n ** 2, eps, sig, self.rc,
                                   precision)
        return loss, g

    def close(self):
        for pl in self.pls:
            pl.close()
        self.pls = []


def snap_beta_true(seed=2024):
    return decks.snap_beta(seed=seed, sigma=0.05)


def snap_beta_perturbed(beta_true, seed=99):
    rng = np.random.default_rng(seed)
    return beta_true + np.concatenate([[0.0], rng.normal(0, 0.01, len(beta_true) - 1)])
