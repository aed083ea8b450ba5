"""LAMMPS-style input decks for the standard workloads.

lj_deck(n)      : FCC LJ melt, density 0.8442, eps = sigma = 1, mass 1,
                  cutoff 2.5, skin 0.3, rebuild every 20 steps, dt 0.005,
                  T* = 1.44 (n = 20 -> 32 000 atoms, n = 100 -> 4 000 000)
snap_deck(n)    : BCC tungsten, a = 3.1803 A, mass 183.84, twojmax 8,
                  rcutfac 4.73442, skin 1.0, T = 300 K, dt = 0.5 fs
                  (n = 10 -> 2 000 atoms)
snap_coeff(...) : a .snapcoeff file with seeded Gaussian beta_1..beta_55
"""
import os

import numpy as np


def lj_deck(n=20, steps=100, seed=87287, temperature=1.44, every=20, thermo=0, newton=False,
            gaussian=True, density=0.8442, skin=0.3, dt=0.005, intensity=None):
    """intensity: None for pair_style lj/cut, else lj/cut/idial with that
    repeat count"""
    dist = " dist gaussian" if gaussian else ""
    pair = ("pair_style      lj/cut 2.5\npair_coeff      1 1 1.0 1.0 2.5" if intensity is None else
            f"pair_style      lj/cut/idial 2.5\npair_coeff      1 1 1.0 1.0 2.5 {intensity}")
    return f"""# LJ melt, FCC {n}^3 cells ({4 * n ** 3} atoms)
units           lj
atom_style      atomic
newton          {'on' if newton else 'off'}
lattice         fcc {density}
region          box block 0 {n} 0 {n} 0 {n}
create_box      1 box
create_atoms    1 box
mass            1 1.0
velocity        all create {temperature} {seed} loop geom{dist}
{pair}
neighbor        {skin} bin
neigh_modify    delay 0 every {every} check no
fix             1 all nve
timestep        {dt}
thermo          {thermo}
run             {steps}
"""


def snap_deck(n=10, steps=100, seed=4928459, temperature=300.0, every=10, thermo=0,
              coeff="W.snapcoeff", param="W.snapparam", lattice="bcc", a=3.1803, dt=0.0005,
              gaussian=True, ny=None, nz=None):
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    dist = " dist gaussian" if gaussian else ""
    return f"""# SNAP W, {lattice.upper()} {n}x{ny}x{nz}
units           metal
newton          on
lattice         {lattice} {a}
region          box block 0 {n} 0 {ny} 0 {nz}
create_box      1 box
create_atoms    1 box
mass            1 183.84
pair_style      snap
pair_coeff      * * {coeff} W {param} W
neighbor        1.0 bin
neigh_modify    once no every {every} delay 0 check yes
velocity        all create {temperature} {seed} loop geom{dist}
fix             1 all nve
timestep        {dt}
thermo          {thermo}
run             {steps}
"""


SNAP_PARAM = """# SNAP parameters: tungsten, twojmax 8
rcutfac 4.73442
twojmax 8
rfac0 0.99363
rmin0 0
diagonalstyle 3
bzeroflag 0
quadraticflag 0
"""


def snap_beta(seed=2024, sigma=0.05, ncoeff=55, beta0=0.0):
    rng = np.random.default_rng(seed)
    return np.concatenate([[beta0], rng.normal(0.0, sigma, ncoeff)])


def snap_coeff_text(beta, elem="W", radius=0.5, weight=1.0):
    lines = ["# SNAP coefficients (seeded Gaussian beta)", "",
             f"1 {len(beta)}", f"{elem} {radius} {weight}"]
    lines += [f" {b:.17g}" for b in beta]
    return "\n".join(lines) + "\n"


def write(path, text):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)
    return path
