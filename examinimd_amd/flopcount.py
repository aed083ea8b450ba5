"""Loop-bound FLOP count of this build's SNAP force (Y / adjoint form), per
atom and step, for the algorithmic side of the SNAP roofline.

Counts the floating-point adds and multiplies (an FMA counts 2; sqrt,
division and transcendentals 1 each) executed by the four kernels of
emd_snap.hip (twojmax <= 8) for an atom with n_in neighbours inside the
cutoff:

  P   k_snap_neigh : per in-cutoff neighbour the Cayley-Klein parameters,
                     their derivative factors, the switching function and
                     the r^-14 repulsion, once;
  U   k_snap_ui_rg : per neighbour, for every half-plane entry of every layer
                     j, the two recursion terms (SNA::compute_uarray,
                     sna_impl.hpp:663-695) and the sfac accumulation
                     (add_uarraytot, sna_impl.hpp:578-604);
  Y   k_snap_yflat : per atom, for every half-plane entry (J, ma, mb) and
                     every (a, b) pair with a non-zero role weight, the
                     Clebsch-Gordan double sum of compute_zi
                     (sna_impl.hpp:221-274) and the weighted accumulation;
  F   k_snap_fij_rev: per neighbour the force as the gradient of
                     sum Re(conj(U) sfac w Y): the U recursion forward, the
                     adjoint recursion backward with the two sums
                     Sa = sum rpa conj(Ubar) P, Sb = sum rpb conj(Ubar) Q, and
                     F_k = dsfac u_k sum Re(conj(U) w Y) + Re(conj(da_k) Sa)
                     - Re(conj(db_k) Sb) (carrying dU/dr forward instead,
                     compute_duarray, sna_impl.hpp:728-893, costs about
                     twice as much).

The reference's Z/dB form (compute_zi over all 125 triples, then dB/dr for
all 55 components per neighbour) does about twice this work; roofline
fractions use this count, so that the algorithmic saving is not reported as
hardware efficiency.

  python -m examinimd_amd.flopcount [--twojmax 8] [--nin 26]
"""
import argparse


def _tdiv(a, b):
    """C integer division (truncates toward zero), as in the kernels"""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def components(twojmax):
    """(j1, j2, j) of the bispectrum components (SNA::build_indexlist,
    diagonalstyle 3: j1 >= j2, j >= j1)"""
    out = []
    for j1 in range(twojmax + 1):
        for j2 in range(j1 + 1):
            for j in range(j1 - j2, min(twojmax, j1 + j2) + 1, 2):
                if j >= j1:
                    out.append((j1, j2, j))
    return out


def y_pairs(twojmax):
    """(a, b, J) keys of the Y table: the three roles of every component"""
    keys = set()
    for j1, j2, j in components(twojmax):
        keys.add((max(j1, j2), min(j1, j2), j))
        keys.add((max(j, j2), min(j, j2), j1))
        keys.add((max(j1, j), min(j1, j), j2))
    return keys


def half_plane(j):
    """(ma, mb) entries of layer j with 2 mb <= j"""
    return [(ma, mb) for mb in range(j // 2 + 1) for ma in range(j + 1)]


def p_flops_per_neighbour(twojmax):
    # r, theta0, sin/cos, z0 and dz0/dr, 1/r0 and its derivative, a, b,
    # c1..c4, the unit vector, sfac and dsfac (times wj), r^-14 repulsion
    return 70


def u_flops_per_neighbour(twojmax):
    f = 0
    for j in range(1, twojmax + 1):
        for ma, mb in half_plane(j):
            if ma < j:
                f += 10   # rp * (conj(a) U) : 4 mul, 2 add, 2 scale, 2 acc
            if ma > 0:
                f += 10   # rp * (conj(b) U)
            f += 4        # U_tot += sfac * U
    return f


def y_flops_per_atom(twojmax):
    keys = y_pairs(twojmax)
    f = 0
    for J in range(twojmax + 1):
        pairs = [(a, b) for (a, b, JJ) in keys if JJ == J]
        for ma, mb in half_plane(J):
            for a, b in pairs:
                ma1lo = max(0, _tdiv(2 * ma - J - b + a, 2))
                ma1hi = min(a, _tdiv(2 * ma - J + b + a, 2))
                mb1lo = max(0, _tdiv(2 * mb - J - b + a, 2))
                mb1hi = min(a, _tdiv(2 * mb - J + b + a, 2))
                for _ in range(ma1lo, ma1hi + 1):
                    f += 10 * max(0, mb1hi - mb1lo + 1)  # c * (U_a U_b), complex
                    f += 4                               # z += cg * s
                f += 4                                   # y += coef * z
    return f


def f_flops_per_neighbour(twojmax):
    f = 15  # da/dr and db/dr from c1..c4 and the unit vector
    for j in range(1, twojmax + 1):
        for ma, mb in half_plane(j):
            weighted = not (2 * mb == j and ma > mb)
            # forward: the recursion terms, Re(conj(U) w Y)
            f += (10 if ma < j else 0) + (10 if ma > 0 else 0) + (4 if weighted else 0)
            # backward: Sa / Sb terms, and (above layer 1) the adjoint sent to
            # layer j-1; the entry's own term sfac w Y
            f += (10 if ma < j else 0) + (10 if ma > 0 else 0)
            if j >= 2:
                f += (10 if ma < j else 0) + (10 if ma > 0 else 0)
            if weighted and j < twojmax:
                f += 4
    f += 3 * 10 + 3 * 8 + 6  # F_k from the sums, the lane reduction, repulsion
    return f


def snap_flops(twojmax=8, n_in=26):
    """per-atom-step FLOPs of the Y form, by kernel"""
    p = n_in * p_flops_per_neighbour(twojmax)
    u = n_in * u_flops_per_neighbour(twojmax)
    y = y_flops_per_atom(twojmax)
    fij = n_in * f_flops_per_neighbour(twojmax)
    return {"P": p, "U": u, "Y": y, "F": fij, "total": p + u + y + fij,
            "twojmax": twojmax, "n_in": n_in, "ncoeff": len(components(twojmax))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--twojmax", type=int, default=8)
    ap.add_argument("--nin", type=int, default=26)
    a = ap.parse_args()
    print(snap_flops(a.twojmax, a.nin))


if __name__ == "__main__":
    main()
