"""Loop-bound FLOP count of this build's SNAP force (Y / adjoint form), per
atom and step, for the algorithmic side of the SNAP roofline.

Counts the floating-point adds and multiplies (an FMA counts 2; sqrt,
division and transcendentals 1 each) executed by the three kernels of
emd_snap.hip for an atom with n_in neighbours inside the cutoff:

  U   k_snap_ui : per neighbour the Cayley-Klein set-up and, for every
                  half-plane entry of every layer j, the two recursion terms
                  (SNA::compute_uarray, sna_impl.hpp:663-695) and the sfac
                  accumulation (add_uarraytot, sna_impl.hpp:578-604);
  Y   k_snap_y  : per atom, for every half-plane entry (J, ma, mb) and every
                  (a, b) pair with a non-zero role weight, the Clebsch-Gordan
                  double sum of compute_zi (sna_impl.hpp:221-274) and the
                  weighted accumulation into Y;
  F   k_snap_fij: per neighbour the U and dU/dr recursion (compute_duarray,
                  sna_impl.hpp:728-893) and the contraction with Y.

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


def u_flops_per_neighbour(twojmax):
    setup = 30  # r, theta0, z0, Cayley-Klein a and b, switching function
    f = setup
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
    setup = 60  # r, theta0 and its derivative, a, b and their gradients, sfac, dsfac
    f = setup
    for j in range(1, twojmax + 1):
        for ma, mb in half_plane(j):
            if ma < j:
                f += 10 + 54   # U term and three dU/dr components
            if ma > 0:
                f += 10 + 54
            if not (2 * mb == j and ma > mb):
                f += 29        # weighted contraction with Y, three components
    f += 18 + 10  # wavefront reduction, r^-14 repulsion
    return f


def snap_flops(twojmax=8, n_in=26):
    """per-atom-step FLOPs of the Y form, by kernel"""
    u = n_in * u_flops_per_neighbour(twojmax)
    y = y_flops_per_atom(twojmax)
    fij = n_in * f_flops_per_neighbour(twojmax)
    return {"U": u, "Y": y, "F": fij, "total": u + y + fij,
            "twojmax": twojmax, "n_in": n_in, "ncoeff": len(components(twojmax))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--twojmax", type=int, default=8)
    ap.add_argument("--nin", type=int, default=26)
    a = ap.parse_args()
    print(snap_flops(a.twojmax, a.nin))


if __name__ == "__main__":
    main()
