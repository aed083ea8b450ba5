"""The SNAP FLOP counter enumerates the same tables as the kernels: the
reference's component counts (55 / 30 / 14 for twojmax 8 / 6 / 4, the
ncoeff of W.snapcoeff and Ta06A.snapcoeff less beta_0), half-plane sizes
sum_j (j+1)(j/2+1), and Y (a, b, J) keys drawn from the 125 full triples."""
from examinimd_amd import flopcount


def test_component_counts_match_reference_files():
    assert len(flopcount.components(8)) == 55
    assert len(flopcount.components(6)) == 30
    assert len(flopcount.components(4)) == 14


def test_half_plane_and_y_keys():
    assert sum(len(flopcount.half_plane(j)) for j in range(9)) == 155
    keys = flopcount.y_pairs(8)
    full = {(j1, j2, j) for j1 in range(9) for j2 in range(j1 + 1)
            for j in range(j1 - j2, min(8, j1 + j2) + 1, 2)}
    assert len(full) == 125 and keys <= full


def test_counts_scale_with_neighbours():
    a = flopcount.snap_flops(8, 26)
    b = flopcount.snap_flops(8, 52)
    assert b["Y"] == a["Y"]                      # per atom, not per neighbour
    assert b["U"] == 2 * a["U"] and b["F"] == 2 * a["F"]
    assert 0.9e6 < a["total"] < 1.2e6
