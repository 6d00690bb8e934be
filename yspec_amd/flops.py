"""Algorithmic work of the spectral path, one convention for every report.

FLOP convention: complex x complex = 6, complex +- complex = 2, real x
complex = 2, complex reciprocal = 6, real op = 1; spline interval weights are
shared per radius.  Per Cash-Karp step: 6 right-hand-side evaluations plus
79 FLOP per complex component for the stage combinations and the
renormalisation test.  The per-step constants below count, for each
system, the spline evaluations, the coefficient matrix and the RHS of the
reference routines (src/module_int.f90:134-184, 261-372, 500-560, 676-715,
818-871, 941-967, 1016-1057); the b-vector constant covers its four source
columns sharing one coefficient matrix per radius.

The b-vector system marched as receiver-side rows with transposed steps
(the default path) does per step the l-independent base of the six stage
radii once (6 x 172), and per row the six transposed RHS and the stage
combinations: the 16100 of the four columns split as 6 x 172 + 4 x 3767,
plus 12 FLOP per component for the row's running sum x' = x + sum zb_i.
"""
FLOP_PER_STEP = {
    "tor": 686,        # 2 complex components
    "solid": 5102,     # 14 solid-region minors
    "fluid": 1697,     # 5 fluid-region minors
    "bvec": 16100,     # 4 columns x 14 b-vector components
    "rad": 1160,       # 2 radial components
}
BROW_SHARED = 6 * 172         # base coefficients of the six stage radii
BROW_PER_ROW = 3767 + 14 * 12  # one transposed row of 14 components


def brow_flop_per_step(rows):
    """b-vector step as `rows` receiver-side rows (2, or 3 with the
    potential/tilt correction)"""
    return BROW_SHARED + rows * BROW_PER_ROW


def ode_flops(counts, brows=None):
    """counts = ysx_step_counts: tor, solid, fluid, bvec, rad steps summed
    over pairs (a b-vector step advances all four source columns, or all
    rows).  brows: rows of the b-vector row march (None: the four columns)"""
    tor, solid, fluid, bvec, rad = (int(c) for c in counts)
    per_b = FLOP_PER_STEP["bvec"] if brows is None else brow_flop_per_step(brows)
    return {
        "tor": FLOP_PER_STEP["tor"] * tor,
        "minors": FLOP_PER_STEP["solid"] * solid + FLOP_PER_STEP["fluid"] * fluid,
        "bvec": per_b * bvec,
        "rad": FLOP_PER_STEP["rad"] * rad,
    }


def sum_flops(nw, lmax, nr):
    """Legendre receiver sum: 88 FLOP per (w, l, receiver) (22 complex x real
    products of the m = -2..2 terms, 4 of them vanish at m = 0)"""
    return 88 * nw * (lmax + 1) * nr


def fft_bytes(nw, nt, nout, nr):
    """HBM bytes of the filtered IFFT + post: read the spectra, write the
    seismograms (3 components per receiver)"""
    return 3 * nr * (16 * nw + 8 * nt)
