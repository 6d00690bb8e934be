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
"""
FLOP_PER_STEP = {
    "tor": 686,        # 2 complex components
    "solid": 5102,     # 14 solid-region minors
    "fluid": 1697,     # 5 fluid-region minors
    "bvec": 16100,     # 4 columns x 14 b-vector components
    "rad": 1160,       # 2 radial components
}


def ode_flops(counts):
    """counts = ysx_step_counts: tor, solid, fluid, bvec, rad steps summed
    over pairs (a b-vector step advances all four source columns)"""
    tor, solid, fluid, bvec, rad = (int(c) for c in counts)
    return {
        "tor": FLOP_PER_STEP["tor"] * tor,
        "minors": FLOP_PER_STEP["solid"] * solid + FLOP_PER_STEP["fluid"] * fluid,
        "bvec": FLOP_PER_STEP["bvec"] * bvec,
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
