"""Algorithmic HBM bytes of the hot-path operators (int32 indices).

One definition per operator, used by bench.py and the profile summaries.
`s_*` are value sizes in bytes: 8 for FP64, 4 for FP32.  The counts are what
the operator must move at minimum (each array read or written once), not what
a particular kernel happens to move; measured traffic above them is re-reading.
"""

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def spmv(n, nnz, s_a=8, s_x=8, s_y=8):
    """y = A x, CSR: row pointers, column indices and values, x and y once."""
    return 4 * (n + 1) + nnz * (4 + s_a) + n * s_x + n * s_y


def spmv_band(n, s_a=8, s_x=8, s_y=8):
    """y = A x on a structured grid, A by band position: five values and a
    one-byte entry mask per row, x and y once."""
    return n * (5 * s_a + 1) + n * s_x + n * s_y


def dilu_apply(n, nnz, s_a=8, s_d=8, s_b=8, s_y=8, s_x=8):
    """ldiv! with DILU, forward then backward sweep: row pointers (both
    sweeps), the strictly lower / upper entries with their columns, Dinv in
    both sweeps, the right-hand side once, the intermediate written and read
    back, the result written."""
    return 8 * (n + 1) + (nnz - n) * (4 + s_a) + 2 * n * s_d + n * s_b + 2 * n * s_y + n * s_x


def dilu_update(n, nnz):
    """update_dilu_diagonal!: pattern + lower/transposed-upper values, D and
    Dinv written (FP64)."""
    return 4 * (n + 1) + (nnz - n) * 12 + 16 * n


def assembly(nfaces, nnz, n, k_f=5, k_c=4):
    """One transport equation: per face the owner/neighbour pair, the two
    matrix slots and k_f face arrays; the matrix values and b written; k_c
    cell arrays read."""
    return nfaces * (8 + 16 + 8 * k_f) + nnz * 8 + n * (8 + 8 * k_c)


def green_gauss(n, nfaces):
    """Scalar Gauss gradient: phi, per face the index pair, weight and area
    vector, the volume, the 2-component gradient written."""
    return n * 8 + nfaces * (8 + 8 + 16) + n * (8 + 16)


def dot(n, s=8):
    return 2 * n * s


def axpy(n, s=8):
    return 3 * n * s


def gbps(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9
