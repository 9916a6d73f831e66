"""Planted record blocks for round-trip checks (test infrastructure).

test-pir.cpp:46-63 plants, in database block k, the values i + 1 + k at rows
i + tpack (i a multiple of r) of column tcol and stores the column through
inverse_encode; the client targeting (tcol, tpack) then recovers the message
1..N.  Here one block is planted for several (column, packing) targets, with
a shift per target, and repeated for every block (block stride 0): block k
then contributes i + 1 + shift at message position i + k, so the recovered
message is m - (m mod r) + 1 + shift.
"""
import numpy as np

N = 4096


def planted_block(oracle, r, cols, packs, shifts):
    """one record block [N][N/2] (values < p) planted for each target"""
    i = np.arange(N, dtype=np.uint64)[:, None]
    j = np.arange(N // 2, dtype=np.uint64)[None, :]
    rec = (i * 31 + j * 17 + 5) % 65536
    rows = np.arange(0, N, r, dtype=np.uint64)
    for c, pk, sh in zip(cols, packs, shifts):
        col = rec[:, c].copy()
        col[rows + pk] = rows + 1 + sh
        rec[:, c] = oracle.inverse_encode(col)
    return rec


def planted_message(r, shift):
    """the message a client targeting a planted column recovers"""
    m = np.arange(N, dtype=np.uint64)
    return m - m % r + 1 + shift
