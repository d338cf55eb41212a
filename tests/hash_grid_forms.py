"""Closed forms of the hash-grid cache passes shared by the oracle pins and
the GPU replays."""
from __future__ import annotations

import numpy as np


def decay_closed_form(decay: np.ndarray, cells: np.ndarray):
    """What one decay pass leaves, stated per bucket without its move order.

    Sets of 4 cells are scanned in order; the scan stops before a set with
    no occupied cell and after a set whose 4th cell was empty or has just
    expired (the early out at hash_grid_decay.hlsl:73-78 tests the last
    decremented value).  In the scanned sets the cells with decay > 1
    survive with decay - 1 and end up, as a multiset, in the bucket's first
    slots; the rest of the scanned sets is cleared; later sets are left as
    they were (so such a bucket keeps a hole before them, as in the
    reference)."""
    d = decay.reshape(65536, 4, 4).astype(np.int64)
    word = d.any(-1)
    cont = d.all(-1) & (d[:, :, 3] > 1)
    reach = np.ones((65536, 4), bool)
    for st in range(1, 4):
        reach[:, st] = reach[:, st - 1] & cont[:, st - 1]
    processed = word & reach
    nproc = processed.sum(1)
    pos = np.arange(16)
    scanned = pos[None, :] < 4 * nproc[:, None]
    alive = scanned & (d.reshape(65536, 16) > 1)
    n_alive = alive.sum(1)
    head = pos[None, :] < n_alive[:, None]
    b = np.repeat(np.arange(65536), 16).reshape(65536, 16)
    rows_in = np.column_stack([b[alive], cells.reshape(65536, 16, 4)[alive], d.reshape(65536, 16)[alive] - 1])
    return scanned, head, rows_in.astype(np.uint64)
