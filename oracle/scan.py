"""Decoupled look-back oracle (test infrastructure only; see oracle/__init__.py).

Restates the tile protocol of ``monoid_r3_kernel`` (cuda/monoid_scan.cu:128-222)
over DFA transition functions instead of monoid indices: every tile (chunk)
reduces its bytes to an aggregate function; a tile's inclusive prefix is its
aggregate applied after the composition of its predecessors' aggregates back
to the nearest tile whose prefix is known; the first tile's prefix is the
state reached from the start state. Functions are tuples f with f[s] = image
of state s. Tiles are visited in an order given by the caller (the look-back
finds prefixes only for tiles already visited, aggregates for the others).
"""

from __future__ import annotations

import numpy as np


def chunk_function(delta: np.ndarray, data, lo: int, hi: int) -> np.ndarray:
    """Aggregate of bytes [lo, hi): f[s] = state reached from s."""
    f = np.arange(delta.shape[0], dtype=np.int64)
    for b in data[lo:hi]:
        f = delta[f, int(b)].astype(np.int64)
    return f


def lookback_final_state(delta: np.ndarray, start: int, data, lo: int, hi: int, tile: int,
                         order: list[int] | None = None) -> int:
    """Final state of bytes [lo, hi) computed tile by tile with the look-back.

    ``order``: the order in which tiles publish (default: reversed, the
    adversarial case in which every tile finds only aggregates behind it
    until the first tile is reached)."""
    n_tiles = max(1, -(-(hi - lo) // tile))
    bounds = [(lo + k * tile, min(hi, lo + (k + 1) * tile)) for k in range(n_tiles)]
    aggs = [chunk_function(delta, data, a, b) for a, b in bounds]
    # phase 1 never waits: every aggregate is published, and the first tile,
    # which starts from the known start state, publishes its prefix at once
    prefix: dict[int, int] = {0: int(aggs[0][start])}
    order = list(range(n_tiles))[::-1] if order is None else order
    for k in order:
        if k == 0:
            continue
        acc = np.arange(delta.shape[0], dtype=np.int64)     # identity
        j = k - 1
        while j not in prefix:
            acc = acc[aggs[j]]                              # earlier tile applies first: acc o agg_j
            j -= 1
        prefix[k] = int(aggs[k][acc[prefix[j]]])
    return prefix[n_tiles - 1]
