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



# ---------------------------------------------------------------------------
# Wave-level model of the chunked kernel's publication protocol
# (scan_kernels.hip / mix_kernels.hip, no reference counterpart: the reference
# looks back tile by tile from one thread, cuda/monoid_scan.cu:128-222)
# ---------------------------------------------------------------------------
def plan_chunks(lengths, chunk_len: int):
    """Chunks of strings of the given lengths in the kernel's processing order
    (the ordering rule of trex_plan_create: longer chunks first, a string's
    first chunk ahead of other chunks of equal length, stable). Returns
    (sid, lo, hi, pred) per chunk in that order, pred = index of the chunk
    before it in its string or -1."""
    raw = []
    for i, n in enumerate(lengths):
        k = 1 if n <= chunk_len else -(-n // chunk_len)
        for j in range(k):
            raw.append((i, j * chunk_len, min(n, (j + 1) * chunk_len), j))
    order = sorted(range(len(raw)), key=lambda c: (-(raw[c][2] - raw[c][1]), 0 if raw[c][3] == 0 else 1))
    pos = {c: k for k, c in enumerate(order)}
    out = []
    for c in order:
        sid, lo, hi, j = raw[c]
        out.append((sid, lo, hi, -1 if j == 0 else pos[c - 1]))
    return out


def _is_const(f: np.ndarray) -> bool:
    return bool(np.all(f == f[0]))


def wave_protocol(delta: np.ndarray, start: int, strings, chunk_len: int, rng, in_wave: bool = True,
                  publish_after_lookback: bool = False, wave: int = 64):
    """Runs the chunked kernel's protocol, waves of ``wave`` chunks scheduled
    in a random interleaving, and returns (final state per string, reads,
    stalled).

    Each wave computes every chunk's function (a string's first chunk: the
    constant function of its walked state), resolves its links in registers,
    publishes, and lets the lanes that must wait look back through the
    published statuses:

    * ``in_wave=True`` (the kernel): a chunk whose predecessor lies in its own
      wave composes that lane's function by pointer jumping (the shuffles);
      only chain roots look back, and a root's predecessor is in an earlier
      wave.
    * ``in_wave=False`` (the previous kernel): a lane is linked only to the
      lane just before it (pred == c - 1); segment heads look back through
      the statuses, possibly to a chunk of their own wave.
    * ``publish_after_lookback``: the wave's stores come after its look-back
      (an order the kernel must not depend on).

    ``reads`` lists (looking chunk, read chunk, published) for every status
    read of a look-back; ``stalled`` is True if the waves left can never
    finish (each waits on a store no running wave will make)."""
    ident = np.arange(delta.shape[0])
    chunks = plan_chunks([len(s) for s in strings], chunk_len)
    n = len(chunks)
    own = []
    for sid, lo, hi, pred in chunks:
        f = chunk_function(delta, strings[sid], lo, hi)
        own.append(np.full_like(f, f[start]) if pred < 0 else f)
    status: dict = {}                  # chunk -> ("P", state) | ("A", function, jump)
    finals: dict = {}
    reads: list = []

    def resolve(base, lanes):
        g = {c: own[c] for c in lanes}
        root = {}
        for c in lanes:
            p = chunks[c][3]
            linked = (p >= base) if in_wave else (p == c - 1 and c > base)
            root[c] = p if linked else c
        closed = {c: root[c] == c for c in lanes}
        for _ in range(6):             # synchronous rounds, like the shuffles
            g0, r0, k0 = dict(g), dict(root), dict(closed)
            for c in lanes:
                if not k0[c]:
                    g[c] = g0[c][g0[r0[c]]]          # the root side applies first
                    closed[c] = k0[r0[c]]
                    root[c] = r0[r0[c]]
        assert all(closed.values())
        return g, root

    def run(w):
        base = w * wave
        lanes = list(range(base, min(n, base + wave)))
        g, root = resolve(base, lanes)

        def publish():
            for c in lanes:
                status[c] = ("P", int(g[c][0])) if _is_const(g[c]) else ("A", g[c], chunks[root[c]][3])

        if not publish_after_lookback:
            publish()
        incoming = {}
        for c in lanes:
            if root[c] != c or _is_const(g[c]):
                continue
            acc, j = ident, chunks[c][3]
            while True:
                st = status.get(j)
                reads.append((c, j, st is not None))
                if st is None:
                    yield False
                    continue
                if st[0] == "P":
                    incoming[c] = int(acc[st[1]])
                    break
                acc = acc[st[1]]                     # the earlier chunks apply first
                j = st[2]
        if publish_after_lookback:
            publish()
        for c in lanes:
            state = int(g[c][0]) if _is_const(g[c]) else int(g[c][incoming[root[c]]])
            status[c] = ("P", state)
            sid, _, hi, _ = chunks[c]
            if hi == len(strings[sid]):
                finals[sid] = state
        yield True

    live = {w: run(w) for w in range(-(-n // wave))}
    stalled = False
    while live:
        before = len(status), len(live)
        for w in list(rng.permutation(list(live))):
            for step in live[w]:
                if step:
                    del live[w]
                break
        if live and (len(status), len(live)) == before:
            # a full pass without a publication or a finished wave: every
            # live wave polls a chunk that no live wave will publish
            stalled = True
            break
    out = [finals.get(i, -1) for i in range(len(strings))]
    return out, reads, stalled
