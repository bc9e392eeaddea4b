"""Probabilistic NFA-64: FP64 oracle pins (CPU) and the MFMA kernels against
the oracle (GPU).

The quantity is LL = log(a . T[c_L] ... T[c_1] s_0) and logmass = log(1 . s_L)
(the reference's V5 kernel only thresholds a . s_L, cuda/nfa_tc_evolution.cu:552-559,
so the oracle is pinned by closed forms, not by reference vectors).

Tolerances (relative LL error |LL_gpu - LL_fp64| / |LL_fp64|):
  * every mode (f16x2 / bf16x2 split state, f16 / bf16 one rounding) against
    FP64 on the FP32 T: <= 1e-3 for every sequence of the Dirichlet(0.1)
    model at 2,048 steps (the stated precision bound); the default split
    mode also over 200,000 steps;
  * exact cases (permutation matrices, identity): 1e-13.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import prob as oprob
from trex_amd.workloads import dirichlet_pnfa

N = 64
SPLIT_TOL = 1e-3
SINGLE_TOL = {"f16": 1e-3, "bf16": 1e-3}


def _rel(x, ref):
    return np.abs(x - ref) / np.abs(ref)


# ---------------------------------------------------------------------------
# closed forms (shared by the oracle pins and the mutation checks)
# ---------------------------------------------------------------------------

def _perm_model(rng, sigma):
    perms = np.stack([rng.permutation(N) for _ in range(sigma)])        # state j -> perms[c][j]
    T = np.zeros((sigma, N, N), dtype=np.float32)
    for c in range(sigma):
        T[c, perms[c], np.arange(N)] = 1.0
    start = (rng.permutation(N) + 1).astype(np.float32) / 64.0          # distinct, exact in FP16/BF16
    accept = (rng.random(N) < 0.3).astype(np.float32)
    accept[0] = 1.0
    return T, perms, start, accept


def _perm_closed_form(perms, start, accept, symbols):
    """s_L[pos[j]] = s0[j] with pos the composed permutation of each sequence."""
    n_seq, L = symbols.shape
    pos = np.tile(np.arange(N), (n_seq, 1))
    for t in range(L):
        pos = perms[symbols[:, t]][np.arange(n_seq)[:, None], pos]
    s = np.zeros((n_seq, N))
    s[np.arange(n_seq)[:, None], pos] = start.astype(np.float64)[None, :]
    return np.log(s @ accept.astype(np.float64)), np.log(s.sum(axis=1))


def _rank_one_model(rng, sigma):
    u = rng.uniform(0.1, 2.0, size=(sigma, N))                           # T[c] = u_c 1^T
    T = np.repeat(u[:, :, None], N, axis=2).astype(np.float32)
    return T, u.astype(np.float32).astype(np.float64)


def _rank_one_closed_form(u, start, accept, symbols):
    """s_t = u_{c_t} (1 . s_{t-1}):  LL = log(a . u_{c_L}) + sum_{t<L} log(1 . u_{c_t}) + log(1 . s0)."""
    L = symbols.shape[1]
    mass = np.log(u.sum(axis=1))
    ll = np.log(u[symbols[:, -1]] @ accept.astype(np.float64)) + np.log(np.sum(start, dtype=np.float64))
    if L > 1:
        ll = ll + mass[symbols[:, :-1]].sum(axis=1)
    lm = ll - np.log(u[symbols[:, -1]] @ accept.astype(np.float64)) + np.log(u[symbols[:, -1]].sum(axis=1))
    return ll, lm


def check_permutation(fwd, seed=1):
    rng = np.random.default_rng(seed)
    T, perms, start, accept = _perm_model(rng, 5)
    sy = rng.integers(0, 5, size=(40, 37))
    ll, lm = fwd(T, start, accept, sy)
    ref_ll, ref_lm = _perm_closed_form(perms, start, accept, sy)
    return np.allclose(ll, ref_ll, rtol=1e-12, atol=1e-12) and np.allclose(lm, ref_lm, rtol=1e-12, atol=1e-12)


def check_rank_one(fwd, seed=2):
    rng = np.random.default_rng(seed)
    T, u = _rank_one_model(rng, 4)
    start = rng.uniform(0.0, 1.0, N).astype(np.float32)
    accept = (rng.random(N) < 0.5).astype(np.float32)
    accept[3] = 1.0
    sy = rng.integers(0, 4, size=(30, 400))          # products far outside FP64 range without rescaling
    ll, lm = fwd(T, start, accept, sy)
    ref_ll, ref_lm = _rank_one_closed_form(u, start, accept, sy)
    return np.allclose(ll, ref_ll, rtol=1e-11) and np.allclose(lm, ref_lm, rtol=1e-11)


def check_scaling(fwd, seed=3):
    rng = np.random.default_rng(seed)
    T, start, accept = dirichlet_pnfa(rng, 3)
    sy = rng.integers(0, 3, size=(10, 300))
    ll1, _ = fwd(T, start, accept, sy)
    ll2, _ = fwd(T * np.float32(4.0), start, accept, sy)
    return np.allclose(ll2 - ll1, 300 * np.log(4.0), rtol=1e-12, atol=1e-9)


# ---------------------------------------------------------------------------
# oracle pins (CPU)
# ---------------------------------------------------------------------------

def test_oracle_matches_bruteforce_product():
    rng = np.random.default_rng(0)
    T, start, accept = dirichlet_pnfa(rng, 3)
    T = T * rng.uniform(0.5, 1.5, size=T.shape).astype(np.float32)   # not stochastic
    sy = rng.integers(0, 3, size=(3, 4))
    ll, lm = oprob.forward_loglik(T, start, accept, sy)
    for i in range(len(sy)):
        bl, bm = oprob.loglik_bruteforce(T, start, accept, list(sy[i]))
        assert ll[i] == pytest.approx(bl, rel=1e-12)
        assert lm[i] == pytest.approx(bm, rel=1e-12)


def test_oracle_identity():
    rng = np.random.default_rng(5)
    T = np.tile(np.eye(N, dtype=np.float32), (4, 1, 1))
    start = rng.random(N).astype(np.float32)
    accept = (rng.random(N) < 0.4).astype(np.float32)
    sy = rng.integers(0, 4, size=(6, 50))
    ll, lm = oprob.forward_loglik(T, start, accept, sy)
    assert np.allclose(ll, np.log(start.astype(np.float64) @ accept), rtol=1e-14)
    assert np.allclose(lm, np.log(start.astype(np.float64).sum()), rtol=1e-14)


def test_oracle_closed_forms():
    assert check_permutation(oprob.forward_loglik)
    assert check_rank_one(oprob.forward_loglik)
    assert check_scaling(oprob.forward_loglik)


def test_oracle_stationary_limit():
    """One symbol repeated: s_L -> lambda^L c pi with (lambda, pi) the Perron pair
    of T and c = (w . s0) / (w . pi) for the left Perron vector w, so
    LL -> log(a . pi) + L log(lambda) + log(c) (lambda = 1 up to the FP32
    rounding of a column-stochastic T)."""
    rng = np.random.default_rng(7)
    T, start, accept = dirichlet_pnfa(rng, 1, alpha=1.0)
    Td = T[0].astype(np.float64)
    w, v = np.linalg.eig(Td)
    k = np.argmax(np.real(w))
    lam, pi = np.real(w[k]), np.real(v[:, k])
    pi = pi / pi.sum()
    wl, vl = np.linalg.eig(Td.T)
    left = np.real(vl[:, np.argmax(np.real(wl))])
    c = (left @ start.astype(np.float64)) / (left @ pi)
    L = 3000
    ll, lm = oprob.forward_loglik(T, start, accept, np.zeros((1, L), dtype=np.int64))
    assert ll[0] == pytest.approx(np.log(pi @ accept) + L * np.log(lam) + np.log(c), rel=1e-10)
    assert lm[0] == pytest.approx(L * np.log(lam) + np.log(c), abs=1e-10)
    assert abs(lm[0]) < 1e-3


def test_oracle_edge_cases():
    rng = np.random.default_rng(8)
    T, start, accept = dirichlet_pnfa(rng, 2)
    ll, lm = oprob.forward_loglik(T, start, accept, np.zeros((2, 0), dtype=np.int64))
    assert np.allclose(ll, np.log(start.astype(np.float64) @ accept))
    ll, lm = oprob.forward_loglik(T, start, accept, np.array([[0, 1, 2, 0], [0, 1, 1, 0]]))
    assert ll[0] == -np.inf and lm[0] == -np.inf and np.isfinite(ll[1])


def test_c_oracle_matches_numpy_oracle():
    from oracle import cpu as ocpu
    rng = np.random.default_rng(9)
    T, start, accept = dirichlet_pnfa(rng, 5)
    start = rng.uniform(0, 3e-5, N).astype(np.float32)
    sy = rng.integers(0, 5, size=(37, 300)).astype(np.uint8)
    sy[3, 100] = 5
    sy[4, 299] = 9
    scale = rng.uniform(0.99, 1.01, size=(5, N))
    T_last = T * np.float32(1.5)
    for kw in ({}, {"col_scale": scale, "T_last": T_last}):
        ref = oprob.forward_loglik(T, start, accept, sy, **kw)
        for threads in (1, 4):
            got = ocpu.pnfa_forward(T, start, accept, sy, threads=threads, **kw)
            for g, r in zip(got, ref):
                fin = np.isfinite(r)
                assert np.array_equal(fin, np.isfinite(g))
                assert np.allclose(g[fin], r[fin], rtol=1e-13)
    assert check_rank_one(ocpu.pnfa_forward) and check_permutation(ocpu.pnfa_forward)


def test_planted_oracle_mistakes_are_caught():
    """Each closed form rejects a plausible mistake planted in the oracle."""
    transposed = lambda T, s, a, sy: oprob.forward_loglik(np.swapaxes(T, 1, 2), s, a, sy)
    assert not check_rank_one(transposed)
    assert not check_permutation(transposed)

    def dropped_scale(T, s, a, sy):      # forget the first step's exponent
        ll, lm = oprob.forward_loglik(T, s, a, sy[:, 1:])
        return ll, lm
    assert not check_rank_one(dropped_scale)
    assert not check_scaling(dropped_scale) or not check_rank_one(dropped_scale)

    def off_by_one(T, s, a, sy):         # symbols read one step late
        return oprob.forward_loglik(T, s, a, np.roll(sy, 1, axis=1))
    assert not check_permutation(off_by_one)


# ---------------------------------------------------------------------------
# GPU: MFMA kernels through the C ABI
# ---------------------------------------------------------------------------

def _engine(T, start, accept, mode=None):
    from trex_amd.engine import ProbNfaEngine
    return ProbNfaEngine(T, start, accept, mode)


MODES = ["f16x2", "bf16x2", "f16", "bf16"]


@pytest.mark.gpu
@pytest.mark.parametrize("waves", [4, 8])
@pytest.mark.parametrize("mode", MODES)
def test_gpu_permutation_layout_probe(gpu_lib, mode, waves):
    """Permutation matrices are exact in every mode: any mistake in the MFMA
    operand or result layout (a transposed fragment, a wrong row block, a lane
    group writing the wrong rows) moves mass to the wrong state and changes LL."""
    rng = np.random.default_rng(11)
    T, perms, start, accept = _perm_model(rng, 32)
    sy = rng.integers(0, 32, size=(300, 100)).astype(np.uint8)   # 4 full groups + 44, crosses a 64-symbol chunk
    eng = _engine(T, start, accept, mode)
    eng.set_waves(waves)
    ll, lm, _, _ = eng.loglik(sy)
    ref_ll, ref_lm = _perm_closed_form(perms, start, accept, sy)
    assert np.allclose(ll, ref_ll, rtol=1e-13, atol=1e-13), np.flatnonzero(~np.isclose(ll, ref_ll))[:10]
    assert np.allclose(lm, ref_lm, rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("waves", [4, 8])
@pytest.mark.parametrize("mode", MODES)
def test_gpu_rescale_probe(gpu_lib, mode, waves):
    """T[c] = 4 P_c and a one-hot start: the single non-zero entry moves between
    the lanes of a result column and grows by 4 per step. LL = log(a[pos]) +
    L log 4 holds exactly only if every lane of the column agrees on the
    power-of-two rescale (the cross-lane maximum) and the exponent is kept."""
    rng = np.random.default_rng(18)
    T, perms, _, _ = _perm_model(rng, 32)
    T = T * np.float32(4.0)
    start = np.zeros(N, dtype=np.float32)
    start[5] = 1.0
    accept = (np.arange(N, dtype=np.float32) + 1) / 64
    sy = rng.integers(0, 32, size=(130, 150)).astype(np.uint8)
    eng = _engine(T, start, accept, mode)
    eng.set_waves(waves)
    ll, lm, _, _ = eng.loglik(sy)
    pos = np.full(len(sy), 5)
    for t in range(sy.shape[1]):
        pos = perms[sy[:, t], pos]
    ref_ll = np.log(accept.astype(np.float64)[pos]) + sy.shape[1] * np.log(4.0)
    assert np.allclose(ll, ref_ll, rtol=1e-13, atol=1e-12)
    assert np.allclose(lm, sy.shape[1] * np.log(4.0), rtol=1e-13)


@pytest.mark.gpu
def test_gpu_rounding_is_nearest_even(gpu_lib):
    rng = np.random.default_rng(12)
    T, start, accept = dirichlet_pnfa(rng, 4)
    eng = _engine(T, start, accept)
    t16, sc16 = eng.rounded("f16x2")
    assert np.array_equal(t16, T.astype(np.float16).astype(np.float32))
    bits = T.view(np.uint32).astype(np.uint64)
    bf = (((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    tb, scb = eng.rounded("bf16")
    assert np.array_equal(tb, bf)
    for tq, sc in ((t16, sc16), (tb, scb)):   # the column factors restore every column's mass
        assert sc.dtype == np.float64
        mass = (tq.astype(np.float64) * sc[:, None, :]).sum(axis=1)
        # to FP64 accuracy: the kernels apply each factor as an FP32 pair; one FP32
        # factor alone would miss by up to 6e-8 on every step with that symbol
        assert np.allclose(mass, T.astype(np.float64).sum(axis=1), rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", MODES)
def test_gpu_dirichlet_precision(gpu_lib, mode):
    """The precision workload's model at 1,024 sequences x 2,048 symbols."""
    rng = np.random.default_rng(13)
    T, start, accept = dirichlet_pnfa(rng, 32)
    sy = rng.integers(0, 32, size=(1024, 2048)).astype(np.uint8)
    eng = _engine(T, start, accept, mode)
    ll, lm, _, _ = eng.loglik(sy)
    ref_ll, ref_lm = oprob.forward_loglik(T, start, accept, sy)
    err = _rel(ll, ref_ll)
    tol = SPLIT_TOL if mode.endswith("x2") else SINGLE_TOL[mode]
    print(f"{mode}: rel LL err median {np.median(err):.2e} max {err.max():.2e}")
    assert err.max() <= tol
    assert np.abs(lm - ref_lm).max() <= tol * np.abs(ref_ll).max()
    if mode.endswith("x2"):
        # against FP64 on the operator the mode applies (rounded T, column factors):
        # what is left is the split state's rounding
        tq, sc = eng.rounded(mode)
        q_ll, _ = oprob.forward_loglik(tq, start, accept, sy, col_scale=sc, T_last=T)
        assert _rel(ll, q_ll).max() <= 2e-4


@pytest.mark.gpu
@pytest.mark.parametrize("waves", [4, 8])
def test_gpu_skewed_symbols_many_tiles(gpu_lib, waves):
    """90% of the sequences on one symbol: up to 4 tiles of 16 columns per step."""
    rng = np.random.default_rng(14)
    T, start, accept = dirichlet_pnfa(rng, 8)
    sy = np.where(rng.random((512, 300)) < 0.9, 3, rng.integers(0, 8, size=(512, 300))).astype(np.uint8)
    eng = _engine(T, start, accept)
    eng.set_waves(waves)
    ll, lm, _, _ = eng.loglik(sy)
    ref_ll, _ = oprob.forward_loglik(T, start, accept, sy)
    assert _rel(ll, ref_ll).max() <= SPLIT_TOL


@pytest.mark.gpu
def test_gpu_nonstochastic_and_edges(gpu_lib):
    rng = np.random.default_rng(15)
    T, start, accept = dirichlet_pnfa(rng, 6)
    T = (T * rng.uniform(0.2, 3.0, size=T.shape)).astype(np.float32)
    start = rng.uniform(0, 1e-3, N).astype(np.float32)
    eng = _engine(T, start, accept)
    for L in (0, 1, 2, 63, 64, 65, 129):
        sy = rng.integers(0, 6, size=(70, L)).astype(np.uint8)
        ll, lm, _, _ = eng.loglik(sy)
        ref_ll, ref_lm = oprob.forward_loglik(T, start, accept, sy)
        assert _rel(ll, ref_ll).max() <= SPLIT_TOL, L
        assert np.allclose(lm, ref_lm, rtol=SPLIT_TOL, atol=1e-5), L
    sy = rng.integers(0, 6, size=(70, 40)).astype(np.uint8)
    sy[5, 17] = 6          # outside the alphabet: the state empties
    sy[9, 39] = 200
    ll, lm, _, _ = eng.loglik(sy)
    assert ll[5] == -np.inf and ll[9] == -np.inf and lm[5] == -np.inf
    assert np.isfinite(np.delete(ll, [5, 9])).all()
    ll, lm, _, _ = eng.loglik(np.zeros((0, 10), dtype=np.uint8))
    assert ll.shape == (0,)


@pytest.mark.gpu
@pytest.mark.parametrize("waves", [4, 8])
def test_gpu_ragged_sequences(gpu_lib, waves):
    """Sequences of different lengths (0, 1, 2, ..., 400) in one launch: each
    matches the FP64 oracle run on its own length."""
    rng = np.random.default_rng(19)
    T, start, accept = dirichlet_pnfa(rng, 32)
    lengths = np.concatenate([[0, 1, 2, 3, 64, 65, 66], rng.integers(0, 400, size=300)])
    offsets = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    data = rng.integers(0, 32, size=int(offsets[-1])).astype(np.uint8)
    bad = np.flatnonzero(lengths > 10)[:3]
    data[offsets[bad] + 5] = 40                        # outside the alphabet
    eng = _engine(T, start, accept)
    eng.set_waves(waves)
    ll, lm = eng.loglik_ragged(data, offsets)
    for L in np.unique(lengths):
        idx = np.flatnonzero(lengths == L)
        sy = np.stack([data[offsets[i]:offsets[i + 1]] for i in idx]) if L else np.zeros((len(idx), 0), np.uint8)
        ref_ll, ref_lm = oprob.forward_loglik(T, start, accept, sy)
        fin = np.isfinite(ref_ll)
        assert np.array_equal(np.isfinite(ll[idx]), fin), L
        assert np.all(_rel(ll[idx][fin], ref_ll[fin]) <= SPLIT_TOL), L
    assert np.all(ll[bad] == -np.inf)


@pytest.mark.gpu
@pytest.mark.parametrize("waves", [4, 8])
def test_gpu_outside_symbols_at_staging_edges(gpu_lib, waves):
    """Symbols outside the alphabet at the positions where the kernel stages
    its next 64 symbols (0, 63, 64, 127, 128, ...), at a sequence's last
    rounded step (length - 2) and at its last symbol (length - 1): those
    sequences give -inf for both outputs, every other sequence of the launch
    matches the FP64 oracle."""
    rng = np.random.default_rng(23)
    T, start, accept = dirichlet_pnfa(rng, 32)
    n = 150
    lengths = rng.integers(200, 330, size=n)
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    data = rng.integers(0, 32, size=int(offsets[-1])).astype(np.uint8)
    cases = [0, 1, 62, 63, 64, 65, 127, 128, 129, 191, 192]
    bad = []
    for i, pos in enumerate(cases):
        data[offsets[2 * i] + pos] = 32 + i          # outside the 32-symbol alphabet
        bad.append(2 * i)
    for i, back in enumerate((2, 1)):                # last rounded step, last symbol
        j = 2 * len(cases) + 2 * i
        data[offsets[j + 1] - back] = 255
        bad.append(j)
    eng = _engine(T, start, accept)
    eng.set_waves(waves)
    ll, lm = eng.loglik_ragged(data, offsets)
    bad = np.array(bad)
    assert np.all(ll[bad] == -np.inf) and np.all(lm[bad] == -np.inf)
    good = np.setdiff1d(np.arange(n), bad)
    for i in good:
        sy = data[offsets[i]:offsets[i + 1]][None, :]
        ref_ll, ref_lm = oprob.forward_loglik(T, start, accept, sy)
        assert _rel(ll[i:i + 1], ref_ll).max() <= SPLIT_TOL, i
        assert abs(lm[i] - ref_lm[0]) <= SPLIT_TOL * abs(ref_ll[0]) + 1e-5, i


@pytest.mark.gpu
@pytest.mark.parametrize("grid", [1, 3])
def test_gpu_grid_cap_same_results(gpu_lib, grid):
    """A grid cap splits a call into consecutive launches of at most `grid`
    workgroups: uniform and ragged results are bit-identical to one launch."""
    rng = np.random.default_rng(23)
    T, start, accept = dirichlet_pnfa(rng, 32)
    eng = _engine(T, start, accept)
    sy = rng.integers(0, 32, size=(64 * 7 + 5, 129)).astype(np.uint8)
    lengths = rng.integers(0, 300, size=64 * 5 + 17)
    offsets = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    data = rng.integers(0, 32, size=int(offsets[-1])).astype(np.uint8)
    ll0, lm0, _, _ = eng.loglik(sy)
    rl0, rm0 = eng.loglik_ragged(data, offsets)
    eng.set_grid(grid)
    ll1, lm1, _, _ = eng.loglik(sy)
    rl1, rm1 = eng.loglik_ragged(data, offsets)
    for a, b in ((ll0, ll1), (lm0, lm1), (rl0, rl1), (rm0, rm1)):
        assert np.array_equal(a, b)
    from trex_amd._lib import TrexError
    with pytest.raises(TrexError):
        eng.set_grid(-1)


# ---------------------------------------------------------------------------
# parallel in time (chunk products chained per sequence)
# ---------------------------------------------------------------------------

def _ragged(rng, lengths, sigma=32):
    offsets = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    return rng.integers(0, sigma, size=int(offsets[-1])).astype(np.uint8), offsets


@pytest.mark.gpu
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("chunk", [1, 7, 64, 0])
def test_gpu_chunked_permutation_exact(gpu_lib, mode, chunk):
    """Permutation matrices: every chunk product is a permutation matrix, so
    LL is exact in every mode and chunk length; a wrong k order between the
    product's result and the next operand, a transposed fragment or a wrong
    chaining order moves mass to the wrong state."""
    rng = np.random.default_rng(41)
    T, perms, start, accept = _perm_model(rng, 32)
    L = 150
    sy = rng.integers(0, 32, size=(70, L)).astype(np.uint8)
    offsets = np.arange(71, dtype=np.int64) * L
    eng = _engine(T, start, accept, mode)
    ll, lm = eng.loglik_ragged(sy.reshape(-1), offsets, chunk_len=chunk)
    ref_ll, ref_lm = _perm_closed_form(perms, start, accept, sy)
    assert np.allclose(ll, ref_ll, rtol=1e-13, atol=1e-13), np.flatnonzero(~np.isclose(ll, ref_ll))[:10]
    assert np.allclose(lm, ref_lm, rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
def test_gpu_chunked_rescale_exact(gpu_lib):
    """T[c] = 4 P_c: each column's power-of-two rescale inside a chunk and the
    exponent alignment of the chaining must add up to exactly L log 4."""
    rng = np.random.default_rng(43)
    T, perms, _, _ = _perm_model(rng, 32)
    T = T * np.float32(4.0)
    start = np.zeros(N, dtype=np.float32)
    start[5] = 1.0
    accept = (np.arange(N, dtype=np.float32) + 1) / 64
    L = 333
    sy = rng.integers(0, 32, size=(20, L)).astype(np.uint8)
    eng = _engine(T, start, accept)
    ll, lm = eng.loglik_ragged(sy.reshape(-1), np.arange(21, dtype=np.int64) * L, chunk_len=50)
    pos = np.full(len(sy), 5)
    for t in range(L):
        pos = perms[sy[:, t], pos]
    assert np.allclose(ll, np.log(accept.astype(np.float64)[pos]) + L * np.log(4.0), rtol=1e-13, atol=1e-12)
    assert np.allclose(lm, L * np.log(4.0), rtol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["f16x2", "bf16x2"])
def test_gpu_chunked_dirichlet_precision(gpu_lib, mode):
    """Dirichlet(0.1) model, ragged lengths (0, 1, 2, a chunk, a chunk + 1, up to
    3,000), bytes outside the alphabet: within the split modes' tolerance of
    FP64, and close to the batch kernel (same operator, other rounding order)."""
    rng = np.random.default_rng(47)
    T, start, accept = dirichlet_pnfa(rng, 32)
    lengths = np.concatenate([[0, 1, 2, 3, 64, 65, 129], rng.integers(0, 3000, size=120)])
    data, offsets = _ragged(rng, lengths)
    bad = np.flatnonzero(lengths > 10)[:2]
    data[offsets[bad] + 7] = 40                       # outside the alphabet
    eng = _engine(T, start, accept, mode)
    ll, lm = eng.loglik_ragged(data, offsets, chunk_len=64)
    bl, bm = eng.loglik_ragged(data, offsets)
    assert np.all(ll[bad] == -np.inf) and np.all(bl[bad] == -np.inf)
    ok = np.setdiff1d(np.arange(len(lengths)), bad)
    for j in ok:
        ref_ll, ref_lm = oprob.forward_loglik(T, start, accept, data[offsets[j]:offsets[j + 1]][None, :])
        assert _rel(ll[j:j + 1], ref_ll).max() <= SPLIT_TOL, (j, lengths[j])
        assert abs(lm[j] - ref_lm[0]) <= SPLIT_TOL * max(abs(ref_ll[0]), 1.0)
    assert _rel(ll[ok], bl[ok]).max() <= SPLIT_TOL


@pytest.mark.gpu
def test_gpu_chunked_long_sequence(gpu_lib):
    """A 200,000-symbol sequence (3,125 chunks of 64): within tolerance of the
    C FP64 oracle and of the batch kernel's sequential walk."""
    rng = np.random.default_rng(53)
    T, start, accept = dirichlet_pnfa(rng, 32)
    sy = rng.integers(0, 32, size=(2, 200_000)).astype(np.uint8)
    offsets = np.array([0, 200_000, 400_000], np.int64)
    from oracle import cpu as ocpu
    eng = _engine(T, start, accept)
    ll, lm = eng.loglik_ragged(sy.reshape(-1), offsets, chunk_len=64)
    bl, bm = eng.loglik_ragged(sy.reshape(-1), offsets)
    ref_ll, ref_lm = ocpu.pnfa_forward(T, start, accept, sy)
    assert _rel(ll, ref_ll).max() <= SPLIT_TOL and _rel(bl, ref_ll).max() <= SPLIT_TOL
    assert _rel(ll, bl).max() <= SPLIT_TOL


@pytest.mark.gpu
def test_gpu_chunked_argument_errors(gpu_lib):
    from trex_amd._lib import TrexError
    from trex_amd.device import DeviceBuffer
    rng = np.random.default_rng(59)
    T, start, accept = dirichlet_pnfa(rng, 32)
    eng = _engine(T, start, accept)
    d = DeviceBuffer(64)
    with pytest.raises(TrexError):   # end before begin
        eng.loglik_chunked_device(d.ptr, np.array([5]), np.array([3]), d.ptr, d.ptr)
    with pytest.raises(TrexError):   # negative chunk length
        eng.loglik_chunked_device(d.ptr, np.array([0]), np.array([3]), d.ptr, d.ptr, chunk_len=-1)
    with pytest.raises(ValueError):
        eng.loglik_chunked_device(d.ptr, np.array([0, 1]), np.array([3]), d.ptr, d.ptr)


@pytest.mark.gpu
def test_gpu_device_pointer_dispatch(gpu_lib):
    from trex_amd.device import DeviceBuffer, Stream
    rng = np.random.default_rng(16)
    T, start, accept = dirichlet_pnfa(rng, 32)
    sy = rng.integers(0, 32, size=(200, 333)).astype(np.uint8)
    eng = _engine(T, start, accept)
    host_ll, host_lm, _, _ = eng.loglik(sy)
    stream = Stream()
    d_sy = DeviceBuffer.from_array(sy)
    d_ll = DeviceBuffer(200 * 8)
    d_lm = DeviceBuffer(200 * 8)
    eng.loglik_device(d_sy.ptr, 200, 333, d_ll.ptr, d_lm.ptr, stream)
    ll = d_ll.download(np.float64, 200, stream)
    lm = d_lm.download(np.float64, 200, stream)
    assert np.array_equal(ll, host_ll) and np.array_equal(lm, host_lm)


@pytest.mark.gpu
def test_gpu_argument_errors(gpu_lib):
    from trex_amd._lib import TrexError
    rng = np.random.default_rng(17)
    T, start, accept = dirichlet_pnfa(rng, 33)
    with pytest.raises(TrexError):
        _engine(T, start, accept)
    T = T[:4].copy()
    T[1, 2, 3] = np.inf
    with pytest.raises(TrexError):
        _engine(T, start, accept)
    T[1, 2, 3] = 1e6        # beyond FP16: BF16 only
    eng = _engine(T, start, accept)
    assert eng.info()["mode"] == "bf16x2" and not eng.info()["f16_ok"]
    assert _engine(T[:, :, :] * 0 + 0.01, start, accept).info()["mode"] == "f16x2"   # FP16 range: split FP16 state
    with pytest.raises(TrexError):
        eng.set_mode("f16")
    with pytest.raises(TrexError):
        eng.set_waves(16)
    with pytest.raises(ValueError):
        _engine(T[:, :32, :], start, accept)


def _emulate_chunked_ll(T, start, accept, syms, clen, fp64_masses):
    """CPU emulation of the chunk products' arithmetic (FP32 state, FP16
    high/low operand split, FP32 products, per-step mass restore and
    power-of-two rescale, FP32 storage) chained in FP64: the LL minus the
    FP64 forward pass's. fp64_masses: the restore's masses in FP64 and its
    factor as an FP32 pair; otherwise both in FP32."""
    T = T.astype(np.float64)
    Tq = T.astype(np.float16).astype(np.float64)
    f = T.sum(1) / Tq.sum(1)
    L = len(syms)
    s = start.astype(np.float64)
    e = 0
    for c in syms:
        s = T[c] @ s
        _, x = np.frexp(s.max())
        s, e = s / 2.0 ** x, e + x
    ref = np.log(accept @ s) + e * np.log(2)
    s, scale = start.astype(np.float64), 0
    for b in range(0, L - 1, clen):
        D = np.eye(64, dtype=np.float32)
        E = np.zeros(64, np.int64)
        for c in syms[b:min(L - 1, b + clen)]:
            x = (D.astype(np.float64) * f[c][:, None]).astype(np.float32)
            hi = x.astype(np.float16).astype(np.float32)
            lo = (x - hi).astype(np.float16).astype(np.float32)
            acc = Tq[c].astype(np.float32) @ hi + Tq[c].astype(np.float32) @ lo
            mi = (Tq[c].sum(0)[:, None] * x.astype(np.float64)).sum(0)
            mo = acc.astype(np.float64).sum(0)
            if fp64_masses:
                fix = np.where((mi > 0) & (mo > 0), mi / mo, 1.0)
                fh = fix.astype(np.float32)
                fl = (fix - fh).astype(np.float32)
                Dn = (acc.astype(np.float64) * fh + (acc * fl).astype(np.float64)).astype(np.float32)
            else:
                mi, mo = mi.astype(np.float32), mo.astype(np.float32)
                fix = np.where((mi > 0) & (mo > 0), mi / mo, 1).astype(np.float32)
                Dn = (acc * fix[None, :]).astype(np.float32)
            _, ex = np.frexp(np.abs(Dn).max(0))
            D = (Dn * (2.0 ** -ex)[None, :]).astype(np.float32)
            E += ex
        v = s * 2.0 ** E
        top = np.frexp(np.abs(v).max())[1]
        s = D.astype(np.float64) @ (v / 2.0 ** top)
        ex2 = np.frexp(np.abs(s).max())[1]
        s, scale = s / 2.0 ** ex2, scale + top + ex2
    s = T[syms[-1]] @ s
    return np.log(accept @ s) + scale * np.log(2) - ref


def test_chunk_mass_restore_in_fp32_drifts():
    """Where the chunked path's chunk-count-dependent error comes from, in a
    round-to-nearest model of the chunk products: each chunk starts from unit
    columns, whose masses sit next to powers of two, so restoring every
    column's mass with FP32 masses and an FP32 factor rounds the same way in
    every chunk, and the LL drifts in one direction by about the same amount
    per chunk. With FP64 masses and the factor as an FP32 pair the drift is
    gone in this model. (The kernel keeps the FP32 restore: the matrix cores'
    FP32 accumulation is biased toward zero, which adds a per-step drift of
    the other sign that the FP32 restore partly cancels.) Emulated on the CPU
    over 4,000 symbols in 1,000 chunks of 4 and 250 chunks of 16."""
    rng = np.random.default_rng(11)
    T, start, accept = dirichlet_pnfa(rng, 32, alpha=1.0)
    drift, fixed = [], []
    for seed in range(2):
        syms = np.random.default_rng(100 + seed).integers(0, 32, 4000)
        drift.append(_emulate_chunked_ll(T, start, accept, syms, 4, False))
        fixed.append(_emulate_chunked_ll(T, start, accept, syms, 4, True))
    assert all(d > 0 for d in drift) or all(d < 0 for d in drift), drift   # one direction
    assert min(abs(d) for d in drift) > 3 * max(abs(x) for x in fixed), (drift, fixed)
    # and about in proportion to the number of chunks
    d16 = _emulate_chunked_ll(T, start, accept, np.random.default_rng(100).integers(0, 32, 4000), 16, False)
    assert abs(drift[0]) > 2 * abs(d16) and np.sign(d16) == np.sign(drift[0]), (drift[0], d16)
