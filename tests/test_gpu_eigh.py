"""The batched eigensolver on its own (nrgk_eigh_batched: qmatrix::diag, nrgcore/include/utils/
qmatrix.hpp:375-404, for a list of blocks) on spectra chosen to be hard for it.

Inputs: H = Q diag(lam) Q^T with Q Haar-distributed (QR of a seeded Gaussian matrix, column signs
fixed) and lam chosen per block -- exact degeneracies of high multiplicity, clusters with gaps of
1e-9, 1e-11 and 1e-13, gaps spread from 1e-11 to 1e-8 (across the refinement's cluster threshold),
values graded over eight orders of magnitude, uniform values -- for block
sizes 1 to 800 in one batch (the widest goes through the crossover to the tridiagonalising solver).

Pins, by construction and independent of LAPACK: the eigenvalues are lam (to the rounding of forming
H, ~n eps |lam|_max); U is orthogonal; H U = U diag(w).  Tolerances (the eigen-error bar of the
chain tests, relative to the block's norm): |w - lam| <= 1e-10 |lam|_max, |U^T U - I|_max <= 1e-10,
|H U - U diag(w)|_max <= 1e-10 |lam|_max.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 16, 17, 31, 33, 64, 100, 257, 619, 800]


def _haar(rng, n):
    Q, R = np.linalg.qr(rng.standard_normal((n, n)))
    return Q * np.sign(np.diag(R))


def _spectrum(rng, n, kind):
    if kind == "degenerate":
        return np.sort(rng.choice([-1.0, -0.3, 0.0, 0.2, 0.7, 1.5], n))
    if kind == "clustered":
        base = np.sort(rng.uniform(-1, 1, max(1, n // 4)))
        lam = []
        for k, b in enumerate(base):
            gap = (1e-9, 1e-11, 1e-13)[k % 3]
            lam += [b + gap * j for j in range(4)]
        return np.sort(np.array(lam[:n] + list(rng.uniform(-1, 1, n - len(lam[:n])))))
    if kind == "threshold":  # gaps from 1e-11 to 1e-8, across the refinement's cluster threshold
        return np.sort(np.cumsum(np.concatenate([[rng.uniform(-1, 0)], 10.0 ** rng.uniform(-11, -8, n - 1)])))
    if kind == "graded":
        return np.sort(rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-8, 0, n))
    return np.sort(rng.uniform(-1, 1, n))


def _case(seed):
    rng = np.random.default_rng(seed)
    kinds = ["degenerate", "clustered", "graded", "uniform", "threshold"]
    blocks, lams = [], []
    for i, n in enumerate(SIZES):
        lam = _spectrum(rng, n, kinds[(i + seed) % 5])
        Q = _haar(rng, n)
        H = (Q * lam) @ Q.T
        blocks.append(0.5 * (H + H.T))
        lams.append(lam)
    return blocks, lams


def _check(blocks, lams, out, tol=1e-10):
    for H, lam, (w, U) in zip(blocks, lams, out):
        n = len(lam)
        scale = max(1.0e-300, np.abs(lam).max()) if n else 1.0
        assert np.all(np.diff(w) >= 0), n
        assert np.abs(w - lam).max() <= tol * scale, (n, np.abs(w - lam).max())
        assert np.abs(U.T @ U - np.eye(n)).max() <= tol, n
        assert np.abs(H @ U - U * w).max() <= tol * scale, (n, np.abs(H @ U - U * w).max())


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_constructed_spectra_mixed(nk, seed):
    blocks, lams = _case(seed)
    _check(blocks, lams, nk.eigh_batched(blocks))


def test_constructed_spectra_plain_and_modes_agree(nk):
    blocks, lams = _case(3)
    plain = nk.eigh_batched(blocks, mode=nk.NRGK_MODE_PLAIN64)
    mixed = nk.eigh_batched(blocks)
    _check(blocks, lams, plain)
    for (wp, _), (wm, _), lam in zip(plain, mixed, lams):
        assert np.abs(wp - wm).max() <= 2e-10 * max(1.0, np.abs(lam).max())


def test_special_blocks(nk):
    """zero blocks, a diagonal block with repeated entries, the identity, a 2x2 with a tiny
    off-diagonal, and empty blocks among them"""
    rng = np.random.default_rng(7)
    d = np.sort(rng.choice([0.5, -0.25, 2.0], 40))
    blocks = [np.zeros((0, 0)), np.zeros((5, 5)), np.diag(d), np.eye(24), np.zeros((0, 0)),
              np.array([[1.0, 1e-14], [1e-14, 1.0 + 1e-13]])]
    out = nk.eigh_batched(blocks)
    assert [len(w) for w, _ in out] == [0, 5, 40, 24, 0, 2]
    np.testing.assert_array_equal(out[1][0], np.zeros(5))
    np.testing.assert_allclose(out[2][0], d, atol=1e-14)
    np.testing.assert_allclose(out[3][0], np.ones(24), atol=1e-14)
    for H, (w, U) in zip(blocks, out):
        n = len(w)
        if n:
            assert np.abs(U.T @ U - np.eye(n)).max() <= 1e-12
            assert np.abs(H @ U - U * w).max() <= 1e-12 * max(1.0, np.abs(w).max())
    assert nk.eigh_batched([]) == []


def test_wide_clusters(nk):
    """clusters wider than the in-LDS Rayleigh-Ritz (64 members): an exactly 100-fold eigenvalue, an
    80-member cluster spread over 1e-12, and two 70-fold degenerate levels 1e-10 apart"""
    rng = np.random.default_rng(11)
    lam1 = np.sort(np.concatenate([np.full(100, 0.25), rng.uniform(-1, 1, 200)]))
    lam2 = np.sort(np.concatenate([0.1 + 1e-12 * np.linspace(0, 1, 80), rng.uniform(-1, 1, 120)]))
    lam3 = np.sort(np.concatenate([np.full(70, -0.5), np.full(70, -0.5 + 1e-10), rng.uniform(-1, 1, 60)]))
    blocks, lams = [], []
    for lam in (lam1, lam2, lam3):
        Q = _haar(rng, len(lam))
        H = (Q * lam) @ Q.T
        blocks.append(0.5 * (H + H.T))
        lams.append(lam)
    _check(blocks, lams, nk.eigh_batched(blocks))


def test_random_batches(nk):
    """eight batches of ten blocks of random sizes 1..899, each with a spectrum of a random kind"""
    rng = np.random.default_rng(123)
    kinds = ["degenerate", "clustered", "graded", "uniform", "threshold"]
    for _ in range(8):
        blocks, lams = [], []
        for n in rng.integers(1, 900, 10):
            lam = _spectrum(rng, int(n), kinds[rng.integers(5)])
            Q = _haar(rng, int(n))
            H = (Q * lam) @ Q.T
            blocks.append(0.5 * (H + H.T))
            lams.append(lam)
        _check(blocks, lams, nk.eigh_batched(blocks))


def test_unconverged_refinement_is_an_error(nk, monkeypatch):
    """with the refinement limited to one step (NRGK_OA_ITERS=1) a 200-wide block cannot reach its
    stopping point: the stand-alone solver reports that instead of returning the vectors"""
    monkeypatch.setenv("NRGK_OA_ITERS", "1")
    rng = np.random.default_rng(5)
    lam = np.sort(rng.uniform(-1, 1, 200))
    Q = _haar(rng, 200)
    with pytest.raises(nk.NrgkError, match="unconverged"):
        nk.eigh_batched([(Q * lam) @ Q.T])


# ---------------------------------------------------------------- wide blocks: Householder path
def _case_sizes(seed, sizes, kinds):
    rng = np.random.default_rng(seed)
    blocks, lams = [], []
    for i, n in enumerate(sizes):
        lam = _spectrum(rng, n, kinds[i % len(kinds)])
        Q = _haar(rng, n)
        H = (Q * lam) @ Q.T
        blocks.append(0.5 * (H + H.T))
        lams.append(lam)
    return blocks, lams


@pytest.mark.parametrize("seed", [21, 22])
def test_wide_blocks_householder(nk, seed):
    """blocks above the crossover (769 .. 1600 rows) through the FP64 Householder tridiagonalisation,
    bisection, inverse iteration and WY back-transformation, beside narrow blocks in the same batch:
    degenerate levels (clusters of ~200 members), clusters at 1e-9 .. 1e-13, graded and uniform
    spectra"""
    blocks, lams = _case_sizes(seed, [1600, 40, 1025, 769, 300, 1200],
                               ["degenerate", "uniform", "clustered", "graded", "uniform", "uniform"])
    _check(blocks, lams, nk.eigh_batched(blocks))


def test_wide_blocks_at_the_headline_width(nk):
    """the widest blocks of the two-channel K=4096 workload (~4 300 rows): a degenerate spectrum
    (six levels, clusters of ~700) and a clustered one (gaps 1e-9 .. 1e-13), Haar-rotated, through the
    Householder path, to the same 1e-10 bars"""
    blocks, lams = _case_sizes(43, [4300, 3100], ["clustered", "degenerate"])
    print("wide blocks 4300 + 3100 built", flush=True)
    _check(blocks, lams, nk.eigh_batched(blocks))


@pytest.mark.parametrize("mixed_max", ["0", "40"])
def test_householder_path_on_small_blocks(nk, monkeypatch, mixed_max):
    """the crossover moved down (NRGK_MIXED_MAX): blocks of 1 .. 300 rows, below and above one 64-row
    tile, and the special blocks (zero, identity, repeated diagonal, tiny off-diagonal) through the
    Householder path"""
    monkeypatch.setenv("NRGK_MIXED_MAX", mixed_max)
    blocks, lams = _case_sizes(31, [1, 2, 3, 41, 63, 64, 65, 127, 129, 300],
                               ["degenerate", "clustered", "graded", "uniform", "threshold"])
    _check(blocks, lams, nk.eigh_batched(blocks))
    rng = np.random.default_rng(7)
    d = np.sort(rng.choice([0.5, -0.25, 2.0], 40))
    special = [np.zeros((5, 5)), np.diag(d), np.eye(24), np.array([[1.0, 1e-14], [1e-14, 1.0 + 1e-13]])]
    out = nk.eigh_batched(special)
    np.testing.assert_array_equal(out[0][0], np.zeros(5))
    np.testing.assert_allclose(out[1][0], d, atol=1e-14)
    np.testing.assert_allclose(out[2][0], np.ones(24), atol=1e-14)
    for H, (w, U) in zip(special, out):
        n = len(w)
        assert np.abs(U.T @ U - np.eye(n)).max() <= 1e-12
        assert np.abs(H @ U - U * w).max() <= 1e-12 * max(1.0, np.abs(w).max())


def test_upper_triangle_only_input(nk, monkeypatch):
    """LAPACKE_dsyevd(ROW_MAJOR, 'U') reads the upper triangle only (qmatrix.hpp:390-391): a block
    given with its strict lower triangle zeroed has the eigenpairs of its symmetric completion, below
    and above the crossover and in the plain FP64 form"""
    rng = np.random.default_rng(41)
    blocks, full = [], []
    for n in (50, 800):
        lam = np.sort(rng.uniform(-1, 1, n))
        Q = _haar(rng, n)
        H = (Q * lam) @ Q.T
        H = 0.5 * (H + H.T)
        full.append((H, lam))
        blocks.append(np.triu(H))
    for mode in (nk.NRGK_MODE_MIXED, nk.NRGK_MODE_PLAIN64):
        out = nk.eigh_batched(blocks, mode=mode)
        for (H, lam), (w, U) in zip(full, out):
            assert np.abs(w - lam).max() <= 1e-10
            assert np.abs(H @ U - U * w).max() <= 1e-10


def test_autotuner_sets_and_caches_the_crossover(nk, tmp_path, monkeypatch):
    """nrgk_tune measures both paths on widths 192..768, picks the widest at which the batched Jacobi is
    not slower, writes the choice and the measured times to its cache, and a context created later
    reads it"""
    path = str(tmp_path / "tune.txt")
    ctx = nk.Context(0)
    cross, table = ctx.tune(path)
    assert cross in [int(w) for w in table[:, 0]]
    assert np.all(table[:, 2] > 0) and np.all(table[:, 3] > 0)
    faster = [int(w) for w, _, tj, th in table if tj <= th]
    assert cross == (max(faster) if faster else int(table[0, 0]))
    assert ctx.crossover() == (cross, True)
    text = open(path).read()
    assert "crossover %d" % cross in text
    monkeypatch.setenv("NRGK_TUNE_CACHE", path)
    assert nk.Context(0).crossover() == (cross, True)
    # a block just above and one just below the tuned crossover, solved with it
    rng = np.random.default_rng(3)
    blocks, lams = [], []
    for n in (cross, cross + 17):
        lam = np.sort(rng.uniform(-1, 1, n))
        Q = _haar(rng, n)
        H = (Q * lam) @ Q.T
        blocks.append(0.5 * (H + H.T))
        lams.append(lam)
    _check(blocks, lams, nk.eigh_batched(blocks, ctx=nk.Context(0)))
