"""Pins of the WNNM oracle (oracle/oracle.py wnnm_*), on the CPU.

- The reference's zero-rank plugin test (tests/test_wnnm_zero_rank_plugin.py:
  sigma 255 with residual 1 removes the centred spectrum, so every patch
  becomes its group's row mean; known answer within 2e-6).
- sigma 0: every singular value is kept whole, so the output is the input.
- A constant plane comes back unchanged at any sigma (rank one with
  residual 0, the first value always kept; zero after centring with
  residual 1).
- sv_shrink on hand-computed spectra.
- A 4 x 2 group built from a chosen spectrum, shrunk by hand.
- The group shrink against a second route through the Gram eigenproblem
  (numpy eigh instead of the SVD) on random groups, m < n and m > n.
"""
import numpy as np
import pytest

import oracle as orc


def _zero_rank_case(block, width):
    height = block
    values = np.tile(0.25 + (np.arange(width, dtype=np.float32) - (width - 1) / 2) / 1024,
                     (height, 1)).astype(np.float32)
    patches = [values[:, x:x + block].astype(np.float64) for x in range(width - block + 1)]
    mean = np.mean(patches, axis=0)
    num = np.zeros_like(values, dtype=np.float64)
    den = np.zeros_like(num)
    for x in range(len(patches)):
        num[:, x:x + block] += mean
        den[:, x:x + block] += 1
    return values, num / den, len(patches)


ZERO_RANK = ((1, 2, 2), (4, 7, 4), (8, 15, 8))


@pytest.mark.parametrize("block,width,group", ZERO_RANK)
@pytest.mark.parametrize("adaptive", (0, 1))
def test_zero_rank_known_answer(block, width, group, adaptive):
    values, expected, npatch = _zero_rank_case(block, width)
    assert npatch == group
    out = orc.wnnm_plane(values, 255.0, block=block, step=block, group=group, bm_range=16, residual=1,
                         adaptive=adaptive)
    assert np.isfinite(out).all()
    assert np.max(np.abs(out - expected)) <= 2e-6


def _noise(h, w, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    clean = 0.5 + 0.3 * np.sin(xx / 7.0) * np.cos(yy / 5.0)
    return (clean + rng.normal(0, 0.05, (h, w))).astype(np.float32)


@pytest.mark.parametrize("block,step,group,residual", [(8, 8, 8, 0), (4, 2, 16, 1), (8, 4, 32, 0), (2, 1, 4, 1)])
def test_sigma_zero_is_identity(block, step, group, residual):
    src = _noise(40, 36, 3)
    out = orc.wnnm_plane(src, 0.0, block=block, step=step, group=group, bm_range=5, residual=residual)
    assert np.max(np.abs(out - src)) <= 1e-6


@pytest.mark.parametrize("residual", (0, 1))
def test_constant_plane_unchanged(residual):
    src = np.full((24, 32), 0.375, np.float32)
    out = orc.wnnm_plane(src, 200.0, block=8, step=4, group=8, bm_range=4, residual=residual)
    assert np.max(np.abs(out - src)) <= 1e-6


def test_sv_shrink_hand_values():
    # 25 - 9 = 16: (5 + 4) / 2 = 4.5; 9 - 9 = 0 stops the walk.
    s, k = orc.sv_shrink([5.0, 3.0, 1.0], 9.0, 0)
    assert k == 1 and np.allclose(s, [4.5, 0.0, 0.0], rtol=0, atol=1e-15)
    # start 1 keeps the first value whole.
    s, k = orc.sv_shrink([5.0, 3.0, 1.0], 9.0, 1)
    assert k == 1 and np.allclose(s, [5.0, 0.0, 0.0], rtol=0, atol=1e-15)
    # 169 - 25 = 144: (13 + 12) / 2 = 12.5; 100 - 25 = 75 > 0: (10 + sqrt 75) / 2.
    s, k = orc.sv_shrink([13.0, 10.0, 5.0], 25.0, 0)
    assert k == 2 and np.allclose(s, [12.5, (10 + 75 ** 0.5) / 2, 0.0], rtol=0, atol=1e-12)
    # A value below the constant stops everything after it.
    s, k = orc.sv_shrink([4.0, 1.0, 0.5], 20.0, 0)
    assert k == 0 and np.all(s == 0)


def test_group_hand_spectrum():
    # A = U diag(5, 3) V^T, U two orthonormal columns of length 4, V a
    # rotation. sigma^2 = 1/2 and n = 2 give c = 8 sqrt(4) / 2 = 8: the first
    # value stays (start 1), the second becomes (3 + sqrt(9 - 8)) / 2 = 2.
    U = np.array([[1, 1], [1, -1], [1, 1], [1, -1]], np.float64) / 2.0
    th = np.pi / 6
    V = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    A = U @ np.diag([5.0, 3.0]) @ V.T
    X, w = orc.wnnm_group(A, np.sqrt(0.5), residual=0, adaptive=1)
    assert np.allclose(X, U @ np.diag([5.0, 2.0]) @ V.T, rtol=0, atol=1e-12)
    assert w == 0.5
    # Non-adaptive weight is 1; with a constant above 9 only the first stays.
    X, w = orc.wnnm_group(A, 1.0, residual=0, adaptive=0)
    assert np.allclose(X, U @ np.diag([5.0, 0.0]) @ V.T, rtol=0, atol=1e-12) and w == 1.0


def _gram_route(A, sigma, residual, adaptive):
    """Second route: eigenpairs of the Gram matrix, gains by place."""
    A = np.array(A, np.float64)
    m, n = A.shape
    mean = A.mean(axis=1, keepdims=True) if residual else np.zeros((m, 1))
    C = A - mean
    lam, V = np.linalg.eigh(C.T @ C)
    s = np.sqrt(np.maximum(lam, 0.0))
    order = sorted(range(n), key=lambda j: (-s[j], j))
    constant = 8.0 * np.sqrt(2.0 * n) * sigma * sigma
    start, rank = (0 if residual else 1), min(m, n)
    gain = np.zeros(n)
    kept = 0
    for place, j in enumerate(order):
        if place < start:
            gain[j] = 1.0
        elif place < rank and s[j] * s[j] - constant > 0:
            gain[j] = (s[j] + np.sqrt(s[j] * s[j] - constant)) / 2 / s[j]
        else:
            break
        kept += 1
    X = C @ (V * gain) @ V.T + mean
    return X, (1.0 / kept if adaptive and kept > 0 else 1.0)


@pytest.mark.parametrize("m,n,residual,adaptive", [(64, 8, 0, 1), (64, 8, 1, 1), (16, 32, 0, 0), (4, 8, 1, 1),
                                                   (256, 16, 0, 1)])
def test_group_matches_gram_route(m, n, residual, adaptive):
    rng = np.random.default_rng(m * 100 + n)
    base = rng.normal(0.5, 0.2, (m, 1)) @ np.ones((1, n))
    A = base + rng.normal(0, 0.04, (m, n)) + 0.1 * rng.normal(size=(m, 2)) @ rng.normal(size=(2, n))
    sigma = 6.0 / 255
    X1, w1 = orc.wnnm_group(A, sigma, residual, adaptive)
    X2, w2 = _gram_route(A, sigma, residual, adaptive)
    assert w1 == w2
    assert np.max(np.abs(X1 - X2)) < 1e-9


def test_match_plane_agrees_with_single_reference_matcher():
    src = _noise(30, 34, 9)
    xy, counts = orc.match_plane(src, 4, 3, 5, 8)
    refs = orc.raster_grid(34, 30, 4, 3)
    assert len(refs) == len(counts)
    for r in (0, 7, len(refs) - 1):
        one = orc.spatial_match(src, refs[r][0], refs[r][1], 4, 5, 8)
        assert counts[r] == len(one)
        assert [tuple(p) for p in xy[r, :counts[r]]] == [(x, y) for x, y, _, _ in one]
