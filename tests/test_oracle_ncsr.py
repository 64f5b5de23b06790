"""Pins of the NCSR oracle (oracle/oracle.py ncsr_*), on the CPU.

- sigma 0 is the exact identity, and a constant plane comes back unchanged
  (the centred group is zero: no code survives, the row mean remains).
- The reference's own NCSR plugin test (tests/test_ncsr_variance_plugin.py):
  a noiseless stroke image at sigma 0.1 moves no pixel by 0.5/255 or more,
  and with bm_range 1 (at most nine matches) groups of 9, 16 and 32 give the
  same output (the second moment counts actual columns only).
- A two-patch group in closed form: one direction d = (a - b) / 2 with codes
  +-|d|, weights 1 and exp(-|a - b|^2 / h), shrunk by hand.
- The group filter against a second route through the Gram eigenproblem.
"""
import numpy as np
import pytest

import oracle as orc


def _stroke_image(size=32):
    y, x = np.mgrid[:size, :size]
    base = (0.12 + ((x * 17 + y * 31) % 97) / 400).astype(np.float32)
    base[((x % 11 == 3) & (y > 3)) | ((y == 9) & (x > 5))] = 0.9
    return base


def test_sigma_zero_identity_and_constant_plane():
    src = _stroke_image() + np.float32(0.01) * np.float32(np.sin(np.arange(32)))[None, :]
    src = src.astype(np.float32)
    assert np.array_equal(orc.ncsr_plane(src, 0.0), src)
    flat = np.full((24, 32), 0.25, np.float32)
    assert np.max(np.abs(orc.ncsr_plane(flat, 50.0, step=4) - flat)) <= 1e-7


def test_low_sigma_keeps_strokes():
    src = _stroke_image()
    out = orc.ncsr_plane(src, 0.1)
    assert np.isfinite(out).all()
    assert float(np.max(np.abs(out - src))) < 0.5 / 255


@pytest.mark.parametrize("block", (1, 4, 8))
def test_padded_group_capacity_changes_nothing(block):
    src = _stroke_image()
    kw = dict(block=block, step=block, bm_range=1, iters=1)
    ref = orc.ncsr_plane(src, 10.0, group=9, **kw)
    for capacity in (16, 32):
        out = orc.ncsr_plane(src, 10.0, group=capacity, **kw)
        np.testing.assert_allclose(out, ref, rtol=2e-4, atol=2e-5)


def test_two_patch_closed_form():
    a = np.array([0.50, 0.52, 0.49, 0.51])
    b = np.array([0.47, 0.505, 0.50, 0.45])
    sigma = 0.01
    m = 4
    mu = (a + b) / 2
    d = (a - b) / 2
    nd = np.linalg.norm(d)
    u = d / nd
    dist = float(np.sum((a - b) ** 2))
    h = 2 * m * sigma * sigma
    w1 = np.exp(-dist / h)
    codes = np.array([nd, -nd])
    beta = (codes[0] + w1 * codes[1]) / (1 + w1)
    moment = np.sum((codes - beta) ** 2)
    tau = 2 * np.sqrt(2) * sigma ** 2 / (np.sqrt(moment / 2) + 1e-12)
    e = codes - beta
    shrunk = np.sign(e) * np.maximum(np.abs(e) - tau, 0) + beta
    expected = np.stack([mu + u * shrunk[0], mu + u * shrunk[1]], axis=1)
    got = orc.ncsr_group(np.stack([a, b], axis=1), [0.0, dist], sigma)
    assert 0 < tau < nd  # the threshold bites but leaves part of the code
    np.testing.assert_allclose(got, expected, rtol=0, atol=1e-12)


def _gram_route(A, dist, sigma):
    A = np.array(A, np.float64)
    m, n = A.shape
    mean = A.mean(axis=1, keepdims=True)
    C = A - mean
    lam, V = np.linalg.eigh(C.T @ C)
    s = np.sqrt(np.maximum(lam, 0))
    order = sorted(range(n), key=lambda j: (-s[j], j))
    h = max(2 * m * sigma * sigma, 1e-12)
    w = np.exp(-np.concatenate([[0.0], np.asarray(dist[1:n], np.float64)]) / h)
    M = np.zeros((n, n))
    for place, i in enumerate(order):
        if place >= min(m, n) or not s[i] > 1e-9:
            continue
        code = s[i] * V[:, i]
        beta = np.dot(w, code) / w.sum()
        tau = 2 * np.sqrt(2) * sigma ** 2 / (np.sqrt(np.sum((code - beta) ** 2) / n) + 1e-12)
        e = code - beta
        M += np.outer(V[:, i], (np.sign(e) * np.maximum(np.abs(e) - tau, 0) + beta) / s[i])
    return C @ M + mean


@pytest.mark.parametrize("m,n", [(64, 8), (16, 32), (4, 8), (256, 16)])
def test_group_matches_gram_route(m, n):
    rng = np.random.default_rng(m + n)
    A = rng.normal(0.5, 0.05, (m, 1)) + rng.normal(0, 0.04, (m, n)) + 0.2 * rng.normal(size=(m, 2)) @ rng.normal(
        size=(2, n))
    dist = np.concatenate([[0.0], np.sort(rng.uniform(0.01, 0.3, n - 1))])
    got = orc.ncsr_group(A, dist, 8 / 255)
    np.testing.assert_allclose(got, _gram_route(A, dist, 8 / 255), rtol=0, atol=1e-9)
