"""Independent FP64 closed forms used to pin the oracle (and the GPU path).

They follow the reference tests' own formulation, not the oracle's code:
- cosine matrices built entry by entry and the orthonormal Haar matrix built by
  the recursion H(2n) = [H(n) x (1, 1); I(n) x (1, -1)] / sqrt(2)
  (tests/test_bm3d_contract.cpp:20-59, tests/test_bm3d_semantics.py:137-159);
- the 3-D transform as matrix products over whole axes (einsum), where the
  oracle runs in-place butterfly and line transforms.
"""
import numpy as np


def dct_matrix(n):
    k = np.arange(n)[:, None]
    x = np.arange(n)[None, :]
    return np.cos(np.pi * (x + .5) * k / n) * np.sqrt(np.where(k == 0, 1., 2.) / n)


def haar_matrix(n):
    h = np.ones((1, 1))
    while h.shape[0] < n:
        m = h.shape[0]
        h = np.vstack([np.kron(h, [1, 1]), np.kron(np.eye(m), [1, -1])]) / np.sqrt(2)
    return h


def forward(cube):
    """3-D coefficients of a (group, block, block) cube: Haar along the group,
    DCT along y and x (the separable form of kron(H, D, D))."""
    g, b, _ = cube.shape
    d, h = dct_matrix(b), haar_matrix(g)
    return np.einsum("kg,yv,xu,gvu->kyx", h, d, d, cube, optimize=True)


def inverse(coef):
    g, b, _ = coef.shape
    d, h = dct_matrix(b), haar_matrix(g)
    return np.einsum("kg,yv,xu,kyx->gvu", h, d, d, coef, optimize=True)


def group_filter(cube, k, sigma, wiener=False, ref_cube=None):
    """(filtered cube, weight) of the BM3D group filter in FP64: the first k
    patches real, the rest zero (tests/test_bm3d_contract.cpp:54-59)."""
    g, b, _ = cube.shape
    v = np.zeros((g, b, b))
    v[:k] = cube[:k]
    c = forward(v)
    if wiener:
        r = np.zeros((g, b, b))
        r[:k] = ref_cube[:k]
        rc = forward(r)
        den = rc * rc + sigma * sigma
        with np.errstate(invalid="ignore", divide="ignore"):
            w = np.where(den > 0, rc * rc / den, 0.)
        w[0, 0, 0] = 1.
        c = c * w
        weight = 1. / max(float(np.sum(w * w)), 1e-12)
    else:
        keep = np.abs(c) >= 2.7 * sigma
        keep[0, 0, 0] = True
        c = c * keep
        weight = 1. / max(int(np.sum(keep)), 1)
    return inverse(c), weight


def single_block(p, ref, group, sigma, wiener):
    """Whole-pipeline output on an image of exactly one block with
    block_step = block: one reference, one match, weight cancels
    (tests/test_bm3d_semantics.py:148-159)."""
    b = p.shape[0]
    cube = np.zeros((group, b, b))
    cube[0] = p
    rc = np.zeros((group, b, b))
    rc[0] = ref
    out, _ = group_filter(cube, 1, sigma, wiener, rc)
    return out[0]


def brute_force_matches(plane, cx, cy, block, rng, group, ts=None):
    """Scalar FP64 SSD over the clamped window, stable sort by (dist, y, x)
    (tests/test_matcher.cpp:51-78). Returns [(x, y, dist)] with the reference
    first."""
    h, w = plane.shape
    max_x, max_y = w - block, h - block
    cx, cy = min(max(cx, 0), max_x), min(max(cy, 0), max_y)
    ref = plane[cy:cy + block, cx:cx + block].astype(np.float64)
    cands = []
    for y in range(max(0, cy - rng), min(max_y, cy + rng) + 1):
        for x in range(max(0, cx - rng), min(max_x, cx + rng) + 1):
            if x == cx and y == cy:
                continue
            d = float(np.sum((ref - plane[y:y + block, x:x + block].astype(np.float64)) ** 2))
            cands.append((d, y, x))
    cands.sort()
    return [(cx, cy, 0.0)] + [(x, y, d) for d, y, x in cands[:group - 1]]


def temporal_single_block(frames, group, sigma, route_to_center=False):
    """Rolling radius-1 output of a clip whose frames are each exactly one
    block (block_step = block): centre c's group is its own patch (slot 1),
    then the patches of frames c - 1 (slot 0) and c + 1 (slot 2) that exist,
    in (distance, slot) order (src/cpu/bm/matcher.hpp:27-57, :319-373); every
    filtered patch goes into the sums of the frame it came from, with the
    centre's weight, and frame n is the ratio of its sums over the centres
    n - 1 .. n + 1 (src/host/filter_bm3d.cpp:730-851). One hard-threshold
    stage. Returns the output frames and each centre's group as frame indices.
    route_to_center sums every patch into the centre's frame instead (a
    routing mistake, for checking that a test would catch one)."""
    n = len(frames)
    b = frames[0].shape[0]
    num = [np.zeros((b, b)) for _ in range(n)]
    den = [0.0] * n
    groups = []
    for c in range(n):
        members = [(0.0, 1, c)]
        for t, f in ((0, c - 1), (2, c + 1)):
            if 0 <= f < n:
                members.append((float(np.sum((frames[c] - frames[f]) ** 2)), t, f))
        members = [members[0]] + sorted(members[1:])
        members = members[:group]
        cube = np.zeros((group, b, b))
        for j, (_, _, f) in enumerate(members):
            cube[j] = frames[f]
        out, w = group_filter(cube, len(members), sigma)
        for j, (_, _, f) in enumerate(members):
            dst = c if route_to_center else f
            num[dst] += w * out[j]
            den[dst] += w
        groups.append([f for _, _, f in members])
    return [num[f] / den[f] for f in range(n)], groups
