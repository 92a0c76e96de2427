"""Pins the oracle's f32 / integer LBVH leg (bh_gravity.wgsl + gpu_gravity.rs:351-388) on the CPU.

Each check is one a plausible restatement mistake would fail: closed-form Morton values, an independent
numpy Morton, brute-force comparison sorts, a top-down radix-tree builder written from the definition
(not from the shader's search), root-moment closed forms, and the reference's own published accuracy of
its LBVH walk on its own seed (docs/37-gpu-barnes-hut-findings.md:21-22).
"""
import numpy as np
import pytest

from oracle import oracle

from helpers import rms_rel, splitmix_cloud


def test_expand_bits_closed_forms():
    assert oracle.lib().orc_expand_bits(0x3FF) == 0x09249249
    assert oracle.lib().orc_expand_bits(0) == 0
    assert oracle.lib().orc_expand_bits(1) == 1
    assert oracle.lib().orc_expand_bits(2) == 8
    for v in [5, 123, 1023, 0x2AA]:
        want = sum(((v >> b) & 1) << (3 * b) for b in range(10))
        assert oracle.lib().orc_expand_bits(v) == want


def test_bbox_key_is_order_preserving_and_lossless():
    vals = np.array([-3.4e38, -1.0, -1e-30, -0.0, 0.0, 1e-30, 1.0, 2.5e6, 3.4e38], np.float32)
    keys = [oracle.lib().orc_bbox_key(float(v)) for v in vals]
    assert keys == sorted(keys)
    for v in vals:
        back = oracle.lib().orc_bbox_decode(oracle.lib().orc_bbox_key(float(v)))
        assert np.float32(back).view(np.uint32) == np.float32(v).view(np.uint32)


def _numpy_morton(pos, lo, hi):
    p = pos.astype(np.float32)
    ext = np.maximum(hi - lo, np.float32(1e-30)).astype(np.float32)
    u = np.clip((p - lo) / ext, 0, 1).astype(np.float32)
    q = np.clip(np.floor(u * np.float32(1024)), 0, 1023).astype(np.uint32)
    code = np.zeros(len(p), np.uint64)
    for b in range(10):
        for a, shift in ((0, 2), (1, 1), (2, 0)):
            code |= ((q[:, a].astype(np.uint64) >> b) & 1) << np.uint64(3 * b + shift)
    return code.astype(np.uint32)


def test_morton_matches_independent_interleave_and_corners():
    pos, _ = splitmix_cloud(1500, 0xABCD1234)
    p = pos.astype(np.float32)
    lo, hi = oracle.bbox_f32(p)
    assert np.array_equal(lo, p.min(0)) and np.array_equal(hi, p.max(0))
    codes = oracle.morton_f32(p, lo, hi)
    assert np.array_equal(codes, _numpy_morton(p, lo, hi))
    # the bbox corners map to 0 and to the all-max code; coincident points share a code
    corners = np.array([lo, hi, p[3], p[3]], np.float32)
    c = oracle.morton_f32(corners, lo, hi)
    assert c[0] == 0 and c[1] == 0x3FFFFFFF and c[2] == c[3]
    # a flat axis (all equal) does not divide by zero (bh_gravity.wgsl:105)
    flat = p.copy()
    flat[:, 2] = 7.0
    lo2, hi2 = oracle.bbox_f32(flat)
    assert np.all(oracle.morton_f32(flat, lo2, hi2) < (1 << 30))


def test_sort_pairs_is_the_code_index_order():
    rng = np.random.default_rng(7)
    codes = rng.integers(0, 64, 5000).astype(np.uint32)  # many duplicates
    sc, order = oracle.sort_pairs(codes)
    want = np.lexsort((np.arange(5000), codes))
    assert np.array_equal(order, want)
    assert np.array_equal(sc, codes[want])
    assert np.all(np.diff(sc.astype(np.int64)) >= 0)


def _topdown_radix_tree(keys):
    """Binary radix tree from the definition: split each range at the highest bit where its first and last
    (code, index) keys differ; node index = split for the left internal child, split + 1 for the right
    (Karras 2012 numbering, root 0, leaves at L-1+i). Independent of the shader's per-node search."""
    L = len(keys)
    left = np.zeros(max(L - 1, 0), np.int64)
    right = np.zeros(max(L - 1, 0), np.int64)
    parent = np.full(2 * L - 1, 0xFFFFFFFF, np.int64)
    ext = [(int(c) << 32) | i for i, c in enumerate(keys)]
    stack = [(0, 0, L - 1)]
    while stack:
        node, f, l = stack.pop()
        hb = (ext[f] ^ ext[l]).bit_length() - 1
        split = f
        for s in range(f, l):
            if (ext[s + 1] >> hb) & 1:
                split = s
                break
        lc = (L - 1) + split if split == f else split
        rc = (L - 1) + split + 1 if split + 1 == l else split + 1
        left[node], right[node] = lc, rc
        parent[lc] = parent[rc] = node
        if split != f:
            stack.append((split, f, split))
        if split + 1 != l:
            stack.append((split + 1, split + 1, l))
    return left, right, parent


@pytest.mark.parametrize("kind", ["unique", "duplicates", "all_equal"])
def test_karras_equals_topdown_definition(kind):
    rng = np.random.default_rng(11)
    if kind == "unique":
        codes = np.sort(rng.choice(1 << 30, 600, replace=False)).astype(np.uint32)
    elif kind == "duplicates":
        codes = np.sort(rng.integers(0, 40, 600)).astype(np.uint32)
    else:
        codes = np.full(300, 12345, np.uint32)
    l, r, p = oracle.karras(codes)
    tl, tr, tp = _topdown_radix_tree(codes)
    assert np.array_equal(l, tl) and np.array_equal(r, tr) and np.array_equal(p, tp)


def test_reference_split_matches_where_valid_and_breaks_on_equal_runs():
    """The shader's midpoint split for all-equal ranges (bh_gravity.wgsl:159-160) equals the Karras split
    whenever first/last codes differ everywhere, and produces nodes with two parents on runs of equal
    codes — the reason the build uses the index-extended split."""
    rng = np.random.default_rng(3)
    uniq = np.sort(rng.choice(1 << 30, 800, replace=False)).astype(np.uint32)
    a = oracle.karras(uniq)
    b = oracle.karras(uniq, reference_split=True)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    dup = np.sort(rng.integers(0, 10, 800)).astype(np.uint32)
    l, r, _ = oracle.karras(dup, reference_split=True)
    counts = np.bincount(np.concatenate([l, r]), minlength=2 * 800 - 1)
    assert (counts > 1).any()


def test_upsweep_root_closed_forms():
    pos, mass = splitmix_cloud(1500, 0xABCD1234)
    p = pos.astype(np.float32)
    m = mass.astype(np.float32)
    lo, hi = oracle.bbox_f32(p)
    sc, order = oracle.sort_pairs(oracle.morton_f32(p, lo, hi))
    for K in (1, 8):
        lc = oracle.cluster_codes(sc, K)
        l, r, _ = oracle.karras(lc)
        sb = np.concatenate([p[order], m[order, None]], axis=1)
        com, bmin, bmax = oracle.upsweep_f32(sb, K, l, r)
        mt = mass.astype(np.float64).sum()
        c = (p.astype(np.float64) * m[:, None]).sum(0) / m.astype(np.float64).sum()
        assert abs(com[0, 3] - mt) / mt < 1e-4
        assert np.linalg.norm(com[0, :3] - c) / np.linalg.norm(c) < 1e-4
        assert np.array_equal(bmin[0], lo) and np.array_equal(bmax[0], hi)


# The reference's own measurement of its LBVH walk on this seed (docs/37-gpu-barnes-hut-findings.md:21-22):
# theta 0.5 -> RMS 0.70 % (max 6.2 %), 0.25 -> 0.12 %, 0.1 -> 0.009 %, theta -> 0 -> 1.8e-6.
@pytest.mark.parametrize("theta,rms,mx", [(0.5, 0.0070, 0.062), (0.25, 0.0012, None), (0.1, 0.00009, None),
                                          (1e-4, 1.8e-6, None)])
def test_lbvh_walk_reproduces_published_accuracy(theta, rms, mx):
    pos, mass = splitmix_cloud(1500, 0xABCD1234)
    exact = oracle.direct_gravity(pos, mass, 5e3)
    acc = oracle.lbvh_gravity_f32(pos, mass, 5e3, theta)
    got_rms, got_max = rms_rel(acc, exact)
    assert got_rms == pytest.approx(rms, rel=0.1)
    if mx is not None:
        assert got_max == pytest.approx(mx, rel=0.1)


def test_lbvh_accel_is_the_same_leg_and_thread_count_free():
    """The one-call LBVH leg used by the CPU step equals the staged Python chain, and the parallel
    per-target loops give the same bits on 1 and 4 threads (direct sum and octree too)."""
    pos, mass = splitmix_cloud(1500, 0xABCD1234)
    staged = oracle.lbvh_gravity_f32(pos, mass, 5e3, 0.5)
    outs = []
    default = oracle.lib().orc_get_threads()
    for t in (1, 4):
        oracle.set_threads(t)
        outs.append((oracle.lbvh_accel(pos, mass, 5e3, 0.5), oracle.direct_gravity(pos, mass, 5e3),
                     oracle.octree_accel(pos, mass, 0.5, 5e3)))
    oracle.set_threads(default)
    assert np.array_equal(outs[0][0].astype(np.float32), staged)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_cluster_codes_and_every_node_moment_by_brute_force():
    """Leaf cluster code = the bucket's first sorted code (gpu_gravity.rs:382-388); every node's mass, COM
    and AABB equal the brute-force moments of the sorted particles under it (bh_gravity.wgsl:200-241)."""
    pos, mass = splitmix_cloud(700, 0x5A5A5A5A)
    p = pos.astype(np.float32)
    m = mass.astype(np.float32)
    lo, hi = oracle.bbox_f32(p)
    sc, order = oracle.sort_pairs(oracle.morton_f32(p, lo, hi))
    for K in (1, 8):
        lc = oracle.cluster_codes(sc, K)
        assert np.array_equal(lc, sc[::K])
        l, r, _ = oracle.karras(lc)
        L = len(lc)
        sb = np.concatenate([p[order], m[order, None]], axis=1)
        com, bmin, bmax = oracle.upsweep_f32(sb, K, l, r)
        first = np.zeros(2 * L - 1, np.int64)
        last = np.zeros(2 * L - 1, np.int64)
        for c in range(L):
            first[L - 1 + c] = c * K
            last[L - 1 + c] = min(c * K + K, len(m)) - 1
        def span(v):
            if v >= L - 1:
                return first[v], last[v]
            a, _ = span(l[v])
            _, b = span(r[v])
            return a, b
        for v in range(2 * L - 1):
            a, b = span(v)
            q = sb[a:b + 1].astype(np.float64)
            mt = q[:, 3].sum()
            c = (q[:, :3] * q[:, 3:4]).sum(0) / mt
            assert abs(com[v, 3] - mt) <= 1e-5 * mt
            assert np.allclose(com[v, :3], c, rtol=0, atol=1e-5 * np.abs(q[:, :3]).max())
            assert np.array_equal(bmin[v], sb[a:b + 1, :3].min(0)) and np.array_equal(bmax[v], sb[a:b + 1, :3].max(0))


def test_walk_counts_closed_forms():
    """The reference walk's interaction count: with theta -> 0 every node opens, so each target sums the
    other N - 1 particles and no cell; the walk that counts is the walk that computes (its forces equal
    lbvh_gravity_f32's, which the published-accuracy test pins); with a far target and an enormous theta the
    root is accepted at once (one cell)."""
    pos, mass = splitmix_cloud(700, 0x5EED)
    n = len(mass)
    c = oracle.lbvh_walk_counts(pos, mass, 5e3, np.arange(n), theta=1e-9)
    assert np.all(c[:, 0] == 0) and np.all(c[:, 1] == n - 1)
    pos2 = np.vstack([pos, [[1e12, 0.0, 0.0]]])
    mass2 = np.append(mass, 1.0)
    c2 = oracle.lbvh_walk_counts(pos2, mass2, 5e3, [n], theta=1e6)
    assert c2[0, 0] >= 1 and c2[0, 0] + c2[0, 1] <= 64
    half = oracle.lbvh_walk_counts(pos, mass, 5e3, np.arange(n), theta=0.5)
    assert np.all(half.sum(1) < n - 1) and np.all(half[:, 0] > 0)
