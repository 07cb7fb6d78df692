"""Graph-walk corpus: the oracle's Barabasi-Albert generator and random walks pinned by
structure and statistics, and the library's host generator equal to the oracle (CPU).

What pins the oracle here (the reference has no graph corpus, so nothing to replay):
  * BA: the closed-form edge count m(m+1)/2 + (n-m-1)m, degree sum 2E, minimum degree m,
    no self loops or repeated edges, symmetric adjacency, and preferential attachment
    itself: the complementary degree distribution of a BA graph falls as k^-2
    (P(k) ~ k^-3), and early nodes collect far more edges than late ones. Attaching to
    uniformly drawn nodes instead gives an exponential tail and fails the slope check.
  * walks: every step follows an edge, every node starts exactly R walks (walk w starts
    at node w mod n), the truncated share and the length distribution (uniform on
    [lmin, L]), untruncated walks have length L, and next steps are uniform over the
    neighbors (chi-square).
"""
import numpy as np
import pytest
from scipy import stats


def edges_of(off, adj):
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off))
    return src, adj


@pytest.mark.parametrize("n,m,seed", [(2000, 1, 3), (20000, 5, 7), (5000, 3, 11)])
def test_ba_structure(orc, n, m, seed):
    off, adj = orc.ba_graph(n, m, seed)
    E = m * (m + 1) // 2 + (n - m - 1) * m
    deg = np.diff(off)
    assert len(adj) == 2 * E and deg.sum() == 2 * E
    assert deg.min() >= m
    src, dst = edges_of(off, adj)
    assert not np.any(src == dst), "self loop"
    key = src.astype(np.int64) * n + dst
    assert len(np.unique(key)) == len(key), "repeated edge"
    rev = dst.astype(np.int64) * n + src
    assert np.array_equal(np.sort(key), np.sort(rev)), "adjacency not symmetric"
    # node v > m attaches to m earlier nodes
    late = src > m
    assert np.all(np.bincount(src[late & (dst < src)], minlength=n)[m + 1:] == m)


def test_ba_preferential_attachment(orc):
    n, m = 200_000, 5
    off, _ = orc.ba_graph(n, m, 5)
    deg = np.diff(off)
    ks = np.arange(20, 201)
    ccdf = np.array([(deg >= k).mean() for k in ks])
    slope = np.polyfit(np.log(ks), np.log(ccdf), 1)[0]
    assert -2.4 < slope < -1.6, slope  # BA: P(K >= k) ~ k^-2
    assert deg.max() > 500  # grows like m sqrt(n)
    assert deg[:100].mean() > 10 * deg[-n // 10:].mean()
    # uniform attachment (same sizes) has an exponential tail: its fit is far steeper
    rng = np.random.default_rng(0)
    udeg = np.zeros(n, np.int64)
    udeg[: m + 1] = m
    for v in range(m + 1, n):
        t = rng.choice(v, m, replace=False) if v < 2000 else rng.integers(0, v, m)
        udeg[t] += 1
        udeg[v] += m
    uccdf = np.array([max((udeg >= k).mean(), 1e-9) for k in ks])
    assert np.polyfit(np.log(ks), np.log(uccdf), 1)[0] < -4.0


@pytest.mark.parametrize("n,m,seed", [(3000, 2, 1), (20000, 5, 9)])
def test_library_ba_equals_oracle(orc, w2vmod, n, m, seed):
    off, adj = orc.ba_graph(n, m, seed)
    loff, ladj = w2vmod.graph_ba(n, m, seed)
    assert np.array_equal(off, loff) and np.array_equal(adj, ladj)


def test_ba_rejects_bad_sizes(w2vmod):
    for n, m in [(5, 5), (3, 0), (10, -1)]:
        with pytest.raises(w2vmod.W2VError):
            w2vmod.graph_ba(n, m, 1)


def test_walks_properties(orc):
    n, R, L, lmin, frac = 3000, 10, 80, 5, 0.3
    off, adj = orc.ba_graph(n, 4, 2)
    tok, so = orc.walks(off, adj, R, L, lmin, frac, 13)
    nw = n * R
    assert len(so) == nw + 1 and so[-1] == len(tok)
    lens = np.diff(so)
    assert np.array_equal(tok[so[:-1]], np.arange(nw) % n)  # walk w starts at node w mod n
    # every step follows an edge
    src, dst = edges_of(off, adj)
    edges = set((src.astype(np.int64) * n + dst).tolist())
    same = np.ones(len(tok), bool)
    same[so[:-1]] = False  # first token of each walk has no predecessor in it
    a = tok[:-1][same[1:]].astype(np.int64)
    b = tok[1:][same[1:]]
    assert all(int(x) in edges for x in a * n + b)
    # truncation share and lengths
    trunc = lens < L
    p = trunc.mean()
    # walks truncated to exactly L are indistinguishable; expected share frac*(L-lmin)/(L-lmin+1)
    expect = frac * (L - lmin) / (L - lmin + 1)
    assert abs(p - expect) < 4 * np.sqrt(expect * (1 - expect) / nw)
    assert lens.min() >= lmin and lens.max() == L
    tl = lens[trunc]
    assert stats.kstest((tl - lmin + 0.5) / (L - lmin), "uniform").pvalue > 1e-4
    # next steps from the highest-degree node are uniform over its neighbors
    hub = int(np.argmax(np.diff(off)))
    nxt = b[a == hub]
    nb = adj[off[hub]:off[hub + 1]]
    counts = np.array([(nxt == v).sum() for v in nb])
    assert counts.sum() > 20 * len(nb)
    assert stats.chisquare(counts).pvalue > 1e-4
