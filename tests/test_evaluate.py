"""Embedding-quality evaluation (word similarity, analogies) — CPU.

Pins: scipy.stats.spearmanr on random data (ties included); analogies on a constructed
embedding whose 3CosAdd answers are known by construction; a hand-computed 3CosMul case;
round trip of the text and binary export formats written by the CLI's format.
"""
import numpy as np
import pytest
from scipy.stats import spearmanr

from w2v import evaluate as ev


def test_spearman_matches_scipy():
    rng = np.random.default_rng(0)
    for n in (5, 50, 353):
        x = rng.normal(size=n)
        y = np.round(x + rng.normal(size=n), 1)  # ties
        assert ev.spearman(x, y) == pytest.approx(spearmanr(x, y)[0], abs=1e-12)


def write_vectors(path, words, mat, binary):
    with open(path, "wb") as f:
        f.write(f"{len(words)} {mat.shape[1]}\n".encode())
        for w, v in zip(words, mat):
            f.write(w.encode() + b" ")
            if binary:
                f.write(v.astype(np.float32).tobytes())
            else:
                f.write("".join(f"{x:f} " for x in v).encode())
            f.write(b"\n")


@pytest.mark.parametrize("binary", [False, True])
def test_load_vectors_round_trip(tmp_path, binary):
    rng = np.random.default_rng(1)
    words = ["</s>", "a", "bb", "ccc"]
    mat = rng.normal(size=(4, 7)).astype(np.float32)
    p = tmp_path / "v.txt"
    write_vectors(p, words, mat, binary)
    w2, m2 = ev.load_vectors(p)
    assert w2 == words
    assert np.allclose(m2, mat, atol=1e-6 if not binary else 0)


def test_similarity_eval_oov_is_zero():
    emb = ev.Embedding(["x", "y", "z"], np.array([[1.0, 0], [1.0, 0.1], [0, 1.0]]))
    assert emb.similarity("x", "nope") == 0.0
    pairs = [("x", "y", 9.0), ("x", "z", 1.0), ("y", "z", 2.0)]
    assert ev.similarity_eval(emb, pairs) == pytest.approx(1.0)


def test_analogy_constructed():
    """Words are offsets a_i + r_j of random directions: king - man + woman = queen holds
    exactly, so 3CosAdd must be perfect."""
    rng = np.random.default_rng(2)
    d = 64
    bases = rng.normal(size=(30, d))
    rel = rng.normal(size=(2, d)) * 0.7
    words, vecs = [], []
    for i in range(30):
        for j in range(2):
            words.append(f"w{i}_{j}")
            vecs.append(bases[i] + rel[j])
    emb = ev.Embedding(words, np.array(vecs))
    qs = [(f"w{i}_0", f"w{i}_1", f"w{k}_0", f"w{k}_1") for i in range(10) for k in range(10, 20)]
    add, mul = ev.analogy_eval(emb, qs)
    assert add == 1.0 and mul > 0.9


def test_analogy_mul_hand_case():
    # unit vectors in 2-D at the given angles; with s -> (s+1)/2:
    #   3CosAdd: good -0.535+0.770+0.939 = 1.174 > bad -0.005+0.100+0.292 = 0.386
    #   3CosMul: good 0.770*0.939/0.545 = 1.33 < bad 0.100*0.292/0.015 = 1.94
    ang = {"a": 0.0, "a_": 0.5, "b": 1.0, "good": 1.5, "bad": 3.0}
    words = list(ang)
    mat = np.array([[np.cos(t), np.sin(t)] for t in ang.values()])
    emb = ev.Embedding(words, mat)
    add, mul = ev.analogy_eval(emb, [("a", "a_", "b", "good")])
    assert add == 1.0 and mul == 0.0
    add, mul = ev.analogy_eval(emb, [("a", "a_", "b", "bad")])
    assert add == 0.0 and mul == 1.0
