"""Host-side tables of the library (InitNet restatement) against the oracle (CPU)."""
import numpy as np
import pytest


@pytest.mark.parametrize("corrected", [False, True])
def test_unigram_table_equals_oracle(w2vmod, orc, corrected):
    rng = np.random.default_rng(0)
    cn = np.concatenate([[500], rng.zipf(1.3, 3000)]).astype(np.uint64)
    n = 1_000_003
    a = w2vmod.unigram_table(cn, n, corrected)
    b = orc.unigram_table(cn, n, corrected)
    assert np.array_equal(a, b)


def test_subsample_equals_oracle(w2vmod, orc):
    cn = np.array([7, 1, 2, 50, 10**5, 10**7, 3], np.uint64)
    for sample in (1e-3, 1e-4, 1e-5):
        assert np.array_equal(w2vmod.subsample_probs(cn, sample, 17_000_000),
                              orc.subsample_probs(cn, sample, 17_000_000))


def test_synth_cluster_labels_equal_oracle(w2vmod, orc):
    import ctypes

    cfg = orc.synth_cfg(V=5000, clusters=64, seed=7)
    words = np.arange(1, 5000)
    a = w2vmod.synth_cluster_of(7, 64, words)
    b = np.array([orc.lib().orc_synth_cluster_of(ctypes.byref(cfg), int(w)) for w in words])
    assert np.array_equal(a, b)
    counts = np.bincount(a, minlength=64)
    assert counts.min() > 40 and counts.max() < 120  # about uniform (78 expected)
