"""Checkpoint/resume on the GPU: training epoch 1 from a checkpoint taken after epoch 0
gives the same model, bit for bit, as training both epochs in one run (deterministic
mode, device sampling: every draw is keyed by seed, epoch and corpus position). The file's
matrices are the downloaded model in the documented layout."""
import os

import numpy as np
import pytest

from test_cli import CLI, parse_contract, run, words_text

pytestmark = pytest.mark.gpu


def setup(w2vmod, kernel, storage, data):
    V, d, W, K, tokens, off, uni, subs, T = data
    tr = w2vmod.Trainer(V, d, W, K, msl=64, storage=storage, kernel=kernel, sigmoid="table",
                        mode="deterministic", seed=5)
    tr.tables_upload(uni, subs)
    tr.set_schedule(0.025, 2, T)
    tr.corpus_upload(tokens, off)
    return tr


@pytest.fixture(scope="module")
def data(orc):
    V, d, W, K = 300, 96, 3, 4
    rng = np.random.default_rng(8)
    lens = rng.integers(1, 40, 60)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    zipf = 1.0 / np.arange(1, V + 1)
    tokens = rng.choice(V, off[-1], p=zipf / zipf.sum()).astype(np.int32)
    cn = np.bincount(tokens, minlength=V).astype(np.uint64) + 1
    T = int(off[-1])
    return V, d, W, K, tokens, off, orc.unigram_table(cn, 10007), orc.subsample_probs(cn, 1e-2, T), T


@pytest.mark.parametrize("kernel,storage", [("plain", "fp32"), ("mfma", "bf16")])
def test_resume_equals_uninterrupted_run(tmp_path, w2vmod, data, kernel, storage):
    n = len(data[5]) - 1
    a = setup(w2vmod, kernel, storage, data)
    a.init_model()
    a.train_device(0, 0, n)
    a.train_device(1, 0, n)
    ref0, ref1 = a.download()
    a.close()

    path = tmp_path / "e1.ckpt"
    b = setup(w2vmod, kernel, storage, data)
    b.init_model()
    b.train_device(0, 0, n)
    words = b.stats()["words_read"]
    b.checkpoint_save(path, 1, words, 0.0125, rng=[3, 4])
    mid0, mid1 = b.download()
    b.close()

    # file layout: 64-byte header, n_rng u64 states, syn0 then syn1neg as FP32 V x d
    V, d = data[0], data[1]
    raw = np.fromfile(path, np.uint8)
    f0 = raw[64 + 16:64 + 16 + 4 * V * d].view(np.float32).reshape(V, d)
    f1 = raw[64 + 16 + 4 * V * d:].view(np.float32).reshape(V, d)
    assert np.array_equal(f0, mid0) and np.array_equal(f1, mid1)
    assert np.array_equal(raw[64:80].view(np.uint64), [3, 4])

    c = setup(w2vmod, kernel, storage, data)
    meta, states = c.checkpoint_load(path, rng_capacity=4)
    assert meta["epochs_done"] == 1 and meta["words_done"] == words and meta["seed"] == 5
    assert list(states) == [3, 4]
    c.train_device(1, 0, n)
    got0, got1 = c.download()
    c.close()
    assert np.array_equal(got0, ref0) and np.array_equal(got1, ref1)


def test_load_rejects_other_sizes(tmp_path, w2vmod, data):
    b = setup(w2vmod, "plain", "fp32", data)
    b.init_model()
    b.checkpoint_save(tmp_path / "x.ckpt", 0, 0, 0.025)
    b.close()
    V, d, W, K = data[:4]
    other = w2vmod.Trainer(V + 1, d, W, K)
    with pytest.raises(w2vmod.W2VError) as e:
        other.checkpoint_load(tmp_path / "x.ckpt")
    other.close()
    assert e.value.code == -1


@pytest.mark.parametrize("extra", [["-precision", "bf16"], ["-precision", "fp32", "-device-sampling", "1"]])
def test_cli_checkpoint_and_resume(tmp_path, w2vmod, extra):
    words_text(tmp_path)
    common = ["-train", "big.txt", "-size", "64", "-window", "3", "-negative", "4", "-threads", "4",
              "-streams", "2", "-cache-eos", "-min-count", "1", "-sample", "1e-3"] + extra
    r = run(common + ["-output", "a.txt", "-epochs", "2", "-save-checkpoint", "ck"], tmp_path, check=0)
    _, words2, _ = parse_contract(r.stdout)
    m = w2vmod.checkpoint_header(tmp_path / "ck")
    assert m["epochs_done"] == 2 and m["layer_size"] == 64
    assert m["n_rng"] == (0 if "-device-sampling" in extra else 4)
    r = run(common + ["-output", "b.txt", "-epochs", "3", "-resume", "ck"], tmp_path, check=0)
    _, words1, _ = parse_contract(r.stdout)
    assert words1 * 2 == words2  # one more epoch of the same corpus
    assert (tmp_path / "b.txt").exists()
    run(common + ["-output", "c.txt", "-epochs", "2", "-resume", "ck"], tmp_path, check=1)
