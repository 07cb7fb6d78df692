"""The full-w2v command line: reference flags, file formats, exit codes, stdout contract.

Expected vocabulary and corpus-cache contents are hand-traced from the reference's
routines (LearnVocabFromTrainFile, SortVocab, createGlobalStream, saveCorpusCache:
improved_driver_functions.cpp:136-164,244-318, legacy_driver_functions.cpp:74-95).
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "full-w2v_amd", "full-w2v")
TEXT = "the cat sat on the mat\nthe dog sat\r\non the log the end\n"


def run(args, cwd, check=None):
    r = subprocess.run([CLI] + args + ["-debug", "0"], cwd=cwd, capture_output=True, text=True, timeout=600)
    if check is not None:
        assert r.returncode == check, r.stdout + r.stderr
    return r


@pytest.fixture
def corpus_dir(tmp_path):
    (tmp_path / "t.txt").write_text(TEXT)
    return tmp_path


def test_vocab_and_cache_formats(corpus_dir):
    run(["-train", "t.txt", "-min-count", "2", "-save-vocab", "v.txt", "-save-corpus-cache", "c.txt",
         "-save-corpus", "s.txt", "-cache-eos"], corpus_dir, check=0)
    # the final </s> reaches end of file and is not counted; ties keep first-seen order
    assert (corpus_dir / "v.txt").read_text() == "</s> 2\nthe 5\nsat 2\non 2\n"
    # header: sentences, train_words (= 11 counted + 1)
    assert (corpus_dir / "c.txt").read_text() == "3 12\n1 2 3 1 \n1 2 \n3 1 1 \n"
    assert (corpus_dir / "s.txt").read_text() == "the sat on the \nthe sat \non the the \n"


def test_msl_split_without_eos(corpus_dir):
    run(["-train", "t.txt", "-min-count", "1", "-save-corpus-cache", "c.txt", "-msl", "4"], corpus_dir, check=0)
    lines = (corpus_dir / "c.txt").read_text().splitlines()
    # 14 known non-</s> tokens in chunks of 4
    assert lines[0].split()[0] == "4"
    assert [len(x.split()) for x in lines[1:]] == [4, 4, 4, 2]


def test_read_vocab_and_cache_round_trip(corpus_dir):
    run(["-train", "t.txt", "-min-count", "1", "-save-vocab", "v.txt", "-save-corpus-cache", "c.txt",
         "-cache-eos"], corpus_dir, check=0)
    run(["-read-vocab", "v.txt", "-read-corpus-cache", "c.txt", "-cache-eos", "-min-count", "1", "-save-vocab", "v2.txt",
         "-save-corpus-cache", "c2.txt"], corpus_dir, check=0)
    assert (corpus_dir / "v.txt").read_text() == (corpus_dir / "v2.txt").read_text()
    a = (corpus_dir / "c.txt").read_text().splitlines()
    b = (corpus_dir / "c2.txt").read_text().splitlines()
    assert a[1:] == b[1:]


@pytest.mark.parametrize("args,msg", [
    (["-train", "missing.txt"], "Training data not found"),
    (["-read-vocab", "missing.txt"], "not found"),
    (["-read-vocab", "v.txt"], "No corpus cache or training file"),
    (["-output", "o.txt"], "No vocabulary or training file"),
    (["-train", "t.txt", "-read-corpus-cache", "missing.txt"], "not found"),
])
def test_error_exits(corpus_dir, args, msg):
    (corpus_dir / "v.txt").write_text("</s> 1\nthe 5\n")
    r = run(args, corpus_dir)
    assert r.returncode == 1 and msg in r.stdout + r.stderr


def test_help_exits_zero(corpus_dir):
    r = subprocess.run([CLI, "-help"], capture_output=True, text=True)
    assert r.returncode == 0 and "-kernel-batch-size" in r.stdout and "-precision" in r.stdout


def tricky_text(path, n_lines=4000, seed=5):
    """CRLF and LF lines, tabs and runs of spaces, empty lines, an over-long word, rare
    words below -min-count, and no newline at the end of the file."""
    rng = np.random.default_rng(seed)
    p = 1.0 / np.arange(1, 401) ** 1.1
    p /= p.sum()
    lines = []
    for i in range(n_lines):
        ws = [f"w{j}" for j in rng.choice(400, int(rng.integers(0, 25)), p=p)]
        if i % 97 == 0:
            ws.append("x" * 1200)
        if i % 13 == 0:
            ws.append(f"rare{i}")
        sep = "\t" if i % 7 == 0 else ("   " if i % 5 == 0 else " ")
        lines.append(sep.join(ws) + ("\r" if i % 3 == 0 else ""))
    path.write_bytes(("\n".join(lines)).encode())


@pytest.mark.parametrize("eos", [True, False])
def test_parallel_ingest_equals_sequential(tmp_path, eos):
    """Vocabulary and corpus stream built on 1 and on 5 threads (chunks of whole lines)
    are byte-identical: same words, counts, order and sentences."""
    tricky_text(tmp_path / "c.txt")
    out = {}
    for t in (1, 5):
        args = ["-train", "c.txt", "-save-vocab", f"v{t}.txt", "-save-corpus-cache", f"k{t}.txt",
                "-min-count", "3", "-msl", "7", "-ingest-threads", str(t)] + (["-cache-eos"] if eos else [])
        run(args, tmp_path, check=0)
        out[t] = ((tmp_path / f"v{t}.txt").read_bytes(), (tmp_path / f"k{t}.txt").read_bytes())
    assert len(out[1][0]) > 100 and len(out[1][1]) > 1000
    assert out[1] == out[5]


def words_text(tmp_path, n_lines=3000, seed=0):
    rng = np.random.default_rng(seed)
    vocab = [f"w{i}" for i in range(300)]
    p = 1.0 / np.arange(1, 301)
    p /= p.sum()
    lines = []
    for _ in range(n_lines):
        L = int(rng.integers(3, 30))
        lines.append(" ".join(vocab[i] for i in rng.choice(300, L, p=p)))
    (tmp_path / "big.txt").write_text("\n".join(lines) + "\n")


def parse_contract(out):
    wps = float(re.search(r"Overall Words Per Second: ([0-9.]+)", out).group(1))
    words = int(re.search(r"Overall Words: ([0-9]+)", out).group(1))
    secs = float(re.search(r"Overall time: ([0-9.]+)", out).group(1))
    return wps, words, secs


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [
    ["-precision", "fp32"],
    ["-precision", "bf16"],
    ["-precision", "fp32", "-device-sampling", "1"],
    ["-precision", "bf16", "-device-sampling", "1"],
    ["-precision", "bf16", "-device-sampling", "1", "-hot-rows", "0"],
    ["-precision", "fp32", "-hot-rows", "50"],
    ["-precision", "bf16", "-device-sampling", "1", "-piece-length", "8"],
])
def test_cli_trains_and_exports(tmp_path, extra):
    words_text(tmp_path)
    r = run(["-train", "big.txt", "-output", "vec.txt", "-size", "64", "-window", "3", "-negative", "4",
             "-epochs", "2", "-threads", "4", "-streams", "2", "-kernel-batch-size", "100",
             "-cache-eos", "-min-count", "1", "-sample", "1e-3"] + extra, tmp_path, check=0)
    wps, words, secs = parse_contract(r.stdout)
    assert words > 0 and wps > 0 and secs > 0
    txt = (tmp_path / "vec.txt").read_text().splitlines()
    V, d = map(int, txt[0].split())
    assert d == 64 and len(txt) == V + 1
    first = txt[1].split()
    assert first[0] == "</s>" and len(first) == 65
    vals = np.array([[float(x) for x in line.split()[1:]] for line in txt[1:]])
    assert np.isfinite(vals).all() and np.abs(vals).max() > 0


@pytest.mark.gpu
def test_cli_binary_export(tmp_path):
    words_text(tmp_path, 500)
    run(["-train", "big.txt", "-output", "vec.bin", "-binary", "1", "-size", "32", "-epochs", "1",
         "-threads", "2", "-min-count", "1"], tmp_path, check=0)
    raw = (tmp_path / "vec.bin").read_bytes()
    head, rest = raw.split(b"\n", 1)
    V, d = map(int, head.split())
    pos = 0
    for _ in range(V):
        sp = rest.index(b" ", pos)
        vec = np.frombuffer(rest[sp + 1:sp + 1 + 4 * d], np.float32)
        assert np.isfinite(vec).all()
        pos = sp + 1 + 4 * d + 1
    assert pos == len(rest)


@pytest.mark.gpu
def test_cli_reference_sweep_streams(tmp_path):
    """The reference's replication sweep runs -threads 40 -streams 4 (160 streams)."""
    words_text(tmp_path, 2000)
    r = run(["-train", "big.txt", "-output", "vec.txt", "-size", "64", "-window", "3", "-negative", "4",
             "-iter", "1", "-threads", "40", "-streams", "4", "-min-count", "1", "-sample", "1e-3",
             "-precision", "bf16"], tmp_path, check=0)
    wps, words, secs = parse_contract(r.stdout)
    assert words > 0 and wps > 0
    txt = (tmp_path / "vec.txt").read_text().splitlines()
    assert len(txt) == int(txt[0].split()[0]) + 1


@pytest.mark.gpu
@pytest.mark.parametrize("sampling", ["0", "1"])
def test_cli_trains_from_binary_corpus_cache(tmp_path, sampling):
    """Training from -read-corpus-bin reads the same corpus as training from the text:
    the same word count (the reference's 'Overall Words'), and a finite export."""
    words_text(tmp_path, 2000)
    common = ["-size", "32", "-window", "3", "-negative", "4", "-epochs", "2", "-threads", "2",
              "-cache-eos", "-min-count", "1", "-sample", "1e-3", "-device-sampling", sampling]
    r1 = run(["-train", "big.txt", "-save-vocab", "v.txt", "-save-corpus-bin", "c.bin", "-output", "a.txt"]
             + common, tmp_path, check=0)
    r2 = run(["-read-vocab", "v.txt", "-read-corpus-bin", "c.bin", "-output", "b.txt"] + common, tmp_path,
             check=0)
    assert parse_contract(r1.stdout)[1] == parse_contract(r2.stdout)[1] > 0
    txt = (tmp_path / "b.txt").read_text().splitlines()
    assert len(txt) == int(txt[0].split()[0]) + 1
    vals = np.array([[float(x) for x in line.split()[1:]] for line in txt[1:]])
    assert np.isfinite(vals).all() and np.abs(vals).max() > 0


def test_binary_corpus_cache_round_trip(tmp_path):
    """-save-corpus-bin then -read-corpus-bin gives the same corpus (compared through the
    reference's text cache); other -msl / -cache-eos values and a damaged file are refused."""
    tricky_text(tmp_path / "c.txt")
    base = ["-min-count", "3", "-msl", "7", "-cache-eos"]
    run(["-train", "c.txt", "-save-vocab", "v.txt", "-save-corpus-cache", "k1.txt", "-save-corpus-bin", "b.bin"]
        + base, tmp_path, check=0)
    run(["-read-vocab", "v.txt", "-read-corpus-bin", "b.bin", "-save-corpus-cache", "k2.txt"] + base, tmp_path,
        check=0)
    assert (tmp_path / "k1.txt").read_bytes() == (tmp_path / "k2.txt").read_bytes()
    r = run(["-read-vocab", "v.txt", "-read-corpus-bin", "b.bin", "-min-count", "3", "-msl", "9", "-cache-eos"],
            tmp_path, check=1)
    assert "built with -msl 7" in r.stdout
    r = run(["-read-vocab", "v.txt", "-read-corpus-bin", "b.bin", "-min-count", "3", "-msl", "7"], tmp_path, check=1)
    assert "-cache-eos 1" in r.stdout
    raw = (tmp_path / "b.bin").read_bytes()
    (tmp_path / "t.bin").write_bytes(raw[:-5])
    r = run(["-read-vocab", "v.txt", "-read-corpus-bin", "t.bin"] + base, tmp_path, check=1)
    assert "size does not match" in r.stdout
