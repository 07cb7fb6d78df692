"""Checkpoint file header (CPU): the documented layout, written here byte by byte, is read
back by the library's header reader; foreign or truncated files are rejected."""
import struct

import pytest


def write_header(path, V=1000, d=300, epochs=2, words=123456789012, seed=7, alpha=0.0125,
                 rng=(11, 22, 33), magic=b"FW2VCKPT", version=1, truncate=None):
    h = magic + struct.pack("<II", version, 64) + struct.pack("<qiiQQfi", V, d, epochs, words, seed,
                                                              alpha, len(rng))
    h += b"\0" * (64 - len(h))
    body = struct.pack(f"<{len(rng)}Q", *rng)
    data = h + body
    if truncate is not None:
        data = data[:truncate]
    path.write_bytes(data)


def test_header_layout_round_trip(tmp_path, w2vmod):
    p = tmp_path / "c.ckpt"
    write_header(p)
    m = w2vmod.checkpoint_header(p)
    assert m == {"vocab_size": 1000, "layer_size": 300, "epochs_done": 2,
                 "words_done": 123456789012, "seed": 7, "alpha": pytest.approx(0.0125),
                 "n_rng": 3}


@pytest.mark.parametrize("kw", [dict(magic=b"FW2VCKPX"), dict(version=2), dict(truncate=40),
                                dict(V=0)])
def test_header_rejects_foreign_files(tmp_path, w2vmod, kw):
    p = tmp_path / "bad.ckpt"
    write_header(p, **kw)
    with pytest.raises(w2vmod.W2VError) as e:
        w2vmod.checkpoint_header(p)
    assert e.value.code == -1


def test_missing_file(tmp_path, w2vmod):
    with pytest.raises(w2vmod.W2VError):
        w2vmod.checkpoint_header(tmp_path / "absent.ckpt")
