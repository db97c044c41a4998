"""Graph input layer: seeded generators and file formats (host only)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import katric_amd as K
from tests.conftest import ROOT, golden


def canonical(g):
    off = g.offsets.astype(np.int64)
    assert off[0] == 0 and (np.diff(off) >= 0).all() and off[-1] == len(g.heads)
    tails = np.repeat(np.arange(g.n), np.diff(off))
    h = g.heads.astype(np.int64)
    assert (h < g.n).all()
    assert (h != tails).all(), "self loop"
    key = tails * g.n + h
    assert (np.diff(key) > 0).all(), "lists ascending without duplicates"
    rev = np.sort(h * g.n + tails)
    np.testing.assert_array_equal(rev, key)  # symmetric


def test_gnm_canonical_and_deterministic():
    a = K.gen_gnm(1000, 20000, seed=3)
    canonical(a)
    b = K.gen_gnm(1000, 20000, seed=3)
    assert a.checksum() == b.checksum()
    c = K.gen_gnm(1000, 20000, seed=4)
    assert a.checksum() != c.checksum()
    # duplicates and self-loops removed: close to the sample count
    assert 19000 < a.m <= 20000


def test_gnm_config1_size():
    g = K.gen_gnm(1 << 16, 1 << 20, seed=1)
    canonical(g)
    assert abs(g.m - (1 << 20)) < 1000  # ~255 pairs collide or loop in expectation


@pytest.mark.parametrize("permute", [True, False])
def test_rmat_canonical(permute):
    g = K.gen_rmat(12, 16, seed=2, permute=permute)
    canonical(g)
    deg = np.diff(g.offsets)
    assert deg.max() > 20 * deg.mean()  # heavy tail


def test_rmat_id_limit():
    g = K.gen_rmat(11, samples=30000, seed=5, n=1500)
    canonical(g)
    assert g.n == 1500


def test_generator_thread_count_independent():
    code = ("import katric_amd as K; print(K.gen_rmat(13, 16, seed=9).checksum(), "
            "K.gen_gnm(5000, 60000, seed=9).checksum())")
    outs = []
    for t in ("1", "5"):
        env = dict(os.environ, KATRIC_THREADS=t, PYTHONPATH=ROOT)
        outs.append(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                                   check=True).stdout)
    assert outs[0] == outs[1]


def test_metis_roundtrip(tmp_path):
    g = K.gen_rmat(9, 8, seed=1)
    p = str(tmp_path / "g.metis")
    K.write_metis(p, g)
    h = K.read_metis(p)
    np.testing.assert_array_equal(g.offsets, h.offsets)
    np.testing.assert_array_equal(g.heads, h.heads)


def test_binary_roundtrip(tmp_path):
    g = K.read_metis(golden("hiv.metis"))
    fo, hd = str(tmp_path / "g.first_out"), str(tmp_path / "g.head")
    K.write_graphio_binary(fo, hd, g)
    h = K.read_graphio_binary(fo, hd)
    np.testing.assert_array_equal(g.offsets, h.offsets)
    np.testing.assert_array_equal(g.heads, h.heads)
    ref = K.read_graphio_binary(golden("hiv.first_out"), golden("hiv.head"))
    np.testing.assert_array_equal(ref.offsets, g.offsets)
    np.testing.assert_array_equal(ref.heads, g.heads)


def test_metis_errors(tmp_path):
    p = tmp_path / "bad.metis"
    p.write_text("3 2\n2\n9\n2\n")
    with pytest.raises(K.KatricError):
        K.read_metis(str(p))
    with pytest.raises(K.KatricError):
        K.read_metis(str(tmp_path / "missing.metis"))


def test_metis_oversized_header(tmp_path):
    # a header promising 10^17 edges must not size an allocation: the file
    # holds one triangle and reads as such
    p = tmp_path / "big_m.metis"
    p.write_text("3 99999999999999999\n2 3\n1 3\n1 2\n")
    g = K.read_metis(str(p))
    assert g.n == 3 and g.nnz == 6


def test_checksum_sensitive():
    g = K.read_metis(golden("hiv.metis"))
    h = K.Graph(g.n, g.offsets, g.heads.copy())
    h.heads[0] ^= 1
    assert g.checksum() != h.checksum()
