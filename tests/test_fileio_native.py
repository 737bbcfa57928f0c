"""The native `.vbr` / `.vbrc` text reader and writer and the binary
container (CPU only).

Parity targets are the reference's own I/O (utils/fileio.py:13-93): each
token parsed with Python's float() / int(), each value written with str().
The reader is checked against exactly that, token by token, on values chosen
to exercise the decimal conversion (17-digit values, subnormals, the largest
and smallest doubles, signed zero, inf, nan), at thread counts that cut the
lines at different commas; the writer is checked character for character
against str()."""
import os

import numpy as np
import pytest

from oracle.fileio_oracle import parse_line_reference
from sable_amd import fileio
from sable_amd.vbr import VBR

_lib = pytest.importorskip("sable_amd._lib")


def tricky_values(n=20000, seed=0):
    rng = np.random.default_rng(seed)
    special = np.array([0.0, -0.0, 1.0, -1.0, 0.1, 0.2, 0.3, 1e-5, 1e-4, 0.0001, 0.00011, 9.999e-5, 1e15,
                        1e16, 1e17, 123456789012345678.0, 1e22, 1e23, 5e-324, -5e-324, 2.2250738585072014e-308,
                        2.2250738585072009e-308, 1.7976931348623157e308, -1.7976931348623157e308,
                        np.inf, -np.inf, np.nan, 100000.0, 1.5, 2.5e-7, 0.5, 3.0e10, 4.35, 0.1 + 0.2])
    mant = rng.uniform(-1, 1, n)
    expo = rng.integers(-1074, 1024, n)
    wide = np.ldexp(mant, expo)
    normal = rng.standard_normal(n)
    ints = rng.integers(-10 ** 6, 10 ** 6, n).astype(np.float64)
    return np.concatenate([special, wide, normal, ints, normal.astype(np.float32).astype(np.float64)])


def python_line(name, a):
    """The reference writer's line: str() of each element."""
    return f"{name}=[{','.join(map(str, a.tolist()))}]\n"


def python_parse(line, conv):
    """The reference reader's parse of one line (utils/fileio.py:56-63)."""
    return parse_line_reference(line, conv)


def same_bits(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_writer_matches_python_str(tmp_path):
    vals = tricky_values()
    ints = np.concatenate([np.array([0, -1, 1, 2 ** 63 - 1, -2 ** 63], dtype=np.int64),
                           np.random.default_rng(1).integers(-2 ** 62, 2 ** 62, 5000)])
    p = tmp_path / "t.vbrc"
    fileio._write_lines(str(p), [("val", vals), ("indx", ints), ("empty", np.zeros(0))], nthreads=4)
    lines = open(p).read().splitlines(keepends=True)
    assert lines[0] == python_line("val", vals)
    assert lines[1] == python_line("indx", ints)
    assert lines[2] == "empty=[]\n"


@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_reader_matches_python_parse(tmp_path, nthreads):
    vals = tricky_values(seed=2)
    ints = np.random.default_rng(3).integers(-2 ** 62, 2 ** 62, 7001)
    p = tmp_path / "t.txt"
    with open(p, "w") as f:
        f.write(python_line("val", vals))
        f.write(python_line("ix", ints))
    got = fileio._read_lines_native(_lib, str(p), [np.float64, np.int64], nthreads)
    lines = open(p).readlines()
    assert same_bits(got[0], python_parse(lines[0], float))
    assert got[1].tolist() == python_parse(lines[1], int)


def test_reader_chunks_a_large_line(tmp_path):
    """A line of several MiB is cut into per-thread chunks at commas; every
    cut position gives the same array as the serial parse."""
    vals = np.random.default_rng(4).standard_normal(600_000) * 10.0 ** np.random.default_rng(5).integers(-30, 30, 600_000)
    p = tmp_path / "big.txt"
    with open(p, "w") as f:
        f.write(python_line("val", vals))
    ref = np.array(open(p).read().split("=")[1][1:-2].split(","), dtype=np.float64)
    for nt in (1, 2, 5, 7, 16):
        got = fileio._read_lines_native(_lib, str(p), [np.float64], nt)[0]
        assert same_bits(got, ref), nt
    assert same_bits(ref, vals)


def test_golden_files_round_trip_text_identically(golden, tmp_path):
    src = os.path.join(golden, "example.vbr")
    arrs = fileio.read_vbr(src)
    lines = open(src).readlines()
    assert same_bits(arrs[0], python_parse(lines[0], float))
    for a, line in zip(arrs[1:], lines[1:]):
        assert a.tolist() == python_parse(line, int)
    fileio.write_vbr_matrix("out", VBR(*arrs), str(tmp_path))
    assert open(tmp_path / "out.vbr").read() == open(src).read()
    srcc = os.path.join(golden, "example.vbrc")
    c = fileio.read_vbrc(srcc)
    val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices, csr_val = c
    fileio.write_vbrc(str(tmp_path / "out.vbrc"), val, csr_val, indptr, indices, indx, bindx, rpntr, cpntr,
                      bpntrb, bpntre, ublocks)
    assert open(tmp_path / "out.vbrc").read() == open(srcc).read()


@pytest.mark.parametrize("text,msg", [
    ("val=[1.0,2.0\n", "is not name"),
    ("val=[1.0,abc]\n", "bad token"),
    ("val=[1.0,]\n", "bad token"),
    ("val=[1.0,,2.0]\n", "bad token"),
    ("val 1.0\n", "is not name"),
])
def test_reader_rejects_malformed_lines(tmp_path, text, msg):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(ValueError, match=msg):
        fileio._read_lines_native(_lib, str(p), [np.float64], 2)


def test_reader_rejects_missing_lines_and_accepts_empty(tmp_path):
    p = tmp_path / "short.vbr"
    p.write_text("val=[]\nindx=[0]\n")
    with pytest.raises(ValueError, match="lines"):
        fileio.read_vbr(str(p))
    a, b = fileio._read_lines_native(_lib, str(p), [np.float64, np.int64], 4)
    assert len(a) == 0 and b.tolist() == [0]


def test_binary_container_round_trip(golden, tmp_path):
    arrs = fileio.read_vbr(os.path.join(golden, "example.vbr"))
    v = VBR(*arrs)
    for mm in (True, False):
        fileio.save_vbr_bin(str(tmp_path / "m.bin"), v)
        w = fileio.load_vbr_bin(str(tmp_path / "m.bin"), mmap=mm)
        for n in ("val", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre"):
            assert np.array_equal(np.asarray(getattr(w, n)), np.asarray(getattr(v, n)))
            assert np.asarray(getattr(w, n)).dtype == np.asarray(getattr(v, n)).dtype
    c = fileio.read_vbrc(os.path.join(golden, "example.vbrc"))
    val, indx, bindx, rpntr, cpntr, bpntrb, bpntre, ublocks, indptr, indices, csr_val = c
    fileio.write_vbrc_bin(str(tmp_path / "c.bin"), val, csr_val, indptr, indices, indx, bindx, rpntr, cpntr,
                          bpntrb, bpntre, ublocks)
    d = fileio.read_vbrc_bin(str(tmp_path / "c.bin"))
    assert all(np.array_equal(x, y) for x, y in zip(c, d))
    # float32 and int32 keep their type; the header is checked
    fileio.save_arrays_bin(str(tmp_path / "t.bin"), 7, [("a", np.arange(5, dtype=np.float32)),
                                                        ("b", np.arange(3, dtype=np.int32)), ("e", [])])
    k, t = fileio.load_arrays_bin(str(tmp_path / "t.bin"))
    assert k == 7 and t["a"].dtype == np.float32 and t["b"].dtype == np.int32 and len(t["e"]) == 0
    (tmp_path / "x.bin").write_bytes(b"NOTSABLE" + bytes(56))
    with pytest.raises(ValueError, match="container"):
        fileio.load_arrays_bin(str(tmp_path / "x.bin"))
    with pytest.raises(ValueError, match="VBR container"):
        fileio.load_vbr_bin(str(tmp_path / "c.bin"))


def test_text_and_binary_agree_on_a_generated_matrix(tmp_path):
    from sable_amd import synth
    vbr, _ = synth.vbr_matrix_gen_seeded(400, 380, 20, 20, 60, seed=8)
    v64 = VBR(np.asarray(vbr.val, dtype=np.float64), *(np.asarray(getattr(vbr, n)) for n in
                                                        ("indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre")))
    fileio.write_vbr_matrix("g", v64, str(tmp_path))
    t = fileio.read_vbr(str(tmp_path / "g.vbr"))
    fileio.save_vbr_bin(str(tmp_path / "g.bin"), v64)
    b = fileio.load_vbr_bin(str(tmp_path / "g.bin"))
    for x, n in zip(t, ("val", "indx", "bindx", "rpntr", "cpntr", "bpntrb", "bpntre")):
        assert np.array_equal(x, np.asarray(getattr(b, n)))
