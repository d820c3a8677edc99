"""GPU parity of CSV ingest (vdf_read_csv): column types inferred from the
first 10 rows, every field against Python's own conversion (float() is
correctly rounded; int()), quoted delimiters, CRLF line ends, an
unterminated last line, wide-exponent doubles, a bad row, and served tables
written to CSV, read back and grouped against the oracle on the original
columns."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from visondf_amd import _native as N
from visondf_amd import datagen

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mixed(n: int, seed: int):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2) + rng.integers(0, 2, n).astype(np.uint64)
    i = rng.integers(-2**62, 2**62, n).astype(np.int64)
    i[1] = -5  # a '-' among the first 10 rows: an INT column
    bits = rng.integers(0, 2**64, n, dtype=np.uint64) & np.uint64(0x7FEFFFFFFFFFFFFF)
    d = bits.view(np.float64) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    d[:5] = [0.5, -0.0, 1e-310, 2.2250738585072014e-308, 1.7976931348623157e308]
    b = rng.integers(0, 2, n)
    short = ["".join(chr(97 + x) for x in rng.integers(0, 26, rng.integers(1, 12))) for _ in range(n)]
    long_ = ["'a,b'" if k % 97 == 0 else "x" * int(rng.integers(17, 40)) for k in range(n)]
    rows = []
    for k in range(n):
        rows.append(",".join([str(int(u[k])), str(int(i[k])), repr(float(d[k])), str(int(b[k])), short[k],
                              long_[k]]))
    return (u, i, d, b, short, long_), rows


def test_read_csv_mixed_types(vdf):
    n = 20_011
    (u, i, d, b, short, long_), rows = _mixed(n, 3)
    text = "u,i,d,b,c,s\r\n" + "\r\n".join(rows)  # CRLF, no newline after the last row
    cols = vdf.read_csv(text.encode())
    assert [c[0] for c in cols] == ["u", "i", "d", "b", "c", "s"]
    kinds = [c[1] for c in cols]
    assert kinds == [N.CSV_UINT, N.CSV_INT, N.CSV_DOUBLE, N.CSV_BOOL, N.CSV_CHAR, N.CSV_STRING], kinds
    got_u = cols[0][2].cpu().numpy().view(np.uint64)
    assert np.array_equal(got_u, u)
    assert np.array_equal(cols[1][2].cpu().numpy(), i)
    got_d = cols[2][2].cpu().numpy()
    exp_d = np.array([float(repr(float(x))) for x in d])
    assert np.array_equal(got_d.view(np.uint64), exp_d.view(np.uint64))
    assert np.array_equal(cols[3][2].cpu().numpy(), b.astype(np.int32))
    c = cols[4][2].cpu().numpy()
    assert c.shape == (n, 16)
    exp_c = np.array([list((s + "0" * 16)[:16].encode()) for s in short], dtype=np.uint8)
    assert np.array_equal(c, exp_c)
    s = cols[5][2].cpu().numpy()
    w = max(len(x) for x in long_)
    assert s.shape == (n, w)
    exp_s = np.zeros((n, w), dtype=np.uint8)
    for k, x in enumerate(long_):
        exp_s[k, :len(x)] = list(x.encode())
    assert np.array_equal(s, exp_s)


def test_read_csv_numbers_and_prefixes(vdf):
    """Doubles of every shape (17 digits, long digit strings, extreme
    exponents that take the wide conversion, subnormals, overflow) and
    integer prefixes / out-of-range values, against Python's conversions."""
    rng = np.random.default_rng(9)
    vals = ["0.5"]
    for k in range(30_000):
        nd = int(rng.integers(1, 45))
        digs = "".join(str(x) for x in rng.integers(0, 10, nd))
        e = int(rng.integers(-340, 320))
        vals.append(f"{digs[:1]}.{digs[1:]}e{e}" if nd > 1 else f"{digs}e{e}")
    vals += ["1e400", "-1e400", "1e-400", "4.9406564584124654e-324", "2.4703282292062328e-324", "inf", "-nan",
             "12.5abc", "abc", "-"]
    text = "x\n" + "\n".join(vals) + "\n"
    cols = vdf.read_csv(text.encode())
    assert cols[0][1] == N.CSV_DOUBLE
    got = cols[0][2].cpu().numpy()

    def py(v):
        try:
            return float(v)
        except ValueError:
            import re
            m = re.match(r"-?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?", v)
            return float(m.group(0)) if m else 0.0

    exp = np.array([py(v) for v in vals])
    both_nan = np.isnan(got) & np.isnan(exp)
    assert np.array_equal(got.view(np.uint64)[~both_nan], exp.view(np.uint64)[~both_nan])
    head = [str(v) for v in range(-5, 5)]  # the first 10 rows decide the type
    ints = ["7", "-3", "9223372036854775807", "-9223372036854775808", "9223372036854775808", "12x", "x", "-"]
    cols = vdf.read_csv(("k\n" + "\n".join(head + ints) + "\n").encode())
    assert cols[0][1] == N.CSV_INT
    assert cols[0][2].cpu().numpy().tolist() == list(range(-5, 5)) + [7, -3, 2**63 - 1, -2**63, 0, 12, 0, 0]
    cols = vdf.read_csv(b"a;b\n1;2\n3;4\n", delim=";", int_bytes=4)
    assert [c[1] for c in cols] == [N.CSV_UINT, N.CSV_UINT]
    assert cols[0][2].dtype.itemsize == 4 and cols[1][2].cpu().numpy().tolist() == [2, 4]


def test_read_csv_no_header_and_errors(vdf):
    cols = vdf.read_csv(b"1,a\n2,b\n", header=False)
    assert [c[0] for c in cols] == ["", ""] and cols[0][2].cpu().numpy().tolist() == [1, 2]
    with pytest.raises(N.VdfError) as e:
        vdf.read_csv(b"a,b\n1,2\n3\n4,5\n")
    assert e.value.status == N.VDF_ERR_INVALID
    assert vdf.read_csv(b"") == []


@pytest.mark.parametrize("kind", ["u64", "char16"])
def test_read_csv_served_table_groupby(vdf, kind):
    """A served table (2M rows of uint64 or 16-byte ASCII keys and float64
    values) written as tab-separated text, read on the device and grouped:
    the same groups, counts, min and max as the oracle on the original
    columns, sums within 1e-12."""
    n = 2_000_003
    k, x = datagen.served_table(n, 7, kind)
    _say("csv served", kind, n)
    if kind == "u64":
        keys_txt = [str(int(v)) for v in k]
    else:
        keys_txt = [bytes(r).decode() for r in k]
    text = "key\tval\n" + "\n".join(f"{a}\t{float(b)!r}" for a, b in zip(keys_txt, x)) + "\n"
    cols = vdf.read_csv(text.encode(), delim="\t", quote="\x7f")
    assert cols[0][1] == (N.CSV_UINT if kind == "u64" else N.CSV_CHAR)
    dk, dv = cols[0][2], cols[1][2]
    assert np.array_equal(dv.cpu().numpy(), x)
    r = vdf.groupby(dk, [dv], key_type=N.KEY_U64 if kind == "u64" else N.KEY_CHAR16)
    o = oracle.groupby(k, [x], threads=16)
    assert r.n_groups == o["n_groups"]
    gk = r.keys.cpu().numpy()
    assert np.array_equal(gk.view(np.uint64) if kind == "u64" else gk, o["keys"])
    assert np.array_equal(r.counts.cpu().numpy().view(np.uint32).astype(np.uint64), o["counts"])
    assert np.array_equal(r.aggs[0]["min"].cpu().numpy(), o["aggs"][0]["min"])
    assert np.array_equal(r.aggs[0]["max"].cpu().numpy(), o["aggs"][0]["max"])
    s = r.aggs[0]["sum"].cpu().numpy()
    assert np.all(np.abs(s - o["aggs"][0]["sum"]) <= 1e-12 * np.abs(o["aggs"][0]["sum"]))


def _vs_oracle(cols, o):
    assert [c[0] for c in cols] == o["names"]
    assert [c[1] for c in cols] == o["kinds"]
    for (name, kind, x), exp, w in zip(cols, o["cols"], o["widths"]):
        got = x.cpu().numpy()
        if kind in (N.CSV_STRING, N.CSV_CHAR):
            assert got.shape[1] == w, name
            assert np.array_equal(got, exp), name
        else:
            assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), name


def test_read_csv_airline_vs_oracle(vdf):
    """Airline-shaped text (21 columns, quoted names holding the delimiter,
    two-decimal numbers; 400K rows, 82 MB) and the mixed-type rows above:
    every column bit-identical to the oracle's CSV ingest."""
    t = datagen.csv_airline(400_000, seed=11)
    _say("csv airline", t.size)
    cols = vdf.read_csv(bytes(t), quote='"', char_width=4)  # 2- and 3-character codes: char columns
    _vs_oracle(cols, oracle.read_csv(t, quote='"', char_width=4, threads=16))
    assert [c[1] for c in cols][:6] == [N.CSV_UINT, N.CSV_UINT, N.CSV_CHAR, N.CSV_STRING, N.CSV_CHAR, N.CSV_STRING]
    _, rows = _mixed(5_003, 5)
    text = ("u,i,d,b,c,s\n" + "\n".join(rows) + "\n").encode()
    _vs_oracle(vdf.read_csv(text, char_width=16), oracle.read_csv(text, char_width=16, threads=4))


def test_read_csv_field_edges_vs_oracle(vdf):
    """Quoted fields holding the delimiter at either end, empty and quoted
    empty fields, and rows of different lengths inside one wave's rows:
    equal to the oracle; a row with an extra field and one with a missing
    field are rejected."""
    rows = []
    for k in range(300):
        rows.append(f'"{k},",{"" if k % 7 == 0 else k},"x{k % 3}"')
        if k % 5 == 0:
            rows.append(f'",{k}",{k * 1000},""')
    text = ("a,b,c\n" + "\n".join(rows) + "\n").encode()
    _vs_oracle(vdf.read_csv(text, quote='"', char_width=0), oracle.read_csv(text, quote='"', threads=2))
    for bad in (b"a,b\n1,2\n3,4,5\n", b"a,b\n1,2\n" + b"3,4\n" * 100 + b"5\n"):
        with pytest.raises(N.VdfError) as e:
            vdf.read_csv(bad)
        assert e.value.status == N.VDF_ERR_INVALID
