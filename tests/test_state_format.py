"""Host-side parse of the serde state layouts (no device): files written by an
independent restatement of bincode 1.x / serde_json (tests/_serde_np.py) are
read back by the library's parser exactly, and malformed files are refused."""
import ctypes as C
import json
import struct

import numpy as np
import pytest

import lqcd
from lqcd import _ffi as F
from _serde_np import (dumps, ef_bincode, ef_obj, state_bincode, state_obj)


def header(path, fmt, with_e, dim):
    size, points, beta, t = C.c_double(), C.c_size_t(), C.c_double(), C.c_uint64()
    rc = F.lib().lqcd_state_read_header(str(path).encode(), fmt, with_e, dim, C.byref(size),
                                        C.byref(points), C.byref(beta), C.byref(t))
    return rc, size.value, points.value, beta.value, t.value


def read_host(path, fmt, with_e, dim, nlink):
    links = np.full((nlink, 18), np.nan)
    e = np.full((nlink, 8), np.nan)
    rc = F.lib().lqcd_state_read_host(str(path).encode(), fmt, with_e, dim, F.ptr(links), nlink,
                                      F.ptr(e) if with_e else None)
    return rc, links, e


def random_state(dim, points, seed=1):
    rng = np.random.default_rng(seed)
    nlink = points ** dim * dim
    return rng.normal(size=(nlink, 18)), rng.normal(size=(nlink, 8)), nlink


@pytest.mark.parametrize("fmt", [F.FORMAT_BINCODE, F.FORMAT_JSON])
@pytest.mark.parametrize("dim,points", [(3, 3), (4, 2), (2, 5)])
def test_state_parse_matches_independent_writer(tmp_path, fmt, dim, points):
    links, _, nlink = random_state(dim, points)
    p = tmp_path / "s"
    if fmt == F.FORMAT_BINCODE:
        p.write_bytes(state_bincode(0.5, points, 6.25, links))
    else:
        p.write_text(dumps(state_obj(0.5, points, 6.25, links)))
    rc, size, pts, beta, _ = header(p, fmt, 0, dim)
    assert rc == F.OK and (size, pts, beta) == (0.5, points, 6.25)
    rc, got, _ = read_host(p, fmt, 0, dim, nlink)
    assert rc == F.OK and np.array_equal(got, links)


@pytest.mark.parametrize("fmt", [F.FORMAT_BINCODE, F.FORMAT_JSON])
def test_efsync_parse_matches_independent_writer(tmp_path, fmt):
    dim, points = 3, 4
    links, e, nlink = random_state(dim, points, 2)
    p = tmp_path / "ef"
    if fmt == F.FORMAT_BINCODE:
        p.write_bytes(ef_bincode(1.0, points, 5.5, links, e, 17, dim))
    else:
        p.write_text(dumps(ef_obj(1.0, points, 5.5, links, e, 17, dim)))
    rc, size, pts, beta, t = header(p, fmt, 1, dim)
    assert rc == F.OK and (size, pts, beta, t) == (1.0, points, 5.5, 17)
    rc, got, got_e = read_host(p, fmt, 1, dim, nlink)
    assert rc == F.OK and np.array_equal(got, links) and np.array_equal(got_e, e)


def test_json_key_order_whitespace_and_unknown_keys(tmp_path):
    dim, points = 2, 3
    links, e, nlink = random_state(dim, points, 3)
    o = ef_obj(0.25, points, 4.0, links, e, 3, dim)
    o2 = {"lattice_state": {"link_matrix": o["lattice_state"]["link_matrix"], "extra": [1, {"a": "x"}],
                            "beta": 4.0, "lattice": {"dim": points, "size": 0.25}},
          "t": 3, "e_field": o["e_field"]}
    p = tmp_path / "j"
    p.write_text(json.dumps(o2, indent=2))
    rc, got, got_e = read_host(p, F.FORMAT_JSON, 1, dim, nlink)
    assert rc == F.OK and np.array_equal(got, links) and np.array_equal(got_e, e)
    # serde_json writes NaN as null; it reads back as NaN here
    links2 = links.copy()
    links2[1, 3] = np.nan
    s = dumps(state_obj(0.25, points, 4.0, links2)).replace("NaN", "null")
    p.write_text(s)
    rc, got, _ = read_host(p, F.FORMAT_JSON, 0, dim, nlink)
    assert rc == F.OK and np.isnan(got[1, 3]) and np.array_equal(np.nan_to_num(got), np.nan_to_num(links2))


def test_malformed_files_are_refused(tmp_path):
    dim, points = 3, 2
    links, _, nlink = random_state(dim, points, 4)
    good = state_bincode(1.0, points, 6.0, links)
    p = tmp_path / "b"
    # wrong D: the link count does not match dim^D * D
    p.write_bytes(good)
    assert header(p, F.FORMAT_BINCODE, 0, 4)[0] == F.ERR_INCOMPATIBLE_SIZE
    for bad in (good[:-8], good + b"\0", good[:20]):
        p.write_bytes(bad)
        assert header(p, F.FORMAT_BINCODE, 0, dim)[0] == F.ERR_IO
    o = state_obj(1.0, points, 6.0, links)
    for mutate in ("drop_beta", "short_matrix", "truncate", "trailing"):
        oo = json.loads(dumps(o))
        s = None
        if mutate == "drop_beta":
            del oo["beta"]
        elif mutate == "short_matrix":
            oo["link_matrix"]["data"][0] = oo["link_matrix"]["data"][0][:8]
        elif mutate == "truncate":
            s = dumps(oo)[:-3]
        else:
            s = dumps(oo) + " 1"
        p.write_text(s if s is not None else dumps(oo))
        assert header(p, F.FORMAT_JSON, 0, dim)[0] == F.ERR_IO, mutate
    assert header(tmp_path / "missing", F.FORMAT_JSON, 0, dim)[0] == F.ERR_IO
    p.write_bytes(good)
    assert header(p, 7, 0, dim)[0] == F.ERR_INVALID_ARGUMENT
    rc, _, _ = read_host(p, F.FORMAT_BINCODE, 0, dim, nlink + 1)
    assert rc == F.ERR_INCOMPATIBLE_SIZE


def test_oversized_length_prefixes_are_refused(tmp_path):
    # a length prefix larger than the file: an I/O error before any
    # allocation, not an exception through the C ABI
    dim, points = 3, 2
    links, e, nlink = random_state(dim, points, 5)
    p = tmp_path / "big"
    for n in (2 ** 40, 2 ** 62, 2 ** 64 - 1, nlink + 1):
        bad = state_bincode(1.0, points, 6.0, links)
        # the link count is the u64 after size (f64), points (u64) and beta (f64)
        bad = bad[:24] + struct.pack("<Q", n) + bad[32:]
        p.write_bytes(bad)
        assert header(p, F.FORMAT_BINCODE, 0, dim)[0] == F.ERR_IO, n
        good_ef = ef_bincode(1.0, points, 6.0, links, e, 3, dim)
        p.write_bytes(struct.pack("<Q", n) + good_ef[8:])
        assert header(p, F.FORMAT_BINCODE, 1, dim)[0] == F.ERR_IO, n


def test_every_truncation_of_a_bincode_state_is_refused(tmp_path):
    # the length prefixes are checked against the file size before any block
    # is read: every proper prefix of a well-formed file is an I/O error, for
    # the header read and for the read into host buffers
    dim, points = 3, 2
    links, e, nlink = random_state(dim, points, 6)
    good = ef_bincode(0.75, points, 5.0, links, e, 9, dim)
    p = tmp_path / "cut"
    e_end = 8 + nlink * 64
    cuts = set(range(0, len(good), 37)) | {0, 7, 8, 9, e_end - 1, e_end, e_end + 7, e_end + 8, e_end + 39,
                                           e_end + 40, len(good) - 144, len(good) - 1}
    for n in sorted(cuts):
        p.write_bytes(good[:n])
        assert header(p, F.FORMAT_BINCODE, 1, dim)[0] == F.ERR_IO, n
        assert read_host(p, F.FORMAT_BINCODE, 1, dim, nlink)[0] == F.ERR_IO, n
    p.write_bytes(good)
    rc, size, pts, beta, t = header(p, F.FORMAT_BINCODE, 1, dim)
    assert rc == F.OK and (size, pts, beta, t) == (0.75, points, 5.0, 9)


def test_bincode_e_field_of_another_lattice_is_refused(tmp_path):
    # an E block whose site count fits the file but not the lattice: the
    # lengths are consistent, the shape is not
    dim, points = 3, 2
    links, e, nlink = random_state(dim, points, 7)
    p = tmp_path / "e"
    p.write_bytes(ef_bincode(1.0, points, 6.0, links, e[: nlink - dim], 1, dim))
    assert header(p, F.FORMAT_BINCODE, 1, dim)[0] == F.ERR_INCOMPATIBLE_SIZE
    assert read_host(p, F.FORMAT_BINCODE, 1, dim, nlink)[0] == F.ERR_INCOMPATIBLE_SIZE
