"""Independent restatement of the serde layouts the reference derives for
LatticeStateDefault<D> and LatticeStateEFSyncDefault (src/simulation/state.rs:
653-659,1046-1062, src/field.rs:26-31,582-586,1022-1027, src/lattice.rs:42-49)
in bincode 1.x (struct.pack, little endian, u64 lengths) and serde_json (the
json module): a matrix is the tuple of its 9 column-major entries, an entry
the tuple (re, im) (nalgebra 0.31, num-complex 0.4)."""
import json
import struct

import numpy as np


def state_obj(size, points, beta, links):
    m = np.asarray(links).reshape(-1, 9, 2)
    return {"lattice": {"size": float(size), "dim": int(points)}, "beta": float(beta),
            "link_matrix": {"data": [[[float(c[0]), float(c[1])] for c in mat] for mat in m]}}


def ef_obj(size, points, beta, links, efield, t, dim):
    e = np.asarray(efield).reshape(-1, dim, 8)
    return {"e_field": {"data": [[{"data": [float(v) for v in su3]} for su3 in site] for site in e]},
            "t": int(t), "lattice_state": state_obj(size, points, beta, links)}


def state_bincode(size, points, beta, links):
    links = np.ascontiguousarray(links, dtype="<f8")
    return (struct.pack("<dQdQ", size, points, beta, links.size // 18) + links.tobytes())


def ef_bincode(size, points, beta, links, efield, t, dim):
    e = np.ascontiguousarray(efield, dtype="<f8")
    nsite = e.size // (8 * dim)
    return struct.pack("<Q", nsite) + e.tobytes() + struct.pack("<Q", t) + state_bincode(size, points, beta, links)


def parse_state_bincode(b):
    size, points, beta, n = struct.unpack_from("<dQdQ", b, 0)
    links = np.frombuffer(b, dtype="<f8", count=n * 18, offset=32).reshape(n, 18)
    assert len(b) == 32 + n * 144
    return size, points, beta, links


def parse_ef_bincode(b, dim):
    (nsite,) = struct.unpack_from("<Q", b, 0)
    e = np.frombuffer(b, dtype="<f8", count=nsite * dim * 8, offset=8).reshape(nsite * dim, 8)
    off = 8 + nsite * dim * 64
    (t,) = struct.unpack_from("<Q", b, off)
    size, points, beta, links = parse_state_bincode(b[off + 8:])
    return size, points, beta, links, e, t


def links_from_obj(o):
    return np.array(o["link_matrix"]["data"], dtype=float).reshape(-1, 18)


def efield_from_obj(o):
    return np.array([[v["data"] for v in site] for site in o["e_field"]["data"]], dtype=float).reshape(-1, 8)


def dumps(o):
    return json.dumps(o, separators=(",", ":"))
