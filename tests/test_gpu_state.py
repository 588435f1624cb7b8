"""State interchange through the device: a state written by the library
(links and E downloaded from HBM) is read by the independent serde
restatement of tests/_serde_np.py and by the library itself, bit for bit."""
import json
import sys

import numpy as np
import pytest

import oracle as O
import lqcd
from lqcd import _ffi as F
from _serde_np import efield_from_obj, links_from_obj, parse_ef_bincode, parse_state_bincode

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", [lqcd.F64, lqcd.F32])
@pytest.mark.parametrize("fmt", [F.FORMAT_BINCODE, F.FORMAT_JSON])
def test_state_roundtrip(tmp_path, precision, fmt):
    announce(f"state write/read precision={precision} format={fmt}")
    st = lqcd.LatticeStateDefault.new_determinist(0.5, 5.75, 4, 11, 4, precision)
    u = st.link_matrix()
    p = tmp_path / "state"
    st.write(p, fmt)
    if fmt == F.FORMAT_BINCODE:
        size, points, beta, links = parse_state_bincode(p.read_bytes())
    else:
        o = json.loads(p.read_text())
        size, points, beta = o["lattice"]["size"], o["lattice"]["dim"], o["beta"]
        links = links_from_obj(o)
    assert (size, points, beta) == (0.5, 4, 5.75)
    assert np.array_equal(links, u)
    back = lqcd.LatticeStateDefault.read(p, 4, fmt, precision)
    assert back.beta() == 5.75 and back.lattice().size() == 0.5 and back.lattice().dim() == 4
    assert np.array_equal(back.link_matrix(), u)
    assert back.average_trace_plaquette() == st.average_trace_plaquette()


@pytest.mark.parametrize("fmt", [F.FORMAT_BINCODE, F.FORMAT_JSON])
def test_efsync_roundtrip_and_resume(tmp_path, fmt):
    # a run written mid-way and resumed continues exactly as the run itself
    announce(f"EFSync state write/read/resume format={fmt}")
    st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 4, 12, 3, lqcd.F64)
    st.set_md_variant(2)   # one work decomposition for both runs (they differ in rounding)
    ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(st, 5, 0, project=False)
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    ef.simulate_symplectic_n(integ, 0.01, 3)
    p = tmp_path / "ef"
    ef.write(p, fmt)
    if fmt == F.FORMAT_BINCODE:
        size, points, beta, links, e, t = parse_ef_bincode(p.read_bytes(), 3)
    else:
        o = json.loads(p.read_text())
        links, e, t = links_from_obj(o["lattice_state"]), efield_from_obj(o), o["t"]
    assert t == ef.t() and np.array_equal(links, ef.link_matrix()) and np.array_equal(e, ef.e_field())
    back = lqcd.LatticeStateEFSyncDefault.read(p, 3, fmt)
    assert back.t() == ef.t() and back.lattice_state().beta() == 6.0
    assert np.array_equal(back.link_matrix(), ef.link_matrix())
    assert np.array_equal(back.e_field(), ef.e_field())
    back.lattice_state().set_md_variant(2)
    ef.simulate_symplectic_n(integ, 0.01, 2)
    back.simulate_symplectic_n(integ, 0.01, 2)
    assert np.array_equal(back.link_matrix(), ef.link_matrix())
    assert np.array_equal(back.e_field(), ef.e_field())


def test_non_hypercubic_state_is_refused(tmp_path):
    announce("non-hypercubic state write")
    ol = O.OracleLattice(3, [4, 4, 6])
    st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, [4, 4, 6], 3), 6.0, ol.cold_start())
    with pytest.raises(lqcd.LqcdError):
        st.write(tmp_path / "x")


@pytest.mark.parametrize("precision", [lqcd.F64, lqcd.F32])
def test_efsync_bincode_through_several_chunks(tmp_path, precision):
    # 16^4: 37.7 MB of links and 16.8 MB of E, each block several 8 MB pinned
    # chunks between the file and HBM, the last one partial
    announce(f"EFSync bincode 16^4 write/read precision={precision}")
    st = lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 16, 21, 4, precision)
    ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(st, 9, 0, project=False)
    u, e = ef.link_matrix(), ef.e_field()
    p = tmp_path / "ef16"
    ef.write(p)
    size, points, beta, links, e2, t = parse_ef_bincode(p.read_bytes(), 4)
    assert (size, points, beta, t) == (1.0, 16, 6.0, ef.t())
    assert np.array_equal(links, u) and np.array_equal(e2, e)
    back = lqcd.LatticeStateEFSyncDefault.read(p, 4, F.FORMAT_BINCODE, precision)
    assert np.array_equal(back.link_matrix(), u) and np.array_equal(back.e_field(), e)
    # refused files leave the lattice as it was: a truncated file (I/O error)
    # and a file of another lattice (incompatible size)
    short = tmp_path / "short"
    short.write_bytes(p.read_bytes()[:-8])
    other = tmp_path / "other"
    lqcd.LatticeStateEFSyncDefault.new_random_e_state(
        lqcd.LatticeStateDefault.new_determinist(1.0, 6.0, 8, 3, 4, precision), 4, 0, project=False).write(other)
    h = back.lattice_state().handle
    for path, code in ((short, F.ERR_IO), (other, F.ERR_INCOMPATIBLE_SIZE)):
        assert F.lib().lqcd_state_read(h, str(path).encode(), F.FORMAT_BINCODE, 1, None, None) == code
        assert np.array_equal(back.link_matrix(), u) and np.array_equal(back.e_field(), e)
