"""Restart files (z4c_amd/restart.py, src/outputs/restart.cpp:43-330) on the
CPU: the mesh header and ID-list records equal the bytes the oracle's C
restatement copies from the reference's structs (RegionSize,
LogicalLocation: src/athena.hpp:58-96), the ID list starts at the header
offset the reference computes (restart.cpp:110-112), the parameter dump is
ParameterDump's layout and parses back, and a file reads back bit for bit."""
import numpy as np
import pytest

import oracle as orc
from z4c_amd import restart as rst
from z4c_amd.initial_data import perturbed_flat
from z4c_amd.z4c import MeshBlockGeom

PARAMS = [("comment", [("problem", "Brill-Lindquist", "# two punctures"),
                       ("configure", "--prob=z4c_two_punctures -z", "")]),
          ("mesh", [("nx1", "32", "# cells"), ("x1min", "-16.0", ""), ("numlevel", "3", "")]),
          ("meshblock", [("nx1", "8", ""), ("nx2", "8", ""), ("nx3", "8", "")]),
          ("z4c", [("lapse_oplog", "2.0", ""), ("shift_eta", "2.0", "# damping")])]
SIZE = dict(x1min=-16.0, x2min=-16.0, x3min=-16.0, x1max=16.0, x2max=16.0, x3max=16.0,
            x1rat=1.0, x2rat=1.0, x3rat=1.0, nx1=32, nx2=32, nx3=32)
SIZE9 = [SIZE[k] for k in ("x1min", "x2min", "x3min", "x1max", "x2max", "x3max", "x1rat",
                           "x2rat", "x3rat")]


def _restart(nb=3, udata=b"", ublock=None):
    g = MeshBlockGeom.cube(8, 1.0, -4.0)
    blocks = [np.ascontiguousarray(perturbed_flat(g, seed=b)) for b in range(nb)]
    locs = [(b, 2 * b + 1, 3, 1 + b % 2) for b in range(nb)]
    return rst.Restart(rst.parameter_dump(PARAMS, "abc123"), 1, dict(SIZE), 12.5, 0.0625, 200,
                       locs, [1.0 + 0.5 * b for b in range(nb)], blocks, udata, ublock or [])


def test_header_and_id_list_bytes_match_the_reference_structs(tmp_path):
    r = _restart()
    p = tmp_path / "z4c.00003.rst"
    datasize = rst.write_restart(p, r)
    raw = p.read_bytes()
    n = len(r.params.encode())
    hdr = orc.rst_header(3, 1, SIZE9, [32, 32, 32], 12.5, 0.0625, 200, datasize)
    assert rst.HEADER_BYTES == len(hdr) == 124
    assert raw[n:n + len(hdr)] == hdr
    off = n + len(hdr)  # headeroffset with udsize = 0
    for b in range(3):
        rec = orc.rst_idrec(r.locs[b][:3], r.locs[b][3], r.costs[b])
        assert rst.LIST_BYTES == len(rec) == 40
        assert raw[off + 40 * b:off + 40 * (b + 1)] == rec
    data0 = off + 40 * 3
    assert datasize == 22 * 17 ** 3 * 8
    assert raw[data0:data0 + datasize] == r.blocks[0].tobytes()
    assert len(raw) == data0 + 3 * datasize


def test_round_trip_with_user_data(tmp_path):
    trackers = [((5.0, 0.0, 0.0), (0.0, 0.1, 0.0)), ((-5.0, 0.0, 0.0), (0.0, -0.1, 0.0))]
    ud = rst.pack_trackers(trackers)
    assert len(ud) == 2 * 2 * 3 * 8  # 2*NDIM Reals per tracker (restart.cpp:108)
    ub = [np.arange(4, dtype=np.float64).tobytes() + bytes([b]) * 8 for b in range(3)]
    r = _restart(udata=ud, ublock=ub)
    p = tmp_path / "z4c.final.rst"
    rst.write_restart(p, r)
    q = rst.read_restart(p, udsize=len(ud))
    assert q.params == r.params and q.root_level == 1 and q.ncycle == 200
    assert (q.time, q.dt) == (12.5, 0.0625)
    assert q.mesh_size == SIZE
    assert q.locs == r.locs and q.costs == r.costs
    assert q.user_mesh_data == ud and q.user_block_data == ub
    for a, b in zip(q.blocks, r.blocks):
        assert a.shape == (22, 17, 17, 17) and np.array_equal(a, b)


def test_parameter_dump_layout_and_parse():
    text = rst.parameter_dump(PARAMS, "abc123")
    lines = text.splitlines()
    assert lines[0].startswith("#---") and lines[1] == "# GR-Athena++ GIT_HASH: abc123"
    assert lines[-1] == "<par_end>"
    mesh = [l for l in lines[lines.index("<mesh>") + 1:] if "=" in l][:3]
    assert len({l.index("=") for l in mesh}) == 1                  # names padded to one column
    assert mesh[0].index("= ") + 2 + len("-16.0") + 1 == mesh[0].index("# cells")
    par = rst.parse_parameters(text)
    assert par["mesh"] == {"nx1": "32", "x1min": "-16.0", "numlevel": "3"}
    assert par["comment"]["configure"] == "--prob=z4c_two_punctures -z"
    assert par["z4c"]["shift_eta"] == "2.0"


def test_bad_files_are_rejected(tmp_path):
    r = _restart()
    r.blocks[1] = r.blocks[1][:, :-1]
    with pytest.raises(ValueError):
        rst.write_restart(tmp_path / "a.rst", r)
    r = _restart()
    p = tmp_path / "b.rst"
    rst.write_restart(p, r)
    p.write_bytes(p.read_bytes()[:-8])
    with pytest.raises(ValueError):
        rst.read_restart(p)
    p.write_bytes(b"no parameters here")
    with pytest.raises(ValueError):
        rst.read_restart(p)


def test_layout_by_explicit_offsets(tmp_path):
    """The file parsed with byte offsets derived by hand from the
    reference's declarations and write order, independent of the ctypes
    structs and of the oracle's struct copies:

    RestartOutput::WriteOutputFile (restart.cpp:138-147) writes, after the
    parameter text: int nbtotal (+0), int root_level (+4), RegionSize (+8;
    athena.hpp:88-96: nine doubles +0..+71, int nx1 +72, nx2 +76, nx3 +80,
    4 bytes of tail padding to the struct's 8-byte alignment: 88 bytes),
    Real time (+96), Real dt (+104), int ncycle (+112), uint64 datasize
    (+116, io_wrapper.hpp:26) -- 124 bytes, which is headeroffset minus the
    text and the user data (restart.cpp:115-117).  Each ID record
    (restart.cpp:215-218) is LogicalLocation (athena.hpp:58-66: int64 lx1
    +0, lx2 +8, lx3 +16, int level +24, tail padding to 32) then double cost
    (+32): listsize = 40 (restart.cpp:119).  Block b's data follows at
    headeroffset + 40 nbtotal + b datasize (restart.cpp:336)."""
    import struct
    r = _restart(nb=2)
    p = tmp_path / "z4c.00001.rst"
    datasize = rst.write_restart(p, r)
    raw = p.read_bytes()
    n = raw.index(b"<par_end>\n") + len(b"<par_end>\n")
    assert raw[:n].decode() == r.params
    nbtotal, root_level = struct.unpack_from("<ii", raw, n)
    assert (nbtotal, root_level) == (2, 1)
    rs = struct.unpack_from("<9d3i", raw, n + 8)
    assert list(rs[:9]) == SIZE9 and list(rs[9:]) == [32, 32, 32]
    time_, dt = struct.unpack_from("<dd", raw, n + 96)
    (ncycle,) = struct.unpack_from("<i", raw, n + 112)
    (dsize,) = struct.unpack_from("<Q", raw, n + 116)
    assert (time_, dt, ncycle, dsize) == (12.5, 0.0625, 200, datasize)
    head = n + 124  # headeroffset, no user mesh data
    for b in range(2):
        lx1, lx2, lx3, level = struct.unpack_from("<3qi", raw, head + 40 * b)
        (cost,) = struct.unpack_from("<d", raw, head + 40 * b + 32)
        assert (lx1, lx2, lx3, level) == tuple(r.locs[b]) and cost == r.costs[b]
    for b in range(2):
        o = head + 40 * 2 + b * datasize
        blk = np.frombuffer(raw, dtype="<f8", count=datasize // 8, offset=o)
        np.testing.assert_array_equal(blk, r.blocks[b].reshape(-1))
    assert len(raw) == head + 40 * 2 + 2 * datasize
