"""Restart files (``.rst``) in the reference's binary layout, for a vacuum
Z4c run (host side, numpy).

Writer: RestartOutput::WriteOutputFile (src/outputs/restart.cpp:43-330).
Reader: ParameterInput::LoadFromFile (src/parameter_input.cpp:309-346) for
the parameter text and the restart branch of the Mesh constructor
(src/mesh/mesh.cpp:860-900) for the header.  The file is, in order:

1. the parameter dump (ParameterInput::ParameterDump,
   src/parameter_input.cpp:1381-1408), ending with the line ``<par_end>``;
2. the mesh header, fields copied back to back: ``int nbtotal``,
   ``int root_level``, ``RegionSize mesh_size`` (src/athena.hpp:88-96: nine
   doubles x1min..x3rat, three ints nx1..nx3, padded to 8 bytes),
   ``Real time``, ``Real dt``, ``int ncycle``, ``uint64 datasize`` (the
   bytes of one MeshBlock's data);
3. the user mesh data (``udsize`` bytes; the puncture trackers add
   ``pos[3], betap[3]`` each, restart.cpp:168-174);
4. the ID list: per MeshBlock its LogicalLocation (src/athena.hpp:58-66:
   int64 lx1, lx2, lx3, int level, padded to 32 bytes) and its double cost;
5. the MeshBlock data in global-ID order: for a vacuum Z4c build only
   ``storage.u`` (22 fields, AthenaArray order (n, k, j, i)) followed by the
   user MeshBlock data (MeshBlock::GetBlockSizeInBytes,
   src/mesh/meshblock.cpp:1051-1099).

Padding bytes are written as zeros.  The library is the product for the
arithmetic; these helpers move a device state to and from such files.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

NGHOST = 4
NVAR = 22
PAR_END = "<par_end>"


class RegionSize(C.Structure):
    """RegionSize (src/athena.hpp:88-96), native alignment."""
    _fields_ = [(n, C.c_double) for n in ("x1min", "x2min", "x3min", "x1max", "x2max", "x3max",
                                          "x1rat", "x2rat", "x3rat")] + [
        ("nx1", C.c_int), ("nx2", C.c_int), ("nx3", C.c_int)]


class LogicalLocation(C.Structure):
    """LogicalLocation (src/athena.hpp:58-66), native alignment."""
    _fields_ = [("lx1", C.c_int64), ("lx2", C.c_int64), ("lx3", C.c_int64), ("level", C.c_int)]


HEADER_BYTES = 3 * C.sizeof(C.c_int) + C.sizeof(RegionSize) + 2 * C.sizeof(C.c_double) + 8
LIST_BYTES = C.sizeof(LogicalLocation) + C.sizeof(C.c_double)


def parameter_dump(blocks, git_hash=""):
    """ParameterInput::ParameterDump: ``blocks`` is a list of
    (block_name, [(name, value, comment), ...]) in input order; names and
    values are padded to the longest of their block plus one space."""
    out = ["#------------------------- PAR_DUMP -------------------------\n",
           f"# GR-Athena++ GIT_HASH: {git_hash}\n"]
    for bname, lines in blocks:
        out.append("\n# =========================================================\n")
        out.append(f"<{bname}>\n")
        ln = max((len(n) for n, _, _ in lines), default=0)
        lv = max((len(v) for _, v, _ in lines), default=0)
        for n, v, c in lines:
            out.append(n + " " * (ln - len(n) + 1) + "= " + v + " " * (lv - len(v) + 1) + c + "\n")
    out.append("#------------------------- PAR_DUMP -------------------------\n")
    out.append(PAR_END + "\n")
    return "".join(out)


def parse_parameters(text):
    """The parameter lines of a dump (ParameterInput::LoadFromStream /
    ParseLine, src/parameter_input.cpp:187-300, 387-430): {block: {name:
    value}}, stopping at ``<par_end>``; comment lines and comments
    (``#`` or ``//``) are dropped."""
    out, cur = {}, None
    for line in text.splitlines():
        s = line.lstrip(" ")
        if not s or s.startswith("#") or s.startswith("//"):
            continue
        if s.startswith(PAR_END):
            break
        if s.startswith("<"):
            cur = out.setdefault(s[1:s.index(">")], {})
            continue
        if cur is None or "=" not in s:
            raise ValueError(f"malformed parameter line: {line!r}")
        cut = min([i for i in (s.find("#"), s.find("//")) if i >= 0], default=len(s))
        eq = s.index("=")
        cur[s[:eq].rstrip(" ")] = s[eq + 1:cut].strip(" ")
    return out


@dataclass
class Restart:
    """Contents of a restart file."""
    params: str                 # parameter dump, ``<par_end>`` line included
    root_level: int
    mesh_size: dict             # RegionSize fields by name
    time: float
    dt: float
    ncycle: int
    locs: list                  # per block (lx1, lx2, lx3, level)
    costs: list                 # per block
    blocks: list                # per block (22, nn3, nn2, nn1) float64 arrays
    user_mesh_data: bytes = b""
    user_block_data: list = field(default_factory=list)  # per block bytes

    @property
    def nbtotal(self):
        return len(self.blocks)


def pack_trackers(trackers):
    """User mesh data of puncture trackers: pos[3] then betap[3] each."""
    return b"".join(np.asarray(list(p) + list(b), dtype=np.float64).tobytes() for p, b in trackers)


def write_restart(path, r: Restart):
    """RestartOutput::WriteOutputFile for one rank holding every block."""
    if not r.blocks:
        raise ValueError("no MeshBlocks")
    ub = r.user_block_data or [b""] * r.nbtotal
    sizes = {a.nbytes + len(u) for a, u in zip(r.blocks, ub)}
    if len(sizes) != 1:
        raise ValueError("every MeshBlock must have the same data size")
    datasize = sizes.pop()
    text = r.params if r.params.rstrip("\n").endswith(PAR_END) else r.params + PAR_END + "\n"
    rs = RegionSize(**r.mesh_size)
    with open(path, "wb") as f:
        f.write(text.encode())
        f.write(np.int32(r.nbtotal).tobytes() + np.int32(r.root_level).tobytes())
        f.write(bytes(rs))
        f.write(np.float64(r.time).tobytes() + np.float64(r.dt).tobytes())
        f.write(np.int32(r.ncycle).tobytes() + np.uint64(datasize).tobytes())
        f.write(r.user_mesh_data)
        for loc, cost in zip(r.locs, r.costs):
            f.write(bytes(LogicalLocation(*[int(v) for v in loc])))
            f.write(np.float64(cost).tobytes())
        for a, u in zip(r.blocks, ub):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            f.write(u)
    return datasize


def read_restart(path, udsize=0, block_shape=None):
    """Reads a restart file.  ``udsize`` is the user mesh data size the
    problem enrolls (it is not stored in the file).  The block shape
    defaults to the VC arrays of <meshblock> nx1..nx3 (or the mesh size),
    nn = nx + 1 + 2*NGHOST; bytes past storage.u are the user block data."""
    raw = open(path, "rb").read()
    loc = raw.find(PAR_END.encode(), 0, 409600)
    if loc < 0:
        raise ValueError("<par_end> is not found in the first 400KBytes")
    hdr = loc + len(PAR_END) + 1
    params = raw[:hdr].decode()
    o = hdr
    nbtotal, root_level = np.frombuffer(raw, np.int32, 2, o)
    o += 8
    rs = RegionSize.from_buffer_copy(raw[o:o + C.sizeof(RegionSize)])
    o += C.sizeof(RegionSize)
    time, dt = np.frombuffer(raw, np.float64, 2, o)
    o += 16
    ncycle = int(np.frombuffer(raw, np.int32, 1, o)[0])
    datasize = int(np.frombuffer(raw, np.uint64, 1, o + 4)[0])
    o += 12
    user_mesh = raw[o:o + udsize]
    o += udsize
    locs, costs = [], []
    for _ in range(int(nbtotal)):
        ll = LogicalLocation.from_buffer_copy(raw[o:o + C.sizeof(LogicalLocation)])
        locs.append((ll.lx1, ll.lx2, ll.lx3, ll.level))
        costs.append(float(np.frombuffer(raw, np.float64, 1, o + C.sizeof(LogicalLocation))[0]))
        o += LIST_BYTES
    if block_shape is None:
        par = parse_parameters(params)
        mb = par.get("meshblock", {})
        nx = [int(mb.get(f"nx{d}", getattr(rs, f"nx{d}"))) for d in (1, 2, 3)]
        block_shape = (NVAR, nx[2] + 1 + 2 * NGHOST, nx[1] + 1 + 2 * NGHOST, nx[0] + 1 + 2 * NGHOST)
    nu = int(np.prod(block_shape)) * 8
    if nu > datasize:
        raise ValueError(f"storage.u of shape {block_shape} exceeds the block data size {datasize}")
    if len(raw) != o + int(nbtotal) * datasize:
        raise ValueError("file size does not match the header")
    blocks, ub = [], []
    for b in range(int(nbtotal)):
        s = o + b * datasize
        blocks.append(np.frombuffer(raw, np.float64, nu // 8, s).reshape(block_shape).copy())
        ub.append(raw[s + nu:s + datasize])
    size = {n: getattr(rs, n) for n, _ in RegionSize._fields_}
    return Restart(params, int(root_level), size, float(time), float(dt), ncycle, locs, costs,
                   blocks, user_mesh, ub)


def from_device(z, u, params, mesh_size, time, dt, ncycle, locs, costs=None, root_level=0,
                user_mesh_data=b""):
    """A Restart of the device state ``u`` of context ``z`` (blocks in the
    context's order, which is the global-ID order of the file)."""
    blocks = [np.ascontiguousarray(a) for a in z.download(u, NVAR)]
    costs = list(costs) if costs is not None else [1.0] * len(blocks)
    return Restart(params, root_level, dict(mesh_size), time, dt, ncycle, list(locs), costs,
                   blocks, user_mesh_data)


def to_device(z, r: Restart):
    """The state of a Restart as a device buffer of context ``z``."""
    if len(r.blocks) != len(z.blocks):
        raise ValueError("the restart holds a different number of MeshBlocks")
    return z.upload(r.blocks, NVAR)
