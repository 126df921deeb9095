"""Text outputs in the reference's formats (host side, numpy).

* ``write_tab`` / ``read_tab``: the formatted-table (``tab``) dump of one
  MeshBlock, as FormattedTableOutput::WriteOutputFile writes it for
  vertex-centred data (src/outputs/formatted_table.cpp:33-184): file name
  ``<basename>.block<gid>.<file_id>.<NNNNN>.tab``, a header line with time,
  cycle and the variable, a column-header line, then one line per vertex
  (k, j, i order) with ``%04d`` indices, the vertex coordinates x1f/x2f/x3f
  and the data, each value printed with ``" " + data_format`` (default
  ``%12.5e``, src/outputs/outputs.cpp:323-326);
* ``write_hst``: the history file with the Z4c columns
  (src/outputs/history.cpp:385-445).

Variable names are the reference's output names (src/z4c/z4c.hpp:54-131),
so dumps of the two codes can be compared column by column.
"""
from __future__ import annotations

import os

import numpy as np

Z4C_NAMES = [
    "geom.z4c.chi", "geom.z4c.gxx", "geom.z4c.gxy", "geom.z4c.gxz", "geom.z4c.gyy",
    "geom.z4c.gyz", "geom.z4c.gzz", "geom.z4c.Khat", "geom.z4c.Axx", "geom.z4c.Axy",
    "geom.z4c.Axz", "geom.z4c.Ayy", "geom.z4c.Ayz", "geom.z4c.Azz", "geom.z4c.Gamx",
    "geom.z4c.Gamy", "geom.z4c.Gamz", "geom.z4c.Theta", "geom.z4c.alpha", "geom.z4c.betax",
    "geom.z4c.betay", "geom.z4c.betaz"]
ADM_NAMES = [
    "geom.adm.alpha", "geom.adm.betax", "geom.adm.betay", "geom.adm.betaz", "geom.adm.gxx",
    "geom.adm.gxy", "geom.adm.gxz", "geom.adm.gyy", "geom.adm.gyz", "geom.adm.gzz",
    "geom.adm.Kxx", "geom.adm.Kxy", "geom.adm.Kxz", "geom.adm.Kyy", "geom.adm.Kyz",
    "geom.adm.Kzz", "geom.adm.psi4"]
CON_NAMES = ["geom.con.C", "geom.con.H", "geom.con.M", "geom.con.Z", "geom.con.Mx",
             "geom.con.My", "geom.con.Mz"]
GROUPS = {"geom.z4c": Z4C_NAMES, "geom.adm": ADM_NAMES, "geom.con": CON_NAMES}
NGHOST = 4


def tab_filename(basename, gid, file_id, file_number):
    return f"{basename}.block{gid}.{file_id}.{file_number:05d}.tab"


def write_tab(directory, basename, file_id, file_number, geom, gid, arrays, variable,
              time=0.0, cycle=0, data_format="%12.5e", ghost_zones=False):
    """Write one block's ``tab`` dump.  ``arrays``: (nvar, nn3, nn2, nn1)
    array of the variable group ``variable`` ("geom.z4c", "geom.adm",
    "geom.con") or a single field (1, nn3, nn2, nn1) named ``variable``."""
    names = GROUPS.get(variable, [variable])
    arrays = np.asarray(arrays)
    if arrays.shape[0] != len(names):
        raise ValueError(f"{variable}: {len(names)} fields expected, got {arrays.shape[0]}")
    fmt = " " + data_format
    lo = [NGHOST] * 3
    hi = [NGHOST + n for n in geom.nx]  # vertex-centred: ie + 1
    if ghost_zones:
        lo = [v - NGHOST for v in lo]
        hi = [v + NGHOST for v in hi]
    x = geom.coords()
    path = os.path.join(directory, tab_filename(basename, gid, file_id, file_number))
    with open(path, "w") as f:
        f.write("# Athena++ data at time=%e" % time)
        f.write("  cycle=%d" % cycle)
        f.write("  variables=%s \n" % variable)
        f.write("#")
        for axis, (l, h) in enumerate(zip(lo, hi)):
            if l != h:
                f.write(" %s       x%df     " % ("ijk"[axis], axis + 1))
        for nm in names:
            f.write("    %s      " % nm)
        f.write("\n")
        for k in range(lo[2], hi[2] + 1):
            for j in range(lo[1], hi[1] + 1):
                for i in range(lo[0], hi[0] + 1):
                    line = []
                    if lo[0] != hi[0]:
                        line.append("%04d" % i + fmt % x[0][i])
                    if lo[1] != hi[1]:
                        line.append(" %04d" % j + fmt % x[1][j])
                    if lo[2] != hi[2]:
                        line.append(" %04d" % k + fmt % x[2][k])
                    line.extend(fmt % arrays[n, k, j, i] for n in range(len(names)))
                    f.write("".join(line) + "\n")
    return path


def read_tab(path):
    """Parse a ``tab`` dump: (header dict, column names, data array)."""
    with open(path) as f:
        head = f.readline()
        cols = f.readline()
        rows = [line.split() for line in f if line.strip()]
    info = {}
    for tok in head.lstrip("#").split():
        if "=" in tok:
            k, v = tok.split("=", 1)
            info[k] = v
    names = cols.lstrip("#").split()
    data = np.array([[float(v) for v in r] for r in rows])
    return info, names, data


HST_Z4C_COLUMNS = ["H-norm2", "M-norm2", "Mx-norm2", "My-norm2", "Mz-norm2", "Z-norm2",
                   "Theta-norm2", "C-norm2"]


def write_hst(path, times, rows, dts, nbtotal, data_format="%12.5e", git_hash="unknown"):
    """History file of a vacuum Z4c run in the reference's layout
    (HistoryOutput::WriteOutputFile, src/outputs/history.cpp:383-447):

    * first line ``# GR-Athena++ (<GIT_HASH>) history data`` (:389-392);
    * the column line ``# [1]=time [2]=dt [3]=N_MeshBlock [4]=H-norm2 ...``,
      each entry followed by one blank (:394-432);
    * per row: time and dt with ``" " + data_format``, the block count with
      ``" %d"``, then the eight Z4c columns with ``" " + data_format``
      (:436-443); ``data_format`` defaults to ``%12.5e`` and gets its
      leading blank as in src/outputs/outputs.cpp:322-325.

    ``dts`` is one dt per row (the mesh dt when the row is written) or a
    single value for every row."""
    fmt = " " + data_format
    dts = np.broadcast_to(np.asarray(dts, dtype=float), (len(times),))
    cols = ["time", "dt", "N_MeshBlock"] + HST_Z4C_COLUMNS
    with open(path, "w") as f:
        f.write(f"# GR-Athena++ ({git_hash}) history data\n")
        f.write("# " + "".join(f"[{i + 1}]={c} " for i, c in enumerate(cols)) + "\n")
        for t, dt, r in zip(times, dts, rows):
            if len(r) != len(HST_Z4C_COLUMNS):
                raise ValueError("a history row holds the eight Z4c columns")
            f.write(fmt % t + fmt % dt + " %d" % int(nbtotal) + "".join(fmt % v for v in r) + "\n")


def read_hst(path):
    """Parse a history file: (column names, data array), names from the
    ``[n]=name`` header entries."""
    names, data = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                toks = [t for t in line[1:].split() if t.startswith("[") and "]=" in t]
                if toks:
                    names = [t.split("]=", 1)[1] for t in toks]
                continue
            if line.strip():
                data.append([float(v) for v in line.split()])
    return names, np.array(data)
