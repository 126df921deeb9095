"""Text outputs in the reference's formats (CPU): the formatted-table dump
of a vertex-centred block (src/outputs/formatted_table.cpp:33-184) and the
history file (src/outputs/history.cpp:385-445)."""
import os

import numpy as np

from z4c_amd import outputs
from z4c_amd.z4c import MeshBlockGeom


def test_tab_dump_layout_and_round_trip(tmp_path):
    g = MeshBlockGeom.cube((3, 2, 2), 0.5, (-1.0, 0.0, 2.0))
    rng = np.random.default_rng(1)
    u = rng.standard_normal((22,) + g.shape)
    path = outputs.write_tab(str(tmp_path), "z4c", "out1", 7, g, 12, u, "geom.z4c",
                             time=0.25, cycle=3)
    assert os.path.basename(path) == "z4c.block12.out1.00007.tab"
    lines = open(path).read().splitlines()
    assert lines[0] == "# Athena++ data at time=2.500000e-01  cycle=3  variables=geom.z4c "
    assert lines[1].startswith("# i       x1f      j       x2f      k       x3f     ")
    assert "    geom.z4c.chi      " in lines[1] and lines[1].rstrip().endswith("geom.z4c.betaz")
    # vertex-centred owned range is..ie+1 per axis: (3+1)(2+1)(2+1) lines
    assert len(lines) == 2 + 4 * 3 * 3
    # first data line: i=4, x1f=-1, j=4, x2f=0, k=4, x3f=2, then %12.5e values
    first = lines[2]
    assert first.startswith("0004 -1.00000e+00 0004  0.00000e+00 0004  2.00000e+00")
    info, names, data = outputs.read_tab(path)
    assert info["time"] == "2.500000e-01" and info["cycle"] == "3"
    assert names[6:] == outputs.Z4C_NAMES
    # values round to %12.5e; order k, j, i
    idx = [(k, j, i) for k in range(4, 7) for j in range(4, 7) for i in range(4, 8)]
    exp = np.array([[u[n, k, j, i] for n in range(22)] for (k, j, i) in idx])
    np.testing.assert_allclose(data[:, 6:], exp, rtol=1e-5, atol=1e-10)
    np.testing.assert_array_equal(data[:, 0], [i for (_, _, i) in idx])


def test_tab_dump_with_ghosts_and_single_field(tmp_path):
    g = MeshBlockGeom.cube(2, 1.0, 0.0)
    u = np.arange(np.prod(g.shape), dtype=float).reshape((1,) + g.shape)
    path = outputs.write_tab(str(tmp_path), "b", "o", 0, g, 0, u, "geom.con.H",
                             data_format="%.3f", ghost_zones=True)
    _, names, data = outputs.read_tab(path)
    assert names[-1] == "geom.con.H"
    assert data.shape[0] == np.prod(g.shape)
    np.testing.assert_array_equal(data[:, -1], u.reshape(-1))


def test_hst_layout(tmp_path):
    """The reference's history layout for a vacuum Z4c run
    (src/outputs/history.cpp:389-443): hash line, column line with
    [3]=N_MeshBlock before the Z4c columns, " %12.5e" values and " %d" for
    the block count."""
    p = tmp_path / "a.hst"
    outputs.write_hst(str(p), [0.0, 0.1], [[1.0] * 8, [2.5e-11] * 8], [0.05, 0.00234375], 8,
                      git_hash="abc123 ")
    text = p.read_text()
    lines = text.splitlines()
    assert lines[0] == "# GR-Athena++ (abc123 ) history data"
    assert lines[1] == ("# [1]=time [2]=dt [3]=N_MeshBlock [4]=H-norm2 [5]=M-norm2 "
                        "[6]=Mx-norm2 [7]=My-norm2 [8]=Mz-norm2 [9]=Z-norm2 [10]=Theta-norm2 "
                        "[11]=C-norm2 ")
    assert lines[2] == "  0.00000e+00  5.00000e-02 8" + "  1.00000e+00" * 8
    assert lines[3] == "  1.00000e-01  2.34375e-03 8" + "  2.50000e-11" * 8
    assert text.endswith("\n")
    names, data = outputs.read_hst(str(p))
    assert names[:3] == ["time", "dt", "N_MeshBlock"] and names[3:] == outputs.HST_Z4C_COLUMNS
    assert data.shape == (2, 11)
    np.testing.assert_array_equal(data[:, 2], [8, 8])
    # a custom data_format replaces %12.5e (outputs.cpp:322-325 prepends the blank)
    outputs.write_hst(str(p), [0.1], [[1.0] * 8], 0.05, 1, data_format="%.3f")
    assert p.read_text().splitlines()[2] == " 0.100 0.050 1" + " 1.000" * 8
