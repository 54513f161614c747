"""The GEO reader (cd3d_read_geo, TDomain::ReadGeo ReadGeo.C:46-200 for
structured macro meshes): host code, no GPU.  The files are the ones the
reference runs of the deformed-mesh fixtures read (oracle/data/make_box_geo.py)."""
import numpy as np
import pytest

import parmoon_cmg_amd as pm
from tests.cases import CD_CASES, _box_geo_module, case_vertices, geo_text


@pytest.mark.parametrize("name", ["m3w_q1_poisson", "m2w_q2_supg_rot"])
def test_reads_the_reference_runs_geo(name, tmp_path):
    c = CD_CASES[name]
    path = tmp_path / "mesh.GEO"
    path.write_text(geo_text(c))
    m, xyz = pm.read_geo(path)
    assert m == c["m"]
    assert np.array_equal(xyz, case_vertices(c))  # the printed coordinates, parsed exactly


def test_brick_box_geo(tmp_path):
    g = _box_geo_module()
    path = tmp_path / "box.GEO"
    path.write_text(g.box_geo(4))
    m, xyz = pm.read_geo(path)
    grid = np.arange(5) / 4.0
    ref = np.stack(np.meshgrid(grid, grid, grid, indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    assert m == 4 and np.abs(xyz - ref).max() == 0.0


def _mutate(text, fn):
    lines = text.split("\n")
    fn(lines)
    return "\n".join(lines)


@pytest.mark.parametrize("how,msg", [
    ("knpr", "KNPR"),          # a vertex given by boundary parameters
    ("order", "lattice order"),  # cell vertices in another order
    ("count", "structured"),   # NEL not a cube
    ("tetra", "hexahedral"),
])
def test_refuses_what_the_lattice_cannot_hold(tmp_path, how, msg):
    g = _box_geo_module()
    text = g.box_geo(2)
    lines = text.split("\n")
    nv = 27
    kv = lines.index("KVERT")
    if how == "knpr":
        lines[kv + 8 + 2] = "   1" + lines[kv + 8 + 2][4:]
    elif how == "order":
        w = lines[kv + 1].split()
        w[2], w[3] = w[3], w[2]
        lines[kv + 1] = " " + " ".join(w)
    elif how == "count":
        lines[2] = lines[2].replace("     8 ", "     7 ", 1)
    elif how == "tetra":
        lines[2] = lines[2].replace("    4    8 ", "    3    4 ", 1)
    assert len(lines) > kv + 8 + nv // 16
    path = tmp_path / "bad.GEO"
    path.write_text("\n".join(lines))
    with pytest.raises(pm.CD3DError) as e:
        pm.read_geo(path)
    assert msg in str(e.value)


def test_missing_file(tmp_path):
    with pytest.raises(pm.CD3DError) as e:
        pm.read_geo(tmp_path / "none.GEO")
    assert "cannot open" in str(e.value)
