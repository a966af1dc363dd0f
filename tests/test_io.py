"""The reference's file formats through the library's host readers (no GPU):
gmsh 2.2 tet meshes, .part files and parameterSetting.txt."""
import os

import numpy as np
import pytest

REF = "/root/reference/Assets"


def test_msh_round_trip(tmp_path, cube4):
    from sgipc import io
    p = tmp_path / "cube.msh"
    io.write_msh(p, cube4.x, cube4.tets)
    x, t = io.read_msh(p)
    assert np.array_equal(x, cube4.x) and np.array_equal(t, cube4.tets.astype(np.uint32))   # %.17g round trips


def test_msh_reader_accepts_reference_layout(tmp_path):
    # the reference writer's layout: no $EndNodes / $EndElements, element
    # lines "id 4 0 a b c d" (1-based), CRLF tolerated; tags before the ids
    from sgipc import io
    p = tmp_path / "ref.msh"
    p.write_text("$MeshFormat\r\n2.2 0 8\r\n$EndMeshFormat\r\n$Nodes\r\n4\r\n1 0 0 0\r\n2 1.5e+00 0 0\r\n"
                 "3 0 1 0\r\n4 0 0 1\r\n$Elements\r\n1\r\n1 4 2 7 7 1 2 3 4\r\n")
    x, t = io.read_msh(p)
    assert x.shape == (4, 3) and x[1, 0] == 1.5 and t.tolist() == [[0, 1, 2, 3]]


def test_io_errors(tmp_path):
    from sgipc import io
    from sgipc.capi import SgipcError
    with pytest.raises(SgipcError, match="cannot open"):
        io.read_msh(tmp_path / "missing.msh")
    bad = tmp_path / "bad.msh"
    bad.write_text("$Nodes\n3\n1 0 0 0\n")
    with pytest.raises(SgipcError, match="truncated"):
        io.read_msh(bad)
    oob = tmp_path / "oob.msh"
    oob.write_text("$Nodes\n1\n1 0 0 0\n$Elements\n1\n1 4 0 1 2 3 4\n")
    with pytest.raises(SgipcError, match="out of range"):
        io.read_msh(oob)
    part = tmp_path / "bad.part"
    part.write_text("0\n1\nx\n")
    with pytest.raises(SgipcError, match="non-numeric"):
        io.read_part(part)
    par = tmp_path / "p.txt"
    par.write_text("volume_mesh_density 1000\n")
    with pytest.raises(SgipcError, match="missing"):
        io.read_parameters(par)


def test_parameters_in_file_order(tmp_path):
    from sgipc import io
    vals = [float(i + 1) * 0.5 for i in range(16)]
    p = tmp_path / "parameterSetting.txt"
    p.write_text("\n".join("name_%d %r" % (i, v) for i, v in enumerate(vals)))
    d = io.read_parameters(p)
    assert [d[n] for n in io.PARAMETER_NAMES] == vals


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference assets not present")
def test_reference_assets_match_fixture():
    """The shipped METIS-sorted mesh and its part file, read by the library,
    equal the committed fixture (written from the same files); the shipped
    parameterSetting.txt parses to its values."""
    from sgipc import io
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "high_mat_sorted16.npz"))
    x, t = io.read_msh(os.path.join(REF, "sorted_mesh", "high_mat_sorted.16.msh"))
    assert np.array_equal(x, d["x_rest"]) and np.array_equal(t.astype(np.int32), d["tets"])
    part = io.read_part(os.path.join(REF, "sorted_mesh", "high_mat_sorted.16.part"))
    assert np.array_equal(part, d["part"])
    prm = io.read_parameters(os.path.join(REF, "scene", "parameterSetting.txt"))
    assert prm["density"] == 1000 and prm["poisson_rate"] == 0.49 and prm["dt"] == 1e-2
    assert prm["pcg_threshold"] == 1e-4 and prm["newton_threshold"] == 1e-2 and prm["relative_dhat"] == 1e-3
