"""Host-side mesh code of libfmbem.so (no GPU needed): OBJ reader
(findNum/readOBJ, octree.c:563-619), icosphere generator against the
reference's sphere_100mm.obj, outward re-winding, perturbed ellipsoid."""
import os

import numpy as np
import pytest

import fmbem_gpu_c_amd as F
from conftest import GOLDEN

OBJ = os.path.join(GOLDEN, "sphere_100mm.obj")


def test_obj_reader_counts_and_ids():
    m = F.Mesh.from_obj(OBJ)
    V, Fc = m.arrays()
    assert V.shape == (642, 3) and Fc.shape == (1280, 3)
    assert Fc.min() == 0 and Fc.max() == 641
    # first records of the file: "v 0.000000 0.000000 -0.100000", "f 20 1 19"
    np.testing.assert_allclose(V[0], [0, 0, -0.1])
    assert Fc[0].tolist() == [19, 0, 18]


def test_obj_reader_errors(tmp_path):
    with pytest.raises(F.FmbError) as e:
        F.Mesh.from_obj(str(tmp_path / "missing.obj"))
    assert e.value.code == -2
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n")
    with pytest.raises(F.FmbError):
        F.Mesh.from_obj(str(bad))


def test_obj_polygons_and_slashes(tmp_path):
    p = tmp_path / "quad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\nf 1/1/1 2/2/1 3/3/1 4/4/1\n")
    V, Fc = F.Mesh.from_obj(str(p)).arrays()
    assert V.shape == (4, 3)
    assert Fc.tolist() == [[0, 1, 2], [0, 2, 3]]


def test_icosphere_reproduces_reference_mesh_vertices():
    V, Fc = F.Mesh.icosphere(3, 0.1).arrays()
    Vr, _ = F.Mesh.from_obj(OBJ).arrays()
    d = np.linalg.norm(V[:, None, :] - Vr[None, :, :], axis=2)
    # the file prints 6 decimals
    assert d.min(axis=1).max() < 2e-6 and d.min(axis=0).max() < 2e-6


def test_icosphere_outward_and_counts():
    for lev in range(0, 5):
        V, Fc = F.Mesh.icosphere(lev, 0.3).arrays()
        assert Fc.shape[0] == 20 * 4 ** lev and V.shape[0] == 10 * 4 ** lev + 2
        t = V[Fc]
        n = np.cross(t[:, 0] - t[:, 2], t[:, 1] - t[:, 2])
        assert np.all(np.einsum("ij,ij->i", n, t.mean(1)) > 0)
        np.testing.assert_allclose(np.linalg.norm(V, axis=1), 0.3, rtol=1e-14)


def test_reference_mesh_normals_point_inward_and_get_flipped():
    m = F.Mesh.from_obj(OBJ)
    V, Fc = m.arrays()
    t = V[Fc]
    n = np.cross(t[:, 0] - t[:, 2], t[:, 1] - t[:, 2])
    assert np.all(np.einsum("ij,ij->i", n, t.mean(1)) < 0)
    assert m.orient_outward() == 1280
    V, Fc = m.arrays()
    t = V[Fc]
    n = np.cross(t[:, 0] - t[:, 2], t[:, 1] - t[:, 2])
    assert np.all(np.einsum("ij,ij->i", n, t.mean(1)) > 0)
    assert m.orient_outward() == 0


def test_perturbed_ellipsoid():
    nu = 20
    m = F.Mesh.ellipsoid(nu, 0.08, 0.07, 0.10, 2, 8, 0.05, seed=3)
    V, Fc = m.arrays()
    assert Fc.shape[0] == 20 * nu * nu and V.shape[0] == 10 * nu * nu + 2
    # the unperturbed ellipsoid radius along each vertex direction
    u = V / np.linalg.norm(V, axis=1)[:, None]
    r0 = 1.0 / np.sqrt((u[:, 0] / 0.08) ** 2 + (u[:, 1] / 0.07) ** 2 + (u[:, 2] / 0.10) ** 2)
    g = np.linalg.norm(V, axis=1) / r0 - 1.0
    assert np.sqrt(np.mean(g ** 2)) == pytest.approx(0.05, rel=1e-6)
    t = V[Fc]
    n = np.cross(t[:, 0] - t[:, 2], t[:, 1] - t[:, 2])
    assert np.all(np.einsum("ij,ij->i", n, t.mean(1)) > 0)
    V2, _ = F.Mesh.ellipsoid(nu, 0.08, 0.07, 0.10, 2, 8, 0.05, seed=3).arrays()
    np.testing.assert_array_equal(V, V2)


def test_mesh_from_arrays_validates():
    with pytest.raises(F.FmbError):
        F.Mesh.from_arrays(np.zeros((3, 3)), np.array([[0, 1, 5]]))


def _obj_error(path):
    with pytest.raises(F.FmbError) as e:
        F.Mesh.from_obj(str(path))
    assert e.value.code == -2
    return str(e.value)


def test_obj_large_polygon_is_refused_not_truncated(tmp_path):
    """A 300-gon exceeds the 256-vertex polygon limit: an error, never a
    partly filled face array."""
    n = 300
    th = np.linspace(0, 2 * np.pi, n, endpoint=False)
    lines = ["v %.6f %.6f 0" % (np.cos(t), np.sin(t)) for t in th]
    lines.append("f " + " ".join(str(i + 1) for i in range(n)))
    p = tmp_path / "poly300.obj"
    p.write_text("\n".join(lines) + "\n")
    assert "polygon" in _obj_error(p)
    # 256 vertices is the largest polygon accepted (254 fan triangles)
    lines = lines[:256] + ["f " + " ".join(str(i + 1) for i in range(256))]
    p.write_text("\n".join(lines) + "\n")
    V, Fc = F.Mesh.from_obj(str(p)).arrays()
    assert Fc.shape == (254, 3) and Fc.min() == 0 and Fc.max() == 255


def test_obj_truncated_and_malformed_files(tmp_path):
    good = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\nf 1 2 4\n"
    p = tmp_path / "t.obj"
    # cut inside the last face record: its third id is missing
    p.write_text(good[:-3])
    F.Mesh.from_obj(str(p))  # "f 1 2" is a degenerate record, skipped
    p.write_text(good[:-8] + "f 1 2 7\n")
    assert "vertex" in _obj_error(p)
    p.write_text(good.replace("f 1 2 4", "f 1 0 4"))  # ids are 1-based
    assert "malformed" in _obj_error(p)
    p.write_text(good.replace("f 1 2 4", "f 1 -9 4"))  # relative id before the first vertex
    _obj_error(p)
    p.write_text("v 0 0 0\nv 1 0\nv 0 1 0\nf 1 2 3\n")
    assert "malformed vertex" in _obj_error(p)
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3 " + "1 " * 3000 + "\n")
    assert "longer" in _obj_error(p)
    p.write_text(good.replace("f 1 2 4", "f -4 -3 -1"))  # relative ids
    V, Fc = F.Mesh.from_obj(str(p)).arrays()
    assert Fc.tolist() == [[0, 1, 2], [0, 1, 3]]
