"""Output writers (rkdg_write_modal, rkdg_write_vtu) in the reference's formats
(swe_post_write_modal.hpp:5-90; swe_post_write_vtu.hpp:5-103 + simulation/writer.hpp:260-350),
read back and checked against the state they were written from.  Point and cell values are
checked against the Dubiner expansion evaluated by the oracle (orc_dubiner) at the master
triangle's vertices and centroid."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from dgswemv2_amd import RKDG_FP64_PLAIN, RkdgError, Simulation
from dgswemv2_amd import workloads as W
from oracle.oracle_py import load as load_oracle
import dgswemv2_amd._native as N

pytestmark = pytest.mark.gpu


def dubiner_at(p, pts):
    nd = (p + 1) * (p + 2) // 2
    z = np.ascontiguousarray(np.array(pts, dtype=np.float64).reshape(-1))
    phi = np.empty(nd * len(pts))
    d1 = np.empty_like(phi)
    d2 = np.empty_like(phi)
    load_oracle().orc_dubiner(p, len(pts), N.dptr(z), N.dptr(phi), N.dptr(d1), N.dptr(d2))
    return phi.reshape(nd, len(pts))


def test_modal_dump(tmp_path):
    mesh, prm = W.bowl(scale=0.1)
    s = Simulation(mesh, prm, RKDG_FP64_PLAIN)
    pre = str(tmp_path / "bowl")
    s.write_modal(pre)
    s.step(2)
    s.write_modal(pre)
    q = s.state()["q"]
    ne, nd = mesh.n_elements, 3
    for v, name in enumerate(("ze", "qx", "qy", "bath")):
        lines = (tmp_path / f"bowl_modal_{name}.txt").read_text().splitlines()
        assert len(lines) == 2 * (1 + ne * nd)
        assert lines[0] == "0.000000" and lines[1 + ne * nd] == "%f" % (2 * prm.dt)
        rec = lines[2 + ne * nd:]
        ids = np.array([int(r.split()[0]) for r in rec])
        vals = np.array([float(r.split()[1]) for r in rec])
        assert np.array_equal(ids, np.repeat(np.arange(ne), nd))
        if v < 3:
            want = q[v].T.reshape(-1)  # element-major, then mode
            assert all(r.split()[1] == "%.6e" % x for r, x in zip(rec[:50], want[:50]))
            assert np.allclose(vals, want, rtol=1e-6, atol=0)
    s.close()


def test_vtu(tmp_path):
    mesh, prm = W.bowl(scale=0.1)
    s = Simulation(mesh, prm, RKDG_FP64_PLAIN)
    s.step(3)
    pre = str(tmp_path / "bowl")
    s.write_vtu(pre)
    st = s.state()
    q, wet = st["q"], st["wet"]
    ne, p = mesh.n_elements, prm.p
    root = ET.parse(tmp_path / "bowl_data_3.vtu").getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == str(3 * ne) and piece.get("NumberOfCells") == str(ne)
    arr = {d.get("Name") or "points": np.array(d.text.split(), dtype=float) for d in piece.iter("DataArray")}
    phi = dubiner_at(p, [(-1, -1), (1, -1), (-1, 1), (-1 / 3, -1 / 3)])
    for v, name in enumerate(("ze", "qx", "qy")):
        pv = np.einsum("de,dk->ek", q[v], phi[:, :3]).reshape(-1)
        cv = q[v].T @ phi[:, 3]
        assert np.allclose(arr[f"{name}_point"], pv.astype(np.float32), rtol=6e-6, atol=1e-30)  # 6 significant digits, as the reference prints
        assert np.allclose(arr[f"{name}_cell"], cv.astype(np.float32), rtol=6e-6, atol=1e-30)  # 6 significant digits, as the reference prints
    assert np.array_equal(arr["ID"], np.arange(ne))
    assert np.array_equal(arr["wet_dry"], wet.astype(float))
    assert np.array_equal(arr["connectivity"], np.arange(3 * ne))
    assert np.array_equal(arr["offsets"], 3 * np.arange(1, ne + 1)) and np.all(arr["types"] == 5)
    # points: each element's three vertices (counter-clockwise), z = 0
    pts = arr["points"].reshape(ne, 3, 3)
    assert np.all(pts[:, :, 2] == 0)
    xy = mesh.node_xyz[mesh.elt_nodes][:, :, :2]
    for e in range(ne):  # the same three vertices, to the 6 significant digits printed
        d = np.abs(pts[e, :, None, :2] - xy[e, None, :, :]).max(axis=2)
        tol = 6e-6 * np.abs(xy[e]).max() + 1e-9
        assert np.all(d.min(axis=1) <= tol) and np.all(d.min(axis=0) <= tol), e
    area2 = ((pts[:, 1, 0] - pts[:, 0, 0]) * (pts[:, 2, 1] - pts[:, 0, 1]) -
             (pts[:, 2, 0] - pts[:, 0, 0]) * (pts[:, 1, 1] - pts[:, 0, 1]))
    assert np.all(area2 > 0)
    # went_completely_dry: mean vertex depth of the linear part at or below h_o's threshold
    # (1e6 h_o eps, swe_pre_init_prob_params.hpp:36)
    h_thr = 1e6 * prm.h_o * np.finfo(float).eps
    ze_v = np.einsum("de,dk->ek", q[0][:3], phi[:3, :3])
    hbar = (ze_v.sum(axis=1) + mesh.node_xyz[mesh.elt_nodes][:, :, 2].sum(axis=1)) / 3
    assert np.array_equal(arr["went_completely_dry"], (hbar <= h_thr).astype(float))
    assert (wet == 0).any()
    s.close()


def test_output_at_step_boundary_only(tmp_path):
    mesh, prm = W.bowl(scale=0.05)
    s = Simulation(mesh, prm, RKDG_FP64_PLAIN)
    s.stage()
    with pytest.raises(RkdgError):
        s.write_vtu(str(tmp_path / "x"))
    with pytest.raises(RkdgError):
        s.write_modal(str(tmp_path / "x"))
    s.stage()  # back at a step boundary (SSP-RK2)
    with pytest.raises(RkdgError):
        s.write_modal(str(tmp_path / "missing_dir" / "x"))
    s.write_vtu(str(tmp_path / "x"))
    assert (tmp_path / "x_data_1.vtu").exists()
    s.close()


def test_vtk_legacy(tmp_path):
    # the VTK legacy file carries the VTU's points, cells and values in the layout of
    # Writer::InitializeMeshGeometryVTK + write_VTK_data (simulation/writer.hpp:195-258,
    # swe_post_write_vtk.hpp:5-95): same numbers, same six-significant-digit printing
    mesh, prm = W.bowl(scale=0.1)
    s = Simulation(mesh, prm, RKDG_FP64_PLAIN)
    s.step(3)
    pre = str(tmp_path / "bowl")
    s.write_vtu(pre)
    s.write_vtk(pre)
    ne = mesh.n_elements
    piece = ET.parse(tmp_path / "bowl_data_3.vtu").getroot().find("UnstructuredGrid/Piece")
    vtu = {d.get("Name") or "points": d.text.split() for d in piece.iter("DataArray")}
    lines = (tmp_path / "bowl_data_3.vtk").read_text().split("\n")
    assert lines[:5] == ["# vtk DataFile Version 3.0", "OUTPUT DATA", "ASCII", "DATASET UNSTRUCTURED_GRID",
                         f"POINTS {3 * ne} double"]
    pts = lines[5:5 + 3 * ne]
    assert [t for l in pts for t in l.split("\t")] == vtu["points"]
    i = 5 + 3 * ne
    assert lines[i] == f"CELLS {ne} {4 * ne}"
    assert lines[i + 1:i + 1 + ne] == [f"3\t{3 * e}\t{3 * e + 1}\t{3 * e + 2}\t" for e in range(ne)]
    i += 1 + ne
    assert lines[i] == f"CELL_TYPES {ne}" and lines[i + 1:i + 1 + ne] == ["5"] * ne
    i += 1 + ne
    assert lines[i] == f"CELL_DATA {ne}"
    i += 1
    kinds = {"ID": "unsigned_int", "wet_dry": "bit", "went_completely_dry": "bit"}
    for name in ("ze_cell", "qx_cell", "qy_cell", "bath_cell", "ID", "wet_dry", "went_completely_dry"):
        assert lines[i] == f"SCALARS {name} {kinds.get(name, 'double')} 1" and lines[i + 1] == "LOOKUP_TABLE default"
        assert lines[i + 2:i + 2 + ne] == vtu[name], name
        i += 2 + ne
    assert lines[i] == f"POINT_DATA {3 * ne}"
    i += 1
    for name in ("ze_point", "qx_point", "qy_point", "bath_point"):
        assert lines[i] == f"SCALARS {name} double 1" and lines[i + 1] == "LOOKUP_TABLE default"
        assert lines[i + 2:i + 2 + 3 * ne] == vtu[name], name
        i += 2 + 3 * ne
    assert lines[i:] == [""]
    s.close()
