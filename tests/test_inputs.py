"""The reference's YAML input file read into run parameters, mesh and output schedule
(dgswemv2_amd.inputs), after the reference's own input tests: test_input_parameters.cpp (a correct
file, a file that does not exist, a file missing a field) and test_swe_inputs.cpp (defaults, every
field set, a negative friction coefficient, function initial conditions).

Fixtures: the reference's weir input (test/files_for_testing/weir/dgswemv2_input.15), its
missing_field_input.15, and the correct_input.15 that its test/CMakeLists.txt:330-378 generates
(written here with the same fields).  Dates are pinned against the C library's own strptime and
timegm on this machine.
"""
import calendar
import ctypes as C
import ctypes.util
import math
from pathlib import Path

import numpy as np
import pytest

from dgswemv2_amd import (FRICTION_CHEZY, FRICTION_MANNING, FRICTION_NONE, PROBLEM_MANUFACTURED, RKDG_FP64_PLAIN,
                          Simulation, read_adcirc)
from dgswemv2_amd import inputs as I

GOLDEN = Path(__file__).resolve().parent / "golden"
WEIR = GOLDEN / "weir"

CORRECT_INPUT = """mesh:
  format: Adcirc
  file_name: {fort14}
  coordinate_system: cartesian

timestepping:
  start_time: 11-05-2015 12:00              #dd-mm-yyyy hh:mm
  end_time: 12-05-2015 00:00                #dd-mm-yyyy hh:mm
  dt: 1                           #in seconds
  order: 2
  nstages: 2
  ramp_duration: 1.5

polynomial_order: 2

problem:
  name: rkdg_swe
  gravity: 9.81
  density_air: 1.225
  density_water: 1000
  h_o: 0.001
  initial_conditions:
    type: Constant
    initial_surface_height: 10.0
    initial_momentum_x: 0.0
    initial_momentum_y: 0.0
  function_source: Enable
  bottom_friction:
    type: Manning
    coefficient: 0.001
    input_file: manning.test
  meteo_forcing:
    type: HWIND
    raw_input_file: weather.test
    input_file: weather.test
    frequency: 900
  tide_potential:
    type: Test
  coriolis: Enable
  wetting_drying:
    h_o: 0.1
  slope_limiting:
    type: Cockburn-Shu
    M: 50
    nu: 1

output:
  logfile:
    name: log
    verbose: true
  path: output/
  vtk:
    frequency: 100

load_balancer:
  name: random
  rebalance_frequency: 10"""


def test_weir_input_file():
    r = I.read_input(WEIR / "dgswemv2_input.15", base_dir=str(WEIR))
    p = r.params
    # dgswemv2_input.15: p = 1, SSP-RK2, dt = 3 s over 2 h, g = 9.81, Cockburn-Shu M = 1e-8, nu = 1.5
    assert (p.p, p.nstages, p.order, p.dt, p.g, p.ramp_duration) == (1, 2, 2, 3.0, 9.81, 0.0)
    assert p.slope_limiting == 1 and p.sl_M == 1e-8 and p.sl_nu == 1.5
    assert p.wetting_drying == 0 and p.friction == FRICTION_NONE and p.function_source == 0
    assert p.ic_constant == 1 and (p.ze0, p.qx0, p.qy0) == (0.0, 0.0, 0.0)
    assert p.spherical_mesh == 0 and p.sp_correction == 0
    assert r.run_time == 7200.0 and r.n_steps == 2400
    assert r.mesh_name == "quadrilateral" and r.output is None and r.warnings == []
    assert r.unused == {"density_air": 1.225, "density_water": 1000.0}
    m = read_adcirc(WEIR / "weir.14", WEIR / "weir.bcis")
    for f in ("node_xyz", "elt_nodes", "bound_edges", "bound_types", "frequency", "node_amplitude", "levee_nodes",
              "levee_param"):
        assert np.array_equal(getattr(r.mesh, f), getattr(m, f)), f


def test_reference_correct_input(tmp_path):
    (tmp_path / "in.15").write_text(CORRECT_INPUT.format(fort14=GOLDEN / "sample_fort.14"))
    # Manning n by node ID, in no particular order, node 7 missing (it gets 0, as std::map's default)
    ids = [i for i in range(1, 131) if i != 7][::-1]
    (tmp_path / "manning.test").write_text("".join(f"{i} {0.01 + 1e-4 * i}\n" for i in ids) + "end\n999 5\n")
    r = I.read_input(tmp_path / "in.15", problem=PROBLEM_MANUFACTURED, base_dir=str(tmp_path))
    p = r.params
    assert (p.p, p.nstages, p.order, p.dt, p.ramp_duration, p.g) == (2, 2, 2, 1.0, 1.5, 9.81)
    assert r.run_time == 43200.0 and r.n_steps == 43200  # 12:00 to 00:00 (the ":%S" field absent)
    assert p.ic_constant == 1 and (p.ze0, p.qx0, p.qy0) == (10.0, 0.0, 0.0)
    assert p.function_source == 1 and p.problem == PROBLEM_MANUFACTURED
    assert p.friction == FRICTION_MANNING and p.friction_cf == 0.001
    n = r.mesh.node_manning_n
    assert n[6] == 0.0 and np.allclose(np.delete(n, 6), [0.01 + 1e-4 * i for i in range(1, 131) if i != 7], rtol=0)
    assert p.wetting_drying == 1 and p.h_o == 0.1  # the problem-level h_o is not read
    assert p.slope_limiting == 1 and p.sl_M == 50.0 and p.sl_nu == 1.0
    o = r.output
    assert o.path == str(tmp_path / "output") + "/" and o.log_name == "log" and o.log_verbose
    assert (o.vtk_step, o.vtu_step, o.modal_step) == (100, 0, 0)
    assert r.unused["load_balancer"] == {"name": "random", "rebalance_frequency": 10.0}
    assert r.unused["meteo_forcing"]["type"] == "HWIND" and r.unused["tide_potential"] == "Test"
    assert r.unused["coriolis"] == "Enable"
    assert r.mesh_name == "ADCIRC" and r.mesh.n_elements == 200


def test_reference_missing_field_input():
    with pytest.raises(ValueError, match="P YAML node not specified"):
        I.read_input(GOLDEN / "missing_field_input.15")


def test_input_file_that_does_not_exist(tmp_path):
    with pytest.raises(OSError):
        I.read_input(tmp_path / "no_such_input.15")


@pytest.mark.parametrize("drop,msg", [("mesh", "Mesh YAML node not specified"),
                                      ("timestepping", "Timestepping YAML node not specified"),
                                      ("problem", "Problem YAML node not specified")])
def test_missing_required_nodes(tmp_path, drop, msg):
    import yaml
    doc = yaml.safe_load((WEIR / "dgswemv2_input.15").read_text())
    del doc[drop]
    (tmp_path / "in.15").write_text(yaml.safe_dump(doc))
    with pytest.raises(ValueError, match=msg):
        I.read_input(tmp_path / "in.15", base_dir=str(WEIR))


@pytest.mark.parametrize("edit,msg", [
    (lambda d: d["mesh"].update(format="Gmsh"), "Unsupported mesh format: Gmsh"),
    (lambda d: d["mesh"].update(coordinate_system="mercator"), "Unsupported coordinate system"),
    (lambda d: d["mesh"].pop("file_name"), "Mesh YAML node is malformatted"),
    (lambda d: d["timestepping"].pop("dt"), "Timestepping YAML node is malformatted"),
    (lambda d: d.update(output={"vtu": {"frequency": 10}}), "Output YAML node is malformatted"),
    (lambda d: d.update(output={"path": "o", "vtu": {}}), "VTU YAML node is malformatted"),
    (lambda d: d.update(output={"path": "o", "logfile": {"name": "x"}}), "Logfile YAML node is malformatted"),
    (lambda d: d.update(load_balancer={"name": "random"}), "Load Balancer YAML node is malformatted"),
    (lambda d: d["problem"].update(name="ehdg_swe"), "rkdg_swe"),
])
def test_malformed_required_fields(tmp_path, edit, msg):
    import yaml
    doc = yaml.safe_load((WEIR / "dgswemv2_input.15").read_text())
    edit(doc)
    (tmp_path / "in.15").write_text(yaml.safe_dump(doc))
    with pytest.raises(ValueError, match=msg):
        I.read_input(tmp_path / "in.15", base_dir=str(WEIR))


# ---- SWE::Inputs, after test_swe_inputs.cpp

def test_swe_inputs_defaults():
    kw, warn, unused, manning = I.swe_inputs({"name": "rkdg_swe"})
    assert kw["g"] == 9.81 and "friction" not in kw and kw["ic_constant"] and (kw["ze0"], kw["qx0"], kw["qy0"]) == (0, 0, 0)
    prm = I.make_params(**kw)
    assert prm.friction == FRICTION_NONE and prm.friction_cf == 0.0 and prm.wetting_drying == 0
    assert prm.slope_limiting == 0 and prm.sl_M == 1e-8 and prm.sl_nu == 1.5 and prm.h_o == 0.1
    assert warn == ["Using default g\n"] and manning is None


def test_swe_inputs_every_field():
    kw, warn, _, _ = I.swe_inputs({"name": "rkdg_swe", "gravity": 1.0, "bottom_friction": {"type": "Chezy", "coefficient": 0.01},
                                   "initial_conditions": {"type": "Constant", "initial_surface_height": 1.0,
                                                          "initial_momentum_x": 2.0, "initial_momentum_y": 3.0}})
    assert kw["g"] == 1.0 and kw["friction"] == FRICTION_CHEZY and kw["friction_cf"] == 0.01
    assert (kw["ze0"], kw["qx0"], kw["qy0"]) == (1.0, 2.0, 3.0) and warn == []


def test_swe_inputs_negative_friction():
    with pytest.raises(ValueError, match="friction coefficient must be postive"):
        I.swe_inputs({"name": "rkdg_swe", "bottom_friction": {"type": "Chezy", "coefficient": -0.01}})


def test_swe_inputs_function_initial_conditions():
    kw, _, _, _ = I.swe_inputs({"name": "rkdg_swe", "initial_conditions": {"type": "Function"}},
                               problem=PROBLEM_MANUFACTURED)
    assert kw["ic_constant"] is False
    with pytest.raises(ValueError, match="need a problem"):
        I.swe_inputs({"name": "rkdg_swe", "initial_conditions": {"type": "Function"}})


def test_swe_inputs_malformed_blocks_fall_back_with_warnings():
    kw, warn, _, _ = I.swe_inputs({"name": "rkdg_swe", "gravity": 9.81, "bottom_friction": {"type": "Darcy"},
                                   "wetting_drying": {}, "slope_limiting": {"type": "Cockburn-Shu", "M": 1},
                                   "spherical_projection": {"R": 1.0}, "initial_conditions": {"type": "Constant"}})
    assert "friction" not in kw and "wetting_drying" not in kw and "slope_limiting" not in kw
    assert not kw["sp_correction"] and kw["ze0"] == 0.0
    assert len(warn) == 5 and all("mal-formatted" in w for w in warn)


def test_swe_inputs_spherical_projection():
    kw, _, _, _ = I.swe_inputs({"name": "rkdg_swe", "spherical_projection": {"longitude_o": -90.0, "latitude_o": 29.5,
                                                                             "R": 6378206.4}})
    assert kw["sp_correction"] and (kw["sph_lon_o"], kw["sph_lat_o"], kw["sph_R"]) == (-90.0, 29.5, 6378206.4)


# ---- dates against the C library

def _libc_seconds(text):
    class Tm(C.Structure):
        _fields_ = [(n, C.c_int) for n in ("sec", "min", "hour", "mday", "mon", "year", "wday", "yday", "isdst")] + \
                   [("gmtoff", C.c_long), ("zone", C.c_char_p)]
    libc = C.CDLL(ctypes.util.find_library("c"))
    libc.strptime.restype = C.c_char_p
    libc.strptime.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Tm)]
    libc.timegm.restype = C.c_longlong
    libc.timegm.argtypes = [C.POINTER(Tm)]
    tm = Tm()
    libc.strptime(text.encode(), b"%d-%m-%Y %H:%M:%S", C.byref(tm))
    return libc.timegm(C.byref(tm))


@pytest.mark.parametrize("text", ["25-11-1987 12:00:00", "25-11-1987 14:00:00", "11-05-2015 12:00", "12-05-2015 00:00",
                                  "01-03-2016 00:00:01", "29-02-2016 23:59:59", "31-12-1999 23:59:59", "1-1-1970 0:0:0",
                                  "15-06-2020", "15-06-2020 7", "07-08-2021 10:30:4", "86400", "35-01-2020",
                                  "2020-05-11", " 5- 6-2020  3:04:05", "05-06-20203:04:05", "00-01-2000", "31-13-2000",
                                  "10-10-10 10:10:10", ""])
def test_dates_as_the_c_library_reads_them(text):
    assert I.run_time_seconds(text) == _libc_seconds(text)


def test_dates_closed_form():
    assert I.run_time_seconds("25-11-1987 12:00:00") == calendar.timegm((1987, 11, 25, 12, 0, 0, 0, 0, 0))
    assert I.run_time_seconds("12-05-2015 00:00") - I.run_time_seconds("11-05-2015 12:00") == 43200


# ---- the serial driver on the device

@pytest.mark.gpu
def test_run_weir_input_file_matches_a_direct_run(tmp_path):
    import yaml
    doc = yaml.safe_load((WEIR / "dgswemv2_input.15").read_text())
    doc["output"] = {"path": str(tmp_path / "out"), "vtu": {"frequency": 15}, "modal": {"frequency": 30},
                     "vtk": {"frequency": 60}, "logfile": {"name": "log", "verbose": True}}
    (tmp_path / "in.15").write_text(yaml.safe_dump(doc))
    r = I.read_input(tmp_path / "in.15", base_dir=str(WEIR))
    assert (r.output.vtu_step, r.output.modal_step) == (5, 10)
    sim = I.run(r, RKDG_FP64_PLAIN, n_steps=20)
    ref = Simulation(r.mesh, r.params, RKDG_FP64_PLAIN)
    ref.step(20)
    a, b = sim.state(), ref.state()
    for f in ("q", "wet", "limited"):
        assert np.array_equal(a[f], b[f]), f
    out = tmp_path / "out"
    assert sorted(x.name for x in out.glob("*.vtu")) == sorted(f"quadrilateral_data_{s}.vtu" for s in (0, 5, 10, 15, 20))
    assert sorted(x.name for x in out.glob("*.vtk")) == sorted(f"quadrilateral_data_{s}.vtk" for s in (0, 20))
    ze = (out / "quadrilateral_modal_ze.txt").read_text().split("\n")
    assert (out / "log").read_text().startswith("Starting simulation with p=1 for quadrilateral mesh")
    assert sum(1 for line in ze if line and len(line.split()) == 1) == 3  # records at steps 0, 10, 20
    assert math.isfinite(float(ze[0]))


def test_meta_format_input(tmp_path):
    # format: Meta reads a MeshMetaData file (mesh_metadata.cpp:87-137) with the .bcis beside it
    import shutil
    import yaml
    from dgswemv2_amd import read_meta, write_meta
    m = read_adcirc(WEIR / "weir.14", WEIR / "weir.bcis")
    write_meta(m, tmp_path / "weir.meta", name="weirmeta")
    shutil.copy(WEIR / "weir.bcis", tmp_path / "weir.bcis")
    doc = yaml.safe_load((WEIR / "dgswemv2_input.15").read_text())
    doc["mesh"].update(format="Meta", file_name="weir.meta")
    (tmp_path / "in.15").write_text(yaml.safe_dump(doc))
    r = I.read_input(tmp_path / "in.15", base_dir=str(tmp_path))
    ref = read_meta(tmp_path / "weir.meta", tmp_path / "weir.bcis")
    assert r.mesh_name == "weirmeta"
    for f in ("node_xyz", "elt_nodes", "bound_edges", "bound_types", "levee_nodes", "levee_param"):
        assert np.array_equal(getattr(r.mesh, f), getattr(ref, f)), f


def test_spherical_coordinates_input(tmp_path):
    import yaml
    doc = yaml.safe_load((WEIR / "dgswemv2_input.15").read_text())
    doc["mesh"]["coordinate_system"] = "spherical"
    (tmp_path / "a.15").write_text(yaml.safe_dump(doc))
    r = I.read_input(tmp_path / "a.15", base_dir=str(WEIR))
    # without a projection block: projected with the defaults, correction off, the reference's warning
    assert r.params.spherical_mesh == 1 and r.params.sp_correction == 0 and r.params.sph_R == 6378200.0
    assert any("Spherical projection input not provided" in w for w in r.warnings)
    doc["problem"]["spherical_projection"] = {"longitude_o": 10.0, "latitude_o": 45.0, "R": 6.4e6}
    (tmp_path / "b.15").write_text(yaml.safe_dump(doc))
    r = I.read_input(tmp_path / "b.15", base_dir=str(WEIR))
    p = r.params
    assert p.spherical_mesh == 1 and p.sp_correction == 1 and (p.sph_lon_o, p.sph_lat_o, p.sph_R) == (10.0, 45.0, 6.4e6)
    assert r.warnings == []


def _params_equal(a, b):
    return all(getattr(a, f) == getattr(b, f) if f != "problem_param" else list(a.problem_param) == list(b.problem_param)
               for f, _ in type(a)._fields_)


@pytest.mark.parametrize("which", ["correct", "weir"])
def test_written_input_reads_back_the_same(tmp_path, which):
    # test_input_parameters.cpp: read a correct file, write it (write_to), read the written file,
    # and the two inputs are equal
    if which == "correct":
        (tmp_path / "in.15").write_text(CORRECT_INPUT.format(fort14=GOLDEN / "sample_fort.14"))
        (tmp_path / "manning.test").write_text("".join(f"{i} 0.02\n" for i in range(1, 131)))
        kw = dict(problem=PROBLEM_MANUFACTURED, base_dir=str(tmp_path))
        src = tmp_path / "in.15"
    else:
        kw = dict(base_dir=str(WEIR))
        src = WEIR / "dgswemv2_input.15"
    a = I.read_input(src, **kw)
    I.write_input(a, tmp_path / "emitted.15")
    b = I.read_input(tmp_path / "emitted.15", **kw)
    assert _params_equal(a.params, b.params)
    assert (a.run_time, a.n_steps, a.mesh_name, a.output, a.unused, a.source) == \
        (b.run_time, b.n_steps, b.mesh_name, b.output, b.unused, b.source)
    for f in ("node_xyz", "elt_nodes", "bound_types"):
        assert np.array_equal(getattr(a.mesh, f), getattr(b.mesh, f))
    assert (tmp_path / "emitted.15").read_text().startswith("#" * 79)
