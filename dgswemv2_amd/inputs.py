"""The reference's YAML input file (`dgswemv2_input.15`) read into this engine's run parameters,
mesh and output schedule, and its serial driver (read, build, step, write).

Follows InputParameters (preprocessor/input_parameters.hpp:96-261: mesh, timestepping,
polynomial order, problem, load balancer, output), SWE::Inputs (problem_input/swe_inputs.cpp:4-210:
gravity, densities, spherical projection, initial conditions, function source, bottom friction,
meteo, tide potential, Coriolis, wetting-drying, slope limiting, with the reference's defaults
of problem_input/swe_inputs.hpp:10-155) and Serial::Simulation (simulation/serial/simulation.hpp:
37-80: n_steps = ceil(run_time / dt), the writer called after every step).  Errors the reference
raises as std::logic_error raise ValueError with the reference's message; its warnings for
mal-formatted optional blocks (which fall back to the defaults) are collected in
`RunInput.warnings`.

Differences, each deliberate:
- the problem functions (initial condition, function source, truth) are compile-time in the
  reference (problem_function_files/*.hpp, chosen by the example built); here they are the
  `problem` argument (PROBLEM_NONE, MANUFACTURED, PARABOLIC_BOWL, DAM_BREAK);
- meteo forcing, tide potential and Coriolis are parsed and reported but have no effect, as in the
  reference, whose RKDG source term for them is commented out (swe_source.hpp:53-96);
- relative file names resolve against `base_dir` (default: the current directory, as in the
  reference).
"""
import math
import os
from dataclasses import dataclass, field

import yaml

from . import (FRICTION_CHEZY, FRICTION_MANNING, PROBLEM_NONE, RKDG_FP64_PLAIN, Simulation, make_params, read_adcirc,
               read_meta)

__all__ = ["OutputSpec", "RunInput", "read_input", "run", "run_time_seconds", "swe_inputs", "write_input"]


@dataclass
class OutputSpec:
    """WriterInput (preprocessor/input_parameters.hpp:198-259): steps between writes,
    ceil(frequency / dt); 0 = that output off."""
    path: str
    log_name: str = ""
    log_verbose: bool = False
    vtk_step: int = 0
    vtu_step: int = 0
    modal_step: int = 0


@dataclass
class RunInput:
    mesh: object
    params: object
    mesh_name: str
    run_time: float         # seconds, from the start and end dates (1 s granularity)
    n_steps: int            # ceil(run_time / dt)
    output: OutputSpec = None
    warnings: list = field(default_factory=list)
    unused: dict = field(default_factory=dict)  # parsed, without effect on the RKDG stage
    source: dict = field(default_factory=dict)  # names and values as the file gave them (write_input)


# ---- dates: strptime(s, "%d-%m-%Y %H:%M:%S", &tm) into a zeroed tm, then timegm

# field, digits, accepted range, the format character after it
_DATE_FIELDS = [("mday", 2, 1, 31, "-"), ("mon", 2, 1, 12, "-"), ("year", 4, 0, 9999, " "), ("hour", 2, 0, 23, ":"),
                ("min", 2, 0, 59, ":"), ("sec", 2, 0, 61, None)]


_DIGITS = "0123456789"


def _days_from_civil(y, m, d):
    # days since 1970-01-01 of the proleptic Gregorian date y-m-d; d may be 0 or exceed the month
    # (timegm normalises out-of-range fields the same way)
    y -= m <= 2
    era = (y if y >= 0 else y - 399) // 400
    yoe = y - era * 400
    doy = (153 * (m + (-3 if m > 2 else 9)) + 2) // 5
    doe = yoe * 365 + yoe // 4 - yoe // 100 + doy
    return era * 146097 + doe - 719468 + (d - 1)


def run_time_seconds(text):
    """Seconds since the epoch of a "dd-mm-yyyy hh:mm:ss" date as the reference reads it: the
    fields are taken in order until the text ends or stops matching (strptime leaves the rest of
    the zeroed tm alone, input_parameters.hpp:150-158), so "11-05-2015 12:00" is 12:00:00; a
    missing day is day 0 (the last day of the month before), a missing month January, a missing
    year 1900.  Each number may follow white space and takes digits while it has fewer than its
    width and ten times its value stays within its upper bound (so "86400" gives day 8); it must
    lie in its range, else it and the fields after it are not set.  A space in the format matches
    any amount of white space, none included."""
    tm = {"mday": 0, "mon": 1, "year": 1900, "hour": 0, "min": 0, "sec": 0}
    s = str(text)
    pos = 0
    for name, width, lo, hi, sep in _DATE_FIELDS:
        while pos < len(s) and s[pos].isspace():
            pos += 1
        if pos >= len(s) or s[pos] not in _DIGITS:
            break
        val, n = 0, 0
        while True:
            val = val * 10 + int(s[pos])
            pos, n = pos + 1, n + 1
            if not (n < width and val * 10 <= hi and pos < len(s) and s[pos] in _DIGITS):
                break
        if not lo <= val <= hi:
            break
        tm[name] = val
        if sep is None:
            break
        if sep == " ":
            while pos < len(s) and s[pos].isspace():
                pos += 1
        else:
            if pos >= len(s) or s[pos] != sep:
                break
            pos += 1
    return (_days_from_civil(tm["year"], tm["mon"], tm["mday"]) * 86400 + tm["hour"] * 3600 + tm["min"] * 60
            + tm["sec"])


# ---- helpers

def _num(node, key):
    v = node[key]
    if isinstance(v, bool):
        raise ValueError(f"Error: {key} is not a number\n")
    return float(v)


def _uint(node, key):
    v = node[key]
    if isinstance(v, bool) or int(v) != float(v) or int(v) < 0:
        raise ValueError(f"Error: {key} is not an unsigned integer\n")
    return int(v)


def _present(node, *keys):
    return isinstance(node, dict) and all(k in node and node[k] is not None for k in keys)


def _resolve(path, base_dir):
    return path if os.path.isabs(path) or base_dir is None else os.path.join(base_dir, path)


def _node_ids(mesh_path, fmt):
    """Node IDs of a .14 or .meta file in the order the readers index them (ascending ID)."""
    ids = []
    with open(mesh_path, "rb") as f:
        f.readline()
        if fmt == "Adcirc":
            nn = int(f.readline().split()[1])
        else:
            ne = int(f.readline().split()[0])
            for _ in range(ne):
                f.readline()
            nn = int(f.readline().split()[0])
        for _ in range(nn):
            ids.append(int(f.readline().split()[0]))
    return sorted(ids)


def _read_manning(path, node_ids):
    """Manning n per node from "node_ID n" lines, read until a line does not parse
    (swe_pre_init_data.hpp:166-184); nodes the file does not list get 0."""
    if not os.path.isfile(path):
        raise ValueError("Fatal Error: manning factor data file " + path + " was not found!\n")
    by_id = {}
    with open(path) as f:
        for line in f:
            tok = line.split()
            try:
                nid, n = int(tok[0]), float(tok[1])
            except (IndexError, ValueError):
                break
            if nid < 0:
                break
            by_id[nid] = n
    return [by_id.get(i, 0.0) for i in node_ids]


# ---- the input file

def swe_inputs(pr, problem=PROBLEM_NONE, problem_param=(0.0, 0.0, 0.0, 0.0), base_dir=None):
    """SWE::Inputs (problem_input/swe_inputs.cpp:4-210) of a `problem` node (a dict): the run
    parameters it sets, as make_params keywords, with the reference's defaults for what is
    absent; the warnings for mal-formatted optional blocks; the parsed values without effect on
    the RKDG stage; and the Manning data file (or None)."""
    if not _present(pr, "name"):
        raise ValueError("Error: problem name not specified\n")
    if str(pr["name"]) != "rkdg_swe":
        raise ValueError("Error: this engine runs the rkdg_swe problem, not " + str(pr["name"]) + "\n")
    warn = []
    unused = {}
    g = 9.81
    if _present(pr, "gravity"):
        g = _num(pr, "gravity")
    else:
        warn.append("Using default g\n")
    unused["density_air"] = _num(pr, "density_air") if _present(pr, "density_air") else 1.2250
    unused["density_water"] = _num(pr, "density_water") if _present(pr, "density_water") else 1000.0

    kw = dict(g=g, problem=problem, problem_param=problem_param)

    sph = {"type": False, "R": 6378200.0, "lon": 0.0, "lat": 0.0}
    if "spherical_projection" in pr:
        sp = pr["spherical_projection"]
        if _present(sp, "longitude_o", "latitude_o", "R"):
            sph = {"type": True, "R": _num(sp, "R"), "lon": _num(sp, "longitude_o"), "lat": _num(sp, "latitude_o")}
        else:
            warn.append("Warning: spherical projection inputs are mal-formatted. Using default parameters.\n")
    kw.update(sp_correction=sph["type"], sph_R=sph["R"], sph_lon_o=sph["lon"], sph_lat_o=sph["lat"])

    ic_bad = "Warning: initial conditions are mal-formatted. Using default parameters.\n"
    kw.update(ic_constant=True, ze0=0.0, qx0=0.0, qy0=0.0)
    if "initial_conditions" in pr:
        ic = pr["initial_conditions"]
        if _present(ic, "type"):
            if ic["type"] == "Constant":
                if _present(ic, "initial_surface_height", "initial_momentum_x", "initial_momentum_y"):
                    kw.update(ze0=_num(ic, "initial_surface_height"), qx0=_num(ic, "initial_momentum_x"),
                              qy0=_num(ic, "initial_momentum_y"))
                else:
                    warn.append(ic_bad)
            elif ic["type"] == "Function":
                if problem == PROBLEM_NONE:
                    raise ValueError("Error: Function initial conditions need a problem (its ic_q)\n")
                kw["ic_constant"] = False
            else:
                warn.append(ic_bad)
        else:
            warn.append(ic_bad)

    if "function_source" in pr:
        if pr["function_source"] == "Enable":
            if problem == PROBLEM_NONE:
                raise ValueError("Error: a function source needs a problem (its source_q)\n")
            kw["function_source"] = True
        else:
            warn.append("Warning: function source is mal-formatted. Using default parameters.\n")

    manning_file = None
    bf_bad = "Warning: bottom friction is mal-formatted. Using default parameters.\n"
    if "bottom_friction" in pr:
        bf = pr["bottom_friction"]
        kind = bf.get("type") if isinstance(bf, dict) else None
        if kind == "Chezy" and _present(bf, "coefficient"):
            kw.update(friction=FRICTION_CHEZY, friction_cf=_num(bf, "coefficient"))
        elif kind == "Manning" and _present(bf, "coefficient", "input_file"):
            kw.update(friction=FRICTION_MANNING, friction_cf=_num(bf, "coefficient"))
            manning_file = _resolve(str(bf["input_file"]), base_dir)
        else:
            warn.append(bf_bad)
        if kw.get("friction") and kw["friction_cf"] < 0.0:
            raise ValueError("Fatal Error: Chezy friction coefficient must be postive!\n")

    if "meteo_forcing" in pr:
        m = pr["meteo_forcing"]
        if _present(m, "type", "raw_input_file", "input_file", "frequency"):
            unused["meteo_forcing"] = dict(m)
        else:
            warn.append("Warning: meteo forcing is mal-formatted. Using default parameters.\n")
    if "tide_potential" in pr:
        t = pr["tide_potential"]
        if _present(t, "type") and t["type"] == "Test":
            unused["tide_potential"] = "Test"
        else:
            warn.append("Warning: tide potential forcing is mal-formatted. Using default parameters.\n")
    if "coriolis" in pr:
        if pr["coriolis"] == "Enable":
            unused["coriolis"] = "Enable"
        else:
            warn.append("Warning: coriolis forcing is mal-formatted. Using default parameters.\n")

    if "wetting_drying" in pr:
        wd = pr["wetting_drying"]
        if _present(wd, "h_o"):
            kw.update(wetting_drying=True, h_o=_num(wd, "h_o"))
        else:
            warn.append("Warning: wet-dry is mal-formatted. Using default parameters.\n")

    if "slope_limiting" in pr:
        sl = pr["slope_limiting"]
        if _present(sl, "type") and sl["type"] == "Cockburn-Shu" and _present(sl, "M", "nu"):
            kw.update(slope_limiting=True, sl_M=_num(sl, "M"), sl_nu=_num(sl, "nu"))
        else:
            warn.append("Warning: slope limiting is mal-formatted. Using default parameters.\n")

    return kw, warn, unused, manning_file


def read_input(path, problem=PROBLEM_NONE, problem_param=(0.0, 0.0, 0.0, 0.0), base_dir=None):
    """Reads a dgswemv2 YAML input file; returns a RunInput whose mesh and params go to
    Simulation (or Batch / PartitionedSimulation) unchanged."""
    with open(path) as f:
        doc = yaml.safe_load(f) or {}
    if not isinstance(doc, dict):
        raise ValueError("Error: the input file is not a YAML map\n")
    warn = []

    if "polynomial_order" not in doc:
        raise ValueError("Error: P YAML node not specified\n")
    p = _uint(doc, "polynomial_order")

    if "mesh" not in doc:
        raise ValueError("Error: Mesh YAML node not specified\n")
    mesh_node = doc["mesh"]
    if not _present(mesh_node, "format", "file_name", "coordinate_system"):
        raise ValueError("Error: Mesh YAML node is malformatted\n")
    fmt = str(mesh_node["format"])
    if fmt not in ("Adcirc", "Meta"):
        raise ValueError("Error: Unsupported mesh format: " + fmt + "\n")
    coord = str(mesh_node["coordinate_system"])
    if coord not in ("cartesian", "polar", "spherical"):
        raise ValueError("Error: Unsupported coordinate system: " + coord + "\n")
    mesh_file = _resolve(str(mesh_node["file_name"]), base_dir)
    stem = mesh_file[:mesh_file.rfind(".")] if "." in os.path.basename(mesh_file) else mesh_file
    bcis_file = stem + ".bcis"

    if "timestepping" not in doc:
        raise ValueError("Error: Timestepping YAML node not specified\n")
    ts = doc["timestepping"]
    if not _present(ts, "start_time", "end_time", "dt", "order", "nstages"):
        raise ValueError("Error: Timestepping YAML node is malformatted\n")
    run_time = float(run_time_seconds(ts["end_time"]) - run_time_seconds(ts["start_time"]))
    dt = _num(ts, "dt")
    order, nstages = _uint(ts, "order"), _uint(ts, "nstages")
    ramp_duration = _num(ts, "ramp_duration") if _present(ts, "ramp_duration") else 0.0

    if "problem" not in doc:
        raise ValueError("Error: Problem YAML node not specified\n")
    pr = doc["problem"]
    kw, pwarn, unused, manning_file = swe_inputs(pr, problem, problem_param, base_dir)
    warn += pwarn
    kw.update(p=p, nstages=nstages, order=order, dt=dt, ramp_duration=ramp_duration,
              spherical_mesh=coord == "spherical")
    if coord == "spherical" and not kw["sp_correction"]:
        warn.append("Warning: Spherical projection input not provided! Using default values to perfrom "
                    "projection!\n")

    if "load_balancer" in doc:
        lb = doc["load_balancer"]
        if not _present(lb, "name", "rebalance_frequency"):
            raise ValueError("Error: Load Balancer YAML node is malformatted\n")
        unused["load_balancer"] = {"name": str(lb["name"]), "rebalance_frequency": _num(lb, "rebalance_frequency")}

    source = {"mesh": {"format": fmt, "file_name": str(mesh_node["file_name"]), "coordinate_system": coord},
              "start_time": run_time_seconds(ts["start_time"]), "end_time": run_time_seconds(ts["end_time"]),
              "manning_file": str(pr["bottom_friction"]["input_file"]) if manning_file else None, "frequency": {}}
    output = None
    if "output" in doc:
        out = doc["output"]
        source["output_path"] = opath_given = str(out["path"]) if _present(out, "path") else None
        if not _present(out, "path"):
            raise ValueError("Error: Output YAML node is malformatted\n")
        opath = opath_given
        output = OutputSpec(path=_resolve(opath if opath.endswith("/") else opath + "/", base_dir))
        if "logfile" in out:
            lf = out["logfile"]
            if not _present(lf, "verbose", "name"):
                raise ValueError("Error: Logfile YAML node is malformatted\n")
            output.log_name, output.log_verbose = str(lf["name"]), bool(lf["verbose"])
        for kind in ("vtk", "vtu", "modal"):
            if kind in out:
                node = out[kind]
                if not _present(node, "frequency"):
                    raise ValueError(f"Error: {kind.upper() if kind != 'modal' else 'Modal'} YAML node is "
                                     "malformatted\n")
                source["frequency"][kind] = _num(node, "frequency")
                step = math.ceil(_num(node, "frequency") / dt)
                if step < 1:
                    raise ValueError(f"Error: {kind} output frequency must be positive\n")
                setattr(output, kind + "_step", step)

    if fmt == "Adcirc":
        mesh = read_adcirc(mesh_file, bcis_file if os.path.isfile(bcis_file) else None)
    else:
        mesh = read_meta(mesh_file, bcis_file if os.path.isfile(bcis_file) else None)
    if manning_file is not None:
        mesh = mesh.with_manning(_read_manning(manning_file, _node_ids(mesh_file, fmt)))
    with open(mesh_file, "rb") as f:  # `ifs >> name` (adcirc_format.cpp:10, mesh_metadata.cpp:94)
        first = f.readline().decode(errors="replace").split()
    mesh_name = first[0] if first else ""

    params = make_params(**kw)
    n_steps = math.ceil(run_time / dt) if dt > 0 else 0
    return RunInput(mesh=mesh, params=params, mesh_name=mesh_name, run_time=run_time, n_steps=n_steps,
                    output=output, warnings=warn, unused=unused, source=source)


def write_input(inp, path):
    """Writes a RunInput as a YAML input file (InputParameters::write_to, input_parameters.hpp:
    296-391, with SWE::Inputs::as_yaml_node, swe_inputs.cpp:316-431): the run parameters as read,
    the file names as the input gave them, the dates normalised as timegm leaves them; reading the
    written file gives the same run."""
    import time
    p, src, un = inp.params, inp.source, inp.unused

    def date(secs):
        return time.strftime("%d-%m-%Y %H:%M:%S", time.gmtime(secs))

    prob = {"name": "rkdg_swe", "gravity": p.g, "density_air": un.get("density_air", 1.2250),
            "density_water": un.get("density_water", 1000.0)}
    if p.sp_correction:
        prob["spherical_projection"] = {"type": "Enable", "longitude_o": p.sph_lon_o, "latitude_o": p.sph_lat_o,
                                        "R": p.sph_R}
    if p.ic_constant:
        prob["initial_conditions"] = {"type": "Constant", "initial_surface_height": p.ze0, "initial_momentum_x": p.qx0,
                                      "initial_momentum_y": p.qy0}
    else:
        prob["initial_conditions"] = {"type": "Function"}
    if p.function_source:
        prob["function_source"] = "Enable"
    if p.friction == FRICTION_CHEZY:
        prob["bottom_friction"] = {"type": "Chezy", "coefficient": p.friction_cf}
    elif p.friction == FRICTION_MANNING:
        prob["bottom_friction"] = {"type": "Manning", "coefficient": p.friction_cf, "input_file": src["manning_file"]}
    for key in ("meteo_forcing", "tide_potential", "coriolis"):
        if key in un:
            prob[key] = {"type": un[key]} if key == "tide_potential" else un[key]
    if p.wetting_drying:
        prob["wetting_drying"] = {"h_o": p.h_o}
    if p.slope_limiting:
        prob["slope_limiting"] = {"type": "Cockburn-Shu", "M": p.sl_M, "nu": p.sl_nu}
    doc = {"polynomial_order": int(p.p), "mesh": dict(src["mesh"]),
           "timestepping": {"start_time": date(src["start_time"]), "end_time": date(src["end_time"]), "dt": p.dt,
                            "order": int(p.order), "nstages": int(p.nstages), "ramp_duration": p.ramp_duration},
           "problem": prob}
    if inp.output is not None:
        out = {"path": src["output_path"]}
        if inp.output.log_name:
            out["logfile"] = {"name": inp.output.log_name, "verbose": bool(inp.output.log_verbose)}
        for kind, f in src["frequency"].items():
            out[kind] = {"frequency": f}
        doc["output"] = out
    if "load_balancer" in un:
        doc["load_balancer"] = dict(un["load_balancer"])
    with open(path, "w") as f:
        f.write("#" * 79 + "\n#\n#  DG_HYPER input file\n#\n" + "#" * 79 + "\n\n")
        f.write(yaml.safe_dump(doc, sort_keys=False))


def run(inp, precision=RKDG_FP64_PLAIN, n_steps=None):
    """Serial::Simulation::Run (simulation.hpp:67-80) on the device: the first output at step 0,
    then n_steps steps, each followed by the writers whose step count divides the step number
    (Writer::WriteOutput, simulation/writer.hpp:121-190).  Returns the Simulation."""
    sim = Simulation(inp.mesh, inp.params, precision)
    out = inp.output
    prefix = None
    log = None
    if out is not None:
        os.makedirs(out.path, exist_ok=True)
        prefix = out.path + inp.mesh_name
        if out.log_name:
            log = open(out.path + out.log_name, "w")
            log.write(f"Starting simulation with p={inp.params.p} for {inp.mesh_name} mesh\n\n")
            log.write("\nLaunching Simulation!\n\n")

    def write(step):
        if out is None:
            return
        if out.vtk_step and step % out.vtk_step == 0:
            sim.write_vtk(prefix)
        if out.vtu_step and step % out.vtu_step == 0:
            sim.write_vtu(prefix)
        if out.modal_step and step % out.modal_step == 0:
            sim.write_modal(prefix)

    try:
        write(0)
        n = inp.n_steps if n_steps is None else n_steps
        for step in range(1, n + 1):
            sim.step(1)
            write(step)
            if log is not None and out.log_verbose:
                log.write(f"Step {step} done\n")
    finally:
        if log is not None:
            log.close()
    return sim
