"""librkdg.so on the host: it loads, exports every entry point of include/rkdg.h, and its
host-only helpers behave (no device work here)."""
import ctypes as C
import re
import sys
from pathlib import Path

import numpy as np
import pytest

from dgswemv2_amd import _native as N

HEADER = Path(__file__).resolve().parent.parent / "include" / "rkdg.h"


def declared():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rkdg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "rkdg_create" in names and "rkdg_step" in names and len(names) >= 18


def test_library_exports_every_declared_symbol():
    lib = N.load_library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) == set(N.SIGNATURES), "ctypes table and header disagree"


def test_create_rejects_bad_arguments():
    lib = N.load_library()
    h = C.c_void_p()
    assert lib.rkdg_create(None, None, 0, C.byref(h)) == N.RKDG_ERR_ARG


def _set_array(desc, field, values, ctype, keep):
    a = np.ascontiguousarray(values)
    keep.append(a)
    setattr(desc, field, a.ctypes.data_as(C.POINTER(ctype)))


def _elements(d):
    return np.ctypeslib.as_array(d.elt_nodes, (d.n_elements * 3,)).copy()


def _nodes(d):
    return np.ctypeslib.as_array(d.node_xyz, (d.n_nodes * 3,)).copy()


def _degenerate(d, keep):
    x, e = _nodes(d), _elements(d)
    x[3 * e[1]:3 * e[1] + 2] = x[3 * e[0]:3 * e[0] + 2]  # vertex 1 of element 0 onto vertex 0
    _set_array(d, "node_xyz", x, C.c_double, keep)


def _node_oob(d, keep):
    e = _elements(d)
    e[7] = d.n_nodes
    _set_array(d, "elt_nodes", e, C.c_int32, keep)


def _nan_depth(d, keep):
    x = _nodes(d)
    x[5] = np.nan
    _set_array(d, "node_xyz", x, C.c_double, keep)


BAD_CREATE = {
    # mesh description: counts against their arrays, indices, geometry
    "negative elements": (lambda d, p, k: setattr(d, "n_elements", -5), "empty mesh"),
    "node index out of range": (lambda d, p, k: _node_oob(d, k), "element node index out of range"),
    "zero-area element": (lambda d, p, k: _degenerate(d, k), "degenerate element"),
    "non-finite depth": (lambda d, p, k: _nan_depth(d, k), "non-finite node coordinate"),
    "boundary edges without arrays": (lambda d, p, k: setattr(d, "bound_edge_nodes", None), "boundary edge count"),
    "negative boundary edges": (lambda d, p, k: setattr(d, "n_bound_edges", -3), "boundary edge count"),
    "constituents without arrays": (lambda d, p, k: setattr(d, "frequency", None), "constituent arrays missing"),
    "negative constituents": (lambda d, p, k: setattr(d, "n_constituents", -2), "negative constituent count"),
    "levee pairs without arrays": (lambda d, p, k: setattr(d, "n_levee_pairs", 2), "levee pair count"),
    "negative levee pairs": (lambda d, p, k: setattr(d, "n_levee_pairs", -4), "levee pair count"),
    "tide without amplitudes": (lambda d, p, k: setattr(d, "node_amplitude", None), "without forcing data"),
    # run parameters
    "unknown problem": (lambda d, p, k: setattr(p, "problem", 99), "unknown problem"),
    "unknown friction law": (lambda d, p, k: setattr(p, "friction", 9), "unknown friction law"),
    "negative stage count": (lambda d, p, k: setattr(p, "nstages", -3), "invalid Runge-Kutta method"),
    "negative order": (lambda d, p, k: setattr(p, "p", -3), "polynomial order"),
}


@pytest.mark.parametrize("case", list(BAD_CREATE))
def test_create_checks_mesh_and_parameters_before_the_device(case):
    # every count is checked against its array and every index against its range before any
    # array is read, and before the device is looked for, so a malformed description gets
    # RKDG_ERR_ARG with a message on any machine (null arrays behind positive counts, for
    # instance, read through a null pointer otherwise)
    from dgswemv2_amd import workloads
    lib = N.load_library()
    mesh, prm = workloads.tidal(scale=0.02)
    desc, keep = mesh.desc, []
    assert desc.n_constituents > 0 and desc.n_bound_edges > 0
    mutate, msg = BAD_CREATE[case]
    mutate(desc, prm, keep)
    h = C.c_void_p()
    assert lib.rkdg_create(C.byref(desc), C.byref(prm), N.RKDG_FP64_PLAIN, C.byref(h)) == N.RKDG_ERR_ARG
    assert msg in lib.rkdg_create_error().decode()
    assert not h.value


def test_unbuilt_order_is_unsupported_before_the_device():
    from dgswemv2_amd import workloads
    lib = N.load_library()
    mesh, prm = workloads.tidal(scale=0.02)
    prm.p = 4
    h = C.c_void_p()
    assert lib.rkdg_create(C.byref(mesh.desc), C.byref(prm), N.RKDG_FP64_PLAIN, C.byref(h)) == N.RKDG_ERR_UNSUPPORTED


def test_c_program_uses_the_abi(tmp_path):
    """include/rkdg.h compiles as strict C11 (-pedantic -Werror) and a C program links librkdg.so and
    drives it (tests/c_abi/lake_at_rest.c); without a HIP device it reports RKDG_ERR_NO_DEVICE
    through rkdg_create and exits with status 2."""
    import subprocess
    import torch  # noqa: F401  (device count only, no HIP initialisation)
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import __graft_entry__ as G
    exe = G.build_c_abi_example(tmp_path / "lake_at_rest")
    if torch.cuda.device_count() > 0:
        pytest.skip("a HIP device is visible (the GPU test runs the program)")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "no HIP device" in r.stdout


def test_no_device_is_reported(monkeypatch):
    # On a host without a HIP device rkdg_create fails loudly instead of running anything.
    from dgswemv2_amd import Simulation, workloads
    import torch  # noqa: F401  (only to probe for a device without initialising HIP here)
    if torch.cuda.device_count() > 0:
        pytest.skip("a HIP device is visible")
    mesh, prm = workloads.manufactured(scale=0.1)
    with pytest.raises(N.RkdgError) as ei:
        Simulation(mesh, prm)
    assert ei.value.status == N.RKDG_ERR_NO_DEVICE


def test_mesh_generator_patterns():
    from dgswemv2_amd import quad_mesh
    for pat in (0, 1, 2):
        m = quad_mesh(0, 0, 3, 2, 3, 2, pattern=pat)
        x, y = m.node_xyz[:, 0], m.node_xyz[:, 1]
        e = m.elt_nodes
        area = (x[e[:, 1]] - x[e[:, 0]]) * (y[e[:, 2]] - y[e[:, 0]]) - (x[e[:, 2]] - x[e[:, 0]]) * (y[e[:, 1]] - y[e[:, 0]])
        assert np.all(area > 0)  # counter-clockwise, as quad_mesh_generator writes them
        assert np.isclose(area.sum() / 2, 6.0)
        assert len(m.bound_types) == 2 * (3 + 2)


def _generate(**fields):
    lib = N.load_library()
    s = N.QuadMeshSpec()
    s.m, s.n, s.x2, s.y2, s.b0 = 4, 3, 1.0, 1.0, 2.0
    for k, v in fields.items():
        if k == "bc":
            for i, b in enumerate(v):
                s.bc[i] = b
        else:
            setattr(s, k, v)
    handle, desc = C.c_void_p(), N.MeshDesc()
    st = lib.rkdg_mesh_generate(C.byref(s), C.byref(handle), C.byref(desc))
    if st == N.RKDG_OK:
        lib.rkdg_mesh_free(handle)
    return st


@pytest.mark.parametrize("fields", [dict(m=0), dict(n=-2), dict(m=1 << 30, n=1 << 30), dict(pattern=3),
                                    dict(x2=0.0), dict(y1=2.0), dict(bathy_kind=4), dict(bathy_kind=-1),
                                    dict(bathy_kind=2, bmin=5.0, bmax=10.0, n_modes=1 << 30), dict(bc=[0, 9, 0, 0]),
                                    dict(bc=[0, 0, 0, 129])])
def test_mesh_generator_rejects_bad_specs(fields):
    # every field the generator sizes or branches on is checked before any work (a huge mode
    # count would otherwise run for hours, an unknown side type reach the model builder)
    assert _generate(**fields) == N.RKDG_ERR_ARG


def test_mesh_generator_accepts_the_documented_ranges():
    assert _generate() == N.RKDG_OK
    assert _generate(bathy_kind=2, bmin=5.0, bmax=10.0, n_modes=4096) == N.RKDG_OK
    assert _generate(bc=[1, 2, 3, 0], frequency=1e-4, forcing_fact=1.0, amplitude=0.1) == N.RKDG_OK


def test_noalloc_forward_rejects_bad_arguments_before_any_device_work():
    lib = N.load_library()
    w = N.MlpWeights()
    X = np.zeros((N.MLP_IN, 4), np.float32)
    out = np.zeros(4, np.float32)
    limit = 16384  # RKDG_MLP_FORWARD_NOALLOC_MAX
    assert "#define RKDG_MLP_FORWARD_NOALLOC_MAX 16384" in HEADER.read_text()
    assert lib.rkdg_mlp_forward_noalloc(None, N.fptr(X), 4, N.fptr(out)) == N.RKDG_ERR_ARG
    assert lib.rkdg_mlp_forward_noalloc(C.byref(w), N.fptr(X), 0, N.fptr(out)) == N.RKDG_ERR_ARG
    assert lib.rkdg_mlp_forward_noalloc(C.byref(w), N.fptr(X), limit + 1, N.fptr(out)) == N.RKDG_ERR_ARG
    X[3, 1] = np.nan
    assert lib.rkdg_mlp_forward_noalloc(C.byref(w), N.fptr(X), 4, N.fptr(out)) == N.RKDG_ERR_ARG
