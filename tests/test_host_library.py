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
