#!/usr/bin/env python3
"""Write the golden-vector fixtures of this directory from the reference's own unit tests.

Reads (as text) and extracts only numbers -- inputs and expected outputs:
  test/test_basis_dubiner_2d.cpp    true_vals (:4-994), true_Dvals (:995-1523), points (:1525-1560)
  test/test_element_triangle.hpp    IntegrationPhi_true, IntegrationDPhiDX_true, IntegrationDPhiDY_true (:17-63)
  test/test_llf_flux.cpp            three LLF configurations (:77-165)
The wetting-drying scenarios of test/test_wetting_drying.cpp (:24-374) are parameterised by
h_o and the vertex depths, so their inputs and expected values are written out below as data.

Usage:  python make_golden.py /root/reference
"""
import json
import re
import sys
from pathlib import Path

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def nested_lists(src, name):
    start = src.index(name)
    i = src.index("{", start)
    depth, j = 0, i
    while True:
        if src[j] == "{":
            depth += 1
        elif src[j] == "}":
            depth -= 1
            if depth == 0:
                break
        j += 1
    body = src[i:j + 1]
    rows = re.findall(r"\{([^{}]*)\}", body)
    if rows and body.count("{") > 1:
        return [[float(x) for x in re.findall(NUM, r)] for r in rows]
    return [float(x) for x in re.findall(NUM, body)]


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    out = Path(__file__).resolve().parent

    src = (ref / "test/test_basis_dubiner_2d.cpp").read_text()
    vals = nested_lists(src, "true_vals")
    dvals = nested_lists(src, "true_Dvals")
    pts = {}
    for i, c, v in re.findall(r"points\[(\d+)\]\[(\d)\]\s*=\s*(" + NUM + ")", src):
        pts.setdefault(int(i), [0.0, 0.0])[int(c)] = float(v)
    points = [pts[i] for i in sorted(pts)]
    assert len(vals) == 66 and all(len(r) == 15 for r in vals)
    assert len(dvals) == 66 and all(len(r) == 30 for r in dvals) and len(points) == 15
    json.dump({"source": "test/test_basis_dubiner_2d.cpp", "p": 10, "points": points, "phi": vals,
               "dphi_interleaved": dvals, "tol_phi": 100, "tol_dphi": 1000}, open(out / "dubiner_p10.json", "w"))

    src = (ref / "test/test_element_triangle.hpp").read_text()
    el = {k: nested_lists(src, k) for k in ("IntegrationPhi_true", "IntegrationDPhiDX_true", "IntegrationDPhiDY_true")}
    assert all(len(v) == 66 for v in el.values())
    el.update({"source": "test/test_element_triangle.hpp", "p": 10,
               "vertices": [[-0.5, 0.0], [0.5, 0.0], [0.0, "sqrt(3)/2"]], "f": "(x+1)^2 + (y-1)^2",
               "tol_factor": 1.0e4})
    json.dump(el, open(out / "element_p10.json", "w"))

    src = (ref / "test/test_llf_flux.cpp").read_text()
    cases = []
    for blk in re.split(r"\{\s*//\s*Configuration \d+", src)[1:]:
        get = lambda k: float(re.search(k + r"\s*=\s*(" + NUM + ")", blk).group(1))
        nrm = re.search(r"normal\{(" + r"[^}]*" + r")\}", blk).group(1)
        sx = -1.0 if nrm.split(",")[0].strip().startswith("-") else 1.0
        sy = -1.0 if nrm.split(",")[1].strip().startswith("-") else 1.0
        expect = [float(x) for x in re.findall(NUM, blk.split("normal,")[1].split(")")[0])]
        cases.append({"q_in": [get("ze_in"), get("qx_in"), get("qy_in")],
                      "q_ex": [get("ze_ex"), get("qx_ex"), get("qy_ex")],
                      "bath": get("bath"), "sp": get("sp"), "normal_sign": [sx, sy],
                      "normal": "sign * 1/sqrt(2)", "F_hat": expect[-3:]})
    assert len(cases) == 3
    json.dump({"source": "test/test_llf_flux.cpp", "g": 9.81, "tol_factor": 100, "cases": cases},
              open(out / "llf_flux.json", "w"))

    # test/test_wetting_drying.cpp: equilateral triangle (-0.5,0), (0.5,0), (0, sqrt(3)/2), p = 1,
    # h_o = 0.01, h_o_threshold left at 1e-6, vertex depths 1, 2, 3, bath_min 1; the scenarios
    # run in sequence on one element (the wet flag carries over unless the test resets it).
    # Vertex inputs/expectations are written in units of h_o with the depth subtracted: an
    # entry [c, "b"] means c*h_o - depth[v]; [c] means the literal c*h_o.
    wd = {
        "source": "test/test_wetting_drying.cpp", "h_o": 0.01, "h_o_threshold": 1.0e-6,
        "depth": [1.0, 2.0, 3.0], "bath_min": 1.0,
        "qx_in": [1.0, 2.0, 3.0], "qy_in": [-1.0, -2.0, -3.0],
        "scenarios": [
            {"name": "completely dry", "wet_in": True, "ze_in": [[0.5, "b"], [0.25, "b"], [1.0 / 6.0, "b"]],
             "wet": False, "ze": "h_avg - b", "qx": [0, 0, 0], "qy": [0, 0, 0], "tol": [1e4, 100]},
            {"name": "completely wet", "wet_in": None, "ze_in": [[1.0], [1.0], [1.0]],
             "wet": True, "ze": [[1.0], [1.0], [1.0]], "qx": [1, 2, 3], "qy": [-1, -2, -3], "tol": [1e4, 1e4]},
            {"name": "dry on average", "wet_in": None, "ze_in": [[0.5, "b"], [0.5, "b"], [2.0, "b"]],
             "wet": False, "ze": "h_avg - b", "qx": [0, 0, 0], "qy": [0, 0, 0], "tol": [1e4, 100]},
            {"name": "flood type", "wet_in": False, "ze_in": [[0.5, "b"], [0.5, "b"], [3.5, "b"]],
             "wet": False, "ze": [[1.0, "b"], [1.0, "b"], [2.5, "b"]], "qx": [0, 0, 0], "qy": [0, 0, 0],
             "tol": [1e5, 100]},
            {"name": "dam-break type", "wet_in": False, "ze_in": [[3.5, "b"], [0.5, "b"], [0.5, "b"]],
             "wet": True, "ze": [[2.5, "b"], [1.0, "b"], [1.0, "b"]], "qx": [6, 0, 0], "qy": [-6, 0, 0],
             "tol": [1e5, 1e5]},
        ],
    }
    json.dump(wd, open(out / "wetting_drying.json", "w"), indent=1)
    print("wrote", [p.name for p in sorted(out.glob("*.json"))])


if __name__ == "__main__":
    main()
