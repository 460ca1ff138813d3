"""Write tests/golden/config_readers.json from the compiled reference's config
readers (read_flow_parameters, hf/mesh_functions.h:37-100; read_grid_filenames,
:102-129).

    make -C oracle ref && python tests/golden/make_config_golden.py

Inputs: the reference's own flow_parameters.csv and grid_filenames.csv (data
files of the reference tree) and a few written here that exercise the
reference's conversions (%hd on a fractional multiplier, %f on omega,
exponent forms, the Neumann flag).  Each file is parsed by the reference in a
subprocess (it exits on a bad file); the fixture holds every input's text
and every field of `parameters` the parse left, or the grid file's level
count and names.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("MGM_REFERENCE", "/root/reference")
sys.path.insert(0, os.path.dirname(HERE))

KEYS = ["domain_dimensions", "poly_deg", "phs_deg", "cloud_size_multiplier", "test_derivative",
        "courant_number", "steady_tolerance", "poisson_solver_tolerance", "sor_parameter", "num_vcycles",
        "num_relax", "num_time_steps", "write_interval", "neumann_flag_boundary", "time_step", "iter_momentum",
        "iter_timple", "Re", "facRe", "facdt"]


def csv(vals):
    return "".join(f"{k},{v}\n" for k, v in zip(KEYS, vals))


def flow_inputs():
    return {
        "reference_flow_parameters": open(os.path.join(REF, "flow_parameters.csv")).read(),
        "cube_3d_multiplier_3.5": csv([3, 3, 3, "3.5", 1, 0.5, "1e-10", "1e-9", 1.15, 20, 3, 2000, 500, 0, 0.001,
                                        7, 2, 1000, 0.9, 1.1]),
        "cavity_2d_exponents": csv([2, 4, 5, "2.0", 0, "5e-1", "1.0E-6", "2.5e-9", "0.7", 1, 5, 100000, 10, 1,
                                     "3.125e-4", 3, 1, "3.2e3", "1", "2.5e-1"]),
        "omega_rounding": csv([2, 3, 3, 2, 0, 0.3, 1e-8, 1e-8, "1.3333333333333333", 10, 2, 1, 1, 1, 0.1, 5, 1,
                               10, 1.0, 1.0]),
    }


def grid_inputs():
    return {
        "reference_grid_filenames": open(os.path.join(REF, "grid_filenames.csv")).read(),
        "five_levels": "num_levels,5\ncube_100.msh\ncube_50.msh\ncube_25.msh\ncube_13.msh\ncube_7.msh\n",
    }


def child(kind, path, out_path):
    from cstructs import Params
    lib = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_heat.so"))
    if kind == "flow":
        p = Params()
        lib.ref_params.argtypes = [C.c_char_p, C.POINTER(Params)]
        lib.ref_params(path.encode(), C.byref(p))
        out = p.as_dict()
    else:
        names = C.create_string_buffer(50 * 16)
        lib.ref_grid_filenames.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        n = lib.ref_grid_filenames(path.encode(), names, 16)
        out = {"num_levels": n, "names": [names.raw[50 * k:50 * (k + 1)].split(b"\0")[0].decode()
                                          for k in range(min(n, 16))]}
    sys.stdout.flush()
    json.dump(out, open(out_path, "w"))


def parse(kind, text):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "in.csv")
        open(path, "w").write(text)
        res = os.path.join(d, "out.json")
        proc = subprocess.run([sys.executable, __file__, "--child", kind, path, res], stdout=subprocess.DEVNULL)
        if proc.returncode != 0:
            raise RuntimeError(f"reference parse failed on {kind} input")
        return json.load(open(res))


def main():
    out = {"flow": {}, "grid": {}}
    for name, text in flow_inputs().items():
        out["flow"][name] = {"text": text, "parameters": parse("flow", text)}
    for name, text in grid_inputs().items():
        out["grid"][name] = {"text": text, **parse("grid", text)}
    path = os.path.join(HERE, "config_readers.json")
    json.dump(out, open(path, "w"), indent=1, sort_keys=True)
    print(f"wrote {path}")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--child":
        child(sys.argv[2], sys.argv[3], sys.argv[4])
    else:
        main()
