"""The reference's drivers compile and link unchanged against include/header_files + libmgm.

Runs where the reference tree exists (its drivers are not shipped with this
repository): each driver is copied into a scratch directory next to a
`header_files` link to this repository's drop-in headers and the reference's
own init/ plug-ins, then compiled with gcc and linked against libmgm.so.
The reference tree is absent where the GPU is, so the drop-in surface is run
on the GPU through tests/dropin/heat_driver.c, a driver of this repository
that follows mg_heat_conduction.c's call sequence (tests/test_gpu_dropin.py).
"""
import os
import shutil
import subprocess
import tempfile

import pytest

from conftest import REFERENCE, ROOT

pytestmark = pytest.mark.skipif(not os.path.isfile(os.path.join(REFERENCE, "mg_heat_conduction.c")),
                                reason="reference drivers not available")


def _scratch(d):
    os.symlink(os.path.join(ROOT, "include", "header_files"), os.path.join(d, "header_files"))
    shutil.copytree(os.path.join(REFERENCE, "init"), os.path.join(d, "init"))


def _compile(d, src, exe):
    lib = os.path.join(ROOT, "multigrid_meshless_amd")
    cmd = ["gcc", "-std=gnu11", "-O2", "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
           src, "-o", exe, f"-L{lib}", "-lmgm", f"-Wl,-rpath,{lib}", "-lm"]
    r = subprocess.run(cmd, cwd=d, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.isfile(os.path.join(d, exe))


@pytest.mark.parametrize("driver", ["mg_heat_conduction.c", "mg_NS_solver.c"])
def test_reference_driver_links_unchanged(driver):
    with tempfile.TemporaryDirectory() as d:
        _scratch(d)
        shutil.copy(os.path.join(REFERENCE, driver), d)
        _compile(d, driver, "a.out")


def test_fractional_step_header_links():
    """mg_NS_solver.c's commented-out alternative: fractional_step_solver.h + init_LDC.c."""
    src = r'''
#include "header_files/fractional_step_solver.h"
#include "header_files/openACC_functions.h"
#include "init/init_LDC.c"
int main(void) {
    PointStructure *P; FieldVariables *F;
    read_flow_parameters("flow_parameters.csv");
    read_grid_filenames(&P, "grid_filenames.csv", &parameters.num_levels);
    read_complete_mesh_data(P, parameters.num_levels);
    for (int l = 0; l < parameters.num_levels; l++) create_derivative_matrices(&P[l]);
    AllocateMemoryFieldVariables(&F, P, parameters.num_levels);
    initial_conditions(P, F, 1); boundary_conditions(P, F, 1);
    parameters.dt = calculate_dt(&P[0]);
    copyin_essentials_to_gpu(P); copyin_field_to_gpu(F, P); copyin_parameters_to_gpu();
    double e = fractional_step_explicit(P, F);
    free_PointStructure(P, parameters.num_levels); free_field(F, parameters.num_levels);
    return e < 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        _scratch(d)
        open(os.path.join(d, "fs_driver.c"), "w").write(src)
        _compile(d, "fs_driver.c", "fs.out")
