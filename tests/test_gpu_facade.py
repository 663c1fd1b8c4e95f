"""The C++ front end (nihu_amd/include/nihu_amd/nihu.hpp) on the GPU: NiHu-style
known-answer programs built against libnbem.so, on the reference's
sphere_1_3.off mesh fixture (218 nodes, 432 triangles): transparent and
pulsating-sphere answers in collocation, constant Galerkin and isoparametric
Galerkin, and NiHu's exceptions; on a 384-element quadrilateral cubed sphere
(OFF file): the collocation answers and the Burton-Miller composition, also
with the gauss field (create_function_space with NiHu's field rows); the
same on a cubed sphere whose every third quadrilateral is split into two
triangles (one OFF file with both face types)."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_facade_known_answers(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "test_facade")
    assert os.path.exists(exe), "build the C++ tests first (__graft_entry__.build())"
    mesh = os.path.join(ROOT, "tests", "golden", "sphere_1_3.off")
    from nihu_amd.mesh import jittered_cubed_sphere, write_off, mixed_cubed_sphere, table_faces
    qmesh = str(tmp_path / "cubed_sphere_8.off")
    write_off(qmesh, *jittered_cubed_sphere(8, seed=6))
    mmesh = str(tmp_path / "mixed_cubed_sphere_8.off")
    mnodes, mtable = mixed_cubed_sphere(8, seed=6)
    write_off(mmesh, mnodes, table_faces(mtable))
    r = subprocess.run([exe, mesh, qmesh, mmesh], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "out_of_range as expected" in r.stdout
    assert "runtime_error as expected: Invalid general singular quadrature" in r.stdout
    assert "logic_error as expected" in r.stdout
    for line in ("Galerkin transparent error", "Pulsating sphere relative error (Galerkin)",
                 "Pulsating sphere relative error (collocation, complex wave number)",
                 "Galerkin isoparametric transparent error", "Quadrilateral collocation transparent error",
                 "Quadrilateral Burton-Miller composition", "Pulsating sphere relative error (quadrilateral Galerkin)",
                 "Pulsating sphere relative error (quadrilateral gauss field)",
                 "Gauss-field Burton-Miller composition", "gauss dof numbering: invalid_argument as expected",
                 "Mixed-mesh collocation transparent error", "Mixed-mesh Burton-Miller composition",
                 "mixed-mesh Galerkin: logic_error as expected", "Mixed-field Galerkin: Gauss identity"):
        assert line in r.stdout
