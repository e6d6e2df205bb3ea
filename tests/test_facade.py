"""The C++ MaterialGrid facade (include/vpfft_materialgrid.hpp): the header
compiles on its own, and its driver (tests/cpp/facade_check.cpp, the shape
of the reference's MPI_VPFFT_Test) gives the same bits as the same run
through the Python mirror of the C ABI."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "facade_check")


def test_facade_header_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "vpfft_materialgrid.hpp"\n'
                   "int main() { VPFFT::LinearAlgebra::EigenRep e(1, 2, 3, 4, 5);\n"
                   "  vpfft_material m = VPFFT::DataStructures::MaterialGrid::ReferenceMaterial();\n"
                   "  return (int)(e[4] - 5) + (m.rate_sensitivity[0] - 10); }\n")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-fsyntax-only", str(src)], check=True)


def _parse(out):
    res = {"avg": [], "iterations": []}
    for line in out.strip().splitlines():
        k, *v = line.split()
        if k in ("avg",):
            res["avg"].append(np.array([float(x) for x in v]))
        elif k == "iterations":
            res["iterations"].append(int(v[0]))
        else:
            res[k] = np.array([float(x) for x in v])
    return res


@pytest.mark.gpu
def test_facade_driver_matches_python_mirror(gpu, vpfft):
    from helpers import edot_tension
    assert os.path.exists(DRIVER), "build first: python -m vpfft_amd.build"
    edge, grains, seed, steps = 12, 5, 0, 2
    print("facade driver", edge, grains, seed, steps, flush=True)
    out = subprocess.run([DRIVER, str(edge), str(grains), str(seed), str(steps)], check=True, capture_output=True,
                         text=True, timeout=120).stdout
    r = _parse(out)
    # the same run through the Python mirror
    R, _, _ = vpfft.generate_voronoi(edge, edge, edge, grains, seed)
    mat = vpfft.Material.uniform(15.5, 10)
    with vpfft.MaterialGrid(edge, edge, edge, mat, crss=11.0) as g:
        g.SetVoxelLength(1.0, 1.0, 1.0)
        g.SendDataAndInitialize(R)
        for s in range(steps):
            avg = g.RunSingleStrainStep(r["edot"], 100, 0.005, 1e-5)
            assert r["iterations"][s] == g.last_info.iterations
            assert np.array_equal(r["avg"][s], avg), s
        v = g.ToIndex(1, 2, 3) if hasattr(g, "ToIndex") else (1 * edge + 2) * edge + 3
        assert np.array_equal(r["stress"], g.Stress()[v])
        assert np.array_equal(r["lambda"], g.LagrangeMultiplier()[v])
        assert np.array_equal(r["schmid"], g.SchmidtTensorList()[v, 7])
        o = g.Orientation()[v]
        assert np.array_equal(r["orientation"], o[[0, 4, 8]])
    assert r["roundtrip"][0] == 1
    # the strain rate built by EigenRep(const SMatrix3x3&) is the tension one
    assert np.max(np.abs(r["edot"] - edot_tension())) < 1e-15
    assert r["slip_systems"][0] == 12 and abs(r["slip_systems"][1] - 1 / (2 * 6 ** 0.5)) < 1e-16
    assert r["degenerate"][0] == -4  # VPFFT_E_DEGENERATE
