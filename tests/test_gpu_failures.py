"""Failure paths forced from the host: the loop's iteration cap, degenerate
and non-finite loading, and a bad orientation field.  Where the reference
exits or stays silent (Src/Error.h:25-32, MaterialGrid.cpp:186-190,
1066-1068), the library reports a status, and the grid stays usable."""
import numpy as np
import pytest

from helpers import edot_tension, polycrystal

pytestmark = pytest.mark.gpu

DEGENERATE, NONFINITE = -4, -5


def _grid(vpfft, dims=(12, 12, 12), grains=6, seed=3):
    R, _ = polycrystal(None, grains, seed, dims)
    g = vpfft.MaterialGrid(*dims, vpfft.Material.uniform(1.0, 10))
    g.SetVoxelLength(1.0, 1.0, 1.0)
    g.SendDataAndInitialize(R)
    return g, R


def test_iteration_cap_reports_not_converged(gpu, vpfft):
    g, _ = _grid(vpfft)
    g.RunSingleStrainStep(edot_tension(), 1, 0.005, 1e-12)
    assert g.last_info.iterations == 1 and g.last_info.converged == 0 and g.last_info.status == 0
    g.close()


def test_degenerate_and_nonfinite_loading_return_errors(gpu, vpfft):
    g, R = _grid(vpfft)
    with pytest.raises(vpfft.VPFFTError) as e:
        g.RunSingleStrainStep(np.zeros(5), 10, 0.005, 1e-6)
    assert e.value.code == DEGENERATE
    bad = edot_tension().copy()
    bad[2] = np.nan
    with pytest.raises(vpfft.VPFFTError) as e:
        g.RunSingleStrainStep(bad, 10, 0.005, 1e-6)
    assert e.value.code == NONFINITE
    # the grid is untouched by the refused calls and still steps
    assert np.array_equal(g.Orientation(), R)
    g.RunSingleStrainStep(edot_tension(), 10, 0.005, 1e-6)
    assert g.last_info.status == 0
    g.close()


def test_nonfinite_orientation_is_reported(gpu, vpfft):
    g, R = _grid(vpfft)
    R2 = R.copy()
    R2[17, 4] = np.nan
    g.SendDataAndInitialize(R2)
    with pytest.raises(vpfft.VPFFTError) as e:
        g.RunSingleStrainStep(edot_tension(), 10, 0.005, 1e-6)
    assert e.value.code in (DEGENERATE, NONFINITE)
    g.SendDataAndInitialize(R)
    g.RunSingleStrainStep(edot_tension(), 10, 0.005, 1e-6)
    assert g.last_info.status == 0
    g.close()
