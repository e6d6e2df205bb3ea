"""A deck-driven run (include/vpfft_io.h: vpfft_run_deck and the vpfft_run
command) against the same run through the MaterialGrid interface and the
oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DECK = """# two-grain tension test, the reference's UnitTester layout
GridDimension 16 12 8
InputStrainRate
 -0.5 0 0
 0 -0.5 0
 0 0 1
MaxNRIter 40
TimeStep 0.005
NumTimeIter 3
NRTolerence 1e-5
TwoGrainInit
CRSS 1 1 1 1 1 1 1 1 1 1 1 1.5
GammaDotBase 1 1 1 1 1 1 1 1 1 1 1 1
RateSensitivity 10 10 10 10 10 10 10 10 10 10 10 10
"""


def _parse_stress_strain(path):
    return np.array([[float(x) for x in line.split()] for line in open(path)])


def test_deck_run_matches_grid_interface_and_oracle(gpu, vpfft, oracle, tmp_path):
    from vpfft_amd import deck as D
    deck_path = tmp_path / "run.config"
    deck_path.write_text(DECK)
    cfg = D.ConfigFile()
    cfg.Read(deck_path)
    dims = (cfg.NumX, cfg.NumY, cfg.NumZ)
    n = int(np.prod(dims))
    ss_path, mat_path = tmp_path / "Stress_Strain.txt", tmp_path / "final.txt"
    print(f"deck run {dims}, {cfg.NumTimeIter} steps", flush=True)
    infos = D.run_deck(deck_path, vpfft.FP64, str(ss_path), str(mat_path))
    assert len(infos) == 3

    # the same steps through the MaterialGrid interface
    R, S = D.two_grain_init(*dims)
    crss = np.tile(cfg.CRSS, (n, 1))
    mat = cfg.material()
    edot = cfg.strain_rate()
    g = vpfft.MaterialGrid(*dims, mat)
    g.SetVoxelLength(1, 1, 1)
    g.SendDataAndInitialize(R, S, crss)
    avgs = []
    for s in range(3):
        state = dict(R=g.Orientation(), crss=g.CRSS(), acc=g.AccumulatedShear(), L=tuple(g.VoxelLength()))
        avgs.append(g.RunSingleStrainStep(edot, cfg.MaxNRIter, cfg.TimeStep, cfg.NRTolerence))
        assert g.last_info.iterations == infos[s]["iterations"]
        assert np.array_equal(avgs[-1], np.array(infos[s]["avg_stress"]))
        if s == 0:  # the oracle on identical inputs
            ref = oracle.strain_step(dims, mat, state["R"], state["crss"], edot, cfg.MaxNRIter, cfg.TimeStep,
                                     cfg.NRTolerence, acc=state["acc"], lengths=state["L"])
            assert ref["iterations"] == infos[0]["iterations"]
            assert rel_err(avgs[0], ref["avg_stress"]) < 1e-8
            assert rel_err(g.Stress(), ref["sigma"]) < 1e-8
    R_end, S_end = g.Orientation(), g.Stress()
    g.close()

    # Stress_Strain.txt: step, E (5), <sigma> (5), |E|, |<sigma>| (Src/MaterialGrid.cpp:1221-1225)
    t = _parse_stress_strain(ss_path)
    assert t.shape == (3, 13)
    assert np.array_equal(t[:, 0], [0, 1, 2])
    E = np.cumsum([edot * cfg.TimeStep] * 3, axis=0)
    assert np.allclose(t[:, 1:6], E, rtol=1e-5, atol=1e-9)
    assert np.allclose(t[:, 6:11], np.array(avgs), rtol=1e-5, atol=1e-9)
    assert np.allclose(t[:, 12], np.linalg.norm(avgs, axis=1), rtol=1e-5)

    # the final microstructure file holds the grid's end state exactly
    R2, S2 = D.read_material_file(mat_path, *dims)
    assert np.array_equal(R2, R_end) and np.array_equal(S2, S_end)

    # the command-line runner gives the same lines
    exe = os.path.join(ROOT, "vpfft_amd", "vpfft_run")
    ss2 = tmp_path / "ss_cli.txt"
    r = subprocess.run([exe, str(deck_path), "--stress-strain", str(ss2)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert ss2.read_text() == ss_path.read_text()
