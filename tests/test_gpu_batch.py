"""Batched strain steps (vpfft_batch_*): many grids of uneven sizes advance
together.  Each grid's result must be the same bits alone and in any batch,
with or without block-shape tuning, agree to rounding when the tuner also
picks radix orders, and match the oracle either way."""
import numpy as np
import pytest

from conftest import rel_err
from helpers import edot_tension, polycrystal, random_unit_edot

pytestmark = pytest.mark.gpu

# (dims, grains, n, precision)
MIX = [((16, 16, 16), 8, 10, 0), ((24, 16, 20), 12, 20, 1), ((12, 18, 10), 5, 5, 0), ((32, 32, 32), 27, 20, 0),
       ((20, 24, 14), 9, 10, 1)]


def _grids(vpfft, cases, seed0=50, seeds=None):
    out = []
    for i, (dims, grains, nexp, prec) in enumerate(cases):
        R, _ = polycrystal(None, grains, seed0 + (i if seeds is None else seeds[i]), dims)
        g = vpfft.MaterialGrid(*dims, vpfft.Material.uniform(1.0, nexp), precision=prec)
        g.SetVoxelLength(1.0, 1.0, 1.0)
        g.SendDataAndInitialize(R)
        out.append(g)
    return out


def _state(g):
    return dict(sigma=g.Stress(), R=g.Orientation(), crss=g.CRSS(), lam=g.LagrangeMultiplier())


def _run_batch(vpfft, cases, idx, steps, tune=False, max_iter=40, tol=1e-6):
    grids = _grids(vpfft, [cases[i] for i in idx], seeds=idx)
    edots = np.array([random_unit_edot(100 + i) for i in idx])
    res = []
    with vpfft.Batch(tune=tune) as b:
        for s in range(steps):
            avg = b.RunSingleStrainStep(grids, edots, max_iter, 0.005, tol)
            res.append((avg.copy(), [inf.iterations for inf in b.last_infos]))
        shapes = [[b.block_shape(g, p) for p in range(3)] for g in grids]
    states = [_state(g) for g in grids]
    for g in grids:
        g.close()
    return res, states, shapes


def test_batch_results_do_not_depend_on_the_mix(gpu, vpfft):
    print("batch of 5 uneven grids, 2 steps, vs each alone", flush=True)
    together, st_all, _ = _run_batch(vpfft, MIX, list(range(len(MIX))), 2)
    for j in range(len(MIX)):
        alone, st_one, _ = _run_batch(vpfft, MIX, [j], 2)
        for s in range(2):
            assert together[s][1][j] == alone[s][1][0]
            assert np.array_equal(together[s][0][j], alone[s][0][0])
        for k in st_one[0]:
            assert np.array_equal(st_all[j][k], st_one[0][k]), (j, k)


def test_batch_tuned_block_shapes_give_the_same_bits(gpu, vpfft):
    print("batch tuned vs default block shapes", flush=True)
    plain, st_plain, shapes_plain = _run_batch(vpfft, MIX, [0, 1, 3], 1)
    tuned, st_tuned, shapes_tuned = _run_batch(vpfft, MIX, [0, 1, 3], 1, tune=True)
    assert np.array_equal(plain[0][0], tuned[0][0])
    assert plain[0][1] == tuned[0][1]
    for a, b in zip(st_plain, st_tuned):
        for k in a:
            assert np.array_equal(a[k], b[k])
    print("default shapes", shapes_plain, "tuned", shapes_tuned, flush=True)


def test_batch_tuned_radix_orders_agree_to_rounding(gpu, vpfft):
    """Level-2 tuning also picks each axis's radix order, which changes the
    transforms' rounding only: same iteration counts, <sigma> and fields
    within 1e-10 of the default schedule for FP64 grids (mixed ones: <sigma>
    1e-6, fields 1e-4); every chosen plan is a factorisation of its axis
    into the supported radices."""
    print("batch tuned radix orders vs default schedule", flush=True)
    idx = [0, 1, 3]
    # one step: across steps the n = 20 Taylor prologue turns rounding-level
    # differences into ~1e-9 (the reference's own sensitivity)
    plain, st_plain, _ = _run_batch(vpfft, MIX, idx, 1)
    grids = _grids(vpfft, [MIX[i] for i in idx], seeds=idx)
    edots = np.array([random_unit_edot(100 + i) for i in idx])
    tuned = []
    with vpfft.Batch(tune="radix") as b:
        for s in range(1):
            avg = b.RunSingleStrainStep(grids, edots, 40, 0.005, 1e-6)
            tuned.append((avg.copy(), [inf.iterations for inf in b.last_infos]))
        for g, i in zip(grids, idx):
            dims = MIX[i][0]
            for p, n in zip(range(3), (dims[2], dims[1], dims[0])):
                stages = b.radix_plan(g, p)
                assert int(np.prod(stages)) == n and set(stages) <= {2, 3, 4, 5, 7, 8}, (dims, p, stages)
            print(dims, "radix plans z/y/x:", [b.radix_plan(g, p) for p in range(3)], flush=True)
    st_tuned = [_state(g) for g in grids]
    for g in grids:
        g.close()
    # FP64 grids to 1e-10; a mixed grid's FP32 transforms round at ~1e-7, so
    # its <sigma> moves by ~1e-9 (bar 1e-6) and single voxels of its fields
    # by ~1e-6 (the n = 20 Newton solve amplifies the FFT's rounding; bar 1e-4)
    bars = [1e-10 if MIX[i][3] == 0 else 1e-6 for i in idx]
    field_bars = [1e-10 if MIX[i][3] == 0 else 1e-4 for i in idx]
    for s in range(1):
        assert tuned[s][1] == plain[s][1]
        for j, bar in enumerate(bars):
            assert rel_err(tuned[s][0][j], plain[s][0][j]) < bar, (s, j)
    for a, b, bar in zip(st_plain, st_tuned, field_bars):
        for k in a:
            assert rel_err(b[k], a[k]) < bar, k


def test_register_form_y_pass_against_the_single_grid_path(gpu, vpfft):
    """The batched y pass of 16- to 64-point axes runs in register form (one
    thread per pencil, the pencil in registers); the single-grid path runs the
    LDS pass.  One step of a batch against the single-grid steps of the same
    grids: FP64 to 1e-10, mixed <sigma> to 1e-6 and fields to 1e-4 (FP32
    transforms round at ~1e-7), same iteration counts."""
    cases = [((12, 48, 20), 9, 10, 1), ((10, 64, 16), 12, 20, 1), ((8, 32, 12), 6, 20, 0), ((10, 24, 8), 5, 10, 0),
             ((6, 16, 10), 4, 5, 0)]
    print("register-form y pass: batch vs single-grid steps", flush=True)
    edots = np.array([random_unit_edot(300 + i) for i in range(len(cases))])
    single = []
    for i, c in enumerate(cases):
        g = _grids(vpfft, [c], seeds=[i], seed0=90)[0]
        a = g.RunSingleStrainStep(edots[i], 30, 0.005, 1e-6)
        single.append((a, g.last_info.iterations, _state(g)))
        g.close()
    grids = _grids(vpfft, cases, seeds=list(range(len(cases))), seed0=90)
    with vpfft.Batch() as b:
        avg = b.RunSingleStrainStep(grids, edots, 30, 0.005, 1e-6)
        iters = [inf.iterations for inf in b.last_infos]
        shapes = [b.block_shape(g, 1) for g in grids]
    assert shapes == [-c[0][1] for c in cases], shapes
    for j, c in enumerate(cases):
        mixed = c[3] == 1
        assert iters[j] == single[j][1], j
        assert rel_err(avg[j], single[j][0]) < (1e-6 if mixed else 1e-10), j
        st = _state(grids[j])
        for k in st:
            assert rel_err(st[k], single[j][2][k]) < (1e-4 if mixed else 1e-10), (j, k)
    for g in grids:
        g.close()


def test_batch_of_one_matches_single_grid_path(gpu, vpfft):
    print("batch of one vs vpfft_run_strain_step", flush=True)
    dims = (24, 24, 24)
    R, _ = polycrystal(None, 27, 9, dims)
    mat = vpfft.Material.uniform(1.0, 20)
    outs = []
    for batched in (False, True):
        with vpfft.MaterialGrid(*dims, mat) as g:
            g.SetVoxelLength(1.0, 1.0, 1.0)
            g.SendDataAndInitialize(R)
            if batched:
                with vpfft.Batch() as b:
                    avg = b.RunSingleStrainStep([g], [edot_tension()], 30, 0.005, 1e-6)[0]
            else:
                avg = g.RunSingleStrainStep(edot_tension(), 30, 0.005, 1e-6)
            outs.append((avg, g.last_info.iterations, g.Stress(), g.Orientation(), g.CRSS()))
    # the reductions are grouped differently (rounding only)
    assert outs[0][1] == outs[1][1]
    assert rel_err(outs[1][0], outs[0][0]) < 1e-10
    for k in (2, 3, 4):
        assert rel_err(outs[1][k], outs[0][k]) < 1e-10


@pytest.mark.parametrize("tune", [False, "radix"])
def test_batch_matches_oracle(gpu, vpfft, oracle, tune):
    print(f"batch of 3 vs the oracle (tuning {tune})", flush=True)
    cases = [((16, 16, 16), 8, 10, 0), ((12, 18, 10), 5, 5, 0), ((20, 16, 12), 6, 20, 0)]
    grids = _grids(vpfft, cases, seed0=70)
    edots = np.array([random_unit_edot(200 + j) for j in range(3)])
    states = [dict(R=g.Orientation(), crss=g.CRSS(), acc=g.AccumulatedShear()) for g in grids]
    with vpfft.Batch(tune=tune) as b:
        avg = b.RunSingleStrainStep(grids, edots, 30, 0.005, 1e-6)
        infos = b.last_infos
    for j, (dims, grains, nexp, prec) in enumerate(cases):
        ref = oracle.strain_step(dims, vpfft.Material.uniform(1.0, nexp), states[j]["R"], states[j]["crss"], edots[j],
                                 30, 0.005, 1e-6, acc=states[j]["acc"])
        assert infos[j].iterations == ref["iterations"]
        assert rel_err(avg[j], ref["avg_stress"]) < 1e-8
        assert rel_err(grids[j].Stress(), ref["sigma"]) < 1e-8
        assert rel_err(grids[j].Orientation(), ref["R"]) < 1e-8
        assert rel_err(grids[j].CRSS(), ref["crss"]) < 1e-8
    for g in grids:
        g.close()


def test_batch_argument_errors(gpu, vpfft):
    grids = _grids(vpfft, MIX[:2])
    with vpfft.Batch() as b:
        with pytest.raises(vpfft.VPFFTError, match="twice"):
            b.RunSingleStrainStep([grids[0], grids[0]], [edot_tension()] * 2, 5, 0.005, 1e-6)
        with pytest.raises(vpfft.VPFFTError, match="zero strain rate"):
            b.RunSingleStrainStep(grids, [edot_tension(), np.zeros(5)], 5, 0.005, 1e-6)
        fresh = vpfft.MaterialGrid(8, 8, 8, vpfft.Material.uniform(1.0, 10))
        with pytest.raises(vpfft.VPFFTError, match="before upload"):
            b.RunSingleStrainStep([grids[0], fresh], [edot_tension()] * 2, 5, 0.005, 1e-6)
        fresh.close()
    for g in grids:
        g.close()
