"""The request loop (vpfft_serve): admission by slots and HBM budget,
continuous batching, FP64 reruns, and results that do not depend on the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _req(arrival, edge, grains, seed, n, k, rng_seed):
    rng = np.random.default_rng(rng_seed)
    ed = rng.standard_normal((k, 5))
    ed /= np.linalg.norm(ed, axis=1, keepdims=True)
    return {"arrival_s": arrival, "edge": edge, "grains": grains, "seed": seed, "n": n, "edots": ed, "dt": 0.005}


MIXREQ = [_req(0.00, 16, 8, 1, 10, 3, 11), _req(0.00, 24, 12, 2, 20, 2, 12), _req(0.02, 20, 6, 3, 10, 4, 13),
          _req(0.05, 32, 20, 4, 20, 1, 14), _req(0.05, 16, 5, 5, 10, 2, 15)]


def test_serve_results_do_not_depend_on_the_mix(gpu):
    from vpfft_amd import serve as S
    print("serve 5 overlapping requests, then each alone", flush=True)
    res, st, avg = S.serve(MIXREQ, max_iter=40, tol=1e-5)
    assert all(r["status"] == 0 for r in res)
    assert st["rve_steps"] == sum(len(q["edots"]) for q in MIXREQ)
    assert st["max_active_seen"] >= 2  # they shared steps
    for i, q in enumerate(MIXREQ):
        alone = dict(q, arrival_s=0.0)
        r1, _, a1 = S.serve([alone], max_iter=40, tol=1e-5)
        assert r1[0]["status"] == 0
        assert r1[0]["iterations"] == res[i]["iterations"]
        assert np.array_equal(a1[0], avg[i]), i
        assert res[i]["latency_s"] >= 0 and res[i]["finished_s"] >= res[i]["admitted_s"]


def test_serve_admits_by_budget_and_slots(gpu, vpfft):
    from vpfft_amd import serve as S
    import ctypes as C
    lib = vpfft.load_library()
    desc = vpfft.vpfft_grid_desc(24, 24, 24, vpfft.MIXED)
    b = C.c_uint64()
    assert lib.vpfft_grid_bytes(C.byref(desc), C.byref(b)) == 0
    per = b.value + 8 * 32 * 24 ** 3  # with the saved state of the FP64 rerun
    reqs = [_req(0.0, 24, 8, 10 + i, 10, 2, 20 + i) for i in range(5)]
    reqs.append(_req(0.0, 96, 8, 99, 10, 1, 99))  # larger than the budget
    res, st, _ = S.serve(reqs, hbm_budget_bytes=2 * per + per // 2, max_iter=20, tol=1e-4)
    assert [r["status"] for r in res[:5]] == [0] * 5
    assert res[5]["status"] == -2  # VPFFT_E_NOMEM
    assert st["max_active_seen"] == 2
    assert st["max_bytes_seen"] <= 2 * per + per // 2
    res2, st2, _ = S.serve(reqs[:5], max_active=1, max_iter=20, tol=1e-4)
    assert st2["max_active_seen"] == 1 and all(r["status"] == 0 for r in res2)


def test_serve_fp64_rerun_of_unconverged_mixed_steps(gpu):
    from vpfft_amd import serve as S
    q = [_req(0.0, 16, 8, 3, 10, 2, 31)]
    # a tolerance the loop cannot meet in 6 iterations, every unconverged step counted as stalled:
    # every mixed step is rerun in FP64
    res, st, avg = S.serve(q, max_iter=6, tol=1e-14, fp64_fallback=True, stall_ratio=0.0)
    assert res[0]["status"] == 0 and res[0]["fp64_reruns"] == 2 and st["fp64_reruns"] == 2
    res64, _, avg64 = S.serve(q, max_iter=6, tol=1e-14, precision=S.FP64, fp64_fallback=False)
    assert res64[0]["fp64_reruns"] == 0
    assert np.array_equal(avg[0], avg64[0])
