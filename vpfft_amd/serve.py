"""Request loop over the batched strain-step engine (include/vpfft_serve.h).

serve(requests, ...) admits requests in arrival order by an HBM budget, keeps
their grids resident between increments and advances every admitted request
by one increment per batched step.  Requests are dicts:
    {"arrival_s", "edge", "grains", "seed", "n", "edots" (k x 5), "dt"}
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import FP64, MIXED, _check, load_library


class vpfft_request(C.Structure):
    _fields_ = [
        ("arrival_s", C.c_double),
        ("edge", C.c_int),
        ("grains", C.c_int),
        ("seed", C.c_uint64),
        ("rate_sensitivity", C.c_int),
        ("num_increments", C.c_int),
        ("edot", C.POINTER(C.c_double)),
        ("dt", C.c_double),
        ("avg_stress", C.POINTER(C.c_double)),
    ]


class vpfft_request_result(C.Structure):
    _fields_ = [
        ("admitted_s", C.c_double),
        ("finished_s", C.c_double),
        ("latency_s", C.c_double),
        ("increments_done", C.c_int),
        ("fp64_reruns", C.c_int),
        ("iterations", C.c_int),
        ("status", C.c_int),
    ]


class vpfft_serve_config(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("max_active", C.c_int),
        ("precision", C.c_int),
        ("fp64_fallback", C.c_int),
        ("stall_ratio", C.c_double),
        ("max_iter", C.c_int),
        ("tol", C.c_double),
        ("tune", C.c_int),
    ]


class vpfft_serve_stats(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double),
        ("rve_steps", C.c_long),
        ("rve_steps_per_s", C.c_double),
        ("voxel_iters_per_s", C.c_double),
        ("p50_latency_s", C.c_double),
        ("p99_latency_s", C.c_double),
        ("mean_latency_s", C.c_double),
        ("max_active_seen", C.c_int),
        ("max_bytes_seen", C.c_uint64),
        ("batched_steps", C.c_int),
        ("fp64_reruns", C.c_int),
    ]


def _struct_dict(s) -> dict:
    return {name: getattr(s, name) for name, _ in s._fields_}


def serve(requests: Sequence[dict], hbm_budget_bytes: int = 64 << 30, max_active: int = 64,
          precision: int = MIXED, fp64_fallback: bool = True, max_iter: int = 100, tol: float = 1e-5,
          tune: bool = True, stall_ratio: float = 0.98):
    """Serve the requests to completion.  Returns (results, stats, avg_stress)
    where avg_stress[i] is request i's <sigma> per increment (k x 5).
    stall_ratio: a mixed step that ends unconverged is rerun in FP64 when its
    error fell by less than 2% over its last five iterations (the FP32
    rounding floor); slow but steady convergence (the same in FP64) is not."""
    lib = load_library()
    lib.vpfft_serve.argtypes = [C.POINTER(vpfft_serve_config), C.POINTER(vpfft_request), C.c_int,
                                C.POINTER(vpfft_request_result), C.POINTER(vpfft_serve_stats)]
    lib.vpfft_serve.restype = C.c_int
    count = len(requests)
    reqs = (vpfft_request * max(count, 1))()
    keep: List[np.ndarray] = []
    outs: List[np.ndarray] = []
    for i, q in enumerate(requests):
        e = np.ascontiguousarray(q["edots"], dtype=np.float64).reshape(-1, 5)
        o = np.zeros_like(e)
        keep.append(e)
        outs.append(o)
        r = reqs[i]
        r.arrival_s = float(q["arrival_s"])
        r.edge = int(q["edge"])
        r.grains = int(q["grains"])
        r.seed = int(q["seed"])
        r.rate_sensitivity = int(q["n"])
        r.num_increments = e.shape[0]
        r.edot = e.ctypes.data_as(C.POINTER(C.c_double))
        r.dt = float(q.get("dt", 0.005))
        r.avg_stress = o.ctypes.data_as(C.POINTER(C.c_double))
    cfg = vpfft_serve_config(int(hbm_budget_bytes), int(max_active), int(precision), 1 if fp64_fallback else 0,
                             float(stall_ratio), int(max_iter), float(tol), 1 if tune else 0)
    res = (vpfft_request_result * max(count, 1))()
    st = vpfft_serve_stats()
    _check(lib.vpfft_serve(C.byref(cfg), reqs, count, res, C.byref(st)))
    return [_struct_dict(res[i]) for i in range(count)], _struct_dict(st), outs


def poisson_requests(count: int, rate: float, seed: int = 0, edges=range(32, 129, 16), grains=(8, 256),
                     exponents=(10, 20), increments=(1, 10), dt: float = 0.005) -> List[dict]:
    """The serving workload: Poisson arrivals at `rate` per second; each request
    one RVE (edge from `edges`, grains log-uniform, n from `exponents`) and
    1-10 increments of random unit traceless strain rates."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    edges = list(edges)
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        k = int(rng.integers(increments[0], increments[1] + 1))
        ed = rng.standard_normal((k, 5))
        ed /= np.linalg.norm(ed, axis=1, keepdims=True)
        out.append({
            "arrival_s": t, "edge": int(rng.choice(edges)),
            "grains": int(round(float(np.exp(rng.uniform(np.log(grains[0]), np.log(grains[1])))))),
            "seed": 7000 + i + 100000 * seed, "n": int(rng.choice(exponents)), "edots": ed, "dt": dt,
        })
    return out


__all__ = ["serve", "poisson_requests", "FP64", "MIXED"]
