#!/usr/bin/env python3
"""Throughput of the variational-MPS inner loop on one MI355X.

Workload: the bulk-site effective-Hamiltonian matvec of the XXZ chain
(J = 1, Delta = 0.8, N = 100, d = 2, MPO bond 5) at bond dimension chi = 1024 --
the operation ARPACK's znaupd loop applies inside every site optimisation
(core.f95:2505-2513, 2529-2556).  One step = one H_eff application; inputs are
resident in HBM before the timed region and the timing uses HIP events on the
stream the kernels run on.

    python bench.py [--steps K] [--warmup W] [--chi 1024] [--cpu-matvecs 3] [--no-cpu]

Prints one JSON line.  The roofline object describes the dominant kernel (the
grouped complex FP64 GEMM, two launches per matvec); cpu_baseline times the
oracle's restatement of the reference's strategy-3 matvec (dense stage-1 tensor
+ two zgemms through OpenBLAS) on a bounded sample on the host cores.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nutcracker-1_amd"))

METRIC = "matvecs/s + sweep sec at chi=1024, mixed vs FP64 on 1 GPU; |dE| vs FP64"
FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def hermitian_env(rng, b, c):
    e = (rng.standard_normal((b, b, c)) + 1j * rng.standard_normal((b, b, c))) * np.sqrt(0.5)
    e = e + e.transpose(1, 0, 2).conj()
    return np.asfortranarray(e)


def flops_matvec(bl, br, d, n_from, n_to, nblk):
    """Minimal block-sparse work of one H_eff application (real flops)."""
    return 8 * d * bl * br * (n_from * bl + n_to * br) + 8 * nblk * bl * br * d * d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--chi", type=int, default=1024)
    ap.add_argument("--delta", type=float, default=0.8)
    ap.add_argument("--cpu-matvecs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one MI355X only: a run on %d GPUs is not supported; rerun with --gpus 1"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from nutcracker_amd import core
    from nutcracker_amd.operators import xxz_operator

    lib = core.lib()
    lib.nc_bench_heff.restype = C.c_int
    lib.nc_bench_heff.argtypes = [C.c_uint32] * 6 + [C.POINTER(C.c_uint32)] + [C.c_void_p] * 5 + \
        [C.c_uint32, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]

    dev = core.device_info()
    chi = args.chi
    d = 2
    op = xxz_operator(3, d, 1.0, args.delta)[1]  # a bulk site: w = 5, 8 blocks
    bl = br = chi
    cl, cr = op.cl, op.cr
    rng = np.random.default_rng(args.seed)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    x = np.asfortranarray((rng.standard_normal((br, bl, d)) + 1j * rng.standard_normal((br, bl, d))) * np.sqrt(0.5))
    x /= np.linalg.norm(x)
    y = np.zeros_like(x)
    idx = np.ascontiguousarray(np.asarray(op.idx, np.uint32).reshape(-1, order="F"))
    mats = np.ascontiguousarray(np.asarray(op.mats).reshape(-1, order="F"))

    total = C.c_double(0)
    phases = (C.c_double * 3)()
    rc = lib.nc_bench_heff(bl, br, cl, cr, d, op.nmat, idx.ctypes.data_as(C.POINTER(C.c_uint32)), mats.ctypes.data,
                           L.ctypes.data, R.ctypes.data, x.ctypes.data, y.ctypes.data, args.warmup, args.steps,
                           C.byref(total), phases)
    if rc != 0:
        raise RuntimeError(core.last_error())
    ms_per_step = total.value / args.steps
    value = 1000.0 / ms_per_step

    n_from = len(set(int(v) for v in op.idx[0]))
    n_to = len(set(int(v) for v in op.idx[1]))
    f_mv = flops_matvec(bl, br, d, n_from, n_to, op.nmat)
    f_gemm_a = 8 * br * bl * bl * d * n_from
    f_gemm_c = 8 * br * br * bl * d * n_to
    t_a, t_mix, t_c = phases[0], phases[1], phases[2]
    gemm_avg_ms = (t_a + t_c) / 2
    achieved = (f_gemm_a + f_gemm_c) / 2 / (gemm_avg_ms * 1e-3) / 1e12

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "XXZ chain J=1 Delta=%g N=100 d=2 MPO bond 5: bulk-site H_eff matvec at chi=%d"
                               % (args.delta, chi),
                   "chi": chi, "d": d, "mpo_bond": cl, "mpo_blocks": op.nmat, "form": "fp64 complex (plain drop-in)"},
        "matvec_tflops": round(f_mv / (ms_per_step * 1e-3) / 1e12, 3),
        "phase_ms": {"gemm_step_a": round(t_a, 4), "mix": round(t_mix, 4), "gemm_step_c": round(t_c, 4)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_MFMA_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4), "traffic": None,
                     "kernel": "zgemm_f64_kernel", "flops_per_launch": (f_gemm_a + f_gemm_c) // 2,
                     "avg_launch_ms": round(gemm_avg_ms, 4)},
        "device": dev,
    }

    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        import oracle.core as oc  # CPU baseline only
        try:
            from threadpoolctl import threadpool_info
            threads = max([p.get("num_threads", 1) for p in threadpool_info() if p.get("internal_api") == "openblas"]
                          or [os.cpu_count()])
        except Exception:
            threads = os.cpu_count()
        t0 = time.perf_counter()
        stage1 = oc.iteration_stage_1(bl, cl, cr, d, L, op.idx, op.mats)
        t_stage1 = time.perf_counter() - t0
        times = []
        yc = None
        for _ in range(max(1, args.cpu_matvecs)):
            t0 = time.perf_counter()
            yc = oc.matvec(bl, br, cl, cr, d, L, op.idx, op.mats, R, x, stage1)
            times.append(time.perf_counter() - t0)
        cpu_ms = float(np.median(times)) * 1e3
        err = float(np.linalg.norm((y - yc).ravel()) / np.linalg.norm(yc.ravel()))
        result["cpu_baseline"] = {
            "value": round(1000.0 / cpu_ms, 4), "unit": "matvecs/s", "cores": int(threads), "kind": "port",
            "sample": "%d oracle strategy-3 matvecs (dense stage-1 tensor built once in %.1f s, then 2 zgemms each, "
                      "OpenBLAS) at the same chi=%d bulk site; median %.1f ms" % (len(times), t_stage1, chi, cpu_ms),
        }
        result["gpu_vs_cpu_matvec_speedup"] = round(cpu_ms / ms_per_step, 2)
        result["max_rel_err_vs_oracle"] = err
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
