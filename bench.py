"""Benchmark: frequencies solved per second for a 256-frequency 1024 x 2048
ESM sweep (plate geometry, 1024 microphones, 2048 equivalent sources,
200-1500 Hz), batched mixed precision vs plain FP64 on one MI355X.

One step = one full sweep (assembly, LU, H_e, SVD, GCV selection over 50
lambdas, Tikhonov solution, surface pressure and sound power for all 256
frequencies) with every input already resident in HBM.  Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--freqs 256]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--freqs", type=int, default=256)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=3, help="frequencies in the CPU sample")
    ap.add_argument("--only", choices=("both", "fp64", "mixed"), default="both",
                    help="profile runs: time one form only")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    return ap.parse_args()


def main():
    a = _parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the sweep shards by "
              "frequency across GPUs but this build measures one GPU's share.", file=sys.stderr)
        sys.exit(2)
    import torch

    import inverseacoustics_amd as ia
    from inverseacoustics_amd import _lib, scenarios as sc

    dev = torch.device("cuda:0")
    L = _lib.lib()
    t0 = time.time()
    P = sc.plate_scenario(n_freq=a.freqs, device="cuda")
    gen_s = time.time() - t0
    ia.set_air(P.c, P.rho)
    m, l, F = P.m, P.l, len(P.k)

    def dten(x, dt):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=dev)

    k_d = dten(P.k, torch.float64)
    rs_d = dten(sc.soa(P.r_s), torch.float64)
    src_d = dten(sc.soa(P.src), torch.float64)
    nrm_d = dten(sc.soa(P.normals), torch.float64)
    rh_d = dten(sc.soa(P.r_h), torch.float64)
    b_d = torch.view_as_real(dten(P.b, torch.complex128)).contiguous()

    def outputs():
        return dict(lam=torch.zeros(F, dtype=torch.float64, device=dev),
                    idx=torch.zeros(F, dtype=torch.int32, device=dev),
                    k=torch.zeros(F, dtype=torch.int32, device=dev),
                    W=torch.zeros(F, dtype=torch.float64, device=dev),
                    smax=torch.zeros(F, dtype=torch.float64, device=dev),
                    v=torch.zeros(F, m, 2, dtype=torch.float64, device=dev),
                    p=torch.zeros(F, m, 2, dtype=torch.float64, device=dev),
                    st=torch.zeros(F, dtype=torch.int32, device=dev))

    stream = torch.cuda.current_stream(dev)

    def run(prec, o, recheck=0.0):
        desc = _lib.ia_sweep_desc(m, l, F, prec, ia._reg("tikhonov", "gcv", 50, 8.0), P.area,
                                  a.chunk, recheck)
        out = _lib.ia_sweep_out(o["lam"].data_ptr(), o["idx"].data_ptr(), o["k"].data_ptr(),
                                o["W"].data_ptr(), o["smax"].data_ptr(), o["v"].data_ptr(),
                                o["p"].data_ptr(), o["st"].data_ptr())
        ia.check(L.ia_esm_sweep_device(desc, k_d.data_ptr(), rs_d.data_ptr(), src_d.data_ptr(),
                                       nrm_d.data_ptr(), rh_d.data_ptr(), b_d.data_ptr(), out,
                                       stream.cuda_stream))

    def timed(prec, recheck=0.0):
        o = outputs()
        for _ in range(a.warmup):
            run(prec, o, recheck)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(a.steps):
            run(prec, o, recheck)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / a.steps
        st = _lib.ia_stats()
        ia.check(L.ia_last_stats(st))
        return dt, o, {n: getattr(st, n) for n, _ in _lib.ia_stats._fields_}

    if a.copy_probe_mib:
        n = a.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if a.only != "both":
        dt, o, st = timed(1 if a.only == "mixed" else 0, -1.0 if a.only == "mixed" else 0.0)
        print(json.dumps({"profile_run": a.only, "ms_per_step": round(dt * 1e3, 2),
                          "phases_ms": {k: round(v, 2) for k, v in st.items()}}))
        return
    # mixed form as is (the headline), the certified mixed form (nearly tied
    # GCV choices solved again in FP64, the library default) and plain FP64
    t_mixed, o_mixed, st_mixed = timed(1, -1.0)
    t_cert, o_cert, st_cert = timed(1, 0.0)
    t_fp64, o_fp64, st_fp64 = timed(0)

    W64 = o_fp64["W"].cpu().numpy()
    idx64 = o_fp64["idx"].cpu().numpy()
    v64 = torch.view_as_complex(o_fp64["v"]).cpu().numpy()

    def errors(o):
        W = o["W"].cpu().numpy()
        v = torch.view_as_complex(o["v"]).cpu().numpy()
        return (float(np.max(np.abs(10 * np.log10(W / W64)))),
                float(np.mean(idx64 == o["idx"].cpu().numpy())),
                float(np.max(np.linalg.norm(v - v64, axis=1) / np.linalg.norm(v64, axis=1))))

    db_err, idx_match, vel_rel = errors(o_mixed)
    db_c, idx_c, vel_c = errors(o_cert)

    # ---- roofline of the chase kernel k_bulge (HBM-bound): the second stage
    # of the two-stage bidiagonalisation, our largest single launch of the
    # sweep (the rocBLAS GEMMs of the factorisations and the band reduction
    # take more time in aggregate, spread over ~600 calls of many shapes).
    # Algorithmic bytes per launch: per task of the chase the overlap block
    # and the block right of it are read and the overlap block and the rows
    # above it written (bulge_alg_bytes), over all frequencies of the launch.
    # The library brackets the launch with HIP events on the sweep's stream
    # and reports its duration, bytes and launch count.
    def roof(st, form, esz):
        n = max(1.0, st["bidiag_kernel_launches"])
        ms = st["bidiag_kernel_ms"] / n
        by = st["bidiag_kernel_bytes"] / n
        ach = by / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        key = (form, F, m, l, a.chunk)
        tr = HBM_TRAFFIC_PER_LAUNCH.get(key)
        return {"bound": "hbm", "kernel": f"k_bulge<{'float' if esz == 8 else 'double'}> ({form} form)",
                "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(ach / 8000.0, 4), "traffic": tr,
                "algorithmic_bytes_per_launch": round(by), "launch_ms_avg": round(ms, 4),
                "launches": int(n)}

    roofline = roof(st_mixed, "mixed", 8)
    roofline_fp64 = roof(st_fp64, "fp64", 16)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(P, a.cpu_sample)

    res = {
        "metric": "frequencies solved/sec, batched mixed-precision vs plain FP64 on GPU; dB error",
        "value": round(F / t_mixed, 3),
        "unit": "frequencies/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_mixed * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64",
        "data": "synthetic (plate, Rayleigh-integral pressures, seeded 30 dB SNR noise)",
        "config": {"workload": "ESM sweep 256 frequencies 200-1500 Hz, 1024 mics x 2048 sources "
                               "(0.6x0.5 m steel plate, d=0.02 m), Tikhonov + GCV over 50 lambdas",
                   "frequencies": F, "l": l, "m": m},
        "fp64_form": {"value": round(F / t_fp64, 3), "ms_per_step": round(t_fp64 * 1e3, 2),
                      "phases_ms": {k: round(v, 2) for k, v in st_fp64.items()}},
        "mixed_phases_ms": {k: round(v, 2) for k, v in st_mixed.items()},
        "mixed_certified": {"value": round(F / t_cert, 3), "ms_per_step": round(t_cert * 1e3, 2),
                            "rechecked_in_fp64": int(st_cert["rechecked"]),
                            "recheck_ms": round(st_cert["recheck_ms"], 2),
                            "db_error_max_vs_fp64": round(db_c, 5),
                            "lambda_index_match_vs_fp64": idx_c,
                            "velocity_rel_l2_max_vs_fp64": vel_c},
        "speedup_mixed_vs_fp64": round(t_fp64 / t_mixed, 3),
        "db_error_max_mixed_vs_fp64": round(db_err, 5),
        "lambda_index_match_mixed_vs_fp64": idx_match,
        "velocity_rel_l2_max_mixed_vs_fp64": vel_rel,
        "roofline": roofline,
        "roofline_fp64_form": roofline_fp64,
        "cpu_baseline": cpu,
        "input_generation_s": round(gen_s, 2),
    }
    print(json.dumps(res))


# HBM bytes per k_bulge launch (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; FETCH scaled by
# 2.0, the WRITE/FETCH ratio of a plain 1 GiB device copy in the same runs) keyed by
# (form, frequencies, m, l, chunk): profiles/bulge_esm_sweep_256f_1024x2048_hbm_counters.txt
HBM_TRAFFIC_PER_LAUNCH = {
    ("fp64", 256, 2048, 1024, 0): 297710959119,
    ("mixed", 256, 2048, 1024, 0): 169788622616,
}


def cpu_baseline(P, nf):
    """The oracle's numpy/LAPACK FP64 pipeline on host cores, on a bounded
    sample of nf frequencies of the same workload."""
    from threadpoolctl import threadpool_info

    from oracle import pipeline

    idx = np.linspace(0, len(P.k) - 1, nf).round().astype(int)
    pipeline.esm_solve(P.k[idx[0]], P.r_s, P.r_h, P.src, P.b[idx[0]], P.area, P.normals, P.rho, P.c)
    t = time.perf_counter()
    for f in idx:
        pipeline.esm_solve(P.k[f], P.r_s, P.r_h, P.src, P.b[f], P.area, P.normals, P.rho, P.c)
    dt = time.perf_counter() - t
    threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    return {"value": round(nf / dt, 4), "unit": "frequencies/s", "cores": threads, "kind": "port",
            "sample": f"{nf} of the 256 frequencies (1024x2048, FP64 numpy/LAPACK oracle), "
                      f"after 1 warm-up frequency"}


if __name__ == "__main__":
    main()
