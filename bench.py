"""Benchmark: frequencies solved per second for a 256-frequency 1024 x 2048
ESM sweep (plate geometry, 1024 microphones, 2048 equivalent sources,
200-1500 Hz), batched mixed precision vs plain FP64 on one MI355X.

One step = one full sweep (assembly, LU, H_e, SVD, GCV selection over 50
lambdas, Tikhonov solution, surface pressure and sound power for all 256
frequencies) with every input already resident in HBM.  The headline is the
certified mixed form (the library default: complex64 bulk, nearly tied
selections solved again in FP64), which keeps the FP64 form's lambda index at
every frequency; the uncertified mixed form and the FP64 form are reported
beside it.  Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--freqs 256]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

# MI355X peaks (AMD datasheet, dense): FP32 / FP64 matrix and vector, HBM3E
PEAK_FP32_TF, PEAK_FP64_TF, PEAK_HBM_GBS = 157.3, 78.6, 8000.0


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--freqs", type=int, default=256)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=8, help="frequencies in the CPU sample")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the IBEM / HELS / radiation-mode / gradient throughputs")
    ap.add_argument("--only", choices=("both", "fp64", "mixed", "certified"), default="both",
                    help="profile runs: time one form only")
    ap.add_argument("--workload", choices=("ibem", "hels", "radiation", "gradient"), default=None,
                    help="profile runs: time this one of the other workloads only")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    return ap.parse_args()


# ------------------------------------------------------------ algorithm work
def sweep_flops(m, l):
    """Real flops of the executed ESM pipeline per frequency (a complex MAC
    is 8 flops): LU of D, the two triangular solves giving Y = H_e^H (m x l),
    the Householder QR of Y (m >= 3l/2) and the two-stage bidiagonalisation of
    its l x l R factor.  Forming U and V is never done (only the vectors a
    solve needs are pushed through the reflectors), so those terms are absent."""
    lu = 8.0 / 3.0 * m ** 3
    trsm = 8.0 * m * m * l
    qr = 8.0 * (m * l * l - l ** 3 / 3.0) if 2 * m >= 3 * l else 0.0
    band = 32.0 / 3.0 * l ** 3
    return {"factor": lu, "solve_he": trsm, "bidiag": qr + band}


def assembly_flops(m, l):
    """Assembly work per frequency (sincos, sqrt, reciprocal count 1 each)."""
    return 20.0 * (l * m + m * m) + 35.0 * m * m


def bulge_compulsory_bytes(C, elem):
    """Bytes the chase k_bulge cannot avoid per matrix: the band in (C x 33
    entries), the stage-two reflectors out (two of 33 entries per task) and
    the real bidiagonal out."""
    tasks = sum(1 + (C - 3 - s) // 32 if C - 3 - s >= 0 else 1 for s in range(C - 1))
    return C * 33 * elem + tasks * 2 * 33 * elem + 2 * C * 8


def main():
    a = _parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the sweep shards by "
              "frequency across GPUs but this build measures one GPU's share.", file=sys.stderr)
        sys.exit(2)
    import torch

    import inverseacoustics_amd as ia
    from inverseacoustics_amd import _lib, scenarios as sc

    if a.workload:
        print(json.dumps({"profile_run": a.workload, **other_workloads(ia, sc, (a.workload,))}))
        return
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
        prec, rc = {"fp64": (0, 0.0), "mixed": (1, -1.0), "certified": (1, 0.0)}[a.only]
        dt, o, st = timed(prec, rc)
        print(json.dumps({"profile_run": a.only, "ms_per_step": round(dt * 1e3, 2),
                          "phases_ms": {k: round(v, 2) for k, v in st.items()}}))
        return
    # the certified mixed form (the library default: the headline), the mixed
    # form without the FP64 recheck, and plain FP64
    t_cert, o_cert, st_cert = timed(1, 0.0)
    # the same certified sweep replayed from a HIP graph (IA_GRAPH=1: the
    # warm-up call captures the complex64 sweep, the timed calls replay it;
    # the FP64 decision of the near ties stays eager), checked bit for bit
    # against the eager outputs
    graph = None
    if a.warmup > 0 and not os.environ.get("IA_GRAPH"):
        os.environ["IA_GRAPH"] = "1"
        try:
            t_g, o_g, _ = timed(1, 0.0)
        finally:
            del os.environ["IA_GRAPH"]
        graph = {"value": round(F / t_g, 3), "ms_per_step": round(t_g * 1e3, 2),
                 "outputs_equal_eager": all(torch.equal(o_g[n], o_cert[n]) for n in o_cert)}
    t_mixed, o_mixed, st_mixed = timed(1, -1.0)
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

    db_c, idx_c, vel_c = errors(o_cert)
    db_m, idx_m, vel_m = errors(o_mixed)

    # ---- roofline of the sweep: the executed algorithm's flops over the
    # sweep's duration, timed by HIP events on the sweep's stream inside the
    # library (ia_stats), against the matrix-core peak of the form's
    # arithmetic: FP32 for the complex64 bulk, FP64 for the FP64 form.  The
    # sweep is matrix-core bound: its HBM traffic (profiles/) moves in a few
    # percent of the sweep time.  Per-phase fractions beside it.
    fl = sweep_flops(m, l)
    total_fl = sum(fl.values()) * F

    def roof(st, peak, form):
        ms = st["total"]
        ach = total_fl / (ms / 1e3) / 1e12
        phases = {}
        # (a graph-replayed sweep, IA_GRAPH=1, reports no phase times)
        for ph in ("factor", "solve_he", "bidiag"):
            t_ph = st[ph]
            if t_ph > 0:
                phases[ph] = {"flops": fl[ph] * F, "ms": round(t_ph, 3),
                              "tflops": round(fl[ph] * F / (t_ph / 1e3) / 1e12, 2),
                              "frac": round(fl[ph] * F / (t_ph / 1e3) / 1e12 / peak, 4)}
        asm = assembly_flops(m, l) * F
        if st["assemble"] > 0:
            phases["assemble"] = {"flops": asm, "ms": round(st["assemble"], 3),
                                  "tflops": round(asm / (st["assemble"] / 1e3) / 1e12, 2),
                                  "frac_valu": round(asm / (st["assemble"] / 1e3) / 1e12 / peak, 4)}
        key = (form, F, m, l, a.chunk)
        return {"bound": "mfma", "kernel": f"ESM sweep, {form} form (all kernels of one sweep)",
                "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": SWEEP_HBM_BYTES.get(key),
                "algorithmic_flops_per_sweep": total_fl, "sweep_ms": round(ms, 3),
                "flop_count": "classical: 8 real flops per complex multiply-add; the batched "
                              "GEMM executes the 3M product (6 of the 8 on the matrix cores, "
                              "the operand sums once per LDS tile)",
                "phases": phases}

    roofline = roof(st_mixed, PEAK_FP32_TF, "mixed")
    roofline_fp64 = roof(st_fp64, PEAK_FP64_TF, "fp64")

    # the chase k_bulge, the largest single launch of the hand-written
    # kernels, against the bytes it cannot avoid (it is latency-bound)
    def chase(st, elem):
        nl = max(1.0, st["bidiag_kernel_launches"])
        ms = st["bidiag_kernel_ms"] / nl
        if not ms > 0:
            return None
        cb = bulge_compulsory_bytes(l, elem) * F
        return {"kernel": f"k_bulge<{'float' if elem == 8 else 'double'}>",
                "launch_ms_avg": round(ms, 3), "compulsory_bytes": cb,
                "achieved_gbs": round(cb / (ms / 1e3) / 1e9, 1),
                "frac_hbm": round(cb / (ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4)}

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(P, a.cpu_sample, o_fp64)

    workloads = None if a.no_workloads else other_workloads(ia, sc)

    res = {
        "metric": "frequencies solved/sec, batched mixed-precision vs plain FP64 on GPU; dB error",
        "value": round(F / t_cert, 3),
        "unit": "frequencies/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_cert * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64",
        "products": "complex64 products of depth >= 64 as six bf16 partial products of "
                    "three-way split FP32 operands with FP32 accumulation (FP32-level); the "
                    "others on FP32 / FP64 matrix cores (3M); selection and eigen-solve in FP64",
        "data": "synthetic (plate, Rayleigh-integral pressures, seeded 30 dB SNR noise)",
        "graph_replay": bool(os.environ.get("IA_GRAPH", "0") not in ("", "0")),
        "config": {"workload": "ESM sweep 256 frequencies 200-1500 Hz, 1024 mics x 2048 sources "
                               "(0.6x0.5 m steel plate, d=0.02 m), Tikhonov + GCV over 50 lambdas, "
                               "certified mixed precision",
                   "frequencies": F, "l": l, "m": m},
        "speedup_certified_vs_fp64": round(t_fp64 / t_cert, 3),
        "certified_graph_replay": graph,
        "db_error_max_vs_fp64": round(db_c, 5),
        "lambda_index_match_vs_fp64": idx_c,
        "velocity_rel_l2_max_vs_fp64": vel_c,
        "rechecked_in_fp64": int(st_cert["rechecked"]),
        "recheck_ms": round(st_cert["recheck_ms"], 2),
        "recheck_full_fp64_sweep": int(st_cert["recheck_full"]),
        "target_speedup": 5.0,
        "fp64_form": {"value": round(F / t_fp64, 3), "ms_per_step": round(t_fp64 * 1e3, 2),
                      "phases_ms": {k: round(v, 2) for k, v in st_fp64.items()}},
        "mixed_uncertified": {"value": round(F / t_mixed, 3), "ms_per_step": round(t_mixed * 1e3, 2),
                              "speedup_vs_fp64": round(t_fp64 / t_mixed, 3),
                              "db_error_max_vs_fp64": round(db_m, 5),
                              "lambda_index_match_vs_fp64": idx_m,
                              "velocity_rel_l2_max_vs_fp64": vel_m,
                              "phases_ms": {k: round(v, 2) for k, v in st_mixed.items()}},
        "roofline": roofline,
        "roofline_fp64_form": roofline_fp64,
        "chase_kernel": {"mixed": chase(st_mixed, 8), "fp64": chase(st_fp64, 16)},
        "cpu_baseline": cpu,
        "other_workloads": workloads,
        "input_generation_s": round(gen_s, 2),
    }
    print(json.dumps(res))


# HBM bytes of one mixed / FP64 sweep (every kernel's FETCH_SIZE + WRITE_SIZE
# from rocprofv3 --pmc, FETCH scaled by the WRITE/FETCH ratio of a plain
# device copy measured in the same runs), keyed by (form, frequencies, m, l,
# chunk): profiles/esm_sweep_256f_1024x2048_hbm_traffic.txt.  Well above the
# algorithmic bytes: the rank-128 / rank-64 trailing updates re-read their
# operand panels per output tile.
SWEEP_HBM_BYTES = {
    ("mixed", 256, 2048, 1024, 0): 838809403715,
    ("fp64", 256, 2048, 1024, 0): 1834993596342,
}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(P, nf, o_fp64):
    """The oracle's numpy/LAPACK FP64 pipeline (the vectorised esmV + csvd +
    GCV + Tikhonov formulation) on the host cores, on a bounded sample of nf
    frequencies spread over the sweep (1 warm-up frequency first), with all
    threads and with one thread; each sampled frequency is also checked
    against the GPU FP64 sweep (same lambda index, W and velocity within
    1e-10)."""
    import torch
    from threadpoolctl import threadpool_info, threadpool_limits

    from oracle import pipeline

    idx = np.linspace(0, len(P.k) - 1, nf).round().astype(int)
    W64 = o_fp64["W"].cpu().numpy()
    I64 = o_fp64["idx"].cpu().numpy()
    V64 = torch.view_as_complex(o_fp64["v"]).cpu().numpy()

    def solve(f):
        return pipeline.esm_solve(P.k[f], P.r_s, P.r_h, P.src, P.b[f], P.area, P.normals, P.rho,
                                  P.c)

    threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    solve(idx[0])
    t = time.perf_counter()
    outs = [solve(f) for f in idx]
    dt_all = time.perf_counter() - t
    with threadpool_limits(limits=1):
        solve(idx[0])
        n1 = max(2, nf // 2)
        t = time.perf_counter()
        for f in idx[:n1]:
            solve(f)
        dt_1 = time.perf_counter() - t
    match = [int(I64[f] == o["idx"]) for f, o in zip(idx, outs)]
    w_err = [abs(W64[f] / o["W"] - 1) for f, o in zip(idx, outs)]
    v_err = [float(np.linalg.norm(V64[f] - o["x"]) / np.linalg.norm(o["x"])) for f, o in zip(idx, outs)]
    return {"value": round(nf / dt_all, 4), "unit": "frequencies/s", "cores": threads,
            "kind": "port", "cpu_model": _cpu_model(), "os_cpu_count": os.cpu_count(),
            "sample": f"{nf} of the {len(P.k)} frequencies (indices {idx.tolist()}), 1024 x 2048, "
                      f"FP64 numpy/LAPACK oracle, after 1 warm-up frequency",
            "one_thread": {"value": round(n1 / dt_1, 4), "cores": 1,
                           "sample": f"the first {n1} of those frequencies"},
            "check_vs_gpu_fp64": {"lambda_index_match": f"{sum(match)}/{nf}",
                                  "W_rel_err_max": float(max(w_err)),
                                  "velocity_rel_l2_max": max(v_err)}}


def other_workloads(ia, sc, which=("ibem", "hels", "radiation", "gradient")):
    """Throughput of the other solves at their full sizes (host inputs and
    outputs, one warm-up call then one timed call each)."""
    import torch

    out = {}

    def clock(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    if "ibem" in which:
        S = sc.ibem_scenario()
        ia.set_air(S.c, S.rho)
        ia.set_quadrature()
    for prec in ("fp64", "mixed") if "ibem" in which else ():
        dt = clock(lambda: ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area, precision=prec))
        out[f"ibem_{prec}"] = {"workload": "IBEM icosphere(3) 642 nodes, 128 mics, 64 frequencies "
                                           "100-2000 Hz, 7-point rule",
                               "frequencies_per_s": round(len(S.k) / dt, 2), "s": round(dt, 4)}
        # the device side captured into a HIP graph (IA_GRAPH=1) and replayed
        os.environ["IA_GRAPH"] = "1"
        try:
            dt = clock(lambda: ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area, precision=prec))
        finally:
            del os.environ["IA_GRAPH"]
        out[f"ibem_{prec}"]["graph_replay"] = {"frequencies_per_s": round(len(S.k) / dt, 2),
                                               "s": round(dt, 4)}
    if "hels" in which:
        H = sc.hels_scenario()
        ia.set_air(H.c, H.rho)
    for prec in ("fp64", "mixed") if "hels" in which else ():
        dt = clock(lambda: ia.hels_sweep(H.k, H.order, H.azelr_s, H.azelr_h, H.b, H.area,
                                         precision=prec))
        out[f"hels_{prec}"] = {"workload": "HELS 256 mics, 256 surface points, 128 frequencies "
                                           "50-3000 Hz, N = ceil(ka) + 4 (5..11)",
                               "frequencies_per_s": round(len(H.k) / dt, 2), "s": round(dt, 4)}
        # the same sweep captured into a HIP graph (IA_GRAPH=1): the warm-up
        # call captures, the timed call replays
        os.environ["IA_GRAPH"] = "1"
        try:
            dt = clock(lambda: ia.hels_sweep(H.k, H.order, H.azelr_s, H.azelr_h, H.b, H.area,
                                             precision=prec))
        finally:
            del os.environ["IA_GRAPH"]
        out[f"hels_{prec}"]["graph_replay"] = {"frequencies_per_s": round(len(H.k) / dt, 2),
                                               "s": round(dt, 4)}
    if "radiation" in which:
        R = sc.radiation_scenario(device="cuda")
        ia.set_air(R.c, R.rho)
    for prec in ("fp64", "mixed") if "radiation" in which else ():
        dt = clock(lambda: ia.radiation_sweep(R.k, R.r_p, R.patch_area, R.r_h, R.b, 128,
                                              precision=prec))
        out[f"radiation_{prec}"] = {
            "workload": "radiation modes, 64 x 64 patches (4096), 1024 mics, 256 frequencies "
                        "50-2500 Hz, 128 modes, FP64 eigendecomposition",
            "frequencies_per_s": round(len(R.k) / dt, 2), "s": round(dt, 4)}
        # captured into a HIP graph (IA_GRAPH=1): the warm-up call runs eagerly
        # and captures, the timed call replays
        os.environ["IA_GRAPH"] = "1"
        try:
            dt = clock(lambda: ia.radiation_sweep(R.k, R.r_p, R.patch_area, R.r_h, R.b, 128,
                                                  precision=prec))
        finally:
            del os.environ["IA_GRAPH"]
        out[f"radiation_{prec}"]["graph_replay"] = {
            "frequencies_per_s": round(len(R.k) / dt, 2), "s": round(dt, 4)}
    if "gradient" not in which:
        return out
    G = sc.plate_scenario(n_freq=32, device="cuda")
    ia.set_air(G.c, G.rho)
    dt = clock(lambda: ia.esm_gradient(G.k, G.r_s, G.normals, 0.02, G.r_h, G.b, G.area))
    out["esm_gradient_fp64"] = {"workload": "ESM adjoint: misfit and power gradients w.r.t. d, "
                                            "1024 mic positions and log lambda, 32 frequencies, "
                                            "1024 x 2048",
                                "frequencies_per_s": round(len(G.k) / dt, 2), "s": round(dt, 4)}
    return out


if __name__ == "__main__":
    main()
