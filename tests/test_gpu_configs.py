"""The IBEM and HELS workloads at their full sizes, each solved in one
batched call, against the oracle.

* IBEM: icosphere(3) of radius 0.1 m (642 nodes, 1280 triangles), 128
  Fibonacci microphones at 0.25 m, 64 frequencies 100-2000 Hz, 7-point rule,
  30 dB SNR: the FP64 form against the oracle (ibemS.m restatement) at six of
  the 64 frequencies, the mixed form at all 64 within 0.1 dB of FP64.
* HELS: 256 microphones at 0.3 m, 128 frequencies 50-3000 Hz, 256 surface
  points on r = a = 0.12 m, N = min(16, ceil(ka) + 4): the FP64 form against
  the oracle (helsV.m restatement) at all 128 frequencies, the mixed form
  within 0.1 dB.

FP64 tolerance: max(1e-10, 10 x the spread between two backward-stable FP64
oracles (LAPACK gesdd vs gesvd)) relative, the same lambda index and TSVD k."""
import numpy as np
import pytest

from inverseacoustics_amd import scenarios as sc
from oracle import ibem as ib
from oracle import pipeline, post

pytestmark = pytest.mark.gpu


def _check(r, f, o, o2, key_v, key_p):
    spread = np.linalg.norm(o2[key_v] - o[key_v]) / np.linalg.norm(o[key_v])
    tol = max(1e-10, 10 * spread)
    assert r.lam_index[f] == o["idx"] and r.k_tsvd[f] == o["k_tsvd"], f
    assert np.linalg.norm(r.velocity[f] - o[key_v]) <= tol * np.linalg.norm(o[key_v]), (f, spread)
    assert np.linalg.norm(r.pressure[f] - o[key_p]) <= tol * np.linalg.norm(o[key_p]), (f, spread)
    assert r.power[f] == pytest.approx(o["W"], rel=tol)
    return spread


def test_ibem_full_size_sweep(ia):
    ia.set_quadrature()
    rule, pages = ib.gauss7(), ib.duffy_pages(40)
    S = sc.ibem_scenario()  # 64 frequencies, icosphere(3), 128 mics
    assert (len(S.k), len(S.X), len(S.E), len(S.r_h)) == (64, 642, 1280, 128)
    ia.set_air(S.c, S.rho)
    r = ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area)
    assert np.all(r.status == 0)
    for f in (0, 13, 27, 40, 52, 63):
        o = pipeline.ibem_solve(S.k[f], S.X, S.E, S.r_h, S.b[f], S.area, rule, pages, rho=S.rho,
                                c=S.c)
        o2 = pipeline.ibem_solve(S.k[f], S.X, S.E, S.r_h, S.b[f], S.area, rule, pages, rho=S.rho,
                                 c=S.c, driver="gesvd")
        spread = _check(r, f, o, o2, "v", "p")
        print(f"ibem f={S.freqs[f]:.1f} Hz: idx {r.lam_index[f]} W {r.power[f]:.9e} "
              f"oracle {o['W']:.9e} spread {spread:.1e}")
    # the reconstruction tracks the analytic pulsating + oscillating sphere
    assert np.max(np.abs(post.power_db(r.power) - post.power_db(S.truth["W"]))) < 1.0
    rm = ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area, precision="mixed")
    d = np.abs(post.power_db(rm.power) - post.power_db(r.power))
    print(f"ibem mixed: max dB error {d.max():.4f}, lambda index match "
          f"{np.mean(rm.lam_index == r.lam_index):.3f}")
    assert d.max() < 0.1


def test_hels_full_size_sweep(ia):
    S = sc.hels_scenario()  # 128 frequencies, 256 mics, 256 surface points
    assert (len(S.k), len(S.azelr_h), len(S.azelr_s)) == (128, 256, 256)
    assert S.order.min() == 5 and S.order.max() == 11  # uneven system sizes
    ia.set_air(S.c, S.rho)
    r = ia.hels_sweep(S.k, S.order, S.azelr_s, S.azelr_h, S.b, S.area)
    worst = 0.0
    for f in range(len(S.k)):
        o = pipeline.hels_solve(S.k[f], S.order[f], S.azelr_s, S.azelr_h, S.b[f], S.area, S.rho,
                                S.c)
        o2 = pipeline.hels_solve(S.k[f], S.order[f], S.azelr_s, S.azelr_h, S.b[f], S.area, S.rho,
                                 S.c, driver="gesvd")
        _check(r, f, o, o2, "v", "p")
        worst = max(worst, np.linalg.norm(r.velocity[f] - o["v"]) / np.linalg.norm(o["v"]))
    print(f"hels: worst velocity rel err {worst:.2e} over 128 frequencies")
    rm = ia.hels_sweep(S.k, S.order, S.azelr_s, S.azelr_h, S.b, S.area, precision="mixed")
    d = np.abs(post.power_db(rm.power) - post.power_db(r.power))
    print(f"hels mixed: max dB error {d.max():.4f}, lambda index match "
          f"{np.mean(rm.lam_index == r.lam_index):.3f}")
    assert d.max() < 0.1
