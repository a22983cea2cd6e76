"""Graph-captured sweep (IA_GRAPH=1): ia_esm_sweep_device captures the whole
sweep once into a HIP graph for its sizes, parameters and buffer addresses
and replays it.  Every replay is checked against the eager sweep on the same
device buffers, with eager work and synchronisations of the program between
the replays; a replay after the pressures in the same buffer were doubled
gives the doubled solution (the sweep is linear in b at a fixed lambda index,
and GCV's choice does not change under scaling), so a replay reads its inputs
afresh.  FP64, uncertified mixed and certified mixed forms (the certified
recheck runs eagerly after the captured complex64 sweep)."""
import numpy as np
import pytest

from inverseacoustics_amd import scenarios as sc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,recheck", [(0, 0.0), (1, -1.0), (1, 0.0)])
def test_graph_replays_match_eager(ia, monkeypatch, prec, recheck):
    import torch

    from inverseacoustics_amd import _lib

    P = sc.plate_scenario(freqs=np.linspace(300.0, 1400.0, 12), mic_n=(16, 16), mic_pitch=0.06,
                          src_n=(32, 16), ref_n=(64, 50))
    ia.set_air(P.c, P.rho)
    dev = torch.device("cuda:0")
    m, l, F = P.m, P.l, len(P.k)

    def dt(x, t):
        return torch.as_tensor(np.ascontiguousarray(x), dtype=t, device=dev)

    k = dt(P.k, torch.float64)
    rs, src = dt(sc.soa(P.r_s), torch.float64), dt(sc.soa(P.src), torch.float64)
    nrm, rh = dt(sc.soa(P.normals), torch.float64), dt(sc.soa(P.r_h), torch.float64)
    b = torch.view_as_real(dt(P.b, torch.complex128)).contiguous()
    o = dict(lam=torch.zeros(F, dtype=torch.float64, device=dev),
             idx=torch.zeros(F, dtype=torch.int32, device=dev),
             k=torch.zeros(F, dtype=torch.int32, device=dev),
             W=torch.zeros(F, dtype=torch.float64, device=dev),
             smax=torch.zeros(F, dtype=torch.float64, device=dev),
             v=torch.zeros(F, m, 2, dtype=torch.float64, device=dev),
             p=torch.zeros(F, m, 2, dtype=torch.float64, device=dev),
             st=torch.zeros(F, dtype=torch.int32, device=dev))
    L = _lib.lib()
    stream = torch.cuda.current_stream(dev)

    def run():
        desc = _lib.ia_sweep_desc(m, l, F, prec, ia._reg("tikhonov", "gcv", 50, 8.0), P.area, 0,
                                  recheck)
        out = _lib.ia_sweep_out(o["lam"].data_ptr(), o["idx"].data_ptr(), o["k"].data_ptr(),
                                o["W"].data_ptr(), o["smax"].data_ptr(), o["v"].data_ptr(),
                                o["p"].data_ptr(), o["st"].data_ptr())
        ia.check(L.ia_esm_sweep_device(desc, k.data_ptr(), rs.data_ptr(), src.data_ptr(),
                                       nrm.data_ptr(), rh.data_ptr(), b.data_ptr(), out,
                                       stream.cuda_stream))
        torch.cuda.synchronize()
        return {n: t.cpu().numpy().copy() for n, t in o.items()}

    monkeypatch.delenv("IA_GRAPH", raising=False)
    ref = run()
    monkeypatch.setenv("IA_GRAPH", "1")
    for rep in range(3):
        for t in o.values():
            t.zero_()
        got = run()
        for n in ("idx", "k", "lam", "W", "smax", "v", "p", "st"):
            np.testing.assert_array_equal(got[n], ref[n], err_msg=f"replay {rep}: {n}")
        # the program's own eager work between replays
        e = ia.esm_sweep(P.k[:2], P.r_s, P.src, P.r_h, P.b[:2], P.area, normals=P.normals)
        assert np.all(e.status == 0)
        torch.cuda.synchronize()
    b.mul_(2.0)
    got = run()
    np.testing.assert_array_equal(got["idx"], ref["idx"])
    np.testing.assert_allclose(got["v"], 2 * ref["v"], rtol=0, atol=1e-9 * np.abs(ref["v"]).max())
    np.testing.assert_allclose(got["W"], 4 * ref["W"], rtol=1e-9)
    b.mul_(0.5)
    got = run()
    np.testing.assert_array_equal(got["v"], ref["v"])
    monkeypatch.delenv("IA_GRAPH")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n_freq,fmax", [(8, 3000.0), (24, 12000.0)])
def test_hels_graph_replays_match_eager(ia, monkeypatch, precision, n_freq, fmax):
    """ia_hels_sweep with IA_GRAPH=1: the order groups' work on the slot
    streams is captured once (fork from and join to one stream) and replayed;
    three replays equal the eager sweep bit for bit with other eager sweeps
    between them, a replay with doubled pressures gives the doubled velocity,
    and 24 frequencies over orders 5..31 put more groups than slots (shared
    slot streams) into the graph."""
    S = sc.hels_scenario(n_freq=n_freq, fmin=50.0, fmax=fmax, l=96, m=64)
    ia.set_air(S.c, S.rho)

    def run(b):
        r = ia.hels_sweep(S.k, S.order, S.azelr_s, S.azelr_h, b, S.area, precision=precision)
        return r

    monkeypatch.delenv("IA_GRAPH", raising=False)
    ref = run(S.b)
    monkeypatch.setenv("IA_GRAPH", "1")
    for rep in range(3):
        got = run(S.b)
        for n in ("lam_index", "k_tsvd", "lam", "power", "velocity", "pressure"):
            np.testing.assert_array_equal(getattr(got, n), getattr(ref, n),
                                          err_msg=f"replay {rep}: {n}")
        monkeypatch.delenv("IA_GRAPH")
        e = ia.hels_sweep(S.k[:3], S.order[:3], S.azelr_s, S.azelr_h, S.b[:3], S.area)
        assert np.all(np.isfinite(e.power))
        monkeypatch.setenv("IA_GRAPH", "1")
    got = run(2 * S.b)
    np.testing.assert_array_equal(got.lam_index, ref.lam_index)
    np.testing.assert_allclose(got.velocity, 2 * ref.velocity, rtol=0,
                               atol=1e-9 * np.abs(ref.velocity).max())
    got = run(S.b)
    np.testing.assert_array_equal(got.velocity, ref.velocity)
    monkeypatch.delenv("IA_GRAPH")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_radiation_graph_replays_match_eager(ia, monkeypatch, precision):
    """ia_radiation_sweep with IA_GRAPH=1: the first call runs eagerly,
    recording the eigen-solver's decisions (subspace-iteration rounds, the
    4000 Hz frequency sent to the full SVD), and captures a graph that
    replays them; later calls with the same wavenumbers and patches replay
    it.  Replays equal the eager sweep bit for bit, also after an eager
    sweep of other frequencies reused the workspace in between, and doubled
    pressures give doubled amplitudes and four times the power."""
    S = sc.radiation_scenario(freqs=[300.0, 1200.0, 2500.0, 4000.0], patches=(32, 32),
                              mic_n=(16, 16), mic_pitch=0.035)
    ia.set_air(S.c, S.rho)
    n = 48

    def run(b):
        return ia.radiation_sweep(S.k, S.r_p, S.patch_area, S.r_h, b, n, precision=precision,
                                  want_modes=True)

    monkeypatch.delenv("IA_GRAPH", raising=False)
    ref = run(S.b)
    assert np.all(ref["status"] == 0)
    for rep in range(3):
        monkeypatch.setenv("IA_GRAPH", "1")
        got = run(S.b)
        for q in ("power", "lambda", "lambda_index", "eigval", "modes", "amplitude", "velocity"):
            np.testing.assert_array_equal(got[q], ref[q], err_msg=f"call {rep}: {q}")
        monkeypatch.delenv("IA_GRAPH")
        other = ia.radiation_sweep(S.k[3:], S.r_p, S.patch_area, S.r_h, S.b[3:], n)
        assert np.all(other["status"] == 0)
    monkeypatch.setenv("IA_GRAPH", "1")
    got = run(2.0 * S.b)
    np.testing.assert_array_equal(got["lambda_index"], ref["lambda_index"])
    np.testing.assert_allclose(got["power"], 4.0 * ref["power"], rtol=1e-9)
    np.testing.assert_allclose(got["amplitude"], 2.0 * ref["amplitude"], rtol=0,
                               atol=1e-9 * np.abs(ref["amplitude"]).max())
    monkeypatch.delenv("IA_GRAPH")


@pytest.mark.parametrize("precision,n_chief", [("fp64", 0), ("mixed", 0), ("fp64", 6)])
def test_ibem_graph_replays_match_eager(ia, monkeypatch, precision, n_chief):
    """ia_ibem_sweep with IA_GRAPH=1 (one chunk): the inputs are uploaded on
    every call and the device side (assembly, the LU with a kept workspace or
    with CHIEF points the least-squares QR, regularised solve, surface
    pressure and power) is captured once and replayed; three calls equal the
    eager sweep bit for bit with an eager sweep of other frequencies between
    them, and doubled pressures give doubled velocities."""
    S = sc.ibem_scenario(n_freq=6, fmin=300.0, fmax=1800.0, subdiv=2, l=64, n_chief=n_chief)
    ia.set_air(S.c, S.rho)
    ia.set_quadrature()
    cp = S.chief if n_chief else None

    def run(b):
        return ia.ibem_sweep(S.k, S.X, S.E, S.r_h, b, S.area, precision=precision, cp=cp)

    monkeypatch.delenv("IA_GRAPH", raising=False)
    ref = run(S.b)
    for rep in range(3):
        monkeypatch.setenv("IA_GRAPH", "1")
        got = run(S.b)
        for q in ("lam_index", "k_tsvd", "lam", "power", "velocity", "pressure"):
            np.testing.assert_array_equal(getattr(got, q), getattr(ref, q),
                                          err_msg=f"call {rep}: {q}")
        monkeypatch.delenv("IA_GRAPH")
        other = ia.ibem_sweep(S.k[:2] * 1.1, S.X, S.E, S.r_h, S.b[:2], S.area)
        assert np.all(np.isfinite(other.power))
    monkeypatch.setenv("IA_GRAPH", "1")
    got = run(2.0 * S.b)
    np.testing.assert_array_equal(got.lam_index, ref.lam_index)
    np.testing.assert_allclose(got.velocity, 2.0 * ref.velocity, rtol=0,
                               atol=1e-9 * np.abs(ref.velocity).max())
    monkeypatch.delenv("IA_GRAPH")
