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
