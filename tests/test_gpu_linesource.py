"""GPU line-source scenario (ia_linesource, MATLAB/linesource.m) against the
oracle, and the IBEM sweep reconstructing its surface velocity from the
field pressures."""
import numpy as np
import pytest

from inverseacoustics_amd import scenarios as sc
from oracle import linesource as ls

pytestmark = pytest.mark.gpu
RHO, C, U0, A = 1.21, 343.0, 0.8, 0.01


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n_ls,L", [(2, 0.1), (257, 0.2), (1000, 0.24)])
def test_linesource_matches_oracle(ia, n_ls, L):
    X, E = sc.icosphere(2, 0.15)
    H = sc.fibonacci_sphere(96, 0.35)
    ks = np.array([0.5, 9.0, 41.0])
    xp, xsf, vsf = ia.linesource(n_ls, X, E, H, RHO, C, U0, ks, A, L)
    assert xp.shape == (3, 96) and xsf.shape == (3, len(X)) and vsf.shape == (3, len(X))
    for f, k in enumerate(ks):
        p0, s0, v0 = ls.linesource(n_ls, X, E, H, RHO, C, U0, k, A, L)
        assert _rel(xp[f], p0) < 1e-12
        assert _rel(xsf[f], s0) < 1e-12
        assert _rel(vsf[f], v0) < 1e-12


def test_scalar_wavenumber_and_optional_outputs(ia):
    X, E = sc.icosphere(1, 0.1)
    H = sc.fibonacci_sphere(10, 0.3)
    xp, xsf, vsf = ia.linesource(33, X, E, H, RHO, C, U0, 12.0, A, 0.1)
    assert xp.shape == (10,) and vsf.shape == (len(X),)
    xp2, xsf2, vsf2 = ia.linesource(33, X, E, H, RHO, C, U0, np.array([3.0, 12.0]), A, 0.1)
    np.testing.assert_array_equal(xp2[1], xp)
    np.testing.assert_array_equal(vsf2[1], vsf)
    # no field points: surface outputs only
    _, xsf3, vsf3 = ia.linesource(33, X, E, np.zeros((0, 3)), RHO, C, U0, 12.0, A, 0.1)
    np.testing.assert_array_equal(xsf3, xsf)
    np.testing.assert_array_equal(vsf3, vsf)


def test_one_sample_is_zero_and_bad_mesh_rejected(ia):
    X, E = sc.icosphere(1, 0.1)
    H = sc.fibonacci_sphere(8, 0.3)
    xp, xsf, vsf = ia.linesource(1, X, E, H, RHO, C, U0, 5.0, A, 0.1)
    assert not xp.any() and not xsf.any() and not vsf.any()
    bad = E.copy()
    bad[3, 1] = len(X)
    with pytest.raises(ia.IAError):
        ia.linesource(16, X, bad, H, RHO, C, U0, 5.0, A, 0.1)
    with pytest.raises(ia.IAError):
        ia.linesource(0, X, E, H, RHO, C, U0, 5.0, A, 0.1)


def test_ibem_sweep_reconstructs_line_source_velocity(ia):
    # closed sphere around the line: the inverse IBEM recovers the normal
    # velocity from the field pressures (2.8 % with linear elements at
    # subdivision 3, measured with the oracle chain; the FP64 and mixed
    # forms must both get there)
    ia.set_quadrature()
    ia.set_air(C, RHO)
    X, E = sc.icosphere(3, 0.15)
    H = sc.fibonacci_sphere(200, 0.3)
    ks = 2 * np.pi * np.array([300.0, 800.0]) / C
    xp, xsf, vsf = ia.linesource(101, X, E, H, RHO, C, 1.0, ks, A, 0.2)
    for prec in ("fp64", "mixed"):
        r = ia.ibem_sweep(ks, X, E, H, xp, 4 * np.pi * 0.15 ** 2, precision=prec)
        for f in range(2):
            ev = np.linalg.norm(r.velocity[f] - vsf[f]) / np.linalg.norm(vsf[f])
            ep = np.linalg.norm(r.pressure[f] - xsf[f]) / np.linalg.norm(xsf[f])
            assert ev < 0.035 and ep < 0.006, (prec, f, ev, ep)
