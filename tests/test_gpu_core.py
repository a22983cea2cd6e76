"""GPU parity of the dense kernels, the SVD and the ESM building blocks
against the oracle / LAPACK (through the C ABI)."""
import numpy as np
import pytest

from inverseacoustics_amd import scenarios as sc
from oracle import esm as oesm
from oracle import regtools as rt

pytestmark = pytest.mark.gpu
RHO, C = 1.21, 343.0


def _cr(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# shapes that reach every tile configuration of the MFMA kernel (32-row,
# 32-column, 32 x 32 and -- from 1024 tiles of 64 x 64 on -- 64 x 64 tiles),
# partial tiles and K not a multiple of the LDS stage depth
GEMM_SHAPES = [(1, 1, 1), (17, 5, 3), (64, 64, 16), (130, 77, 45), (5, 200, 1), (300, 257, 96),
               (32, 300, 70), (300, 20, 33), (40, 100, 9), (129, 65, 200), (2048, 2047, 40)]


@pytest.mark.parametrize("opa,opb", [("N", "N"), ("C", "N"), ("N", "C"), ("C", "C")])
@pytest.mark.parametrize("mnk", GEMM_SHAPES)
def test_zgemm(ia, opa, opb, mnk):
    m, n, k = mnk
    A = _cr((m, k) if opa == "N" else (k, m), 1)
    B = _cr((k, n) if opb == "N" else (n, k), 2)
    C0 = _cr((m, n), 3)
    opA = A if opa == "N" else A.conj().T
    opB = B if opb == "N" else B.conj().T
    ref = (0.5 - 2j) * opA @ opB + (1.5 + 0.25j) * C0
    got = ia.zgemm(A, B, opa, opb, alpha=0.5 - 2j, beta=1.5 + 0.25j, Cm=C0)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * np.sqrt(k)
    # beta = 0 reads nothing of C
    got0 = ia.zgemm(A, B, opa, opb, alpha=1.0, beta=0.0, Cm=np.full((m, n), np.nan + 0j))
    assert np.abs(got0 - opA @ opB).max() <= 1e-13 * max(1.0, np.abs(opA @ opB).max()) * np.sqrt(k)


@pytest.mark.parametrize("opa,opb", [("N", "N"), ("C", "N"), ("N", "C"), ("C", "C")])
@pytest.mark.parametrize("mnk", [(17, 5, 3), (130, 77, 45), (300, 257, 96), (32, 300, 70),
                                 (300, 20, 33), (2048, 2047, 40), (200, 150, 700), (65, 97, 1030),
                                 (130, 200, 129)])
def test_cgemm_fp32(ia, opa, opb, mnk):
    """complex64 products; from K = 128 on they run on the bf16 matrix cores
    with three-way split operands, to the same FP32-level bound."""
    m, n, k = mnk
    A = _cr((m, k) if opa == "N" else (k, m), 1).astype(np.complex64)
    B = _cr((k, n) if opb == "N" else (n, k), 2).astype(np.complex64)
    C0 = _cr((m, n), 3).astype(np.complex64)
    opA = A.astype(complex) if opa == "N" else A.astype(complex).conj().T
    opB = B.astype(complex) if opb == "N" else B.astype(complex).conj().T
    ref = (0.5 - 2j) * opA @ opB + (1.5 + 0.25j) * C0
    got = ia.cgemm(A, B, opa, opb, alpha=0.5 - 2j, beta=1.5 + 0.25j, Cm=C0)
    assert np.abs(got - ref).max() <= 2e-6 * max(1.0, np.abs(ref).max()) * np.sqrt(k)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (5, 3), (31, 7), (33, 40), (100, 64), (257, 10)])
@pytest.mark.parametrize("op", ["N", "C"])
def test_zgesv(ia, n, nrhs, op):
    D = _cr((n, n), 4) + 0.1 * n * np.eye(n)
    B = _cr((n, nrhs), 5)
    X = ia.zgesv(D, B, op)
    Dop = D if op == "N" else D.conj().T
    np.testing.assert_allclose(Dop @ X, B, atol=1e-11 * np.abs(B).max())


@pytest.mark.parametrize("n", [2, 63, 64, 65, 130, 1000, 3000])
@pytest.mark.parametrize("op", ["N", "C"])
def test_zgesv_single_rhs_general(ia, n, op):
    # one right-hand side (the in-LDS solves of op N and op C), no diagonal
    # dominance so that rows are interchanged throughout; backward-error
    # bound of LU with partial pivoting, and agreement with the multi-RHS path
    rng = np.random.default_rng(n)
    D = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = ia.zgesv(D, B, op)
    assert X.shape == (n,)
    Dop = D if op == "N" else D.conj().T
    res = np.abs(Dop @ X - B).max()
    assert res <= 1e-13 * n * np.abs(D).max() * np.abs(X).max()
    X2 = ia.zgesv(D, np.stack([B, B], axis=1), op)
    np.testing.assert_allclose(X, X2[:, 0], rtol=0, atol=1e-9 * np.abs(X).max())


def test_zgesv_pivoting_needed(ia):
    # zero leading entry forces a row interchange in the first panel
    D = np.array([[0, 2, 1], [1, 1, 0], [3, 0, 1]], dtype=complex)
    B = np.array([1, 2, 3], dtype=complex)
    np.testing.assert_allclose(ia.zgesv(D, B), np.linalg.solve(D, B), rtol=1e-13)
    with pytest.raises(ia.IAError):
        ia.zgesv(np.zeros((3, 3)), B)


@pytest.mark.parametrize("shape", [(1, 1), (4, 6), (6, 4), (37, 23), (23, 37), (64, 100),
                                   (100, 64), (128, 642)])
def test_csvd_against_lapack(ia, shape):
    A = _cr(shape, 6)
    U, s, V = ia.csvd(A)
    U0, s0, V0 = rt.csvd(A)
    assert U.shape == U0.shape and V.shape == V0.shape
    np.testing.assert_allclose(s, s0, rtol=1e-12, atol=1e-13 * s0[0])
    np.testing.assert_allclose((U * s) @ V.conj().T, A, atol=1e-12 * s0[0] * np.sqrt(max(shape)))
    p = min(shape)
    np.testing.assert_allclose(U.conj().T @ U, np.eye(p), atol=1e-12)
    np.testing.assert_allclose(V.conj().T @ V, np.eye(p), atol=1e-12)


@pytest.mark.parametrize("shape", [(80, 60), (300, 80)])  # (300, 80): QR-first path
def test_csvd_graded_spectrum(ia, shape):
    # exponentially decaying singular values down to ~1e-15 (the ESM regime)
    mr, nc = shape
    U0, _ = np.linalg.qr(_cr((mr, nc), 8))
    V0, _ = np.linalg.qr(_cr((nc, nc), 9))
    sv = 10.0 ** (-15 * np.arange(nc) / (nc - 1))
    A = (U0 * sv) @ V0.conj().T
    U, s, V = ia.csvd(A)
    s_np = np.linalg.svd(A, compute_uv=False)
    assert np.abs(s - s_np).max() <= 4e-15 * s_np[0] * nc
    np.testing.assert_allclose(s[: nc // 2], sv[: nc // 2], rtol=1e-6)
    np.testing.assert_allclose((U * s) @ V.conj().T, A, atol=1e-13 * np.sqrt(mr))


def test_esm_legacy_matches_oracle(ia):
    S = sc.sphere_scenario()
    ia.set_air(S.c, S.rho)
    m, l = S.m, S.l
    M_e, H_e = ia.esm(m, l, S.r_s, S.r_h, S.k[0], 0.7)
    M_e0, H_e0 = oesm.esm(S.k[0], S.r_s, S.r_h, 0.7, S.rho, S.c)
    assert np.linalg.norm(H_e - H_e0) <= 1e-11 * np.linalg.norm(H_e0)
    assert np.linalg.norm(M_e - M_e0) <= 1e-11 * np.linalg.norm(M_e0)
    Mg, Hg = ia.esmgpu(m, l, S.r_s, S.r_h, S.k[0], 0.7)
    np.testing.assert_array_equal(Hg, H_e)


def test_esm_matrices_general_geometry(ia):
    P = sc.plate_scenario(freqs=[700.0], mic_n=(8, 8), mic_pitch=0.07, src_n=(12, 10), ref_n=(24, 20))
    ia.set_air(P.c, P.rho)
    got = ia.esm_matrices(P.r_s, P.r_h, P.src, P.k[0], P.normals)
    D0, Msq0, Mhq0 = oesm.esm_assemble(P.k[0], P.r_s, P.r_h, P.src, P.normals, P.rho, P.c)
    np.testing.assert_allclose(got["D"], D0, rtol=1e-13, atol=1e-14 * np.abs(D0).max())
    np.testing.assert_allclose(got["M_sq"], Msq0, rtol=1e-13)
    np.testing.assert_allclose(got["M_hq"], Mhq0, rtol=1e-13)
    He0 = np.linalg.solve(D0.T, Mhq0.T).T
    assert np.linalg.norm(got["H_e"] - He0) <= 1e-10 * np.linalg.norm(He0)


def test_filters_and_tsvd_match_oracle(ia):
    A = _cr((40, 70), 10)
    b = _cr(40, 11)
    U, s, V = rt.csvd(A)
    for k in (1, 17, 40):
        np.testing.assert_allclose(ia.tsvd(U, s, V, b, k), rt.tsvd(U, s, V, b, k), rtol=1e-12,
                                   atol=1e-13)
    lam = 0.3 * s[0]
    np.testing.assert_allclose(ia.tikhonov(U, s, V, b, lam), rt.tikhonov(U, s, V, b, lam),
                               rtol=1e-12)
    np.testing.assert_allclose(ia.dsvd(U, s, V, b, lam), rt.dsvd(U, s, V, b, lam), rtol=1e-12)


def test_tsvd_reads_only_k_columns_of_v(ia):
    # tsvd.cuh:145-146 copies n*k entries of V: V truncated to k columns is valid
    A = _cr((40, 70), 16)
    b = _cr(40, 17)
    U, s, V = rt.csvd(A)
    for k in (1, 9, 40):
        Vk = np.asfortranarray(V[:, :k])
        np.testing.assert_allclose(ia.tsvd(U, s, Vk, b, k), rt.tsvd(U, s, V, b, k), rtol=1e-12,
                                   atol=1e-13)


def test_tsvd_k_out_of_range_raises(ia):
    A = _cr((10, 8), 12)
    U, s, V = rt.csvd(A)
    with pytest.raises(ia.IAError):
        ia.tsvd(U, s, V, _cr(10, 13), 9)
    with pytest.raises(ia.IAError):
        ia.tsvd(U, s, V, _cr(10, 13), 0)


@pytest.mark.parametrize("shape", [(64, 100), (100, 64)])
def test_regularize_selection_matches_oracle(ia, shape):
    S = sc.sphere_scenario()
    rng = np.random.default_rng(14)
    A = _cr(shape, 15)
    U0, s0, V0 = np.linalg.svd(A, full_matrices=False)
    s0 = 10.0 ** (-8 * np.arange(len(s0)) / len(s0)) * 50
    A = (U0 * s0) @ V0
    x_true = _cr(shape[1], 16)
    b = A @ x_true + 1e-3 * _cr(shape[0], 17)
    U, s, V = rt.csvd(A)
    beta = U.conj().T @ b
    rows = shape[0]
    d0 = max(np.linalg.norm(b) ** 2 - np.linalg.norm(beta) ** 2, 0) if rows > len(s) else 0.0
    j0, lam0, _ = rt.gcv_tikhonov_grid(s, beta, rows, d0)
    x, lam, idx = ia.regularize(U, s, V, b, "tikhonov", "gcv")
    assert idx == j0 and lam == pytest.approx(lam0[j0], rel=1e-14)
    np.testing.assert_allclose(x, rt.tikhonov(U, s, V, b, lam0[j0]), rtol=1e-11)
    k0, _ = rt.gcv_tsvd(beta, rows, len(s), d0)
    x, _, k = ia.regularize(U, s, V, b, "tsvd", "gcv")
    assert k == k0
    np.testing.assert_allclose(x, rt.tsvd(U, s, V, b, k0), rtol=1e-11)
    jl, laml, _ = rt.l_curve_grid(s, beta, rows, "t", b_norm2=np.linalg.norm(b) ** 2)
    x, lam, idx = ia.regularize(U, s, V, b, "tikhonov", "lcurve")
    assert idx == jl
    reg_min, _, _, gi = rt.gcv_regtools(U, s, b, "t")
    x, lam, idx = ia.regularize(U, s, V, b, "tikhonov", "gcv_regtools")
    assert idx == gi and lam == pytest.approx(reg_min, rel=1e-6)
    reg_min, _, _, gi = rt.gcv_regtools(U, s, b, "d")
    x, lam, idx = ia.regularize(U, s, V, b, "dsvd", "gcv_regtools")
    assert idx == gi and lam == pytest.approx(reg_min, rel=1e-6)
    np.testing.assert_allclose(x, rt.dsvd(U, s, V, b, lam), rtol=1e-10)
    rq, _, _ = rt.quasiopt(U, s, b, "t")
    x, lam, idx = ia.regularize(U, s, V, b, "tikhonov", "quasiopt")
    assert lam == pytest.approx(rq, rel=1e-6)
    rc, _, _, gi = rt.l_curve(U, s, b, "t")
    x, lam, idx = ia.regularize(U, s, V, b, "tikhonov", "lcurve_regtools")
    assert idx == gi and lam == pytest.approx(rc, rel=1e-6)


def test_utils_products(ia):
    M = _cr((30, 20), 18)
    x = _cr(20, 19)
    np.testing.assert_allclose(ia.to_pressure(M, x), M @ x, rtol=1e-13)
    np.testing.assert_allclose(ia.to_velocity(M, x), M @ x, rtol=1e-13)
    p, v = _cr(30, 20), _cr(30, 21)
    assert ia.sound_power(p, v, 3.0) == pytest.approx(0.5 * np.real(np.vdot(v, p)) * 3.0 / 30, rel=1e-13)


@pytest.mark.parametrize("shape", [(96, 40), (300, 257), (64, 64)])
def test_bidiag_split_and_fused_forms_agree(ia, shape, monkeypatch):
    """Small batches run the bidiagonalisation as four launches per column
    (products spread over many workgroups), large ones as one fused
    workgroup per matrix; both must give the oracle's solution."""
    rng = np.random.default_rng(11)
    B = 3
    A = rng.standard_normal((B,) + shape) + 1j * rng.standard_normal((B,) + shape)
    b = rng.standard_normal((B, shape[0])) + 1j * rng.standard_normal((B, shape[0]))
    monkeypatch.setenv("IA_ONE_STAGE", "1")
    monkeypatch.setenv("IA_GEBRD_SPLIT_BELOW", "1000")
    xs, *_ = ia.regsolve(A, b, filter="tikhonov", select="fixed", lam=0.05)
    monkeypatch.setenv("IA_GEBRD_SPLIT_BELOW", "0")
    xf, *_ = ia.regsolve(A, b, filter="tikhonov", select="fixed", lam=0.05)
    for i in range(B):
        U, s, V = rt.csvd(A[i])
        x0 = rt.tikhonov(U, s, V, b[i], 0.05)
        assert np.linalg.norm(xs[i] - x0) <= 1e-12 * np.linalg.norm(x0)
        assert np.linalg.norm(xf[i] - x0) <= 1e-12 * np.linalg.norm(x0)


@pytest.mark.parametrize("shape", [(128, 128), (200, 129), (129, 200), (300, 257), (257, 700),
                                   (600, 600)])
def test_two_stage_csvd_against_lapack(ia, shape, monkeypatch):
    """Columns >= 128: band reduction (QR / LQ panels) + bulge chasing."""
    monkeypatch.delenv("IA_ONE_STAGE", raising=False)
    A = _cr(shape, 12)
    U, s, V = ia.csvd(A)
    s0 = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(s, s0, rtol=1e-12, atol=1e-13 * s0[0])
    np.testing.assert_allclose((U * s) @ V.conj().T, A, atol=1e-12 * s0[0] * np.sqrt(max(shape)))
    p = min(shape)
    np.testing.assert_allclose(U.conj().T @ U, np.eye(p), atol=1e-12)
    np.testing.assert_allclose(V.conj().T @ V, np.eye(p), atol=1e-12)


@pytest.mark.parametrize("shape", [(200, 200), (500, 160)])
def test_two_stage_graded_spectrum(ia, shape, monkeypatch):
    monkeypatch.delenv("IA_ONE_STAGE", raising=False)
    mr, nc = shape
    U0, _ = np.linalg.qr(_cr((mr, nc), 13))
    V0, _ = np.linalg.qr(_cr((nc, nc), 14))
    sv = 10.0 ** (-15 * np.arange(nc) / (nc - 1))
    A = (U0 * sv) @ V0.conj().T
    U, s, V = ia.csvd(A)
    s_np = np.linalg.svd(A, compute_uv=False)
    assert np.abs(s - s_np).max() <= 4e-15 * s_np[0] * nc
    np.testing.assert_allclose(s[: nc // 2], sv[: nc // 2], rtol=1e-6)
    np.testing.assert_allclose((U * s) @ V.conj().T, A, atol=1e-13 * np.sqrt(mr))


@pytest.mark.parametrize("shape", [(300, 257), (160, 400), (700, 200)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_and_two_stage_solutions_agree(ia, shape, precision, monkeypatch):
    rng = np.random.default_rng(15)
    B = 3
    A = rng.standard_normal((B,) + shape) + 1j * rng.standard_normal((B,) + shape)
    b = rng.standard_normal((B, shape[0])) + 1j * rng.standard_normal((B, shape[0]))
    monkeypatch.setenv("IA_ONE_STAGE", "1")
    x1, *_ = ia.regsolve(A, b, filter="tikhonov", select="fixed", lam=0.05, precision=precision)
    monkeypatch.delenv("IA_ONE_STAGE")
    x2, *_ = ia.regsolve(A, b, filter="tikhonov", select="fixed", lam=0.05, precision=precision)
    tol = 1e-12 if precision == "fp64" else 2e-5
    for i in range(B):
        U, s, V = rt.csvd(A[i])
        x0 = rt.tikhonov(U, s, V, b[i], 0.05)
        assert np.linalg.norm(x1[i] - x0) <= tol * np.linalg.norm(x0)
        assert np.linalg.norm(x2[i] - x0) <= tol * np.linalg.norm(x0)


@pytest.mark.parametrize("cand", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("opa,opb", [("N", "N"), ("C", "N"), ("N", "C"), ("C", "C")])
def test_gemm_every_candidate(ia, monkeypatch, cand, opa, opb):
    """Every kernel the GEMM autotuner may pick (IA_GEMM_FORCE: 0 = bf16
    split, complex64 only, 1..4 = 3M with 32x32 / 32x64 / 64x32 / 64x64
    tiles) on shapes with ragged edges, shallow and deep, to the FP32 /
    FP64 bounds of the built-in choice's tests."""
    monkeypatch.setenv("IA_GEMM_FORCE", str(cand))
    for m, n, k in [(17, 5, 3), (130, 77, 45), (65, 97, 300), (33, 200, 129)]:
        A = _cr((m, k) if opa == "N" else (k, m), 1)
        B = _cr((k, n) if opb == "N" else (n, k), 2)
        C0 = _cr((m, n), 3)
        opA = A if opa == "N" else A.conj().T
        opB = B if opb == "N" else B.conj().T
        ref = (0.5 - 2j) * opA @ opB + (1.5 + 0.25j) * C0
        got = ia.zgemm(A, B, opa, opb, alpha=0.5 - 2j, beta=1.5 + 0.25j, Cm=C0)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * np.sqrt(k)
        A4, B4, C4 = A.astype(np.complex64), B.astype(np.complex64), C0.astype(np.complex64)
        opA4 = A4.astype(complex) if opa == "N" else A4.astype(complex).conj().T
        opB4 = B4.astype(complex) if opb == "N" else B4.astype(complex).conj().T
        ref4 = (0.5 - 2j) * opA4 @ opB4 + (1.5 + 0.25j) * C4
        got4 = ia.cgemm(A4, B4, opa, opb, alpha=0.5 - 2j, beta=1.5 + 0.25j, Cm=C4)
        assert np.abs(got4 - ref4).max() <= 2e-6 * max(1.0, np.abs(ref4).max()) * np.sqrt(k)


@pytest.mark.parametrize("cand", [0, 4])
@pytest.mark.parametrize("opa,opb", [("N", "N"), ("C", "N"), ("N", "C"), ("C", "C")])
def test_gemm_complex64_scaled_rows(ia, monkeypatch, cand, opa, opb):
    """The complex64 products (0 = bf16 three-part split, 4 = FP32 3M)
    against the FP64 product of the same complex64 operands, entry by
    entry: rows of op(A) and columns of op(B)
    spread over 1e-9 .. 1e9, a zero row and column, rows whose magnitude
    jumps by 1e10 between stages, and ragged edges (every product and sum
    stays inside the FP32 normal range).  Bound per entry:
    (3 2^-22 + K 2^-23) (|op(A)| |op(B)|)_ij, the split (or 3M) error plus
    FP32 accumulation."""
    monkeypatch.setenv("IA_GEMM_FORCE", str(cand))
    rng = np.random.default_rng(7)
    for m, n, k in [(70, 90, 100), (130, 65, 257)]:
        opA = _cr((m, k), 11) * 10.0 ** rng.uniform(-9, 9, (m, 1))
        opB = _cr((k, n), 12) * 10.0 ** rng.uniform(-9, 9, (1, n))
        opA[3] = 0
        opB[:, 5] = 0
        opA[7, :32] *= 1e-10   # first stage tiny, later stages O(1)
        opA[9, 32:] *= 1e-10   # first stage O(1), later stages tiny
        opB[40:, 11] *= 1e10
        A4 = (opA if opa == "N" else opA.conj().T).astype(np.complex64)
        B4 = (opB if opb == "N" else opB.conj().T).astype(np.complex64)
        a = A4.astype(complex) if opa == "N" else A4.astype(complex).conj().T
        bm = B4.astype(complex) if opb == "N" else B4.astype(complex).conj().T
        ref = a @ bm
        bound = (3 * 2.0 ** -22 + k * 2.0 ** -23) * (np.abs(a) @ np.abs(bm))
        got = ia.cgemm(A4, B4, opa, opb).astype(complex)
        err = np.abs(got - ref)
        assert np.all(err <= bound), (m, n, k, float(np.max(err / np.maximum(bound, 1e-300))))
        assert np.all(got[3] == 0) and np.all(got[:, 5] == 0)
