"""Gradients of the ESM regularised solve -- test infrastructure (the checker
of ``ia_esm_gradient``; nothing in the product imports it).

The reference has no derivative code.  The forward chain it differentiates
is the reference's own:

* sources q_j = r_s,j - d n_j (retreat d along the unit normals), and D, M_sq,
  M_hq as in esm.cuh:1-36 / esmV.m:20-29 (``oracle.esm.esm_assemble``);
* H_e = M_hq D^-1 (esm.cuh:59-92);
* the standard-form Tikhonov solution at a fixed lambda,
  x = argmin ||H_e x - b||^2 + lambda^2 ||x||^2 (tikhonovG.m:29-30);
* the misfit J = ||H_e x - b||^2 and the sound power
  W = 1/2 Re(x^H M_sq D^-1 x) S/m (to_pressure / sound_power,
  utils.cuh:199-289).

``gradient`` differentiates that chain with torch autograd (complex128, CPU)
with respect to d, the microphone coordinates and log lambda.  The
least-squares solve is the QR factorisation of [H_e; lambda I], whose
rounding error scales with cond = s_1 / lambda rather than its square.
Pinned by central finite differences of the numpy forward
(tests/test_oracle_adjoint.py).
"""
from __future__ import annotations

import math

import numpy as np

from . import esm as oesm


def forward(k, r_s, normals, d, r_h, b, lam, area, rho, c):
    """(J, W, x) of the chain above in numpy (LAPACK QR least squares)."""
    r_s = np.asarray(r_s, np.float64)
    normals = np.asarray(normals, np.float64)
    src = r_s - d * normals
    D, M_sq, M_hq = oesm.esm_assemble(k, r_s, r_h, src, normals, rho, c)
    A = np.linalg.solve(D.T, M_hq.T).T
    m = A.shape[1]
    Q, R = np.linalg.qr(np.concatenate([A, lam * np.eye(m)], axis=0))
    rhs = np.concatenate([b, np.zeros(m, complex)])
    x = np.linalg.solve(R, Q.conj().T @ rhs)
    r = A @ x - b
    J = float(np.vdot(r, r).real)
    p_s = M_sq @ np.linalg.solve(D, x)
    W = 0.5 * (area / m) * float(np.vdot(x, p_s).real)
    return J, W, x


def gradient(k, r_s, normals, d, r_h, b, lam, area, rho, c):
    """Reference gradients by reverse-mode autograd: a dict with J, W and
    dJ_dd, dW_dd, dJ_dmic (l, 3), dW_dmic (l, 3), dJ_dloglam, dW_dloglam."""
    import torch

    f64, c128 = torch.float64, torch.complex128
    rs = torch.as_tensor(np.asarray(r_s, np.float64), dtype=f64)
    nr = torch.as_tensor(np.asarray(normals, np.float64), dtype=f64)
    dd = torch.tensor(float(d), dtype=f64, requires_grad=True)
    mics = torch.tensor(np.asarray(r_h, np.float64), dtype=f64, requires_grad=True)
    loglam = torch.tensor(math.log(lam), dtype=f64, requires_grad=True)
    bb = torch.as_tensor(np.asarray(b, np.complex128), dtype=c128)
    src = rs - dd * nr
    kk = float(k)
    kap = 1j * rho * c * kk / (4 * math.pi)

    def green(P):
        diff = P[:, None, :] - src[None, :, :]
        R = torch.sqrt(torch.sum(diff * diff, dim=2))
        return diff, R, torch.exp(1j * kk * R.to(c128))

    diff, R, e = green(rs)
    cos_t = torch.sum(nr[:, None, :] * diff, dim=2) / R
    D = (1j * kk * R - 1.0) / (4 * math.pi * R * R) * e * cos_t
    M_sq = kap * e / R
    _, Rh, eh = green(mics)
    M_hq = kap * eh / Rh
    A = torch.linalg.solve(D.transpose(0, 1), M_hq.transpose(0, 1)).transpose(0, 1)
    m = A.shape[1]
    lamt = torch.exp(loglam)
    Aug = torch.cat([A, lamt.to(c128) * torch.eye(m, dtype=c128)], dim=0)
    Q, Rq = torch.linalg.qr(Aug)
    rhs = torch.cat([bb, torch.zeros(m, dtype=c128)])
    x = torch.linalg.solve_triangular(Rq, (Q.conj().transpose(0, 1) @ rhs)[:, None],
                                      upper=True)[:, 0]
    r = A @ x - bb
    J = torch.sum(torch.abs(r) ** 2)
    p_s = M_sq @ torch.linalg.solve(D, x)
    W = 0.5 * (area / m) * torch.real(torch.sum(torch.conj(x) * p_s))
    gJ = torch.autograd.grad(J, [dd, mics, loglam], retain_graph=True)
    gW = torch.autograd.grad(W, [dd, mics, loglam])
    return dict(J=float(J.detach()), W=float(W.detach()), dJ_dd=float(gJ[0]), dW_dd=float(gW[0]),
                dJ_dmic=gJ[1].numpy(), dW_dmic=gW[1].numpy(), dJ_dloglam=float(gJ[2]),
                dW_dloglam=float(gW[2]), x=x.detach().numpy())
