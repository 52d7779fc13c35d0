"""FP64 oracle of the explicit-matrix radiosity system.

TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the benchmark's
CPU-baseline leg use it as the checker; the engine never imports it.

The reference (morris821028/hw-radiosity) solves radiosity by progressive
shooting: each step takes the source with the largest unshot energy and adds
rho_i * ff_ij * dB_j to every receiver's B and dB (CalRadiosity,
shade.cpp:244-259, with rho = Frgb * 0.5 / 255, shade.cpp:246).  Written
over an explicit gathering matrix F (F[i][j] = form factor from patch i to
patch j), the limit of that process is the linear system

    B = E + diag(rho) F B        (per colour channel / right-hand side)

which this module solves densely in FP64:

  dense_solve       LU of M = I - diag(rho_c) F per distinct rho column
  residual          E + rho o (F B) - B
  adjoint           lambda = M^-T dL/dB, dL/dE = lambda, dL/drho = lambda o (F B)
  fd_grad_rho/_E    central finite differences of L = ||B - B_target||^2,
                    each perturbed solve by Sherman-Morrison on the one LU
  shoot32           the engine's plain progressive shooting restated in
                    float32 numpy, operation for operation (bit-exact target)
  shoot64           the same schedule in FP64 (pure Python; small n)

Pinned by tests/test_linear_oracle.py against closed forms (uniform rank-1 F,
the 2-patch system), brute-force re-solves and brute-force finite
differences on small systems.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def emission_reflectance(frgb, brgb, light: float = 1.0):
    """E and rho (n x 3) of a .tri scene: rho = Frgb * 0.5 / 255
    (shade.cpp:246); a triangle emits iff a Brgb channel is 255 (rad.h:47-49)
    and then E = light * Brgb / 255."""
    frgb = np.asarray(frgb, np.float64)
    brgb = np.asarray(brgb, np.float64)
    rho = frgb * 0.5 / 255.0
    emits = (brgb == 255).any(axis=1)
    E = np.where(emits[:, None], light * brgb / 255.0, 0.0)
    return E, rho


def _groups(rho):
    """Columns of rho grouped by identical vectors: {key: [cols]}."""
    groups = {}
    for c in range(rho.shape[1]):
        key = rho[:, c].tobytes()
        groups.setdefault(key, []).append(c)
    return groups


def _lu(F, rho_col):
    M = np.eye(F.shape[0]) - rho_col[:, None] * F
    return sla.lu_factor(M, check_finite=False)


def dense_solve(F, rho, E):
    """B with B = E + rho o (F B), column by column, FP64 LU."""
    F = np.asarray(F, np.float64)
    rho = np.asarray(rho, np.float64)
    E = np.asarray(E, np.float64)
    B = np.zeros_like(E)
    for cols in _groups(rho).values():
        lu = _lu(F, rho[:, cols[0]])
        B[:, cols] = sla.lu_solve(lu, E[:, cols], check_finite=False)
    return B


def residual(F, rho, E, B):
    F = np.asarray(F, np.float64)
    return E + rho * (F @ B) - B


def adjoint(F, rho, B, g):
    """(dL/dE, dL/drho) for dL/dB = g: lambda = (I - diag(rho) F)^-T g."""
    F = np.asarray(F, np.float64)
    lam = np.zeros_like(g)
    for cols in _groups(rho).values():
        lu = _lu(F, rho[:, cols[0]])
        lam[:, cols] = sla.lu_solve(lu, g[:, cols], trans=1, check_finite=False)
    return lam, lam * (F @ B)


def loss(B, Bt):
    return float(np.sum((B - Bt) ** 2))


def fd_grad(F, rho, E, Bt, idx, h=1e-6, wrt="rho"):
    """Central differences of L = ||B(rho, E) - Bt||^2 with respect to
    rho[i, c] (wrt="rho") or E[i, c] (wrt="E") for (i, c) in idx.  Each
    perturbed solve is exact in FP64 via Sherman-Morrison on the LU of the
    unperturbed M_c = I - diag(rho_c) F:
      rho_i += d:  B' = B + d z (F_i . B) / (1 - d F_i . z),  z = M^-1 e_i
      E_i   += d:  B' = B + d z."""
    F = np.asarray(F, np.float64)
    B = dense_solve(F, rho, E)
    out = np.zeros(len(idx))
    by_col = {}
    for k, (i, c) in enumerate(idx):
        by_col.setdefault(c, []).append((k, i))
    for c, items in by_col.items():
        lu = _lu(F, rho[:, c])
        rhs = np.zeros((F.shape[0], len(items)))
        for q, (_, i) in enumerate(items):
            rhs[i, q] = 1.0
        Z = sla.lu_solve(lu, rhs, check_finite=False)
        base = loss(B, Bt) - loss(B[:, c], Bt[:, c])
        for q, (k, i) in enumerate(items):
            z = Z[:, q]
            vals = []
            for d in (h, -h):
                if wrt == "rho":
                    fb, fz = F[i] @ B[:, c], F[i] @ z
                    bc = B[:, c] + d * z * fb / (1.0 - d * fz)
                else:
                    bc = B[:, c] + d * z
                vals.append(base + loss(bc, Bt[:, c]))
            out[k] = (vals[0] - vals[1]) / (2.0 * h)
    return out


def _key(p, j):
    return p, -j  # larger power first, then the lower index


def shoot32(F, rho, E, area, shots):
    """The engine's progressive shooting in float32, operation for operation:
    source j = argmax over patches of area_j * (sum_c dB_j[c]) (float32, c in
    order; only positive powers; ties to the lowest index); then
    v = (rho_i * F_ij) * dBsrc_c, B += v, dB += v for every i (the source
    included: its dB was zeroed first).  Returns (B, dB, sources)."""
    F = np.asarray(F, np.float32)
    rho = np.asarray(rho, np.float32)
    B = np.asarray(E, np.float32).copy()
    dB = B.copy()
    area = np.asarray(area, np.float32)
    sources = []

    def pick():
        p = np.zeros(len(area), np.float32)
        for c in range(dB.shape[1]):
            p = (p + dB[:, c]).astype(np.float32)
        p = (p * area).astype(np.float32)
        if not (p > 0).any():
            return -1
        m = p[p > 0].max()
        return int(np.nonzero(p == m)[0][0])

    for _ in range(shots):
        j = pick()
        if j < 0:
            break
        src = dB[j].copy()
        dB[j] = 0
        sources.append(j)
        f = F[:, j]
        for c in range(dB.shape[1]):
            v = ((rho[:, c] * f).astype(np.float32) * src[c]).astype(np.float32)
            B[:, c] = (B[:, c] + v).astype(np.float32)
            dB[:, c] = (dB[:, c] + v).astype(np.float32)
    return B, dB, sources


def shoot64(F, rho, E, area, shots):
    """shoot32's schedule in FP64 (the converged limit is dense_solve)."""
    F = np.asarray(F, np.float64)
    B = np.asarray(E, np.float64).copy()
    dB = B.copy()
    for _ in range(shots):
        p = dB.sum(axis=1) * area
        if not (p > 0).any():
            break
        j = int(np.argmax(np.where(p > 0, p, -np.inf)))
        src = dB[j].copy()
        dB[j] = 0
        v = rho * F[:, j:j + 1] * src[None, :]
        B += v
        dB += v
    return B, dB
