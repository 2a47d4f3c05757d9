"""Independent FP64 solve of rhoC lap_h phi + d/dz(rhoW dphi/dz) = rhs with scipy's DST-II
(phi = 0 outside the lateral faces, Neumann at ground and lid): a library check of the
oracle and of the GPU solve at sizes where the dense oracle would be slow."""
import numpy as np
from scipy.fft import dstn, idstn


def solve_scipy(rhs, dx, dz, rhoC, rhoW):
    rhs = np.asarray(rhs, np.float64)
    nz, ny, nx = rhs.shape
    hat = dstn(rhs, type=2, norm="ortho", axes=(1, 2))
    lx = -4.0 / dx ** 2 * np.sin(np.pi * (np.arange(nx) + 1) / (2 * nx)) ** 2
    ly = -4.0 / dx ** 2 * np.sin(np.pi * (np.arange(ny) + 1) / (2 * ny)) ** 2
    shift = ly[:, None] + lx[None, :]
    a = 1.0 / dz ** 2
    lo = np.array([rhoW[k] * a if k > 0 else 0.0 for k in range(nz)])
    up = np.array([rhoW[k + 1] * a if k < nz - 1 else 0.0 for k in range(nz)])
    diag = -(lo + up)
    cp = np.empty_like(hat)
    dp = np.empty_like(hat)
    cprev = np.zeros_like(shift)
    dprev = np.zeros_like(shift)
    for k in range(nz):
        den = diag[k] + shift * rhoC[k] - lo[k] * cprev
        cprev = up[k] / den
        dprev = (hat[k] - lo[k] * dprev) / den
        cp[k] = cprev
        dp[k] = dprev
    x = np.empty_like(hat)
    x[nz - 1] = dp[nz - 1]
    for k in range(nz - 2, -1, -1):
        x[k] = dp[k] - cp[k] * x[k + 1]
    return idstn(x, type=2, norm="ortho", axes=(1, 2))


def apply_operator(phi, dx, dz, rhoC, rhoW):
    """The discrete operator with the solve's boundary conditions (antisymmetric ghosts in x, y;
    zero-flux in z): rhoC (phi_e + phi_w + phi_n + phi_s - 4 phi)/dx^2 + (rhoW_{k+1}(phi_{k+1}-phi_k)
    - rhoW_k (phi_k - phi_{k-1}))/dz^2."""
    phi = np.asarray(phi, np.float64)
    nz = phi.shape[0]
    p = np.pad(phi, ((0, 0), (1, 1), (1, 1)))
    p[:, 0, :] = -p[:, 1, :]
    p[:, -1, :] = -p[:, -2, :]
    p[:, :, 0] = -p[:, :, 1]
    p[:, :, -1] = -p[:, :, -2]
    lap = (p[:, 1:-1, 2:] + p[:, 1:-1, :-2] + p[:, 2:, 1:-1] + p[:, :-2, 1:-1] - 4 * phi) / dx ** 2
    out = np.asarray(rhoC, np.float64)[:, None, None] * lap
    fl = np.zeros((nz + 1,) + phi.shape[1:])
    for k in range(1, nz):
        fl[k] = rhoW[k] * (phi[k] - phi[k - 1]) / dz ** 2
    out += fl[1:] - fl[:-1]
    return out
