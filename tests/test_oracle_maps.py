"""The oracle of the 2-D maps kernel (model.js:1028-1100) against closed forms (CPU only)."""
import numpy as np
import pytest
from scipy.integrate import quad

from _cases import host_model, uniforms
from _oracle import Oracle

OMEGA = 1e-3


def setup(nx=24, ny=20, nz=40, dx=2000.0, scheme="lin"):
    m = host_model(nx, ny, nz, dx, scheme=scheme)
    U = uniforms(m, 6.0)
    o = Oracle(U, m.B, "f64")
    B = np.asarray(m.B, np.float64)
    V = np.zeros(o.vshape())
    th0 = B[0:8 * nz:8][:nz]
    V[:nz, :, :, 3] = th0[:, None, None]
    V[nz, :, :, 3] = th0[-1]
    Q1 = np.zeros(o.cshape())
    Q2 = np.zeros(o.cshape())
    phi = np.zeros((nz, m.nyp, m.nxp))
    return m, U, o, B, V, Q1, Q2, phi


def temperature(B, k, V, i=0, j=0):
    return V[k, j, i, 3] * B[8 * k + 4]


def test_calm_clear_column():
    m, U, o, B, V, Q1, Q2, phi = setup()
    M = o.maps(V, Q1, Q2, phi)
    assert np.all(M[0, ..., 0] == -30) and np.all(M[0, ..., 1] == -30)
    assert np.all(M[0, ..., 2] == 0) and np.all(M[0, ..., 3] == 0)
    assert np.all(M[1] == 0) and np.all(M[2] == 0)
    # a cloud-free column is seen at its lowest level's temperature; no pressure, no wind
    assert np.allclose(M[3, ..., 1], temperature(B, 0, V), rtol=1e-12)
    assert np.all(M[3, ..., 0] == 0) and np.all(M[3, ..., 2] == 0) and np.all(M[3, ..., 3] == 0)


def test_solid_body_rotation_gives_twice_omega():
    m, U, o, B, V, Q1, Q2, phi = setup()
    nz, ny, nx = m.nz, m.ny, m.nx
    dx = m.dx
    yc, xc = 0.5 * ny * dx, 0.5 * nx * dx
    jj = (np.arange(ny + 1) + 0.5) * dx  # u-faces sit at cell-centre y
    ii = (np.arange(nx + 1) + 0.5) * dx  # v-faces sit at cell-centre x
    V[:, :, :, 0] = (-OMEGA * (jj - yc))[None, :, None]
    V[:, :, :, 1] = (OMEGA * (ii - xc))[None, None, :]
    W = 3.0
    V[1:nz, :, :, 2] = W
    M = o.maps(V, Q1, Q2, phi)
    inner = (slice(1, ny - 2), slice(1, nx - 2))
    assert np.allclose(M[0][inner][..., 3], 2 * OMEGA, rtol=1e-10)
    assert np.allclose(M[3][inner][..., 3], 2 * OMEGA, rtol=1e-10)
    # updraft helicity: w * zeta * dz over the cell centres between 2 and 5 km
    levels = [k for k in range(nz) if 2000 < (k + 0.5) * m.dz < 5000]
    assert np.allclose(M[1][inner][..., 1], len(levels) * W * 2 * OMEGA * m.dz, rtol=1e-10)
    assert np.allclose(M[1][inner][..., 0], W)


def test_uniform_divergence_and_wind():
    m, U, o, B, V, Q1, Q2, phi = setup()
    a, u0, v0 = 2e-4, 7.0, -4.0
    x = np.arange(m.nx + 1) * m.dx
    V[:, :, :, 0] = u0 + a * x[None, None, :]
    V[:, :, :, 1] = v0
    M = o.maps(V, Q1, Q2, phi)
    assert np.allclose(M[1][:, :-1][..., 2][:, :-1], -a * 1000, rtol=1e-9)
    # the lowest level's cell velocity and speed; ~1 km the same field
    uc = u0 + a * (np.arange(m.nx) + 0.5) * m.dx
    assert np.allclose(M[2][..., 0], uc[None, :]) and np.allclose(M[2][..., 1], v0)
    assert np.allclose(M[2][..., 2], uc[None, :]) and np.allclose(M[2][..., 3], v0)
    assert np.allclose(M[1][..., 3], np.hypot(uc[None, :], v0))


def test_brightness_temperature_of_cloud_layers():
    m, U, o, B, V, Q1, Q2, phi = setup()
    kc = 24
    # opaque: the layer's own temperature
    Q1[kc, :, :, 1] = 0.05
    M = o.maps(V, Q1, Q2, phi)
    assert np.allclose(M[3][..., 1], temperature(B, kc, V), rtol=1e-12)
    # thin: emissivity a = 1 - exp(-80 qc rho dz) over the lowest level's temperature
    qc = 2e-6
    Q1[kc, :, :, 1] = qc
    M = o.maps(V, Q1, Q2, phi)
    a = 1 - np.exp(-80 * qc * B[8 * kc + 2] * m.dz)
    assert np.allclose(M[3][..., 1], a * temperature(B, kc, V) + (1 - a) * temperature(B, 0, V), rtol=1e-12)


def test_lin_reflectivity_laws():
    """Power laws Z = 3.63e9 (rho qr)^1.75 (rain) and 9.8e8 (rho qs)^1.75 x 0.224 x 4 (snow); the
    reference converts with ln Z / 2.302585, within 1e-7 of log10."""
    m, U, o, B, V, Q1, Q2, phi = setup(scheme="lin")
    k = 6
    rho = B[8 * k + 2]
    Q1[k, :, :, 3] = 2e-3  # rain
    M = o.maps(V, Q1, Q2, phi)
    z = 3.63e9 * (rho * 2e-3) ** 1.75
    assert np.allclose(M[0][..., 0], 10 * np.log10(z), rtol=1e-6)
    Q1[k] = 0
    Q2[k, :, :, 0] = 1e-3  # snow
    M = o.maps(V, Q1, Q2, phi)
    assert np.allclose(M[0][..., 0], 10 * np.log10(9.80e8 * (rho * 1e-3) ** 1.75 * 0.224 * 4.0), rtol=1e-6)


@pytest.mark.parametrize("qr,nr", [(1e-3, 5e3), (3e-4, 2e5)])
def test_thompson_rain_reflectivity_is_the_sixth_moment(qr, nr):
    """Z of the exponential size distribution n(D) = N0 exp(-lambda D) holding the rain mass: the
    integral of D^6 n(D) (mm^6 m^-3), evaluated numerically, with lambda from the mass and number and
    the number within the size bounds, N = rho qr lambda^3 / (pi rho_w); pi rho_w as the model's
    3141.59 (and pi as 3.14159, the cube root as ** 0.33333 in the slope)."""
    m, U, o, B, V, Q1, Q2, phi = setup(scheme="thompson")
    k = 6
    rho = B[8 * k + 2]
    Q1[k, :, :, 3] = qr
    Q2[k, :, :, 2] = nr
    M = o.maps(V, Q1, Q2, phi)
    lam = min(max((3.14159 * 1000.0 * nr / qr) ** 0.33333, 1e3), 1e5)
    n_tot = rho * qr * lam ** 3 / 3141.59
    n0 = n_tot * lam
    zz, _ = quad(lambda D: n0 * D ** 6 * np.exp(-lam * D), 0, 60 / lam, epsabs=0, epsrel=1e-12)
    assert np.isclose(M[0][0, 0, 0], 10 * np.log10(zz * 1e18), rtol=1e-6)
