"""Pins of the oracle's restatement of the step kernels (oracle/orc_step.c) by closed forms and
invariants that a plausible transcription mistake would break (CPU only)."""
import math

import numpy as np
import pytest

import _oracle
from _cases import host_model, uniforms
from _oracle import Oracle
from _util import rel_l2


def pcg_py(v):
    """PCG-RXS-M-XS 32-bit hash (Jarzynski & Olano 2020), the function noise.js:5 names."""
    s = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & 0xFFFFFFFF
    return ((w >> 22) ^ w) & 0xFFFFFFFF


def test_hash_known_values():
    L = _oracle.lib()
    for v in [0, 1, 2, 17, 0x7f4a7c15, 0xFFFFFFFF, 123456789]:
        assert L.orc_pcg(v) == pcg_py(v)
    for x, y, z in [(0, 0, 0), (17, 3, 5), (-1, 2, -3), (1000, -7, 65541)]:
        e = pcg_py((x & 0xFFFFFFFF) ^ pcg_py((y & 0xFFFFFFFF) ^ pcg_py(((z & 0xFFFFFFFF) + 0x7F4A7C15) & 0xFFFFFFFF)))
        assert L.orc_hash3i(x, y, z) == e


def small(nx=24, ny=20, nz=12, dx=2000.0, scheme="lin", dt=6.0):
    m = host_model(nx, ny, nz, dx, scheme=scheme)
    return m, uniforms(m, dt)


def base_V(m, rng=None, wind=None):
    """Faces with theta = base theta0 and the given (u, v, w) arrays (default 0)."""
    V = np.zeros((m.nz + 1, m.ny + 1, m.nx + 1, 4))
    th0 = np.concatenate([m.B[0:8 * m.nz:8][: m.nz], [m.B[8 * (m.nz - 1)]]])
    V[..., 3] = th0[:, None, None]
    if wind is not None:
        V[..., 0], V[..., 1], V[..., 2] = wind
    return V


@pytest.mark.parametrize("sfx", ["f32", "f64"])
def test_advection_identity_at_zero_wind(sfx):
    m, U = small()
    o = Oracle(U, m.B, sfx)
    rng = np.random.default_rng(0)
    V = base_V(m)
    V[..., 3] += rng.standard_normal(V.shape[:3])
    V = o.a(V)
    Vt = o.advv1(V)
    # physical texels (the dummy faces beyond the last cell are traced back into the domain)
    ph = np.s_[: m.nz, : m.ny, : m.nx]
    assert np.array_equal(Vt[ph], V[ph])
    assert np.array_equal(o.advv2(V, Vt)[ph], V[ph])
    Q1 = np.zeros((m.nz, m.ny, m.nx, 4))
    Q1[..., :3] = rng.uniform(0, 1e-2, Q1.shape[:3] + (3,))
    Q2 = np.zeros_like(Q1)
    a, b = o.advq1(V, Q1, Q2)
    assert np.array_equal(a[..., :3], o.a(Q1)[..., :3])
    c, d = o.advq2(V, Q1, Q2, a, b)
    assert np.array_equal(c[..., :3], o.a(Q1)[..., :3])


def test_advection_preserves_constants_with_wind():
    m, U = small()
    o = Oracle(U, m.B, "f32")
    rng = np.random.default_rng(1)
    shape = (m.nz + 1, m.ny + 1, m.nx + 1)
    V = base_V(m, wind=(rng.uniform(-40, 40, shape), rng.uniform(-40, 40, shape), rng.uniform(-10, 10, shape)))
    V[..., 3] = 301.25
    V = o.a(V)
    Vt = o.advv1(V)
    assert np.all(Vt[..., 3] == np.float32(301.25))
    assert np.all(o.advv2(V, Vt)[..., 3] == np.float32(301.25))
    Q1 = np.zeros((m.nz, m.ny, m.nx, 4))
    Q1[..., 0] = 0.0125
    Q2 = np.zeros_like(Q1)
    a, b = o.advq1(V, Q1, Q2)
    assert np.all(a[..., 0] == np.float32(0.0125))
    c, _ = o.advq2(V, Q1, Q2, a, b)
    assert np.all(c[..., 0] == np.float32(0.0125))


def test_advection_integer_courant_shift_is_exact():
    """u * dt / dx = 1 exactly: every carried quantity moves one cell in x, MacCormack adds nothing."""
    m = host_model(24, 16, 10, 512.0)
    U = uniforms(m, 4.0)  # dt/dx = 1/128, u = 128 m/s
    o = Oracle(U, m.B, "f32")
    rng = np.random.default_rng(2)
    shape = (m.nz + 1, m.ny + 1, m.nx + 1)
    V = base_V(m, wind=(np.full(shape, 128.0), np.zeros(shape), np.zeros(shape)))
    V[..., 3] += rng.standard_normal(shape)
    V = o.a(V)
    Vt = o.advv1(V)
    Vo = o.advv2(V, Vt)
    sl, src = np.s_[: m.nz, : m.ny, 2:-2], np.s_[: m.nz, : m.ny, 1:-3]
    assert np.array_equal(Vt[..., 3][sl], V[..., 3][src])
    assert np.array_equal(Vo[..., 3][sl], V[..., 3][src])
    Q1 = np.zeros((m.nz, m.ny, m.nx, 4))
    Q1[..., 0] = rng.uniform(0, 1e-2, Q1.shape[:3])
    Q2 = np.zeros_like(Q1)
    a, b = o.advq1(V, Q1, Q2)
    c, _ = o.advq2(V, Q1, Q2, a, b)
    Q1f = o.a(Q1)
    assert np.array_equal(a[:, :, 2:-2, 0], Q1f[:, :, 1:-3, 0])
    assert np.array_equal(c[:, :, 2:-2, 0], Q1f[:, :, 1:-3, 0])


def test_advection_creates_no_new_extrema_and_keeps_moisture_positive():
    m, U = small()
    o = Oracle(U, m.B, "f32")
    rng = np.random.default_rng(3)
    shape = (m.nz + 1, m.ny + 1, m.nx + 1)
    V = base_V(m, wind=(rng.uniform(-60, 60, shape), rng.uniform(-60, 60, shape), rng.uniform(-20, 20, shape)))
    V[..., 3] += 3 * rng.standard_normal(shape)
    V = o.a(V)
    Vo = o.advv2(V, o.advv1(V))
    for ch in range(4):
        assert Vo[..., ch].min() >= V[..., ch].min() and Vo[..., ch].max() <= V[..., ch].max()
    Q1 = rng.uniform(0, 1e-2, (m.nz, m.ny, m.nx, 4)) * (rng.random((m.nz, m.ny, m.nx, 4)) < 0.3)
    Q2 = rng.uniform(0, 1e-2, (m.nz, m.ny, m.nx, 4)) * (rng.random((m.nz, m.ny, m.nx, 4)) < 0.3)
    a, b = o.advq1(V, Q1, Q2)
    c, d = o.advq2(V, Q1, Q2, a, b)
    assert c.min() >= 0 and d.min() >= 0
    assert c.max() <= o.a(Q1).max() and d[..., :3].max() <= o.a(Q2)[..., :3].max()


def test_trilinear_sample_brute_force():
    """A displaced sample equals the brute-force trilinear interpolation at texel coordinate p - 1/2."""
    m = host_model(16, 16, 8, 1000.0)
    U = uniforms(m, 5.0)  # k = 5/1000 horizontally
    o = Oracle(U, m.B, "f64")
    rng = np.random.default_rng(4)
    shape = (m.nz + 1, m.ny + 1, m.nx + 1)
    V = base_V(m, wind=(np.full(shape, 37.0), np.full(shape, -23.0), np.zeros(shape)))
    th = rng.standard_normal(shape)
    V[..., 3] = th
    Vt = o.advv1(V)
    x, y, z = 7, 5, 3
    # theta at the cell centre (x+.5, y+.5, z+.5), departure (x+.5-0.185, y+.5+0.115, z+.5) -> texel coords - 0.5
    tx, ty, tz = x - 0.185, y + 0.115, z
    i, j, k = int(np.floor(tx)), int(np.floor(ty)), int(np.floor(tz))
    ax, ay, az = tx - i, ty - j, tz - k
    ref = 0.0
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                w = (ax if di else 1 - ax) * (ay if dj else 1 - ay) * (az if dk else 1 - az)
                ref += w * th[k + dk, j + dj, i + di]
    assert Vt[z, y, x, 3] == pytest.approx(ref, abs=1e-12)


def test_nut_closed_forms():
    m, U = small(dt=6.0)
    o = Oracle(U, m.B, "f64")
    B = m.B.astype(np.float64).reshape(-1, 8)
    nz = m.nz
    u0 = np.concatenate([B[:nz, 6], [B[nz - 1, 6]]])
    v0 = np.concatenate([B[:nz, 7], [B[nz - 1, 7]]])
    shape = (nz + 1, m.ny + 1, m.nx + 1)
    V = base_V(m, wind=(np.broadcast_to(u0[:, None, None], shape), np.broadcast_to(v0[:, None, None], shape),
                        np.zeros(shape)))
    V[..., 3] = 300.0
    Q1 = np.zeros((nz, m.ny, m.nx, 4))
    assert np.all(o.nut(V, Q1) == 0)
    a = 4e-3  # linear shear du/dz relative to the base wind
    zc = (np.arange(nz + 1) + 0.5) * m.dz
    V[..., 0] = (u0 + a * zc)[:, None, None]
    aux = o.nut(V, Q1)
    dx, dz, dt = m.dx, m.dz, 6.0
    cs = float(np.float32(0.18))  # the uniform block carries cs in FP32
    nuh = min((cs * dx) ** 2 * a, 0.08 * dx * dx / dt)
    nuv = min((0.18 * dz) ** 2 * a, 0.15 * dz * dz / dt)
    # rtol 1e-7: dz reaches the kernel as FP32 (20000/12 is not representable)
    assert np.allclose(aux[1:-1, :, :, 0], nuh, rtol=1e-7)
    assert np.allclose(aux[1:-1, :, :, 1], nuv, rtol=1e-7)


def _micro_states(rng, n, warm):
    st = np.zeros((n, 8))
    st[:, 0] = rng.uniform(285, 300, n) if warm else rng.uniform(215, 265, n)  # theta, converted below
    st[:, 1] = rng.uniform(1e-4, 2e-2, n)
    st[:, 2:7] = rng.uniform(0, 4e-3, (n, 5)) * (rng.random((n, 5)) < 0.6)
    st[:, 7] = rng.uniform(0, 3e5, n)
    return st


@pytest.mark.parametrize("scheme", ["lin", "thompson"])
@pytest.mark.parametrize("warm", [True, False])
def test_microphysics_conserves_total_water_and_positivity(scheme, warm):
    m, U = small(scheme=scheme)
    o = Oracle(U, m.B, "f64")
    B = m.B.astype(np.float64).reshape(-1, 8)
    rng = np.random.default_rng(5)
    for k in (1, 4, 8):
        for st in _micro_states(rng, 60, warm):
            st = st.copy()
            st[0] = st[0] / B[k, 4]  # T -> theta at this level
            out = o.micro_point(k, st)
            w0, w1 = st[1:7].sum(), out[1:7].sum()
            assert abs(w1 - w0) <= 1e-12 * max(w0, 1e-3)
            assert np.all(out[1:7] >= -1e-15)
            assert np.isfinite(out).all()


def test_saturation_adjustment_reaches_saturation():
    m, U = small()
    o = Oracle(U, m.B, "f64")
    B = m.B.astype(np.float64).reshape(-1, 8)
    k = 2
    pi0, p0 = B[k, 4], B[k, 5]
    T = 295.0

    def qvs(T):
        e = 611.2 * math.exp(17.67 * (T - 273.15) / (T - 29.65))
        return 0.622 * e / max(p0 - e, 100)

    st = np.array([T / pi0, 1.05 * qvs(T), 0, 0, 0, 0, 0, 0])
    out = o.micro_point(k, st)
    T1 = out[0] * pi0
    assert out[2] > 0  # condensed
    assert out[1] == pytest.approx(qvs(T1), rel=2e-4)
    assert T1 > T  # latent heating
    # a subsaturated cloud-free point is left alone (no condensate appears)
    st = np.array([T / pi0, 0.5 * qvs(T), 0, 0, 0, 0, 0, 0])
    assert np.array_equal(o.micro_point(k, st), st)


@pytest.mark.parametrize("scheme", ["lin", "thompson"])
def test_microphysics_is_the_identity_in_clear_subsaturated_air(scheme):
    """The block leaves a condensate-free point below saturation unchanged but for Thompson's Nr (set to 0) and
    theta's round trip through T: the property the physics kernel's clear-air branch relies on. Points range from
    warm to far below freezing and from dry to 0.999 of saturation (the kernel's margin)."""
    m, U = small(scheme=scheme)
    o = Oracle(U, m.B, "f32")
    B = m.B.astype(np.float32).reshape(-1, 8)
    rng = np.random.default_rng(11)
    f32 = np.float32
    n = 0
    for k in range(B.shape[0] - 1):
        pi0, p0 = B[k, 4], B[k, 5]
        for T in rng.uniform(200.0, 310.0, 40):
            f = min(max((273.15 - T) / 40.0, 0.0), 1.0)
            esl = 611.2 * math.exp(17.67 * (T - 273.15) / (T - 29.65))
            esi = 611.2 * math.exp(21.8745584 * (T - 273.16) / (T - 7.66))
            e = esl + (esi - esl) * f
            qs = 0.622 * e / max(p0 - e, 100.0)
            qv = qs * rng.uniform(0.0, 0.997)
            th = f32(f32(T) / pi0)
            st = np.array([th, qv, 0, 0, 0, 0, 0, rng.uniform(0, 1e5)], np.float32)
            out = o.micro_point(k, st)
            assert out[0] == f32(f32(th * pi0) / pi0)
            assert np.array_equal(out[1:7], st[1:7])
            assert out[7] == (0.0 if scheme == "thompson" else st[7])
            n += 1
    assert n >= 400
    # a point with a trace of any one species, or at saturation, is not left alone
    k, T = 4, 285.0
    pi0 = B[k, 4]
    st = np.array([T / pi0, 0.03, 0, 0, 0, 0, 0, 0], np.float32)
    assert not np.array_equal(o.micro_point(k, st)[1:3], st[1:3])


def test_microphysics_warm_branches():
    m, U = small(dt=6.0)
    o = Oracle(U, m.B, "f64")
    B = m.B.astype(np.float64).reshape(-1, 8)
    k = 1
    pi0 = B[k, 4]
    # cloud water below qc0 = 2 g/kg and no rain: Kessler makes no rain
    st = np.array([300.0 / pi0, 0.05, 1e-3, 0, 0, 0, 0, 0])
    assert o.micro_point(k, st)[4] == 0
    # above qc0 in saturated air: autoconversion auto (qc - qc0) dt (FP32 parameters), no accretion at qr = 0
    st = np.array([300.0 / pi0, 0.05, 3e-3, 0, 0, 0, 0, 0])
    out = o.micro_point(k, st)
    auto, qc0 = float(np.float32(1e-3)), float(np.float32(2e-3))
    assert out[4] == pytest.approx(auto * (3e-3 - qc0) * 6.0, rel=1e-12)
    # melting above 0 C: snow loses min(0.012 dTc qs dt, qs) (saturated air: no evaporation)
    T = 280.0
    st = np.array([T / pi0, 0.05, 0, 0, 0, 2e-3, 0, 0])
    out = o.micro_point(k, st)
    assert 2e-3 - out[5] == pytest.approx(0.012 * (T - 273.15) * 2e-3 * 6.0, rel=1e-9)


def test_divergence_vanishes_after_projection():
    m, U = small(nx=24, ny=16, nz=10)
    o = Oracle(U, m.B, "f64")
    rng = np.random.default_rng(6)
    shape = (m.nz + 1, m.ny + 1, m.nx + 1)
    V = base_V(m, wind=(rng.standard_normal(shape), rng.standard_normal(shape), rng.standard_normal(shape)))
    rhs = o.div(V)
    phi = o.solve(rhs, "f64")
    d1 = o.div(o.proj(V, phi))
    p = m.pad
    inner = d1[:, p:p + m.ny, p:p + m.nx]
    assert np.abs(inner).max() <= 1e-9 * np.abs(rhs).max()


def test_phys_speed_cap_and_nan_guard():
    m, U = small()
    o = Oracle(U, m.B, "f64")
    B = m.B.astype(np.float64).reshape(-1, 8)
    V = base_V(m)
    V[3, 5, 6, 0] = 300.0
    V[4, 6, 7, 3] = np.nan
    Q = np.zeros((m.nz, m.ny, m.nx, 4))
    A = np.zeros((m.ny, m.nx, 4)); A[..., 1] = 100; A[..., 3] = 100
    Bl = np.zeros((m.ny, m.nx, 4)); Bl[..., 2] = 100
    ci = np.zeros((m.ny, m.nx, 4)); ci[..., 2:] = 1500
    aux = np.zeros((m.nz, m.ny, m.nx, 2))
    Vo, _, _ = o.phys(V, Q, Q, aux, A, Bl, np.zeros((m.nz, m.ny, m.nx)), ci)
    assert np.linalg.norm(Vo[3, 5, 6, :3]) == pytest.approx(110.0, rel=1e-12)
    assert Vo[3, 5, 6, 0] > 0
    assert np.array_equal(Vo[4, 6, 7], [B[4, 6], B[4, 7], 0.0, B[4, 0]])


def test_lsf_without_boundaries_and_ci_without_events():
    m, U = small()
    o = Oracle(U, m.B, "f32")
    A, B, ph = o.lsf(np.zeros(8 * 148, np.float32))
    assert np.all(A == np.array([0, 100, 0, 100], np.float32))
    assert np.all(B == np.array([0, 0, 100, 0], np.float32))
    assert np.all(ph == 0)
    assert np.all(o.ci(None) == np.array([0, 0, 1500, 1500], np.float32))


def test_lsf_straight_front_closed_form():
    m = host_model(40, 24, 12, 2000.0)
    m.bounds = [None]
    U = uniforms(m, 6.0)
    o = Oracle(U, m.B, "f64")
    bd = np.zeros(8 * 148, np.float32)
    x0 = 0.5 * m.Lx
    bd[0:20] = [2, 0, 0, 5000, -5, 2000, 0.05, 500, 0.1, 1.0, 0, 0, 1.0, 0, 0, 1, 0, 0, 60000, 0]
    bd[20:28] = [x0, -1e6, 0, 0, x0, 1e6, 0, 0]
    A, Bm, ph = o.lsf(bd)
    # west of the line (inside the air mass, far beyond the half width): full anomaly
    assert np.allclose(A[:, 2], [-5.0, 2000.0, 0.1, 2000.0])
    # east of it: nothing
    assert np.allclose(A[:, -3], [0.0, 100.0, 0.0, 100.0])
    # phT is the downward trapezoid integral of the target buoyancy (d phi / dz = b, 0 at the lid)
    Bs = m.B.astype(np.float64).reshape(-1, 8)
    z = (np.arange(m.nz) + 0.5) * m.dz

    def tprof(zz, H):
        t = np.clip((zz - 0.7 * H) / (0.45 * H), 0, 1)
        return 1 - t * t * (3 - 2 * t)

    bt = 9.81 * ((-5.0 / Bs[:m.nz, 4] * tprof(z, 2000.0)) / Bs[:m.nz, 0] + 0.608 * (-Bs[:m.nz, 1] * 0.1 * tprof(z, 2000.0)))
    phi = np.zeros(m.nz)
    phi[-1] = -bt[-1] * 0.5 * m.dz
    for k in range(m.nz - 2, -1, -1):
        phi[k] = phi[k + 1] - 0.5 * (bt[k] + bt[k + 1]) * m.dz
    assert np.allclose(ph[:, 5, 2], phi, rtol=1e-6, atol=1e-9)
    assert ph[0, 5, 2] > 0  # a cold air mass raises the surface pressure


def test_ci_single_tower_closed_form():
    m = host_model(40, 40, 12, 2000.0)
    m.ciActive = 1
    m.time = 400.0
    U = uniforms(m, 6.0)
    o = Oracle(U, m.B, "f64")
    xc, yc = 20.5 * m.dx, 19.5 * m.dx
    ev = np.array([xc, yc, 1500, 12000, 1500, 10, 0, 3600, 0, 0, 1.37, 150], np.float32)
    cm = o.ci(ev)
    assert np.allclose(cm[19, 20], [1.0, 10.0, 1500, 1500], rtol=1e-12)
    assert np.allclose(cm[2, 2], [0, 0, 1500, 1500])


def test_init_target_noise_and_humidity_cap():
    m, U = small(nz=40)
    o = Oracle(U, m.B, "f64")
    A, B, _ = o.lsf(np.zeros(8 * 148, np.float32))
    V, Q1, Q2 = o.init(A, B, 0)
    Bs = m.B.astype(np.float64).reshape(-1, 8)
    th0 = Bs[: m.nz, 0]
    zc = (np.arange(m.nz) + 0.5) * m.dz
    dth = V[: m.nz, :, :, 3] - th0[:, None, None]
    noise = 0.1
    assert np.abs(dth).max() <= noise * (1 + 1e-9)
    assert np.all(dth[zc >= 1500] == 0)
    assert np.abs(dth[zc < 800]).max() > 0.5 * noise
    assert np.allclose(V[: m.nz, : m.ny, :, 0], Bs[: m.nz, 6][:, None, None])
    assert np.all(V[..., 2] == 0) and np.all(Q2 == 0)
    assert np.all(Q1[..., 1:] == 0) and np.all(Q1[..., 0] <= Bs[: m.nz, 1][:, None, None] * (1 + 1e-12))
    # another seed changes the noise, seed 0 is the reference's (offset 17, 3, 5)
    V1, _, _ = o.init(A, B, 1)
    assert not np.array_equal(V1, V)


def test_fp32_and_fp64_oracles_agree():
    m, U = small(scheme="thompson")
    rng = np.random.default_rng(7)
    shape = (m.nz + 1, m.ny + 1, m.nx + 1)
    V = base_V(m, wind=(rng.uniform(-20, 20, shape), rng.uniform(-20, 20, shape), rng.uniform(-5, 5, shape)))
    V[..., 3] += rng.standard_normal(shape)
    o32, o64 = Oracle(U, m.B, "f32"), Oracle(U, m.B, "f64")
    V32 = V.astype(np.float32)
    a = o32.advv2(V32, o32.advv1(V32))
    b = o64.advv2(V32.astype(np.float64), o64.advv1(V32.astype(np.float64)))
    assert rel_l2(a[..., :3], b[..., :3]) < 1e-6


def test_split_fp64_step_equals_whole_step():
    """The FP64 step in two parts around the solve (the bench's full-size FP64 report) is the whole FP64 step when
    the caller supplies the oracle's own FP64 solve; with scipy's FP64 DST solve in its place the state agrees to
    FP64 rounding."""
    import numpy as np
    from _cases import host_model, uniforms
    from _oracle import Oracle, solve, step_f64, step_f64_split
    from _ref_solve import solve_scipy

    m = host_model(24, 16, 12, 2000.0, scheme="thompson")
    o = Oracle(uniforms(m, 1.0), m.B, "f64")
    A, B, ph = o.lsf(np.zeros(8 * 148, np.float32))
    V, Q1, Q2 = o.init(A, B, 3)
    V = o.bubble(V, 0.5 * m.Lx, 0.5 * m.Ly, 1500.0, 10000.0, 1500.0, 2.0)
    c = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
    def state():
        return dict(V2=c(np.stack([V, V])), Q12=c(np.stack([Q1, Q1])), Q22=c(np.stack([Q2, Q2])),
                    Vt=np.zeros_like(c(V)), Q1t=np.zeros_like(c(Q1)), Q2t=np.zeros_like(c(Q1)),
                    aux=np.zeros(Q1.shape[:3] + (2,)), phT=c(ph), lsfA=c(A), lsfB=c(B),
                    ciMap=np.zeros(Q1.shape[1:3] + (4,)), rhs=np.zeros((m.nz, m.nyp, m.nxp)),
                    phi=np.zeros((m.nz, m.nyp, m.nxp)))
    U = uniforms(m, 6.0)
    bd = np.zeros(8 * 148, np.float32)
    whole = state()
    step_f64(U, m.B, bd, None, True, 0, **whole)
    rc, rw = o.densities()
    for solver in ("oracle", "scipy"):
        s = state()
        step_f64_split(U, m.B, bd, None, True, 0, **s, part=0)
        if solver == "oracle":
            s["phi"][...] = solve(m.nxp, m.nyp, m.nz, float(U[3]), float(U[4]), rc, rw, s["rhs"], "f64")
        else:
            s["phi"][...] = solve_scipy(s["rhs"], float(U[3]), float(U[4]), rc, rw)
        step_f64_split(U, m.B, bd, None, True, 0, **s, part=1)
        for k in whole:
            if solver == "oracle":
                assert np.array_equal(s[k], whole[k]), k
            else:
                scale = np.abs(whole[k]).max() or 1.0
                assert np.abs(s[k] - whole[k]).max() <= 1e-11 * scale, k
