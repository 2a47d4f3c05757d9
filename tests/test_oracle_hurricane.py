"""Pins of the hurricane-mode oracle (model.js:88-119, 284-389, 392-576, 713-774, 895, 948-1026, 1033-1098)
and of the host's vortex table (setVortex, model.js:1318-1365), on the CPU.

Each check is fixed by something other than the restated formula, mostly on the FP64 oracle:
- the vortex table: the Rankine-type profile reaches vmax at the radius of maximum wind, the absolute
  angular momentum never decreases outward, the core is warm aloft and the centre's pressure low;
- the vortex in the initial state: the vortex's winds are the table's tangential winds (the difference from
  the vortex-free state), and the moist core's humidity is the asked-for fraction of saturation;
- the K-profile boundary layer: a uniform wind over a calm sea or land gives K = kappa u* z (1 - z/h)^2 with
  u* from the sea's Donelan-type drag or the land's log-law drag, in closed form;
- PREP: a pressure linear in x takes cn.x cn.y dt/dx of its gradient off u only;
- FILT: a 2-dx wave along x loses 16 a, a checkerboard 64 a, constants and linear fields nothing; w stays 0 at
  the ground and lid; hydrometeors stay non-negative;
- midpoint trajectories: identical to first-order ones in a uniform wind; in solid-body rotation they keep the
  speed that the first-order trace loses by drawing air from larger radius;
- the sea's fluxes: two seas differing only in SST change the lowest level's theta by dt C_k U (theta_s1 -
  theta_s2) and its vapour by dt C_k U (q_s1 - q_s2), with theta_s and q_s from the Exner function and the
  saturation mixing ratio at the surface pressure;
- the maps over the sea: a clear column's brightness temperature is the SST; maritime graupel's reflectivity
  is its power law.
"""
import math

import numpy as np
import pytest

from _cases import host_model
from _oracle import Oracle
from outbreaklab_amd.configs import HURRICANE_SETTINGS
from outbreaklab_amd.model import vortex_table
from outbreaklab_amd.uniforms import write_uniforms

N, NZ, DX = 40, 20, 2000.0


def hur_model(dt=6.0, steps=1, calm=False, vortex=None, **hur):
    m = host_model(N, N, NZ, DX, scheme="lin", noise=0.0)
    m.params["lat"] = 20.0
    h = dict(HURRICANE_SETTINGS, sfcP=101000.0, centre=[m.Lx / 2, m.Ly / 2], envR=1e6)
    h.update(hur)
    m.hur, m.steps, m.prep, m.vortex, m.phiScale = h, steps, True, vortex, 1 / 0.6
    if calm:
        m.B = m.B.copy()
        m.B[6::8] = 0.0  # no base wind
        m.B[7::8] = 0.0
    return m


def uniforms_of(m, dt=6.0):
    return write_uniforms(m.U, m, dt)


def qvsat64(T, p):
    e = 611.2 * math.exp(17.67 * (T - 273.15) / (T - 29.65))
    return 0.622 * e / max(p - e, 100.0)


VX = {"x": N * DX / 2, "y": N * DX / 2, "vmax": 30.0, "rmw": 20000.0, "r0": 60000.0, "ztop": 16000.0,
      "moistR": 30000.0, "moistRH": 0.9}


def test_vortex_table():
    m = hur_model()
    tab, info = vortex_table(NZ, m.dz, m.dx, 20.0, m.B, VX)
    dr, nr = info["dr"], info["nr"]
    assert dr == 1000.0 and nr == math.ceil(VX["r0"] * 2 / dr) + 3
    z0 = 0.5 * m.dz
    F0 = 1 - (z0 / VX["ztop"] - math.sin(2 * math.pi * z0 / VX["ztop"]) / (2 * math.pi))
    i = int(VX["rmw"] / dr)
    assert abs(tab[0, i, 0] - VX["vmax"] * F0) < 1e-4  # x = 1: (1 + a) / (a + 1) = 1, no taper inside r0 / 2
    assert np.argmax(tab[0, :, 0]) == i
    f = 2 * 7.292e-5 * math.sin(20.0 * math.pi / 180)
    r = np.arange(nr) * dr
    for k in range(NZ):
        M = r * tab[k, :, 0].astype(np.float64) + 0.5 * f * r * r
        assert np.all(np.diff(M) >= -1e-6 * np.abs(M).max()), k
    assert info["pc"] < -5 and info["warm"] > 0.5  # low pressure at the centre, a warm core aloft
    kmid = int(8000 / m.dz)
    assert tab[kmid, 0, 1] > tab[kmid, -1, 1] + 0.5


def test_init_vortex_and_moist_core():
    m = hur_model()
    tab, info = vortex_table(NZ, m.dz, m.dx, 20.0, m.B, VX)
    m.vortex = info
    U = uniforms_of(m, 1.0)
    o = Oracle(U, m.B, "f64")
    A, B, _ = o.lsf(np.zeros(8 * 148, np.float32))
    U0 = U.copy()
    U0[70] = 0.0  # vtx2.z: the same block without the vortex
    V0, Q0, _ = Oracle(U0, m.B, "f64").init(A, B)
    V1, Q1, _ = o.init2(A, B, tab)
    k = 2
    du, dv = V1[k, :, :, 0] - V0[k, :, :, 0], V1[k, :, :, 1] - V0[k, :, :, 1]
    # the u face at (i, j + 1/2): -v(r) (y - yc) / r
    xs = np.arange(N + 1) * DX
    ys = (np.arange(N + 1) + 0.5) * DX
    X, Y = np.meshgrid(xs - VX["x"], ys - VX["y"])
    R = np.maximum(np.hypot(X, Y), 1.0)
    vt = np.interp(R, np.arange(info["nr"]) * info["dr"], tab[k, :, 0].astype(np.float64), right=0.0)
    exp_u = -vt * Y / R
    assert np.max(np.abs(du[:N, :] - exp_u[:N, :])) < 1e-9 + 1e-6 * np.abs(exp_u).max()
    assert np.abs(dv).max() > 0.5 * VX["vmax"]
    # the moist core: qv = max(the environment's qv, RH qvs(T)) within moistR, between 1.5 and 7 km fully
    j = i = N // 2
    rc = math.hypot((i + 0.5) * DX - VX["x"], (j + 0.5) * DX - VX["y"])
    moistened = 0
    for kz in range(NZ):
        z = (kz + 0.5) * m.dz
        if not 1500 <= z <= 7000:
            continue
        th = V0[kz, j, i, 3] + np.interp(rc, np.arange(info["nr"]) * info["dr"], tab[kz, :, 1].astype(np.float64))
        T = th * float(m.B[kz * 8 + 4])
        if T < 273.15:
            continue
        qs = qvsat64(T, float(m.B[kz * 8 + 5]))
        # the environment's target humidity: the vortex-free state's, where it was not capped at 0.95 qvs
        T0 = V0[kz, j, i, 3] * float(m.B[kz * 8 + 4])
        if Q0[kz, j, i, 0] >= 0.95 * qvsat64(T0, float(m.B[kz * 8 + 5])) * (1 - 1e-12):
            continue
        env = min(Q0[kz, j, i, 0], 0.95 * qs)
        assert abs(Q1[kz, j, i, 0] - max(env, 0.9 * qs)) < 1e-7 * qs, kz
        moistened += 0.9 * qs > env
    assert moistened >= 1


@pytest.mark.parametrize("land", [0.0, 1.0])
def test_k_profile(land):
    m = hur_model(calm=True)
    U = uniforms_of(m)
    o = Oracle(U, m.B, "f64")
    th0 = np.asarray(m.B[0:8 * NZ:8], np.float64)
    V = np.zeros(o.vshape())
    V[..., 0] = 12.0
    V[..., 3] = 300.0  # uniform theta and wind: no strain, no stability
    Q1 = np.zeros(o.cshape())
    sfc = np.zeros((N, N, 4), np.float32)
    sfc[..., 0] = land
    sfc[..., 1] = 302.0
    aux = o.nut2(V, Q1, sfc)
    # the block's float32 settings: U10 factor, drag cap, land drag (log law for z0 = 0.3 m), gust, BL depth
    f, cap_cd, cd_land, gust, h = (float(U[i]) for i in (51, 60, 52, 53, 50))
    assert abs(cd_land - (0.41 / math.log(0.5 * m.dz / 0.3 + 1)) ** 2) < 1e-9 and h == 1000.0 and abs(f - 0.8) < 1e-7
    cd = min((0.75 + 0.067 * f * 12.0) * 1e-3, cap_cd) * f * f if land == 0 else cd_land
    ust = math.sqrt(cd) * max(12.0, gust)
    for k in range(NZ):
        z = (k + 0.5) * m.dz
        cap = 0.15 * m.dz * m.dz / 6.0
        exp = min(0.4 * ust * z * (1 - z / h) ** 2, cap) if z < h else 0.0
        assert abs(aux[k, 5, 7, 1] - exp) <= 1e-12 + 1e-9 * exp, (k, aux[k, 5, 7, 1], exp)
    assert np.all(aux[..., 0] == 0)
    del th0


def test_prep_linear_pressure():
    m = hur_model(steps=3)
    U = uniforms_of(m)
    assert abs(U[72] - 0.4) < 1e-7 and abs(U[73] - 1 / 0.6) < 1e-6  # cn: 40 % of the old force, phi scale
    o = Oracle(U, m.B, "f64")
    rng = np.random.default_rng(1)
    V = rng.normal(size=o.vshape())
    alpha = 3.0
    phi = np.broadcast_to(alpha * np.arange(o.nxp, dtype=np.float64), (o.nz, o.nyp, o.nxp)).copy()
    Vp = o.prep(V, phi)
    kx = float(U[72]) * float(U[73]) * 6.0 / DX
    assert np.allclose(Vp[:NZ, :N, :, 0], V[:NZ, :N, :, 0] - kx * alpha, rtol=0, atol=1e-12)
    assert np.array_equal(Vp[..., 1:], V[..., 1:])


def test_filter_waves():
    m = hur_model()
    U = uniforms_of(m)
    a = float(U[62])
    assert abs(a - 4 * 0.003) < 1e-9
    o = Oracle(U, m.B, "f64")
    x = np.arange(N + 1)
    y = np.arange(N + 1)[:, None]
    V = np.zeros(o.vshape())
    V[..., 0] = 5.0 + 2.0 * (-1.0) ** x                       # 2-dx wave along x
    V[..., 1] = 3.0 * (-1.0) ** (x + y)                       # checkerboard
    V[..., 2] = 1.0
    V[..., 3] = 300.0 + 0.01 * x                              # linear
    Q1 = np.zeros(o.cshape())
    Q1[:, 5, N // 2, 3] = 1e-3                                # a spike: its diagonal neighbours would go negative
    Q2 = np.zeros(o.cshape())
    Vo, Q1o, Q2o = o.filt(V, Q1, Q2)
    s = (slice(None), slice(2, N - 2), slice(2, N - 2))
    assert np.allclose(Vo[s + (0,)], (5.0 + 2.0 * (1 - 16 * a) * (-1.0) ** x)[2:N - 2][None, None, :], atol=1e-12)
    assert np.allclose(Vo[s + (1,)], (3.0 * (1 - 64 * a) * (-1.0) ** (x + y))[2:N - 2, 2:N - 2][None], atol=1e-12)
    assert np.allclose(Vo[s + (3,)], V[s + (3,)], atol=1e-12)
    assert np.all(Vo[0, ..., 2] == 0) and np.all(Vo[NZ, ..., 2] == 0) and np.allclose(Vo[1:NZ, ..., 2], 1.0)
    assert Q1o.min() >= 0 and Q1o[0, 5, N // 2, 3] == pytest.approx(1e-3 * (1 - 20 * a))
    assert Q1o[0, 6, N // 2 + 1, 3] == 0 and Q1o[0, 5, N // 2 + 1, 3] == pytest.approx(8 * a * 1e-3)


def test_midpoint_trajectories():
    m = hur_model()
    U = uniforms_of(m)
    assert U[75] == 1.0
    U0 = U.copy()
    U0[75] = 0.0
    o1, o0 = Oracle(U, m.B, "f64"), Oracle(U0, m.B, "f64")
    V = np.zeros(o1.vshape())
    V[..., 0], V[..., 1], V[..., 3] = 7.0, -4.0, 300.0
    assert np.array_equal(o1.advv1_2(V, V), o0.advv1_2(V, V))
    # solid-body rotation (u = -W y, v = W x about the centre, linear so trilinear samples are exact): the steady
    # field's advected u at a u face is -W y_d at the true departure point, rotated back by W dt
    W = 2e-3
    c = N * DX / 2
    yc = (np.arange(N + 1) + 0.5) * DX - c
    V[..., 0] = -W * yc[None, :, None]
    V[..., 1] = W * ((np.arange(N + 1) + 0.5) * DX - c)[None, None, :]
    r1, r0 = o1.advv1_2(V, V), o0.advv1_2(V, V)
    th = W * 6.0
    X = (np.arange(N + 1) * DX - c)[None, :]
    Y = yc[:, None]
    ref = -W * (-X * math.sin(th) + Y * math.cos(th))
    sl = (5, slice(10, N - 10), slice(10, N - 10))
    e1 = np.abs(r1[sl + (0,)] - ref[10:N - 10, 10:N - 10]).mean()
    e0 = np.abs(r0[sl + (0,)] - ref[10:N - 10, 10:N - 10]).mean()
    assert e1 < 0.05 * e0, (e1, e0)


def test_sea_fluxes():
    outs = []
    for sst in (300.0, 305.0):
        m = hur_model(calm=True, thermals=0.0)
        U = uniforms_of(m, 60.0)
        o = Oracle(U, m.B, "f64")
        A, B, ph = o.lsf(np.zeros(8 * 148, np.float32))
        V, Q1, Q2 = o.init(A, B)
        aux = np.zeros(o.cshape(2))
        ci = np.zeros((N, N, 4))
        ci[..., 2] = 1500
        ci[..., 3] = 1500
        sfc = np.zeros((N, N, 4), np.float32)
        sfc[..., 1] = sst
        phi = np.zeros((NZ, o.nyp, o.nxp))
        outs.append(o.phys2(V, Q1, Q2, aux, A, B, ph, ci, sfc, phi))
        ck, u10f, gust = U[45], U[51], U[53]
    (Va, Qa, _), (Vb, Qb, _) = outs
    j = i = N // 2
    ch = float(ck) * float(u10f) * max(0.0, float(gust)) / m.dz
    ps = 101000.0
    ex = (1.0e5 / ps) ** 0.2854
    dth = Vb[0, j, i, 3] - Va[0, j, i, 3]
    assert dth == pytest.approx(60.0 * ch * (305.0 - 300.0) * ex, rel=1e-9)
    dq = Qb[0, j, i, 0] - Qa[0, j, i, 0]
    assert dq == pytest.approx(60.0 * ch * (qvsat64(305.0, ps) - qvsat64(300.0, ps)), rel=1e-9)
    assert np.array_equal(Va[1:, j, i], Vb[1:, j, i])  # only the lowest level feels the sea


def test_maps_over_the_sea():
    m = hur_model()
    U = uniforms_of(m)
    o = Oracle(U, m.B, "f64")
    V = np.zeros(o.vshape())
    V[..., 3] = 300.0
    Q1 = np.zeros(o.cshape())
    Q2 = np.zeros(o.cshape())
    k = 2
    Q2[k, 3, 4, 1] = 2e-3  # graupel in one cell, warm (300 K theta, pi < 1 ... T above freezing at 1.25 km)
    sfc = np.zeros((N, N, 4), np.float32)
    sfc[..., 1] = 301.5
    phi = np.zeros((NZ, o.nyp, o.nxp))
    M = o.maps2(V, Q1, Q2, phi, sfc)
    clear = np.ones((N, N), bool)
    clear[3, 4] = False
    assert np.allclose(M[3, ..., 1][clear], 301.5, rtol=0, atol=1e-9)
    rho = float(m.B[k * 8 + 2])
    T = 300.0 * float(m.B[k * 8 + 4])
    assert T > 273.15
    assert M[0, 3, 4, 0] == pytest.approx(10 * math.log10(1.0e10 * (rho * 2e-3) ** 1.75), rel=1e-6)
