"""GPU parity on the full-size workloads and on the boundary and convective-initiation paths.

- Boundaries and towers, 128 x 128 x 40 at 2 km: two moving boundaries in bndModel's record layout
  (app.js:527-541), one a bent three-point outflow with along-front variation, hot spots and nudging, the other a
  drying line; four towers (one moving, one fading out, two overlapping, one not started yet). The large-scale
  forcing (model.js:142-244, noise.js:19-30), the CI map (model.js:249-281) and the step with its CI push
  (model.js:683-694) against the oracle.
- The 1024 x 1024 x 80 workload at 500 m (Thompson, 8 bubbles) after 20 minutes, and the 1024 x 1024 x 100
  workload at 250 m with its cold-pool boundary after 5 minutes: one step kernel by kernel against the oracle fed
  the GPU's own inputs at full size, the solve against an FP64 solve of the GPU's RHS, and the post-projection
  divergence. The 250 m workload is checked in the plain form with the split-BF16 transforms its combined form
  uses (the surrogate has its own tests).

Tolerances are the stencil kernels' FP32 bound (relative L2 1e-5 per field, theta' RMS 1e-4 K), as in the
step tests on the 128 x 128 x 40 workload. At full size two outputs carry more FP32 sensitivity than that, and the
FP32 restatement itself differs from the FP64 oracle there by more than 1e-5: the boundary's target winds (lsfB),
where a grid row passes within metres of a polyline vertex and FP32 cannot tell which of the two segments is
nearer (the along-front position, and with it the hot-spot lift, then differs by the metres between them), and
the physics kernel's scarce cloud ice, whose saturation-adjustment branches flip at a few points. For those two the
bound is twice the FP32 restatement's own error against FP64 from the same inputs, or 1e-5 if larger.
"""
import numpy as np
import pytest

from _oracle import Oracle, step_f32
from _ref_solve import solve_scipy
from _util import rel_l2, theta_pert

pytestmark = pytest.mark.gpu

TOL = 1e-5
THETA_RMS_K = 1e-4
THREADS = 16  # the host's CPU share per GPU
V_CH = ["u", "v", "w", "theta"]
Q1_CH = ["qv", "qc", "qi", "qr"]
Q2_CH = ["qs", "qg", "Nr"]


def _cmp(name, gpu, ref, chans, errs, BS=None, nz=None):
    for i, c in enumerate(chans):
        errs[f"{name}.{c}"] = rel_l2(gpu[..., i], ref[..., i])
        if c == "theta":
            d = theta_pert(gpu, BS, nz) - theta_pert(ref, BS, nz)
            errs[f"{name}.theta_rms_K"] = float(np.sqrt(np.mean(d * d)))


def _bad(errs, tols=None):
    tols = tols or {}
    return {k: v for k, v in errs.items() if not v <= (THETA_RMS_K if k.endswith("_rms_K") else tols.get(k, TOL))}


def _report(tag, errs):
    for k, v in errs.items():
        print(f"{tag} {k:18s} {v:.3e}", flush=True)


# ---------------------------------------------------------------- boundaries and towers

def _setup_fronts(m):
    from outbreaklab_amd.sounding import wk_sounding_text

    m.allocate("custom", {"key": "custom:2000:128x128x40", "dx": 2000, "n": [128, 128, 40]})
    m.params["scheme"] = "thompson"
    m.params["noise"] = 0.1
    m.params["lat"] = 35.9
    m.setSounding(wk_sounding_text(th0=300.0, thtr=343.0, qv0=0.014, shear6=30.0, hodo_r=5.0))
    L, W = m.Lx, m.Ly
    ws = m.baseWind(150.0)
    outflow = {"pts": [[0.35 * L, 0.1 * W], [0.3 * L, 0.5 * W], [0.38 * L, 0.9 * W]], "motion": [6.0, 1.0],
               "halfW": 6000.0, "dT": -5.0, "Hmax": 1800.0, "slope": 0.05, "Hmin": 300.0, "dq": 0.1, "qRatio": 1.0,
               "du": 8.0 - ws[0], "dv": -ws[1], "wRatio": 1.0, "nudge": True, "dFree": 25000.0, "side": 1,
               "vary": 0.8, "seed": 17.3, "varyL": 40000.0}
    dryline = {"pts": [[0.7 * L, 0.2 * W], [0.75 * L, 0.8 * W]], "motion": [3.0, 2.0], "halfW": 4000.0, "dT": 1.5,
               "Hmax": 1200.0, "slope": 0.03, "Hmin": 200.0, "dq": 0.35, "qRatio": 1.5, "du": 2.0, "dv": -3.0,
               "wRatio": 0.8, "nudge": False, "dFree": 25000.0, "side": -1, "vary": 0.4, "seed": 5.1,
               "varyL": 60000.0}
    m.setBoundaries([outflow, dryline])
    towers = [
        {"x": 0.5 * L, "y": 0.5 * W, "zc": 1500, "rh": 12000, "rz": 1500, "wt": 8, "t0": 0, "t1": 3600, "vx": 4,
         "vy": 2, "ramp": 150},
        {"x": 0.6 * L, "y": 0.3 * W, "zc": 2000, "rh": 8000, "rz": 1800, "wt": 6, "t0": 30, "t1": 240, "ramp": 60},
        {"x": 0.55 * L, "y": 0.52 * W, "zc": 1000, "rh": 15000, "rz": 1200, "wt": 10, "t0": 0, "t1": 3600,
         "ramp": 300},
        {"x": 0.2 * L, "y": 0.7 * W, "zc": 1500, "rh": 10000, "rz": 1500, "wt": 8, "t0": 500, "t1": 2000},
    ]
    m.setSurface(None, None)
    m.reset()
    m.setCI(towers)
    m.add_bubble(0.5 * L, 0.5 * W, 1500.0, 10000.0, 1500.0, 2.0)
    return 6.0


@pytest.fixture(scope="module")
def fronts():
    from outbreaklab_amd import RegionalModel

    m = RegionalModel()
    dt = _setup_fronts(m)
    yield m, dt
    m.destroy()


def test_boundary_forcing_at_reset_matches_oracle(fronts):
    m, _ = fronts
    o = Oracle(m._u_reset, m._B)
    A, B, ph = o.lsf(m._bd)
    gA, gB, gph = m.lsfA.download(), m.lsfB.download(), m.phT.download()
    # the boundaries are live: an air mass, a varying depth along the front, nudging and a target phi
    assert A[..., 0].min() < -1 and np.ptp(A[A[..., 0] < -1][:, 1]) > 100 and B[..., 3].max() > 0
    assert np.abs(ph).max() > 1
    errs = {f"lsfA.{i}": rel_l2(gA[..., i], A[..., i]) for i in range(4)}
    errs.update({f"lsfB.{i}": rel_l2(gB[..., i], B[..., i]) for i in range(4)})
    errs["phT"] = rel_l2(gph, ph)
    _report("reset", errs)
    assert not _bad(errs), _bad(errs)


def test_boundaries_and_towers_in_the_step_match_oracle(fronts):
    m, dt = fronts
    while not (m.steps >= 20 and m.lsf_due()):
        m.step(dt)
    m.sync()
    assert m.time == pytest.approx(120.0)
    p = m.parity
    V2 = np.stack([m.V[0].download(), m.V[1].download()])
    Q12 = np.stack([m.Q1[0].download(), m.Q1[1].download()])
    Q22 = np.stack([m.Q2[0].download(), m.Q2[1].download()])
    lsfA, lsfB, phT, ciMap = m.lsfA.download(), m.lsfB.download(), m.phT.download(), m.ciMap.download()
    aux, Vt, Q1t, Q2t = m.aux.download(), m.Vt.download(), m.Q1t.download(), m.Q2t.download()
    rhs, phi = m.rhs.download(), m.phi.download()
    V0, Q10, Q20 = V2[p].copy(), Q12[0].copy(), Q22[0].copy()
    m.step(dt)
    m.sync()
    U = m._u_last
    o = Oracle(U, m._B)
    errs = {}
    # the moved boundaries' forcing and the towers' map at t = 120 s
    A, B, ph = o.lsf(m._bd)
    gA, gB, gph = m.lsfA.download(), m.lsfB.download(), m.phT.download()
    errs.update({f"lsfA.{i}": rel_l2(gA[..., i], A[..., i]) for i in range(4)})
    errs.update({f"lsfB.{i}": rel_l2(gB[..., i], B[..., i]) for i in range(4)})
    errs["phT"] = rel_l2(gph, ph)
    ci = o.ci(m._ci)
    gci = m.ciMap.download()
    assert m.ciActive == 4 and (ci[..., 0] > 0).sum() > 50
    errs.update({f"ci.{i}": rel_l2(gci[..., i], ci[..., i]) for i in range(4)})
    # the physics kernel, with the CI push, from the GPU's own advected state
    Vadv = o.advv2(V0, o.advv1(V0))
    Q1a, Q2a = m.Q1[1].download(), m.Q2[1].download()
    Vp, Q1p, Q2p = o.phys(Vadv, Q1a, Q2a, m.aux.download(), gA, gB, gph, gci)
    Vphys = m.V[p].download()
    _cmp("phys", Vphys, Vp, V_CH, errs, m._B, m.nz)
    _cmp("phys.q1", m.Q1[0].download(), Q1p, Q1_CH, errs)
    _cmp("phys.q2", m.Q2[0].download(), Q2p, Q2_CH, errs)
    # the CI push acts: w below the tower centres exceeds what the step without the towers gives
    Vnp, _, _ = o.phys(Vadv, Q1a, Q2a, m.aux.download(), gA, gB, gph, o.ci(None))
    assert (Vp[..., 2] - Vnp[..., 2]).max() > 0.05
    # the whole step chained on the oracle
    step_f32(U, m._B, m._bd, m._ci, True, p, V2, Q12, Q22, Vt, Q1t, Q2t, aux, phT, lsfA, lsfB, ciMap, rhs, phi)
    _cmp("step.V", m.V[1 - p].download(), V2[1 - p], V_CH, errs, m._B, m.nz)
    _cmp("step.Q1", m.Q1[0].download(), Q12[0], Q1_CH, errs)
    _cmp("step.Q2", m.Q2[0].download(), Q22[0], Q2_CH, errs)
    _report("fronts", errs)
    assert not _bad(errs), _bad(errs)


# ---------------------------------------------------------------- full-size workloads

def _capture_step(m, dt):
    """Download the state before one step and every intermediate after it (the step the checks compare)."""
    p = m.parity
    st = {"V0": m.V[p].download(), "Q10": m.Q1[0].download(), "Q20": m.Q2[0].download(), "p": p,
          "run_lsf": m.lsf_due()}
    m.step(dt)
    m.sync()
    st.update(U=m._u_last, aux=m.aux.download(),
              Q1a=m.Q1[1].download(), Q2a=m.Q2[1].download(), Vphys=m.V[p].download(), Q1n=m.Q1[0].download(),
              Q2n=m.Q2[0].download(), rhs=m.rhs.download(), phi=m.phi.download(), Vn=m.V[1 - p].download(),
              lsfA=m.lsfA.download(), lsfB=m.lsfB.download(), phT=m.phT.download(), ciMap=m.ciMap.download())
    return st


def _kernels(m, st, tag):
    """Every kernel of the step against the oracle fed the GPU's inputs; returns (errors, per-key tolerances)."""
    o = Oracle(st["U"], m._B, threads=THREADS)
    o64 = Oracle(st["U"], m._B, "f64", threads=THREADS)
    errs, tols = {}, {}
    if st["run_lsf"]:
        A, B, ph = o.lsf(m._bd)
        A6, B6, ph6 = o64.lsf(m._bd)
        for nm, g, r, r6 in (("lsfA", st["lsfA"], A, A6), ("lsfB", st["lsfB"], B, B6)):
            for i in range(4):
                errs[f"{nm}.{i}"] = rel_l2(g[..., i], r[..., i])
                tols[f"{nm}.{i}"] = max(TOL, 2 * rel_l2(r[..., i], r6[..., i]))
        errs["phT"] = rel_l2(st["phT"], ph)
        tols["phT"] = max(TOL, 2 * rel_l2(ph, ph6))
        del A6, B6, ph6
    _cmp("nut", st["aux"], o.nut(st["V0"], st["Q10"]), ["nuh", "nuv"], errs)
    # the moisture advection runs as one kernel (predictor and corrector): its output against the oracle's two
    r1, r2 = o.advq2(st["V0"], st["Q10"], st["Q20"], *o.advq1(st["V0"], st["Q10"], st["Q20"]))
    _cmp("advQ.q1", st["Q1a"], r1, Q1_CH, errs)
    _cmp("advQ.q2", st["Q2a"], r2, Q2_CH, errs)
    del r1, r2
    # the wind advection (one kernel) is checked with the physics kernel that consumes it
    Vadv = o.advv2(st["V0"], o.advv1(st["V0"]))
    phys_in = (st["Q1a"], st["Q2a"], st["aux"], st["lsfA"], st["lsfB"], st["phT"], st["ciMap"])
    Vp, Q1p, Q2p = o.phys(Vadv, *phys_in)
    _cmp("advV+phys", st["Vphys"], Vp, V_CH, errs, m._B, m.nz)
    _cmp("phys.q1", st["Q1n"], Q1p, Q1_CH, errs)
    _cmp("phys.q2", st["Q2n"], Q2p, Q2_CH, errs)
    _, Q1p6, Q2p6 = o64.phys(Vadv, *phys_in)
    for nm, r, r6, chans in (("phys.q1", Q1p, Q1p6, Q1_CH), ("phys.q2", Q2p, Q2p6, Q2_CH)):
        for i, c in enumerate(chans):
            tols[f"{nm}.{c}"] = max(TOL, 2 * rel_l2(r[..., i], r6[..., i]))
    del Vadv, Vp, Q1p, Q2p, Q1p6, Q2p6
    errs["div"] = rel_l2(st["rhs"], o.div(st["Vphys"]))
    _cmp("proj", st["Vn"], o.proj(st["Vphys"], st["phi"]), V_CH, errs, m._B, m.nz)
    _report(tag, errs)
    for k, t in tols.items():
        if t > TOL:
            print(f"{tag} {k:18s} tolerance {t:.3e} (twice the FP32 restatement's error vs FP64)", flush=True)
    return errs, tols


def _solve_and_divergence(m, st, tol_solve, tol_div):
    """The GPU's phi against the FP64 solve of its own RHS, and the divergence left after the projection at the
    model's cells, relative to the divergence the solve removed (bounded by tol_div)."""
    o = Oracle(st["U"], m._B, "f64", threads=THREADS)
    rc, rw = o.densities()
    ref = solve_scipy(st["rhs"], m.dx, m.dz, rc, rw)
    e_phi = rel_l2(st["phi"], ref)
    p = m.pad
    d1 = o.div(st["Vn"].astype(np.float64))[:, p:p + m.ny, p:p + m.nx]
    d0 = st["rhs"][:, p:p + m.ny, p:p + m.nx].astype(np.float64)
    # the same projection with the FP64 phi: the level the FP32 state itself leaves
    d64 = o.div(o.proj(st["Vphys"].astype(np.float64), ref))[:, p:p + m.ny, p:p + m.nx]
    res, res64 = np.linalg.norm(d1) / np.linalg.norm(d0), np.linalg.norm(d64) / np.linalg.norm(d0)
    print(f"solve rel err vs FP64 {e_phi:.3e}; divergence after projection {res:.3e} (FP64 phi {res64:.3e})",
          flush=True)
    assert e_phi <= tol_solve
    assert res <= tol_div


@pytest.fixture(scope="module")
def plain_1024():
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_plain_1024

    m = RegionalModel()
    dt = setup_plain_1024(m)
    for _ in range(599):
        m.step(dt)
    m.sync()
    st = _capture_step(m, dt)
    yield m, st
    m.destroy()


def test_plain_1024_state_after_20_minutes(plain_1024):
    m, st = plain_1024
    for k in ("V0", "Q10", "Q20", "Vn", "Q1n", "Q2n"):
        assert np.isfinite(st[k]).all(), k
    assert st["Q10"][..., 1].max() > 1e-4, "the storms should have cloud by 20 minutes"
    assert st["V0"][..., 2].max() > 5.0


def test_plain_1024_step_kernels_match_oracle(plain_1024):
    m, st = plain_1024
    errs, tols = _kernels(m, st, "1024x80")
    assert not _bad(errs, tols), _bad(errs, tols)


def test_plain_1024_solve_and_divergence(plain_1024):
    m, st = plain_1024
    # FP32 transforms: the divergence left is the FP32 rounding of the transforms (measured 9e-5)
    _solve_and_divergence(m, st, 1e-4, 2e-4)


@pytest.fixture(scope="module")
def combined_1024():
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_combined_1024

    m = RegionalModel()
    dt = setup_combined_1024(m)
    m.set_solver_mode("bf16x3", "mixed")
    while not (m.steps >= 300 and m.lsf_due()):
        m.step(dt)
    m.sync()
    st = _capture_step(m, dt)
    yield m, st
    m.destroy()


def test_combined_1024_boundary_step_kernels_match_oracle(combined_1024):
    m, st = combined_1024
    assert st["run_lsf"] and st["lsfA"][..., 0].min() < -3, "the cold-pool boundary's forcing should be live"
    for k in ("V0", "Q10", "Vn", "Q1n"):
        assert np.isfinite(st[k]).all(), k
    errs, tols = _kernels(m, st, "1024x100")
    assert not _bad(errs, tols), _bad(errs, tols)


def test_combined_1024_split_bf16_solve_and_divergence(combined_1024):
    m, st = combined_1024
    # split-BF16 transforms: each operand keeps 16 of its 24 bits (hi + lo), so the projection leaves about 30 times
    # the FP32 form's divergence (measured 2.7e-3) while phi stays within 1e-5 of FP64
    _solve_and_divergence(m, st, 1e-5, 5e-3)
