"""Error of one GPU step against the FP64 oracle of the same step (shared by the FP64 test and the CPU leg of
bench.py): per prognostic field, the pressure and the post-projection divergence, each next to the error of the FP32
restatement of the reference's arithmetic from the same inputs."""
import numpy as np

from _oracle import Oracle, step_f32, step_f64
from _util import rel_l2, theta_pert

FIELDS = {"u": ("V", 0), "v": ("V", 1), "w": ("V", 2), "qv": ("Q1", 0), "qc": ("Q1", 1), "qi": ("Q1", 2),
          "qr": ("Q1", 3), "qs": ("Q2", 0), "qg": ("Q2", 1), "Nr": ("Q2", 2)}


def fp64_errors(m, dt, threads=None):
    """Run one GPU step of model m and the FP32 and FP64 oracle steps from the same state. Returns
    {field: (gpu vs FP64, FP32 restatement vs FP64)}, theta' RMS in kelvin likewise, phi, and
    "divergence": (gpu, FP32 restatement, FP64) post-projection divergence relative to the divergence removed,
    over the model's cells."""
    p = m.parity
    dl = lambda f: np.stack([f[0].download(), f[1].download()])  # noqa: E731
    ins = dict(V2=dl(m.V), Q12=dl(m.Q1), Q22=dl(m.Q2), Vt=m.Vt.download(), Q1t=m.Q1t.download(), Q2t=m.Q2t.download(),
               aux=m.aux.download(), phT=m.phT.download(), lsfA=m.lsfA.download(), lsfB=m.lsfB.download(),
               ciMap=m.ciMap.download(), rhs=m.rhs.download(), phi=m.phi.download())
    run_lsf = m.lsf_due()
    m.step(dt)
    m.sync()
    U, ci = m._u_last, (m._ci if m.ciActive else None)
    gpu = {"V": m.V[1 - p].download(), "Q1": m.Q1[0].download(), "Q2": m.Q2[0].download(), "phi": m.phi.download()}
    s32 = {k: np.ascontiguousarray(v, np.float32) for k, v in ins.items()}
    step_f32(U, m._B, m._bd, ci, run_lsf, p, **s32, threads=threads)
    f32 = {"V": s32["V2"][1 - p], "Q1": s32["Q12"][0], "Q2": s32["Q22"][0], "phi": s32["phi"]}
    rhs64 = np.ascontiguousarray(ins["rhs"], np.float64)
    s64 = {k: np.ascontiguousarray(v, np.float64) for k, v in ins.items()}
    del ins, s32
    step_f64(U, m._B, m._bd, ci, run_lsf, p, **s64, threads=threads)
    f64 = {"V": s64["V2"][1 - p], "Q1": s64["Q12"][0], "Q2": s64["Q22"][0], "phi": s64["phi"]}
    out = {}
    for name, (fld, ch) in FIELDS.items():
        out[name] = (rel_l2(gpu[fld][..., ch], f64[fld][..., ch]), rel_l2(f32[fld][..., ch], f64[fld][..., ch]))
    th = lambda V: theta_pert(V, m._B, m.nz)  # noqa: E731
    rms = lambda d: float(np.sqrt(np.mean(d * d)))  # noqa: E731
    out["theta_rms_K"] = (rms(th(gpu["V"]) - th(f64["V"])), rms(th(f32["V"]) - th(f64["V"])))
    out["phi"] = (rel_l2(gpu["phi"], f64["phi"]), rel_l2(f32["phi"], f64["phi"]))
    o = Oracle(U, m._B, "f64", threads=threads)
    pd = m.pad
    cut = np.s_[:, pd:pd + m.ny, pd:pd + m.nx]
    d0 = np.linalg.norm(rhs64[cut])
    out["divergence"] = tuple(float(np.linalg.norm(o.div(np.asarray(V, np.float64))[cut]) / d0)
                              for V in (gpu["V"], f32["V"], f64["V"]))
    return out


def fp64_report(m, dt, forms=("bf16x6", "fp32"), threads=None):
    """Error against FP64 of one GPU step per solver form, all from model m's current state (restored between the
    forms), at any size: the FP64 step runs the oracle's FP64 kernels around an FP64 scipy solve of the same rhs
    (the dense FP64 oracle solve costs O(N^3) per level). Returns {form: {field: rel L2 vs FP64, "theta_rms_K",
    "phi", "divergence"}} with "divergence" the post-projection divergence relative to the divergence removed, and
    the FP64 step's own divergence under "fp64"."""
    from _oracle import step_f64_split
    from _ref_solve import solve_scipy

    p = m.parity
    names = ["V0", "V1", "Q10", "Q11", "Q20", "Q21", "Vt", "Q1t", "Q2t", "aux", "phT", "lsfA", "lsfB", "ciMap", "rhs",
             "phi"]
    fields = [m.V[0], m.V[1], m.Q1[0], m.Q1[1], m.Q2[0], m.Q2[1], m.Vt, m.Q1t, m.Q2t, m.aux, m.phT, m.lsfA, m.lsfB,
              m.ciMap, m.rhs, m.phi]
    saved = [f.download() for f in fields]
    run_lsf = m.lsf_due()
    clock = (m.time, m.steps, m.parity, m.lsfT, m.phiScale)
    gpu = {}
    for form in forms:
        for f, a in zip(fields, saved):
            f.upload(a)
        m.time, m.steps, m.parity, m.lsfT, m.phiScale = clock
        m.set_solver_mode(form, "mixed")
        m.step(dt)
        m.sync()
        gpu[form] = {"V": m.V[1 - p].download(), "Q1": m.Q1[0].download(), "Q2": m.Q2[0].download(),
                     "phi": m.phi.download()}
    U, ci = m._u_last, (m._ci if m.ciActive else None)
    d = dict(zip(names, saved))
    c = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
    s64 = dict(V2=c(np.stack([d["V0"], d["V1"]])), Q12=c(np.stack([d["Q10"], d["Q11"]])),
               Q22=c(np.stack([d["Q20"], d["Q21"]])), Vt=c(d["Vt"]), Q1t=c(d["Q1t"]), Q2t=c(d["Q2t"]), aux=c(d["aux"]),
               phT=c(d["phT"]), lsfA=c(d["lsfA"]), lsfB=c(d["lsfB"]), ciMap=c(d["ciMap"]), rhs=c(d["rhs"]),
               phi=c(d["phi"]))
    del saved, d
    step_f64_split(U, m._B, m._bd, ci, run_lsf, p, **s64, part=0, threads=threads)
    o = Oracle(U, m._B, "f64", threads=threads)
    rc, rw = o.densities()
    s64["phi"][...] = solve_scipy(s64["rhs"], float(U[3]), float(U[4]), rc, rw)
    step_f64_split(U, m._B, m._bd, ci, run_lsf, p, **s64, part=1, threads=threads)
    f64 = {"V": s64["V2"][1 - p], "Q1": s64["Q12"][0], "Q2": s64["Q22"][0], "phi": s64["phi"]}
    pd = m.pad
    cut = np.s_[:, pd:pd + m.ny, pd:pd + m.nx]
    d0 = np.linalg.norm(s64["rhs"][cut])
    div = lambda V: float(np.linalg.norm(o.div(np.asarray(V, np.float64))[cut]) / d0)  # noqa: E731
    th = lambda V: theta_pert(V, m._B, m.nz)  # noqa: E731
    out = {}
    for form, g in gpu.items():
        e = {name: rel_l2(g[fld][..., ch], f64[fld][..., ch]) for name, (fld, ch) in FIELDS.items()}
        dth = th(g["V"]) - th(f64["V"])
        e["theta_rms_K"] = float(np.sqrt(np.mean(dth * dth)))
        e["phi"] = rel_l2(g["phi"], f64["phi"])
        e["divergence"] = div(g["V"])
        out[form] = e
    out["fp64"] = {"divergence": div(f64["V"])}
    return out
