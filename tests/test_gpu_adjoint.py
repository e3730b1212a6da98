"""GPU adjoint (VJP) tests: hand-written adjoint kernels against the oracle.

The reference differentiates with torch autograd.  The adjoint of every
kernel and of the whole large step is checked against float64 central finite
differences of the CPU oracle (a restatement of the reference forward that
matches it bit for bit) through the dot-product identity
<ybar, J t> = <J^T ybar, t>, for random cotangents ybar and tangents t.
Inputs are perturbed with noise first so that no value sits on a switching
surface (an upwind sign, a limiter or min/max selection), where the hard
primal has no derivative.
"""

from __future__ import annotations

import zlib

import numpy as np
import pytest
import torch

from conftest import perturbed_state, srk3_case

pytestmark = pytest.mark.gpu

CASE = "l2_L20_tracer"


def _noisy_state(seed: int = 0, case: str = CASE):
    st, _, meta = srk3_case(case)
    return perturbed_state(st, seed), meta


def _len_disp(st, meta):
    return float(st["nominalMinDc"]) if "nominalMinDc" in st else float(meta["len_disp"])


def _cfg(st, meta):
    from mpas_dycore_pytorch_amd.time_integration import resolve_config

    c = {"config_time_integration_order": 3, "config_len_disp": _len_disp(st, meta)}
    return resolve_config(c, st, st, meta["dt"], int(np.asarray(st["scalars"]).shape[2]))


def _capture_oracle_calls(st, meta):
    from oracle import oracle as O, srk3 as OS

    names = ["moist_coefficients", "vert_imp_coefs", "solve_diagnostics", "dyn_tend", "smlstep_pert",
             "acoustic_step", "divergence_damping", "recover_large_step", "advance_scalars",
             "advance_scalars_mono", "reconstruct"]
    calls, orig = [], {}
    for n in names:
        f = getattr(O, n)
        orig[n] = f

        def g(*a, _f=f, _n=n, **k):
            r = _f(*a, **k)
            calls.append((_n, a, k, r))
            return r

        setattr(O, n, g)
    try:
        om = O.Mesh(st)
        OS.srk3(st, om, meta["dt"], len_disp=_len_disp(st, meta))
    finally:
        for n in names:
            setattr(O, n, orig[n])
    return om, calls


def _dot(a: dict, b: dict) -> float:
    return float(sum(np.sum(np.asarray(a[k], dtype=np.float64) * np.asarray(b[k], dtype=np.float64)) for k in a))


def _tangents(rng, X: dict, rows: dict) -> dict:
    t = {}
    for k, v in X.items():
        v = np.asarray(v, dtype=np.float64)
        s = float(np.std(v)) or 1.0
        tt = rng.standard_normal(v.shape) * s
        if v.ndim >= 1 and v.shape[0] in rows:
            tt[-1] = 0.0
        t[k] = tt
    return t


def _fd(fn, X: dict, t: dict, ybar: dict, h: float) -> float:
    def J(sgn):
        X2 = {k: np.asarray(v, dtype=np.float64) + sgn * h * t[k] for k, v in X.items()}
        Y = fn(X2)
        return sum(float(np.sum(ybar[k] * np.asarray(Y[k], dtype=np.float64))) for k in ybar)

    return (J(1.0) - J(-1.0)) / (2.0 * h)


def _check_identity(name, fn_oracle, gpu_op, X: dict, out_names, rng, rows, h=1e-7, tol=2e-6):
    """<ybar, J t> (oracle FD) against <J^T ybar, t> (GPU adjoint of one kernel call)."""
    from mpas_dycore_pytorch_amd.adjoint import Tape, vjp_scratch

    Y = fn_oracle(X)
    ybar = {}
    for k in out_names:
        y = np.asarray(Y[k], dtype=np.float64)
        s = float(np.std(y)) or 1.0
        yb = rng.standard_normal(y.shape) / s
        if y.ndim >= 1 and y.shape[0] in rows:
            yb[-1] = 0.0
        ybar[k] = yb
    t = _tangents(rng, X, rows)
    # the better of two steps: a step can cross a limiter switch of the hard primal
    fds = [_fd(fn_oracle, X, t, ybar, hh) for hh in (h, 0.1 * h)]
    dm, tape, ins, outs = gpu_op(X)
    seeds = {id(outs[k]): torch.as_tensor(ybar[k], device=dm.device) for k in out_names}
    bars = tape.backward(seeds, vjp_scratch(dm, 4))
    torch.cuda.synchronize()
    ad = 0.0
    for k, tin in ins.items():
        b = bars.get(id(tin))
        if b is not None:
            ad += float(np.sum(b.cpu().numpy() * t[k]))
    rel = min(abs(fd - ad) / max(abs(fd), abs(ad), 1e-300) for fd in fds)
    print(f"[adjoint] {name}: fd={fds} adjoint={ad:.12e} rel={rel:.2e}", flush=True)
    assert rel < tol, (name, fds, ad, rel)


@pytest.fixture(scope="module")
def case():
    st, meta = _noisy_state(0)
    om, calls = _capture_oracle_calls(st, meta)
    from mpas_dycore_pytorch_amd.device import DeviceMesh

    dm = DeviceMesh(st, torch.device("cuda"))
    rows = {om.nc + 1, om.ne + 1, om.nv + 1}
    return st, meta, om, calls, dm, rows


def _first(calls, name, idx=0):
    return [c for c in calls if c[0] == name][idx]


def _gpu(x, dm):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(x, dtype=np.float64)), device=dm.device)


@pytest.mark.parametrize("idx", [0, 2])
def test_dyn_tend_adjoint(case, idx):
    from oracle import oracle as O
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd.adjoint import Tape

    st, meta, om, calls, dm, rows = case
    _, a, _, r = _first(calls, "dyn_tend", idx)
    params, f, carried = a[1], a[2], (a[3] if len(a) > 3 else None)
    rng = np.random.default_rng(10 + idx)
    X = {n: np.array(f[n], dtype=np.float64) for n in K.DYN_INPUTS}
    for n in ("u", "v", "ru"):
        X[n] = X[n] + 2.0 * rng.standard_normal(X[n].shape)
    X["pp"] = X["pp"] + 10.0 * rng.standard_normal(X["pp"].shape)
    X["rw"] = X["rw"] + 0.01 * rng.standard_normal(X["rw"].shape)
    for n in K.DYN_CARRIED:
        X[n + "_in"] = np.array(carried[n]) if carried is not None else np.zeros_like(np.asarray(r[n]))
    rk = params.rk_step

    def fn(X):
        cc = {n: X[n + "_in"] for n in K.DYN_CARRIED} if rk != 1 else None
        return O.dyn_tend(om, params, {n: X[n] for n in K.DYN_INPUTS}, cc)

    def gpu_op(X):
        tape = Tape(dm)
        ins = {k: _gpu(v, dm) for k, v in X.items()}
        p = K.dyn_params(rk, params.dt, 0.0)
        for fld, _ in type(params)._fields_:
            setattr(p, fld, getattr(params, fld))
        cin = {n: ins[n + "_in"] for n in K.DYN_CARRIED}
        outs = K.dyn_tend(dm, p, {n: ins[n] for n in K.DYN_INPUTS}, cin if rk != 1 else None)
        tape.kernel("dyn_tend", (p,), ins, outs)
        return dm, tape, ins, outs

    _check_identity(f"dyn_tend rk{rk}", fn, gpu_op, X, K.DYN_CARRIED + K.DYN_OUTPUTS, rng, rows)


def _vjp_twice(dm, op, rng):
    """Input adjoints of one recorded op, computed twice from the same output seeds."""
    import ctypes as C

    from mpas_dycore_pytorch_amd import _native as N
    from mpas_dycore_pytorch_amd.adjoint import vjp_scratch

    lib = N.load()
    seeds = []
    for o in op.outs:
        sd = torch.as_tensor(rng.standard_normal(tuple(o.shape)), device=dm.device)
        sd[-1] = 0.0
        seeds.append(sd)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    inv = C.byref(dm.inverse_tables())
    res = []
    for _ in range(2):
        bars = [torch.zeros_like(x) for x in op.ins] + [sd.clone() for sd in seeds]
        n = len(bars)
        fwd = (C.c_void_p * n)(*[x.data_ptr() for x in op.ins + op.outs])
        bar = (C.c_void_p * n)(*[b.data_ptr() for b in bars])
        # scratch from the allocator as it comes: every element a pass reads is written first
        scratch = vjp_scratch(dm, 4)
        sp = C.c_void_p(scratch.data_ptr())
        if op.kind == "solve_diagnostics":
            rc = lib.mpas_hip_solve_diagnostics_vjp(dm.ref, *op.args, fwd, bar, sp, inv, stream)
        else:
            rc = lib.mpas_hip_dyn_tend_vjp(dm.ref, C.byref(op.args[0]), fwd, bar, sp, inv, stream)
        N.check(rc, op.kind + "_vjp")
        torch.cuda.synchronize()
        res.append([b.cpu().numpy() for b in bars[: len(op.ins)]])
        del scratch
    return res


@pytest.mark.parametrize("kind,idx", [("solve_diagnostics", 1), ("solve_diagnostics", 2),
                                      ("solve_diagnostics", 3), ("dyn_tend", 0), ("dyn_tend", 1)])
def test_op_adjoint_is_bitwise_reproducible(case, kind, idx):
    """The diagnostics and dynamics-tendency adjoints gather every transpose in
    slot order (no atomics): two runs from the same seeds agree bit for bit."""
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd.adjoint import Tape

    st, meta, om, calls, dm, rows = case
    _, a, _, r = _first(calls, kind, idx)
    tape = Tape(dm)
    if kind == "solve_diagnostics":
        dt, apvm, rk = float(a[1]), float(a[2]), int(a[3])
        ins = {"u": _gpu(a[4], dm), "h": _gpu(a[5], dm), "v_in": _gpu(a[6], dm)}
        outs = K.solve_diagnostics(dm, dt, apvm, rk, ins["u"], ins["h"], ins["v_in"])
        tape.kernel(kind, (dt, apvm, rk), ins, outs)
    else:
        params, f, carried = a[1], a[2], (a[3] if len(a) > 3 else None)
        rk = params.rk_step
        ins = {n: _gpu(f[n], dm) for n in K.DYN_INPUTS}
        for n in K.DYN_CARRIED:
            ins[n + "_in"] = _gpu(carried[n] if carried is not None else np.zeros_like(np.asarray(r[n])), dm)
        p = K.dyn_params(rk, params.dt, 0.0)
        for fld, _ in type(params)._fields_:
            setattr(p, fld, getattr(params, fld))
        cin = {n: ins[n + "_in"] for n in K.DYN_CARRIED}
        outs = K.dyn_tend(dm, p, {n: ins[n] for n in K.DYN_INPUTS}, cin if rk != 1 else None)
        tape.kernel(kind, (p,), ins, outs)
    print(f"[reproducible] {kind} call {idx} rk={rk}", flush=True)
    first, second = _vjp_twice(dm, tape.ops[0], np.random.default_rng(100 + idx))
    for x, y in zip(first, second):
        assert np.array_equal(x, y), (kind, idx, float(np.max(np.abs(x - y))))
        assert np.all(np.isfinite(x)), (kind, idx)


@pytest.mark.parametrize("idx", [0, 3])
def test_acoustic_adjoint(case, idx):
    from oracle import oracle as O
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd.adjoint import Tape

    st, meta, om, calls, dm, rows = case
    _, a, _, _ = _first(calls, "acoustic_step", idx)
    ap, f = a[1], a[2]
    rng = np.random.default_rng(20 + idx)
    X = {n: np.array(f[n], dtype=np.float64) for n in K.ACOUSTIC_IN if n not in ("cofrz", "cofwr")}
    for n in K.ACOUSTIC_IO:
        X[n + "_in"] = np.array(f[n], dtype=np.float64)
    const = {"cofrz": np.array(f["cofrz"]), "cofwr": np.array(f["cofwr"])}

    def fn(X):
        ff = dict(const)
        ff.update({n: X[n] for n in K.ACOUSTIC_IN if n in X})
        ff.update({n: X[n + "_in"] for n in K.ACOUSTIC_IO})
        return O.acoustic_step(om, ap, ff)

    def gpu_op(X):
        tape = Tape(dm)
        ins = {k: _gpu(v, dm) for k, v in X.items()}
        allin = dict(ins)
        allin.update({k: _gpu(v, dm) for k, v in const.items()})
        p = K.acoustic_params(ap.small_step, ap.dts, ap.epssm)
        ff = {n: allin[n] for n in K.ACOUSTIC_IN}
        ff.update({n: allin[n + "_in"] for n in K.ACOUSTIC_IO})
        outs = K.acoustic_step(dm, p, ff)
        tape.kernel("acoustic_step", (p,), allin, outs)
        return dm, tape, ins, outs

    _check_identity(f"acoustic small_step={ap.small_step}", fn, gpu_op, X, K.ACOUSTIC_IO, rng, rows)


@pytest.mark.parametrize("which", ["vimp", "diag", "recover3", "moist", "mono", "scal", "divdamp", "recon"])
def test_small_kernel_adjoints(case, which):
    from oracle import oracle as O
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd.adjoint import Tape

    st, meta, om, calls, dm, rows = case
    rng = np.random.default_rng(zlib.crc32(which.encode()))
    tape = Tape(dm)
    if which == "vimp":
        _, a, _, _ = _first(calls, "vert_imp_coefs", 1)
        dts, eps = a[1], a[2]
        X = dict(zip(("cqw", "p", "t", "rt", "qtot"), (np.array(x, dtype=np.float64) for x in a[3:8])))
        fn = lambda X: O.vert_imp_coefs(om, dts, eps, *X.values())  # noqa: E731

        def gpu_op(X):
            ins = {k: _gpu(v, dm) for k, v in X.items()}
            outs = K.vert_imp_coefs(dm, dts, eps, *ins.values())
            tape.kernel("vert_imp_coefs", (float(dts), float(eps)), ins, outs)
            return dm, tape, ins, outs
        outn = K.VIMP_OUT[2:]
    elif which == "diag":
        _, a, _, _ = _first(calls, "solve_diagnostics", 3)
        dt, apvm, rk = a[1], a[2], a[3]
        X = {"u": np.array(a[4]), "h": np.array(a[5]), "v_in": np.array(a[6])}
        fn = lambda X: O.solve_diagnostics(om, dt, apvm, rk, X["u"], X["h"], X["v_in"])  # noqa: E731

        def gpu_op(X):
            ins = {k: _gpu(v, dm) for k, v in X.items()}
            outs = K.solve_diagnostics(dm, dt, apvm, rk, ins["u"], ins["h"], ins["v_in"])
            tape.kernel("solve_diagnostics", (float(dt), float(apvm), int(rk)), ins, outs)
            return dm, tape, ins, outs
        outn = K.DIAG_OUT
    elif which == "recover3":
        _, a, k, _ = _first(calls, "recover_large_step", 2)
        rk, dt, ns, f = a[1], a[2], a[3], a[4]
        X = {n: np.array(f[n], dtype=np.float64) for n in K.RECOVER_IN}
        X["rt_diabatic_tend"] = X["rt_diabatic_tend"] + 1e-4 * rng.standard_normal(X["rt_diabatic_tend"].shape)
        X.update({"wwAvg_in": np.array(f["wwAvg"]), "ruAvg_in": np.array(f["ruAvg"]),
                  "exner_in": np.array(k["exner_in"]), "pressure_p_in": np.array(k["pressure_p_in"])})

        def fn(X):
            ff = {n: X[n] for n in K.RECOVER_IN}
            ff.update({"wwAvg": X["wwAvg_in"], "ruAvg": X["ruAvg_in"]})
            return O.recover_large_step(om, rk, dt, ns, ff, exner_in=X["exner_in"], pressure_p_in=X["pressure_p_in"])

        def gpu_op(X):
            ins = {k2: _gpu(v, dm) for k2, v in X.items()}
            ff = {n: ins[n] for n in K.RECOVER_IN}
            ff.update({"wwAvg": ins["wwAvg_in"], "ruAvg": ins["ruAvg_in"]})
            outs = K.recover_large_step(dm, rk, dt, ns, ff, exner_in=ins["exner_in"],
                                        pressure_p_in=ins["pressure_p_in"])
            tape.kernel("recover_large_step", (int(rk), float(dt), int(ns)), ins, outs)
            return dm, tape, ins, outs
        outn = ("wwAvg", "ruAvg", "rho_p", "rho_zz", "rtheta_p", "theta_m", "exner", "pressure_p", "rw", "w", "ru",
                "u")
    elif which == "moist":
        _, a, _, _ = _first(calls, "moist_coefficients")
        m0, m1 = a[2], a[3]
        X = {"scalars": np.array(a[1], dtype=np.float64)}
        S = X["scalars"].shape[2]
        fn = lambda X: O.moist_coefficients(om, X["scalars"], m0, m1)  # noqa: E731

        def gpu_op(X):
            ins = {"scalars": _gpu(X["scalars"], dm)}
            outs = K.moist_coefficients(dm, ins["scalars"], m0, m1)
            tape.kernel("moist_coefficients", (S, int(m0), int(m1)), ins, outs)
            return dm, tape, ins, outs
        outn = ("cqw", "cqu", "qtot")
    elif which == "mono":
        _, a, _, _ = _first(calls, "advance_scalars_mono")
        dt, c3 = a[1], a[2]
        names = ("scalars_old", "scalars_new_in", "rho_zz_old", "uhAvg", "wwAvg")
        X = {n: np.array(v, dtype=np.float64) for n, v in zip(names, a[3:8])}
        X["uhAvg"] = X["uhAvg"] + 5.0 * rng.standard_normal(X["uhAvg"].shape)
        X["wwAvg"] = X["wwAvg"] + 0.05 * rng.standard_normal(X["wwAvg"].shape)
        X["scalars_old"] = X["scalars_old"] + 0.1 * rng.random(X["scalars_old"].shape)
        X["scalars_new_in"] = X["scalars_new_in"] + 0.1 * rng.random(X["scalars_new_in"].shape)
        S = X["scalars_old"].shape[2]

        def fn(X):
            out, rzi = O.advance_scalars_mono(om, dt, c3, *[X[n] for n in names])
            return {"scalars_new_out": out, "rho_zz_int": rzi}

        def gpu_op(X):
            ins = {k: _gpu(v, dm) for k, v in X.items()}
            out, rzi = K.advance_scalars_mono(dm, dt, c3, *[ins[n] for n in names])
            outs = {"scalars_new_out": out, "rho_zz_int": rzi}
            tape.kernel("advance_scalars_mono", (S, float(dt), float(c3)), ins, outs)
            return dm, tape, ins, outs
        outn = ("scalars_new_out", "rho_zz_int")
    elif which == "scal":
        _, a, _, _ = _first(calls, "advance_scalars", 1)
        dt, c3, wo, wn = a[1], a[2], a[3], a[4]
        names = ("scalar_old", "scalar_new_in", "rho_zz_old", "rho_zz_new", "uhAvg", "wwAvg")
        X = {n: np.array(v, dtype=np.float64) for n, v in zip(names, a[5:11])}
        X["uhAvg"] = X["uhAvg"] + 5.0 * rng.standard_normal(X["uhAvg"].shape)
        S = X["scalar_old"].shape[2]

        def fn(X):
            out, hf = O.advance_scalars(om, dt, c3, wo, wn, *[X[n] for n in names])
            return {"scalar_new_out": out, "horiz_flux": hf}

        def gpu_op(X):
            ins = {k: _gpu(v, dm) for k, v in X.items()}
            out, hf = K.advance_scalars(dm, dt, c3, wo, wn, *[ins[n] for n in names])
            outs = {"scalar_new_out": out, "horiz_flux": hf}
            tape.kernel("advance_scalars", (S, float(dt), float(c3), float(wo), float(wn)), ins, outs)
            return dm, tape, ins, outs
        outn = ("scalar_new_out",)
    elif which == "divdamp":
        _, a, _, _ = _first(calls, "divergence_damping", 3)
        dts, smdiv, ld = a[1], a[2], a[3]
        X = dict(zip(("theta_m", "rtheta_pp", "rtheta_pp_old", "ru_p_in"), (np.array(x) for x in a[4:8])))
        fn = lambda X: {"ru_p_out": O.divergence_damping(om, dts, smdiv, ld, *X.values())}  # noqa: E731

        def gpu_op(X):
            ins = {k: _gpu(v, dm) for k, v in X.items()}
            out = K.divergence_damping(dm, dts, smdiv, ld, *ins.values())
            outs = {"ru_p_out": out}
            tape.kernel("divergence_damping", (float(dts), float(smdiv), float(ld)), ins, outs)
            return dm, tape, ins, outs
        outn = ("ru_p_out",)
    else:
        _, a, _, _ = _first(calls, "reconstruct")
        X = {"u": np.array(a[1], dtype=np.float64)}
        fn = lambda X: O.reconstruct(om, X["u"])  # noqa: E731

        def gpu_op(X):
            ins = {"u": _gpu(X["u"], dm)}
            outs = K.reconstruct(dm, ins["u"])
            tape.kernel("reconstruct", (), ins, outs)
            return dm, tape, ins, outs
        outn = K.RECON_OUT
    _check_identity(which, fn, gpu_op, X, outn, rng, rows)


def _step_inputs(st):
    from mpas_dycore_pytorch_amd.adjoint import STEP_INPUTS

    out = {}
    for k in STEP_INPUTS:
        v = st.get(k)
        if v is None and k == "rho_edge":
            v = st.get("h_edge")
        out[k] = v
    return out


def test_taped_step_matches_fused_step():
    """The per-function chain the adjoint differentiates is the forward the fused step runs."""
    from mpas_dycore_pytorch_amd.adjoint import STEP_INPUTS, Tape, taped_step
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    st, meta = _noisy_state(1)
    cfg = _cfg(st, meta)
    stepper = Srk3Stepper(st, st, {"config_time_integration_order": 3, "config_len_disp": _len_disp(st, meta)},
                          meta["dt"])
    dm = stepper.mesh
    state = {k: stepper.t[k].clone() for k in STEP_INPUTS}
    stepper.step(1)
    fused = stepper.state_dict()
    taped = taped_step(dm, state, cfg, meta["dt"], Tape(dm))
    torch.cuda.synchronize()
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "ru", "rw", "rho_p", "rtheta_p",
              "pv_edge", "ke", "divergence", "vorticity", "ruAvg", "wwAvg"):
        a, b = fused[k].cpu().numpy(), taped[k].cpu().numpy()
        scale = max(float(np.max(np.abs(a))), 1e-300)
        err = float(np.max(np.abs(a - b))) / scale
        print(f"[adjoint] taped vs fused {k}: {err:.2e}", flush=True)
        assert err <= 1e-13, k


@pytest.mark.parametrize("case", [CASE, "l2_L20_terrain"])
def test_step_vjp_against_oracle_finite_differences(case):
    """Whole large step: <ybar, J t> by oracle central differences vs the GPU VJP."""
    from oracle import oracle as O, srk3 as OS
    from mpas_dycore_pytorch_amd.adjoint import STEP_INPUTS, step_vjp
    from mpas_dycore_pytorch_amd.device import DeviceMesh

    st, meta = _noisy_state(2, case)
    cfg = _cfg(st, meta)
    dm = DeviceMesh(st, torch.device("cuda"))
    om = O.Mesh(st)
    rng = np.random.default_rng(3)
    rows = {om.nc + 1, om.ne + 1, om.nv + 1}
    wrt = ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "rtheta_p", "rho_p", "ru", "rw", "pressure_p")
    outk = ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "rtheta_p")
    nc1, L = int(st["nCells"]) + 1, int(st["nVertLevels"])
    X = {k: np.array(st[k] if k in st else np.zeros((nc1, L)), dtype=np.float64) for k in STEP_INPUTS}
    t = _tangents(rng, {k: X[k] for k in wrt}, rows)
    t["w"][:, 0] = 0.0
    t["w"][:, -1] = 0.0
    t["rw"][:, 0] = 0.0
    t["rw"][:, -1] = 0.0
    ld = _len_disp(st, meta)

    def fn(Xp):
        s2 = dict(st)
        s2.update(Xp)
        return OS.srk3(s2, om, meta["dt"], len_disp=ld)

    Y = fn({k: X[k] for k in wrt})
    ybar = {}
    for k in outk:
        y = np.asarray(Y[k], dtype=np.float64)
        yb = rng.standard_normal(y.shape) / (float(np.std(y)) or 1.0)
        yb[-1] = 0.0
        ybar[k] = yb
    fds = [_fd(fn, {k: X[k] for k in wrt}, t, ybar, h) for h in (1e-7, 1e-8)]
    state = {k: _gpu(X[k], dm) for k in STEP_INPUTS}
    _, grads = step_vjp(dm, state, cfg, meta["dt"], {k: _gpu(v, dm) for k, v in ybar.items()})
    torch.cuda.synchronize()
    ad = sum(float(np.sum(grads[k].cpu().numpy() * t[k])) for k in wrt)
    rels = [abs(fd - ad) / max(abs(fd), abs(ad)) for fd in fds]
    print(f"[adjoint] whole step: fd={fds} adjoint={ad:.12e} rel={rels}", flush=True)
    assert min(rels) < 1e-5


def test_run_simulation_vjp_two_steps():
    """run_simulation.vjp through two checkpointed large steps vs oracle finite differences."""
    from oracle import oracle as O, srk3 as OS
    from mpas_dycore_pytorch_amd.atmosphere_model import run_simulation
    from mpas_dycore_pytorch_amd.boundary import mark_complete_step_boundary

    st, meta = _noisy_state(4)
    om = O.Mesh(st)
    ld = _len_disp(st, meta)
    config = {"config_time_integration_order": 3, "config_len_disp": ld}
    dev = torch.device("cuda")
    state = {k: (torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)
                 if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v) for k, v in st.items()}
    mark_complete_step_boundary(state, mesh=st, config=config, dt=meta["dt"])
    nc = int(st["nCells"])
    area = torch.as_tensor(1.0 / np.asarray(st["invAreaCell"], dtype=np.float64)[:nc], device=dev)
    theta_ref = state["theta_m"].detach().clone()

    def objective(out):
        d = out["theta_m"][:nc] - theta_ref[:nc]
        return torch.sum(area[:, None] * d * d)

    grads = run_simulation.vjp(state, st, objective, track_keys=("u", "theta_m", "w"), config=config,
                               dt=meta["dt"], nsteps=2, device=dev)
    rng = np.random.default_rng(5)
    rows = {om.nc + 1, om.ne + 1, om.nv + 1}
    X = {k: np.array(st[k], dtype=np.float64) for k in ("u", "theta_m", "w")}
    t = _tangents(rng, X, rows)
    t["w"][:, 0] = 0.0
    t["w"][:, -1] = 0.0
    area_np = area.cpu().numpy()
    th0 = theta_ref.cpu().numpy()

    def loss(Xp):
        s2 = dict(st)
        s2.update(Xp)
        s2 = OS.srk3(s2, om, meta["dt"], len_disp=ld)
        s2 = OS.srk3(s2, om, meta["dt"], len_disp=ld)
        d = s2["theta_m"][:nc] - th0[:nc]
        return float(np.sum(area_np[:, None] * d * d))

    fds = []
    for h in (1e-7, 1e-8):
        Xp = {k: X[k] + h * t[k] for k in X}
        Xm = {k: X[k] - h * t[k] for k in X}
        fds.append((loss(Xp) - loss(Xm)) / (2 * h))
    ad = sum(float(np.sum(grads[k].cpu().numpy() * t[k])) for k in X)
    rels = [abs(fd - ad) / max(abs(fd), abs(ad)) for fd in fds]
    print(f"[adjoint] run_simulation.vjp 2 steps: fd={fds} adjoint={ad:.12e} rel={rels}", flush=True)
    assert min(rels) < 1e-5


def test_run_simulation_vjp_is_bitwise_reproducible():
    """Two run_simulation.vjp calls on the same state give the same gradients
    bit for bit (no adjoint kernel uses atomics): two large steps with a
    tracer under monotonic transport, every tracked key."""
    from mpas_dycore_pytorch_amd.atmosphere_model import run_simulation
    from mpas_dycore_pytorch_amd.boundary import mark_complete_step_boundary

    st, meta = _noisy_state(4)
    config = {"config_time_integration_order": 3, "config_len_disp": _len_disp(st, meta)}
    dev = torch.device("cuda")
    state = {k: (torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)
                 if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v) for k, v in st.items()}
    mark_complete_step_boundary(state, mesh=st, config=config, dt=meta["dt"])
    nc = int(st["nCells"])
    area = torch.as_tensor(1.0 / np.asarray(st["invAreaCell"], dtype=np.float64)[:nc], device=dev)
    theta_ref = state["theta_m"].detach().clone()

    def objective(out):
        d = out["theta_m"][:nc] - theta_ref[:nc]
        return torch.sum(area[:, None] * d * d) + torch.sum(out["scalars"][:nc] ** 2)

    keys = ("u", "w", "theta_m", "rho_zz", "scalars")
    runs = [run_simulation.vjp(state, st, objective, track_keys=keys, config=config, dt=meta["dt"], nsteps=2,
                               device=dev) for _ in range(2)]
    for k in keys:
        a, b = runs[0][k], runs[1][k]
        print(f"[reproducible] run_simulation.vjp d/d{k}: max|g| {float(a.abs().max()):.3e}", flush=True)
        assert bool(torch.isfinite(a).all()), k
        assert torch.equal(a, b), (k, float((a - b).abs().max()))


def test_kept_tape_matches_recompute():
    """The step that records its tape in the forward (the last step, or every
    step with checkpoint_timesteps=False) returns the fused step's state bit
    for bit, and its gradients equal the recomputing backward's."""
    from mpas_dycore_pytorch_amd.atmosphere_model import run_simulation
    from mpas_dycore_pytorch_amd.boundary import mark_complete_step_boundary

    st, meta = _noisy_state(6)
    config = {"config_time_integration_order": 3, "config_len_disp": _len_disp(st, meta)}
    dev = torch.device("cuda")
    state = {k: (torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)
                 if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v) for k, v in st.items()}
    mark_complete_step_boundary(state, mesh=st, config=config, dt=meta["dt"])
    with torch.no_grad():  # no backward can follow: every step runs fused
        plain = run_simulation(dict(state), st, config=config, dt=meta["dt"], nsteps=2, device=dev,
                               differentiable=True)
    diff_in = dict(state)
    diff_in["u"] = state["u"].detach().clone().requires_grad_(True)
    diff = run_simulation(diff_in, st, config=config, dt=meta["dt"], nsteps=2, device=dev, differentiable=True)
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "ru", "rw", "rho_p", "rtheta_p",
              "v", "pv_edge", "ke", "divergence", "vorticity", "h_edge", "uReconstructZonal", "ruAvg", "wwAvg"):
        assert torch.equal(plain[k], diff[k].detach()), k
    nc = int(st["nCells"])
    w8 = torch.as_tensor(np.random.default_rng(8).standard_normal((nc, int(st["nVertLevels"]))), device=dev)

    def objective(out):
        return torch.sum(w8 * out["theta_m"][:nc]) + torch.sum(out["u"][:-1] ** 2)

    keys = ("u", "theta_m", "w", "rho_zz")
    g_ck = run_simulation.vjp(dict(state), st, objective, track_keys=keys, config=config, dt=meta["dt"],
                              nsteps=2, device=dev)
    g_all = run_simulation.vjp(dict(state), st, objective, track_keys=keys, config=config, dt=meta["dt"],
                               nsteps=2, device=dev, checkpoint_timesteps=False)
    for k in keys:
        # the recomputed tape is bitwise the kept one and every adjoint kernel is atomic-free,
        # so the two gradients are the same bits
        same = torch.equal(g_ck[k], g_all[k])
        print(f"[adjoint] kept tape vs recompute d/d{k}: bitwise equal {same}", flush=True)
        assert same, k


@pytest.mark.parametrize("perturbed", [False, True])
def test_vjp_matches_reference_gradients(perturbed):
    """run_simulation.vjp (two checkpointed steps, hand-written adjoints) against
    torch autograd of the reference (golden vectors of make_golden_vjp.py), to
    1e-9 relative per gradient field.  The unperturbed state exercises the tie
    conventions (the dry water-vapour slot makes every limiter bound tie)."""
    from conftest import load_npz
    from mpas_dycore_pytorch_amd.atmosphere_model import run_simulation
    from mpas_dycore_pytorch_amd.boundary import mark_complete_step_boundary

    z, meta = load_npz(f"vjp_{CASE}{'_perturbed' if perturbed else ''}.npz")
    st, _, smeta = srk3_case(CASE)
    if perturbed:
        st = perturbed_state(st, meta["perturb_seed"])
    nc, ne, L = int(st["nCells"]), int(st["nEdges"]), int(st["nVertLevels"])
    dev = torch.device("cuda")
    state = {k: (torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)
                 if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v) for k, v in st.items()}
    state.setdefault("rt_diabatic_tend", torch.zeros((nc + 1, L), dtype=torch.float64, device=dev))
    area = torch.as_tensor(1.0 / np.asarray(st["invAreaCell"], dtype=np.float64)[:nc], device=dev)
    ae = torch.as_tensor((np.asarray(st["dcEdge"], dtype=np.float64) * np.asarray(st["dvEdge"]))[:ne], device=dev)
    th0 = torch.as_tensor(np.asarray(st["theta_m"], dtype=np.float64), device=dev)
    consts = [float(c) for c in z["loss_constants"]]

    def objective(out):
        d = out["theta_m"][:nc] - th0[:nc]
        terms = [torch.sum(area[:, None] * d * d), torch.sum(ae[:, None] * out["u"][:ne] * out["u"][:ne]),
                 torch.sum(area[:, None] * out["scalars"][:nc, :, 1] * out["scalars"][:nc, :, 1])]
        return sum(t / c for t, c in zip(terms, consts))

    keys = tuple(meta["tracked"])
    config = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    # the recorded gradients are those of atm_srk3 from the prepared state: bind it as a boundary
    mark_complete_step_boundary(state, mesh=st, config=config, dt=smeta["dt"])
    grads = run_simulation.vjp(state, st, objective, track_keys=keys, config=config, dt=smeta["dt"],
                               nsteps=meta["steps"], device=dev)
    torch.cuda.synchronize()
    worst = 0.0
    for k in keys:
        ref = z["grad__" + k]
        got = grads[k].cpu().numpy() if k in grads else np.zeros_like(ref)
        scale = float(np.max(np.abs(ref)))
        err = float(np.max(np.abs(got[: ref.shape[0] - 1] - ref[: ref.shape[0] - 1]))) / scale
        print(f"[adjoint] reference gradient {'perturbed' if perturbed else 'as prepared'} {k}: "
              f"max|g| {scale:.3e} rel err {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst <= 1e-9, worst
