"""The reference's gradients (golden VJP vectors) checked against the oracle, on the CPU.

``tests/golden/make_golden_vjp.py`` recorded torch autograd gradients of the
reference through two large steps.  On the perturbed state (no value on a
switching surface) they must satisfy <g, t> = d loss(x + h t)/dh, with the
derivative taken by central differences of the oracle's two steps; this pins
the golden gradients that the GPU adjoint is compared with.
"""

from __future__ import annotations

import numpy as np

from conftest import load_npz, perturbed_state, srk3_case

CASE = "l2_L20_tracer"
KEYS = ("u", "w", "theta_m", "rho_zz", "scalars", "rt_diabatic_tend")


def vjp_loss_oracle(st: dict, consts, steps: int):
    from oracle import oracle as O, srk3 as OS

    om = O.Mesh(st)
    nc, ne, L = om.nc, om.ne, om.L
    area = 1.0 / np.asarray(st["invAreaCell"], dtype=np.float64)[:nc]
    ae = (np.asarray(st["dcEdge"], dtype=np.float64) * np.asarray(st["dvEdge"], dtype=np.float64))[:ne]
    th0 = np.asarray(st["theta_m"], dtype=np.float64)

    def loss(X: dict) -> float:
        s2 = dict(st)
        s2.setdefault("rt_diabatic_tend", np.zeros((nc + 1, L)))
        s2.update(X)
        for _ in range(steps):
            s2 = OS.srk3(s2, om, float(dt), len_disp=float(st["nominalMinDc"]))
        d = s2["theta_m"][:nc] - th0[:nc]
        terms = (np.sum(area[:, None] * d * d), np.sum(ae[:, None] * s2["u"][:ne] ** 2),
                 np.sum(area[:, None] * s2["scalars"][:nc, :, 1] ** 2))
        return float(sum(t / c for t, c in zip(terms, consts)))

    dt = srk3_case(CASE)[2]["dt"]
    return loss


def test_perturbed_golden_gradients_match_oracle_finite_differences():
    z, meta = load_npz(f"vjp_{CASE}_perturbed.npz")
    st = perturbed_state(srk3_case(CASE)[0], meta["perturb_seed"])
    loss = vjp_loss_oracle(st, z["loss_constants"], meta["steps"])
    nc, L = int(st["nCells"]), int(st["nVertLevels"])
    X = {k: np.array(st[k] if k in st else np.zeros((nc + 1, L)), dtype=np.float64) for k in KEYS}
    assert abs(loss({}) - float(z["loss"])) <= 1e-12 * abs(float(z["loss"]))
    rng = np.random.default_rng(2)
    # per leaf, the best of several steps: large steps cross limiter switches,
    # small ones lose digits to rounding of the O(1) loss
    for k in KEYS:
        t = rng.standard_normal(X[k].shape) * (float(np.std(X[k])) or 1.0)
        t[-1] = 0.0
        if k == "w":
            t[:, 0] = 0.0
            t[:, -1] = 0.0
        ad = float(np.sum(z["grad__" + k] * t))
        errs = []
        for h in (1e-6, 1e-7, 1e-8, 1e-9):
            fd = (loss({k: X[k] + h * t}) - loss({k: X[k] - h * t})) / (2 * h)
            errs.append(abs(fd - ad))
        # rounding of a loss of ~3 limits the difference quotient to ~1e-6 absolute
        assert min(errs) <= max(1e-3 * abs(ad), 2e-6), (k, ad, errs)
