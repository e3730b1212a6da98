"""Write the reference's gradients of a multi-step loss (golden VJP vectors).

Run in the build container only (the reference tree is not on the GPU side):

    python tests/golden/make_golden_vjp.py

For the prepared input state of the ``l2_L20_tracer`` whole-step fixture
(``srk3_l2_L20_tracer.npz``), the reference advances two large steps with
``atm_srk3(..., differentiable=True)``
(``mpas_pytorch/dynamics/time_integration.py:4897``) from leaves for the
inter-step prognostics ``u, w, theta_m, rho_zz, scalars, rt_diabatic_tend``
(``atmosphere_model.py:157-168``), evaluates

    loss = T_theta / c_theta + T_u / c_u + T_q / c_q
    T_theta = sum_cells area * sum_k (theta_m - theta_m0)^2
    T_u     = sum_edges dc dv * sum_k u^2
    T_q     = sum_cells area * sum_k scalars[..., 1]^2

(theta perturbation, kinetic-energy-like and tracer terms, so that the
dynamics, the transport and the monotonic limiter all carry adjoint; the
constants c are the terms' values, recorded, so each term weighs about 1), and
records ``torch.autograd.grad`` of it with respect to the leaves, as
``run_simulation.vjp`` (``atmosphere_model.py:2101-2155``) does.  The outputs
``vjp_l2_L20_tracer.npz`` (the fixture state as prepared, where the dry
water-vapour slot makes every limiter bound tie, so torch's tie conventions
are exercised) and ``vjp_l2_L20_tracer_perturbed.npz`` (the state moved off
the switching surfaces by ``conftest.perturbed_state`` with seed
``PERTURB_SEED``, which finite differences can check) hold the loss, its
normalising constants and the six gradients.
"""

from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

from reference_harness import import_reference  # noqa: E402

import numpy as np  # noqa: E402
import torch  # noqa: E402

from make_golden import reference_config  # noqa: E402
from conftest import perturbed_state, srk3_case  # noqa: E402

R = import_reference()

CASE = "l2_L20_tracer"
STEPS = 2
TRACKED = ("u", "w", "theta_m", "rho_zz", "scalars", "rt_diabatic_tend")
PERTURB_SEED = 11


def loss_terms(out: dict, st0: dict, nc: int, ne: int) -> list:
    area = torch.as_tensor(1.0 / np.asarray(st0["invAreaCell"], dtype=np.float64)[:nc])
    ae = torch.as_tensor((np.asarray(st0["dcEdge"], dtype=np.float64) * np.asarray(st0["dvEdge"]))[:ne])
    th0 = torch.as_tensor(np.asarray(st0["theta_m"], dtype=np.float64))
    d = out["theta_m"][:nc] - th0[:nc]
    l_theta = torch.sum(area[:, None] * d * d)
    l_u = torch.sum(ae[:, None] * out["u"][:ne] * out["u"][:ne])
    q = out["scalars"][:nc, :, 1]
    l_q = torch.sum(area[:, None] * q * q)
    return [l_theta, l_u, l_q]


def write(perturb: bool) -> None:
    st, _, meta = srk3_case(CASE)
    if perturb:
        st = perturbed_state(st, PERTURB_SEED)
    nc, ne, L = int(st["nCells"]), int(st["nEdges"]), int(st["nVertLevels"])
    mesh = {k: (torch.from_numpy(np.ascontiguousarray(v).copy()) if isinstance(v, np.ndarray) else v)
            for k, v in st.items()}
    state = dict(mesh)
    if "rt_diabatic_tend" not in state:
        state["rt_diabatic_tend"] = torch.zeros((nc + 1, L), dtype=torch.float64)
    leaves = {}
    for k in TRACKED:
        leaf = state[k].detach().clone().requires_grad_(True)
        state[k] = leaf
        leaves[k] = leaf
    cfg = reference_config(st, meta["dt"])
    t0 = time.perf_counter()
    cur = state
    for _ in range(STEPS):
        cur = R.ti.atm_srk3(cur, mesh, cfg, float(meta["dt"]), differentiable=True)
    terms = loss_terms(cur, st, nc, ne)
    consts = [float(t.detach()) for t in terms]
    loss = sum(t / c for t, c in zip(terms, consts))
    grads = torch.autograd.grad(loss, list(leaves.values()), allow_unused=True)
    secs = time.perf_counter() - t0
    payload = {"loss": np.asarray(float(loss.detach())), "loss_constants": np.asarray(consts)}
    for k, g in zip(leaves, grads):
        payload["grad__" + k] = (np.zeros(tuple(leaves[k].shape)) if g is None else g.detach().numpy().copy())
    info = {"case": CASE, "steps": STEPS, "dt": meta["dt"], "tracked": list(TRACKED),
            "perturb_seed": PERTURB_SEED if perturb else None,
            "reference_seconds": secs, "torch_threads": torch.get_num_threads()}
    path = os.path.join(HERE, f"vjp_{CASE}{'_perturbed' if perturb else ''}.npz")
    np.savez_compressed(path, __meta__=np.frombuffer(json.dumps(info).encode(), dtype=np.uint8), **payload)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB) loss={float(loss.detach()):.17e} terms={consts} in {secs:.1f} s")


def main() -> None:
    torch.set_num_threads(8)
    write(False)
    write(True)


if __name__ == "__main__":
    main()
