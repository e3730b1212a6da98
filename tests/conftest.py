"""Shared test set-up: the ``gpu`` marker, import paths and golden loaders."""

from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def golden_path(name: str) -> str:
    return os.path.join(GOLDEN, name)


def load_npz(name: str):
    z = np.load(golden_path(name), allow_pickle=False)
    meta = json.loads(bytes(z["__meta__"]).decode())
    return z, meta


def srk3_case(name: str):
    """(prepared input state dict, outputs per step, meta) of a whole-step fixture."""
    z, meta = load_npz(f"srk3_{name}.npz")
    st = {}
    for k in z.files:
        if k.startswith("in__"):
            v = z[k]
            st[k[4:]] = v.item() if v.ndim == 0 else v
    outs = []
    for step in range(meta["steps"]):
        pre = f"out{step + 1}__"
        outs.append({k[len(pre):]: z[k] for k in z.files if k.startswith(pre)})
    return st, outs, meta


def kernel_calls():
    """(merged input mapping, list of (variant, inputs, outputs)) of the per-kernel fixture."""
    z, meta = load_npz("kernels_l1_L10.npz")

    def unpack(d):
        return {k: (z[v["array"]] if "array" in v else v["value"]) for k, v in d.items()}

    calls = [(c["variant"], unpack(c["inputs"]), unpack(c["outputs"])) for c in meta["calls"]]
    merged = {}
    for _, ins, _ in calls:
        for k, v in ins.items():
            merged.setdefault(k, v)
    for a, b in {"rho_base": "rb", "rtheta_base": "rtb", "exner_base": "pb"}.items():
        merged.setdefault(a, merged.get(b))
    return merged, calls


def perturbed_state(st: dict, seed: int) -> dict:
    """A prepared state moved off the hard primal's switching surfaces.

    Seeded noise on u, w (interior interfaces), theta_m and a small positive
    amount on every scalar (the dry water-vapour slot is identically zero,
    where every limiter bound ties), so that finite differences see the same
    branch on both sides of each evaluation."""
    st = dict(st)
    rng = np.random.default_rng(seed)
    nc, ne = int(st["nCells"]), int(st["nEdges"])
    u = np.array(st["u"], dtype=np.float64)
    u[:ne] += 0.3 * rng.standard_normal(u[:ne].shape)
    w = np.array(st["w"], dtype=np.float64)
    w[:nc, 1:-1] += 1e-3 * rng.standard_normal(w[:nc, 1:-1].shape)
    sc = np.array(st["scalars"], dtype=np.float64)
    sc[:nc] += 1e-4 * rng.random(sc[:nc].shape)
    th = np.array(st["theta_m"], dtype=np.float64)
    th[:nc] += 0.05 * rng.standard_normal(th[:nc].shape)
    st.update({"u": u, "w": w, "scalars": sc, "theta_m": th})
    return st


def rel_err(got, ref) -> float:
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    got = got[: ref.shape[0]]
    scale = float(np.max(np.abs(ref))) if ref.size else 0.0
    diff = float(np.max(np.abs(got - ref))) if ref.size else 0.0
    if scale == 0.0:
        return diff
    return diff / scale


@pytest.fixture(scope="session")
def gpu_available():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    return True
