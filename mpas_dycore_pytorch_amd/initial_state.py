"""Synthetic initial conditions and cold-start preparation.

The reference takes raw initial fields from MPAS ``init.nc`` files and
prepares them before the first step (``atmosphere_model.py:782``
``_prepare_timestep_state``): ``atm_init_coupled_diagnostics``
(``dynamics/init_coupled_diagnostics.py:147``), the padding sentinel
``rho_zz[nCells] = 1`` (``atmosphere_model.py:529``), the diagnostic solve
(``atmosphere_model.py:1186``) and the vector reconstruction
(``atmosphere_model.py:1339``).

``raw_state`` builds the raw fields of the synthetic workload:

* dry constant-stratification base state, N = 0.01 1/s, theta_s = 300 K:
  theta_b(z) = theta_s exp(N^2 z / g),
  pi_b(z) = 1 + g^2 / (cp theta_s N^2) (exp(-N^2 z / g) - 1),
  rho_b = p0 pi_b^(cv/R) / (R theta_b), evaluated at mass-level mid-heights;
* u = u0 cos(lat) projected on the edge normals, w = 0;
* theta = theta_b + a seeded N(0, sigma) perturbation smoothed by two passes
  of cell-neighbour averaging; rho = rho_b theta_b / theta (unchanged pressure);
* scalars: slot 0 is water vapour (zero: dry), further slots are passive
  tracers with seeded smooth values in [0, 1].

``prepare_state`` is the host-side restatement of the cold-start chain; the
time-stepping path never calls it.
"""

from __future__ import annotations

import numpy as np

GRAVITY = 9.80616
RGAS = 287.0
CP = 1004.5
CV = 717.5
P0 = 100000.0
RVORD = 1.6083623693379792


def _cell_neighbour_average(field: np.ndarray, mesh: dict) -> np.ndarray:
    """Average of each cell with its edge neighbours (active rows only)."""
    nc = int(mesh["nCells"])
    coc = mesh["cellsOnCell"][:nc]
    valid = coc > 0
    idx = np.where(valid, coc - 1, 0)
    s = field[:nc].copy()
    for j in range(coc.shape[1]):
        s = s + np.where(valid[:, j, None], field[idx[:, j]], 0.0)
    cnt = 1.0 + valid.sum(axis=1)
    out = field.copy()
    out[:nc] = s / cnt[:, None]
    return out


def raw_state(
    mesh: dict,
    *,
    seed: int,
    n_tracers: int = 0,
    theta_s: float = 300.0,
    brunt_vaisala: float = 0.01,
    u0: float = 20.0,
    theta_sigma: float = 0.5,
) -> dict:
    """Raw (unprepared) prognostic and base-state fields, padded layout."""
    nc, ne = int(mesh["nCells"]), int(mesh["nEdges"])
    L = int(mesh["nVertLevels"])
    rng = np.random.default_rng(seed)

    zgrid = mesh["zgrid"]
    zmid = 0.5 * (zgrid[:, :-1] + zgrid[:, 1:])
    n2 = brunt_vaisala * brunt_vaisala
    theta_b = theta_s * np.exp(n2 * zmid / GRAVITY)
    exner_b = 1.0 + GRAVITY * GRAVITY / (CP * theta_s * n2) * (np.exp(-n2 * zmid / GRAVITY) - 1.0)
    rho_b = P0 * exner_b ** (CV / RGAS) / (RGAS * theta_b)
    theta_b[nc] = 0.0
    rho_b[nc] = 0.0

    pert = np.zeros((nc + 1, L))
    pert[:nc] = rng.normal(0.0, theta_sigma, size=(nc, L))
    pert = _cell_neighbour_average(pert, mesh)
    pert = _cell_neighbour_average(pert, mesh)
    pert[nc] = 0.0
    theta = theta_b + pert
    rho = np.zeros_like(rho_b)
    rho[:nc] = rho_b[:nc] * theta_b[:nc] / theta[:nc]

    lat_e, lon_e = mesh["latEdge"][:ne], mesh["lonEdge"][:ne]
    east = np.stack([-np.sin(lon_e), np.cos(lon_e), np.zeros_like(lon_e)], axis=1)
    # edge normal recovered from angleEdge (angle from local east)
    north = np.stack(
        [-np.sin(lat_e) * np.cos(lon_e), -np.sin(lat_e) * np.sin(lon_e), np.cos(lat_e)], axis=1
    )
    ang = mesh["angleEdge"][:ne]
    normal = np.cos(ang)[:, None] * east + np.sin(ang)[:, None] * north
    un = u0 * np.cos(lat_e) * np.sum(east * normal, axis=1)
    u = np.zeros((ne + 1, L))
    u[:ne] = un[:, None]

    n_scalars = 1 + n_tracers
    scalars = np.zeros((nc + 1, L, n_scalars))
    if n_tracers:
        lat_c, lon_c = mesh["latCell"][:nc], mesh["lonCell"][:nc]
        zfrac = zmid[:nc] / zgrid[:nc, -1:]
        for t in range(n_tracers):
            ph = rng.uniform(0.0, 2.0 * np.pi, size=3)
            wave = (
                np.sin(2.0 * lon_c + ph[0]) * np.cos(lat_c) ** 2
                + 0.5 * np.sin(3.0 * lat_c + ph[1])
            )
            vert = np.cos(np.pi * zfrac + ph[2])
            q = 0.5 + 0.3 * wave[:, None] * (0.75 + 0.25 * vert)
            scalars[:nc, :, 1 + t] = np.clip(q, 0.0, 1.0)

    t_init = theta_b.copy()
    return {
        "u": u,
        "w": np.zeros((nc + 1, L + 1)),
        "theta": theta,
        "rho": rho,
        "qv": scalars[:, :, 0].copy(),
        "rho_base": rho_b,
        "theta_base": theta_b,
        "scalars": scalars,
        "t_init": t_init,
        "u_init": np.zeros(L),
        "v_init": np.zeros(L),
        "qv_init": np.zeros(L),
        "num_scalars": n_scalars,
        "moist_start": 1,
        "moist_end": 1,
    }


# ---------------------------------------------------------------------------
# cold-start preparation (host restatement)
# ---------------------------------------------------------------------------
def init_coupled_diagnostics(mesh: dict, raw: dict) -> dict:
    """``atm_init_coupled_diagnostics`` (``init_coupled_diagnostics.py:147-380``).

    Flat-terrain form: the zb/zb3 terrain terms of rw vanish because
    ``zb_cell`` and ``zb3_cell`` are zero on this mesh.
    """
    nc, ne = int(mesh["nCells"]), int(mesh["nEdges"])
    L = int(mesh["nVertLevels"])
    zz = mesh["zz"]
    fzm, fzp = mesh["fzm"], mesh["fzp"]
    theta_m = raw["theta"] * (1.0 + RVORD * raw["qv"])
    rho_zz = np.zeros_like(raw["rho"])
    rho_zz[:nc] = raw["rho"][:nc] / zz[:nc]
    coe = mesh["cellsOnEdge"][:ne] - 1
    u = raw["u"]
    ru = np.zeros((ne + 1, L))
    ru[:ne] = 0.5 * u[:ne] * (rho_zz[coe[:, 0]] + rho_zz[coe[:, 1]])
    rw = np.zeros((nc + 1, L + 1))
    if L > 1:
        rho_i = fzp[1:] * rho_zz[:nc, :-1] + fzm[1:] * rho_zz[:nc, 1:]
        zz_i = fzp[1:] * zz[:nc, :-1] + fzm[1:] * zz[:nc, 1:]
        rw[:nc, 1:L] = raw["w"][:nc, 1:L] * rho_i * zz_i
    rho_base, theta_base = raw["rho_base"], raw["theta_base"]
    rgas_over_p0 = RGAS / P0
    rcv = RGAS / (CP - RGAS)
    rho_p = rho_zz - rho_base
    rtheta_base = theta_base * rho_base
    rtheta_p = theta_m * rho_p + rho_base * (theta_m - theta_base)
    exner = (zz * rgas_over_p0 * (rtheta_p + rtheta_base)) ** rcv
    exner_base = (zz * rgas_over_p0 * rtheta_base) ** rcv
    pressure_p = zz * RGAS * (exner * rtheta_p + rtheta_base * (exner - exner_base))
    pressure_base = zz * RGAS * exner_base * rtheta_base
    out = {
        "theta_m": theta_m,
        "rho_zz": rho_zz,
        "ru": ru,
        "rw": rw,
        "rho_p": rho_p,
        "rtheta_base": rtheta_base,
        "rtheta_p": rtheta_p,
        "exner": exner,
        "exner_base": exner_base,
        "pressure_p": pressure_p,
        "pressure_base": pressure_base,
    }
    for key in ("rho_p", "rtheta_base", "rtheta_p", "exner", "exner_base", "pressure_p", "pressure_base"):
        out[key][nc] = 0.0
    out["theta_m"][nc] = 0.0
    out["rho_zz"][nc] = 1.0  # padding sentinel, atmosphere_model.py:529-545
    return out


DIAGNOSTIC_KEYS = ("h_edge", "v", "divergence", "vorticity", "ke", "pv_edge")
RECONSTRUCT_KEYS = ("ur_cell", "vr_cell", "uReconstructX", "uReconstructY", "uReconstructZ",
                    "uReconstructZonal", "uReconstructMeridional")


def prepare_state(mesh: dict, raw: dict, *, diagnostics, reconstruct,
                  coef_3rd_order: float = 0.25, c_s: float = 0.125) -> dict:
    """Cold-start chain of the reference (``atmosphere_model.py:782-1339``).

    ``diagnostics(u, rho_zz)`` must return the first diagnostic solve
    (``rk_step=None``: h_edge, v, divergence, vorticity, ke, pv_edge) and
    ``reconstruct(u)`` the cell vector reconstruction; the caller picks the
    implementation (GPU kernels in production).  Returns one mapping that
    holds the mesh, the raw fields and the prepared state, as the reference
    state dict does.
    """
    st = dict(mesh)
    st.update(raw)
    st.update(init_coupled_diagnostics(mesh, raw))
    st["coef_3rd_order"] = coef_3rd_order
    st["c_s"] = c_s
    d = diagnostics(st["u"], st["rho_zz"])
    for k in DIAGNOSTIC_KEYS:
        st[k] = np.asarray(d[k])
    st["rho_edge"] = st["h_edge"]
    r = reconstruct(st["u"])
    for k in RECONSTRUCT_KEYS:
        st[k] = np.asarray(r[k])
    return st
