"""Problem definitions (contexts and initial conditions) for the configurations this package runs.

Mirrors HONEE's problem setup: NS_NEWTONIAN_IG defaults (problems/newtonian.c:588-613), the Taylor-Green YAML
(examples/taylor_green_vortex.yaml) and ICsTaylorGreen (qfunctions/taylorgreen.h:23-51).  Host-side setup only.
"""
from __future__ import annotations

import numpy as np

from .driver import STAB_SUPG, VAR_CONS, VAR_ENTROPY, VAR_PRIM, context_from_dict
from .mesh import BoxMesh

# examples/taylor_green_vortex.yaml: Re=1600, M=0.1, Pr=0.71
TGV_GAS = {"lambda": -2.0 / 3.0, "mu": 6.25e-4, "k": 0.8836267606, "cv": 717.0, "cp": 1003.8}
TGV_REFERENCE_Y = (71.42857143, 1.0, 0.0, 0.0, 0.2490535963)  # pressure, velocity, temperature


def default_tau(degree: int) -> dict:
    """Default Tau coefficients of NS_NEWTONIAN_IG (problems/newtonian.c:601-608)."""
    cv_func = (36.0, 60.0, 128.0)
    return {"Ctau_t": 1.0, "Ctau_v": cv_func[min(3, degree) - 1], "Ctau_C": 0.25 / degree, "Ctau_M": 0.25 / degree, "Ctau_E": 0.125}


TGV_TAU = {"Ctau_t": 1.0, "Ctau_v": 60.0, "Ctau_C": 0.5, "Ctau_M": 0.5, "Ctau_E": 0.5}


def tgv_context_dict(state_var=VAR_CONS, stabilization=STAB_SUPG, implicit=True, dt=2e-3, tau=None) -> dict:
    return {"gas": dict(TGV_GAS), "tau": dict(tau or TGV_TAU), "state_var": state_var, "stabilization": stabilization,
            "is_implicit": implicit, "dt": dt}


def tgv_context(**kw):
    return context_from_dict(tgv_context_dict(**kw))


# ---- state conversions needed to express an initial condition in the chosen variables (newtonian_state.h) ----
def _rgas(gas):
    return gas["cp"] - gas["cv"]


def cons_from_prim(gas, Y):
    rho = Y[..., 0] / (_rgas(gas) * Y[..., 4])
    u = Y[..., 1:4]
    U = np.empty_like(Y)
    U[..., 0] = rho
    U[..., 1:4] = rho[..., None] * u
    U[..., 4] = rho * (gas["cv"] * Y[..., 4] + 0.5 * np.sum(u * u, axis=-1))
    return U


def entropy_from_prim(gas, Y):
    gamma = gas["cp"] / gas["cv"]
    rho = Y[..., 0] / (_rgas(gas) * Y[..., 4])
    entropy = np.log(Y[..., 0]) - gamma * np.log(rho)
    rdp = rho / Y[..., 0]
    ek = 0.5 * np.sum(Y[..., 1:4] ** 2, axis=-1)
    V = np.empty_like(Y)
    V[..., 0] = (gamma - entropy) / (gamma - 1) - rdp * ek
    V[..., 1:4] = rdp[..., None] * Y[..., 1:4]
    V[..., 4] = -rdp
    return V


def q_from_prim(gas, Y, state_var):
    if state_var == VAR_PRIM:
        return np.array(Y, dtype=np.float64)
    if state_var == VAR_CONS:
        return cons_from_prim(gas, Y)
    return entropy_from_prim(gas, Y)


def taylor_green_prim(xyz, ref_Y=TGV_REFERENCE_Y, gas=TGV_GAS, background=(0.0, 0.0, 0.0), lx=2 * np.pi):
    """Primitive TGV state at points xyz[..., 3] (qfunctions/taylorgreen.h:35-45); all directions scale by 2 pi / lx."""
    R = _rgas(gas)
    V0 = float(np.linalg.norm(ref_Y[1:4]))
    rho0 = ref_Y[0] / (ref_Y[4] * R)
    x = np.asarray(xyz, dtype=np.float64) * (2 * np.pi / lx)
    Y = np.empty(x.shape[:-1] + (5,))
    Y[..., 0] = ref_Y[0] + (rho0 * V0**2 / 16) * (np.cos(2 * x[..., 0]) + np.cos(2 * x[..., 1])) * np.cos(2 * x[..., 2] + 2)
    Y[..., 1] = background[0] + V0 * np.sin(x[..., 0]) * np.cos(x[..., 1]) * np.cos(x[..., 2])
    Y[..., 2] = background[1] - V0 * np.cos(x[..., 0]) * np.sin(x[..., 1]) * np.cos(x[..., 2])
    Y[..., 3] = background[2]
    Y[..., 4] = ref_Y[4]
    return Y


def nodal_field(mesh: BoxMesh, per_node_fn, average: bool = True) -> np.ndarray:
    """Evaluate a pointwise function at every element's nodes and average over node multiplicity, as HONEE's
    ICs_FixMultiplicity does (src/misc.c:16-48).  Returns an L-vector [nnodes][5].  average=False keeps one element's
    value per node instead (the same for a continuous, periodic function up to rounding; for meshes too large for the
    averaging's unbuffered accumulation)."""
    if not average:  # one coordinate per node (a periodic image's on shared periodic nodes), evaluated once
        xyz = mesh.element_node_coords(fast=True)
        xn = np.empty((mesh.nnodes, 3))
        xn[mesh.offsets.ravel()] = xyz.reshape(-1, 3)
        return per_node_fn(xn).reshape(mesh.nnodes, 5)
    xyz = mesh.element_node_coords()
    vals = per_node_fn(xyz.reshape(-1, 3)).reshape(mesh.nelem, -1, 5)
    q = np.zeros((mesh.nnodes, 5))
    np.add.at(q, mesh.offsets.ravel(), vals.reshape(-1, 5))
    return q / mesh.multiplicity()[:, None]


def taylor_green_ic(mesh: BoxMesh, state_var=VAR_CONS, gas=TGV_GAS, ref_Y=TGV_REFERENCE_Y, background=(0.0, 0.0, 0.0), lx=2 * np.pi):
    return nodal_field(mesh, lambda x: q_from_prim(gas, taylor_green_prim(x, ref_Y, gas, background, lx), state_var))


def random_fourier_modes(seed: int, kmax: int = 8, k0: float = 4.0, u_rms: float = 1.0, lx: float = 2 * np.pi):
    """Random-phase solenoidal Fourier modes of a periodic box [0, lx]^3 with energy spectrum E(k) ~ k^4 exp(-2 (k/k0)^2),
    wavevectors with 1 <= |k| <= kmax (one of each +-k pair), normalized to the given per-component RMS velocity.  Returns the
    wavevectors [m][3] (physical), unit directions perpendicular to them [m][3], amplitudes [m] and phases [m]."""
    rng = np.random.default_rng(seed)
    rng_k = np.arange(-kmax, kmax + 1)
    kk = np.stack(np.meshgrid(rng_k, rng_k, rng_k, indexing="ij"), -1).reshape(-1, 3)
    kn = np.linalg.norm(kk, axis=1)
    half = (kk[:, 0] > 0) | ((kk[:, 0] == 0) & (kk[:, 1] > 0)) | ((kk[:, 0] == 0) & (kk[:, 1] == 0) & (kk[:, 2] > 0))
    sel = half & (kn >= 1) & (kn <= kmax)
    kk, kn = kk[sel].astype(np.float64), kn[sel]
    shell = np.rint(kn).astype(int)
    counts = np.bincount(shell)
    E = kn ** 4 * np.exp(-2 * (kn / k0) ** 2)
    amp = np.sqrt(E / counts[shell])
    d = rng.normal(size=kk.shape)
    d -= (np.sum(d * kk, axis=1) / kn ** 2)[:, None] * kk
    d /= np.linalg.norm(d, axis=1)[:, None]
    phase = rng.uniform(0, 2 * np.pi, size=len(kk))
    amp *= u_rms / np.sqrt(np.sum(amp ** 2) / 6.0)  # per-component mean square: sum amp^2 / 2 over three components
    return kk * (2 * np.pi / lx), d, amp, phase


def random_fourier_velocity(xyz, modes) -> np.ndarray:
    """Velocity of random_fourier_modes at points xyz [n][3]: sum_m amp_m d_m cos(k_m . x + phase_m)."""
    kk, d, amp, phase = modes
    u = np.zeros((len(xyz), 3))
    for s in range(0, len(kk), 256):
        arg = xyz @ kk[s:s + 256].T + phase[s:s + 256]
        u += (np.cos(arg) * amp[s:s + 256]) @ d[s:s + 256]
    return u


def turbulence_prim(mesh: BoxMesh, modes, ref_Y=TGV_REFERENCE_Y) -> np.ndarray:
    """Primitive state [nnodes][5]: reference pressure and temperature with the random Fourier velocity."""
    def f(x):
        Y = np.tile(np.asarray(ref_Y, dtype=np.float64), (len(x), 1))
        Y[:, 1:4] = random_fourier_velocity(x, modes)
        return Y
    return nodal_field(mesh, f)


# ---- channel (problems/channel.c:75-140, qfunctions/channel.h:24-53) ----
def channel_context_dict(mu, umax=1.0, theta0=TGV_REFERENCE_Y[4], P0=TGV_REFERENCE_Y[0], H=1.0, body_force_scale=1.0, state_var=VAR_PRIM,
                         stabilization=STAB_SUPG, dt=1e-2, tau=None):
    """NS_CHANNEL's context on the Taylor-Green gas with viscosity mu: the body force B = 2 umax mu / H^2 enters as the
    gravity vector g = (B / rho0, 0, 0) with rho0 = P0 / (R theta0) (problems/channel.c:126-137)."""
    gas = dict(TGV_GAS)
    gas["mu"] = float(mu)
    rho0 = P0 / (_rgas(gas) * theta0)
    B = body_force_scale * 2 * umax * gas["mu"] / (H * H)
    d = {"gas": gas, "tau": dict(tau or TGV_TAU), "state_var": state_var, "stabilization": stabilization, "is_implicit": True, "dt": dt,
         "g": (B / rho0, 0.0, 0.0)}
    return d


def channel_prim(xyz, gas, umax=1.0, theta0=TGV_REFERENCE_Y[4], P0=TGV_REFERENCE_Y[0], center=1.0, H=1.0):
    """Exact_Channel (qfunctions/channel.h:24-53): Poiseuille velocity and the reference's temperature profile
    theta0 (1 + (Pr Ec / 3)(1 - ((y - center) / H)^4)) with Pr = mu / (cp k) and Ec = umax^2 / (cp theta0), as the
    reference writes them."""
    y = np.asarray(xyz, dtype=np.float64)[..., 1]
    eta2 = ((y - center) / H) ** 2
    Pr = gas["mu"] / (gas["cp"] * gas["k"])
    Ec = umax * umax / (gas["cp"] * theta0)
    Y = np.zeros(np.shape(xyz)[:-1] + (5,))
    Y[..., 0] = P0
    Y[..., 1] = umax * (1 - eta2)
    Y[..., 4] = theta0 * (1 + (Pr * Ec / 3) * (1 - eta2 * eta2))
    return Y


def channel_perturbation(xyz, seed, lengths, n_modes=20, amplitude=0.1, center=1.0, H=1.0):
    """Seeded velocity perturbation of the channel: per component, n_modes random Fourier modes in each periodic
    direction (x and z, integer wavenumbers 1..n_modes on the box lengths, random phases and unit amplitudes), damped by
    (1 - ((y - center) / H)^2)^2 so it vanishes at the walls, scaled to a maximum of `amplitude` (the centerline
    velocity's fraction)."""
    rng = np.random.default_rng(seed)
    x = np.asarray(xyz, dtype=np.float64)
    damp = (1 - ((x[:, 1] - center) / H) ** 2) ** 2
    u = np.zeros((len(x), 3))
    for c in range(3):
        for d, L in ((0, lengths[0]), (2, lengths[2])):
            k = rng.integers(1, n_modes + 1, size=n_modes)
            ph = rng.uniform(0, 2 * np.pi, size=n_modes)
            u[:, c] += np.sin(np.outer(x[:, d], 2 * np.pi * k / L) + ph).sum(axis=1)
    u *= damp[:, None]
    return u * (amplitude / max(np.abs(u).max(), 1e-300))


def channel_ic(mesh: BoxMesh, gas, seed=None, umax=1.0, amplitude=0.1, lengths=(4 * np.pi, 2.0, 4 * np.pi / 3)):
    """Primitive channel state [nnodes][5]: Exact_Channel plus, with a seed, the perturbation at amplitude x umax."""
    def f(x):
        Y = channel_prim(x, gas, umax=umax)
        if seed is not None:
            Y[:, 1:4] += umax * channel_perturbation(x, seed, lengths, amplitude=amplitude)
        return Y
    return nodal_field(mesh, f)
