"""Host-side channel setup (no GPU): the laminar profile of Exact_Channel (qfunctions/channel.h:24-53) and the body
force of NS_CHANNEL (problems/channel.c:126-137) balance (mu u'' + rho0 g = 0 for the quadratic profile), the walls are
at rest, and the seeded perturbation vanishes at the walls with the requested amplitude."""
import numpy as np

from pyhonee import mesh as M
from pyhonee import problems as PR


def test_poiseuille_balances_body_force():
    d = PR.channel_context_dict(1e-2, umax=1.5)
    gas = d["gas"]
    y = np.linspace(0.0, 2.0, 41)
    xyz = np.stack([np.zeros_like(y), y, np.zeros_like(y)], axis=1)
    Y = PR.channel_prim(xyz, gas, umax=1.5)
    assert Y[0, 1] == 0.0 and Y[-1, 1] == 0.0 and abs(Y[20, 1] - 1.5) < 1e-15
    u_yy = np.polyfit(y, Y[:, 1], 2)[0] * 2
    rho0 = PR.TGV_REFERENCE_Y[0] / ((gas["cp"] - gas["cv"]) * PR.TGV_REFERENCE_Y[4])
    assert abs(gas["mu"] * u_yy + rho0 * d["g"][0]) < 1e-12
    assert np.all(Y[:, 4] >= PR.TGV_REFERENCE_Y[4]) and np.all(Y[:, 0] == PR.TGV_REFERENCE_Y[0])


def test_perturbation_vanishes_at_walls():
    mesh = M.box_mesh((4, 6, 2), 2, lengths=(2 * np.pi, 2.0, 2 * np.pi / 3), periodic=(True, False, True), stretch=(1.0, 1.1, 1.0),
                      symmetric_stretch=(False, True, False))
    x = mesh.element_node_coords().reshape(-1, 3)
    u = PR.channel_perturbation(x, 7, (2 * np.pi, 2.0, 2 * np.pi / 3))
    wall = (np.abs(x[:, 1]) < 1e-12) | (np.abs(x[:, 1] - 2.0) < 1e-12)
    assert wall.any() and np.abs(u[wall]).max() < 1e-15
    assert abs(np.abs(u).max() - 0.1) < 1e-12
    # the stretched mesh grows geometrically from both walls
    h = np.diff(mesh.vertex_grid[0, :, 0, 1])
    assert np.allclose(h[1:3] / h[0:2], 1.1) and np.allclose(h, h[::-1])
