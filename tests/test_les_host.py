"""Host-side pieces of the a-posteriori LES study (no GPU): the FFT synthesis of the random Fourier initial velocity
against the direct sum of its modes, the shell spectrum against Parseval and the designed E(k) ~ k^4 exp(-2 (k/4)^2),
and the nested coarse sampling."""
import numpy as np

from pyhonee import les as LES
from pyhonee import problems as PR


def test_grid_velocity_equals_mode_sum():
    n = 24
    modes = PR.random_fourier_modes(7)
    u = LES.grid_velocity(n, modes)
    x = np.arange(n) * (2 * np.pi / n)
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    xyz = np.stack([xx, y, z], axis=-1).reshape(-1, 3)
    ref = PR.random_fourier_velocity(xyz, modes)
    assert np.abs(u.reshape(-1, 3) - ref).max() < 1e-12 * np.abs(ref).max()
    # divergence-free: k . u_hat = 0 for every mode
    uh = np.fft.fftn(u, axes=(0, 1, 2))
    k = np.fft.fftfreq(n, 1.0 / n)
    div = k[None, None, :] * uh[..., 0] + k[None, :, None] * uh[..., 1] + k[:, None, None] * uh[..., 2]
    assert np.abs(div).max() < 1e-9 * np.abs(uh).max()


def test_spectrum_parseval_and_shape():
    n = 24
    u = LES.grid_velocity(n, PR.random_fourier_modes(8))
    E = LES.energy_spectrum(u)
    assert abs(E.sum() - 0.5 * np.mean(np.sum(u * u, axis=-1))) < 1e-12
    assert abs(E.sum() - 1.5) < 1e-12  # unit RMS per component
    # each shell holds the mean of k^4 exp(-2 (k/4)^2) over its wavevectors (amplitudes sqrt(E(|k|) / shell count))
    r = np.arange(-8, 9)
    kk = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    kn = np.linalg.norm(kk, axis=1)
    shell = np.rint(kn).astype(int)
    sel = (kn >= 1) & (kn <= 8)
    Ek = kn[sel] ** 4 * np.exp(-2 * (kn[sel] / 4.0) ** 2)
    design = np.bincount(shell[sel], Ek)[1:9] / np.bincount(shell[sel])[1:9]
    ratio = E[1:9] / design
    assert np.allclose(ratio, ratio[0], rtol=1e-10)
    assert E[0] < 1e-28 and E[9:].max() < 1e-28


def test_single_mode_spectrum():
    n = 16
    modes = (np.array([[3.0, 4.0, 0.0]]), np.array([[0.0, 0.0, 1.0]]), np.array([2.0]), np.array([0.3]))
    E = LES.energy_spectrum(LES.grid_velocity(n, modes))
    assert abs(E[5] - 1.0) < 1e-13 and abs(E.sum() - 1.0) < 1e-13  # amp^2 / 4 * 2 components of cos^2 ... = 2^2 / 4


def test_nested_sampling_and_offsets():
    nf, stride = 16, 4
    g = np.arange(nf ** 3 * 2, dtype=float).reshape(nf, nf, nf, 2)
    s = LES._sample(g, stride, (1, 2, 3))
    assert s.shape == ((nf // stride) ** 3, 2)
    assert np.array_equal(s[0], g[1, 2, 3]) and np.array_equal(s[1], g[1, 2, 7])
    off = LES.sample_offsets(stride, 8, 0)
    assert off[0] == (0, 0, 0) and len(set(off)) == 8


def test_prim_cons_round_trip():
    rng = np.random.default_rng(1)
    Y = np.column_stack([71 + rng.uniform(-1, 1, 50), rng.normal(size=(50, 3)), 0.249 + 0.01 * rng.uniform(-1, 1, 50)])
    gas = LES._gas(5e-3)
    assert np.allclose(LES._prim_from_cons(gas, PR.cons_from_prim(gas, Y)), Y, rtol=1e-13, atol=1e-13)
