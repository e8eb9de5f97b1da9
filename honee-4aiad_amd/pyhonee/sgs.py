"""Data-driven subgrid-stress model parameters: the PHASTA-style .dat files HONEE reads
(src/honee-file.c:277-301) and the layout SgsDDContextFill gives them (problems/sgs_dd_model.c:602-650)."""
from __future__ import annotations

import os

import numpy as np

NUM_INPUTS, NUM_OUTPUTS, NUM_NEURONS = 6, 6, 20


def read_dat(path: str) -> np.ndarray:
    """A .dat array: a first line "rows cols", then one row of `cols` values per line."""
    with open(path) as f:
        rows, cols = (int(v) for v in f.readline().split()[:2])
        vals = [[float(v) for v in f.readline().split()] for _ in range(rows)]
    a = np.array(vals, dtype=np.float64)
    if a.shape != (rows, cols):
        raise ValueError(f"{path}: expected {rows}x{cols} values")
    return a


def load_dd_model(directory: str, alpha: float = 0.0) -> dict:
    """Biases, weights and output bounds of the 6-20-6 model.  The weight files hold the matrices column-major
    (w1.dat [input][neuron], w2.dat [neuron][output]); the returned W1 is [neuron][input] and W2 [output][neuron],
    as the model evaluates them (problems/sgs_dd_model.c:632-645)."""
    b1 = read_dat(os.path.join(directory, "b1.dat")).ravel()
    b2 = read_dat(os.path.join(directory, "b2.dat")).ravel()
    w1 = read_dat(os.path.join(directory, "w1.dat")).ravel()
    w2 = read_dat(os.path.join(directory, "w2.dat")).ravel()
    bounds = read_dat(os.path.join(directory, "OutScaling.dat")).ravel()
    if not (b1.size == NUM_NEURONS and b2.size == NUM_OUTPUTS and w1.size == NUM_INPUTS * NUM_NEURONS and w2.size == NUM_NEURONS * NUM_OUTPUTS
            and bounds.size == 2 * NUM_OUTPUTS):
        raise ValueError(f"{directory}: parameter sizes do not match a {NUM_INPUTS}-{NUM_NEURONS}-{NUM_OUTPUTS} model")
    W1 = w1.reshape(NUM_INPUTS, NUM_NEURONS).T.copy()
    W2 = w2.reshape(NUM_NEURONS, NUM_OUTPUTS).T.copy()
    return {"b1": b1, "b2": b2, "W1": W1, "W2": W2, "bounds": bounds.reshape(NUM_OUTPUTS, 2).copy(), "alpha": float(alpha)}
