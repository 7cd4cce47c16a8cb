"""Readers for the golden fixtures written by make_golden.js (reference outputs, stored as data)."""
import gzip
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_json(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def mask_index(N):
    """Flat indices (reference layout (kz*N + ky)*NH + kx) of the modes inside the 2/3 mask, in loop order."""
    NH = N // 2 + 1
    kc = N // 3
    kw = np.where(np.arange(N) <= N // 2, np.arange(N), np.arange(N) - N)
    l, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(NH), indexing="ij")
    m = (np.abs(kw[l]) <= kc) & (np.abs(kw[j]) <= kc) & (i <= kc)
    return ((l * N + j) * NH + i)[m]


def load_state(name, N):
    """A compact state fixture expanded to the reference's 6 x NS layout (zeros outside the mask)."""
    with gzip.open(os.path.join(HERE, name), "rb") as f:
        raw = np.frombuffer(f.read(), dtype="<f8")
    idx = mask_index(N)
    comp = raw.reshape(6, idx.size)
    NS = N * N * (N // 2 + 1)
    S = np.zeros((6, NS))
    S[:, idx] = comp
    return S


def load_f64(name):
    with gzip.open(os.path.join(HERE, name), "rb") as f:
        return np.frombuffer(f.read(), dtype="<f8").copy()
