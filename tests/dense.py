"""Dense statevector simulation (numpy) of small circuits: the brute-force oracle of the tests.

Qubit 0 is the most significant axis of the state vector (as in the reference tests, test_tableau.py:13-24).
Gate matrices are the qelib1 definitions; rz(theta) = diag(e^{-i theta/2}, e^{i theta/2}).
"""
import re

import numpy as np

G1 = {
    "h": np.array([[1, 1], [1, -1]]) / np.sqrt(2),
    "s": np.diag([1, 1j]),
    "sdg": np.diag([1, -1j]),
    "sx": np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]) / 2,
    "sxdg": np.array([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]]) / 2,
    "x": np.array([[0, 1], [1, 0]]),
    "z": np.diag([1, -1]),
    "t": np.diag([1, np.exp(1j * np.pi / 4)]),
    "tdg": np.diag([1, np.exp(-1j * np.pi / 4)]),
}


def gate(name):
    if name in G1:
        return G1[name]
    m = re.match(r"rz\((.*)\)", name)
    if m:
        th = float(eval(m.group(1), {"pi": np.pi}))
        return np.diag([np.exp(-1j * th / 2), np.exp(1j * th / 2)])
    raise ValueError(name)


def apply1(psi, U, q, n):
    psi = psi.reshape([2] * n)
    psi = np.moveaxis(np.tensordot(U, psi, axes=([1], [q])), 0, q)
    return psi.reshape(-1)


def cz(psi, a, b, n):
    psi = psi.reshape([2] * n).copy()
    idx = [slice(None)] * n
    idx[a] = 1
    idx[b] = 1
    psi[tuple(idx)] *= -1
    return psi.reshape(-1)


def simulate(n, ops):
    psi = np.zeros(2 ** n, complex)
    psi[0] = 1
    for name, qs in ops:
        psi = cz(psi, qs[0], qs[1], n) if name == "cz" else apply1(psi, gate(name), qs[0], n)
    return psi


def probabilities(n, ops):
    return np.abs(simulate(n, ops)) ** 2


def step_order_index(bits_step, step_vertices, n):
    """Statevector index of a step-order string."""
    b = [0] * n
    for k, v in enumerate(step_vertices):
        b[v] = int(bits_step[k])
    return int("".join(map(str, b)), 2)
