"""Host side of the adjoint: the angle model reproduces the planner's matrices, its derivatives match finite
differences, and the chain rule from G_k(x) = dZ/dw_k(x) to d ln Pr / d angle matches finite differences of a
brute-force Ising sum (no device)."""
import itertools

import numpy as np

from _pkg import egs


def brute_Z(A, w):
    n = len(A)
    tot = 0j
    for x in itertools.product((0, 1), repeat=n):
        ph = sum(int(A[i][j]) * x[i] * x[j] for i in range(n) for j in range(i + 1, n)) & 1
        p = (-1) ** ph
        for j in range(n):
            p *= w[j][x[j]]
        tot += p
    return tot


def test_angle_model_matrices_match_plan():
    inst = egs.instances.Instance(12, 5, 7)
    model = egs.adjoint.AngleModel(inst)
    M, dA, dB = model.matrices(model.alpha0, model.beta0)
    for k, q in enumerate(inst.step_vertices):
        assert np.allclose(M[k], inst.am.M[q], atol=1e-14)
    h = 1e-6
    for q in (0, 5, 11):
        a2 = model.alpha0.copy()
        a2[q] += h
        a1 = model.alpha0.copy()
        a1[q] -= h
        k = inst.step_vertices.index(q)
        fd = (model.matrices(a2, model.beta0)[0][k] - model.matrices(a1, model.beta0)[0][k]) / (2 * h)
        assert np.allclose(fd, dA[k], atol=1e-8)
        b2 = model.beta0.copy()
        b2[q] += h
        b1 = model.beta0.copy()
        b1[q] -= h
        fd = (model.matrices(model.alpha0, b2)[0][k] - model.matrices(model.alpha0, b1)[0][k]) / (2 * h)
        assert np.allclose(fd, dB[k], atol=1e-8)


def test_chain_rule_matches_brute_force_finite_differences():
    inst = egs.instances.Instance(8, 4, 3)
    model = egs.adjoint.AngleModel(inst)
    A = inst.am.A
    rng = np.random.default_rng(2)
    bits = rng.integers(0, 2, size=(4, 8))                     # step order = qubit order (natural)
    sv = inst.step_vertices

    def lnpr(alpha, beta, b):
        M, _, _ = model.matrices(alpha, beta)
        w = [None] * 8
        for k, q in enumerate(sv):
            w[q] = (M[k][b[k], 0], M[k][b[k], 1])
        return np.log(abs(brute_Z(A, w)) ** 2)

    M, dA, dB = model.matrices(model.alpha0, model.beta0)
    Z = np.zeros(4, complex)
    G = np.zeros((4, 8, 2), complex)
    for i, b in enumerate(bits):
        w = [None] * 8
        for k, q in enumerate(sv):
            w[q] = (M[k][b[k], 0], M[k][b[k], 1])
        Z[i] = brute_Z(A, w)
        for k, q in enumerate(sv):                          # G_k(x) = Z with w_q(x) = 1, w_q(1-x) = 0
            for x in (0, 1):
                w2 = list(w)
                w2[q] = (1.0, 0.0) if x == 0 else (0.0, 1.0)
                G[i, k, x] = brute_Z(A, w2)
    ga, gb = model.grad_ln_pr(bits, Z, G, dA, dB)
    h = 1e-5
    for i in range(4):
        for q in range(8):
            b2 = model.beta0.copy()
            b2[q] += h
            b1 = model.beta0.copy()
            b1[q] -= h
            fd = (lnpr(model.alpha0, b2, bits[i]) - lnpr(model.alpha0, b1, bits[i])) / (2 * h)
            assert abs(fd - gb[i, q]) < 1e-6 * max(1, abs(fd))
        assert np.max(np.abs(ga[i])) < 1e-12             # a final RZ does not change Z-basis probabilities
