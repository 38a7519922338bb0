"""Gradients of log Pr(b) with respect to the final rotation layer (no counterpart in the reference).

For an instance whose qubit q ends with the tail ``rz(beta_q) h rz(alpha_q)`` (``instances.py``), the measurement
matrix of q is M_q = W_q U_q^dagger with W_q = RZ(alpha_q) H RZ(beta_q) (amp.py:26 with the tail product of
amp.py:16) and U_q from the graph form of the Clifford body.  RZ(alpha_q) is then the last gate before the
measurement in the Z basis: it multiplies each outcome's amplitude by a phase, so d ln Pr / d alpha_q = 0 for every
string, and only the 60 beta angles of the n = 60 workload carry a gradient.  ``AngleModel(inst, last_h=True)``
models the tail ``rz(beta_q) h rz(alpha_q) h`` instead (W_q = H RZ(alpha_q) H RZ(beta_q): alpha acts as an X rotation
before the measurement), where both angles are live.  The engine's adjoint returns, per string and step k,
G_k(x) = dZ/dM_k[b_k, x] (``gs_batch_grad``); with Pr(b) proportional to |Z|^2,

    d ln Pr(b) / d theta = 2 Re( sum_x dM_q[b_q, x]/d theta * G_{k(q)}(x) / Z ),

for theta in {alpha_q, beta_q}, dRZ(t)/dt = RZ(t) diag(-i/2, i/2).

``LogProb`` wraps this as a ``torch.autograd.Function`` over CPU tensors of angles (torch is imported lazily and only
handles the host-side bookkeeping; the sweeps run on the GPU through the C ABI).
"""
import re

import numpy as np

from .amp import tail_matrix
from .graphform import graph_form

_RZ = re.compile(r"rz\((.*)\)")
_H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)


def _rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


def _drz(t):
    return np.diag([-0.5j * np.exp(-0.5j * t), 0.5j * np.exp(0.5j * t)])


class AngleModel:
    """Measurement matrices of an ``instances.Instance`` as functions of its tail angles (alpha, beta per data
    qubit; ancilla tails, if any, are fixed).  The angles are float32 when the instance's are (``fp32_angles``);
    the matrices and derivatives are computed in float64 from those values.  ``last_h``: the tail is followed by
    one more ``h`` (see the module docstring)."""

    def __init__(self, inst, last_h=False):
        self.inst = inst
        self.last_h = last_h
        _, U = graph_form(_tableau(inst))
        self.Udag = [u.conj().T for u in U]
        self.n = inst.n
        self.N = inst.N
        self.step_vertices = inst.step_vertices
        dtype = np.float32 if getattr(inst, "fp32_angles", False) else np.float64
        self.alpha0 = np.zeros(self.n, dtype)
        self.beta0 = np.zeros(self.n, dtype)
        per = {}
        for name, qs in inst.tail:
            per.setdefault(qs[0], []).append(name)
        for q in range(self.n):
            names = per[q]
            self.beta0[q] = float(_RZ.match(names[0]).group(1))
            self.alpha0[q] = float(_RZ.match(names[2]).group(1))
        self.fixed = {q: tail_matrix([(nm, [q]) for nm in per[q]]) @ self.Udag[q] for q in range(self.n, self.N)}

    def matrices(self, alpha, beta):
        """``(M [ns, 2, 2], dM/dalpha [ns, 2, 2], dM/dbeta [ns, 2, 2])`` in step order (zero derivatives for ancillas)."""
        ns = len(self.step_vertices)
        M = np.empty((ns, 2, 2), complex)
        dA = np.zeros((ns, 2, 2), complex)
        dB = np.zeros((ns, 2, 2), complex)
        for k, q in enumerate(self.step_vertices):
            if q >= self.n:
                M[k] = self.fixed[q]
                continue
            a, b = float(alpha[q]), float(beta[q])
            M[k] = _rz(a) @ _H @ _rz(b) @ self.Udag[q]
            dA[k] = _drz(a) @ _H @ _rz(b) @ self.Udag[q]
            dB[k] = _rz(a) @ _H @ _drz(b) @ self.Udag[q]
            if self.last_h:
                M[k], dA[k], dB[k] = _H @ M[k], _H @ dA[k], _H @ dB[k]
        return M, dA, dB

    def grad_ln_pr(self, bits, Z, G, dA, dB):
        """Per-string gradients of ln Pr(b) w.r.t. (alpha [B, n], beta [B, n]) from the engine's G [B, ns, 2]."""
        bits = np.asarray(bits)
        B = bits.shape[0]
        ga = np.zeros((B, self.n))
        gb = np.zeros((B, self.n))
        for k, q in enumerate(self.step_vertices):
            if q >= self.n:
                continue
            bk = bits[:, k].astype(int)
            # sum_x dM[b_k, x] G_k(x)
            da = dA[k][bk, 0] * G[:, k, 0] + dA[k][bk, 1] * G[:, k, 1]
            db = dB[k][bk, 0] * G[:, k, 0] + dB[k][bk, 1] * G[:, k, 1]
            ga[:, q] = 2 * np.real(da / Z)
            gb[:, q] = 2 * np.real(db / Z)
        return ga, gb


def _tableau(inst):
    from .instances import tableau_of
    return tableau_of(inst.N, inst.body)


def make_log_prob(engine, model, plan_index=0):
    """A ``torch.autograd.Function`` ``LogProb.apply(alpha, beta, bits) -> ln Pr [B]`` evaluated by a
    ``gsamp.BatchEngine`` (instance ``plan_index``) and differentiated by its adjoint."""
    import torch

    class LogProb(torch.autograd.Function):
        @staticmethod
        def forward(ctx, alpha, beta, bits):
            a = alpha.detach().cpu().numpy().astype(np.float64)
            b = beta.detach().cpu().numpy().astype(np.float64)
            M, dA, dB = model.matrices(a, b)
            engine.set_matrices(plan_index, M)
            bb = bits.cpu().numpy().astype(np.int8)
            lp, _, _ = engine.score([bb], want_cond=False, want_amp=False)
            ctx.save_for_backward(alpha, beta, bits)
            return torch.as_tensor(lp[0] * np.log(2.0), dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            alpha, beta, bits = ctx.saved_tensors
            a = alpha.detach().cpu().numpy().astype(np.float64)
            b = beta.detach().cpu().numpy().astype(np.float64)
            M, dA, dB = model.matrices(a, b)
            engine.set_matrices(plan_index, M)
            bb = bits.cpu().numpy().astype(np.int8)
            res, _ = engine.grad([bb])
            _, Z, G = res[0]
            ga, gb = model.grad_ln_pr(bb, Z, G, dA, dB)
            w = gout.detach().cpu().numpy()[:, None]
            return (torch.as_tensor((w * ga).sum(0), dtype=alpha.dtype),
                    torch.as_tensor((w * gb).sum(0), dtype=beta.dtype), None)

    return LogProb
