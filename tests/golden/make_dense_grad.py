"""Dense-statevector FP64 reference gradient (torch autograd on the CPU) for the adjoint test:
the n = 20, band 8 banded instance (seed 44), 64 seeded strings, loss = mean -ln Pr(b) over them, gradients with
respect to every tail angle alpha_q, beta_q.  The tail of qubit q is rz(beta_q) h rz(alpha_q) (grad_n20.npz; alpha
only phases the outcomes, its gradient is 0) and rz(beta_q) h rz(alpha_q) h (grad_n20_live.npz; both angles live).
Written to tests/golden/c4dense/ (numeric arrays only)."""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _pkg import egs  # noqa: E402
import dense  # noqa: E402


def apply1(psi, U, q, n):
    psi = psi.reshape([2] * n)
    psi = torch.movedim(torch.tensordot(U, psi, dims=([1], [q])), 0, q)
    return psi.reshape(-1)


def main(last_h=False):
    n, band, seed = 20, 8, 44
    inst = egs.instances.Instance(n, band, seed)
    model = egs.adjoint.AngleModel(inst, last_h=last_h)
    psi0 = torch.from_numpy(dense.simulate(inst.N, inst.body))
    alpha = torch.tensor(model.alpha0, dtype=torch.float64, requires_grad=True)
    beta = torch.tensor(model.beta0, dtype=torch.float64, requires_grad=True)
    H = torch.tensor([[1, 1], [1, -1]], dtype=torch.complex128) / np.sqrt(2)

    def rz(t):
        e = torch.exp(-0.5j * t.to(torch.complex128))
        return torch.diag(torch.stack([e, 1 / e]))

    psi = psi0
    for q in range(n):
        psi = apply1(psi, rz(beta[q]), q, n)
        psi = apply1(psi, H, q, n)
        psi = apply1(psi, rz(alpha[q]), q, n)
        if last_h:
            psi = apply1(psi, H, q, n)
    rng = np.random.default_rng(9)
    bits = rng.integers(0, 2, size=(64, n)).astype(np.int8)      # step order = qubit order (natural order)
    idx = torch.tensor([int("".join(map(str, b)), 2) for b in bits])
    lnp = torch.log(torch.abs(psi[idx]) ** 2)
    loss = -lnp.mean()
    loss.backward()
    np.savez(os.path.join(HERE, "c4dense", "grad_n20_live.npz" if last_h else "grad_n20.npz"), alpha=model.alpha0,
             beta=model.beta0, bits=bits,
             lnp=lnp.detach().numpy(), grad_alpha=alpha.grad.numpy(), grad_beta=beta.grad.numpy(),
             n=n, band=band, seed=seed)
    print("loss", float(loss), "max |grad|", float(alpha.grad.abs().max()), float(beta.grad.abs().max()))


if __name__ == "__main__":
    main()
    main(last_h=True)          # the tail rz(beta) h rz(alpha) h: both angles live
