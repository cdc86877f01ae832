"""Child process of tests/test_gpu_adjoint.py: wraps the device compliance / adjoint gradient
in a torch.autograd.Function (the only place torch is used) and runs torch.autograd.gradcheck.
torch is imported before libmgx.so is loaded, so one HIP runtime serves both."""
import os
import sys

import torch  # noqa: E402  (first: see compressiontests_amd/_lib.py)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from compressiontests_amd.adjoint import compliance_and_gradient  # noqa: E402

DIMS = (4, 4, 4)


class Compliance(torch.autograd.Function):
    """C(rho) = f^T K(rho)^-1 f with the device adjoint as its backward."""

    @staticmethod
    def forward(ctx, rho):
        C, g, _, st = compliance_and_gradient(DIMS, rho.detach().cpu().numpy(), tolerance=1e-14, levels=2)
        if not st.HasConverged:
            raise RuntimeError("solve did not converge")
        ctx.save_for_backward(torch.from_numpy(g))
        return rho.new_tensor(C)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g


def main():
    rng = np.random.default_rng(7)
    rho = torch.tensor(0.1 + 0.9 * rng.random(int(np.prod(DIMS))), dtype=torch.float64, requires_grad=True)
    ok = torch.autograd.gradcheck(Compliance.apply, (rho,), eps=1e-6, atol=1e-7, rtol=1e-6)
    C = Compliance.apply(rho)
    C.backward()
    assert torch.all(rho.grad <= 0)
    print("gradcheck ok" if ok else "gradcheck FAILED", float(C))


if __name__ == "__main__":
    main()
