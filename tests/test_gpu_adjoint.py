"""The parameter adjoint (mgpu_adjoint) against central finite differences of
the FP64 forward (h = 1e-6 relative), for every potential form and for the
charges: dE/dtheta, the cotangent product sum_i v_i . dF_i/dtheta and the
directional derivative of the site forces.  A torch.autograd.Function over
the C ABI (test harness only) gives the force-matching gradient, checked the
same way."""
import numpy as np
import pytest

import systems_small as S
from moldy_amd import system

pytestmark = pytest.mark.gpu

H_REL = 1e-6
TOL = 2e-6


def _cases():
    return {
        "silica": lambda: S.silica(),
        "water": lambda: S.water(n=64, cutoff=6.0),
        "salt": lambda: S.salt(),
        "argon": lambda: S.argon(),
        "mcy": lambda: S.two_species_generic(2, n=60),
        "generic": lambda: S.two_species_generic(3, n=60),
        "hiw": lambda: S.two_species_generic(4, n=60),
        "morse": lambda: S.two_species_generic(6, n=60),
    }


class Forward:
    """E and site forces of eval_forces at a given theta (FP64)."""

    def __init__(self, s):
        from moldy_amd import engine as E
        self.E = E
        self.s = s
        self.eng = E.ForceEngine(s)

    def __call__(self, theta):
        s2 = self.E.set_theta(self.s, theta)
        self.eng.rebind(s2)
        self.eng.upload_state()
        pe, _, _ = self.eng.eval_forces_device()
        _, _, _, sf = self.eng.download()
        return pe.sum(), sf

    def adjoint(self, theta, v=None, dtheta=None):
        s2 = self.E.set_theta(self.s, theta)
        self.eng.rebind(s2)
        self.eng.upload_state()
        return self.eng.adjoint(v=v, dtheta=dtheta)


def _steps(theta):
    # 1e-6 relative; parameters that are zero (pair types without that term,
    # uncharged ids) enter linearly or quadratically, where any step is exact
    return np.where(theta != 0.0, H_REL * np.abs(theta), 1e-2)


@pytest.mark.parametrize("case", list(_cases()))
def test_adjoint_vs_central_differences(case):
    from moldy_amd import engine as E
    s = _cases()[case]()
    fw = Forward(s)
    th = E.get_theta(s)
    rng = np.random.Generator(np.random.Philox(17))
    v = rng.standard_normal((3, s.nsites))
    dth = rng.standard_normal(th.size) * np.abs(th)
    got = fw.adjoint(th, v=v, dtheta=dth)
    h = _steps(th)
    fd_e = np.zeros(th.size)
    fd_v = np.zeros(th.size)
    for t in range(th.size):
        tp, tm = th.copy(), th.copy()
        tp[t] += h[t]
        tm[t] -= h[t]
        (ep, fp), (em, fm) = fw(tp), fw(tm)
        fd_e[t] = (ep - em) / (2 * h[t])
        fd_v[t] = (v * (fp - fm)).sum() / (2 * h[t])
    eps = 1e-6
    (_, fp), (_, fm) = fw(th + eps * dth), fw(th - eps * dth)
    fd_j = (fp - fm) / (2 * eps)
    ee = np.abs(got["dE"] - fd_e).max() / np.abs(fd_e).max()
    ev = np.abs(got["vjp"] - fd_v).max() / np.abs(fd_v).max()
    ej = np.linalg.norm(got["jvp"] - fd_j) / np.linalg.norm(fd_j)
    print(f"{case}: ntheta {th.size}  dE {ee:.1e}  vjp {ev:.1e}  jvp {ej:.1e}")
    assert ee < TOL and ev < TOL and ej < TOL
    fw.eng.close()


def test_force_matching_gradient_through_autograd():
    """Silica-like crystal, Buckingham: loss = |F(theta) - F_target|^2 with
    the target forces of a perturbed parameter set."""
    import torch
    from moldy_amd import engine as E
    s = S.silica()
    fw = Forward(s)
    th0 = E.get_theta(s)
    rng = np.random.Generator(np.random.Philox(5))
    th_true = th0 * (1 + 0.02 * rng.standard_normal(th0.size))
    _, F_target = fw(th_true)

    class SiteForces(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            t = theta.detach().cpu().numpy()
            ctx.theta = t
            return torch.from_numpy(fw(t)[1])

        @staticmethod
        def backward(ctx, g):
            out = fw.adjoint(ctx.theta, v=g.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(out["vjp"])

    theta = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    target = torch.from_numpy(F_target)
    loss = ((SiteForces.apply(theta) - target) ** 2).sum()
    loss.backward()
    g = theta.grad.numpy()
    h = _steps(th0)
    fd = np.zeros(th0.size)
    for t in range(th0.size):
        tp, tm = th0.copy(), th0.copy()
        tp[t] += h[t]
        tm[t] -= h[t]
        fd[t] = (((fw(tp)[1] - F_target) ** 2).sum() - ((fw(tm)[1] - F_target) ** 2).sum()) / (2 * h[t])
    err = np.abs(g - fd).max() / np.abs(fd).max()
    print(f"force matching: |grad| {np.abs(fd).max():.3e}  rel err {err:.1e}")
    assert err < TOL
    fw.eng.close()
