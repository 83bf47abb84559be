"""Exact central differences for the differentiable-solve loss.

The anisotropic diffusion problem (generators.anisotropic2d) builds A from
per-edge conductivities theta.  Each theta_e enters A as a rank-one term
c u u^T, with u = e_a - e_b for an interior edge between cells a and b, or
u = e_a for an edge to the Dirichlet exterior; an upwind convection part of
A is constant in theta.  So A(theta +- h e_i) = A0 +- h c u u^T, and by the
Sherman-Morrison identity

    x(theta +- h e_i) = x0 - (+-hc) (u^T x0) w / (1 +- hc s),
    w = A0^{-1} u,  s = u^T w.

For L = 1/2 ||X - X_t||^2 over the right-hand-side columns k, the central
difference [L(theta + h e_i) - L(theta - h e_i)] / 2h then has the closed
form

    -(c / D) sum_k (u^T x0_k) [w . (x0_k - xt_k) + (hc)^2 s (u^T x0_k) w.w / D],
    D = 1 - (hcs)^2,

which needs no symmetry of A0.  Subtracting two separately solved losses
instead loses about eps |L| / h to rounding, which at h = 1e-6 and |L| ~ 1e10
swamps the difference even with exact solves.  The w come from forward
solves with A0 (not A0^T), so the check is independent of the adjoint.
"""
import numpy as np
import torch

from . import generators as gen


def edge_rank_one(rows_of, vmap, coef, e):
    """The entries of A that theta[e] scales, as (support rows, signs, c):
    an interior edge gives (a,a,+c), (b,b,+c), (a,b,-c), (b,a,-c); an edge to
    the exterior gives (a,a,+c).  Raises if theta[e] enters A otherwise.
    vmap rows are (value index, theta index, column of the value)."""
    ks = np.where(vmap[:, 1] == e)[0]
    ent = sorted((int(rows_of[vmap[k, 0]]), int(vmap[k, 2]), float(coef[k])) for k in ks)
    diag = [(r, c, w) for r, c, w in ent if r == c]
    off = [(r, c, w) for r, c, w in ent if r != c]
    cval = diag[0][2]
    if len(ent) == 1 and cval > 0:
        return [diag[0][0]], [1.0], cval
    a, b = diag[0][0], diag[1][0]
    ok = (len(ent) == 4 and len(diag) == 2 and diag[1][2] == cval and
          sorted((r, c) for r, c, _ in off) == sorted([(a, b), (b, a)]) and
          all(w == -cval for _, _, w in off))
    if not ok:
        raise ValueError(f"theta[{e}] does not enter A as a rank-one term")
    return [a, b], [1.0, -1.0], cval


class AdjointProblem:
    """The nx^2 anisotropic diffusion problem with nrhs right-hand sides on a
    device: conductivities theta = exp(g), g ~ N(0, 0.5^2) (seed 6), b and
    x_target ~ U[-1, 1] (seed 7), optional upwind convection (constant in
    theta).  values_of(theta) scatters theta into the CSR value array with
    torch ops, so gradients reach theta."""

    def __init__(self, nx, nrhs, device, convection=0.0):
        rp, ci, vals, theta, vmap, coef = gen.anisotropic2d(nx, seed=6, convection=convection)
        self.nx, self.n, self.nnz = nx, nx * nx, len(ci)
        self.rp, self.ci, self.vmap, self.coef, self.theta = rp, ci, vmap, coef, theta
        self.device = device
        self.vmap_t = torch.as_tensor(vmap).to(device)
        self.coef_t = torch.as_tensor(coef).to(device)
        self.theta0 = torch.as_tensor(theta).to(device)
        rng = np.random.default_rng(7)
        self.b = torch.as_tensor(rng.uniform(-1, 1, (self.n, nrhs))).to(device)
        self.xt = torch.as_tensor(rng.uniform(-1, 1, (self.n, nrhs))).to(device)
        self.rp_t = torch.as_tensor(rp).to(device)
        self.ci_t = torch.as_tensor(ci).to(device)
        # the part of the values that theta does not scale (the convection);
        # without convection it is zero (the generator's values differ from
        # the scatter only by the rounding of another summation order), so A
        # stays exactly symmetric
        self.const = torch.as_tensor(vals).to(device) - self._scatter(self.theta0, None)
        if convection == 0.0:
            self.const.zero_()

    def _scatter(self, t, base):
        z = torch.zeros(self.nnz, dtype=torch.float64, device=self.device) if base is None \
            else base
        return z.index_add(0, self.vmap_t[:, 0], self.coef_t * t[self.vmap_t[:, 1]])

    def values_of(self, theta):
        return self._scatter(theta, self.const)

    @property
    def symmetric(self):
        return bool(torch.all(self.const == 0).item())

    def loss(self, x):
        return 0.5 * ((x - self.xt) ** 2).sum()

    def central_differences(self, solve, x0, entries, h):
        """Exact central differences of the loss at theta0 for the theta
        indices `entries`; solve(U) must return A0^{-1} U for an n x m
        block U (a forward solve)."""
        rows_of = np.repeat(np.arange(self.n), np.diff(self.rp))
        vm = np.stack([self.vmap[:, 0], self.vmap[:, 1], self.ci[self.vmap[:, 0]]], axis=1)
        U = torch.zeros((self.n, len(entries)), dtype=torch.float64, device=self.device)
        shapes = []
        for m, e in enumerate(entries):
            sup, sgn, c = edge_rank_one(rows_of, vm, self.coef, int(e))
            for r_, s_ in zip(sup, sgn):
                U[r_, m] = s_
            shapes.append((sup, sgn, c))
        W = solve(U)
        x0 = x0.detach()
        fd = []
        for m, (sup, sgn, c) in enumerate(shapes):
            w = W[:, m]
            s_ = float(sum(sg * w[r_].item() for r_, sg in zip(sup, sgn)))
            ux = sum(sg * x0[r_, :] for r_, sg in zip(sup, sgn))        # u^T x0_k
            wr = (w[:, None] * (x0 - self.xt)).sum(0)                   # w.(x0_k - xt_k)
            ww = float((w * w).sum().item())
            hc = h * c
            den = 1.0 - (hc * s_) ** 2
            fd.append(float((-(c / den) * (ux * (wr + hc * hc * s_ * ux * ww / den))).sum()))
        return np.array(fd)
