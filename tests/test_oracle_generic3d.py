"""Pins of the oracle's full nonlinear RHS on a generic 3-D state.

The other oracle pins use one-dimensional diagonal data (gauge and linear
waves), conformally flat data (punctures) or flat-space term checks, which
leave the off-diagonal, three-dimensional parts of W^2 R_ij, the Gamma~
RHS, the momentum constraint and its derivative open.  Here the state is a
smooth periodic one with every component varying along x, y and z:

* gamma~ = M det(M)^(-1/3) for a non-diagonal symmetric M, so
  det gamma~ = 1 exactly;
* A~ = P - gamma~ (gamma~^kl P_kl) / 3 for a non-diagonal P, so tr A~ = 0
  exactly;
* Gamma~^i = gamma~^jk Gamma~^i_jk + delta^i with a non-zero delta^i
  (the evolved connection off its constraint);
* W, alpha, K, beta^i non-constant, beta^i of both signs.

The exact continuum values are computed from the ADM definitions on the
physical metric gamma_ij = gamma~_ij / W^2, differentiated by torch autograd
in float64 -- not from the W-form term assembly the oracle restates:

* R_ij from the physical Christoffel symbols, R_ij = d_k G^k_ij -
  d_j G^k_ik + G^k_kl G^l_ij - G^k_jl G^l_ik, plus the one place where
  the reference uses the evolved connection instead of the one computed
  from the metric (bssn/geometry.py:67-71: gamma~_k(i d_j) Gamma~^k and
  Gamma~^k Gamma~_(ij)k), i.e. W^2 (gamma~_k(i d_j) delta^k +
  delta^k Gamma~_(ij)k);
* D_i D_j alpha with the physical Christoffel symbols;
* H = R + K^2 - K_ij K^ij and M_i = D_j K^j_i - D_i K with the physical
  K_ij = (A~_ij + gamma~_ij K / 3) / W^2 (bssn/constraints.py:22-97,
  271-303 use the W-form of both), and the constraint-damping term
  kappa/2 alpha (D~_j M_i + D~_i M_j) with the conformal connection
  (bssn/extrinsic_curvature.py:204-218);
* the evolution equations of the W-form BSSN system with the
  single-variable Gamma driver (bssn/spatial_metric.py, bssn/
  conformal_connection.py:19-162, bssn/shift_and_lapse.py:65-132).

The oracle's RHS (fourth-order stencils, KO off) must converge to these
values at fourth order per field: observed orders above 3.5 from 16^3 to
32^3.  A transposed index in any term of the nonlinear assembly leaves an
O(1) error that does not converge.
"""

import math

import numpy as np
import pytest
import torch

from oracle import bssn as OB

TP = 2.0 * math.pi
ORDER_BAR = 3.5


def _modes(rng, nmodes=2):
    """Integer wave vectors in {-1, 0, 1}^3 (non-zero) with random phases."""
    out = []
    while len(out) < nmodes:
        k = rng.integers(-1, 2, 3)
        if np.any(k):
            out.append((k.astype(float), rng.uniform(0, TP)))
    return out


def _field(X, amp, modes):
    f = torch.zeros(X.shape[0], dtype=X.dtype)
    for k, ph in modes:
        f = f + amp * torch.sin(TP * (X @ torch.tensor(k, dtype=X.dtype)) + ph)
    return f / len(modes)


class State:
    """The analytic state as torch functions of the coordinates X (n, 3)."""

    def __init__(self, seed=3):
        rng = np.random.default_rng(seed)
        self.e = {(i, j): (0.08 if i == j else 0.06, _modes(rng)) for i in range(3) for j in range(i, 3)}
        self.P = {(i, j): (0.1, _modes(rng)) for i in range(3) for j in range(i, 3)}
        self.W = (0.1, _modes(rng))
        self.alpha = (0.1, _modes(rng))
        self.K = (0.15, _modes(rng))
        self.beta = [(0.1, _modes(rng)) for _ in range(3)]
        self.delta = [(0.05, _modes(rng)) for _ in range(3)]

    def _sym(self, X, table, diag=0.0):
        n = X.shape[0]
        out = torch.zeros((n, 3, 3), dtype=X.dtype)
        for (i, j), (amp, modes) in table.items():
            v = _field(X, amp, modes) + (diag if i == j else 0.0)
            out[:, i, j] = v
            out[:, j, i] = v
        return out

    def fields(self, X):
        M = self._sym(X, self.e, diag=1.0)
        gt = M * torch.linalg.det(M)[:, None, None] ** (-1.0 / 3.0)
        W = 1.0 + _field(X, *self.W)
        alpha = 1.0 + _field(X, *self.alpha)
        K = _field(X, *self.K)
        beta = torch.stack([_field(X, *b) for b in self.beta], dim=1)
        P = self._sym(X, self.P)
        gti = torch.linalg.inv(gt)
        At = P - gt * torch.einsum('nkl,nkl->n', gti, P)[:, None, None] / 3.0
        delta = torch.stack([_field(X, *b) for b in self.delta], dim=1)
        return dict(gt=gt, gti=gti, W=W, alpha=alpha, K=K, beta=beta, At=At, delta=delta)


def _d(T, X):
    """Gradient of every component of T (n, ...): (n, ..., 3), index last."""
    flat = T.reshape(T.shape[0], -1)
    cols = []
    for c in range(flat.shape[1]):
        g = torch.autograd.grad(flat[:, c].sum(), X, create_graph=True, allow_unused=True)[0]
        cols.append(torch.zeros_like(X) if g is None else g)
    return torch.stack(cols, dim=1).reshape(tuple(T.shape) + (3,))


def _christoffel(g, gi, X):
    """G[n, k, i, j] = Gamma^k_ij of the metric g (n, 3, 3)."""
    dg = _d(g, X)  # dg[n, a, b, c] = d_c g_ab
    t = dg.permute(0, 1, 3, 2) + dg - dg.permute(0, 3, 1, 2)  # [l, i, j] = d_i g_lj + d_j g_li - d_l g_ij
    return 0.5 * torch.einsum('nkl,nlij->nkij', gi, t)


def exact(state, X, kappa, eta, g_driver):
    """Exact RHS fields and constraints at the points X (n, 3)."""
    X = X.clone().requires_grad_(True)
    f = state.fields(X)
    gt, gti, W, alpha, K, beta, At, delta = (f[k] for k in ('gt', 'gti', 'W', 'alpha', 'K', 'beta', 'At', 'delta'))
    # physical metric and its curvature
    g = gt / W[:, None, None] ** 2
    gi = gti * W[:, None, None] ** 2
    G = _christoffel(g, gi, X)
    dG = _d(G, X)  # [n, k, i, j, l] = d_l Gamma^k_ij
    ricci = (torch.einsum('nkijk->nij', dG) - torch.einsum('nkikj->nij', dG)
             + torch.einsum('nkkl,nlij->nij', G, G) - torch.einsum('nkjl,nlik->nij', G, G))
    # conformal connection: computed from the metric, evolved = computed + delta
    Gt = _christoffel(gt, gti, X)
    Gam = torch.einsum('njk,nijk->ni', gti, Gt) + delta
    Gt_low = torch.einsum('nil,nljk->nijk', gt, Gt)  # Gamma~_ijk, first index lowered
    ddelta = _d(delta, X)  # [n, k, j] = d_j delta^k
    corr = 0.5 * (torch.einsum('nki,nkj->nij', gt, ddelta) + torch.einsum('nkj,nki->nij', gt, ddelta)
                  + torch.einsum('nk,nijk->nij', delta, Gt_low) + torch.einsum('nk,njik->nij', delta, Gt_low))
    W2 = (W ** 2)[:, None, None]
    W2R = W2 * (ricci + corr)
    # lapse Hessian
    da = _d(alpha, X)
    dda = _d(da, X)
    W2DDa = W2 * (dda - torch.einsum('nkij,nk->nij', G, da))
    # extrinsic curvature and constraints
    Kp = (At + gt * K[:, None, None] / 3.0) / W2
    Kup = torch.einsum('nik,njl,nkl->nij', gi, gi, Kp)
    H = torch.einsum('nij,nij->n', gti, W2R) + K ** 2 - torch.einsum('nij,nij->n', Kp, Kup)
    Kmix = torch.einsum('njk,nki->nji', gi, Kp)  # K^j_i
    dKmix = _d(Kmix, X)  # [n, j, i, l]
    dK = _d(K, X)
    Mom = (torch.einsum('njij->ni', dKmix) + torch.einsum('njjl,nli->ni', G, Kmix)
           - torch.einsum('nlji,njl->ni', G, Kmix) - dK)
    dM = _d(Mom, X)  # [n, i, l] = d_l M_i
    DM = dM - torch.einsum('nkij,nk->nij', Gt, Mom)  # D~_j M_i at [i, j]
    kap = 0.5 * kappa * alpha[:, None, None] * (DM + DM.transpose(1, 2))
    # evolution equations
    db = _d(beta, X)  # [n, i, j] = d_j beta^i
    div = torch.einsum('nii->n', db)
    dgt = _d(gt, X)
    dAt = _d(At, X)
    dW = _d(W, X)
    dGam = _d(Gam, X)  # [n, i, j] = d_j Gamma~^i
    ddb = _d(db, X)  # [n, i, j, k] = d_k d_j beta^i
    Aup = torch.einsum('nik,njl,nkl->nij', gti, gti, At)
    adv = lambda dF: torch.einsum('nk,n...k->n...', beta, dF)  # noqa: E731

    def tf(Xij):
        return Xij - gt * torch.einsum('nkl,nkl->n', gti, Xij)[:, None, None] / 3.0

    r_gt = (adv(dgt) + torch.einsum('nik,nkj->nij', gt, db) + torch.einsum('njk,nki->nij', gt, db)
            - 2.0 / 3.0 * gt * div[:, None, None] - 2.0 * alpha[:, None, None] * At)
    r_W = adv(dW) + W * (alpha * K - div) / 3.0
    r_K = (adv(dK) - torch.einsum('nij,nij->n', gti, W2DDa)
           + alpha * (torch.einsum('nij,nij->n', At, Aup) + K ** 2 / 3.0))
    r_A = (adv(dAt) + torch.einsum('nik,nkj->nij', At, db) + torch.einsum('njk,nki->nij', At, db)
           - 2.0 / 3.0 * At * div[:, None, None] + tf(alpha[:, None, None] * W2R - W2DDa)
           + alpha[:, None, None] * (K[:, None, None] * At - 2.0 * torch.einsum('nik,nkl,nlj->nij', At, gti, At))
           + kap)
    r_G = (adv(dGam) - torch.einsum('nj,nij->ni', Gam, db) + 2.0 / 3.0 * Gam * div[:, None]
           + torch.einsum('njk,nijk->ni', gti, ddb) + torch.einsum('nij,nkjk->ni', gti, ddb) / 3.0
           - 2.0 * torch.einsum('nij,nj->ni', Aup, da)
           + 2.0 * alpha[:, None] * (torch.einsum('nijk,njk->ni', Gt, Aup)
                                     - 3.0 / W[:, None] * torch.einsum('nij,nj->ni', Aup, dW)
                                     - 2.0 / 3.0 * torch.einsum('nij,nj->ni', gti, dK)))
    r_alpha = adv(da) - 2.0 * alpha * K
    r_beta = g_driver * Gam + torch.einsum('nj,nij->ni', beta, db) - eta * beta
    # detached, so that the autograd graph is freed when this returns
    return dict(fields={k: t.detach() for k, t in f.items()}, Gam=Gam.detach(),
                rhs=tuple(t.detach() for t in (r_gt, r_W, r_A, r_K, r_G, r_alpha, r_beta)),
                W2R=W2R.detach(), W2DDa=W2DDa.detach(), H=H.detach(), M=Mom.detach(), kap=kap.detach())


def _grid(n):
    x = np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    return torch.tensor(np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1))


def _to_field(t, n):
    """(n^3, ...) torch -> (..., n, n, n) numpy."""
    a = t.detach().numpy()
    a = a.reshape((n, n, n) + a.shape[1:])
    return np.moveaxis(a, (0, 1, 2), (-3, -2, -1))


def _exact_chunked(state, X, kappa, eta, g_driver, chunk=2048):
    """exact() over blocks of points (its autograd graph takes about 2 MB
    per point), results joined."""
    parts = [exact(state, X[i:i + chunk], kappa, eta, g_driver) for i in range(0, X.shape[0], chunk)]

    def join(get):
        return torch.cat([get(p) for p in parts])
    return dict(fields={k: join(lambda p, k=k: p['fields'][k]) for k in parts[0]['fields']},
                Gam=join(lambda p: p['Gam']),
                rhs=tuple(join(lambda p, i=i: p['rhs'][i]) for i in range(len(parts[0]['rhs']))),
                **{k: join(lambda p, k=k: p[k]) for k in ('W2R', 'W2DDa', 'H', 'M', 'kap')})


def _case(state, n, kappa=0.3, eta=0.4, g_driver=0.75):
    X = _grid(n)
    ex = _exact_chunked(state, X, kappa, eta, g_driver)
    f = ex['fields']
    v = OB.BSSNVariables(_to_field(f['gt'], n), _to_field(f['W'], n), _to_field(f['At'], n),
                         _to_field(f['K'], n), _to_field(ex['Gam'], n), _to_field(f['alpha'], n),
                         _to_field(f['beta'], n))
    p = OB.BSSNParameters(eta=eta, kappa=kappa, nu=0.0, g=g_driver, dx=1.0 / n, dt=0.1 / n, zero_shift=0,
                          gauge=1)
    return v, p, ex


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def cases():
    torch.set_grad_enabled(True)
    state = State()
    return {n: _case(state, n) for n in (16, 32)}


def _orders(errs):
    return {k: math.log2(errs[16][k] / errs[32][k]) for k in errs[16]}


def test_analytic_state_is_generic(cases):
    """The state exercises what the pins are for: off-diagonal metric and
    A~, exact algebraic constraints, a connection off its constraint."""
    v, p, ex = cases[32]
    g = v.conformal_metric
    assert np.abs(g[0, 1]).max() > 0.01 and np.abs(g[1, 2]).max() > 0.01
    assert np.abs(np.linalg.det(np.moveaxis(g, (0, 1), (-2, -1))) - 1).max() < 1e-13
    assert np.abs(OB.compute_trace_A_violation(v)).max() < 1e-13
    assert np.abs(v.traceless_K[0, 2]).max() > 0.01
    assert np.abs(v.shift).max() > 0.02 and v.shift.min() < 0 < v.shift.max()
    dGam = OB.compute_gamma_constraint(v, p)
    assert np.abs(dGam).max() > 0.01


def test_rhs_converges_to_exact_at_fourth_order(cases):
    """Every field of compute_bssn_rhs (kappa = 0.3, so the momentum
    constraint derivative enters the A~ RHS)."""
    errs = {}
    for n, (v, p, ex) in cases.items():
        got = OB.compute_bssn_rhs(v, p)
        errs[n] = {name: _rel(a, _to_field(b, n)) for name, a, b in zip(v._fields, got, ex['rhs'])}
    orders = _orders(errs)
    print("errors", errs, "orders", orders)
    for k, o in orders.items():
        assert o > ORDER_BAR, (k, errs[16][k], errs[32][k], o)


def test_geometry_and_constraints_converge_to_exact(cases):
    """W^2 R_ij, W^2 D_i D_j alpha, H, M_i, the constraint-damping term and
    the Gamma~ constraint, each from its own oracle entry point."""
    errs = {}
    for n, (v, p, ex) in cases.items():
        ricci = OB.compute_W2_ricci(v, p)
        hess = OB.compute_W2_covariant_lapse_hessian(v, p)
        M, dM = OB.compute_momentum_constraint_and_derivative(v, p)
        c2 = OB.christoffel_symbols_second_kind(
            OB.invert_3x3_metric(v.conformal_metric),
            np.stack([OB.diff1_physical(v.conformal_metric, d + 2, p) for d in range(3)]))
        DM = dM - np.einsum('kij...,k...->ij...', c2, M)
        kap = p.kappa / 2 * v.lapse * (DM + np.swapaxes(DM, 0, 1))
        errs[n] = {
            "W2R": _rel(ricci, _to_field(ex['W2R'], n)),
            "W2DDalpha": _rel(hess, _to_field(ex['W2DDa'], n)),
            "H": _rel(OB.compute_hamiltonian_constraint(v, p), _to_field(ex['H'], n)),
            "M": _rel(OB.compute_momentum_constraint(v, p), _to_field(ex['M'], n)),
            "kappa_term": _rel(kap, _to_field(ex['kap'], n)),
            "gamma_constraint": _rel(OB.compute_gamma_constraint(v, p), _to_field(ex['fields']['delta'], n)),
        }
    orders = _orders(errs)
    print("errors", errs, "orders", orders)
    for k, o in orders.items():
        assert o > ORDER_BAR, (k, errs[16][k], errs[32][k], o)
