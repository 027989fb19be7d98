"""The oracle's gauge-link functions pinned by closed forms.

The reference's tests hold no displaced or smeared data (test_all.ini:38 has
the displacement commented out), so displace, stout_step and the displaced
branch of build_vdaggerv are pinned here by what the mathematics fixes:
unit links reduce a displacement to a lattice shift, a forward step undone by
a backward one is the identity, VdaggerV of displaced operators is gauge
invariant, stout smearing is gauge covariant, keeps links in SU(3), is the
identity at rho = 0, and on commuting (diagonal) links reduces to the phase
update theta' = theta + Im w - (1/3) sum Im w with w = (rho / 4) sum of the
staple phases e^{i(phi - theta)}, computed here by explicit coordinate loops.
The SU(3) projection and the HYP step (hyp_step) likewise: fixed points,
unitarity and determinant 1, gauge covariance, zero parameters as the
identity, and the per-component phase form on diagonal links.
"""
import itertools

import numpy as np
import pytest

from oracle import slaph_oracle as o

L = (3, 4, 2)
SITES = L[0] * L[1] * L[2]


def _su3(rng, n, eps=1.0):
    """n random SU(3) matrices exp(i eps H), H traceless Hermitian."""
    import scipy.linalg
    x = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    h = 0.5 * (x + x.conj().swapaxes(1, 2))
    h -= (np.trace(h, axis1=1, axis2=2) / 3.0)[:, None, None] * np.eye(3)
    return np.stack([scipy.linalg.expm(1j * eps * m) for m in h])


def _links(seed, eps=0.4):
    rng = np.random.default_rng(seed)
    return _su3(rng, SITES * 3, eps).reshape(SITES, 3, 3, 3)


def _vecs(seed, N=5):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((3 * SITES, N)) + 1j * rng.standard_normal((3 * SITES, N))


def _site(x, y, z):
    return (x * L[1] + y) * L[2] + z


def test_unit_links_displacement_is_a_lattice_shift():
    V = _vecs(1)
    unit = np.broadcast_to(np.eye(3), (SITES, 3, 3, 3)).astype(complex)
    grid = V.reshape(L[0], L[1], L[2], 3, -1)
    for mu, axis in enumerate("xyz"):
        fwd = o.displace(V, unit, ((">", axis),), L)
        bwd = o.displace(V, unit, (("<", axis),), L)
        # W(x) = V(x + mu) and V(x - mu)
        assert np.array_equal(fwd, np.roll(grid, -1, axis=mu).reshape(V.shape))
        assert np.array_equal(bwd, np.roll(grid, 1, axis=mu).reshape(V.shape))


def test_displacement_explicit_loops():
    U, V = _links(2), _vecs(3)
    Vs = V.reshape(SITES, 3, -1)
    for mu, axis in enumerate("xyz"):
        fwd = o.displace(V, U, ((">", axis),), L).reshape(SITES, 3, -1)
        bwd = o.displace(V, U, (("<", axis),), L).reshape(SITES, 3, -1)
        for x, y, z in itertools.product(range(L[0]), range(L[1]), range(L[2])):
            c = [x, y, z]
            up, dn = list(c), list(c)
            up[mu] = (c[mu] + 1) % L[mu]
            dn[mu] = (c[mu] - 1) % L[mu]
            s, su, sd = _site(*c), _site(*up), _site(*dn)
            for r in range(3):
                want_f = sum(U[s, mu, r, k] * Vs[su, k] for k in range(3))
                want_b = sum(np.conj(U[sd, mu, k, r]) * Vs[sd, k] for k in range(3))
                assert np.allclose(fwd[s, r], want_f, atol=1e-14)
                assert np.allclose(bwd[s, r], want_b, atol=1e-14)


def test_forward_then_backward_is_identity():
    U, V = _links(4), _vecs(5)
    for axis in "xyz":
        W = o.displace(V, U, ((">", axis), ("<", axis)), L)
        assert np.max(np.abs(W - V)) < 1e-13


def _gauge_transform(U, g):
    """U'_mu(x) = g(x) U_mu(x) g(x + mu)^+."""
    out = np.empty_like(U)
    for mu in range(3):
        nb = o.neighbour(L, mu, 1)
        out[:, mu] = g @ U[:, mu] @ g[nb].conj().swapaxes(1, 2)
    return out


@pytest.mark.parametrize("disp", [((">", "x"),), (("<", "y"),), ((">", "z"), ("<", "x")), ((">", "y"), (">", "y"))])
def test_displaced_vdaggerv_is_gauge_invariant(disp):
    U, V = _links(6), _vecs(7)
    g = _su3(np.random.default_rng(8), SITES)
    Vg = (g @ V.reshape(SITES, 3, -1)).reshape(V.shape)
    a = V.conj().T @ o.displace(V, U, disp, L)
    b = Vg.conj().T @ o.displace(Vg, _gauge_transform(U, g), disp, L)
    assert np.max(np.abs(a - b)) < 1e-12
    # a wrong link orientation breaks it
    a_bad = V.conj().T @ o.displace(V, U.conj().swapaxes(2, 3), disp, L)
    b_bad = Vg.conj().T @ o.displace(Vg, _gauge_transform(U, g).conj().swapaxes(2, 3), disp, L)
    assert np.max(np.abs(a_bad - b_bad)) > 1e-3


def test_build_vdaggerv_displaced_branch_drops_the_phase():
    cfg = {"Lx": L[0], "Ly": L[1], "Lz": L[2], "number_of_eigen_vec": 4}
    V = np.linalg.qr(_vecs(9, 4))[0]
    U = _links(10)
    lut = {"vdv_lookup": [(0, (0, 0, 0), ()), (1, (1, 0, 0), ((">", "x"),)), (2, (0, 0, 0), ((">", "x"),))],
           "index_of_unity": 0}
    got = o.build_vdaggerv(cfg, lut, [V], links=[U])
    assert np.array_equal(got[(0, 0)], np.eye(4))
    want = V.conj().T @ o.displace(V, U, ((">", "x"),), L)
    assert np.array_equal(got[(1, 0)], want)
    assert np.array_equal(got[(2, 0)], want)
    with pytest.raises(RuntimeError):
        o.build_vdaggerv(cfg, lut, [V])


def test_stout_rho_zero_is_identity_and_keeps_su3():
    U = _links(11)
    assert np.max(np.abs(o.stout_step(U, 0.0, L) - U)) < 1e-14
    S = U
    for _ in range(3):
        S = o.stout_step(S, 0.1, L)
    m = S.reshape(-1, 3, 3)
    assert np.max(np.abs(m @ m.conj().swapaxes(1, 2) - np.eye(3))) < 1e-13
    assert np.max(np.abs(np.linalg.det(m) - 1.0)) < 1e-13
    assert np.max(np.abs(S - U)) > 1e-3


def test_stout_is_gauge_covariant():
    U = _links(12)
    g = _su3(np.random.default_rng(13), SITES)
    a = _gauge_transform(o.stout_step(U, 0.1, L), g)
    b = o.stout_step(_gauge_transform(U, g), 0.1, L)
    assert np.max(np.abs(a - b)) < 1e-13


def test_stout_on_diagonal_links_matches_the_phase_update():
    rng = np.random.default_rng(14)
    th = rng.uniform(-np.pi, np.pi, (SITES, 3, 2))
    theta = np.concatenate([th, -th.sum(axis=2, keepdims=True)], axis=2)  # traceless: det 1
    U = np.zeros((SITES, 3, 3, 3), complex)
    for k in range(3):
        U[:, :, k, k] = np.exp(1j * theta[:, :, k])
    rho = 0.3
    got = o.stout_step(U, rho, L)

    def at(c, mu, k):
        return theta[_site(*[c[i] % L[i] for i in range(3)]), mu, k]

    for x, y, z in itertools.product(range(L[0]), range(L[1]), range(L[2])):
        c = (x, y, z)
        for mu in range(3):
            e = [0, 0, 0]
            e[mu] = 1
            im_w = np.zeros(3)
            for k in range(3):
                w = 0j
                for nu in range(3):
                    if nu == mu:
                        continue
                    f = [0, 0, 0]
                    f[nu] = 1
                    cp = [c[i] + f[i] for i in range(3)]
                    cm = [c[i] + e[i] for i in range(3)]
                    cd = [c[i] - f[i] for i in range(3)]
                    cdm = [c[i] - f[i] + e[i] for i in range(3)]
                    up = at(c, nu, k) + at(cp, mu, k) - at(cm, nu, k)
                    dn = -at(cd, nu, k) + at(cd, mu, k) + at(cdm, nu, k)
                    w += np.exp(1j * (up - at(c, mu, k))) + np.exp(1j * (dn - at(c, mu, k)))
                im_w[k] = (rho / 4.0 * w).imag
            for k in range(3):
                want = at(c, mu, k) + im_w[k] - im_w.sum() / 3.0
                assert abs(got[_site(*c), mu, k, k] - np.exp(1j * want)) < 1e-13
    # off-diagonal parts stay zero
    off = got.copy()
    for k in range(3):
        off[:, :, k, k] = 0
    assert np.max(np.abs(off)) < 1e-15


def test_proj_su3_closed_forms():
    """proj_su3: an SU(3) matrix is fixed; c U (c > 0) -> U; the result is
    unitary with determinant 1; diagonal diag(z) -> diag(z/|z|) / (prod z/|z|)^(1/3)."""
    U = _links(15).reshape(-1, 3, 3)[:20]
    assert np.max(np.abs(o.proj_su3(U) - U)) < 1e-14
    assert np.max(np.abs(o.proj_su3(2.5 * U) - U)) < 1e-14
    rng = np.random.default_rng(16)
    A = rng.standard_normal((20, 3, 3)) + 1j * rng.standard_normal((20, 3, 3))
    P = o.proj_su3(A)
    assert np.max(np.abs(P @ P.conj().swapaxes(1, 2) - np.eye(3))) < 1e-13
    assert np.max(np.abs(np.linalg.det(P) - 1.0)) < 1e-13
    z = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    ph = z / np.abs(z)
    want = np.diag(ph) / np.prod(ph) ** (1.0 / 3.0)
    assert np.max(np.abs(o.proj_su3(np.diag(z)[None])[0] - want)) < 1e-14


def test_hyp_zero_alphas_identity_su3_and_covariance():
    U = _links(17)
    assert np.max(np.abs(o.hyp_step(U, 0.0, 0.0, L) - U)) < 1e-13
    S = o.hyp_step(U, 0.75, 0.6, L)
    m = S.reshape(-1, 3, 3)
    assert np.max(np.abs(m @ m.conj().swapaxes(1, 2) - np.eye(3))) < 1e-13
    assert np.max(np.abs(np.linalg.det(m) - 1.0)) < 1e-13
    assert np.max(np.abs(S - U)) > 1e-2
    g = _su3(np.random.default_rng(18), SITES)
    a = _gauge_transform(o.hyp_step(U, 0.75, 0.6, L), g)
    b = o.hyp_step(_gauge_transform(U, g), 0.75, 0.6, L)
    assert np.max(np.abs(a - b)) < 1e-12


def test_hyp_on_diagonal_links_matches_the_phase_form():
    """Diagonal links diag(e^{i theta_k}): every product is diagonal, so each
    level acts per component k: w_k = (1 - alpha) e^{i theta_k} + (alpha / n)
    (sum of the staple phases), projected to w_k / |w_k| times the common
    factor (prod_k w_k / |w_k|)^(-1/3).  Explicit loops over coordinates."""
    rng = np.random.default_rng(19)
    th = rng.uniform(-np.pi, np.pi, (SITES, 3, 2))
    theta = np.concatenate([th, -th.sum(axis=2, keepdims=True)], axis=2)
    U = np.zeros((SITES, 3, 3, 3), complex)
    for k in range(3):
        U[:, :, k, k] = np.exp(1j * theta[:, :, k])
    a1, a2 = 0.7, 0.5
    got = o.hyp_step(U, a1, a2, L)

    def unit_and_fix(w):
        ph = w / np.abs(w)
        return ph / np.prod(ph) ** (1.0 / 3.0)

    def site(c):
        return _site(*[c[i] % L[i] for i in range(3)])

    def shift(c, d, s):
        c = list(c)
        c[d] += s
        return c

    coords = list(itertools.product(range(L[0]), range(L[1]), range(L[2])))
    others = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
    dec = {}  # (site, mu, rho) -> diagonal (3,)
    for c in coords:
        for mu in range(3):
            for rho in others[mu]:
                up = (np.exp(1j * theta[site(c), rho]) * np.exp(1j * theta[site(shift(c, rho, 1)), mu])
                      * np.exp(-1j * theta[site(shift(c, mu, 1)), rho]))
                cd = shift(c, rho, -1)
                dn = (np.exp(-1j * theta[site(cd), rho]) * np.exp(1j * theta[site(cd), mu])
                      * np.exp(1j * theta[site(shift(cd, mu, 1)), rho]))
                w = (1 - a2) * np.exp(1j * theta[site(c), mu]) + a2 / 2 * (up + dn)
                dec[(site(c), mu, rho)] = unit_and_fix(w)
    for c in coords:
        for mu in range(3):
            st = np.zeros(3, complex)
            for nu in range(3):
                if nu == mu:
                    continue
                p = 3 - mu - nu
                cd = shift(c, nu, -1)
                st += (dec[(site(c), nu, p)] * dec[(site(shift(c, nu, 1)), mu, p)]
                       * np.conj(dec[(site(shift(c, mu, 1)), nu, p)]))
                st += (np.conj(dec[(site(cd), nu, p)]) * dec[(site(cd), mu, p)]
                       * dec[(site(shift(cd, mu, 1)), nu, p)])
            want = unit_and_fix((1 - a1) * np.exp(1j * theta[site(c), mu]) + a1 / 4 * st)
            assert np.max(np.abs(np.diag(got[site(c), mu]) - want)) < 1e-13
