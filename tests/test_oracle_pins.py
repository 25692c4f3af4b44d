"""Pins of the oracle functions that the conservation and projector checks
of test_oracle.py leave open (CPU only).

Every check is a property the mathematics fixes, chosen so that a plausible
slip in the restatement fails it:
  * the filter's inner-product rescale (dycore_mod.F90:212-251): a filtered
    row is a multiple of the projected row, lies in the projector's range and
    keeps sum(tend * w) with w = U, V or gd + ghs;
  * diffusion of u and v (diffusion_mod.F90:134-171): the spherical-harmonic
    eigenvalue -n(n+1)/a^2 on the u and v points with second-order
    convergence;
  * diffusion's pole rows (diffusion_mod.F90:116-133): exact mass
    conservation for a state whose zonal mean varies next to the poles, with
    and without the filter;
  * diffusion's filter (diffusion_mod.F90:182-193): the increments of the
    flagged rows lie in the projector's range;
  * update_state's back transform (dycore_mod.F90:636-645): every u, v of a
    new state pairs with its U, V as U = (G_i + G_i+1)/2 u and
    V = (G_j + G_j+1)/2 v;
  * csp2_splitting's time weights (dycore_mod.F90:671-687): second-order
    convergence to the unsplit solution;
  * the ghs term of the energy (diag_mod.F90:115-119): a lake at rest over
    mountains;
  * isp_splitting (dycore_mod.F90:689-752): the final update's identity
    <D, D> = 2 <D, F> (D = F_new - F) under the qcon beta, its combined
    tendency R equal to twice the midpoint tendency L(F + dt/2 L(F)) to
    O(dt^2), and, for a small perturbation of the resting state, the closed
    form of the fast subcycles with the linearised fast operator.  The sign
    of the first slow correction stays unpinned: it acts only at third order
    around rest and at O(dt^2) in R.
"""
import numpy as np
import pytest

from test_oracle import A, G, rh_state, smooth_state


def _ready(om, N, M, st, **kw):
    o = om.Oracle(N, M, 600.0, **kw)
    o.set_initial_state(*st)
    o.start()
    return o


def _grid(N, M):
    dlon, dlat = 2 * np.pi / N, np.pi / (M - 1)
    flat = -0.5 * np.pi + np.arange(M) * dlat
    hlat = flat[:-1] + 0.5 * dlat
    flon = np.arange(N) * dlon
    return flat, hlat, flon, flon + 0.5 * dlon


def _weights(N, M):
    """cos of the full and half latitudes with the pole rows reset to a
    quarter of the neighbouring half-latitude cos (dycore_mod.F90:159-173)."""
    flat, hlat, _, _ = _grid(N, M)
    fc, hc = np.cos(flat), np.cos(hlat)
    fc[0], fc[-1] = 0.25 * hc[0], 0.25 * hc[-1]
    return fc, hc


def _ip(N, M, a, b):
    """The energy inner product of (du, dv, dgd)-shaped triples: du over the
    rows between the poles with the full-row cos, dv with the half-row cos,
    dgd over every row with the full-row cos."""
    fc, hc = _weights(N, M)
    return ((a[0][1:-1] * b[0][1:-1] * fc[1:-1, None]).sum() + (a[1] * b[1] * hc[:, None]).sum()
            + (a[2] * b[2] * fc[:, None]).sum())


def _noisy_state(N, M, seed):
    """A smooth state (ghs != 0) plus small-scale waves, so that the polar
    filter changes the tendency rows."""
    u, v, gd, ghs = smooth_state(N, M, seed)
    rng = np.random.default_rng(seed + 50)
    flat, hlat, flon, hlon = _grid(N, M)
    for k in (9, 13, 17):
        u[1:-1] += 2.0 * np.cos(flat[1:-1, None]) * np.cos(k * hlon[None, :] + rng.uniform(0, 6))
        v += 2.0 * np.cos(hlat[:, None]) * np.sin(k * flon[None, :] + rng.uniform(0, 6))
        gd += G * 10.0 * np.cos(flat[:, None]) ** 2 * np.cos(k * flon[None, :] + rng.uniform(0, 6))
    return u, v, gd, ghs


# ---------------------------------------------------------------------------
# polar filter rescale
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("pass_", [0, 1, 2])
def test_filter_rescale_keeps_weighted_row_sum(oracle_mod, pass_):
    T = oracle_mod
    N, M = 72, 37
    cut = [8, 7, 6, 5, 4, 3, 2]
    st = _noisy_state(N, M, 5)
    assert np.abs(st[3]).max() > 100.0  # mountains: the dgd weight is gd + ghs, not gd
    of = _ready(oracle_mod, N, M, st, cutoff=cut)
    on = _ready(oracle_mod, N, M, st, cutoff=cut, use_zonal_tend_filter=False)
    of.space_operators(of.old, 0, pass_)
    on.space_operators(on.old, 0, pass_)
    ff, hf, _, _ = of.filter_table()
    assert ff.sum() >= 10 and hf.sum() >= 10
    gd = of.state(of.old, T.GD)
    weights = {T.T_DU: of.state(of.old, T.IAP_U), T.T_DV: of.state(of.old, T.IAP_V),
               T.T_DGD: gd + st[3]}
    checked = moved = weight_matters = 0
    for comp, flags, lo, hi in ((T.T_DU, ff, 1, M - 1), (T.T_DV, hf, 0, M - 1), (T.T_DGD, ff, 0, M)):
        t0, t1, w = on.tend(0, comp), of.tend(0, comp), weights[comp]
        filt = of.filter_half if comp == T.T_DV else of.filter_full
        for j in range(lo, hi):
            if not flags[j]:
                np.testing.assert_array_equal(t1[j], t0[j])
                continue
            s1, mag = (t0[j] * w[j]).sum(), np.abs(t0[j] * w[j]).sum()
            if mag == 0.0:
                np.testing.assert_array_equal(t1[j], t0[j])
                continue
            assert abs(s1) > 1e-6 * mag  # a well-conditioned rescale
            p = filt(j + 1, t0[j])
            # the projection moves the row, so the rescale has work to do
            moved += np.abs(p - t0[j]).max() > 1e-6 * np.abs(t0[j]).max()
            # the rescaled row keeps the weighted sum ...
            assert abs((t1[j] * w[j]).sum() - s1) <= 1e-10 * mag, (comp, j)
            # ... is a multiple of the projected row ...
            alpha = (t1[j] * p).sum() / (p * p).sum()
            assert np.abs(t1[j] - alpha * p).max() <= 1e-12 * np.abs(t1[j]).max(), (comp, j)
            # ... and lies in the projector's range
            assert np.abs(filt(j + 1, t1[j]) - t1[j]).max() <= 1e-12 * np.abs(t1[j]).max(), (comp, j)
            if comp == T.T_DGD:
                # gd alone as the weight would give another multiple
                weight_matters += abs((t1[j] * gd[j]).sum() - (t0[j] * gd[j]).sum()) > 1e-8 * mag
            checked += 1
    assert checked >= 20 and moved >= 10
    if pass_ != 2:
        assert weight_matters >= 5


# ---------------------------------------------------------------------------
# diffusion: u and v operators, pole rows, filter
# ---------------------------------------------------------------------------

def test_diffusion_eigenvalue_on_u_and_v_points(oracle_mod):
    """The operator on u (full rows, half longitudes) and on v (half rows,
    full longitudes) applied to Y_3^2 gives -12/a^2 Y to O(dlat^2)."""
    T = oracle_mod
    errs = []
    for N, M in ((72, 37), (144, 73)):
        flat, hlat, flon, hlon = _grid(N, M)
        Y = lambda lat, lon: (np.cos(lat) ** 2 * np.sin(lat))[:, None] * np.cos(2 * lon)[None, :]
        u0 = 10.0 * Y(flat, hlon)
        u0[0] = u0[-1] = 0.0
        v0 = 10.0 * Y(hlat, flon)
        gd0 = np.full((M, N), 1.0e5)
        o = _ready(oracle_mod, N, M, (u0, v0, gd0, np.zeros((M, N))), use_diffusion=True,
                   diffusion_coef=A ** 2, use_zonal_tend_filter=False)
        o.ordinary_diffusion(1.0, o.old)  # x <- x + dt nu lap(x), dt nu = a^2
        du = o.state(o.old, T.U) - u0
        dv = o.state(o.old, T.V) - v0
        # a fixed latitude band (the rows next to the poles converge later)
        bu, bv = np.abs(flat) <= np.radians(60.0), np.abs(hlat) <= np.radians(60.0)
        eu = np.abs(du[bu] + 12.0 * u0[bu]).max() / np.abs(12.0 * u0).max()
        ev = np.abs(dv[bv] + 12.0 * v0[bv]).max() / np.abs(12.0 * v0).max()
        errs.append((eu, ev))
    (eu0, ev0), (eu1, ev1) = errs
    assert eu0 < 2e-2 and eu1 < eu0 / 3.0
    assert ev0 < 2e-2 and ev1 < ev0 / 3.0


@pytest.mark.parametrize("filt", [False, True])
def test_diffusion_conserves_mass_with_polar_zonal_mean(oracle_mod, filt):
    """gd whose zonal mean changes between the pole and the next row: the pole
    rows' values close the mass budget exactly (sum_j cos_j sum_i gdd = 0).
    The filter keeps each row's zonal mean, so it changes nothing here."""
    T = oracle_mod
    N, M = 72, 37
    flat, _, flon, _ = _grid(N, M)
    u, v, _, ghs = _noisy_state(N, M, 3)
    gd = G * (8000.0 + 300.0 * np.sin(flat) + 200.0 * np.sin(flat) ** 2)[:, None] + np.zeros((M, N))
    gd += G * 40.0 * np.cos(flat[:, None]) * np.cos(5 * flon[None, :])
    o = _ready(oracle_mod, N, M, (u, v, gd, ghs), use_diffusion=True, diffusion_coef=1e6,
               use_zonal_tend_filter=filt, cutoff=[6, 5, 4, 3])
    m0, _ = o.diag(o.old)
    o.ordinary_diffusion(600.0, o.old)
    m1, _ = o.diag(o.old)
    g1 = o.state(o.old, T.GD)
    # the pole rows moved, uniformly along the row
    for j in (0, M - 1):
        assert np.abs(g1[j] - gd[j]).max() > 1e-3 * G
        assert np.all(g1[j] == g1[j][0])
    assert abs(m1 - m0) <= 2e-15 * m0


def test_diffusion_filter_keeps_increments_in_filter_range(oracle_mod):
    """The filtered rows' Laplacians (no rescale) are projections: each
    flagged row's increment of gd and u (full rows between the poles) and of
    v (half rows) is left unchanged by the filter, while an unflagged row's
    increment has the small-scale waves the filter would remove."""
    T = oracle_mod
    N, M = 72, 37
    cut = [6, 5, 4, 3, 2]
    st = _noisy_state(N, M, 9)
    o = _ready(oracle_mod, N, M, st, use_diffusion=True, diffusion_coef=2e5, cutoff=cut)
    ff, hf, _, _ = o.filter_table()
    x0 = [o.state(o.old, f).copy() for f in (T.GD, T.U, T.V)]
    o.ordinary_diffusion(600.0, o.old)
    x1 = [o.state(o.old, f) for f in (T.GD, T.U, T.V)]
    inc = [b - a for a, b in zip(x0, x1)]
    # an increment is a difference of two stored values: it carries their
    # rounding, 2 eps |x| (gd is ~8e4 where its increment is ~0.05)
    tol = [1e-12 * np.abs(d).max() + 4.5e-16 * np.abs(x).max() for d, x in zip(inc, x1)]
    n = 0
    for j in range(1, M - 1):
        if ff[j]:
            for d, t in zip(inc[:2], tol[:2]):
                assert np.abs(o.filter_full(j + 1, d[j]) - d[j]).max() <= t, j
                n += 1
    for j in range(M - 1):
        if hf[j]:
            d = inc[2]
            assert np.abs(o.filter_half(j + 1, d[j]) - d[j]).max() <= tol[2], j
            n += 1
    assert n >= 15
    # the first unflagged row: its increment is not in that row's range of a
    # low cutoff (the waves of the state are there)
    j = next(j for j in range(1, M - 1) if not ff[j])
    d = inc[1][j]
    fl = o.filter_full(2, d)  # the projector of the row next to the pole (cutoff 6)
    assert np.abs(fl - d).max() > 1e-3 * np.abs(d).max()


# ---------------------------------------------------------------------------
# update_state back transform, csp2 weights, energy
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("pass_", [0, 1, 2])
def test_update_state_pairs_u_v_with_U_V(oracle_mod, pass_):
    """After an integrator call with dt != 0 every stored u, v is the wind of
    its IAP pair: U = (G_i + G_i+1)/2 u on the rows between the poles and
    V = (G_j + G_j+1)/2 v, G = sqrt(gd)."""
    T = oracle_mod
    N, M = 72, 37
    o = _ready(oracle_mod, N, M, _noisy_state(N, M, 4), cutoff=[5, 3])
    g0 = o.state(o.old, T.GD).copy()
    o.predict_correct(300.0, o.old, o.new, pass_, True)
    s = o.new
    gd, u, v = o.state(s, T.GD), o.state(s, T.U), o.state(s, T.V)
    U, V, Gs = o.state(s, T.IAP_U), o.state(s, T.IAP_V), o.state(s, T.IAP_GD)
    if pass_ != 2:
        assert np.abs(gd - g0).max() > 1e-3  # the state moved
    np.testing.assert_array_equal(Gs, np.sqrt(gd))
    Gr = np.sqrt(gd)
    Uc = 0.5 * (Gr[1:-1] + np.roll(Gr[1:-1], -1, axis=1)) * u[1:-1]
    Vc = 0.5 * (Gr[:-1] + Gr[1:]) * v
    assert np.abs(Uc - U[1:-1]).max() <= 4e-16 * np.abs(U).max()
    assert np.abs(Vc - V).max() <= 4e-16 * np.abs(V).max()


def _run_to(oracle_mod, st, N, M, dt, nsteps, split):
    o = oracle_mod.Oracle(N, M, dt, split_scheme=split, subcycles=3, qcon_modified=False,
                          use_zonal_tend_filter=False)
    o.set_initial_state(*st)
    o.start()
    _, _, _, ns = o.run(nsteps)
    assert ns == 0
    T = oracle_mod
    return [o.state(o.old, f) for f in (T.GD, T.IAP_U, T.IAP_V)]


def test_csp2_converges_to_unsplit_solution_at_second_order(oracle_mod):
    """csp2 (slow dt/2, three fast dt/3 calls, slow dt/2) and the unsplit
    integrator approximate the same solution: at a fixed time the splitting
    error falls 4x when dt halves, against an unsplit run at dt/16."""
    N, M = 72, 37
    st = rh_state(N, M)
    Tend = 2400.0
    ref = _run_to(oracle_mod, st, N, M, 300.0 / 16, 128, "none")
    x1 = _run_to(oracle_mod, st, N, M, 300.0, 8, "csp2")
    x2 = _run_to(oracle_mod, st, N, M, 150.0, 16, "csp2")
    g0 = st[2]
    change = np.abs(ref[0] - g0).max()
    e1 = np.abs(x1[0] - ref[0]).max()
    e2 = np.abs(x2[0] - ref[0]).max()
    assert change > 10.0 and Tend == 8 * 300.0
    assert e1 < 2e-2 * change
    assert 3.0 < e1 / e2 < 5.5
    for k in (1, 2):  # the momentum converges too
        d1 = np.abs(x1[k] - ref[k]).max()
        d2 = np.abs(x2[k] - ref[k]).max()
        assert d1 < 2e-2 * np.abs(ref[k] - 0).max() and 3.0 < d1 / d2 < 5.5


def test_energy_of_lake_at_rest_over_mountains(oracle_mod):
    """u = v = 0 and a flat free surface gd + ghs = K over mountains: the
    total energy is K^2 N sum_j cos_j (pole rows with the reset cos)."""
    N, M = 72, 37
    _, _, _, ghs = smooth_state(N, M, 2)
    K = G * 8000.0
    assert ghs.max() > 100.0
    o = _ready(oracle_mod, N, M, (np.zeros((M, N)), np.zeros((M - 1, N)), K - ghs, ghs))
    _, energy = o.diag(o.old)
    fc, _ = _weights(N, M)
    assert np.isclose(energy, K * K * N * fc.sum(), rtol=1e-13)


# ---------------------------------------------------------------------------
# isp_splitting
# ---------------------------------------------------------------------------

def _isp_step(oracle_mod, st, N, M, dt, qcon):
    """One isp step: (F, F_new, beta); F and F_new as (U, V, gd) triples."""
    T = oracle_mod
    o = oracle_mod.Oracle(N, M, dt, split_scheme="isp", subcycles=3, qcon_modified=qcon,
                          use_zonal_tend_filter=False)
    o.set_initial_state(*st)
    o.start()
    F = [o.state(o.old, f).copy() for f in (T.IAP_U, T.IAP_V, T.GD)]
    beta = o.time_integrate()
    Fn = [o.state(o.new, f) for f in (T.IAP_U, T.IAP_V, T.GD)]
    return F, Fn, beta


def _midpoint_tendency(oracle_mod, st, N, M, dt):
    """L(F + dt/2 L(F)) through the (pinned) operator and update primitives."""
    T = oracle_mod
    o = oracle_mod.Oracle(N, M, dt, use_zonal_tend_filter=False)
    o.set_initial_state(*st)
    o.start()
    o.space_operators(o.old, 0, T.ALL_PASS)
    o.update_state(0.5 * dt, 0, o.old, o.new)
    o.space_operators(o.new, 1, T.ALL_PASS)
    return [o.tend(1, f) for f in (T.T_DU, T.T_DV, T.T_DGD)]


def test_isp_final_update_identity(oracle_mod):
    """With qcon_modified, beta = <R,F>/<R,R> * 4/dt and F_new = F + dt/2 beta R
    (dycore_mod.F90:743-750), so D = F_new - F = 2 <R,F>/<R,R> R satisfies
    <D, D> = 2 <D, F> whatever R is."""
    N, M = 36, 19
    u, v, gd, ghs = smooth_state(N, M, 6)  # mountains: <R, F> does not vanish
    F, Fn, beta = _isp_step(oracle_mod, (u, v, gd, ghs), N, M, 300.0, True)
    D = [b - a for a, b in zip(F, Fn)]
    lhs, rhs = _ip(N, M, D, D), 2.0 * _ip(N, M, D, F)
    assert lhs > 0 and abs(lhs - rhs) <= 1e-8 * abs(lhs)
    # the reported beta is the one applied: D = dt/2 beta R with R = D/(dt/2 beta)
    # reproducing beta = <R,F>/<R,R> * 4/dt
    h = 0.5 * 300.0 * beta
    R = [d / h for d in D]
    assert np.isclose(_ip(N, M, R, F) / _ip(N, M, R, R) * 4.0 / 300.0, beta, rtol=1e-8)


def test_isp_tendency_is_twice_the_midpoint_tendency(oracle_mod):
    """Without qcon_modified the final update is F + 2 R (beta = 4/dt), where
    R = L2(Q2) + L2(F) + 2/m sum_k L1(P2^{k+1/2}) (slow L2, fast L1).  As dt
    shrinks R = 2 L(F + dt/2 L(F)) + O(dt^2): the fast subcycles average L1
    over the step's midpoints and the slow corrections evaluate L2 at its end
    and start.  The error falls 4x when dt halves."""
    N, M = 36, 19
    st = rh_state(N, M)
    errs = []
    for dt in (240.0, 120.0):
        F, Fn, beta = _isp_step(oracle_mod, st, N, M, dt, False)
        assert np.isclose(beta, 4.0 / dt, rtol=1e-15)
        R = [(b - a) / 2.0 for a, b in zip(F, Fn)]
        L = _midpoint_tendency(oracle_mod, st, N, M, dt)
        diff = [r - 2.0 * l for r, l in zip(R, L)]
        errs.append(np.sqrt(_ip(N, M, diff, diff) / _ip(N, M, L, L)))
    assert errs[0] < 1e-2
    assert 3.0 < errs[0] / errs[1] < 5.5


def _fast_jacobian(oracle_mod, N, M, gd0, eps):
    """The fast operator's Jacobian at rest (U = V = 0, gd = gd0 everywhere,
    no topography), column by column: central differences of
    space_operators(FAST_PASS) at rest +- eps e_k.  Each perturbed state is
    formed by update_state from a tendency holding eps e_k, so its u, v are
    derived as in the model; the columns carry an O(eps^2) error."""
    T = oracle_mod
    o = T.Oracle(N, M, 600.0, use_zonal_tend_filter=False)
    o.set_initial_state(np.zeros((M, N)), np.zeros((M - 1, N)), np.full((M, N), gd0), np.zeros((M, N)))
    o.start()
    tf = (T.T_DU, T.T_DV, T.T_DGD)
    rows = (M, M - 1, M)
    n = N * sum(rows)

    def fast(x):
        k = 0
        for f, r in zip(tf, rows):
            part = np.ascontiguousarray(x[k:k + r * N].reshape(r, N))
            T.lib().og_set_field(o.h, 1, 0, f, part)
            k += r * N
        o.update_state(1.0, 0, o.old, o.new)
        o.space_operators(o.new, 1, T.FAST_PASS)
        return np.concatenate([o.tend(1, f).ravel() for f in tf])

    J = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = eps
        J[:, k] = (fast(e) - fast(-e)) / (2.0 * eps)
    return J


def test_isp_fast_subcycles_match_their_linear_closed_form(oracle_mod):
    """A small perturbation x of the resting state: the slow operator
    (advection) is quadratic in it and drops out, the fast one is linear,
    L1 = J x.  Without qcon_modified one isp step then gives, to first
    order in x (dycore_mod.F90:689-752; h = dt/m):
      P2_k = (I + h/2 J + h^2/4 J^2) P_k,   P_k+1 = P_k + h J P2_k,
      F_new - F = 2 R,   R = 2/m sum_k J P2_k,   P_0 = x.
    The fast stages' fractions (h/2, h/2, h) each change the result at the
    scale (h |J|)^2 of the residual allowed here."""
    T = oracle_mod
    N, M, m, dt = 12, 7, 2, 3000.0
    gd0 = G * 8000.0
    J = _fast_jacobian(oracle_mod, N, M, gd0, 1e-3)
    rng = np.random.default_rng(4)
    lat = np.linspace(-np.pi / 2, np.pi / 2, M)
    lon = np.arange(N) * 2 * np.pi / N
    hl = (np.linspace(-np.pi / 2, np.pi / 2, M)[:-1] + np.pi / (2 * (M - 1)))
    amp = 1e-3
    u = amp * np.cos(lat)[:, None] * (np.sin(lon + rng.uniform(0, 6))[None, :] + 0.5)
    u[0] = u[-1] = 0.0
    v = amp * np.cos(hl)[:, None] * np.cos(2 * lon + rng.uniform(0, 6))[None, :]
    gd = gd0 + G * amp * np.cos(lat)[:, None] ** 2 * np.cos(lon + rng.uniform(0, 6))[None, :]
    o = T.Oracle(N, M, dt, split_scheme="isp", subcycles=m, qcon_modified=False, use_zonal_tend_filter=False)
    o.set_initial_state(u, v, gd, np.zeros((M, N)))
    o.start()
    F = [o.state(o.old, f).copy() for f in (T.IAP_U, T.IAP_V, T.GD)]
    o.time_integrate()
    Fn = [o.state(o.new, f) for f in (T.IAP_U, T.IAP_V, T.GD)]
    x = np.concatenate([F[0].ravel(), F[1].ravel(), (F[2] - gd0).ravel()])
    step = np.concatenate([(b - a).ravel() for a, b in zip(F, Fn)])
    h = dt / m
    I = np.eye(len(x))
    A2 = I + 0.5 * h * J + 0.25 * h * h * (J @ J)
    P = x.copy()
    R = np.zeros_like(x)
    for _ in range(m):
        P2 = A2 @ P
        R += J @ P2
        P = P + h * (J @ P2)
    pred = 2.0 * (2.0 / m) * R
    # the same closed form with the first stage at h instead of h/2
    P, Rs = x.copy(), np.zeros_like(x)
    A2s = I + 0.5 * h * J + 0.5 * h * h * (J @ J)
    for _ in range(m):
        P2 = A2s @ P
        Rs += J @ P2
        P = P + h * (J @ P2)
    slipped = 2.0 * (2.0 / m) * Rs
    err = np.linalg.norm(step - pred) / np.linalg.norm(pred)
    gap = np.linalg.norm(slipped - pred) / np.linalg.norm(pred)
    assert gap > 1e-3            # the closed form resolves the stage fraction
    assert err < 1e-2 * gap, (err, gap)
