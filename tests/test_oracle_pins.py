"""Pins of the CPU oracle that do not need the reference build: closed forms,
invariants, limits and brute-force recomputations.  Each is chosen so that a
plausible slip in the restatement (a dropped term, a wrong sign or index, a
swapped operand) fails it.  Runs on the CPU."""
import math

import numpy as np
import pytest

import cases
import oracle

R_UNIV = 8.31434
RU_CAL = 1.987
P0 = 101325.0
# atoms (H, O, N) of the mechanism's species, in file order (input/H2_JA919.txt:5)
SPECIES = ["H", "O", "H2O", "OH", "O2", "H2", "H2O2", "HO2", "N2"]
ATOMS = np.array([[1, 0, 0], [0, 1, 0], [2, 1, 0], [1, 1, 0], [0, 2, 0], [2, 0, 0], [2, 2, 0], [1, 2, 0],
                  [0, 0, 2]], dtype=float)


@pytest.fixture(scope="module")
def orc(tmp_path_factory):
    d = tmp_path_factory.mktemp("pins")
    return oracle.Oracle(cases.rsbi(str(d)))


def thermo(o, T, s, R=R_UNIV):
    c0, c1 = o.field("coeff0"), o.field("coeff1")
    L = oracle.lib()
    args = (oracle.ptr(c0), oracle.ptr(c1), o.ns)
    return (L.or_getcpi(T, R, s, *args), L.or_gethi(T, R, s, *args), L.or_getsi(T, R, s, *args))


def test_thermo_closed_form_at_probe_temperatures(orc):
    """NASA-9 polynomials (McBride, Zehe & Gordon 2002) evaluated independently
    with Horner's rule on the thermo.dat coefficients."""
    c0, c1 = orc.field("coeff0"), orc.field("coeff1")
    for T in [300.0, 999.0, 1000.0, 1001.0, 2000.0, 3500.0]:
        a = c0 if T < 1000.0 else c1
        for s in range(orc.ns):
            k = a[:, s]
            cp = k[0] / T ** 2 + k[1] / T + k[2] + T * (k[3] + T * (k[4] + T * (k[5] + T * k[6])))
            h = -k[0] / T + k[1] * math.log(T) + T * (k[2] + T * (k[3] / 2 + T * (k[4] / 3 + T * (k[5] / 4 + T * k[6] / 5)))) + k[7]
            sv = -k[0] / T ** 2 / 2 - k[1] / T + k[2] * math.log(T) + T * (k[3] + T * (k[4] / 2 + T * (k[5] / 3 + T * k[6] / 4))) + k[8]
            got = thermo(orc, T, s)
            for g, e in zip(got, (R_UNIV * cp, R_UNIV * h, R_UNIV * sv)):
                assert abs(g - e) <= 1e-12 * (abs(e) + R_UNIV * 1e3), (T, SPECIES[s])


def test_thermo_derivative_consistency(orc):
    """dh/dT = cp and ds/dT = cp/T inside each range (catches a wrong power or divisor)."""
    for T in [400.0, 700.0, 1500.0, 2800.0]:
        d = 1e-3 * T
        for s in range(orc.ns):
            cp, _, _ = thermo(orc, T, s)
            _, hp, sp = thermo(orc, T + d, s)
            _, hm, sm = thermo(orc, T - d, s)
            assert abs((hp - hm) / (2 * d) - cp) <= 1e-6 * abs(cp), (T, SPECIES[s])
            assert abs((sp - sm) / (2 * d) - cp / T) <= 1e-6 * abs(cp / T), (T, SPECIES[s])


def test_atomic_hydrogen_low_range_is_monatomic(orc):
    """H below 1000 K: cp = 2.5 R exactly (input/thermo.dat:16-17)."""
    for T in [300.0, 600.0, 999.0]:
        cp, h, _ = thermo(orc, T, 0)
        assert cp == 2.5 * R_UNIV
        b1 = orc.field("coeff0")[7, 0]
        assert abs(h - R_UNIV * (2.5 * T + b1)) <= 1e-12 * abs(h)


def _ausm(direction, L, R, ns):
    F = np.zeros(ns + 4)
    oracle.lib().or_ausm_face(direction, oracle.ptr(np.ascontiguousarray(L)), oracle.ptr(np.ascontiguousarray(R)),
                              ns, oracle.ptr(F))
    return F


@pytest.mark.parametrize("mach", [0.0, 0.3, -0.55, 0.99, 1.5, -2.2])
@pytest.mark.parametrize("direction", [1, 2, 3])
def test_ausm_uniform_state_is_exact_euler_flux(mach, direction):
    """For equal left/right states AUSM+ gives the exact Euler flux:
    B1+B2 = M and G1+G2 = 1 (Liou 1996), supersonic limits included."""
    ns = 9
    P, D, g = 2.3e5, 1.7, 1.33
    c = math.sqrt(g * P / D)
    vel = [17.0, -5.0, 3.0]
    vel[direction - 1] = mach * c
    H = 2.1e6
    Y = np.linspace(0.01, 0.2, ns)
    st = np.array([P, D] + vel + [H, g] + list(Y))
    F = _ausm(direction, st, st, ns)
    un = vel[direction - 1]
    exact = np.zeros(ns + 4)
    exact[:ns] = D * un * Y
    exact[ns:ns + 3] = [D * un * v for v in vel]
    exact[ns + direction - 1] += P
    exact[ns + 3] = D * un * (H + 0.5 * sum(v * v for v in vel))
    scale = np.abs(exact) + abs(D * c) * (1 + np.abs(np.r_[Y, vel, H]))
    assert np.all(np.abs(F - exact) <= 1e-13 * scale), (F, exact)


def test_ausm_mirror_symmetry():
    """Reflecting the problem (swap sides, flip the normal velocity) flips the mass flux."""
    ns = 9
    rng = np.random.default_rng(3)
    for _ in range(20):
        L = np.r_[rng.uniform(5e4, 5e5), rng.uniform(0.2, 3), rng.uniform(-300, 300, 3), rng.uniform(1e5, 3e6),
                  rng.uniform(1.2, 1.6), rng.dirichlet(np.ones(ns))]
        R = np.r_[rng.uniform(5e4, 5e5), rng.uniform(0.2, 3), rng.uniform(-300, 300, 3), rng.uniform(1e5, 3e6),
                  rng.uniform(1.2, 1.6), rng.dirichlet(np.ones(ns))]
        F = _ausm(1, L, R, ns)
        Lm, Rm = R.copy(), L.copy()
        Lm[2] *= -1
        Rm[2] *= -1
        G = _ausm(1, Lm, Rm, ns)
        assert np.allclose(G[:ns], -F[:ns], rtol=1e-12, atol=1e-12 * np.abs(F[:ns]).max())
        assert abs(G[ns] - F[ns]) <= 1e-10 * abs(F[ns])  # momentum flux is even


def _uniform(o):
    P, T, D, Hh, g = 1.2e5, 900.0, 0.5, 1.3e6, 1.35
    Y = np.linspace(0.02, 0.2, o.ns)
    Y /= Y.sum()
    for n, v in [("P", P), ("D", D), ("U", 120.0), ("V", -40.0), ("W", 15.0), ("H", Hh), ("Gamma", g), ("T", T)]:
        o.field(n)[...] = v
    o.field("Yi")[...] = Y


def test_ee_free_stream_preservation(tmp_path):
    """A uniform state gives RHS = 0 exactly (MUSCL differences vanish, equal fluxes)."""
    o = oracle.Oracle(cases.rsbi(str(tmp_path), ni=9, nj=7, nk=6))
    _uniform(o)
    o.set_dt(1e-7)
    o.advection()
    assert np.all(o.interior(o.field("RHS")) == 0.0)


def test_ee_stencil_footprint(tmp_path):
    """Perturbing one cell changes RHS only within the MUSCL/AUSM stencil
    (|di| + |dj| + |dk| <= 2 along one axis at a time, i.e. a 2-cell cross)."""
    o = oracle.Oracle(cases.rsbi(str(tmp_path), ni=13, nj=11, nk=9))
    _uniform(o)
    b = o.bc
    ci, cj, ck = b + 6, b + 5, b + 4
    o.field("Yi")[ci, cj, ck, 3] += 1e-3
    o.set_dt(1e-7)
    o.advection()
    rhs = o.field("RHS")[..., 3]
    nz = np.argwhere(rhs != 0.0)
    assert len(nz) > 0
    for i, j, k in nz:
        d = (abs(i - ci), abs(j - cj), abs(k - ck))
        assert sorted(d)[:2] == [0, 0] and max(d) <= 2, d


def test_rates_arrhenius_and_detailed_balance(orc):
    """kf = A T^b exp(-E/(Ru T)); kb = kf / Kc with Kc from the Gibbs energy of
    reaction (brute force from the thermo); equilibrium constant in
    concentration units (R T)^-dn, p0 = 0.101325 in the reference's units."""
    nf = np.ctypeslib.as_array(orc.L.or_field(orc.h, b"stoi_f"), shape=(orc.ns * orc.nr,)).reshape(orc.ns, orc.nr)
    nb = np.ctypeslib.as_array(orc.L.or_field(orc.h, b"stoi_b"), shape=(orc.ns * orc.nr,)).reshape(orc.ns, orc.nr)
    A, B, E = (orc.field_raw(n, None) for n in ("af", "bf", "eaf"))
    Mc = np.full(orc.ns, 1e-6)
    for T in [800.0, 1200.0, 2500.0]:
        r = orc.rates(T, Mc)
        G = np.array([thermo(orc, T, s)[1] - T * thermo(orc, T, s)[2] for s in range(orc.ns)])
        for k in range(orc.nr):
            kf = A[k] * T ** B[k] * math.exp(-E[k] / (RU_CAL * T))
            assert abs(r["kf"][k] - kf) <= 1e-12 * kf
            dn = (nb[:, k] - nf[:, k]).sum()
            dg = ((nb[:, k] - nf[:, k]) * G).sum()
            Kc = math.exp(-dg / (R_UNIV * T)) * (P0 * 1e-6) ** dn * (R_UNIV * T) ** (-dn)
            assert abs(r["kb"][k] - kf / Kc) <= 1e-10 * abs(kf / Kc), k


def test_rates_conserve_elements(orc):
    """sum_s w_s a_es / Mw_s = 0 for H, O, N (each reaction balances atoms)."""
    mw = orc.field("mw")
    rng = np.random.default_rng(11)
    for T in [700.0, 1100.0, 1900.0, 3100.0]:
        Mc = 10.0 ** rng.uniform(-9, -5, orc.ns)
        wi = orc.rates(T, Mc)["wi"]
        molar = wi / mw
        for e in range(3):
            terms = molar * ATOMS[:, e]
            assert abs(terms.sum()) <= 1e-11 * np.abs(terms).max() + 1e-300, (T, e)


def test_trapezoid_small_step_limit(orc):
    """(rho' - rho)/dt -> w as dt -> 0 (production and loss terms consistent with w)."""
    rng = np.random.default_rng(5)
    T = 1500.0
    Mc_si = 10.0 ** rng.uniform(-3, 1, orc.ns)
    Di = Mc_si * orc.field("mw") / 1000.0
    cms = np.zeros(orc.ns)
    errs = []
    for dt in [1e-11, 1e-12, 1e-13, 1e-15]:
        D2 = Di.copy()
        orc.L.or_trapezoid_cell(orc.h, T, oracle.ptr(Mc_si.copy()), oracle.ptr(D2), dt, oracle.ptr(cms))
        rate = (D2 - Di) / dt
        errs.append(np.max(np.abs(rate - cms)) / np.abs(cms).max())
    # first order in dt until round-off, then small
    assert errs[1] < 0.2 * errs[0] and errs[2] < 0.2 * errs[1]
    assert errs[3] < 1e-4


def test_nchem_brute_force(orc):
    rng = np.random.default_rng(9)
    mw = orc.field("mw")
    for _ in range(30):
        T = rng.uniform(900, 2800)
        Mc_si = 10.0 ** rng.uniform(-4, 1.5, orc.ns)
        Di = Mc_si * mw / 1000.0
        Y = Di / Di.sum()
        dtm = 10.0 ** rng.uniform(-8, -5)
        wi = orc.rates(T, Mc_si * 1e-6)["wi"]
        allowed = dtm
        for s in range(orc.ns):
            if Y[s] >= 1e-6 and wi[s] != 0:
                allowed = min(allowed, abs(Di[s] * 1e-3 / wi[s]))
        expect = math.ceil(dtm / allowed)
        got = orc.L.or_predict_cell(orc.h, T, oracle.ptr(Mc_si), oracle.ptr(Di), oracle.ptr(Y), dtm)
        assert got == expect


def test_newton_temperature_recovery(orc):
    """Get_temp returns the T whose mixture internal energy is E; for pure H in
    the low range the answer is analytic: T = (e/R_H - b1)/1.5."""
    ri = orc.field("ri")
    rng = np.random.default_rng(2)
    for _ in range(20):
        Y = rng.dirichlet(np.ones(orc.ns))
        Tt = rng.uniform(400, 2900)
        e = sum(Y[s] * (thermo(orc, Tt, s, ri[s])[1] - ri[s] * Tt) for s in range(orc.ns)) * 1e3
        T = orc.get_temp(Tt * 1.05, e, Y)
        assert abs(T - Tt) <= 1e-7 * Tt
    Y = np.zeros(orc.ns)
    Y[0] = 1.0
    b1 = orc.field("coeff0")[7, 0]
    for Tt in [350.0, 600.0, 950.0]:
        e = ri[0] * (1.5 * Tt + b1) * 1e3  # J/kg, low range
        T = orc.get_temp(500.0, e, Y)
        assert abs(T - (e * 1e-3 / ri[0] - b1) / 1.5) <= 1e-9 * T


def test_ghost_maps_by_enumeration(tmp_path):
    """FieldBoundary_3dRSBI: x mirror (j,k interior), y-front symmetry with
    v -> -v, y-back copy, z mirror with w -> -w, filled in x, y, z order."""
    o = oracle.Oracle(cases.rsbi(str(tmp_path), ni=7, nj=6, nk=5))
    rng = np.random.default_rng(4)
    for n in ["U", "V", "W", "P", "D", "T", "H", "Gamma", "E"]:
        o.field(n)[...] = rng.uniform(1, 2, o.field(n).shape)
    o.field("Yi")[...] = rng.dirichlet(np.ones(o.ns), size=o.field("Yi").shape[:3])
    before = {n: o.field(n).copy() for n in ["U", "V", "W", "P", "T", "Yi"]}
    o.field_boundary()
    b, ni, nj, nk = o.bc, o.ni, o.nj, o.nk
    exp = {n: v.copy() for n, v in before.items()}
    J, K = slice(b, nj + b), slice(b, nk + b)
    for n, v in exp.items():
        for i in range(b):
            v[i, J, K] = v[2 * b - 1 - i, J, K]
        for i in range(ni + b, ni + 2 * b):
            v[i, J, K] = v[2 * (ni + b) - 1 - i, J, K]
        for j in range(b):
            v[:, j, K] = (-1 if n == "V" else 1) * v[:, 2 * b - 1 - j, K]
        for j in range(nj + b, nj + 2 * b):
            v[:, j, K] = v[:, 2 * (nj + b) - 1 - j, K]
        for k in range(b):
            v[:, :, k] = (-1 if n == "W" else 1) * v[:, :, 2 * b - 1 - k]
        for k in range(nk + b, nk + 2 * b):
            v[:, :, k] = (-1 if n == "W" else 1) * v[:, :, 2 * (nk + b) - 1 - k]
    for n in exp:
        assert np.array_equal(exp[n], o.field(n)), n


def test_rk3_is_three_ee_stages_in_shu_osher_form(tmp_path):
    """One SSP-RK3 advance equals the composition of EE-type stages, each with
    its own ghost fill, flux evaluation and lagged-E state recovery:
    u1 = u0 + L(u0); u2 = 3/4 u0 + 1/4 (u1 + L(u1)); u3 = 1/3 u0 + 2/3 (u2 + L(u2))."""
    ctrl = cases.rsbi(str(tmp_path), ni=11, nj=9, nk=5)
    a, b = oracle.Oracle(ctrl), oracle.Oracle(ctrl)
    dt = 3e-8
    a.set_dt(dt)
    a.advance_rk3()
    b.set_dt(dt)
    I = b.interior
    u0 = I(b.field("CS")).copy()
    for stage in range(3):
        b.field_boundary()
        b.advection()
        u, l = I(b.field("CS")), I(b.field("RHS"))
        if stage == 0:
            u[...] = u + l
        elif stage == 1:
            u[...] = 0.75 * u0 + 0.25 * (u + l)
        else:
            u[...] = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (u + l)
        b.update_after_cs()
    for n in ["CS", "T", "E", "P", "Yi"]:
        assert np.array_equal(I(a.field(n)), I(b.field(n))), n
    # and the stages are not trivial: the result differs from one EE step
    c = oracle.Oracle(ctrl)
    c.set_dt(dt)
    c.field_boundary()
    c.advection()
    c.update_after_adv()
    assert not np.array_equal(I(c.field("CS")), I(a.field("CS")))


# ---------------------------------------------------------------------------
# the surrogate network's CPU restatement (oracle/mlp_oracle.py)
# ---------------------------------------------------------------------------
import mlp_oracle as mo  # noqa: E402


def test_bf16_rounding_exact_cases():
    one = 1.0
    cases_ = [(one, one), (one + 2.0 ** -8, one),                    # tie -> even (down)
              (one + 3 * 2.0 ** -8, one + 2.0 ** -6),                 # tie -> even (up)
              (one + 2.0 ** -8 + 2.0 ** -20, one + 2.0 ** -7),        # above the tie
              (-(one + 3 * 2.0 ** -8), -(one + 2.0 ** -6)), (0.0, 0.0), (2.0 ** -130, 2.0 ** -130)]
    for x, want in cases_:
        assert mo.bf16(np.array([x]))[0] == want, x
    assert np.isinf(mo.bf16(np.array([np.inf]))[0])
    assert np.isnan(mo.bf16(np.array([np.nan]))[0])
    assert np.isinf(mo.bf16(np.array([3.4e38]))[0])  # rounds past the largest BF16


def test_gelu_identities():
    x = np.linspace(-6, 6, 241)
    assert np.allclose(mo.gelu(x) - mo.gelu(-x), x, rtol=0, atol=1e-14)   # Phi(x) + Phi(-x) = 1
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.allclose(mo.gelu_grad(x), fd, rtol=0, atol=1e-8)


def test_mlp_oracle_gradients_match_central_differences():
    H, n_in, n_out, B = 64, 12, 10, 16
    rng = np.random.default_rng(5)
    p = mo.init_params(H, n_in, n_out, seed=3).astype(np.float64)
    layers = mo.unpack(p, H)
    # nonzero biases on the real block so that every term of the chain is exercised
    for l, (W, b) in enumerate(layers):
        nout = n_out if l == 3 else b.size
        b[:nout] = rng.normal(0, 0.3, nout)
    p = mo.pack(layers)
    x = rng.normal(size=(B, n_in))
    t = rng.normal(size=(B, n_out))
    norm = (rng.normal(size=32) * 0.1, 1 + rng.random(32), rng.normal(size=32) * 0.1, 1 + rng.random(32))
    loss, g = mo.loss_and_grads(p, H, x, t, norm, n_in, n_out, emulate=False)
    assert loss > 0
    # real (non-padded) parameter indices of every layer
    real, o = [], 0
    for l, (k, n) in enumerate(mo.shapes(H)):
        kin = n_in if l == 0 else k
        nout = n_out if l == 3 else n
        real += [o + a * n + c for a in range(kin) for c in range(nout)]
        o += k * n
        real += [o + c for c in range(nout)]
        o += n
    real = np.array(real)
    pad = np.setdiff1d(np.arange(p.size), real)
    assert np.all(g[pad] == 0.0)   # padded weights never move
    pick = rng.choice(real, size=60, replace=False)
    h = 1e-6
    for i in pick:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grads(pp, H, x, t, norm, n_in, n_out, emulate=False)[0] -
              mo.loss_and_grads(pm, H, x, t, norm, n_in, n_out, emulate=False)[0]) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, g[i], fd)


# ---------------------------------------------------------------------------
# alternative chemistry schemes
# ---------------------------------------------------------------------------
def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def test_fixed_trapezoid_equals_adaptive_when_nchem_is_one(tmp_path):
    """Where the adaptive prediction gives Nchem = 1 everywhere, the
    non-adaptive Trapezoid (one sub-step of dt) is the same computation."""
    a = oracle.Oracle(cases.rsbi(str(tmp_path / "a"), dt=1e-9))
    b = oracle.Oracle(cases.rsbi(str(tmp_path / "b"), dt=1e-9, reaction="Trapezoid"))
    a.run(2)
    b.run(2)
    assert np.all(a.interior(a.nchem()) == 1)
    for n in ["CS", "T", "P", "Yi", "CMS"]:
        assert np.array_equal(_bits(a.interior(a.field(n))), _bits(b.interior(b.field(n)))), n


def test_imex_is_ee_update_for_an_inert_mixture(tmp_path):
    """Pure nitrogen has no reaction (every rate of progress is 0, so W = 0
    and the Jacobian diagonal is 0): the IMEX update is the EE update
    CS += RHS and the same recovery, bit for bit."""
    ctrl = cases.sbi(str(tmp_path), ni=16, nj=8, nk=8, r_frac=-100.0, time_adv="EE", cfl=0.5)
    a, b = oracle.Oracle(ctrl), oracle.Oracle(ctrl)
    assert np.all(a.interior(a.field("Yi"))[..., 8] == 1.0)
    for o in (a, b):
        o.set_dt(o.cfl(0.5, 1.0))
        o.field_boundary()
        o.advection()
    a.update_after_adv()
    b.imex()
    for n in ["CS", "T", "P", "D", "U", "Yi"]:
        assert np.array_equal(_bits(a.interior(a.field(n))), _bits(b.interior(b.field(n)))), n
    assert np.all(b.interior(b.field("CMS")) == 0.0)


def test_muscl1_is_first_order_copy():
    """MUSCL_1 (include/Differ.hpp:110-137): the face between i and i+1 gets
    f(i) on the left and f(i+1) on the right, in every direction, over the
    faces bc-1 .. n+bc-1 of each axis (brute force on a small random field)."""
    NI, NJ, NK, bc = 7, 6, 5, 2
    rng = np.random.default_rng(3)
    f = rng.standard_normal(NI * NJ * NK)
    F = f.reshape(NI, NJ, NK)
    for d in (1, 2, 3):
        out = np.full(2 * f.size, np.nan)
        oracle.lib().or_muscl1(d, NI, NJ, NK, bc, oracle.ptr(f), oracle.ptr(out))
        o = out.reshape(NI, NJ, NK, 2)
        sl = tuple(slice(bc - 1, n - bc) for n in (NI, NJ, NK))
        nxt = [np.roll(F, -1, axis=a) for a in range(3)][d - 1]
        assert np.array_equal(o[sl + (0,)], F[sl])
        assert np.array_equal(o[sl + (1,)], nxt[sl])
