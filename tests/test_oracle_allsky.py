"""Pins of the all-sky oracle (cloud optics, AtmOptics combination, cloud
cover, phase matrix, layer R/T, ADA adding) on the CPU.

Each function is checked against the reference routines that build here
(oracle/_ref: CloudCoeff reader, Find_Index/LPoly/Interp_nD,
CRTM_AtmOptics_Combine, ASYMTX/matinv, Double_Gauss_Quadrature,
Legendre_M) or against what the mathematics fixes:
  * Matrix_Exp + doubling vs the reference's other layer solver, the
    eigen-decomposition MOM (ADA_Module.f90:262-354) on ASYMTX/matinv;
  * layer splitting: one homogeneous layer equals two halves added;
  * an isothermal column inside an isothermal enclosure radiates B(T) for
    any scattering (the discrete scheme conserves energy exactly);
  * non-scattering limit of ADA = the Schwarzschild solution;
  * interpolation exact for quadratics, table nodes reproduced;
  * closed forms of the overlap assumptions.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from crtmv3_amd import coeffio, synthetic

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def P(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(_dp)


@pytest.fixture(scope="module")
def ref(oracle_mod):
    if not os.path.exists(oracle_mod.REF_PATH):
        pytest.skip("oracle/_ref not built")
    R = ctypes.CDLL(oracle_mod.REF_PATH)
    R.ref_matinv.argtypes = [ctypes.c_int, _dp, _dp]
    R.ref_asymtx.argtypes = [ctypes.c_int, _dp, _dp, _dp]
    R.ref_gauss.argtypes = [ctypes.c_int, _dp, _dp]
    R.ref_legendre_m.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
    R.ref_interp.argtypes = [ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
    R.ref_combine.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
    R.ref_cloudcoeff_load.argtypes = [ctypes.c_char_p, ctypes.c_int, _ip]
    R.ref_cloudcoeff_fetch.argtypes = [_dp] * 16
    R.ref_aerosolcoeff_load.argtypes = [ctypes.c_char_p, ctypes.c_int, _ip]
    R.ref_aerosolcoeff_fetch.argtypes = [_ip] + [_dp] * 10
    R.ref_aerosol_type_index.argtypes = [ctypes.c_int]
    return R


@pytest.fixture(scope="module")
def cloudcoeff(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("cc"))
    path = synthetic.write_cloud_coeff(d)
    return path, coeffio.read_cloudcoeff(path)


def ref_inv(R, a):
    n = a.shape[0]
    out = np.zeros((n, n), order="F")
    af = np.asfortranarray(a)
    assert R.ref_matinv(n, af.ctypes.data_as(_dp), out.ctypes.data_as(_dp)) == 0
    return out


def ref_eig(R, a):
    n = a.shape[0]
    ev = np.zeros((n, n), order="F")
    ew = np.zeros(n)
    af = np.asfortranarray(a)
    assert R.ref_asymtx(n, af.ctypes.data_as(_dp), ev.ctypes.data_as(_dp), ew.ctypes.data_as(_dp)) == 0
    return ev, ew


# ---------------------------------------------------------------- CloudCoeff
def test_cloudcoeff_reader_matches_reference(cloudcoeff, ref):
    path, cc = cloudcoeff
    dims = (ctypes.c_int * 9)()
    b = path.encode()
    assert ref.ref_cloudcoeff_load(b, len(b), dims) == 0
    nfm, nrm, nfi, nri, nt, ndm, ndi, nl, npe = list(dims)
    assert (nfm, nrm, nfi, nri, nt, ndm, ndi, nl, npe) == (
        cc.frequency_mw.size, cc.reff_mw.size, cc.frequency_ir.size, cc.reff_ir.size, cc.temperature.size,
        cc.density_mw.size, cc.n_ir_densities, cc.n_legendre_terms, cc.n_phase_elements)
    bufs = [np.zeros(n) for n in (nfm, nfi, nrm, nri, nt, ndm, ndi)]
    bufs += [np.zeros(cc.ke_l_mw.size), np.zeros(cc.w_l_mw.size), np.zeros(cc.pcoeff_l_mw.size),
             np.zeros(cc.ke_s_mw.size), np.zeros(cc.w_s_mw.size), np.zeros(cc.pcoeff_s_mw.size),
             np.zeros(cc.ke_ir.size), np.zeros(cc.w_ir.size), np.zeros(cc.pcoeff_ir.size)]
    ref.ref_cloudcoeff_fetch(*[x.ctypes.data_as(_dp) for x in bufs])
    mine = [cc.frequency_mw, cc.frequency_ir, cc.reff_mw, cc.reff_ir, cc.temperature, cc.density_mw, None,
            cc.ke_l_mw, cc.w_l_mw, cc.pcoeff_l_mw, cc.ke_s_mw, cc.w_s_mw, cc.pcoeff_s_mw, cc.ke_ir, cc.w_ir,
            cc.pcoeff_ir]
    for a, b_ in zip(mine, bufs):
        if a is not None:
            np.testing.assert_array_equal(a.ravel(order="F"), b_)
    np.testing.assert_array_equal(bufs[6], 10.0)   # Density_IR fixed by the reader


# ---------------------------------------------------------- interpolation
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_interpolation_matches_reference(cloudcoeff, ref, oracle_mod, nd):
    _, cc = cloudcoeff
    rng = np.random.default_rng(7 + nd)
    axes = [cc.frequency_mw, cc.reff_mw, cc.temperature][:nd]
    n = np.array([a.size for a in axes] + [1] * (3 - nd), np.int32)
    z = rng.standard_normal(tuple(n[:nd]))
    xs = np.concatenate(axes)
    # oracle: cloud_opt_mw on a table whose ke is z (RAIN 3-D, SNOW 2-D over (f, r), ICE 1-D over f)
    for _ in range(20):
        pt = np.array([rng.uniform(a[0] - 0.1 * (a[-1] - a[0]), a[-1] * 1.05) for a in axes] + [0.0] * (3 - nd))
        out = ctypes.c_double()
        ref.ref_interp(nd, n.ctypes.data_as(_ip), P(xs), P(pt), P(z.ravel(order="F")), ctypes.byref(out))
        got = _oracle_interp(oracle_mod, cc, nd, z, pt)
        assert got == pytest.approx(out.value, rel=1e-13, abs=1e-13)


def _oracle_interp(orc, cc, nd, z, pt):
    """Drive oracle_cloud_optics so that its ke lookup interpolates table z."""
    c = coeffio.CloudCoeff(**{f: getattr(cc, f) for f in cc.__dataclass_fields__})
    nf, nr, nt = cc.frequency_mw.size, cc.reff_mw.size, cc.temperature.size
    if nd == 3:
        c.ke_l_mw = np.asarray(z).reshape(nf, nr, nt)
        typ, reff, temp = synthetic.RAIN_CLOUD, pt[1], pt[2]
    elif nd == 2:
        ks = np.zeros((nf, nr, cc.density_mw.size))
        ks[:, :, 0] = z
        c.ke_s_mw = ks
        typ, reff, temp = synthetic.SNOW_CLOUD, pt[1], 270.0
    else:
        ks = np.zeros((nf, nr, cc.density_mw.size))
        ks[:, 0, 2] = z
        c.ke_s_mw = ks
        typ, reff, temp = synthetic.ICE_CLOUD, 50.0, 270.0
    oc = orc.CloudCoeff(c)
    ke, w = ctypes.c_double(), ctypes.c_double()
    pc = np.zeros(17)
    assert orc.lib().oracle_cloud_optics(ctypes.byref(oc.c), 1, float(pt[0]), typ, float(reff), float(temp), 4,
                                         ctypes.byref(ke), ctypes.byref(w), P(pc)) == 0
    return ke.value


def test_interpolation_exact_for_quadratics_and_nodes(cloudcoeff, oracle_mod):
    _, cc = cloudcoeff
    f, r, t = cc.frequency_mw, cc.reff_mw, cc.temperature
    F, Rr, T = np.meshgrid(f, r, t, indexing="ij")
    z = (1 + 0.01 * F - 2e-5 * F ** 2) * (2 - 1e-3 * Rr + 3e-7 * Rr ** 2) * (1 + 0.02 * (T - 270) + 1e-4 * (T - 270) ** 2)
    rng = np.random.default_rng(3)
    for _ in range(25):
        pt = np.array([rng.uniform(f[0], f[-1]), rng.uniform(r[0], r[-1]), rng.uniform(t[0], t[-1])])
        exact = (1 + 0.01 * pt[0] - 2e-5 * pt[0] ** 2) * (2 - 1e-3 * pt[1] + 3e-7 * pt[1] ** 2) * \
            (1 + 0.02 * (pt[2] - 270) + 1e-4 * (pt[2] - 270) ** 2)
        assert _oracle_interp(oracle_mod, cc, 3, z, pt) == pytest.approx(exact, rel=1e-12)
    # table nodes are reproduced; outside the table the value is clamped
    for i, j, k in [(0, 0, 0), (5, 3, 2), (len(f) - 1, len(r) - 1, len(t) - 1)]:
        assert _oracle_interp(oracle_mod, cc, 3, z, np.array([f[i], r[j], t[k]])) == pytest.approx(z[i, j, k], rel=1e-13)
    assert _oracle_interp(oracle_mod, cc, 3, z, np.array([f[-1] * 2, r[0] / 3, t[0] - 40])) == \
        pytest.approx(z[-1, 0, 0], rel=1e-13)


def test_cloud_type_table_selection(cloudcoeff, oracle_mod):
    """WATER ignores Reff (radius index 1), ICE ignores Reff and T, SNOW /
    GRAUPEL / HAIL read density slices 1 / 2 / 3, RAIN and WATER read T."""
    _, cc = cloudcoeff
    orc = oracle_mod
    c = coeffio.CloudCoeff(**{f: getattr(cc, f) for f in cc.__dataclass_fields__})
    ks = np.array(cc.ke_s_mw)
    for d in range(ks.shape[2]):
        ks[:, :, d] = 1.0 + d
    c.ke_s_mw = ks
    oc = orc.CloudCoeff(c)

    def opt(typ, reff, temp, freq=89.0, nl=4):
        ke, w = ctypes.c_double(), ctypes.c_double()
        pc = np.zeros(17)
        orc.lib().oracle_cloud_optics(ctypes.byref(oc.c), 1, freq, typ, reff, temp, nl, ctypes.byref(ke),
                                      ctypes.byref(w), P(pc))
        return ke.value, w.value, pc[:nl + 1].copy()
    assert opt(synthetic.WATER_CLOUD, 5.0, 280.0)[0] == opt(synthetic.WATER_CLOUD, 900.0, 280.0)[0]
    assert opt(synthetic.WATER_CLOUD, 5.0, 250.0)[0] != opt(synthetic.WATER_CLOUD, 5.0, 290.0)[0]
    assert opt(synthetic.WATER_CLOUD, 5.0, 280.0)[1] == 0.0
    assert opt(synthetic.RAIN_CLOUD, 500.0, 250.0)[0] != opt(synthetic.RAIN_CLOUD, 500.0, 290.0)[0]
    assert opt(synthetic.SNOW_CLOUD, 300.0, 0.0)[0] == pytest.approx(1.0)
    assert opt(synthetic.GRAUPEL_CLOUD, 300.0, 0.0)[0] == pytest.approx(2.0)
    assert opt(synthetic.HAIL_CLOUD, 300.0, 0.0)[0] == pytest.approx(3.0)
    assert opt(synthetic.ICE_CLOUD, 30.0, 250.0)[0] == pytest.approx(3.0)
    assert opt(synthetic.ICE_CLOUD, 30.0, 250.0)[1] == 0.0
    # Legendre offset: 8 streams read slots 13..20, 4 streams slots 1..4
    k8 = opt(synthetic.RAIN_CLOUD, 800.0, 280.0, nl=8)[2]
    k4 = opt(synthetic.RAIN_CLOUD, 800.0, 280.0, nl=4)[2]
    assert k8[0] == k4[0] == 0.5
    g = k4[1] / 1.5
    np.testing.assert_allclose(k8[1:], [(2 * l + 1) / 2 * g ** l for l in range(1, 9)], rtol=2e-2)


# -------------------------------------------------------------- combine
def test_combine_matches_reference(ref, oracle_mod):
    rng = np.random.default_rng(11)
    for nl in (4, 8):
        K = 12
        od = rng.uniform(0.01, 2.0, K)
        ssa = od * rng.uniform(0, 0.9, K)
        ssa[[2, 7]] = [0.0, 5e-11]                  # below BS_THRESHOLD: untouched
        g = rng.uniform(0, 0.5, K)
        pc = np.zeros((nl + 1, K))
        for l in range(1, nl + 1):
            pc[l] = (2 * l + 1) / 2 * g ** l * ssa
        od_r, ssa_r, pc_r, sod = od.copy(), ssa.copy(), np.asfortranarray(pc.copy()), ctypes.c_double()
        ref.ref_combine(nl, K, P(od_r), P(ssa_r), pc_r.ctypes.data_as(_dp), ctypes.byref(sod))
        od_o, ssa_o, pc_o = od.copy(), ssa.copy(), np.ascontiguousarray(pc.T.copy())
        sod_o = oracle_mod.lib().oracle_combine(nl, K, P(od_o), P(ssa_o), P(pc_o))
        np.testing.assert_allclose(od_o, od_r, rtol=1e-15)
        np.testing.assert_allclose(ssa_o, ssa_r, rtol=1e-15)
        np.testing.assert_allclose(pc_o.T, pc_r, rtol=1e-15)
        assert sod_o == pytest.approx(sod.value, rel=1e-15)
        sc = ssa > 1e-10
        np.testing.assert_array_equal(od_o[~sc], od[~sc])      # below BS_THRESHOLD: untouched
        np.testing.assert_array_equal(ssa_o[~sc], ssa[~sc])


# -------------------------------------------------- quadrature, Legendre
@pytest.mark.parametrize("n", [2, 3, 4, 5, 9])
def test_gauss_quadrature(ref, oracle_mod, n):
    mu, w = np.zeros(n), np.zeros(n)
    oracle_mod.lib().oracle_gauss(n, P(mu), P(w))
    x, wl = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(mu, np.sort((x + 1) / 2), rtol=1e-13)
    np.testing.assert_allclose(w, wl[np.argsort(x)] / 2, rtol=1e-12)
    rmu, rw = np.zeros(n), np.zeros(n)
    ref.ref_gauss(n, P(rmu), P(rw))
    np.testing.assert_allclose(mu, rmu, rtol=1e-15)
    np.testing.assert_allclose(w, rw, rtol=1e-14)


def _phase(orc, ns, cos_sensor, nl, pl):
    nzm = ns + 1
    mu, wt = np.zeros(nzm), np.zeros(nzm)
    ff, bb = np.zeros(nzm * nzm), np.zeros(nzm * nzm)
    nz = orc.lib().oracle_phase_matrix(ns, cos_sensor, nl, P(pl), P(mu), P(wt), P(ff), P(bb))
    return nz, mu[:nz], wt[:nz], ff[:nz * nz].reshape(nz, nz), bb[:nz * nz].reshape(nz, nz)


@pytest.mark.parametrize("ns,nl,g", [(3, 4, 0.3), (5, 8, 0.45), (2, 2, 0.2), (9, 16, 0.5)])
def test_phase_matrix_normalisation(ref, oracle_mod, ns, nl, g):
    pl = np.array([(2 * l + 1) / 2 * g ** l for l in range(nl + 1)])
    nz, mu, w, ff, bb = _phase(oracle_mod, ns, 0.8, nl, pl)
    assert nz == ns + 1 and mu[-1] == 0.8 and w[-1] == 0.0
    # raw sums with the reference Legendre_M
    leg = np.zeros((nz, nl + 1))
    for i in range(nz):
        ref.ref_legendre_m(nl, mu[i], leg[i].ctypes.data_as(_dp))
    sgn = (-1.0) ** np.arange(nl)
    raw_f = np.maximum(np.einsum("l,il,jl->ij", pl[:nl], leg[:, :nl], leg[:, :nl]), 0) + \
        (np.einsum("l,il,jl->ij", pl[:nl], leg[:, :nl], leg[:, :nl]) < 0) * 1e-7
    rb = np.einsum("l,il,jl->ij", pl[:nl] * sgn, leg[:, :nl], leg[:, :nl])
    raw_b = np.where(rb < 0, 1e-7, rb)
    # normalised: symmetric, energy-conserving rows, each upper-triangle row a
    # common multiple of the raw sums
    np.testing.assert_allclose(ff, ff.T, rtol=0, atol=0)
    np.testing.assert_allclose(bb, bb.T, rtol=0, atol=0)
    np.testing.assert_allclose((ff + bb) @ np.where(np.arange(nz) < ns, w, 0.0), 1.0, rtol=1e-12)
    for i in range(ns):
        s = ff[i, i:] / raw_f[i, i:]
        np.testing.assert_allclose(s, s[0], rtol=1e-12)
        np.testing.assert_allclose(bb[i, i:] / raw_b[i, i:], s[0], rtol=1e-12)
    last = ff[:, -1] / raw_f[:, -1]
    assert np.all(np.isfinite(last))


def test_isotropic_phase_is_untouched(oracle_mod):
    nz, mu, w, ff, bb = _phase(oracle_mod, 5, 0.55, 8, np.r_[0.5, np.zeros(8)])
    np.testing.assert_allclose(ff[:5, :5], 0.5, rtol=1e-13)
    np.testing.assert_allclose(bb[:5, :5], 0.5, rtol=1e-13)


# ------------------------------------------------------------ layer R, T
def _ab(orc, ns, cos_sensor, nl, g, omega):
    pl = np.array([(2 * l + 1) / 2 * g ** l for l in range(nl + 1)])
    nz, mu, w, ff, bb = _phase(orc, ns, cos_sensor, nl, pl)
    s = min(omega, 0.999999)
    A = s / mu[:, None] * ff * w[None, :] - np.diag(1 / mu)
    B = s / mu[:, None] * bb * w[None, :]
    return A, B, pl


def _matrix_exp(orc, A, B, tau):
    n = A.shape[0]
    r, t = np.zeros(n * n), np.zeros(n * n)
    nd = orc.lib().oracle_matrix_exp(n, tau, P(A), P(B), P(r), P(t))
    return nd, r.reshape(n, n), t.reshape(n, n)


def _mom(R, A, B, tau):
    """MOM layer solution (ADA_Module.f90:262-354) on the reference ASYMTX/matinv."""
    ppm, ppp = A - B, A + B
    i_ppm = ref_inv(R, ppm)
    ev, ew = ref_eig(R, ppm @ ppp)
    lam = np.sqrt(np.maximum(ew, 0.0))
    evf = i_ppm @ (ev * lam[None, :])
    gp, gm = (ev + evf) / 2, (ev - evf) / 2
    ex = np.exp(-lam * tau)
    a1, a4 = gp * ex[None, :], gm * ex[None, :]
    a2 = ref_inv(R, gm) @ a1
    i_gm_a5 = ref_inv(R, gm - a1 @ a2)
    return (gp - a4 @ a2) @ i_gm_a5, (a4 - gp @ a2) @ i_gm_a5


@pytest.mark.parametrize("tau", [1e-3, 0.05, 0.7, 3.0, 25.0])
@pytest.mark.parametrize("ns,nl,omega", [(3, 4, 0.6), (5, 8, 0.95), (9, 16, 0.999)])
def test_matrix_exp_equals_mom(ref, oracle_mod, tau, ns, nl, omega):
    A, B, _ = _ab(oracle_mod, ns, 0.7, nl, 0.35, omega)
    nd, r, t = _matrix_exp(oracle_mod, A, B, tau)
    rm, tm = _mom(ref, A, B, tau)
    scale = max(np.abs(t).max(), np.abs(r).max())
    np.testing.assert_allclose(r, rm, rtol=0, atol=2e-10 * scale)
    np.testing.assert_allclose(t, tm, rtol=0, atol=2e-10 * scale)


def test_matrix_exp_doubling_semigroup_and_thin_limit(oracle_mod):
    A, B, _ = _ab(oracle_mod, 5, 0.62, 8, 0.4, 0.9)
    n = A.shape[0]
    _, r, t = _matrix_exp(oracle_mod, A, B, 2.4)
    _, r1, t1 = _matrix_exp(oracle_mod, A, B, 1.2)
    g = np.linalg.inv(np.eye(n) - r1 @ r1)
    np.testing.assert_allclose(r, r1 + t1 @ g @ r1 @ t1, atol=1e-12)
    np.testing.assert_allclose(t, t1 @ g @ t1, atol=1e-12)
    tau = 1e-6
    _, r, t = _matrix_exp(oracle_mod, A, B, tau)
    # first order in tau; the second-order term is ~ tau |A| relative
    second = tau * np.abs(A).max()
    np.testing.assert_allclose(r, tau * B, rtol=0, atol=2 * second * tau * np.abs(B).max())
    np.testing.assert_allclose(t, np.eye(n) + tau * A, rtol=0, atol=2 * (tau * np.abs(A).max()) ** 2)


def test_amom_layer_emission_limits(oracle_mod):
    orc = oracle_mod
    ns, nl = 5, 8
    nz = ns + 1
    pl = np.array([(2 * l + 1) / 2 * 0.3 ** l for l in range(nl + 1)])
    r, t, s = np.zeros(nz * nz), np.zeros(nz * nz), np.zeros(nz)
    # nearly non-scattering: source = B (1 - exp(-tau/mu))
    orc.lib().oracle_amom_layer(ns, 0.66, nl, P(pl), 0.8, 1e-9, 2.0, P(r), P(t), P(s))
    mu = np.zeros(nz); w = np.zeros(nz)
    _phase(orc, ns, 0.66, nl, pl)
    orc.lib().oracle_phase_matrix(ns, 0.66, nl, P(pl), P(mu), P(w), P(np.zeros(nz * nz)), P(np.zeros(nz * nz)))
    np.testing.assert_allclose(s, 2.0 * (1 - np.exp(-0.8 / mu)), rtol=1e-7)
    # conservative scattering (albedo capped at 0.999999): almost no emission
    orc.lib().oracle_amom_layer(ns, 0.66, nl, P(pl), 0.8, 1.0, 2.0, P(r), P(t), P(s))
    assert np.all(np.abs(s) < 2.0 * 5e-6)


# ------------------------------------------------------------ ADA column
def _ada(orc, ns, cos_sensor, nl, tau, ssa, pl, B, Bs, eps, cosmic):
    K = len(tau)
    out = ctypes.c_double()
    pl2 = np.ascontiguousarray(np.broadcast_to(pl[:nl], (K, nl)))
    rc = orc.lib().oracle_ada(ns, cos_sensor, nl, K, P(tau), P(ssa), P(pl2), P(B), Bs, eps, cosmic,
                              ctypes.byref(out))
    assert rc == 0
    return out.value


@pytest.mark.parametrize("ns,nl", [(3, 4), (5, 8), (9, 16)])
def test_ada_isothermal_enclosure(oracle_mod, ns, nl):
    """Isothermal column, cosmic background at the same B, any emissivity and
    any scattering: the radiance is B exactly."""
    rng = np.random.default_rng(ns)
    K = 30
    tau = rng.uniform(0.001, 2.0, K)
    ssa = np.where(rng.uniform(size=K) < 0.5, rng.uniform(0.0, 0.999, K), 0.0)
    pl = np.array([(2 * l + 1) / 2 * 0.4 ** l for l in range(nl + 1)])
    for eps in (1.0, 0.6, 0.1):
        v = _ada(oracle_mod, ns, 0.83, nl, tau, ssa, pl, np.full(K, 3.7), 3.7, eps, 3.7)
        assert v == pytest.approx(3.7, rel=1e-12)


def test_ada_non_scattering_equals_schwarzschild(oracle_mod):
    rng = np.random.default_rng(5)
    K = 40
    tau = rng.uniform(0.0, 0.3, K)
    B = rng.uniform(1.0, 3.0, K)
    u, eps, Bs, cosmic = 0.74, 0.55, 2.5, 0.05
    ssa = np.zeros(K)
    ssa[17] = 2e-10                       # one layer just above the threshold
    pl = np.r_[0.5, np.zeros(8)]
    v = _ada(oracle_mod, 5, u, 8, tau, ssa, pl, B, Bs, eps, cosmic)
    t = np.exp(-tau / u)
    above = np.r_[1.0, np.cumprod(t)]     # transmittance from the top of layer k to space
    total = above[-1]
    up_atm = np.sum(B * (1 - t) * above[:-1])
    down = cosmic * total + np.sum(B * (1 - t) * (total / above[1:]))
    exact = up_atm + total * (eps * Bs + (1 - eps) * down)
    assert v == pytest.approx(exact, rel=1e-8)


def test_ada_layer_splitting(oracle_mod):
    rng = np.random.default_rng(9)
    K = 12
    tau = rng.uniform(0.05, 1.5, K)
    ssa = rng.uniform(0.2, 0.95, K)
    B = rng.uniform(1.0, 3.0, K)
    pl = np.array([(2 * l + 1) / 2 * 0.45 ** l for l in range(9)])
    v = _ada(oracle_mod, 5, 0.6, 8, tau, ssa, pl, B, 2.2, 0.7, 0.1)
    k = 6
    tau2 = np.r_[tau[:k], tau[k] / 2, tau[k] / 2, tau[k + 1:]]
    ssa2 = np.r_[ssa[:k], ssa[k], ssa[k], ssa[k + 1:]]
    B2 = np.r_[B[:k], B[k], B[k], B[k + 1:]]
    v2 = _ada(oracle_mod, 5, 0.6, 8, tau2, ssa2, pl, B2, 2.2, 0.7, 0.1)
    assert v2 == pytest.approx(v, rel=1e-11)


# ------------------------------------------------------------ cloud cover
def test_cloud_cover_closed_forms(oracle_mod):
    rng = np.random.default_rng(1)
    K, nc = 20, 2
    cf = np.where(rng.uniform(size=K) < 0.5, rng.uniform(0.1, 0.9, K), 0.0)
    wc = np.where(cf[None, :] > 0, rng.uniform(0.0, 0.1, (nc, K)), 0.0)
    lwc = np.where(wc > 1e-6, wc, 0).sum(0)
    expect = {1: np.max(cf), 2: 1 - np.prod(1 - cf), 4: np.sum(cf * lwc) / np.sum(lwc), 5: 1.0}
    for ovl, e in expect.items():
        w = wc.copy()
        tcc = oracle_mod.lib().oracle_cloud_cover(K, nc, P(w), P(cf), ovl)
        assert tcc == pytest.approx(e, rel=1e-14)
        np.testing.assert_allclose(w, wc / e, rtol=1e-14)
    # maximum-random: contiguous blocks overlap maximally, separated blocks randomly
    cf2 = np.zeros(K)
    cf2[3:6] = [0.2, 0.5, 0.3]
    cf2[12:14] = [0.4, 0.1]
    w = wc.copy()
    tcc = oracle_mod.lib().oracle_cloud_cover(K, nc, P(w), P(cf2), 3)
    assert tcc == pytest.approx(1 - (1 - 0.5) * (1 - 0.4), rel=1e-14)


# ------------------------------------------------------------ full profile
@pytest.fixture(scope="module")
def atms(coeff_dir, oracle_mod, cloudcoeff):
    from conftest import load_sensor
    tau, spc = load_sensor(coeff_dir, "atms_syn")
    return oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cloudcoeff[1])


def test_allsky_without_clouds_is_clear(atms, oracle_mod):
    S, C = atms
    ps = synthetic.make_profiles(3, 60, seed=4)
    eps = np.full(22, 0.6)
    cl = synthetic.add_clouds(ps, seed=5)
    cl.cloud_fraction = np.zeros_like(cl.cloud_fraction)      # cf < 1e-12 removes the clouds
    for m in range(3):
        clr = oracle_mod.run_profile(S, ps, m, eps)
        for p_ in (ps, cl):
            o = oracle_mod.run_allsky(S, C, p_, m, eps, n_streams=4)
            np.testing.assert_allclose(o["bt"], clr["bt"], rtol=0, atol=1e-12)
            np.testing.assert_array_equal(o["n_streams"], 0)


def test_allsky_fractional_combination(atms, oracle_mod):
    """Partial cover: R = (1 - c) R_clear + c R_overcast(wc / c)."""
    S, C = atms
    ps = synthetic.add_clouds(synthetic.make_profiles(2, 80, seed=8), seed=9, cover="random")
    eps = np.full(22, 0.6)
    for m in range(2):
        o = oracle_mod.run_allsky(S, C, ps, m, eps, n_streams=8)
        c = o["total_cloud_cover"][0]
        assert 0 < c < 1
        clr = oracle_mod.run_profile(S, ps, m, eps)
        np.testing.assert_allclose(o["r_clear"], clr["radiance"], rtol=1e-14)
        ov = synthetic.add_clouds(synthetic.make_profiles(2, 80, seed=8), seed=9, cover="random")
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        ov.cloud_water_content = ov.cloud_water_content / c
        ov.cloud_fraction = np.where(ov.cloud_fraction > 0, 1.0, 0.0)
        ovo = oracle_mod.run_allsky(S, C, ov, m, eps, n_streams=8, overlap=5)
        np.testing.assert_allclose(o["radiance"], (1 - c) * clr["radiance"] + c * ovo["radiance"], rtol=1e-12)


def test_allsky_stream_choice(atms, oracle_mod):
    S, C = atms
    ps = synthetic.add_clouds(synthetic.make_profiles(2, 60, seed=12), seed=13)
    eps = np.full(22, 0.6)
    o = oracle_mod.run_allsky(S, C, ps, 0, eps, n_streams=0)
    wn = synthetic.sensor_def("atms_syn").wavenumber
    rmax = ps.cloud_effective_radius[:, :60].max()
    mie = 2 * np.pi * rmax * wn / 1e4
    expect = np.where(mie < 0.01, 2, np.where(mie < 1, 4, 6))
    np.testing.assert_array_equal(o["n_streams"][o["n_streams"] > 0], expect[o["n_streams"] > 0])


# ------------------------------------------------- all-sky K-matrix (complex step)
def test_allsky_complex_step_matches_central_differences(atms, oracle_mod):
    """Complex-step Jacobians of the all-sky forward against central
    differences for every input kind: temperature, water vapour, cloud water
    content and effective radius (scattering rain and snow), cloud fraction
    (partial cover), surface emissivity."""
    import dataclasses
    S, C = atms
    ps = synthetic.add_clouds(synthetic.make_profiles(1, 50, seed=91), seed=92, cover="random", reff_scale=8.0)
    eps = np.full(22, 0.6)
    f, k = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, n_streams=4)

    def bt(p_, e=eps):
        return oracle_mod.run_allsky(S, C, p_, 0, e, n_streams=4)["bt"]

    def cdiff(field, idx, h):
        a, b = dataclasses.replace(ps), dataclasses.replace(ps)
        va, vb = getattr(ps, field).copy(), getattr(ps, field).copy()
        va[idx] += h
        vb[idx] -= h
        setattr(a, field, va)
        setattr(b, field, vb)
        return (bt(a) - bt(b)) / (2 * h)

    wc = ps.cloud_water_content
    checked = 0
    for n in (2, 3):      # rain, snow: radius matters
        ks = np.nonzero(wc[n] > 1e-6)[0]
        for kk in ks[:2]:
            fd = cdiff("cloud_water_content", (n, kk), 1e-5 * wc[n, kk])
            assert np.abs(fd - k["wc_k"][:, n, kk]).max() < 1e-5 * max(1.0, np.abs(fd).max())
            fd = cdiff("cloud_effective_radius", (n, kk), 1e-5 * ps.cloud_effective_radius[n, kk])
            assert np.abs(fd - k["reff_k"][:, n, kk]).max() < 1e-5 * max(1e-3, np.abs(fd).max())
            fd = cdiff("cloud_fraction", kk, 1e-6)
            assert np.abs(fd - k["cf_k"][:, kk]).max() < 1e-5 * max(1.0, np.abs(fd).max())
            checked += 1
    assert checked >= 2
    for kk in (10, 30, 45):
        fd = cdiff("temperature", kk, 1e-4)
        assert np.abs(fd - k["t_k"][:, kk]).max() < 1e-6
    # water vapour in the troposphere (above, the ODPS absorber clamps are
    # active and the reference AD passes straight through them by design)
    for kk in (35, 45):
        fd = cdiff("absorber", (0, kk), 1e-4 * ps.absorber[0, kk])
        assert np.abs(fd - k["abs_k"][:, 0, kk]).max() < 1e-4 * max(1.0, np.abs(fd).max())
    fd = (bt(ps, eps + 1e-6) - bt(ps, eps - 1e-6)) / 2e-6
    np.testing.assert_allclose(fd, k["emis_k"], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("sensor,cover,n_streams,overlap,thick", [
    ("atms_syn", "random", 4, 3, False),
    ("atms_syn", "overcast", 8, 0, False),
    ("atms_syn", "random", 0, 1, True),
    ("hirs_syn", "random", 6, 2, False),
    ("hirs_syn", "random", 16, 4, False),
    ("atms_syn", "random", 16, 3, True),
])
def test_hand_adjoint_matches_complex_step(coeff_dir, oracle_mod, cloudcoeff, sensor, cover, n_streams, overlap,
                                           thick):
    """The hand-written all-sky adjoint (ADA, AMOM/Matrix_Exp doubling, phase
    normalisation, Combine, cloud-table derivatives, cloud cover) equals the
    complex-step Jacobians of the same forward to rounding: any dropped term or
    transposed operand in the reverse sweep shows up here at O(1) relative."""
    from conftest import load_sensor
    tau, spc = load_sensor(coeff_dir, sensor)
    S, C = oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cloudcoeff[1])
    mw = sensor == "atms_syn"
    ps = synthetic.add_clouds(synthetic.make_profiles(1, 60, seed=93), seed=94, cover=cover,
                              reff_scale=(25.0 if thick else 8.0) if mw else 1.0,
                              median_gm3=(0.3 if thick else 0.05) if mw else 0.005)
    L = S.n_channels
    rng = np.random.default_rng(5)
    eps = rng.uniform(0.5, 0.9, L)
    btk, rk = rng.normal(size=L), rng.normal(size=L) * 1e3
    f1, k1 = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, n_streams, overlap, btk, rk)
    f2, k2 = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, n_streams, overlap, btk, rk, method="adjoint")
    np.testing.assert_allclose(f2["bt"], f1["bt"], rtol=0, atol=1e-10)
    for key in k1:
        scale = max(np.abs(k1[key]).max(), 1e-300)
        assert np.abs(k1[key] - k2[key]).max() < 1e-11 * scale, key
    assert np.abs(k1["wc_k"]).max() > 0 and np.abs(k1["reff_k"]).max() > 0


# ------------------------------------------------------------- aerosols
@pytest.fixture(scope="module")
def aerosolcoeff(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("ac"))
    path = synthetic.write_aerosol_coeff(d)
    return path, coeffio.read_aerosolcoeff(path)


def test_aerosolcoeff_reader_matches_reference(aerosolcoeff, ref):
    """The synthetic AerosolCoeff file read by the reference's
    AerosolCoeff_Binary_ReadFile ('CRTM' scheme) and AerosolCoeff_Frequency
    gives the arrays the package reader returns, the frequencies the lookup
    interpolates in, and the reference's type-id -> table-index mapping."""
    path, ac = aerosolcoeff
    dims = np.zeros(10, np.int32)
    b = path.encode()
    assert ref.ref_aerosolcoeff_load(b, len(b), dims.ctypes.data_as(_ip)) == 0
    nW, nRH, nR, nS, nT = ac.ke.shape
    assert list(dims) == [nW, nR, nS, nT, nRH, ac.rh.size, ac.n_legendre_terms, ac.pcoeff.shape[6],
                          ac.release, ac.version]
    tid = np.zeros(nT, np.int32)
    out = [np.zeros(a.size) for a in (ac.wavelength, ac.frequency, ac.reff, ac.rsig, ac.rh, ac.ke, ac.w, ac.g,
                                       ac.kb, ac.pcoeff)]
    ref.ref_aerosolcoeff_fetch(tid.ctypes.data_as(_ip), *[o.ctypes.data_as(_dp) for o in out])
    np.testing.assert_array_equal(tid, ac.type_id)
    for o, a in zip(out, (ac.wavelength, ac.frequency, ac.reff, ac.rsig, ac.rh, ac.ke, ac.w, ac.g, ac.kb,
                          ac.pcoeff)):
        np.testing.assert_allclose(o, np.asarray(a).ravel(order="F"), rtol=1e-15, atol=0)
    assert np.all(np.diff(out[1]) > 0)     # the lookup needs ascending frequencies
    for q, t in enumerate(ac.type_id):
        assert ref.ref_aerosol_type_index(int(t)) == q + 1


def test_aerosol_table_indexing(aerosolcoeff, oracle_mod):
    """'CRTM' scheme lookup on tables linear in frequency and radius (the
    piecewise-quadratic interpolation is exact for them): the type's radius
    grid, the first RH and size-variance entries, and Legendre slot l + offset
    (4 / 6 / 8 / 16 streams -> 0 / 5 / 12 / 21) are the ones read."""
    _, ac0 = aerosolcoeff
    nW, nR, nT, nL1 = ac0.wavelength.size, ac0.reff.shape[0], ac0.type_id.size, ac0.pcoeff.shape[5]
    nRH, nS = 2, 2
    fr = ac0.frequency
    ke = np.full((nW, nRH, nR, nS, nT), -7.0)     # RH / sigma entries 2 hold poison
    w = np.full_like(ke, -7.0)
    pc = np.full((nW, nRH, nR, nS, nT, nL1, 1), -7.0)
    for t in range(nT):
        r = ac0.reff[:, t]
        ke[:, 0, :, 0, t] = 100.0 * (t + 1) + 0.01 * fr[:, None] + 3.0 * r[None, :]
        w[:, 0, :, 0, t] = 0.5 + 1e-5 * fr[:, None] + 1e-3 * r[None, :]
        for l in range(nL1):
            pc[:, 0, :, 0, t, l, 0] = 1000.0 * l + t + 1e-6 * fr[:, None]
    ac = coeffio.AerosolCoeff(ac0.type_id, ac0.type_name, ac0.wavelength, ac0.reff, np.ones((nS, nT)),
                              np.zeros(nRH), ke, w, ke, ke, pc)
    cc = oracle_mod.CloudCoeff(coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(os.path.dirname(
        aerosolcoeff[0]))), ac)
    for t_id in (synthetic.DUST_AEROSOL, synthetic.SULFATE_AEROSOL):
        t = int(np.nonzero(ac.type_id == t_id)[0][0])
        r = float(np.sqrt(ac0.reff[3, t] * ac0.reff[4, t]))
        for nl, off in ((4, 0), (6, 5), (8, 12), (16, 21)):
            k_, w_ = ctypes.c_double(), ctypes.c_double()
            p_ = np.zeros(17)
            assert oracle_mod.lib().oracle_aerosol_optics(cc.c.aerosol, 987.6, t_id, r, nl, ctypes.byref(k_),
                                                          ctypes.byref(w_), P(p_)) == 0
            assert k_.value == pytest.approx(100.0 * (t + 1) + 0.01 * 987.6 + 3.0 * r, rel=1e-12)
            assert w_.value == pytest.approx(0.5 + 1e-5 * 987.6 + 1e-3 * r, rel=1e-12)
            assert p_[0] == 0.5
            np.testing.assert_allclose(p_[1:nl + 1], [1000.0 * (l + off) + t + 1e-6 * 987.6 for l in range(1, nl + 1)],
                                       rtol=1e-12)
    k_, w_ = ctypes.c_double(), ctypes.c_double()
    p_ = np.zeros(17)
    assert oracle_mod.lib().oracle_aerosol_optics(cc.c.aerosol, 987.6, 1, 1.0, 2, ctypes.byref(k_), ctypes.byref(w_),
                                                  P(p_)) == -1      # two streams: no aerosol phase function
    assert oracle_mod.lib().oracle_aerosol_optics(cc.c.aerosol, 987.6, 99, 1.0, 4, ctypes.byref(k_), ctypes.byref(w_),
                                                  P(p_)) == -6      # type id not in the table


@pytest.fixture(scope="module")
def hirs_aer(coeff_dir, oracle_mod, cloudcoeff, aerosolcoeff):
    from conftest import load_sensor
    tau, spc = load_sensor(coeff_dir, "hirs_syn")
    return oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cloudcoeff[1], aerosolcoeff[1])


@pytest.mark.parametrize("clouds", [False, True])
@pytest.mark.parametrize("n_streams", [0, 8])
def test_aerosol_adjoint_matches_complex_step(hirs_aer, oracle_mod, clouds, n_streams):
    """Aerosol concentration / radius Jacobians (and every other block) of the
    hand-written adjoint equal the complex-step Jacobians, with and without
    clouds in the column."""
    S, C = hirs_aer
    ps = synthetic.make_profiles(1, 60, seed=201)
    if clouds:
        ps = synthetic.add_clouds(ps, seed=202, cover="random", median_gm3=0.005)
    ps = synthetic.add_aerosols(ps, seed=203)
    eps = np.full(S.n_channels, 0.97)
    btk = np.random.default_rng(3).normal(size=S.n_channels)
    f1, k1 = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, n_streams, 0, btk)
    f2, k2 = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, n_streams, 0, btk, method="adjoint")
    for key in k1:
        scale = max(np.abs(k1[key]).max(), 1e-300)
        assert np.abs(k1[key] - k2[key]).max() < 1e-11 * scale, key
    assert np.abs(k1["aconc_k"]).max() > 0 and np.abs(k1["areff_k"]).max() > 0


def test_aerosol_complex_step_matches_central_differences(hirs_aer, oracle_mod):
    import dataclasses
    S, C = hirs_aer
    ps = synthetic.add_aerosols(synthetic.make_profiles(1, 60, seed=211), seed=212)
    eps = np.full(S.n_channels, 0.97)
    _, k = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, 4, 0)

    def bt(p_):
        return oracle_mod.run_allsky(S, C, p_, 0, eps, 4)["bt"]
    for n, kk in ((0, 55), (1, 50), (2, 58)):
        for field, key in (("aerosol_concentration", "aconc_k"), ("aerosol_effective_radius", "areff_k")):
            v = getattr(ps, field)
            h = 1e-5 * v[n, kk]
            a, b = dataclasses.replace(ps), dataclasses.replace(ps)
            va, vb = v.copy(), v.copy()
            va[n, kk] += h
            vb[n, kk] -= h
            setattr(a, field, va)
            setattr(b, field, vb)
            fd = (bt(a) - bt(b)) / (2 * h)
            assert np.abs(fd - k[key][:, n, kk]).max() < 1e-5 * np.abs(fd).max(), (field, n, kk)


def test_aerosols_skip_microwave_and_need_four_streams(atms, hirs_aer, oracle_mod, aerosolcoeff, cloudcoeff):
    """Aerosol scattering is infrared-only (CRTM_AerosolScatter.f90:200); an
    infrared channel at two streams has no aerosol phase function and the
    reference fails it (:208-225)."""
    import dataclasses
    S, _ = atms
    C = oracle_mod.CloudCoeff(cloudcoeff[1], aerosolcoeff[1])
    ps = synthetic.make_profiles(1, 60, seed=221)
    pa = synthetic.add_aerosols(ps, seed=222)
    eps = np.full(22, 0.6)
    np.testing.assert_array_equal(oracle_mod.run_allsky(S, C, pa, 0, eps, 4)["bt"],
                                  oracle_mod.run_allsky(S, C, ps, 0, eps, 4)["bt"])
    Sh, Ch = hirs_aer
    with pytest.raises(RuntimeError, match="-5"):
        oracle_mod.run_allsky(Sh, Ch, synthetic.add_aerosols(ps, seed=223), 0, np.full(Sh.n_channels, 0.97), 2)


def test_empty_aerosol_slots_are_no_aerosols(hirs_aer, oracle_mod):
    """Aerosol type 0 marks an unused slot: a profile whose slots are all
    unused has n_Aerosols = 0, so two streams are allowed (the failure of
    CRTM_AerosolScatter.f90:208-225 needs an aerosol) and the result equals
    the aerosol-free one."""
    import dataclasses
    Sh, Ch = hirs_aer
    ps = synthetic.make_profiles(1, 60, seed=231)
    ps = synthetic.add_clouds(ps, seed=232, median_gm3=0.005)
    pa = synthetic.add_aerosols(ps, seed=233, aerosol_fraction=0.0)
    assert (pa.aerosol_type == 0).all()
    eps = np.full(Sh.n_channels, 0.97)
    for ns in (2, 4):
        np.testing.assert_array_equal(oracle_mod.run_allsky(Sh, Ch, pa, 0, eps, ns)["bt"],
                                      oracle_mod.run_allsky(Sh, Ch, ps, 0, eps, ns)["bt"])


# ------------------------------------------------ Include_Scattering = .FALSE.
def _emission(tau, B, Bs, eps, cosmic, u):
    """Specular-surface emission along the sensor angle (numpy): downwelling
    from the top with the cosmic background, reflected, then upwelling."""
    t = np.exp(-tau / u)
    rd = cosmic
    for k in range(tau.size):
        rd = rd * t[k] + B[k] * (1.0 - t[k])
    ru = eps * Bs + (1.0 - eps) * rd
    for k in range(tau.size - 1, -1, -1):
        ru = ru * t[k] + B[k] * (1.0 - t[k])
    return ru


def test_no_scattering_is_emission_with_absorption_depths(coeff_dir, oracle_mod, hirs_aer):
    """Include_Scattering = .FALSE. (CRTM_CloudScatter.f90:986-997,
    CRTM_AerosolScatter.f90:845-848, Common_RTSolution.f90:263): the radiance
    is the emission solution of the gas optical depth plus, per cloud and
    aerosol, the amount times the absorption ke (1 - w) of the raw table
    lookup.  Restated here in numpy from the white-box table lookups and the
    clear-sky gas depths; a (1 + w), a clamp applied before the scaling or
    an albedo left in the column fails it."""
    from conftest import load_sensor
    S, C = hirs_aer
    tau_c, spc = load_sensor(coeff_dir, "hirs_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(1, 60, seed=231), seed=232, median_gm3=0.005)
    ps = synthetic.add_aerosols(ps, seed=233)
    L, K = S.n_channels, 60
    eps = np.full(L, 0.97)
    clr = oracle_mod.run_profile(S, ps, 0, eps)
    assert clr["tau"].shape == (L, K)      # the profile reaches the top: no added layers
    o = oracle_mod.run_allsky(S, C, ps, 0, eps, n_streams=4, include_scattering=False)
    scat = oracle_mod.run_allsky(S, C, ps, 0, eps, n_streams=4)
    assert np.all(o["n_streams"] == 0)     # the emission solution, not ADA
    u = np.cos(np.radians(ps.sensor_zenith[0]))
    t = ps.temperature[:K]
    tsfc = float(np.dot([ps.land_coverage[0], ps.water_coverage[0], ps.snow_coverage[0], ps.ice_coverage[0]],
                        [ps.land_temperature[0], ps.water_temperature[0], ps.snow_temperature[0],
                         ps.ice_temperature[0]]))
    ke, w = ctypes.c_double(), ctypes.c_double()
    pc = np.zeros(17)
    absorbed = 0
    for l in range(L):
        wn = float(spc.wavenumber[l])
        tau = clr["tau"][l].copy()
        for n, typ in enumerate(ps.cloud_type[0]):
            for k in np.nonzero(ps.cloud_water_content[n, :K] > 1e-6)[0]:
                assert oracle_mod.lib().oracle_cloud_optics(ctypes.byref(C.c), 0, wn, int(typ),
                                                            float(ps.cloud_effective_radius[n, k]), float(t[k]), 4,
                                                            ctypes.byref(ke), ctypes.byref(w), P(pc)) == 0
                tau[k] += ps.cloud_water_content[n, k] * max(ke.value * (1.0 - w.value), 0.0)
                absorbed += 1
        for n, typ in enumerate(ps.aerosol_type[0]):
            for k in np.nonzero(ps.aerosol_concentration[n, :K] > 1e-15)[0]:
                assert oracle_mod.lib().oracle_aerosol_optics(C.c.aerosol, wn, int(typ),
                                                              float(ps.aerosol_effective_radius[n, k]), 4,
                                                              ctypes.byref(ke), ctypes.byref(w), P(pc)) == 0
                tau[k] += ps.aerosol_concentration[n, k] * max(ke.value * (1.0 - w.value), 0.0)
        c1, c2 = spc.planck_c1[l], spc.planck_c2[l]
        bc1, bc2 = spc.band_c1[l], spc.band_c2[l]
        B = c1 / (np.exp(c2 / (bc1 + bc2 * t)) - 1.0)
        Bs = c1 / (np.exp(c2 / (bc1 + bc2 * tsfc)) - 1.0)
        want = _emission(tau, B, Bs, eps[l], spc.cosmic_background_radiance[l], u)
        np.testing.assert_allclose(o["radiance"][l], want, rtol=1e-12)
    assert absorbed > 0
    assert np.abs(o["bt"] - scat["bt"]).max() > 1e-3      # the switch changes the answer


def test_no_scattering_leaves_absorbing_microwave_clouds_unchanged(atms, oracle_mod):
    """Microwave liquid and ice clouds have no albedo table (absorption only,
    CRTM_CloudScatter.f90:1363-1367, 1389-1392): the switch changes nothing
    for them, forward and K-matrix, to the last bit."""
    S, C = atms
    ps = synthetic.add_clouds(synthetic.make_profiles(1, 70, seed=241), seed=242,
                              types=(synthetic.WATER_CLOUD, synthetic.ICE_CLOUD))
    eps = np.full(22, 0.6)
    assert ps.cloud_water_content.max() > 0
    for ns in (0, 4):
        f1, k1 = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, ns)
        f2, k2 = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, ns, include_scattering=False)
        np.testing.assert_array_equal(f1["radiance"], f2["radiance"])
        for key in k1:
            np.testing.assert_array_equal(k1[key], k2[key])


@pytest.mark.parametrize("sensor", ["atms_syn", "hirs_syn"])
def test_no_scattering_adjoint_matches_complex_step(coeff_dir, oracle_mod, cloudcoeff, aerosolcoeff, sensor):
    """The hand-written adjoint of the absorption-only optics,
    d[ke (1 - w)] = (1 - w) dke - ke dw, equals the complex-step Jacobians
    (clouds of every type; aerosols on the infrared sensor)."""
    from conftest import load_sensor
    tau, spc = load_sensor(coeff_dir, sensor)
    S, C = oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cloudcoeff[1], aerosolcoeff[1])
    mw = sensor == "atms_syn"
    ps = synthetic.add_clouds(synthetic.make_profiles(1, 60, seed=251), seed=252, cover="random",
                              reff_scale=8.0 if mw else 1.0, median_gm3=0.05 if mw else 0.005)
    if not mw:
        ps = synthetic.add_aerosols(ps, seed=253)
    L = S.n_channels
    rng = np.random.default_rng(6)
    eps = rng.uniform(0.5, 0.9, L)
    btk = rng.normal(size=L)
    f1, k1 = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, 4, 0, btk, include_scattering=False)
    f2, k2 = oracle_mod.k_matrix_allsky(S, C, ps, 0, eps, 4, 0, btk, method="adjoint", include_scattering=False)
    np.testing.assert_allclose(f2["bt"], f1["bt"], rtol=0, atol=1e-10)
    for key in k1:
        scale = max(np.abs(k1[key]).max(), 1e-300)
        assert np.abs(k1[key] - k2[key]).max() < 1e-11 * scale, key
    assert np.abs(k1["wc_k"]).max() > 0 and np.abs(k1["reff_k"]).max() > 0
    if not mw:
        assert np.abs(k1["areff_k"]).max() > 0
