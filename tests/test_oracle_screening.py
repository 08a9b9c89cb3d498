"""The oracle's screened Froehlich kinds (emcScreenedFroehlichInteraction.hpp)
and the C-ABI's plasmon-screening state.

Pinned three ways:
* bit-exact against emcScreenedFroehlich{Absorption,Emission}3D themselves:
  the screened_polar rate tables and the 64-particle trajectory recorded by
  the harness (test_oracle_golden.py, parametrised over every model);
* the log factor against the integral it stands for, the q-weighted vertex
  int_{q-}^{q+} q / (q^2 + q_s^2) dq (numerical quadrature), and the angle
  sampler against the cumulative distribution of P(cos) ~ 1 / (A' - B cos)
  (quadrature again), so that a dropped q_s^2 or a swapped q+/q- fails;
* q_s^2 = 0 against the bare Froehlich kinds (rates of polar_gamma, the
  textbook angle (1 + f - (1 + 2f)^(1-r)) / f of a parabolic band).
The C-ABI's tables for screened_polar are bit-exact against the reference's
(test_capi_host.py); here its screening state: validation, the disabled
form and the rebuild that a new q_s^2 needs.
"""
import copy

import numpy as np
import pytest
from scipy.integrate import quad

from emc_models import MODELS
from oracle import oracle_py as O


@pytest.fixture(scope="module")
def L():
    return O.load()


@pytest.mark.parametrize("kI,kF", [(1.3e9, 0.9e9), (0.4e9, 1.1e9), (2.0e9, 1.95e9)])
@pytest.mark.parametrize("qs2", [0.0, 1e15, 1e17, 4e18])
def test_log_factor_is_the_screened_vertex_integral(L, kI, kF, qs2):
    lo, hi = abs(kI - kF), kI + kF
    ref, _ = quad(lambda q: q / (q * q + qs2), lo, hi, epsabs=0, epsrel=1e-13, limit=200)
    got = L.orc_scr_log_factor(kI, kF, qs2)
    assert abs(got - ref) <= 1e-11 * abs(ref)
    # q_s^2 = 0: the bare Froehlich ln(q+ / q-)
    if qs2 == 0.0:
        assert abs(got - np.log(hi / lo)) <= 1e-14 * abs(got)


def test_log_factor_degenerate_and_limits(L):
    # bare vertex with q- = 0 has no finite integral; the reference returns 0
    assert L.orc_scr_log_factor(1e9, 1e9, 0.0) == 0.0
    # screened with q- = 0 stays finite: 1/2 ln((4k^2 + qs2) / qs2)
    k, qs2 = 1e9, 1e17
    assert abs(L.orc_scr_log_factor(k, k, qs2) - 0.5 * np.log((4 * k * k + qs2) / qs2)) < 1e-14
    # strong screening: 2 kI kF / qs2 to first order
    kI, kF, big = 1.2e9, 0.8e9, 1e24
    assert abs(L.orc_scr_log_factor(kI, kF, big) / (2 * kI * kF / big) - 1) < 1e-5
    # monotone in q_s^2
    vals = [L.orc_scr_log_factor(kI, kF, q) for q in (0.0, 1e14, 1e16, 1e18, 1e20)]
    assert all(a > b for a, b in zip(vals, vals[1:]))


@pytest.mark.parametrize("kI,kF", [(1.3e9, 0.9e9), (0.5e9, 0.7e9)])
@pytest.mark.parametrize("qs2", [0.0, 1e16, 1e17, 3e18])
def test_cos_theta_inverts_the_screened_distribution(L, kI, kF, qs2):
    Ap, B = kI * kI + kF * kF + qs2, 2 * kI * kF
    pdf = lambda c: 1.0 / (Ap - B * c)  # noqa: E731
    norm, _ = quad(pdf, -1, 1, epsabs=0, epsrel=1e-13)
    for r in (1e-6, 0.05, 0.3, 0.5, 0.77, 0.999):
        c = L.orc_scr_cos_theta(kI, kF, qs2, r)
        assert -1.0 <= c <= 1.0
        cdf, _ = quad(pdf, -1, c, epsabs=0, epsrel=1e-13)
        assert abs(cdf / norm - r) < 1e-10, (r, c)
    # r -> 0 is backward, r -> 1 forward scattering
    assert L.orc_scr_cos_theta(kI, kF, qs2, 0.0) == pytest.approx(-1.0, abs=1e-12)
    assert L.orc_scr_cos_theta(kI, kF, qs2, 1.0) == pytest.approx(1.0, abs=1e-12)


def test_cos_theta_unscreened_is_the_froehlich_angle(L):
    # parabolic band: k ~ sqrt(E), f = 2 sqrt(Ei Ef) / (sqrt(Ei) - sqrt(Ef))^2
    # = 2 kI kF / (kI - kF)^2 (emcFroehlichInteraction.hpp:109-128).  The
    # bare sampler sends r -> 1 backward, the screened one r -> 0: the same
    # distribution with the uniform mirrored, c_scr(r) = c_bare(1 - r)
    for kI, kF in ((1.3e9, 0.9e9), (0.3e9, 0.8e9)):
        f = 2 * kI * kF / (kI - kF) ** 2
        for r in (0.01, 0.4, 0.93):
            ref = (1 + f - (1 + 2 * f) ** (1 - r)) / f
            assert abs(L.orc_scr_cos_theta(kI, kF, 0.0, r) - ref) < 1e-13


def test_cos_theta_isotropic_fallbacks(L):
    # B = 0 (one wavevector zero) and A' - B = 0 (kI = kF, no screening)
    for args in ((0.0, 1e9, 1e17), (1e9, 1e9, 0.0)):
        for r in (0.1, 0.6):
            assert L.orc_scr_cos_theta(*args, r) == 1.0 - 2.0 * r


def _with_qs2(qs2):
    spec = copy.deepcopy(MODELS["screened_polar"])
    spec["screening"] = dict(qs2=qs2)
    return spec


def test_unscreened_rates_are_the_froehlich_rates():
    """q_s^2 = 0: the screened kinds reduce to emcFroehlich{Absorption,
    Emission}3D (polar_gamma, same mode and valley; golden-pinned)."""
    scr = O.OracleModel(_with_qs2(0.0))
    bare = O.OracleModel(MODELS["polar_gamma"])
    E = np.linspace(0.001, 1.9, 200)
    for m in range(2):
        a = np.array([scr.rate(m, e) for e in E])
        b = np.array([bare.rate(m, e) for e in E])
        nz = b > 0
        assert (a[~nz] == 0).all()
        assert np.max(np.abs(a[nz] - b[nz]) / b[nz]) < 1e-13


def test_screening_lowers_every_rate():
    E = np.linspace(0.07, 1.9, 50)
    prev = None
    for qs2 in (0.0, 1e16, 1e17, 1e18):
        M = O.OracleModel(_with_qs2(qs2))
        r = np.array([[M.rate(m, e) for e in E] for m in range(2)])
        if prev is not None:
            assert (r < prev).all()
        prev = r


def _host_ctx():
    from viennaemc_amd import emcgpu
    return emcgpu, emcgpu.Context(precision=emcgpu.FP64, n_instances=2, device=-1)


def test_capi_screening_state_and_validation():
    emcgpu, ctx = _host_ctx()
    assert ctx.get_screening(0) == 0.0
    ctx.set_screening(1, 2.5e17)
    assert ctx.get_screening(1) == 2.5e17 and ctx.get_screening(0) == 0.0
    # emcPlasmonScreening(eps, false): setQs2 leaves 0
    ctx.set_screening(1, 2.5e17, enabled=False)
    assert ctx.get_screening(1) == 0.0
    for bad in (dict(qs2=-1.0), dict(qs2=float("nan")), dict(qs2=1e17, eps_static=0.0),
                dict(qs2=1e17, mode=7)):
        with pytest.raises(emcgpu.EmcGpuError):
            ctx.set_screening(0, **bad)
    with pytest.raises(emcgpu.EmcGpuError):
        ctx.set_screening(2, 1e17)


def test_capi_tables_follow_qs2_at_rebuild():
    """The C-ABI's screened tables at several q_s^2 equal the oracle's, and a
    new q_s^2 reaches the rates only through the next table build (the
    reference's reinitScatterTables)."""
    emcgpu, ctx = _host_ctx()
    spec = MODELS["screened_polar"]
    nl = spec["n_levels"]
    ctx.set_model(0, spec)
    ctx.set_screening(0, 1e17)
    ctx.build_tables(0, nl, spec["max_energy"], spec["T"])
    t17, _ = ctx.get_tables(0, 0, nl)
    ctx.set_screening(0, 1e18)
    same, _ = ctx.get_tables(0, 0, nl)
    assert (same == t17).all()
    ctx.build_tables(0, nl, spec["max_energy"], spec["T"])
    t18, tau18 = ctx.get_tables(0, 0, nl)
    M = O.OracleModel(_with_qs2(1e18))
    assert np.array_equal(t18.reshape(-1), M.cum(0).reshape(-1))
    assert tau18 == M.tau0(0)
    assert not np.array_equal(t18, t17)
