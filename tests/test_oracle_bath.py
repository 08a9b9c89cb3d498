"""The oracle's LO phonon bath and hot-phonon Froehlich mechanisms.

Pinned three ways:
* bit-exact against emcPhononBath itself (golden KATs recorded by the
  harness: event rounds + update for the plain bath, the acoustic
  reservoir, acoustic + Ridley and per-bin lifetime profiles);
* against closed forms of the update equation (emcPhononBath.hpp:296-319):
  pure relaxation N - N0 -> (N - N0)(1 - dt/tau)^n, one event adding
  1/D_ph(q), the clamp at zero, the q^2-weighted mean, one step of the
  acoustic reservoir;
* the hot-phonon rates against the equilibrium Froehlich rates of the same
  material (pinned separately) at N = N0, and their linear dependence on N.
"""
import numpy as np
import pytest

from emc_models import MODELS, NONPARA_ISO
from oracle import oracle_py as O

PI = 3.14159265358979323846


def _bath_model(bath, n_levels=10):
    spec = dict(n_levels=n_levels, max_energy=1.0, T=300.0, field_kvcm=0.0, field_dir=(1, 0, 0),
                valleys=[dict(kind=NONPARA_ISO, mass=(0.28,) * 3, deg=1, alpha=0.3, bottom=0.0)],
                baths=[bath], mechs=[])
    return O.OracleModel(spec)


def _kat_bath(golden, name):
    nb, dq, tau, hw, T, V, ac_hw, tau_ac, to_hw, tau_to, dt = golden["bath.params"]
    cfg = dict(plain=(False, 0.0, False), acoustic=(True, 0.0, False), ridley=(True, 0.3, False),
               profile=(False, 0.0, True), acoustic_profile=(True, 0.25, True))[name]
    b = dict(n_bins=int(nb), dq=dq, tau_lo=tau, hw=hw, T=T, volume=V, acoustic=cfg[0], ac_hw=ac_hw,
             tau_ac=tau_ac, w_ridley=cfg[1], to_hw=to_hw, tau_to=tau_to)
    if cfg[2]:
        q = (np.arange(int(nb)) + 0.5) * dq
        b["tau_profile"] = 1e-12 * (0.5 + 1.5 * q / (nb * dq))
    return b, dt


@pytest.mark.parametrize("name", ["plain", "acoustic", "ridley", "profile", "acoustic_profile"])
def test_bath_update_matches_reference(name, golden):
    bath, dt = _kat_bath(golden, name)
    M = _bath_model(bath)
    g = "bath." + name + "."
    q, em, counts = golden[g + "q"], golden[g + "emission"], golden[g + "n"]
    nq = golden[g + "nq"].reshape(len(counts), -1)
    scal = golden[g + "scal"].reshape(len(counts), 6)
    pos = 0
    for r, n in enumerate(counts):
        for i in range(pos, pos + int(n)):
            M.bath_record(0, q[i], em[i])
        pos += int(n)
        M.bath_update(0, dt)
        st = M.bath(0)
        assert (st["nq"] == nq[r]).all(), (name, r)
        got = [st["mean"], st["n_ac"], st["n_to"], st["n0"], st["t_ac"], st["t_to"]]
        assert got == list(scal[r]), (name, r, got, scal[r])
    # the KAT rounds drive some bins to the clamp
    if name == "plain":
        assert (nq == 0).any()


def test_bin_of_truncates_and_clamps():
    nb, dq = 16, 5e7
    M = _bath_model(dict(n_bins=nb, dq=dq, tau_lo=1e-12, hw=0.06, volume=1e-22))
    for q, want in ((0.0, 0), (0.999 * dq, 0), (dq, 1), (3.5 * dq, 3), (15.999 * dq, 15),
                    (16 * dq, 15), (1e3 * dq, 15)):
        before = M.bath(0)["n_em"].copy()
        M.bath_record(0, q, True)
        after = M.bath(0)["n_em"]
        assert list(np.nonzero(after != before)[0]) == [want], q
    M.bath_record(0, 2.5 * dq, False)
    assert M.bath(0)["n_abs"][2] == 1.0


def test_pure_relaxation_closed_form():
    nb, dq, tau, dt = 32, 5e7, 1e-12, 3e-14
    M = _bath_model(dict(n_bins=nb, dq=dq, tau_lo=tau, hw=0.06, volume=1e-22))
    n0 = M.bath(0)["n0"]
    x = 0.06 * 1.60219e-19 / (1.38066e-23 * 300.0)
    assert abs(n0 - 1.0 / np.expm1(x)) < 1e-14 * n0
    excess = np.linspace(0.5, 3.0, nb)
    M.set_bath(0, n0 + excess)
    for _ in range(25):
        M.bath_update(0, dt)
    want = n0 + excess * (1 - dt / tau) ** 25
    assert np.allclose(M.bath(0)["nq"], want, rtol=1e-12, atol=0)


def test_single_event_adds_inverse_density_of_states():
    nb, dq, V, dt = 32, 5e7, 3e-26, 1e-15
    M = _bath_model(dict(n_bins=nb, dq=dq, tau_lo=1e-12, hw=0.06, volume=V))
    n0 = M.bath(0)["n0"]
    M.bath_record(0, 7.25 * dq, True)   # bin 7, emission
    M.bath_record(0, 20.5 * dq, False)  # bin 20, absorption
    M.bath_update(0, dt)
    nq = M.bath(0)["nq"]
    q = (np.arange(nb) + 0.5) * dq
    dph = q * q * dq * V / (2 * PI * PI)
    want = np.full(nb, n0)
    want[7] += 1 / dph[7]
    want[20] = max(0.0, n0 - 1 / dph[20])
    assert np.allclose(nq, want, rtol=1e-12, atol=1e-15)
    assert nq[20] == 0.0  # one absorption in a small volume empties the bin: clamp
    # tallies reset by the update
    st = M.bath(0)
    assert not st["n_em"].any() and not st["n_abs"].any()


def test_mean_is_q_squared_weighted():
    nb, dq = 40, 5e7
    M = _bath_model(dict(n_bins=nb, dq=dq, tau_lo=1e-12, hw=0.06, volume=1e-22))
    rng = np.random.default_rng(3)
    nq = rng.uniform(0, 5, nb)
    M.set_bath(0, nq)
    q = (np.arange(nb) + 0.5) * dq
    assert abs(M.bath(0)["mean"] - (q * q * nq).sum() / (q * q).sum()) < 1e-13 * nq.mean()


def test_acoustic_reservoir_one_step():
    nb, dq, tau, tau_ac, dt = 24, 5e7, 1e-12, 1e-11, 2e-14
    M = _bath_model(dict(n_bins=nb, dq=dq, tau_lo=tau, hw=0.06, volume=1e-22, acoustic=True,
                         ac_hw=0.03, tau_ac=tau_ac))
    st = M.bath(0)
    n0, nac_eq = st["n0"], st["n_ac"]
    x = 0.03 * 1.60219e-19 / (1.38066e-23 * 300.0)
    assert abs(nac_eq - 1 / np.expm1(x)) < 1e-14 * nac_eq
    excess = np.linspace(0.1, 1.0, nb)
    M.set_bath(0, n0 + excess, n_ac=nac_eq)
    M.bath_update(0, dt)
    q = (np.arange(nb) + 0.5) * dq
    dn_mean = (q * q * excess).sum() / (q * q).sum()
    want = nac_eq + dt * dn_mean / tau  # N_ac = N_ac_eq: no loss term yet
    assert abs(M.bath(0)["n_ac"] - want) < 1e-12 * want
    # the reservoir is now above equilibrium: the LO target follows its
    # temperature, T_ac = hw_ac / (kB ln(1 + 1/N_ac))
    t_ac = M.bath(0)["t_ac"]
    assert t_ac > 300.0
    assert abs(t_ac - 0.03 * 1.60219e-19 / (1.38066e-23 * np.log1p(1 / want))) < 1e-9 * t_ac


def test_hot_rates_equal_equilibrium_froehlich_at_n0():
    hot = O.OracleModel(MODELS["hot_polar"])
    eq = O.OracleModel(MODELS["polar_gamma"])  # same material, equilibrium N
    n_mean = hot.bath(0)["mean"]
    n0 = hot.bath(0)["n0"]
    assert abs(n_mean - n0) < 1e-14 * n0  # a few ulp from the weighted sum
    for E in np.linspace(0.001, 1.9, 97):
        for j in (0, 1):
            a, b = hot.rate(j, E), eq.rate(j, E)
            assert abs(a - b) <= 1e-14 * max(abs(b), 1e-300), (j, E, a, b)


def test_hot_rates_follow_the_bath():
    hot = O.OracleModel(MODELS["hot_polar"])
    n0 = hot.bath(0)["n0"]
    E = np.linspace(0.07, 1.9, 40)
    r0 = np.array([[hot.rate(j, e) for e in E] for j in (0, 1)])
    nb = MODELS["hot_polar"]["baths"][0]["n_bins"]
    hot.set_bath(0, np.full(nb, 3 * n0))
    r1 = np.array([[hot.rate(j, e) for e in E] for j in (0, 1)])
    # absorption ~ N, emission ~ N + 1
    assert np.allclose(r1[0], 3 * r0[0], rtol=1e-13)
    assert np.allclose(r1[1], r0[1] * (3 * n0 + 1) / (n0 + 1), rtol=1e-13)
    # the tables follow after a rebuild (tau0 = 1 / max total rate)
    tau_before = hot.tau0(0)
    hot.set_bath_mode(update=False)
    hot.bath_step(1e-15, rebuild=True)  # update relaxes toward N0, then rebuild
    assert hot.tau0(0) < tau_before


def test_energy_initialiser_statistics():
    M = O.OracleModel(MODELS["hot_polar"])
    st = M.init_philox_energy(200_000, 0.5, 0.05, 12345)
    E = st["E"]
    assert abs(E.mean() - 0.5) < 5 * 0.05 / np.sqrt(E.size)
    assert abs(E.std() - 0.05) < 0.01 * 0.05
    assert E.min() > 0.5 - 5.3 * 0.05 and E.max() < 0.5 + 5.3 * 0.05
    # |k| on the Kane dispersion and isotropic directions
    k = st["k"].reshape(-1, 3)
    kn = np.linalg.norm(k, axis=1)
    assert np.allclose(kn[:50], [M.normk(0, e) for e in E[:50]], rtol=1e-14)
    assert np.all(np.abs((k / kn[:, None]).mean(axis=0)) < 5 / np.sqrt(E.size))
    fixed = M.init_philox_energy(1000, 0.5, 0.0, 12345)
    assert (fixed["E"] == 0.5).all()


# ---- window sums (q-resolved coupling) -------------------------------------

@pytest.mark.parametrize("name", ["plain", "acoustic", "ridley", "profile", "acoustic_profile"])
def test_window_sums_and_queries_match_reference(name, golden):
    """rebuildWindowSums after update and setScreeningQ2, getNqInWindow and
    sampleQ (emcPhononBath.hpp:109-124, 203-208, 378-458), bit-exact against
    emcPhononBath on the KAT rounds' occupations, unscreened and screened."""
    bath, dt = _kat_bath(golden, name)
    M = _bath_model(bath)
    g = "bath." + name + "."
    q, em, counts = golden[g + "q"], golden[g + "emission"], golden[g + "n"]
    qw = golden[g + "qwin"].reshape(-1, 3)
    nb = bath["n_bins"]
    win, nqwin, sampq = [], [], []

    def queries():
        w, wn = M.bath_window(0)
        win.extend([w, wn])
        for qmin, qmax, r in qw:
            nqwin.append(M.nq_in_window(0, qmin, qmax))
            sampq.append(M.sample_q(0, qmin, qmax, False, r))
            sampq.append(M.sample_q(0, qmin, qmax, True, r))

    pos = 0
    for r, n in enumerate(counts):
        for i in range(pos, pos + int(n)):
            M.bath_record(0, q[i], em[i])
        pos += int(n)
        M.bath_update(0, dt)
        if r in (1, 5):
            queries()
    M.set_qs2(3e16)
    queries()
    M.set_qs2(3e16)
    M.set_qs2(0.0)
    queries()
    assert (np.concatenate(win) == golden[g + "win"]).all()
    assert (np.array(nqwin) == golden[g + "nqwin"]).all()
    assert (np.array(sampq) == golden[g + "sampq"]).all()
    assert len(win) == 8 and win[0].size == nb + 1


def _uneven_bath(qs2):
    """A 32-bin bath whose occupation falls off with q (events recorded at
    small q, one update), with screening q_s^2."""
    b = dict(n_bins=32, dq=5e7, tau_lo=1e-12, hw=0.06, volume=1e-22)
    M = _bath_model(b)
    rng = np.random.default_rng(5)
    for x in rng.random(3000) ** 2 * 32 * 5e7:
        M.bath_record(0, x, True)
    M.bath_update(0, 1e-14)
    M.set_qs2(qs2)
    return M, b


@pytest.mark.parametrize("qs2", [0.0, 3e16, 1e18])
def test_window_mean_is_the_coupling_weighted_mean(qs2):
    """getNqInWindow = sum w N / sum w over the bins the window touches, with
    w = q/(q^2 + q_s^2) at the bin centres (the Froehlich dq/q weight)."""
    M, b = _uneven_bath(qs2)
    nq = M.bath(0)["nq"]
    dq = b["dq"]
    qc = (np.arange(32) + 0.5) * dq
    w = qc / (qc ** 2 + qs2)
    for qmin, qmax in ((0.2 * dq, 3.7 * dq), (5 * dq, 20 * dq), (10.5 * dq, 31.2 * dq),
                       (2 * dq, 40 * dq)):
        lo, hi = int(np.floor(qmin / dq)), min(int(np.ceil(qmax / dq)), 32)
        want = (w[lo:hi] * nq[lo:hi]).sum() / w[lo:hi].sum()
        # differences of prefix sums: a window high in q loses the digits of
        # the sum below it
        assert abs(M.nq_in_window(0, qmin, qmax) - want) <= 1e-10 * want
    # outside the grid or degenerate: the DOS-weighted mean
    mean = M.bath(0)["mean"]
    assert M.nq_in_window(0, 3 * dq, 3 * dq) == mean
    assert M.nq_in_window(0, 40 * dq, 50 * dq) == mean


@pytest.mark.parametrize("emission", [False, True])
@pytest.mark.parametrize("qs2", [0.0, 1e18])
def test_sample_q_follows_the_transition_weight(emission, qs2):
    """Sampled q fall into each bin of a bin-aligned window with probability
    w f / sum w f, f = N (absorption) or N + 1 (emission); q stays in the
    window and grows with r."""
    M, b = _uneven_bath(qs2)
    nq = M.bath(0)["nq"]
    dq = b["dq"]
    qc = (np.arange(32) + 0.5) * dq
    w = qc / (qc ** 2 + qs2) * (nq + 1.0 if emission else nq)
    lo, hi = 2, 27
    n = 20000
    qs = np.array([M.sample_q(0, lo * dq, hi * dq, emission, (j + 0.5) / n) for j in range(n)])
    assert (np.diff(qs) >= 0).all() and qs.min() >= lo * dq and qs.max() <= hi * dq
    frac = np.bincount((qs / dq).astype(int), minlength=32)[lo:hi] / n
    want = w[lo:hi] / w[lo:hi].sum()
    assert np.abs(frac - want).max() < 2.0 / n + 1e-12


def test_qres_angle_maps_q_to_cos_theta():
    """sampleQResolvedCosTheta: cos = (kI^2 + kF^2 - q^2) / (2 kI kF) of the
    sampled q, so the momentum transfer of the new direction is q."""
    M, b = _uneven_bath(3e16)
    kI, kF = 7.0e8, 5.5e8
    for r in (0.05, 0.5, 0.95):
        c = M.qres_cos_theta(0, kI, kF, True, r)
        qq = M.sample_q(0, abs(kI - kF), kI + kF, True, r)
        assert abs(np.sqrt(kI ** 2 + kF ** 2 - 2 * kI * kF * c) - qq) < 1e-6 * qq
    assert M.qres_cos_theta(0, 0.0, kF, False, 0.3) == 1.0 - 2.0 * 0.3
