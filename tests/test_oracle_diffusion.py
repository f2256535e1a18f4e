"""CPU tests that pin the oracle's diffusion step (no GPU).

The reference declares the diffusive topology (src/Atoms/Neighbors.cpp:139-277,
r_cutoff_Eb = 2.6 A and max_chemical_neighs = 50, src/Macros.hpp:860,863) and
the diffusion parameters (src/Variables.hpp:66-90) but no rate law, so the
rate law is a definition of this library and is pinned by closed forms only:

* the diffusive lists equal a numpy enumeration of the H-H pairs within 2.6 A;
* without interactions a dimer exchanges by Fick's law with the Arrhenius rate
  D0 exp(-Eb / k_B T) / r^2, computed here from the physical constants;
* a dimer in detailed balance, ln(x/(1-x)) + mu equal on both sites, is
  stationary, and a chemical-potential gradient drives H downhill;
* a masked site (x < 5e-4) exchanges by Fick's law only;
* pair fluxes are antisymmetric (mass conservation), a perfect crystal with
  uniform occupancy is stationary, and the implicit step conserves H exactly
  and satisfies its own backward-Euler equation.
"""
import numpy as np
import pytest

from conftest import RC

import dmdgpu.lattice as lat

K_B = 8.617332478e-5   # eV/K, src/Macros.hpp:849
D0, EB = 1e-7, 0.4      # m^2/s, eV


def dimer(x1, x2, r=2.0, T=300.0):
    b = 1.0 / (K_B * T)
    return {"n": 2, "q": np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]]),
            "sigma": np.array([0.15, 0.15]), "x": np.array([x1, x2]),
            "species": np.array([1, 1], dtype=np.int32), "beta": np.array([b, b]),
            "gamma": np.zeros(2), "lo": np.zeros(3), "hi": np.full(3, 30.0),
            "periodic": np.zeros(3, dtype=np.int32)}


def arrhenius(r, T=300.0):
    """Jump rate of one pair (1/s): D0 exp(-Eb / k_B T) / r^2, D0 in A^2/s."""
    return D0 * 1e20 * np.exp(-EB / (K_B * T)) / r ** 2


def rates(oracle, sd, mu):
    st, cnt, lst = oracle.diff_neighbors(sd)
    assert st == 0
    return oracle.diff_rates(sd, mu, cnt, lst, D0, EB)


def test_diffusive_lists_match_enumeration(oracle):
    for sd in (lat.hcp_mgh(4, 3, 3, seed=1), lat.nanowire(15, seed=3)):
        st, cnt, lst = oracle.diff_neighbors(sd)
        assert st == 0
        q, L = sd["q"], sd["hi"] - sd["lo"]
        h = np.where(sd["species"] == 1)[0]
        d = q[h][:, None, :] - q[h][None, :, :]
        per = sd["periodic"].astype(bool)
        d[..., per] -= L[per] * np.rint(d[..., per] / L[per])
        r = np.sqrt((d ** 2).sum(-1))
        for a, i in enumerate(h):
            want = set(h[(r[a] <= 2.6) & (h != i)].tolist())
            assert set(lst[i, :cnt[i]].tolist()) == want
        assert (cnt[sd["species"] == 12] == 0).all()
        # symmetric lists
        for i in h:
            for j in lst[i, :cnt[i]]:
                assert i in lst[j, :cnt[j]]


def test_fick_dimer_has_arrhenius_rate(oracle):
    for r, (x1, x2) in ((2.0, (0.1, 0.3)), (1.3, (0.7, 0.2)), (2.5, (0.05, 0.05))):
        sd = dimer(x1, x2, r)
        xd = rates(oracle, sd, np.zeros(2))
        g = arrhenius(r)
        assert xd[0] == pytest.approx(g * (x2 - x1), rel=1e-12, abs=1e-300)
        assert xd[1] == -xd[0]
    # the rate follows exp(-Eb / k_B T) in temperature
    lo = dimer(0.1, 0.3)
    hi = dimer(0.1, 0.3, T=600.0)
    ratio = rates(oracle, hi, np.zeros(2))[0] / rates(oracle, lo, np.zeros(2))[0]
    assert ratio == pytest.approx(np.exp(EB / K_B * (1 / 300.0 - 1 / 600.0)), rel=1e-12)


def test_detailed_balance_and_downhill_flow(oracle):
    ell = lambda x: np.log(x / (1 - x))
    x1, mu1, mu2 = 0.2, -120.0, -123.5
    x2 = 1.0 / (1.0 + np.exp(-(ell(x1) + mu1 - mu2)))  # l(x2) + mu2 = l(x1) + mu1
    sd = dimer(x1, x2)
    xd = rates(oracle, sd, np.array([mu1, mu2]))
    assert abs(xd[0]) <= 1e-12 * arrhenius(2.0)
    # equal occupancies, higher chemical potential on site 0: H leaves site 0
    xd = rates(oracle, dimer(0.2, 0.2), np.array([-100.0, -101.0]))
    assert xd[0] < 0 and xd[1] == -xd[0]
    # the flux equals G L dmu with L = x(1-x) for equal occupancies
    assert xd[0] == pytest.approx(arrhenius(2.0) * 0.2 * 0.8 * (-1.0), rel=1e-12)


def test_masked_site_exchanges_by_fick_only(oracle):
    sd = dimer(2e-4, 0.3)
    xd = rates(oracle, sd, np.array([0.0, -120.0]))
    assert xd[0] == pytest.approx(arrhenius(2.0) * (0.3 - 2e-4), rel=1e-12)


def test_mass_conservation_and_stationarity(oracle):
    sd = lat.hcp_mgh(4, 3, 3, seed=9, x_h=(0.0, 0.3))
    st, nnb, nbr = oracle.neighbors(sd, RC)
    res = oracle.fact_residual(oracle.system(sd, nnb, nbr))
    mu = oracle.diff_mu(sd, res["Rx"])
    xd = rates(oracle, sd, mu)
    assert abs(xd.sum()) <= 1e-12 * np.abs(xd).sum()
    # perfect crystal, uniform occupancy: every T site is equivalent, mu is
    # uniform and nothing flows
    pc = lat.hcp_mgh(4, 3, 3, seed=1, pos_noise=0.0, sigma_jitter=0.0, x_h=(0.1, 0.1))
    st, nnb, nbr = oracle.neighbors(pc, RC)
    res = oracle.fact_residual(oracle.system(pc, nnb, nbr))
    mu = oracle.diff_mu(pc, res["Rx"])
    h = pc["species"] == 1
    assert np.ptp(mu[h]) <= 1e-9 * np.abs(mu[h]).max()
    xd = rates(oracle, pc, mu)
    scale = arrhenius(2.3) * 0.09 * np.abs(mu[h]).max()
    assert np.abs(xd).max() <= 1e-9 * scale


def test_implicit_step_conserves_and_solves_backward_euler(oracle):
    sd = lat.hcp_mgh(4, 3, 3, seed=1, x_h=(0.0, 0.3))
    st, nnb, nbr = oracle.neighbors(sd, RC)
    x, sweeps, dmax = oracle.diffusion_step(sd, nnb, nbr, dt=1e-9, D0=D0, Eb=EB, tol=1e-12)
    assert dmax <= 1e-12 and sweeps < 20
    h = sd["species"] == 1
    assert abs(x[h].sum() - sd["x"][h].sum()) <= 1e-12 * sd["x"][h].sum()
    assert np.abs(x - sd["x"]).max() > 1e-4  # something moved
    # backward Euler: x - x_n = dt xdot(x), with the activity of x_n
    s = dict(sd)
    s["x"] = x
    act = sd["x"] >= 5e-4
    s2 = dict(s)
    s2["x"] = np.where(act, x, 0.0)
    res = oracle.fact_residual(oracle.system(s2, nnb, nbr))
    mu = oracle.diff_mu(s2, res["Rx"])
    st, cnt, lst = oracle.diff_neighbors(sd)
    xd = oracle.diff_rates(s, mu, cnt, lst, D0, EB, xa=sd["x"])
    assert np.abs(x - sd["x"] - 1e-9 * xd)[h].max() <= 1e-11
