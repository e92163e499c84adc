"""Synthetic particle sets for the five benchmark configurations (seeded, numpy).

  C1  net14, ideal gas, 1000 particles, rho log-U[1e6,1e9], T log-U[1e9,6e9], X(12C)=X(16O)=0.5
  C2  net14, Helmholtz, rho log-U[1e5,5e9], T log-U[5e8,7e9], X ~ Dirichlet(0.5) over 4He/12C/16O/20Ne
  C3  net86, Helmholtz, 70% cool (T[1e8,1e9], rho[1e4,1e7]) / 30% hot (T[3e9,8e9], rho[1e7,5e9]),
      X(12C)=0.3, X(16O)=0.6, X(20Ne)=0.1, shuffled
  C4  net87, Helmholtz, Ye U[0.45,0.5], T log-U[2e9,9e9], rho log-U[1e7,1e10],
      X ~ Dirichlet(0.3) over the 14 alpha-chain nuclei
  C5  C2 + C3 + C4 particle sets in one call

Molar abundances are Y_i = X_i / A_i (as the reference's drivers set them,
test/net14_test.cpp:108-109); the initial nuclear timestep is 1e-5 (nuclear_net.hpp:51).
Species of net14 are placed in net86/net87 through the (Z, A) match of
net86_constants.hpp:129-151.
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SEED0 = 20261202
INITIAL_DT = 1e-5


def _table(name):
    with open(os.path.join(HERE, "csrc", "nnet_tables.h")) as f:
        txt = f.read()
    m = re.search(r"static const (?:double|int) " + name + r"\[\d+\] = \{([^}]*)\}", txt)
    return np.array([float(x) for x in m.group(1).split(",")])


Z14, A14 = _table("NNET14_Z"), _table("NNET14_A")
Z86, A86 = _table("NNET86_Z"), _table("NNET86_A")  # 87 entries, the last one is the electron


def net14_species_order():
    """Index in net86 of each net14 species (net86_constants.hpp:129-151)."""
    out = []
    for z, a in zip(Z14, A14):
        out.append(int(np.nonzero((Z86[:86] == z) & (A86[:86] == a))[0][0]))
    return np.array(out)


def helmholtz_Z(net):
    """Z vector handed to HelmholtzFunctor: net14 Z, net86 Z (86), net87 Z incl. the electron (87)."""
    return Z14.copy() if net == 14 else Z86[:86].copy() if net == 86 else Z86.copy()


def binding_energies(net):
    """Constant binding energies in erg/mol (BE in MeV x MeV-to-erg per mole), as the rate
    functors use them before their Coulomb/thermal corrections."""
    mev_to_erg = 9.648529392e17
    if net == 14:
        return np.array(_table("NNET14_BE_MEV")) * mev_to_erg
    be = np.array(_table("NNET86_BE_MEV")) * mev_to_erg
    if net == 87:
        be = np.append(be, np.array(_table("NNET87_ELECTRON_BE_MEV")) * mev_to_erg)
    return be


def mass_numbers(net):
    return A14.copy() if net == 14 else A86[:86].copy() if net == 86 else A86.copy()


def _logu(rng, lo, hi, n):
    return 10 ** rng.uniform(np.log10(lo), np.log10(hi), n)


def _dirichlet(rng, alpha, k, n):
    g = rng.gamma(alpha, 1.0, size=(k, n))
    return g / g.sum(axis=0, keepdims=True)


def config(cid, n=None, seed=None):
    """Returns dict(net, eos, rho, temp, Y (d x n), dt, hydro_dt, steps, mu, Z, label)."""
    rng = np.random.default_rng(SEED0 + cid if seed is None else seed)
    if cid == 1:
        n = 1000 if n is None else n
        rho, T = _logu(rng, 1e6, 1e9, n), _logu(rng, 1e9, 6e9, n)
        Y = np.zeros((14, n))
        Y[1], Y[2] = 0.5 / 12, 0.5 / 16
        return dict(net=14, eos=0, mu=10.0, Z=None, rho=rho, temp=T, Y=Y, dt=np.full(n, INITIAL_DT),
                    hydro_dt=1e-3, steps=1, label="net14 ideal-gas(mu=10) C1")
    if cid == 2:
        n = 1048576 if n is None else n
        rho, T = _logu(rng, 1e5, 5e9, n), _logu(rng, 5e8, 7e9, n)
        X = _dirichlet(rng, 0.5, 4, n)
        Y = np.zeros((14, n))
        Y[:4] = X / A14[:4, None]
        return dict(net=14, eos=1, mu=0.0, Z=helmholtz_Z(14), rho=rho, temp=T, Y=Y, dt=np.full(n, INITIAL_DT),
                    hydro_dt=1e-4, steps=10, label="net14 Helmholtz C2")
    if cid == 3:
        n = 2097152 if n is None else n
        nc = int(round(0.7 * n))
        hot = np.zeros(n, bool)
        hot[nc:] = True
        rng.shuffle(hot)
        rho = np.where(hot, _logu(rng, 1e7, 5e9, n), _logu(rng, 1e4, 1e7, n))
        T = np.where(hot, _logu(rng, 3e9, 8e9, n), _logu(rng, 1e8, 1e9, n))
        o = net14_species_order()
        Y = np.zeros((86, n))
        Y[o[1]], Y[o[2]], Y[o[3]] = 0.3 / 12, 0.6 / 16, 0.1 / 20
        return dict(net=86, eos=1, mu=0.0, Z=helmholtz_Z(86), rho=rho, temp=T, Y=Y, dt=np.full(n, INITIAL_DT),
                    hydro_dt=1e-4, steps=20, hot=hot, label="net86 Helmholtz C3")
    if cid == 4:
        n = 2097152 if n is None else n
        rho, T = _logu(rng, 1e7, 1e10, n), _logu(rng, 2e9, 9e9, n)
        ye = rng.uniform(0.45, 0.5, n)
        X = _dirichlet(rng, 0.3, 14, n)
        o = net14_species_order()
        Y = np.zeros((87, n))
        Y[o] = X / A14[:, None]
        Y[86] = ye
        return dict(net=87, eos=1, mu=0.0, Z=helmholtz_Z(87), rho=rho, temp=T, Y=Y, dt=np.full(n, INITIAL_DT),
                    hydro_dt=1e-5, steps=20, label="net87 Helmholtz C4")
    raise ValueError(cid)


def config5(n=None, seed=None, steps=100):
    """Mixed-network batch in one call: n14 net14 (C2 distribution) + n86 net86 (C3) + n87 net87 (C4)
    particles with n14 = n86 = 2 n87 (n = n14 + n86 + n87, default 2,621,440), and the per-step
    hydro dt, log-uniform in [1e-6, 1e-3], for `steps` steps.  Returns (list of three configs,
    dts)."""
    n = 2621440 if n is None else n
    n87 = n // 5
    n14 = n86 = (n - n87) // 2
    base = SEED0 + 5 if seed is None else seed
    parts = [config(2, n=n14, seed=base + 2), config(3, n=n86, seed=base + 3), config(4, n=n87, seed=base + 4)]
    rng = np.random.default_rng(base)
    return parts, _logu(rng, 1e-6, 1e-3, steps)
