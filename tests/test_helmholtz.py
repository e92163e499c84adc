"""Helmholtz EOS: the generated table's physics, the interpolation, and parity with the
reference's helmholtzEos compiled against the same table."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import workloads
from oracle import D, _p

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "sph-exa-nnet_amd", "data", "helm_table.bin")
IMAX, JMAX = 541, 201

needs_table = pytest.mark.skipif(not os.path.exists(TABLE), reason="Helmholtz table not generated (build())")


def grid():
    return 10 ** (-12 + np.arange(IMAX) * 0.05), 10 ** (3 + np.arange(JMAX) * 0.05)


@needs_table
def test_table_degenerate_limit():
    """T = 1e3 K, rho Ye >= 1e3: the electron pressure rho^2 df/drho is the T = 0 Chandrasekhar
    pressure (closed form)."""
    tab = np.fromfile(TABLE).reshape(21, IMAX, JMAX)
    h, me, c, avo = 6.6260689633e-27, 9.1093821545e-28, 2.99792458e10, 6.0221417930e23
    din, _ = grid()
    for i in range(300, IMAX, 40):
        n = din[i] * avo
        x = (3 * n / (8 * np.pi)) ** (1 / 3) * h / (me * c)
        P0 = np.pi * me ** 4 * c ** 5 / (3 * h ** 3) * (x * (2 * x * x - 3) * np.sqrt(1 + x * x) + 3 * np.arcsinh(x))
        assert tab[1, i, 0] * din[i] ** 2 == pytest.approx(P0, rel=1e-8)


@needs_table
def test_table_nondegenerate_limit():
    """Low density, T = 1e7..3e7 K (no pairs): P = n k T exactly (classical gas, relativistic or
    not) and cv = 3/2 k N_A (1 + 5/2 b - 15/4 b^2 + O(b^3)) per gram of Ye = 1 gas,
    b = kT / m_e c^2 (asymptotic expansion of the Juettner-gas energy, K1/K2 Bessel ratio)."""
    tab = np.fromfile(TABLE).reshape(21, IMAX, JMAX)
    kerg, avo = 1.380650424e-16, 6.0221417930e23
    din, T = grid()
    mec2 = 9.1093821545e-28 * 2.99792458e10 ** 2
    for i in (0, 40, 80):
        for j in (80, 85, 90):
            b = kerg * T[j] / mec2
            P = tab[1, i, j] * din[i] ** 2
            assert P == pytest.approx(din[i] * avo * kerg * T[j], rel=1e-9)
            assert -tab[4, i, j] * T[j] == pytest.approx(1.5 * kerg * avo * (1 + 2.5 * b - 3.75 * b * b), rel=20 * b ** 3)


@needs_table
def test_table_maxwell_relations(oracle_lib):
    """Thermodynamic consistency of the full EOS on the generated table (helmholtz.cpp:944-948)."""
    oracle_lib.helm_load(TABLE)
    rng = np.random.default_rng(1)
    worst = 0
    for _ in range(1500):
        T, rho = 10 ** rng.uniform(7, 10.5), 10 ** rng.uniform(0, 11)
        abar = rng.uniform(1, 60)
        zbar = rng.uniform(0.4, 0.5) * abar
        out = oracle_lib.helm_raw(1 / abar, zbar / abar, T, rho)
        worst = max(worst, np.max(np.abs(out[6:9])))
        assert out[0] > 0 and out[1] > 0
    assert worst < 1e-10


def test_biquintic_reproduces_quintic_polynomials(oracle_lib):
    """A table filled from f = P(d) Q(T) with P, Q of degree <= 5 (values and exact derivatives)
    is reproduced exactly (to rounding) by the biquintic Hermite interpolation."""
    din, T = grid()
    rng = np.random.default_rng(4)
    # work in local scaled coordinates so the polynomial is well conditioned on each cell scale
    pc, qc = rng.standard_normal(6), rng.standard_normal(6)
    d0, t0 = 1e3, 1e8
    P = np.polynomial.Polynomial(pc)
    Q = np.polynomial.Polynomial(qc)
    dd = lambda k: P.deriv(k)(din / d0) / d0 ** k if k else P(din / d0)
    tt = lambda k: Q.deriv(k)(T / t0) / t0 ** k if k else Q(T / t0)
    arr = np.zeros((21, IMAX, JMAX))
    # f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt
    orders = [(0, 0), (1, 0), (0, 1), (2, 0), (0, 2), (1, 1), (2, 1), (1, 2), (2, 2)]
    for q, (a, b) in enumerate(orders):
        arr[q] = np.outer(dd(a), tt(b))
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        arr.tofile(f)
        path = f.name
    try:
        oracle_lib.helm_load(path)
        oracle_lib.L.oracle_helm_interp.argtypes = [C.c_double, C.c_double, D]
        out = np.zeros(5)
        for x in [(3.7e2, 2.2e7), (1.1e3, 9.1e7), (5.5e3, 4.3e8), (2.0e2, 1.3e8)]:
            oracle_lib.L.oracle_helm_interp(x[0], x[1], _p(out))
            u, v = x[0] / d0, x[1] / t0
            exp = [P(u) * Q(v), P.deriv(1)(u) / d0 * Q(v), P(u) * Q.deriv(1)(v) / t0, P(u) * Q.deriv(2)(v) / t0 ** 2,
                   P.deriv(1)(u) / d0 * Q.deriv(1)(v) / t0]
            for got, want in zip(out, exp):
                assert got == pytest.approx(want, rel=1e-8, abs=1e-10 * abs(exp[0]) + 1e-300)
    finally:
        os.unlink(path)
        if os.path.exists(TABLE):
            oracle_lib.helm_load(TABLE)


@needs_table
def test_oracle_helmholtz_vs_reference(oracle_lib, ref_lib):
    if not ref_lib.has_helmholtz:
        pytest.skip("reference built without a Helmholtz table")
    oracle_lib.helm_load(TABLE)
    rng = np.random.default_rng(3)
    for k in range(900):
        net = (14, 86, 87)[k % 3]
        Z = workloads.helmholtz_Z(net)
        Y = rng.random(len(workloads.mass_numbers(net))) * 0.03
        T, rho = 10 ** rng.uniform(7.5, 10.2), 10 ** rng.uniform(2, 11)
        a, b = oracle_lib.eos(1, Y, T, rho, Z=Z), ref_lib.eos(1, Y, T, rho, Z=Z)
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), (net, T, rho)


def _fixture(name):
    path = os.path.join(ROOT, "tests", "golden", name)
    g = np.load(path)
    from make_golden_digest import digest
    if not os.path.exists(TABLE) or digest(TABLE) != str(g["digest"]):
        pytest.skip("Helmholtz fixtures were minted from a different table")
    return g


@needs_table
def test_helmholtz_golden_eos(oracle_lib):
    g = _fixture("helm_eos.npz")
    oracle_lib.helm_load(TABLE)
    for p, out, net in zip(g["pts"], g["out"], g["net"]):
        d = len(workloads.mass_numbers(int(net)))
        got = oracle_lib.eos(1, p[2:2 + d], p[0], p[1], Z=workloads.helmholtz_Z(int(net)))
        assert np.array_equal(got.view(np.int64), out.view(np.int64))


@needs_table
@pytest.mark.parametrize("cid", [2, 3, 4])
def test_helmholtz_golden_particles(oracle_lib, cid):
    g = _fixture(f"helm_particles_c{cid}.npz")
    oracle_lib.helm_load(TABLE)
    net = int(g["net"])
    for k in range(len(g["rho"])):
        Y, T, dt, st, _ = oracle_lib.particle(net, g["Y0"][k], g["T0"][k], g["rho"][k], float(g["hydro_dt"]), 1e-5,
                                              eos_kind=1, Z=g["Z"], max_iter=10000)
        assert np.array_equal(Y.view(np.int64), g["Y"][k].view(np.int64)), (cid, k)
        assert T == g["T"][k] and dt == g["dt"][k] and np.array_equal(st[:4], g["stats"][k][:4]), (cid, k)


@needs_table
def test_helm_eos_batch_golden(oracle_lib):
    """computeHelmholtzImpl restated (oracle_helm_eos_batch) against the reference's outputs,
    bit for bit, including the out-of-table corners and a NaN sound speed."""
    g = _fixture("helm_eos_batch.npz")
    oracle_lib.helm_load(TABLE)
    for net in (14, 86, 87):
        got = np.array(oracle_lib.helm_eos_batch(g[f"Z{net}"], g[f"T{net}"], g[f"rho{net}"], g[f"Y{net}"]))
        assert np.array_equal(got.view(np.int64), g[f"out{net}"].view(np.int64)), net


@needs_table
def test_helm_eos_batch_vs_reference(oracle_lib, ref_lib):
    if not ref_lib.has_helmholtz:
        pytest.skip("reference built without a Helmholtz table")
    oracle_lib.helm_load(TABLE)
    for cid in (2, 3, 4):
        cfg = workloads.config(cid, n=2000, seed=99 + cid)
        a = np.array(oracle_lib.helm_eos_batch(cfg["Z"], cfg["temp"], cfg["rho"], cfg["Y"]))
        b = np.array(ref_lib.helm_eos_batch(cfg["Z"], cfg["temp"], cfg["rho"], cfg["Y"]))
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), cid


@needs_table
def test_helm_eos_batch_is_the_pointwise_eos(oracle_lib):
    """The batch form equals the single-point EOS (oracle_eos, pinned separately) at abar_ =
    sum(Y), zbar_ = sum(Y Z): the sums and the field mapping u, cv, p, c, dpdT are right."""
    oracle_lib.helm_load(TABLE)
    cfg = workloads.config(4, n=50, seed=5)
    out = np.array(oracle_lib.helm_eos_batch(cfg["Z"], cfg["temp"], cfg["rho"], cfg["Y"]))
    for i in range(50):
        e = oracle_lib.eos(1, cfg["Y"][:, i], cfg["temp"][i], cfg["rho"][i], Z=cfg["Z"])
        # oracle_eos order: cv, dpdT, p, cp, c, u, ...
        assert np.array_equal(out[:, i], [e[5], e[0], e[2], e[4], e[1]], equal_nan=True)


TEXT_TABLE = os.path.join(ROOT, "sph-exa-nnet_amd", "data", "helm_table.dat")


@pytest.mark.skipif(not os.path.exists(TEXT_TABLE) or not os.path.exists(TABLE), reason="tables not generated")
def test_library_reads_the_reference_text_table(tmp_path):
    """nnet_helm_table_read (host code of the library, no device) parses the reference's text
    table -- STR( ... )-wrapped as the reference #includes it, or bare -- into exactly the binary
    table's values, and rejects a truncated or over-long file."""
    import nnet_mi355x as nm
    want = np.fromfile(TABLE, dtype=np.float64).reshape(21, 541, 201)
    assert np.array_equal(nm.helm_table_read(TABLE).view(np.int64), want.view(np.int64))
    got = nm.helm_table_read(TEXT_TABLE)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    text = open(TEXT_TABLE).read()
    body = text[text.index("STR(") + 4:text.rindex(")")]
    bare = tmp_path / "bare.dat"
    bare.write_text(body)
    assert np.array_equal(nm.helm_table_read(str(bare)).view(np.int64), want.view(np.int64))
    short = tmp_path / "short.dat"
    short.write_text(body[:len(body) // 2])
    with pytest.raises(nm.NnetError):
        nm.helm_table_read(str(short))
    extra = tmp_path / "extra.dat"
    extra.write_text(body + " 1.0\n")
    with pytest.raises(nm.NnetError):
        nm.helm_table_read(str(extra))
    with pytest.raises(nm.NnetError):
        nm.helm_table_read(str(tmp_path / "missing.dat"))
