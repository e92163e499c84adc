"""TEST INFRASTRUCTURE — ctypes binding of the CPU oracle (liboracle.so) and, where it was built,
of the reference build (oracle/_ref/libnnet_ref.so).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module, and only as the checker / CPU baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libnnet_ref.so")
DIMS = {14: 14, 86: 86, 87: 87}

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def _p(a):
    return a.ctypes.data_as(D)


def _pi(a):
    return a.ctypes.data_as(I)


class _Lib:
    def __init__(self, path, prefix):
        self.L = C.CDLL(path)
        self.pre = prefix
        L, pre = self.L, prefix
        g = lambda n: getattr(L, pre + n)
        g("rates").argtypes = [C.c_int, D, C.c_double, C.c_double, D, D, D]
        g("solve").argtypes = [D, D, D, C.c_int]
        g("eos").argtypes = [C.c_int, C.c_double, D, C.c_int, D, C.c_double, C.c_double, D]
        g("prepare").argtypes = [C.c_int, C.c_int, C.c_double, D, C.c_int, D, C.c_double, D, C.c_double, C.c_double,
                                 C.c_double, C.c_double, D, D]
        g("prepare").restype = C.c_int
        g("particle").argtypes = [C.c_int, C.c_int, C.c_double, D, C.c_int, D, D, C.c_double, C.c_double, C.c_double,
                                  D, C.c_long, D, D, C.c_int]
        g("particle").restype = C.c_long
        g("reaction_list").argtypes = [C.c_int, I, I, I, I]
        g("num_reactions").argtypes = [C.c_int]

    def fn(self, name):
        return getattr(self.L, self.pre + name)

    def num_reactions(self, net):
        return self.fn("num_reactions")(net)

    def reaction_list(self, net):
        nR = self.num_reactions(net)
        rb = np.zeros(nR + 1, np.int32)
        pb = np.zeros(nR, np.int32)
        ids = np.zeros(4 * nR + 8, np.int32)
        nums = np.zeros(4 * nR + 8, np.int32)
        k = self.fn("reaction_list")(net, _pi(rb), _pi(pb), _pi(ids), _pi(nums))
        return rb, pb, ids[:k], nums[:k]

    def rates(self, net, Y, T, rho):
        d = DIMS[net]
        nR = self.num_reactions(net)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        BE, r, dr = np.zeros(d), np.zeros(nR), np.zeros(nR)
        self.fn("rates")(net, _p(Y), T, rho, _p(BE), _p(r), _p(dr))
        return BE, r, dr

    def solve(self, M, b):
        n = len(b)
        Mf = np.array(M, dtype=np.float64, order="F", copy=True)
        bb = np.array(b, dtype=np.float64)
        x = np.zeros(n)
        self.fn("solve")(Mf.ctypes.data_as(D), _p(bb), _p(x), n)
        return x

    def eos(self, kind, Y, T, rho, mu=10.0, Z=None):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        Zc = np.ascontiguousarray(Z if Z is not None else np.zeros(1), dtype=np.float64)
        out = np.zeros(13)
        self.fn("eos")(kind, mu, _p(Zc), len(Zc) if Z is not None else 0, _p(Y), T, rho, _p(out))
        return out

    def prepare(self, net, Y, T, Yg, Tg, rho, dt, eos_kind=0, mu=10.0, Z=None, drho_dt=0.0):
        d = DIMS[net]
        n = d + 1
        Zc = np.ascontiguousarray(Z if Z is not None else np.zeros(1), dtype=np.float64)
        M, R = np.zeros(n * n), np.zeros(n)
        rc = self.fn("prepare")(net, eos_kind, mu, _p(Zc), len(Zc) if Z is not None else 0,
                                _p(np.ascontiguousarray(Y, dtype=np.float64)), T,
                                _p(np.ascontiguousarray(Yg, dtype=np.float64)), Tg, rho, drho_dt, dt, _p(M), _p(R))
        return rc, M.reshape(n, n, order="F"), R

    def helm_eos_batch(self, Z, temp, rho, Y, nthreads=0):
        """computeHelmholtzImpl over the columns of Y (dim x n, C-contiguous rows): returns the
        arrays u, cv, p, c, dpdT."""
        f = self.fn("helm_eos_batch")
        f.argtypes = [C.c_size_t, C.c_int, D, D, D, C.POINTER(D), D, D, D, D, D, C.c_int]
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        dim, n = Y.shape
        Zc = np.ascontiguousarray(Z, dtype=np.float64)
        assert Zc.shape == (dim,)
        temp = np.ascontiguousarray(temp, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        rows = (D * dim)(*[Y[s].ctypes.data_as(D) for s in range(dim)])
        out = [np.zeros(n) for _ in range(5)]
        f(n, dim, _p(Zc), _p(temp), _p(rho), rows, *[_p(o) for o in out], nthreads)
        return tuple(out)

    def particle(self, net, Y, T, rho, dt_tot, dt, eos_kind=0, mu=10.0, Z=None, max_iter=10000, trace_cap=0,
                 drho_dt=0.0):
        """One particle through the substep solver.  Returns (Y, T, dt, stats, trace); stats = status,
        accepted substeps, Newton iterations, rejected attempts, longest run of rejected attempts."""
        Y = np.array(Y, dtype=np.float64)
        Tt, dtt = np.array([T], np.float64), np.array([dt], np.float64)
        st = np.zeros(5)
        tr = np.zeros(6 * max(trace_cap, 1))
        Zc = np.ascontiguousarray(Z if Z is not None else np.zeros(1), dtype=np.float64)
        n = self.fn("particle")(net, eos_kind, mu, _p(Zc), len(Zc) if Z is not None else 0, _p(Y), _p(Tt), rho,
                                drho_dt, dt_tot, _p(dtt), max_iter, _p(st), _p(tr) if trace_cap else None, trace_cap)
        return Y, Tt[0], dtt[0], st, tr[:6 * min(n, trace_cap)].reshape(-1, 6)


class Oracle(_Lib):
    def __init__(self, path=ORACLE_SO):
        super().__init__(path, "oracle_")
        L = self.L
        L.oracle_helm_load.argtypes = [C.c_char_p]
        L.oracle_helm_raw.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, D]
        L.oracle_batch.argtypes = [C.c_int, C.c_int, C.c_double, D, C.c_int, C.c_size_t, D, D, C.POINTER(D), D, D,
                                   C.c_double, C.c_double, C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32), C.c_int, C.c_long, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint8)]
        L.oracle_set_solver_model.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int]
        L.oracle_solver_counts.argtypes = [C.POINTER(C.c_long)]
        L.oracle_dynamic_batch_size.argtypes = [C.c_size_t, C.c_int]

    def helm_load(self, path):
        rc = self.L.oracle_helm_load(path.encode())
        if rc != 0:
            raise RuntimeError(f"oracle_helm_load({path}) = {rc}")

    def helm_raw(self, abar_, zbar_, T, rho):
        out = np.zeros(13)
        self.L.oracle_helm_raw(abar_, zbar_, T, rho, _p(out))
        return out

    def batch(self, net, rho, temp, Y, dt, hydro_dt, previous_dt=None, eos_kind=0, mu=10.0, Z=None, rho_m1=None,
              use_drhodt=False, nthreads=0, max_iter=10000):
        """In-place batch integration of numpy arrays (Y: d x n, C-contiguous rows).
        Returns (status, substeps, newton); self.last_max_run holds, per particle, the longest run
        of consecutive rejected attempts."""
        n = len(rho)
        d = DIMS[net]
        assert Y.shape == (d, n) and Y.flags["C_CONTIGUOUS"] and Y.dtype == np.float64
        rows = (D * d)(*[Y[s].ctypes.data_as(D) for s in range(d)])
        status = np.zeros(n, np.uint8)
        sub = np.zeros(n, np.uint32)
        newt = np.zeros(n, np.uint32)
        mrun = np.zeros(n, np.uint32)
        redone = np.zeros(n, np.uint8)
        rm1 = rho_m1 if rho_m1 is not None else rho.copy()
        Zc = np.ascontiguousarray(Z if Z is not None else np.zeros(1), dtype=np.float64)
        self.L.oracle_batch(net, eos_kind, mu, _p(Zc), len(Zc) if Z is not None else 0, n, _p(rho), _p(rm1), rows,
                            _p(temp), _p(dt), hydro_dt, hydro_dt if previous_dt is None else previous_dt,
                            int(use_drhodt), status.ctypes.data_as(C.POINTER(C.c_uint8)),
                            sub.ctypes.data_as(C.POINTER(C.c_uint32)), newt.ctypes.data_as(C.POINTER(C.c_uint32)),
                            nthreads, max_iter, mrun.ctypes.data_as(C.POINTER(C.c_uint32)),
                            redone.ctypes.data_as(C.POINTER(C.c_uint8)))
        self.last_max_run = mrun
        self.last_redone = redone
        return status, sub, newt


    def set_solver_model(self, mode, tol=1e-11, sweeps=1, inner=8):
        """Linear-solve model of batch(): 0 FP64 (the reference), 1 FP32 LU + FP64 refinement (the
        device's mixed path), 2 GMRES-IR; a particle whose mixed solve does not converge is
        integrated again in FP64 (self.last_redone marks it)."""
        self.L.oracle_set_solver_model(mode, tol, sweeps, inner)

    def solver_counts(self):
        out = (C.c_long * 3)()
        self.L.oracle_solver_counts(out)
        return {"solves": out[0], "converged": out[1], "gmres_iterations": out[2]}


class Reference(_Lib):
    def __init__(self, path=REF_SO):
        super().__init__(path, "ref_")
        self.L.ref_has_helmholtz.restype = C.c_int
        self.L.ref_batch_guarded.argtypes = [C.c_int, C.c_int, C.c_double, D, C.c_int, C.c_size_t, D, C.POINTER(D), D,
                                             D, C.c_double, C.c_long, C.c_int, C.POINTER(C.c_uint8),
                                             C.POINTER(C.c_uint32)]

    def batch_guarded(self, net, rho, temp, Y, dt, hydro_dt, eos_kind=0, mu=10.0, Z=None, nthreads=0,
                      max_iter=10000):
        """The reference's compiled substep solver in its OpenMP particle loop, with the oracle's
        replay detection and Newton cap; in place on numpy arrays (Y: d x n, C-contiguous rows).
        Returns (status, newton)."""
        n = len(rho)
        d = DIMS[net]
        assert Y.shape == (d, n) and Y.flags["C_CONTIGUOUS"] and Y.dtype == np.float64
        rows = (D * d)(*[Y[s].ctypes.data_as(D) for s in range(d)])
        status = np.zeros(n, np.uint8)
        newt = np.zeros(n, np.uint32)
        Zc = np.ascontiguousarray(Z if Z is not None else np.zeros(1), dtype=np.float64)
        self.L.ref_batch_guarded(net, eos_kind, mu, _p(Zc), len(Zc) if Z is not None else 0, n, _p(rho), rows,
                                 _p(temp), _p(dt), hydro_dt, max_iter, nthreads,
                                 status.ctypes.data_as(C.POINTER(C.c_uint8)), newt.ctypes.data_as(C.POINTER(C.c_uint32)))
        return status, newt

    @property
    def has_helmholtz(self):
        return bool(self.L.ref_has_helmholtz())


def reference_available():
    return os.path.exists(REF_SO)
