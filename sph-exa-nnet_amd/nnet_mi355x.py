"""ctypes binding of libnnet_mi355x.so (the C-ABI in include/nnet_mi355x.h).

Used by the tests, the benchmark and the smoke check.  Device buffers are torch tensors on
``cuda`` (HIP) devices; only their data pointers cross the C-ABI.  There is no CPU fallback:
if the library is missing or cannot initialise a device, the calls raise.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# NNET_MI355X_LIB selects another build of the library (e.g. to compare two builds on one GPU)
LIB_PATH = os.environ.get("NNET_MI355X_LIB") or os.path.join(HERE, "lib", "libnnet_mi355x.so")

RATES = {14: 14, 86: 86, 87: 87}
EOS_IDEAL, EOS_HELMHOLTZ = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1
DRHODT_OFF, DRHODT_CPU, DRHODT_INTEGRATE = 0, 1, 2
STATUS = {0: "OK", 1: "SKIPPED", 2: "STALLED", 3: "ZERO_DT", 4: "MAX_ITER"}
N_SPECIES = {14: 14, 86: 86, 87: 87}
MAXSP = 88


class NnetError(RuntimeError):
    pass


class Batch(C.Structure):
    _fields_ = [
        ("network", C.c_void_p),
        ("eos", C.c_int),
        ("mu", C.c_double),
        ("Z", C.POINTER(C.c_double)),
        ("z_dim", C.c_int),
        ("precision", C.c_int),
        ("n", C.c_size_t),
        ("rho", C.c_void_p),
        ("rho_m1", C.c_void_p),
        ("Y", C.POINTER(C.c_void_p)),
        ("temp", C.c_void_p),
        ("dt", C.c_void_p),
        ("status", C.c_void_p),
        ("substeps", C.c_void_p),
        ("newton", C.c_void_p),
        ("fp64_solves", C.c_void_p),
        ("cost_hint", C.c_void_p),
        ("fp64_hint", C.c_void_p),
    ]


class Tuning(C.Structure):
    """nnet_tuning: ppw[net14, net86, net87][eos][precision], fused."""
    _fields_ = [("ppw", ((C.c_int * 2) * 2) * 3), ("fused", C.c_int)]

    def as_dict(self):
        nets = (14, 86, 87)
        return {"ppw": {(nets[k], e, p): self.ppw[k][e][p] for k in range(3) for e in range(2) for p in range(2)},
                "fused": bool(self.fused)}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NnetError(f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.nnet_ctx_create.argtypes = [C.POINTER(vp), C.c_int, C.c_char_p]
        L.nnet_ctx_destroy.argtypes = [vp]
        L.nnet_last_error.argtypes = [vp]
        L.nnet_last_error.restype = C.c_char_p
        L.nnet_ctx_load_helmholtz.argtypes = [vp, C.c_char_p]
        L.nnet_ctx_set_max_iterations.argtypes = [vp, C.c_long]
        L.nnet_network_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, ip, ip, ip, C.POINTER(vp)]
        L.nnet_network_builtin.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.nnet_network_destroy.argtypes = [vp]
        L.nnet_network_reactions.argtypes = [vp, ip, ip, ip, ip, ip, ip]
        L.nnet_compute.argtypes = [vp, C.POINTER(Batch), C.c_int, C.c_double, C.c_double, C.c_int, vp]
        L.nnet_compute_sync.argtypes = [vp, C.POINTER(Batch), C.c_int, C.c_double, C.c_double, C.c_int]
        L.nnet_compute_noalloc.argtypes = [vp, C.POINTER(Batch), C.c_int, C.c_double, C.c_double, C.c_int, vp]
        eos_args = [vp, C.c_size_t, C.c_int, dp, vp, vp, C.POINTER(vp), vp, vp, vp, vp, vp]
        L.nnet_helm_eos.argtypes = eos_args + [vp]
        L.nnet_helm_eos_sync.argtypes = eos_args
        L.nnet_helm_table_read.argtypes = [C.c_char_p, dp]
        L.nnet_solve_slots.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip]
        L.nnet_ctx_set_particles_per_workgroup.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.nnet_ctx_get_particles_per_workgroup.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip]
        L.nnet_ctx_set_fused_launch.argtypes = [vp, C.c_int]
        L.nnet_ctx_get_fused_launch.argtypes = [vp, ip]
        L.nnet_tuning_read.argtypes = [C.c_char_p, C.POINTER(Tuning)]
        L.nnet_ctx_load_tuning.argtypes = [vp, C.c_char_p]
        L.nnet_ctx_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        _lib = L
    return _lib


EXPORTS = [
    "nnet_ctx_create", "nnet_ctx_destroy", "nnet_last_error", "nnet_ctx_load_helmholtz",
    "nnet_network_create", "nnet_network_builtin", "nnet_network_destroy", "nnet_network_reactions",
    "nnet_compute", "nnet_compute_sync", "nnet_ctx_set_max_iterations", "nnet_helm_eos", "nnet_helm_eos_sync",
    "nnet_helm_table_read", "nnet_solve_slots", "nnet_compute_noalloc",
    "nnet_ctx_set_particles_per_workgroup", "nnet_ctx_get_particles_per_workgroup",
    "nnet_ctx_set_fused_launch", "nnet_ctx_get_fused_launch", "nnet_tuning_read", "nnet_ctx_load_tuning",
    "nnet_ctx_get_tuning",
]


def tuning_read(path):
    """Parse an autotuner table on the host (no device): the launch parameters as a dict."""
    t = Tuning()
    rc = lib().nnet_tuning_read(path.encode(), C.byref(t))
    if rc != 0:
        raise NnetError(f"nnet_tuning_read({path}): {rc}")
    return t.as_dict()


def helm_table_read(path):
    """The 21 Helmholtz table arrays (21 x 541 x 201) of a binary or reference-text table file,
    parsed by the library on the host (no device needed)."""
    import numpy as np
    out = np.empty((21, 541, 201))
    rc = lib().nnet_helm_table_read(path.encode(), out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise NnetError(f"nnet_helm_table_read({path}): {rc}")
    return out


class Context:
    """One HIP device context (tables uploaded once)."""

    def __init__(self, device=0, helm_table=None):
        L = lib()
        h = C.c_void_p()
        rc = L.nnet_ctx_create(C.byref(h), device, helm_table.encode() if helm_table else None)
        if rc != 0:
            msg = L.nnet_last_error(h).decode() if h.value else "device init failed"
            if h.value:
                L.nnet_ctx_destroy(h)
            raise NnetError(f"nnet_ctx_create: {rc}: {msg}")
        self.h = h
        self.networks = {}

    def check(self, rc, what):
        if rc != 0:
            raise NnetError(f"{what}: {rc}: {lib().nnet_last_error(self.h).decode()}")

    def network(self, net):
        if net not in self.networks:
            h = C.c_void_p()
            self.check(lib().nnet_network_builtin(self.h, net, C.byref(h)), "nnet_network_builtin")
            self.networks[net] = h
        return self.networks[net]

    def reactions(self, net):
        import numpy as np
        h = self.network(net)
        nr, ne = C.c_int(), C.c_int()
        self.check(lib().nnet_network_reactions(h, C.byref(nr), C.byref(ne), None, None, None, None), "reactions")
        rb = np.zeros(nr.value + 1, np.int32)
        pb = np.zeros(nr.value, np.int32)
        ids = np.zeros(ne.value, np.int32)
        nums = np.zeros(ne.value, np.int32)
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        self.check(lib().nnet_network_reactions(h, None, None, P(rb), P(pb), P(ids), P(nums)), "reactions")
        return rb, pb, ids, nums

    def resident_slots(self, net, eos=EOS_HELMHOLTZ, precision=PRECISION_FP64):
        """Particles one solve launch keeps in flight (resident workgroups of the persistent kernel)."""
        out = C.c_int()
        self.check(lib().nnet_solve_slots(self.h, net, eos, precision, C.byref(out)), "nnet_solve_slots")
        return out.value

    def set_max_iterations(self, m):
        self.check(lib().nnet_ctx_set_max_iterations(self.h, m), "set_max_iterations")

    def set_particles_per_workgroup(self, net, eos, precision, ppw):
        """Particles per workgroup of the solve launches (net14: 1, 2 or 4; net86/net87: 1)."""
        self.check(lib().nnet_ctx_set_particles_per_workgroup(self.h, net, eos, precision, ppw),
                   "set_particles_per_workgroup")

    def load_tuning(self, path):
        self.check(lib().nnet_ctx_load_tuning(self.h, path.encode()), "load_tuning")

    def tuning(self):
        t = Tuning()
        self.check(lib().nnet_ctx_get_tuning(self.h, C.byref(t)), "tuning")
        return t.as_dict()

    def set_fused_launch(self, fused):
        """Batches of several networks in one call: one fused launch (True) or one launch per batch."""
        self.check(lib().nnet_ctx_set_fused_launch(self.h, int(bool(fused))), "set_fused_launch")

    def fused_launch(self):
        out = C.c_int()
        self.check(lib().nnet_ctx_get_fused_launch(self.h, C.byref(out)), "fused_launch")
        return bool(out.value)

    def particles_per_workgroup(self, net, eos, precision):
        out = C.c_int()
        self.check(lib().nnet_ctx_get_particles_per_workgroup(self.h, net, eos, precision, C.byref(out)),
                   "particles_per_workgroup")
        return out.value

    def make_batch(self, net, state, eos=EOS_IDEAL, mu=10.0, Z=None, precision=PRECISION_FP64):
        """state: dict of torch device tensors rho, rho_m1, Y (d x n), temp, dt, status, substeps, newton
        and (optional) fp64_solves, cost_hint and fp64_hint."""
        b = Batch()
        b.network = self.network(net).value
        b.eos = eos
        b.mu = mu
        if Z is not None:
            zarr = (C.c_double * len(Z))(*[float(z) for z in Z])
            b.Z = C.cast(zarr, C.POINTER(C.c_double))
            b.z_dim = len(Z)
            b._zkeep = zarr
        else:
            b.Z = None
            b.z_dim = 0
        b.precision = precision
        Y = state["Y"]
        n = Y.shape[1]
        b.n = n
        ptr = lambda k: state[k].data_ptr() if state.get(k) is not None else None
        b.rho, b.rho_m1, b.temp, b.dt = ptr("rho"), ptr("rho_m1"), ptr("temp"), ptr("dt")
        b.status, b.substeps, b.newton = ptr("status"), ptr("substeps"), ptr("newton")
        b.fp64_solves = ptr("fp64_solves")
        b.cost_hint = ptr("cost_hint")
        b.fp64_hint = ptr("fp64_hint")
        yp = (C.c_void_p * Y.shape[0])(*[Y[s].data_ptr() for s in range(Y.shape[0])])
        b.Y = C.cast(yp, C.POINTER(C.c_void_p))
        b._ykeep = yp
        return b

    def compute(self, batches, hydro_dt, previous_dt=None, use_drhodt=DRHODT_OFF, stream=None, sync=True):
        """use_drhodt: DRHODT_OFF / DRHODT_CPU / DRHODT_INTEGRATE (True means DRHODT_CPU)."""
        arr = (Batch * len(batches))(*batches)
        prev = hydro_dt if previous_dt is None else previous_dt
        if sync and stream is None:
            rc = lib().nnet_compute_sync(self.h, arr, len(batches), hydro_dt, prev, int(use_drhodt))
        else:
            rc = lib().nnet_compute(self.h, arr, len(batches), hydro_dt, prev, int(use_drhodt), stream)
        self.check(rc, "nnet_compute")

    def helm_eos(self, Z, temp, rho, Y, out, stream=None, sync=True):
        """Post-step EOS refresh (computeHelmEOS): Y is a (dim x n) device tensor (or a list of dim
        per-species device tensors), temp and rho device tensors of n doubles, `out` a dict whose
        keys among u, cv, p, c, dpdT name the device tensors to fill (others are skipped)."""
        rows = [Y[s] for s in range(len(Y))]
        n = temp.shape[0]
        zarr = (C.c_double * len(Z))(*[float(z) for z in Z])
        yp = (C.c_void_p * len(rows))(*[r.data_ptr() for r in rows])
        o = [out[k].data_ptr() if out.get(k) is not None else None for k in ("u", "cv", "p", "c", "dpdT")]
        args = [self.h, n, len(rows), zarr, temp.data_ptr(), rho.data_ptr(), yp] + o
        if sync and stream is None:
            rc = lib().nnet_helm_eos_sync(*args)
        else:
            rc = lib().nnet_helm_eos(*args, stream)
        self.check(rc, "nnet_helm_eos")

    def close(self):
        if self.h is not None and self.h.value:
            for h in self.networks.values():
                lib().nnet_network_destroy(h)
            self.networks = {}
            lib().nnet_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_state(torch, net, rho, temp, Y, dt, device="cuda:0"):
    """Upload numpy particle arrays (Y as d x n) into device tensors."""
    import numpy as np
    n = len(rho)
    st = {
        "rho": torch.tensor(np.ascontiguousarray(rho), dtype=torch.float64, device=device),
        "rho_m1": torch.tensor(np.ascontiguousarray(rho), dtype=torch.float64, device=device),
        "temp": torch.tensor(np.ascontiguousarray(temp), dtype=torch.float64, device=device),
        "Y": torch.tensor(np.ascontiguousarray(Y), dtype=torch.float64, device=device),
        "dt": torch.tensor(np.ascontiguousarray(dt), dtype=torch.float64, device=device),
        "status": torch.zeros(n, dtype=torch.uint8, device=device),
        "substeps": torch.zeros(n, dtype=torch.int32, device=device),
        "newton": torch.zeros(n, dtype=torch.int32, device=device),
        "fp64_solves": torch.zeros(n, dtype=torch.int32, device=device),
    }
    return st
