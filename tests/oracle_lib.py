"""ctypes binding of the oracle library (oracle/_build/liboracle.so).

Test infrastructure only: the oracle is the CPU restatement of the reference step that the GPU path is
checked against.  The product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


def _p(a):
    """numpy float64/int32 array -> ctypes pointer (None passes through)."""
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(DP)
    if a.dtype == np.int32:
        return a.ctypes.data_as(IP)
    raise TypeError(a.dtype)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(LIB_PATH)
        L.or_last_error.restype = C.c_char_p
        L.or_set_threads.argtypes = [C.c_int]
        L.or_thomas.argtypes = [DP, DP, DP, DP, DP, C.c_int]
        L.or_make_uniform.argtypes = [C.c_double, C.c_double, DP, C.c_int]
        L.or_make_regions.argtypes = [C.c_double, DP, DP, C.c_int, C.c_double, DP, C.c_int]
        L.or_sim_record_chi.argtypes = [C.c_void_p, C.c_int]
        L.or_laser_gaussian.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 8 + [DP]
        L.or_laser_step.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                    DP, DP, DP, DP]
        L.or_laser_sums.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, DP, DP]
        L.or_solver_create.restype = C.c_void_p
        L.or_solver_create.argtypes = [DP, C.c_int]
        L.or_solver_destroy.argtypes = [C.c_void_p]
        L.or_solver_solve.argtypes = [C.c_void_p, C.c_int, DP, DP, DP, DP, C.c_double]
        L.or_solver_gradient.argtypes = [C.c_void_p, DP, DP, C.c_int]
        L.or_solver_over_r.argtypes = [C.c_void_p, DP, DP]
        L.or_solver_coeffs.argtypes = [C.c_void_p, C.c_int, DP, DP, DP]
        L.or_solver_flux_coef.restype = C.c_double
        L.or_solver_flux_coef.argtypes = [C.c_void_p]
        L.or_grid_arrays.argtypes = [C.c_void_p, DP, DP, DP]
        L.or_grid_locate.argtypes = [C.c_void_p, C.c_int, DP, IP, DP]
        L.or_uniform_deposit.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, DP]
        L.or_push_batch.argtypes = [C.c_int, C.c_int, DP, DP, DP, DP, DP]
        L.or_make_ab.argtypes = [C.c_int, DP]
        L.or_make_ab_variable.argtypes = [C.c_int, DP, C.c_double, DP]
        L.or_mt19937_64.restype = C.c_ulonglong
        L.or_mt19937_64.argtypes = [C.c_ulonglong, C.c_int]
        L.or_beam_deposit.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_int] + [DP] * 7 + [DP, C.c_int]
        L.or_beam_advance.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_int, C.c_int] + [DP] * 7 + [DP, C.c_int, C.c_double, C.c_double, C.c_int]
        L.or_sim_create.restype = C.c_void_p
        L.or_sim_create.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.or_sim_destroy.argtypes = [C.c_void_p]
        L.or_sim_info.argtypes = [C.c_void_p, IP]
        L.or_sim_step.argtypes = [C.c_void_p]
        L.or_sim_step_fields.argtypes = [C.c_void_p]
        L.or_sim_sweep_partial.argtypes = [C.c_void_p, C.c_int]
        L.or_sim_timing.argtypes = [C.c_void_p, DP]
        L.or_sim_lost.restype = C.c_long
        L.or_sim_lost.argtypes = [C.c_void_p]
        L.or_sim_time.restype = C.c_double
        L.or_sim_time.argtypes = [C.c_void_p]
        L.or_sim_get.restype = C.c_long
        L.or_sim_get.argtypes = [C.c_void_p, C.c_char_p, DP, C.c_long]
        L.or_sim_beam_sums.argtypes = [C.c_void_p, C.c_int, DP, DP]
        L.or_sim_beam_diag.argtypes = [C.c_void_p, C.c_int, C.c_int, DP]
        L.or_sim_write_fields.argtypes = [C.c_void_p, C.c_char_p]
        L.or_sim_spread_grad.argtypes = [C.c_void_p, C.c_int, DP, DP, C.c_int]
        L.or_sim_set_delta.argtypes = [C.c_void_p, C.c_int, DP, C.c_long]
        L.or_sim_adjoint_get.restype = C.c_long
        L.or_sim_adjoint_get.argtypes = [C.c_void_p, C.c_int, DP, C.c_long]
        L.or_sim_write_axis.argtypes = [C.c_void_p, C.c_char_p]
        _lib = L
    return _lib


def err():
    return lib().or_last_error().decode()


def check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + err())
    return rc


def set_threads(n):
    return lib().or_set_threads(int(n))


def thomas(a, b, c, d):
    x = np.zeros_like(d)
    lib().or_thomas(_p(a), _p(b), _p(c), _p(d), _p(x), len(d))
    return x


def make_uniform(R, dr):
    out = np.zeros(2_000_000)
    n = lib().or_make_uniform(R, dr, _p(out), len(out))
    if n <= 0:
        raise RuntimeError(err())
    return out[:n].copy()


def make_regions(R, regions, max_ratio):
    ends = np.array([a for a, _ in regions], dtype=np.float64)
    hs = np.array([b for _, b in regions], dtype=np.float64)
    out = np.zeros(2_000_000)
    n = lib().or_make_regions(R, _p(ends), _p(hs), len(regions), max_ratio, _p(out), len(out))
    if n <= 0:
        raise RuntimeError(err())
    return out[:n].copy()


class Solver:
    """Grid + FieldSolver of the oracle on a node list."""

    def __init__(self, nodes):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.M = len(self.nodes)
        self.h = lib().or_solver_create(_p(self.nodes), self.M)
        if not self.h:
            raise RuntimeError(err())

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_solver_destroy(self.h)
            self.h = None

    def solve(self, op, rhs, chi=None, wall=None, scale=1.0):
        f = np.zeros(self.M)
        check(lib().or_solver_solve(self.h, op, _p(rhs), _p(f), _p(chi), _p(wall), scale))
        return f

    def gradient(self, f, odd=False):
        df = np.zeros(self.M)
        check(lib().or_solver_gradient(self.h, _p(f), _p(df), int(odd)))
        return df

    def over_r(self, f):
        out = np.zeros(self.M)
        check(lib().or_solver_over_r(self.h, _p(f), _p(out)))
        return out

    def coeffs(self, op):
        a, b, c = np.zeros(self.M), np.zeros(self.M), np.zeros(self.M)
        check(lib().or_solver_coeffs(self.h, op, _p(a), _p(b), _p(c)))
        return a, b, c

    def flux_coef(self):
        return lib().or_solver_flux_coef(self.h)

    def grid_arrays(self):
        V, iV, ih = np.zeros(self.M), np.zeros(self.M), np.zeros(self.M - 1)
        check(lib().or_grid_arrays(self.h, _p(V), _p(iV), _p(ih)))
        return V, iV, ih

    def locate(self, xs):
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        j = np.zeros(len(xs), dtype=np.int32)
        t = np.zeros(len(xs))
        check(lib().or_grid_locate(self.h, len(xs), _p(xs), _p(j), _p(t)))
        return j, t

    def laser_gaussian(self, K, xi_min, dxi, a0, w0, L0, xi0, focus, k0):
        """The Gaussian pulse (l = 0) on this grid, (K, M, 2); xi, xi0 internal."""
        a = np.zeros((K, self.M, 2))
        check(lib().or_laser_gaussian(self.h, K, xi_min, dxi, a0, w0, L0, xi0, focus, k0, _p(a)))
        return a

    def laser_step(self, a, dxi, dt, k0, scheme, circular=False, chi=None):
        """One envelope step over the whole box: (a^(n+1), pond (K, M, 3)).  scheme 0 trapezoid, 1 bdf2, 2 euler."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        K = a.shape[0]
        assert a.shape == (K, self.M, 2)
        if chi is not None:
            chi = np.ascontiguousarray(chi, dtype=np.float64)
            assert chi.shape == (K, self.M)
        an = np.zeros_like(a)
        pond = np.zeros((K, self.M, 3))
        check(lib().or_laser_step(self.h, K, dxi, dt, k0, scheme, int(circular), _p(chi), _p(a), _p(an), _p(pond)))
        return an, pond

    def laser_sums(self, a, xi_min, dxi):
        a = np.ascontiguousarray(a, dtype=np.float64)
        out = np.zeros(5)
        check(lib().or_laser_sums(self.h, a.shape[0], xi_min, dxi, _p(a), _p(out)))
        return out

    def uniform_deposit(self, ppc, ntheta=1, mode1=False):
        out = np.zeros(self.M)
        check(lib().or_uniform_deposit(self.h, ppc, ntheta, int(mode1), _p(out)))
        return out


PUSHERS = {"vay": 0, "boris": 1, "hc": 2, "imp": 3, "imp_rr0": 4}


def push(mode, u, E, B, eps):
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
    n = len(u)
    E = np.ascontiguousarray(E, dtype=np.float64).reshape(-1, 3)
    B = np.ascontiguousarray(B, dtype=np.float64).reshape(-1, 3)
    eps = np.ascontiguousarray(np.broadcast_to(eps, (n,)), dtype=np.float64)
    out = np.zeros_like(u)
    lib().or_push_batch(PUSHERS[mode] if isinstance(mode, str) else mode, n, _p(u), _p(E), _p(B), _p(eps), _p(out))
    return out


def make_ab(order):
    c = np.zeros(5)
    check(lib().or_make_ab(order, _p(c)))
    return c[:order]


def make_ab_variable(order, t, h):
    c = np.zeros(5)
    t = np.ascontiguousarray(t, dtype=np.float64)
    check(lib().or_make_ab_variable(order, _p(t), h, _p(c)))
    return c[:order]


def mt19937_64(seed, n):
    return lib().or_mt19937_64(seed, n)


BEAM_ARRAYS = ("x", "y", "px", "py", "pz", "xi", "w")


def beam_deposit(solver, xi_min, dxi, nxi, ngp, m1, q, parts, nc, into=None):
    """Beam::deposit of one beam, added to `into` (a (nxi, M, nc) array) or to zeros."""
    src = np.zeros(nxi * solver.M * nc) if into is None else np.ascontiguousarray(into, dtype=np.float64).reshape(-1).copy()
    p = [np.ascontiguousarray(parts[k], dtype=np.float64) for k in BEAM_ARRAYS]
    check(lib().or_beam_deposit(solver.h, xi_min, dxi, nxi, int(ngp), int(m1), q, len(p[0]), *[_p(a) for a in p],
                                _p(src), nc))
    return src.reshape(nxi, solver.M, nc)


def beam_advance(solver, xi_min, dxi, nxi, ngp, m1, qm, mode, parts, fld, dt_kick, dt, move):
    p = {k: np.array(parts[k], dtype=np.float64) for k in BEAM_ARRAYS}
    fld = np.ascontiguousarray(fld, dtype=np.float64)
    nc = fld.shape[-1]
    check(lib().or_beam_advance(solver.h, xi_min, dxi, nxi, int(ngp), int(m1), qm, mode, len(p["x"]),
                                *[_p(p[k]) for k in BEAM_ARRAYS], _p(fld), nc, dt_kick, dt, int(move)))
    return p


class Sim:
    """The oracle's whole step on a deck (text) with key=value overrides."""

    def __init__(self, deck, overrides=None, parallel=False):
        if os.path.exists(deck):
            deck = open(deck).read()
        ov = "\n".join(f"{k}={v}" for k, v in (overrides or {}).items())
        self.h = lib().or_sim_create(deck.encode(), ov.encode(), int(parallel))
        if not self.h:
            raise RuntimeError(err())
        info = (C.c_int * 10)()
        lib().or_sim_info(self.h, info)
        self.M, self.K, self.m1, self.NCF, self.NCB, self.NCD, self.nspecies, self.nbeams, self.nsteps = info[:9]

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sim_destroy(self.h)
            self.h = None

    def step(self):
        check(lib().or_sim_step(self.h))

    def step_fields(self):
        check(lib().or_sim_step_fields(self.h))

    def spread_grad(self, b, nseg):
        """One step; J = gamma_rms / gamma_mean of beam b after it and dJ/dv of its weight profile (the
        oracle's reverse sweep)."""
        J = np.zeros(1)
        g = np.zeros(nseg)
        check(lib().or_sim_spread_grad(self.h, b, _p(J), _p(g), nseg))
        return J[0], g

    def set_delta(self, which, d):
        """Test hook: d is added to the beam sources (which 0) or the fields (1) of every later step."""
        d = np.ascontiguousarray(d, dtype=np.float64).ravel()
        check(lib().or_sim_set_delta(self.h, which, _p(d) if d.size else None, d.size))

    def adjoint(self, which):
        """The adjoints of the last spread_grad: 0 fields (K, M, NCF), 1 beam sources (K, M, NCB), 2 weights."""
        n = lib().or_sim_adjoint_get(self.h, which, None, 0)
        out = np.zeros(n)
        lib().or_sim_adjoint_get(self.h, which, _p(out), n)
        return out

    def sweep_partial(self, nk):
        check(lib().or_sim_sweep_partial(self.h, nk))

    def timing(self):
        """(sweep s, beam push s, particle-slices of the last sweep)"""
        out = np.zeros(3)
        lib().or_sim_timing(self.h, _p(out))
        return out

    def lost(self):
        return lib().or_sim_lost(self.h)

    @property
    def t(self):
        return lib().or_sim_time(self.h)

    def get(self, what):
        n = lib().or_sim_get(self.h, what.encode(), None, 0)
        if n < 0:
            raise KeyError(what)
        out = np.zeros(n)
        lib().or_sim_get(self.h, what.encode(), _p(out), n)
        return out

    @property
    def fld(self):
        return self.get("fld").reshape(self.K, self.M, self.NCF)

    @property
    def diag(self):
        return self.get("diag").reshape(self.K, self.M, self.NCD)

    @property
    def bsrc(self):
        return self.get("bsrc").reshape(self.K, self.M, self.NCB)

    @property
    def nodes(self):
        return self.get("nodes")

    def beam(self, b):
        return {k: self.get(f"beam{b}.{k}") for k in BEAM_ARRAYS}

    def beam_sums(self, b):
        out, refs = np.zeros(17), np.zeros(2)
        check(lib().or_sim_beam_sums(self.h, b, _p(out), _p(refs)))
        return out, refs

    def beam_diag(self, b, order="index"):
        """Beam log values from the index-order sums (the reference's) or, order="tree", from the sums in the
        product's on-device reduction order."""
        out = np.zeros(15)
        check(lib().or_sim_beam_diag(self.h, b, int(order == "tree"), _p(out)))
        keys = ["t", "npart", "charge", "gamma_mean", "gamma_rms", "xi_mean", "xi_rms", "r_rms", "emit_nx", "x_mean",
                "y_mean", "emit_ny", "sigma_x", "sigma_y", "alive"]
        return dict(zip(keys, out))

    def write_fields(self, path):
        check(lib().or_sim_write_fields(self.h, path.encode()))

    def write_axis(self, path):
        check(lib().or_sim_write_axis(self.h, path.encode()))


def read_fields(path):
    """A fields_*.bin file of format version 3 (src/Simulation.cpp:873-993) as a dict: "r" (M nodes), "xi" (K user
    xi values, increasing: the head is the last row) and one (K, M) array per named component."""
    with open(path, "rb") as f:
        ver, M, K, _m1 = (int(v) for v in np.frombuffer(f.read(16), np.int32))
        if ver != 3:
            raise ValueError(f"{path}: format version {ver}")
        f.read(8)   # time
        out = {"r": np.frombuffer(f.read(8 * M), np.float64).copy(),
               "xi": np.frombuffer(f.read(8 * K), np.float64).copy()}
        nc = int(np.frombuffer(f.read(4), np.int32)[0])
        for _ in range(nc):
            name = f.read(16).split(b"\0")[0].decode()
            out[name] = np.frombuffer(f.read(8 * K * M), np.float64).reshape(K, M).copy()
    return out
