"""ctypes binding of the CPU FP64 oracle (oracle/ibpm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the ibpm_amd product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libibpm_oracle.so")
SRC = os.path.join(HERE, "ibpm_oracle.c")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile the oracle with gcc (FP64, no fast-math, no FMA contraction)."""
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC):
        subprocess.check_call(["gcc", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fPIC",
                               "-shared", "-o", LIB_PATH, SRC, "-lm"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.og_sim_create.restype = ctypes.c_void_p
        L.og_sim_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_double,
                                    ctypes.c_int]
        for name in ["og_sim_destroy", "og_sim_init", "og_sim_init_timing", "og_sim_advance",
                     "og_sim_num_points", "og_sim_uses_cg", "og_sim_set_ubf"]:
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.og_sim_error.restype = ctypes.c_char_p
        L.og_sim_time.restype = ctypes.c_double
        L.og_sim_time.argtypes = [ctypes.c_void_p]
        L.og_sim_beta.restype = ctypes.c_double
        L.og_sim_beta.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.og_reg_build.restype = ctypes.c_int
        L.og_ip_scalar.restype = ctypes.c_double
        L.og_ip_flux.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def scalar_size(nx, ny, ngrid):
    return ngrid * (nx - 1) * (ny - 1)


def flux_size(nx, ny, ngrid):
    return ngrid * (2 * nx * ny + nx + ny)


def dst2d(x):
    x = _c(x)
    out = np.empty_like(x)
    lib().og_dst2d(x.shape[0], x.shape[1], _p(x), _p(out))
    return out


def dst2d_direct(x):
    x = _c(x)
    out = np.empty_like(x)
    lib().og_dst2d_direct(x.shape[0], x.shape[1], _p(x), _p(out))
    return out


def coarsify(nx, ny, ngrid, s):
    s = _c(s).copy()
    lib().og_coarsify(nx, ny, ngrid, _p(s))
    return s


def getbc(nx, ny, ngrid, s, lev):
    s = _c(s)
    bc = np.zeros(2 * (nx + ny))
    lib().og_getbc(nx, ny, ngrid, _p(s), lev, _p(bc))
    return bc


def elliptic2d_solve(kind, nx, ny, dx, alpha, f, bc=None):
    f = _c(f)
    u = np.empty_like(f)
    bcp = _p(_c(bc)) if bc is not None else None
    lib().og_elliptic2d_solve(kind, nx, ny, ctypes.c_double(dx), ctypes.c_double(alpha), _p(f), bcp, _p(u))
    return u


def elliptic_solve(kind, nx, ny, ngrid, dx, alpha, f):
    f = _c(f)
    u = np.empty_like(f)
    lib().og_elliptic_solve(kind, nx, ny, ngrid, ctypes.c_double(dx), ctypes.c_double(alpha), _p(f), _p(u))
    return u


def _op_q_to_s(name, nx, ny, ngrid, dx, q):
    q = _c(q)
    out = np.zeros(scalar_size(nx, ny, ngrid))
    getattr(lib(), name)(nx, ny, ngrid, ctypes.c_double(dx), _p(q), _p(out))
    return out


def _op_s_to_q(name, nx, ny, ngrid, dx, s, q=None):
    s = _c(s)
    out = np.zeros(flux_size(nx, ny, ngrid)) if q is None else _c(q).copy()
    getattr(lib(), name)(nx, ny, ngrid, ctypes.c_double(dx), _p(s), _p(out))
    return out


def curl_flux(nx, ny, ngrid, dx, q):
    return _op_q_to_s("og_curl_flux", nx, ny, ngrid, dx, q)


def curl_scalar(nx, ny, ngrid, w):
    w = _c(w)
    out = np.zeros(flux_size(nx, ny, ngrid))
    lib().og_curl_scalar(nx, ny, ngrid, _p(w), _p(out))
    return out


def laplacian(nx, ny, ngrid, dx, w):
    w = _c(w)
    out = np.zeros(scalar_size(nx, ny, ngrid))
    lib().og_laplacian(nx, ny, ngrid, ctypes.c_double(dx), _p(w), _p(out))
    return out


def laplacian2d_bc(nx, ny, dx, f, bc):
    f = _c(f)
    bc = _c(bc)
    g = np.zeros_like(f)
    lib().og_laplacian2d_bc(nx, ny, ctypes.c_double(dx), _p(f), _p(bc), _p(g))
    return g


def flux_to_xvel(nx, ny, ngrid, dx, q):
    return _op_q_to_s("og_flux_to_xvel", nx, ny, ngrid, dx, q)


def flux_to_yvel(nx, ny, ngrid, dx, q):
    return _op_q_to_s("og_flux_to_yvel", nx, ny, ngrid, dx, q)


def xvel_to_flux(nx, ny, ngrid, dx, u, q=None):
    return _op_s_to_q("og_xvel_to_flux", nx, ny, ngrid, dx, u, q)


def yvel_to_flux(nx, ny, ngrid, dx, v, q=None):
    return _op_s_to_q("og_yvel_to_flux", nx, ny, ngrid, dx, v, q)


def nonlinear(nx, ny, ngrid, dx, q, w):
    q = _c(q)
    w = _c(w)
    out = np.zeros(scalar_size(nx, ny, ngrid))
    lib().og_nonlinear(nx, ny, ngrid, ctypes.c_double(dx), _p(q), _p(w), _p(out))
    return out


def cross_product(nx, ny, ngrid, dx, q, w):
    q = _c(q)
    w = _c(w)
    out = np.zeros(flux_size(nx, ny, ngrid))
    lib().og_cross_product(nx, ny, ngrid, ctypes.c_double(dx), _p(q), _p(w), _p(out))
    return out


def linear_N(nx, ny, ngrid, dx, q0, w0, q, w):
    """LinearizedIBSolver::N: Curl(q0 x omega + q x omega0)."""
    q0, w0, q, w = _c(q0), _c(w0), _c(q), _c(w)
    out = np.zeros(scalar_size(nx, ny, ngrid))
    lib().og_linear_N(nx, ny, ngrid, ctypes.c_double(dx), _p(q0), _p(w0), _p(q), _p(w), _p(out))
    return out


def adjoint_N(nx, ny, ngrid, dx, q0, w0, q):
    """AdjointIBSolver::N: Laplacian(q0 x q) - Curl(q x omega0)."""
    q0, w0, q = _c(q0), _c(w0), _c(q)
    out = np.zeros(scalar_size(nx, ny, ngrid))
    lib().og_adjoint_N(nx, ny, ngrid, ctypes.c_double(dx), _p(q0), _p(w0), _p(q), _p(out))
    return out


def cross_flux_flux(nx, ny, ngrid, dx, q1, q2):
    q1 = _c(q1)
    q2 = _c(q2)
    out = np.zeros(scalar_size(nx, ny, ngrid))
    lib().og_cross_flux_flux(nx, ny, ngrid, ctypes.c_double(dx), _p(q1), _p(q2), _p(out))
    return out


def reg_build(nx, ny, dx, xoff, yoff, px, py):
    px = _c(px)
    py = _c(py)
    n = len(px)
    cap = 36 * n + 1
    bi = np.zeros(cap, dtype=np.int32)
    fi = np.zeros(cap, dtype=np.int32)
    w = np.zeros(cap)
    na = lib().og_reg_build(nx, ny, ctypes.c_double(dx), ctypes.c_double(xoff), ctypes.c_double(yoff), n,
                            _p(px), _p(py), bi.ctypes.data_as(_I), fi.ctypes.data_as(_I), _p(w))
    return bi[:na].copy(), fi[:na].copy(), w[:na].copy()


def reg_toflux(nx, ny, ngrid, dx, assoc, f):
    bi, fi, w = assoc
    f = _c(f)
    q = np.zeros(flux_size(nx, ny, ngrid))
    lib().og_reg_toflux(nx, ny, ngrid, ctypes.c_double(dx), len(w), bi.ctypes.data_as(_I),
                        fi.ctypes.data_as(_I), _p(w), _p(f), _p(q))
    return q


def reg_toboundary(npts, dx, assoc, q):
    bi, fi, w = assoc
    q = _c(q)
    f = np.zeros(2 * npts)
    lib().og_reg_toboundary(npts, ctypes.c_double(dx), len(w), bi.ctypes.data_as(_I),
                            fi.ctypes.data_as(_I), _p(w), _p(q), _p(f))
    return f


SCHEMES = {"euler": 0, "ab2": 1, "rk3": 2, "rk3b": 3}
MODELS = {"nonlinear": 0, "linear": 1, "adjoint": 2, "linearperiodic": 3, "sfd": 4}


class Sim:
    """Whole-solver oracle: one flow advanced by the restated IBSolver."""

    def __init__(self, nx, ny, ngrid, length, xoffset, yoffset, Re, dt, geom,
                 scheme="rk3", alpha_deg=0.0, tol=1e-7, force_cg=False, ubf=False,
                 model="nonlinear", period=1, chi=0.02, delta=15.0):
        L = lib()
        self.nx, self.ny, self.ngrid = nx, ny, ngrid
        self.h = L.og_sim_create(nx, ny, ngrid, length, xoffset, yoffset, Re, dt, SCHEMES[scheme],
                                 alpha_deg, geom.encode(), tol, int(force_cg))
        if not self.h:
            raise ValueError("oracle: " + L.og_sim_error().decode())
        self.nb = L.og_sim_num_points(self.h)
        self.ubf = 0
        if ubf:   # ibpm -ubf: 1 if the body's motion moved to the base flow
            self.ubf = L.og_sim_set_ubf(self.h)
            if self.ubf < 0:
                raise ValueError("oracle: " + L.og_sim_error().decode())
        self.nsub = 1 if scheme in ("euler", "ab2") else (4 if scheme == "rk3b" else 3)
        if model != "nonlinear":
            if L.og_sim_set_model(ctypes.c_void_p(self.h), MODELS[model], period, ctypes.c_double(chi),
                                  ctypes.c_double(delta)) != 0:
                raise ValueError("oracle: " + L.og_sim_error().decode())

    def set_filtered(self, w):
        """SFD: set the filtered vorticity (loadFilteredState); after init()."""
        w = _c(w)
        lib().og_sim_set_filtered(ctypes.c_void_p(self.h), _p(w))

    def filtered(self):
        out = np.zeros(scalar_size(self.nx, self.ny, self.ngrid))
        lib().og_sim_get_filtered(ctypes.c_void_p(self.h), _p(out))
        return out

    def set_base_state(self, omega, q, k=0):
        """Base state k (vorticity, flux) of the linear models; before init()."""
        w, qq = _c(omega), _c(q)
        lib().og_sim_set_base_state(ctypes.c_void_p(self.h), k, _p(w), _p(qq))

    def __del__(self):
        if getattr(self, "h", None):
            lib().og_sim_destroy(self.h)
            self.h = None

    @property
    def uses_cg(self):
        return bool(lib().og_sim_uses_cg(self.h))

    def init(self):
        lib().og_sim_init(self.h)

    def init_timing(self):
        """Set-up for timing only: identity Cholesky factors (no build)."""
        lib().og_sim_init_timing(self.h)

    def advance(self, n=1):
        for _ in range(n):
            lib().og_sim_advance(self.h)

    def forces(self):
        cd = ctypes.c_double()
        cl = ctypes.c_double()
        lib().og_sim_forces(ctypes.c_void_p(self.h), ctypes.byref(cd), ctypes.byref(cl))
        return cd.value, cl.value

    def omega(self):
        out = np.zeros(scalar_size(self.nx, self.ny, self.ngrid))
        lib().og_sim_get_omega(ctypes.c_void_p(self.h), _p(out))
        return out

    def set_omega(self, w):
        w = _c(w)
        lib().og_sim_set_omega(ctypes.c_void_p(self.h), _p(w))

    def q(self):
        out = np.zeros(flux_size(self.nx, self.ny, self.ngrid))
        lib().og_sim_get_q(ctypes.c_void_p(self.h), _p(out))
        return out

    def f(self):
        out = np.zeros(2 * self.nb)
        lib().og_sim_get_f(ctypes.c_void_p(self.h), _p(out))
        return out

    def points(self):
        out = np.zeros(2 * self.nb)
        lib().og_sim_get_points(ctypes.c_void_p(self.h), _p(out))
        return out

    def time(self):
        return lib().og_sim_time(self.h)

    def cg_iters(self):
        it = np.zeros(4, dtype=np.int32)
        n = lib().og_sim_last_cg_iters(ctypes.c_void_p(self.h), it.ctypes.data_as(_I))
        return it[:n].copy()

    def beta(self, i):
        return lib().og_sim_beta(self.h, i)

    def projection_solve(self, i, a, b, f0):
        a = _c(a)
        b = _c(b)
        f = _c(f0).copy()
        w = np.zeros(scalar_size(self.nx, self.ny, self.ngrid))
        lib().og_sim_projection_solve(ctypes.c_void_p(self.h), i, _p(a), _p(b), _p(w), _p(f))
        return w, f

    def M(self, i, f):
        f = _c(f)
        y = np.zeros(2 * self.nb)
        lib().og_sim_M(ctypes.c_void_p(self.h), i, _p(f), _p(y))
        return y

    def B(self, i, f):
        f = _c(f)
        w = np.zeros(scalar_size(self.nx, self.ny, self.ngrid))
        lib().og_sim_B(ctypes.c_void_p(self.h), i, _p(f), _p(w))
        return w

    def C(self, w):
        w = _c(w)
        f = np.zeros(2 * self.nb)
        lib().og_sim_C(ctypes.c_void_p(self.h), _p(w), _p(f))
        return f

    def Ainv(self, i, a):
        a = _c(a)
        w = np.zeros(scalar_size(self.nx, self.ny, self.ngrid))
        lib().og_sim_Ainv(ctypes.c_void_p(self.h), i, _p(a), _p(w))
        return w

    def cholesky(self, i):
        n = 2 * self.nb
        d = np.zeros(n)
        L = np.zeros((n, n))
        rc = lib().og_sim_cholesky(ctypes.c_void_p(self.h), i, _p(d), _p(L))
        if rc != 0:
            raise ValueError("no Cholesky factor (CG solver in use)")
        return d, L


def geom_points(geom, t, cap=100000):
    """Body points [x.., y..] and velocities [u.., v..] after moveBodies(t)."""
    xy = np.zeros(2 * cap)
    uv = np.zeros(2 * cap)
    n = lib().og_geom_points(geom.encode(), ctypes.c_double(t), _p(xy), _p(uv), cap)
    if n < 0:
        raise ValueError("oracle: bad geometry")
    return xy[:2 * n].copy(), uv[:2 * n].copy()


def geom_sequence(geom, times):
    """moveBodies at each time in order (stateful motions integrate over the
    calls); returns points (ntimes, 2n) [x.., y..] and velocities (ntimes, 2n)."""
    times = np.ascontiguousarray(times, dtype=np.float64)
    cap = 1 << 22
    xy = np.zeros(cap)
    uv = np.zeros(cap)
    n = lib().og_geom_sequence(geom.encode(), len(times), _p(times), _p(xy), _p(uv), cap)
    if n < 0:
        raise ValueError("oracle: bad geometry" if n == -1 else "oracle: sequence too long")
    k = 2 * n * len(times)
    return xy[:k].reshape(len(times), 2 * n).copy(), uv[:k].reshape(len(times), 2 * n).copy()


def baseflow_sequence(geom, nx, ny, ngrid, length, xoff, yoff, alpha_deg, times):
    """Unsteady base flow of a one-body geometry: the base flux (all levels)
    after moveFlow at each time, and (alpha, mag) after it."""
    times = np.ascontiguousarray(times, dtype=np.float64)
    nf = ngrid * (2 * nx * ny + nx + ny)
    q = np.zeros(nf * len(times))
    am = np.zeros(2 * len(times))
    D = ctypes.c_double
    rc = lib().og_baseflow_sequence(geom.encode(), nx, ny, ngrid, D(length), D(xoff), D(yoff), D(alpha_deg),
                                    len(times), _p(times), _p(q), _p(am))
    if rc != 0:
        raise ValueError("oracle: base flow needs one moving body")
    return q.reshape(len(times), nf), am.reshape(len(times), 2)


def ip_scalar(nx, ny, ngrid, dx, f, g):
    f = _c(f)
    g = _c(g)
    return lib().og_ip_scalar(nx, ny, ngrid, ctypes.c_double(dx), _p(f), _p(g))


def ip_flux(nx, ny, ngrid, p, q):
    p = _c(p)
    q = _c(q)
    return lib().og_ip_flux(nx, ny, ngrid, _p(p), _p(q))
