"""ibpm_amd -- MI355X (gfx950) projection-step library of the immersed-boundary
projection method, driven through the C ABI of include/ibpm_gpu.h.

This module is a thin ctypes layer over ibpm_amd/libibpm_gpu.so.  It never
falls back to a CPU path: if the library is missing or no GPU is visible the
calls raise.
"""
import ctypes
import os
import threading
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# IBPM_GPU_LIB selects another build of the library (e.g. a tuning variant).
LIB_PATH = os.environ.get("IBPM_GPU_LIB", os.path.join(HERE, "libibpm_gpu.so"))

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None

SCHEMES = {"euler": 0, "ab2": 1, "rk3": 2, "rk3b": 3}
MODELS = {"nonlinear": 0, "linear": 1, "adjoint": 2, "linearperiodic": 3, "sfd": 4}
WARM = {"zero": 0, "previous": 1, "learned": 2}
NN_IN, NN_HID = 13, 32
NN_PARAMS = NN_HID * NN_IN + NN_HID + 2 * NN_HID + 2


class IbpmError(RuntimeError):
    pass


class FlowParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("ngrid", ctypes.c_int),
                ("length", ctypes.c_double), ("xoffset", ctypes.c_double), ("yoffset", ctypes.c_double),
                ("reynolds", ctypes.c_double), ("dt", ctypes.c_double), ("scheme", ctypes.c_int),
                ("alpha_deg", ctypes.c_double), ("cg_tol", ctypes.c_double), ("force_cg", ctypes.c_int),
                ("geom", ctypes.c_char_p), ("ubf", ctypes.c_int), ("model", ctypes.c_int),
                ("period", ctypes.c_int), ("sfd_chi", ctypes.c_double), ("sfd_delta", ctypes.c_double)]


def load():
    """Loads libibpm_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IbpmError("libibpm_gpu.so not built: run `python -m ibpm_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.ibpm_last_error.restype = ctypes.c_char_p
    L.ibpm_batch_time.restype = ctypes.c_double
    L.ibpm_batch_time.argtypes = [ctypes.c_void_p]
    vp = ctypes.c_void_p
    for name, args in {
        "ibpm_ctx_destroy": [vp], "ibpm_ctx_synchronize": [vp],
        "ibpm_malloc": [vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)],
        "ibpm_free": [vp, vp],
        "ibpm_memcpy_h2d": [vp, vp, vp, ctypes.c_size_t],
        "ibpm_memcpy_d2h": [vp, vp, vp, ctypes.c_size_t],
        "ibpm_dst2d": [vp, ctypes.c_int, ctypes.c_int, vp, vp],
        "ibpm_elliptic2d_solve": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, vp, vp, vp],
        "ibpm_elliptic_solve": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_double, vp, vp],
        "ibpm_elliptic_solve_noalloc": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, vp, vp],
        "ibpm_elliptic_solve_precision": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, vp, vp, ctypes.c_int],
        "ibpm_batch_set_precision": [vp, ctypes.c_int],
        "ibpm_batch_precision_report": [vp, vp, vp, vp, vp, ctypes.c_int],
        "ibpm_curl_flux": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp, vp],
        "ibpm_curl_scalar": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                             ctypes.c_double, ctypes.c_double, vp, vp],
        "ibpm_nonlinear": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp, vp, vp],
        "ibpm_laplacian": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp, vp],
        "ibpm_cross_product": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp, vp, vp],
        "ibpm_cross_flux_flux": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp, vp, vp],
        "ibpm_batch_create": [vp, ctypes.c_int, ctypes.POINTER(FlowParams), ctypes.POINTER(ctypes.c_void_p)],
        "ibpm_batch_destroy": [vp], "ibpm_batch_init": [vp],
        "ibpm_batch_advance": [vp, ctypes.c_int],
        "ibpm_batch_num_points": [vp, ctypes.c_int],
        "ibpm_batch_get_omega": [vp, ctypes.c_int, vp], "ibpm_batch_set_omega": [vp, ctypes.c_int, vp],
        "ibpm_batch_get_q": [vp, ctypes.c_int, vp], "ibpm_batch_get_f": [vp, ctypes.c_int, vp],
        "ibpm_batch_get_points": [vp, ctypes.c_int, vp],
        "ibpm_batch_get_forces": [vp, ctypes.c_int, _D, _D],
        "ibpm_batch_force_history": [vp, ctypes.c_int, vp, ctypes.c_int],
        "ibpm_batch_base_flow": [vp, ctypes.c_int, _D, _D],
        "ibpm_geom_points": [ctypes.c_char_p, ctypes.c_int, vp, vp, vp, ctypes.c_long],
        "ibpm_batch_cg_iters": [vp, ctypes.c_int, vp, ctypes.c_int],
        "ibpm_batch_use_graph": [vp, ctypes.c_int],
        "ibpm_batch_set_base_state": [vp, ctypes.c_int, ctypes.c_int, vp, vp],
        "ibpm_batch_get_filtered": [vp, ctypes.c_int, vp],
        "ibpm_batch_set_filtered": [vp, ctypes.c_int, vp],
        "ibpm_batch_kernel_stats": [vp, ctypes.c_int, vp, vp, vp, ctypes.c_int],
        "ibpm_batch_set_warm_start": [vp, ctypes.c_int],
        "ibpm_batch_set_state": [vp, ctypes.c_int, vp, vp, ctypes.c_double, ctypes.c_int],
        "ibpm_batch_collect": [vp, ctypes.c_int, ctypes.c_long],
        "ibpm_batch_train": [vp, ctypes.c_int, ctypes.c_double, _D],
        "ibpm_batch_nn_params": [vp, vp, vp],
        "ibpm_nn_loss_grad": [vp, vp, vp, vp, ctypes.c_long, vp, _D],
        "ibpm_batch_set_precond": [vp, ctypes.c_int, ctypes.c_int],
        "ibpm_batch_set_precond_build": [vp, ctypes.c_int, ctypes.c_double],
        "ibpm_batch_precond_matrices": [vp, ctypes.c_int, ctypes.c_int, vp, vp],
        "ibpm_batch_set_local_helmholtz": [vp, ctypes.c_int],
        "ibpm_batch_cholesky_cache": [vp, ctypes.c_int, ctypes.c_char_p],
        "ibpm_batch_cholesky_loaded": [vp, ctypes.c_int],
        "ibpm_rhs": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                     ctypes.c_double, ctypes.c_double, ctypes.c_double, vp, vp, vp, vp, vp],
        "ibpm_flux_from_vorticity": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, vp, vp],
        "ibpm_reg_create": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_void_p)],
        "ibpm_reg_update": [vp, vp], "ibpm_reg_destroy": [vp],
        "ibpm_reg_to_flux": [vp, vp, vp], "ibpm_reg_to_boundary": [vp, vp, vp],
        "ibpm_batch_solver_stats": [vp, vp, ctypes.c_int],
    }.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = ctypes.c_int
    L.ibpm_batch_num_samples.restype = ctypes.c_long
    L.ibpm_batch_num_samples.argtypes = [ctypes.c_void_p]
    L.ibpm_kernel_class_name.restype = ctypes.c_char_p
    L.ibpm_kernel_class_name.argtypes = [ctypes.c_int]
    L.device_count = L.ibpm_device_count
    _lib = L
    return L


def _check(rc):
    if rc < 0:
        raise IbpmError(load().ibpm_last_error().decode())
    return rc


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class DeviceArray:
    """A device buffer of doubles owned by a Context."""

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, int(n)
        self._L = load()   # kept: module globals may be gone when __del__ runs at shutdown
        self.ptr = ctypes.c_void_p()
        _check(load().ibpm_malloc(ctx.h, 8 * max(self.n, 1), ctypes.byref(self.ptr)))

    @classmethod
    def of(cls, ctx, host):
        host = _c(host).ravel()
        d = cls(ctx, host.size)
        _check(load().ibpm_memcpy_h2d(ctx.h, d.ptr, host.ctypes.data, 8 * host.size))
        return d

    def get(self):
        out = np.empty(self.n)
        _check(load().ibpm_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, 8 * self.n))
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self._L.ibpm_free(self.ctx.h, self.ptr)
            self.ptr = None


class Context:
    """One GPU and one HIP stream (ibpm_ctx)."""

    def __init__(self, device=0):
        self._L = load()   # kept: module globals may be gone when __del__ runs at shutdown
        self.h = ctypes.c_void_p()
        _check(self._L.ibpm_ctx_create(device, ctypes.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self._L.ibpm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- EllipticSolver2d / EllipticSolver ------------------------------
    def dst2d(self, x):
        x = _c(x)
        n1, n2 = x.shape
        d_in = DeviceArray.of(self, x)
        d_out = DeviceArray(self, x.size)
        _check(load().ibpm_dst2d(self.h, n1, n2, d_in.ptr, d_out.ptr))
        return d_out.get().reshape(n1, n2)

    def elliptic2d_solve(self, kind, nx, ny, dx, alpha, f, bc=None):
        d_f = DeviceArray.of(self, f)
        d_u = DeviceArray(self, (nx - 1) * (ny - 1))
        d_bc = DeviceArray.of(self, bc) if bc is not None else None
        _check(load().ibpm_elliptic2d_solve(self.h, kind, nx, ny, dx, alpha, d_f.ptr,
                                            d_bc.ptr if d_bc else None, d_u.ptr))
        return d_u.get()

    def elliptic_solve(self, kind, nx, ny, ngrid, dx, alpha, f):
        d_f = DeviceArray.of(self, f)
        d_u = DeviceArray(self, ngrid * (nx - 1) * (ny - 1))
        _check(load().ibpm_elliptic_solve(self.h, kind, nx, ny, ngrid, dx, alpha, d_f.ptr, d_u.ptr))
        return d_u.get()

    def elliptic_solve_precision(self, kind, nx, ny, ngrid, dx, alpha, f, precision):
        """ibpm_elliptic_solve with FP64 (0) or FP32 (1) transforms."""
        d_f = DeviceArray.of(self, f)
        d_u = DeviceArray(self, ngrid * (nx - 1) * (ny - 1))
        _check(load().ibpm_elliptic_solve_precision(self.h, kind, nx, ny, ngrid, dx, alpha, d_f.ptr, d_u.ptr,
                                                    precision))
        return d_u.get()

    def elliptic_solve_noalloc(self, kind, nx, ny, ngrid, dx, alpha, f):
        """ibpm_elliptic_solve through the allocation-free entry point (the
        library's own buffers come from its fixed device scratch array)."""
        d_f = DeviceArray.of(self, f)
        d_u = DeviceArray(self, ngrid * (nx - 1) * (ny - 1))
        _check(load().ibpm_elliptic_solve_noalloc(self.h, kind, nx, ny, ngrid, dx, alpha, d_f.ptr, d_u.ptr))
        return d_u.get()

    # -- learned warm start: loss and gradient by the training kernels -----
    def nn_loss_grad(self, params, X, Y):
        params, X, Y = _c(params), _c(X), _c(Y)
        g = np.zeros(NN_PARAMS)
        loss = ctypes.c_double()
        _check(load().ibpm_nn_loss_grad(self.h, params.ctypes.data, X.ctypes.data, Y.ctypes.data,
                                        X.shape[0], g.ctypes.data, ctypes.byref(loss)))
        return loss.value, g

    # -- VectorOperations ------------------------------------------------
    def curl_flux(self, nx, ny, ngrid, dx, q):
        d_q = DeviceArray.of(self, q)
        d_w = DeviceArray(self, ngrid * (nx - 1) * (ny - 1))
        _check(load().ibpm_curl_flux(self.h, nx, ny, ngrid, dx, d_q.ptr, d_w.ptr))
        return d_w.get()

    def curl_scalar(self, nx, ny, ngrid, dx, w, u0=0.0, v0=0.0):
        d_w = DeviceArray.of(self, w)
        d_q = DeviceArray(self, ngrid * (2 * nx * ny + nx + ny))
        _check(load().ibpm_curl_scalar(self.h, nx, ny, ngrid, dx, u0, v0, d_w.ptr, d_q.ptr))
        return d_q.get()

    def laplacian(self, nx, ny, ngrid, dx, w):
        d_w = DeviceArray.of(self, w)
        d_o = DeviceArray(self, ngrid * (nx - 1) * (ny - 1))
        _check(load().ibpm_laplacian(self.h, nx, ny, ngrid, dx, d_w.ptr, d_o.ptr))
        return d_o.get()

    def nonlinear(self, nx, ny, ngrid, dx, q, w):
        d_q = DeviceArray.of(self, q)
        d_w = DeviceArray.of(self, w)
        d_n = DeviceArray(self, ngrid * (nx - 1) * (ny - 1))
        _check(load().ibpm_nonlinear(self.h, nx, ny, ngrid, dx, d_q.ptr, d_w.ptr, d_n.ptr))
        return d_n.get()

    def cross_product(self, nx, ny, ngrid, dx, q, w):
        """CrossProduct(Flux, Scalar) -> Flux."""
        d_q = DeviceArray.of(self, q)
        d_w = DeviceArray.of(self, w)
        d_o = DeviceArray(self, ngrid * (2 * nx * ny + nx + ny))
        _check(load().ibpm_cross_product(self.h, nx, ny, ngrid, dx, d_q.ptr, d_w.ptr, d_o.ptr))
        return d_o.get()

    def cross_flux_flux(self, nx, ny, ngrid, dx, q1, q2):
        """CrossProduct(Flux, Flux) -> Scalar."""
        d_1 = DeviceArray.of(self, q1)
        d_2 = DeviceArray.of(self, q2)
        d_o = DeviceArray(self, ngrid * (nx - 1) * (ny - 1))
        _check(load().ibpm_cross_flux_flux(self.h, nx, ny, ngrid, dx, d_1.ptr, d_2.ptr, d_o.ptr))
        return d_o.get()

    def rhs(self, nx, ny, ngrid, dx, dt, nu, an, bn, q, w, nprev=None):
        """(a, N) of one substep (IBSolver::advanceSubstep)."""
        ns = ngrid * (nx - 1) * (ny - 1)
        d_q = DeviceArray.of(self, q)
        d_w = DeviceArray.of(self, w)
        d_p = DeviceArray.of(self, nprev) if nprev is not None else None
        d_a = DeviceArray(self, ns)
        d_n = DeviceArray(self, ns)
        _check(load().ibpm_rhs(self.h, nx, ny, ngrid, dx, dt, nu, an, bn, d_q.ptr, d_w.ptr,
                               d_p.ptr if d_p is not None else None, d_a.ptr, d_n.ptr))
        return d_a.get(), d_n.get()

    def flux_from_vorticity(self, nx, ny, ngrid, dx, w, u0=0.0, v0=0.0):
        """NavierStokesModel::computeFlux."""
        d_w = DeviceArray.of(self, w)
        d_q = DeviceArray(self, ngrid * (2 * nx * ny + nx + ny))
        _check(load().ibpm_flux_from_vorticity(self.h, nx, ny, ngrid, dx, u0, v0, d_w.ptr, d_q.ptr))
        return d_q.get()


class Regularizer:
    """Regularizer of one set of points on the finest level
    (Regularizer::update / toFlux / toBoundary)."""

    def __init__(self, ctx, nx, ny, ngrid, length, xoffset, yoffset, xy):
        self.ctx, self.nx, self.ny, self.ngrid = ctx, nx, ny, ngrid
        xy = _c(xy)
        self.npts = len(xy) // 2
        self._reg_lib = load()
        self.h = ctypes.c_void_p()
        _check(load().ibpm_reg_create(ctx.h, nx, ny, ngrid, length, xoffset, yoffset, self.npts,
                                      xy.ctypes.data, ctypes.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self._reg_lib.ibpm_reg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def update(self, xy):
        xy = _c(xy)
        _check(load().ibpm_reg_update(self.h, xy.ctypes.data))
        return self

    def to_flux(self, f):
        d_f = DeviceArray.of(self.ctx, f)
        d_q = DeviceArray(self.ctx, self.ngrid * (2 * self.nx * self.ny + self.nx + self.ny))
        _check(load().ibpm_reg_to_flux(self.h, d_f.ptr, d_q.ptr))
        return d_q.get()

    def to_boundary(self, q):
        d_q = DeviceArray.of(self.ctx, q)
        d_f = DeviceArray(self.ctx, 2 * self.npts)
        _check(load().ibpm_reg_to_boundary(self.h, d_q.ptr, d_f.ptr))
        return d_f.get()


@dataclass
class Flow:
    """Parameters of one flow, named as the ibpm command line names them."""
    nx: int = 200
    ny: int = 200
    ngrid: int = 1
    length: float = 4.0
    xoffset: float = -2.0
    yoffset: float = -2.0
    Re: float = 100.0
    dt: float = 0.02
    geom: str = ""
    scheme: str = "rk3"
    alpha: float = 0.0
    tol: float = 1e-7
    force_cg: bool = False
    ubf: bool = False          # unsteady base flow (ibpm -ubf): solve in the frame of the only body
    model: str = "nonlinear"   # ibpm -model: nonlinear, linear, adjoint, linearperiodic, sfd
    period: int = 1            # linearperiodic: number of base states
    chi: float = 0.02          # sfd gain (ibpm -chi)
    Delta: float = 15.0        # sfd filter width (ibpm -Delta)

    def params(self):
        return FlowParams(self.nx, self.ny, self.ngrid, self.length, self.xoffset, self.yoffset,
                          self.Re, self.dt, SCHEMES[self.scheme], self.alpha, self.tol,
                          int(self.force_cg), self.geom.encode(), int(self.ubf), MODELS[self.model],
                          self.period, self.chi, self.Delta)


class Batch:
    """Independent flows advanced together (IBSolver::advance per flow)."""

    def __init__(self, ctx, flows):
        self.ctx = ctx
        self.flows = list(flows)
        arr = (FlowParams * len(self.flows))(*[f.params() for f in self.flows])
        self._keep = arr
        self._L = load()   # kept: module globals may be gone when __del__ runs at shutdown
        self.h = ctypes.c_void_p()
        _check(load().ibpm_batch_create(ctx.h, len(self.flows), arr, ctypes.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self._L.ibpm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self):
        _check(load().ibpm_batch_init(self.h))
        return self

    def cholesky_cache(self, basename, k=0):
        """Before init(): read the flow's Cholesky factors from the reference's
        <basename>_0i.cholesky files, or build and write them."""
        _check(load().ibpm_batch_cholesky_cache(self.h, k, basename.encode() if basename else None))
        return self

    def cholesky_loaded(self, k=0):
        return _check(load().ibpm_batch_cholesky_loaded(self.h, k))

    def set_base_state(self, omega, q, phase=0, k=0):
        """Base state (vorticity, flux) of a linear model's flow k: x0 of the
        linear and adjoint models, phase `phase` of the linear periodic one.
        Before init()."""
        omega, q = _c(omega), _c(q)
        _check(load().ibpm_batch_set_base_state(self.h, k, phase, omega.ctypes.data, q.ctypes.data))
        return self

    def filtered(self, k=0):
        """SFD: the filtered vorticity."""
        f = self.flows[k]
        return self._get("ibpm_batch_get_filtered", k, f.ngrid * (f.nx - 1) * (f.ny - 1))

    def set_filtered(self, w, k=0):
        w = _c(w)
        _check(load().ibpm_batch_set_filtered(self.h, k, w.ctypes.data))
        return self

    def use_graph(self, on=True):
        _check(load().ibpm_batch_use_graph(self.h, int(on)))
        return self

    def advance(self, nsteps=1):
        _check(load().ibpm_batch_advance(self.h, nsteps))

    def num_points(self, k=0):
        return _check(load().ibpm_batch_num_points(self.h, k))

    def _get(self, fn, k, n):
        out = np.empty(n)
        _check(getattr(load(), fn)(self.h, k, out.ctypes.data))
        return out

    def omega(self, k=0):
        f = self.flows[k]
        return self._get("ibpm_batch_get_omega", k, f.ngrid * (f.nx - 1) * (f.ny - 1))

    def set_omega(self, w, k=0):
        w = _c(w)
        _check(load().ibpm_batch_set_omega(self.h, k, w.ctypes.data))

    def set_state(self, omega, f, time, timestep, k=0):
        """Restart: vorticity, force, clock, then refreshState."""
        omega, f = _c(omega), _c(f)
        _check(load().ibpm_batch_set_state(self.h, k, omega.ctypes.data, f.ctypes.data, time, timestep))

    def q(self, k=0):
        f = self.flows[k]
        return self._get("ibpm_batch_get_q", k, f.ngrid * (2 * f.nx * f.ny + f.nx + f.ny))

    def f(self, k=0):
        return self._get("ibpm_batch_get_f", k, 2 * self.num_points(k))

    def points(self, k=0):
        return self._get("ibpm_batch_get_points", k, 2 * self.num_points(k))

    def forces(self, k=0):
        cd, cl = ctypes.c_double(), ctypes.c_double()
        _check(load().ibpm_batch_get_forces(self.h, k, ctypes.byref(cd), ctypes.byref(cl)))
        return cd.value, cl.value

    def force_history(self, k=0, max_steps=1 << 16):
        out = np.empty(2 * max_steps)
        n = _check(load().ibpm_batch_force_history(self.h, k, out.ctypes.data, max_steps))
        return out[:2 * n].reshape(n, 2)

    def base_flow(self, k=0):
        """(alpha in radians, magnitude) that rotate the current force into
        drag and lift, and whether the base flow is unsteady."""
        al, mag = ctypes.c_double(), ctypes.c_double()
        unsteady = _check(load().ibpm_batch_base_flow(self.h, k, ctypes.byref(al), ctypes.byref(mag)))
        return al.value, mag.value, bool(unsteady)

    def cg_iters(self, k=0):
        it = np.zeros(4, dtype=np.int32)
        n = _check(load().ibpm_batch_cg_iters(self.h, k, it.ctypes.data, 4))
        return it[:n].copy()

    # -- preconditioned CG force solve (moving bodies) ----------------------
    def set_precond(self, mode="lowrank", pairs=8):
        """mode "none": plain CG (the reference's solver); "dense": CG
        preconditioned by the dense FP64 P alone (its source and refresh:
        set_precond_build); "stale": the same (P from an earlier body position
        unless refreshed); "lowrank": P plus a low-rank update from the
        previous solve's search directions (`pairs` of them, at most 8)."""
        m = {"none": 0, "dense": 1, "stale": 1, "lowrank": 1}[mode]
        k = pairs if mode == "lowrank" else 0
        _check(load().ibpm_batch_set_precond(self.h, m, k))
        return self

    def set_precond_build(self, levels="all", rebuild_cost=1.0):
        """How the preconditioner's P is built: from M ("all" levels of the
        L^-1 solve, the default), from M's finest-level form ("finest": L^-1
        on level 0 under zero boundary values, a fraction of the cost per
        column; needs the local Helmholtz path), from the kernel table
        ("table": M = beta E G E^T with G read off its response to a unit flux
        at the grid's centre, no M applications), or from the table and
        refreshed before every solve by a Newton-Schulz step toward the
        table's M at the substep's points ("tracked").  rebuild_cost: a
        build's cost in M applications per unknown, the rebuild rule's
        break-even."""
        code = {"all": 0, "finest": 1, "table": 2, "tracked": 3}[levels]
        _check(load().ibpm_batch_set_precond_build(self.h, code, float(rebuild_cost)))
        return self

    def precond_matrices(self, flow=0, substep=2):
        """(kernel-table M of `substep` at the flow's current points, the P in
        use scaled to that substep or None), both 2 nb x 2 nb.  After advance,
        the points are those of the scheme's last substep (2 for RK3)."""
        n = 2 * self.num_points(flow)
        sub = substep
        m = np.zeros((n, n))
        p = np.zeros((n, n))
        _check(load().ibpm_batch_precond_matrices(self.h, flow, sub, m.ctypes.data, None))
        try:
            _check(load().ibpm_batch_precond_matrices(self.h, flow, sub, None, p.ctypes.data))
        except IbpmError:
            p = None
        return m, p

    PRECISION = {"fp64": 0, "fp32_step": 1, "fp32": 2, "reduced_tuned": 3, "gemm_step": 4}

    def precision_report(self):
        """{N: (uses GEMM-DST, radix us, GEMM us)} of the step's solves under
        modes reduced_tuned / gemm_step."""
        n = np.zeros(16, dtype=np.int32)
        g = np.zeros(16, dtype=np.int32)
        tr = np.zeros(16)
        tg = np.zeros(16)
        k = _check(load().ibpm_batch_precision_report(self.h, n.ctypes.data, g.ctypes.data, tr.ctypes.data,
                                                      tg.ctypes.data, 16))
        return {int(n[i]): (bool(g[i]), float(tr[i]), float(tg[i])) for i in range(k)}

    def set_precision(self, mode="fp64"):
        """Transform precision of the elliptic solves: "fp64" (the reference's),
        "fp32_step" (FP32 transforms in the step's own solves), "fp32" (also
        inside the force solve's M applications), "reduced_tuned" (step solves
        by FP32 radix or split-BF16 GEMM-DST on MFMA, whichever the autotuner
        times faster per length N <= 511) or "gemm_step" (GEMM-DST wherever it
        applies); fields stay FP64."""
        _check(load().ibpm_batch_set_precision(self.h, self.PRECISION[mode]))
        return self

    def set_local_helmholtz(self, on=True):
        """Inside the moving-body force solve, A^-1 by Jacobi sweeps over a box
        around the body on each level where that equals the DST solve to
        rounding (the default), or always by the DST solve (on=False)."""
        _check(load().ibpm_batch_set_local_helmholtz(self.h, 1 if on else 0))
        return self

    SOLVER_STATS = ("iterations", "solves", "m_applications", "build_m_applications", "builds",
                    "fallbacks", "abandoned", "host_us_kinematics", "host_us_polls", "local_helmholtz_m_applications",
                    "build_us", "build_host_us", "refreshes")

    def solver_stats(self):
        out = np.zeros(len(self.SOLVER_STATS), dtype=np.int64)
        n = _check(load().ibpm_batch_solver_stats(self.h, out.ctypes.data, len(out)))
        return dict(zip(self.SOLVER_STATS[:n], (int(v) for v in out[:n])))

    # -- learned CG warm start (moving bodies) ------------------------------
    def set_warm_start(self, mode):
        _check(load().ibpm_batch_set_warm_start(self.h, WARM[mode]))
        return self

    def collect(self, enable=True, capacity=1 << 20):
        _check(load().ibpm_batch_collect(self.h, int(enable), capacity))
        return self

    def num_samples(self):
        return load().ibpm_batch_num_samples(self.h)

    def train(self, iters=500, lr=3e-3):
        loss = ctypes.c_double()
        _check(load().ibpm_batch_train(self.h, iters, lr, ctypes.byref(loss)))
        return loss.value

    def nn_params(self, new=None):
        out = np.zeros(NN_PARAMS)
        src = _c(new) if new is not None else None
        _check(load().ibpm_batch_nn_params(self.h, out.ctypes.data, src.ctypes.data if src is not None else None))
        return out

    def kernel_stats(self, nsteps=1):
        """Advances nsteps with a HIP event pair around every launch; returns
        {kernel class: (launches, total ms, algorithmic bytes)}."""
        n = 32
        counts = np.zeros(n, dtype=np.int32)
        ms = np.zeros(n)
        by = np.zeros(n)
        k = _check(load().ibpm_batch_kernel_stats(self.h, nsteps, counts.ctypes.data, ms.ctypes.data,
                                                  by.ctypes.data, n))
        return {load().ibpm_kernel_class_name(c).decode(): (int(counts[c]), float(ms[c]), float(by[c]))
                for c in range(k) if counts[c]}

    @property
    def time(self):
        return load().ibpm_batch_time(self.h)


def advance_concurrently(batches, nsteps):
    """Advances batches that live on different contexts (one HIP stream each)
    by nsteps each at the same time, one host thread per batch, and returns
    when all have finished.  Independent groups of flows on separate streams
    let one group's bandwidth-bound stencils fill the HBM that another group's
    occupancy-bound DST passes leave idle.  Raises the first error a batch
    reported (after every thread has ended)."""
    batches = list(batches)
    ctxs = [b.ctx.h.value for b in batches]
    if len(set(ctxs)) != len(ctxs):
        raise ValueError("advance_concurrently: each batch needs a context (stream) of its own")
    errs = [None] * len(batches)

    def run(i):
        try:
            batches[i].advance(nsteps)
        except Exception as e:   # reported after the join
            errs[i] = e

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(batches))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for e in errs:
        if e is not None:
            raise e
