"""ctypes binding of libflame3d.so (include/flame3d.h).

Device arrays are torch tensors on a ROCm device.  The reference's
column-major [i, j, k, var] layout is a C-contiguous torch tensor of shape
(NV, Nz+2NG, Ny+2NG, Nx+2NG) (variable slowest, i fastest); face arrays are
(NV, nz, ny, nx) with the face direction's extent one larger.  Every call
checks shapes, dtypes and devices on the host before the launch, and raises
F3DError on a non-zero return code.  There is no CPU fallback: if the
library is missing, loading fails.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflame3d.so")

F3D_FP64, F3D_FP32 = 0, 1
CHEM_NONE, CHEM_STIFF, CHEM_EXPLICIT, CHEM_NEWTON, CHEM_SURROGATE = 0, 1, 2, 3, 4
BC_KEEP, BC_PERIODIC, BC_COPY, BC_INFLOW = 0, 1, 2, 3
NG = 4

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p


class F3DError(RuntimeError):
    pass


class Dims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("Nx", "Ny", "Nz", "NG", "Ns", "Nr")]


class Thermo(C.Structure):
    _fields_ = [("Ru", C.c_double), ("min_temp", C.c_double), ("max_temp", C.c_double),
                ("mw", _dp), ("coeffs_sep", _dp), ("coeffs_lo", _dp), ("coeffs_hi", _dp),
                ("visc_poly", _dp), ("conduct_poly", _dp), ("binarydiff_poly", _dp)]


class Kinetics(C.Structure):
    _fields_ = [("atm", C.c_double), ("reaction_type", _i64p), ("sgm", _i64p), ("vf", _i64p),
                ("vr", _i64p), ("Arr", _dp), ("ef", _dp), ("loP", _dp), ("Troe", _dp)]


class Scheme(C.Structure):
    _fields_ = [("SWeps2", C.c_double), ("WENOeps", C.c_double), ("hybrid_phi1", C.c_double),
                ("hybrid_phi2", C.c_double), ("UP7", C.c_double * 7), ("T_criteria", C.c_double),
                ("precision", C.c_int32), ("bug_compat", C.c_int32), ("sensor_kind", C.c_int32),
                ("recon_kind", C.c_int32), ("stage_kind", C.c_int32), ("reserved", C.c_int32 * 3)]


NN_NPARAM = 12512  # F3D_NN_NPARAM
SENSOR_PRESSURE, SENSOR_DUCROS = 0, 1  # F3D_SENSOR_*
RECON_REFERENCE, RECON_WENO7 = 0, 1    # F3D_RECON_*
STAGE_FUSED, STAGE_GRANULAR = 0, 1     # F3D_STAGE_*
KIN_RUNTIME, KIN_H2_LIDRYER, KIN_AIR5 = 0, 1, 2  # F3D_KIN_*
# Ducros switch (UP7 below theta = 0.1, WENO5-Z above, never NND)
DUCROS_SCHEME = dict(sensor_kind=SENSOR_DUCROS, hybrid_phi1=0.1, hybrid_phi2=2.0)


class NNConfig(C.Structure):
    _fields_ = [("in_mean", C.c_float * 11), ("in_std", C.c_float * 11),
                ("out_mean", C.c_float * 9), ("out_std", C.c_float * 9),
                ("mahal_mean", C.c_float * 11), ("mahal_inv", C.c_float * 121),
                ("d_max", C.c_float), ("elem_tol", C.c_float), ("neg_tol", C.c_float),
                ("res_tol", C.c_float), ("res_T_tol", C.c_float), ("dT_max", C.c_float),
                ("chi_max", C.c_float), ("chain_max", C.c_float), ("chain_species", C.c_int32),
                ("boxcox_lambda", C.c_float), ("label_mode", C.c_int32),
                ("label_scale", C.c_float * 9),
                ("label_rel", C.c_float), ("label_floor", C.c_float), ("y_floor", C.c_float),
                ("n_elem", C.c_int32),
                ("elem", (C.c_float * 9) * 4), ("dt", C.c_double)]


# exported symbols and their argument types (all return int unless noted)
_PROTOS = {
    "f3d_create": [C.POINTER(Dims), C.POINTER(Thermo), C.POINTER(Kinetics), C.POINTER(Scheme),
                   C.POINTER(_vp)],
    "f3d_destroy": [_vp],
    "f3d_real_size": [_vp],
    "f3d_kinetics_source": [_vp],
    "f3d_flux_split": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_species_split": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_shock_sensor": [_vp, _vp, _vp, _vp],
    "f3d_weno": [_vp, C.c_int32, _vp, _vp, _vp, _vp, C.c_int32, _vp, _vp],
    "f3d_viscous_flux": [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_spec_viscous_flux": [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_div": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp, _vp],
    "f3d_div_specs": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp, _vp],
    "f3d_mixture": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_c2prim": [_vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_prim2c": [_vp, _vp, _vp, _vp],
    "f3d_lincomb": [_vp, _vp, _vp, C.c_int32, C.c_double, C.c_double, _vp, _vp],
    "f3d_get_y": [_vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_set_boundary": [_vp, _vp, _vp, _vp],
    "f3d_fill": [_vp, _vp, _vp, _vp, _vp],
    "f3d_chem_stiff": [_vp, _vp, _vp, _vp, C.c_double, _vp, _vp],
    "f3d_chem_explicit": [_vp, _vp, _vp, _vp, C.c_double, _vp, _vp],
    "f3d_chem_newton": [_vp, _vp, _vp, _vp, C.c_double, C.c_int32, C.c_double, C.c_double,
                        _vp, _vp, _vp],
    "f3d_nn_init": [_vp, C.POINTER(NNConfig)],
    "f3d_nn_set_params": [_vp, _vp],
    "f3d_nn_get_params": [_vp, _vp],
    "f3d_nn_train_step": [_vp, _vp, _vp, C.c_int32, C.c_float, _vp, _vp],
    "f3d_nn_train_step_static": [_vp, _vp, _vp, C.c_int32, C.c_float, _vp, _vp],
    "f3d_nn_grad": [_vp, _vp, _vp, C.c_int32, _vp, _vp],
    "f3d_nn_forward": [_vp, _vp, _vp, C.c_int32, _vp],
    "f3d_nn_chem": [_vp, _vp, _vp, _vp, C.c_double, C.c_int32, C.c_double, C.c_double, _vp,
                    C.POINTER(C.c_int64), _vp, _vp],
    "f3d_set_metrics": [_vp, _vp, _vp],
    "f3d_mesh_ghosts": [_vp, _vp, _vp, _vp, _vp],
    "f3d_mesh_metrics": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_rk3_stage": [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp],
    "f3d_workspace": [_vp, C.c_char_p],
    "f3d_time_step": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                      C.c_double, C.c_double, C.POINTER(C.c_int64), _vp, _vp],
    "f3d_nan_check": [_vp, _vp, C.POINTER(C.c_int32), _vp],
    "f3d_nn_timing": [_vp, C.c_int32, C.POINTER(C.c_double)],
    "f3d_sweep_conv": [_vp, C.c_int32, _vp, _vp, _vp, _vp, C.c_int32, _vp],
    "f3d_sweep_diff": [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "f3d_sweep_stage_end": [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp],
    "f3d_last_error": [],
}

_lib = None


def exported_symbols():
    return list(_PROTOS)


def load_library(path: str = LIB_PATH):
    """Load libflame3d.so and declare every prototype.  Raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise F3DError(f"{path} is not built; run `python flame3d.jl_amd/build.py` "
                       "(there is no fallback path)")
    lib = C.CDLL(path)
    for name, args in _PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.f3d_last_error.restype = C.c_char_p
    lib.f3d_workspace.restype = _vp
    _lib = lib
    return lib


def _rc(rc: int, what: str):
    if rc != 0:
        msg = _lib.f3d_last_error().decode()
        raise F3DError(f"{what} failed ({rc}): {msg}")


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def _i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64).ravel(order="F"))


UP7 = (0.0, 17 / 600, -23 / 120, 22 / 30, 0.5, -0.075, 0.005)  # run.jl:57


class Context:
    """One f3d_ctx: block geometry, parameters and flux precision."""

    def __init__(self, Nx: int, Ny: int, Nz: int, mech, precision: str = "fp64",
                 scheme: Optional[dict] = None, bug_compat: int = 0):
        import torch
        self.torch = torch
        self.lib = load_library()
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.mech = mech
        self.Ns, self.Nr = mech.Ns, mech.Nr
        self.shape3 = (Nz + 2 * NG, Ny + 2 * NG, Nx + 2 * NG)
        self.NP = int(np.prod(self.shape3))
        self.precision = precision
        self.rdtype = torch.float64 if precision == "fp64" else torch.float32
        s = dict(SWeps2=0.0, WENOeps=1e-26, hybrid_phi1=1e-3, hybrid_phi2=0.1, UP7=UP7,
                 T_criteria=500.0, sensor_kind=SENSOR_PRESSURE, recon_kind=RECON_REFERENCE,
                 stage_kind=STAGE_FUSED)
        if scheme:
            s.update(scheme)
        self.scheme = s
        m = mech
        keep = [_f64(x) for x in (m.mw, m.coeffs_sep, m.coeffs_lo, m.coeffs_hi, m.visc_poly,
                                  m.conduct_poly, m.binarydiff_poly)]
        th = Thermo(m.Ru, m.min_temp, m.max_temp, *[k.ctypes.data_as(_dp) for k in keep])
        if m.Nr:
            ki = [_i64(x) for x in (m.reaction_type, m.sgm, m.vf, m.vr)]
            kf = [_f64(x) for x in (m.Arr, m.ef, m.loP, m.Troe)]
            kin = Kinetics(m.atm, *[k.ctypes.data_as(_i64p) for k in ki],
                           *[k.ctypes.data_as(_dp) for k in kf])
        else:
            kin = Kinetics(m.atm)
        sc = Scheme(s["SWeps2"], s["WENOeps"], s["hybrid_phi1"], s["hybrid_phi2"],
                    (C.c_double * 7)(*s["UP7"]), s["T_criteria"],
                    F3D_FP64 if precision == "fp64" else F3D_FP32, bug_compat,
                    int(s["sensor_kind"]), int(s["recon_kind"]), int(s["stage_kind"]))
        dims = Dims(Nx, Ny, Nz, NG, m.Ns, m.Nr)
        h = _vp()
        _rc(self.lib.f3d_create(C.byref(dims), C.byref(th), C.byref(kin), C.byref(sc),
                                C.byref(h)), "f3d_create")
        self.h = h
        self._metrics_keep = None

    def kinetics_source(self) -> int:
        """F3D_KIN_*: compiled-in reaction tables of a shipped mechanism, or
        the run-time parameter block."""
        k = self.lib.f3d_kinetics_source(self.h)
        if k < 0:
            _rc(k, "f3d_kinetics_source")
        return k

    def close(self):
        if getattr(self, "h", None):
            self.lib.f3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- helpers -------------------------------------------------------
    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def zeros(self, nv=None, dtype=None):
        t = self.torch
        sh = self.shape3 if nv is None else (nv,) + self.shape3
        return t.zeros(sh, dtype=dtype or t.float64, device="cuda")

    def rzeros(self, nv=None):
        return self.zeros(nv, self.rdtype)

    def face_shape(self, d, nv):
        n = [self.Nz, self.Ny, self.Nx]
        n[2 - d] += 1
        return (nv, *n)

    def face_zeros(self, d, nv):
        return self.torch.zeros(self.face_shape(d, nv), dtype=self.rdtype, device="cuda")

    def _chk(self, t, shape, dtype, name):
        if t is None:
            raise F3DError(f"{name} is None")
        if not t.is_cuda or not t.is_contiguous():
            raise F3DError(f"{name} must be a contiguous device tensor")
        if tuple(t.shape) != tuple(shape):
            raise F3DError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
        if t.dtype != dtype:
            raise F3DError(f"{name} has dtype {t.dtype}, expected {dtype}")
        return t.data_ptr()

    def _pad(self, t, nv, name, dtype=None):
        sh = self.shape3 if nv is None else (nv,) + self.shape3
        return self._chk(t, sh, dtype or self.torch.float64, name)

    def _rpad(self, t, nv, name):
        return self._pad(t, nv, name, self.rdtype)

    def _face(self, t, d, nv, name):
        return self._chk(t, self.face_shape(d, nv), self.rdtype, name)

    def _tag(self, tag):
        if tag is None:
            return None
        return self._chk(tag, self.shape3, self.torch.int64, "tag")

    def _mptrs(self, metrics):
        if len(metrics) != 9:
            raise F3DError("metrics must hold 9 arrays")
        arr = (C.c_void_p * 9)(*[self._pad(m, None, f"metrics[{q}]") for q, m in enumerate(metrics)])
        return arr

    # ---- kernel-granular API -------------------------------------------
    def flux_split(self, Q, Fp, Fm, Ax, Ay, Az, tag=None):
        _rc(self.lib.f3d_flux_split(self.h, self._pad(Q, 7, "Q"), self._rpad(Fp, 5, "Fp"),
                                    self._rpad(Fm, 5, "Fm"), self._pad(Ax, None, "Ax"),
                                    self._pad(Ay, None, "Ay"), self._pad(Az, None, "Az"),
                                    self._tag(tag), self.stream()), "f3d_flux_split")

    def species_split(self, rhoi, Q, Fp, Fm, Ax, Ay, Az, tag=None):
        Ns = self.Ns
        _rc(self.lib.f3d_species_split(self.h, self._pad(rhoi, Ns, "rhoi"), self._pad(Q, 7, "Q"),
                                       self._rpad(Fp, Ns, "Fp"), self._rpad(Fm, Ns, "Fm"),
                                       self._pad(Ax, None, "Ax"), self._pad(Ay, None, "Ay"),
                                       self._pad(Az, None, "Az"), self._tag(tag), self.stream()),
            "f3d_species_split")

    def shock_sensor(self, phi, Q):
        _rc(self.lib.f3d_shock_sensor(self.h, self._rpad(phi, None, "phi"), self._pad(Q, 7, "Q"),
                                      self.stream()), "f3d_shock_sensor")

    def weno(self, d, F, phi, Fp, Fm, NV, tag=None):
        _rc(self.lib.f3d_weno(self.h, d, self._face(F, d, NV, "F"), self._rpad(phi, None, "phi"),
                              self._rpad(Fp, NV, "Fp"), self._rpad(Fm, NV, "Fm"), NV,
                              self._tag(tag), self.stream()), "f3d_weno")

    def viscous_flux(self, d, Fv, Q, metrics, J, lam, mu, Fh, tag=None):
        _rc(self.lib.f3d_viscous_flux(self.h, d, self._face(Fv, d, 4, "Fv"), self._pad(Q, 7, "Q"),
                                      self._mptrs(metrics), self._pad(J, None, "J"),
                                      self._rpad(lam, None, "lambda"), self._rpad(mu, None, "mu"),
                                      self._face(Fh, d, 3, "Fh"), self._tag(tag), self.stream()),
            "f3d_viscous_flux")

    def spec_viscous_flux(self, d, Fd, Q, Yi, metrics, J, D, Fh, tag=None):
        Ns = self.Ns
        _rc(self.lib.f3d_spec_viscous_flux(self.h, d, self._face(Fd, d, Ns, "Fd"),
                                           self._pad(Q, 7, "Q"), self._rpad(Yi, Ns, "Yi"),
                                           self._mptrs(metrics), self._pad(J, None, "J"),
                                           self._rpad(D, Ns, "D"), self._face(Fh, d, 3, "Fh"),
                                           self._tag(tag), self.stream()),
            "f3d_spec_viscous_flux")

    def div(self, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J, tag=None):
        _rc(self.lib.f3d_div(self.h, self._pad(U, 5, "U"), self._face(Fx, 0, 5, "Fx"),
                             self._face(Fy, 1, 5, "Fy"), self._face(Fz, 2, 5, "Fz"),
                             self._face(Fvx, 0, 4, "Fvx"), self._face(Fvy, 1, 4, "Fvy"),
                             self._face(Fvz, 2, 4, "Fvz"), dt, self._pad(J, None, "J"),
                             self._tag(tag), self.stream()), "f3d_div")

    def div_specs(self, rhoi, Fx, Fy, Fz, Fdx, Fdy, Fdz, dt, J, tag=None):
        Ns = self.Ns
        _rc(self.lib.f3d_div_specs(self.h, self._pad(rhoi, Ns, "rhoi"), self._face(Fx, 0, Ns, "Fx"),
                                   self._face(Fy, 1, Ns, "Fy"), self._face(Fz, 2, Ns, "Fz"),
                                   self._face(Fdx, 0, Ns, "Fdx"), self._face(Fdy, 1, Ns, "Fdy"),
                                   self._face(Fdz, 2, Ns, "Fdz"), dt, self._pad(J, None, "J"),
                                   self._tag(tag), self.stream()), "f3d_div_specs")

    def mixture(self, Q, rhoi, Yi, lam, mu, D, tag=None):
        Ns = self.Ns
        _rc(self.lib.f3d_mixture(self.h, self._pad(Q, 7, "Q"), self._pad(rhoi, Ns, "rhoi"),
                                 self._rpad(Yi, Ns, "Yi"), self._rpad(lam, None, "lambda"),
                                 self._rpad(mu, None, "mu"), self._rpad(D, Ns, "D"),
                                 self._tag(tag), self.stream()), "f3d_mixture")

    def c2prim(self, U, Q, rhoi, tag=None):
        _rc(self.lib.f3d_c2prim(self.h, self._pad(U, 5, "U"), self._pad(Q, 7, "Q"),
                                self._pad(rhoi, self.Ns, "rhoi"), self._tag(tag), self.stream()),
            "f3d_c2prim")

    def prim2c(self, U, Q):
        _rc(self.lib.f3d_prim2c(self.h, self._pad(U, 5, "U"), self._pad(Q, 7, "Q"), self.stream()),
            "f3d_prim2c")

    def lincomb(self, U, Un, NV, a, b, tag=None):
        _rc(self.lib.f3d_lincomb(self.h, self._pad(U, NV, "U"), self._pad(Un, NV, "Un"), NV, a, b,
                                 self._tag(tag), self.stream()), "f3d_lincomb")

    def get_y(self, Yi, rhoi, Q, tag=None):
        Ns = self.Ns
        _rc(self.lib.f3d_get_y(self.h, self._pad(Yi, Ns, "Yi"), self._pad(rhoi, Ns, "rhoi"),
                               self._pad(Q, 7, "Q"), self._tag(tag), self.stream()), "f3d_get_y")

    def set_boundary(self, kinds: Sequence[int], inflow_Q=None, inflow_rhoi=None):
        k = (C.c_int32 * 6)(*kinds)
        iq = None if inflow_Q is None else (C.c_double * 7)(*inflow_Q)
        ir = None if inflow_rhoi is None else (C.c_double * self.Ns)(*inflow_rhoi)
        _rc(self.lib.f3d_set_boundary(self.h, k, iq, ir), "f3d_set_boundary")

    def fill(self, Q, U, rhoi):
        _rc(self.lib.f3d_fill(self.h, self._pad(Q, 7, "Q"), self._pad(U, 5, "U"),
                              self._pad(rhoi, self.Ns, "rhoi"), self.stream()), "f3d_fill")

    def chem_stiff(self, U, Q, rhoi, dt, tag=None):
        _rc(self.lib.f3d_chem_stiff(self.h, self._pad(U, 5, "U"), self._pad(Q, 7, "Q"),
                                    self._pad(rhoi, self.Ns, "rhoi"), dt, self._tag(tag),
                                    self.stream()), "f3d_chem_stiff")

    def chem_explicit(self, U, Q, rhoi, dt, tag=None):
        _rc(self.lib.f3d_chem_explicit(self.h, self._pad(U, 5, "U"), self._pad(Q, 7, "Q"),
                                       self._pad(rhoi, self.Ns, "rhoi"), dt, self._tag(tag),
                                       self.stream()), "f3d_chem_explicit")

    def chem_newton(self, U, Q, rhoi, dt, maxiter=50, rtol=1e-10, ttol=1e-6, iters=None,
                    tag=None):
        it = None
        if iters is not None:
            it = self._chk(iters, (self.Nz, self.Ny, self.Nx), self.torch.int32, "iters")
        _rc(self.lib.f3d_chem_newton(self.h, self._pad(U, 5, "U"), self._pad(Q, 7, "Q"),
                                     self._pad(rhoi, self.Ns, "rhoi"), dt, maxiter, rtol, ttol,
                                     it, self._tag(tag), self.stream()), "f3d_chem_newton")

    # ---- learned chemistry surrogate ------------------------------------
    def nn_init(self, cfg: "NNConfig"):
        _rc(self.lib.f3d_nn_init(self.h, C.byref(cfg)), "f3d_nn_init")
        self._nn_cfg = cfg

    def nn_set_params(self, params):
        a = np.ascontiguousarray(np.asarray(params, dtype=np.float32).ravel())
        if a.size != NN_NPARAM:
            raise F3DError(f"params: expected {NN_NPARAM} floats, got {a.size}")
        _rc(self.lib.f3d_nn_set_params(self.h, a.ctypes.data), "f3d_nn_set_params")

    def nn_get_params(self):
        a = np.zeros(NN_NPARAM, np.float32)
        _rc(self.lib.f3d_nn_get_params(self.h, a.ctypes.data), "f3d_nn_get_params")
        return a

    def _nn_batch(self, X, Y, ny):
        t = self.torch
        nin = self.Ns + 2
        if X.dim() != 2 or X.shape[1] != nin:
            raise F3DError(f"X: expected (B, {nin}), got {tuple(X.shape)}")
        B = X.shape[0]
        self._chk(X, (B, nin), t.float32, "X")
        self._chk(Y, (B, ny), t.float32, "Y")
        return B

    def nn_train_step(self, X, Y, lr, loss=None):
        """One Adam step; loss: optional device float32 tensor of one element."""
        B = self._nn_batch(X, Y, self.Ns)
        lp = None
        if loss is not None:
            lp = self._chk(loss, (1,), self.torch.float32, "loss")
        _rc(self.lib.f3d_nn_train_step(self.h, X.data_ptr(), Y.data_ptr(), B, lr, lp,
                                       self.stream()), "f3d_nn_train_step")

    def nn_train_step_static(self, X, Y, lr, loss=None):
        """nn_train_step on the library's static scratch (no device
        allocation; batch <= 8192)."""
        B = self._nn_batch(X, Y, self.Ns)
        lp = None
        if loss is not None:
            lp = self._chk(loss, (1,), self.torch.float32, "loss")
        _rc(self.lib.f3d_nn_train_step_static(self.h, X.data_ptr(), Y.data_ptr(), B, lr, lp,
                                              self.stream()), "f3d_nn_train_step_static")

    def nn_grad(self, X, Y):
        """Gradient (NN_NPARAM floats) and loss of the batch, as a device tensor."""
        B = self._nn_batch(X, Y, self.Ns)
        g = self.torch.empty(NN_NPARAM + 1, dtype=self.torch.float32, device=X.device)
        _rc(self.lib.f3d_nn_grad(self.h, X.data_ptr(), Y.data_ptr(), B, g.data_ptr(),
                                 self.stream()), "f3d_nn_grad")
        return g

    def nn_forward(self, X):
        Y = self.torch.empty((X.shape[0], self.Ns), dtype=self.torch.float32, device=X.device)
        B = self._nn_batch(X, Y, self.Ns)
        _rc(self.lib.f3d_nn_forward(self.h, X.data_ptr(), Y.data_ptr(), B, self.stream()),
            "f3d_nn_forward")
        return Y

    def nn_chem(self, U, Q, rhoi, dt, maxiter=50, rtol=1e-10, ttol=1e-6, status=None,
                tag=None):
        """Surrogate chemistry step with Newton fallback; returns the fallback count."""
        st = None
        if status is not None:
            st = self._chk(status, (self.Nz, self.Ny, self.Nx), self.torch.int8, "status")
        nfb = C.c_int64(0)
        _rc(self.lib.f3d_nn_chem(self.h, self._pad(U, 5, "U"), self._pad(Q, 7, "Q"),
                                 self._pad(rhoi, self.Ns, "rhoi"), dt, maxiter, rtol, ttol, st,
                                 C.byref(nfb), self._tag(tag), self.stream()), "f3d_nn_chem")
        return int(nfb.value)

    # ---- stage level -----------------------------------------------------
    def set_metrics(self, metrics, J):
        ptrs = self._mptrs(metrics)
        _rc(self.lib.f3d_set_metrics(self.h, ptrs, self._pad(J, None, "J")), "f3d_set_metrics")
        self._metrics_keep = (list(metrics), J)

    def mesh_ghosts(self, x, y, z):
        """Fill the ghost nodes of padded coordinates in place (parse_mesh.jl:52-113)."""
        _rc(self.lib.f3d_mesh_ghosts(self.h, self._pad(x, None, "x"), self._pad(y, None, "y"),
                                     self._pad(z, None, "z"), self.stream()), "f3d_mesh_ghosts")

    def mesh_metrics(self, x, y, z):
        """(9 cofactor metrics, J) at every padded point from padded node
        coordinates (parse_mesh.jl:132-221), as new device arrays."""
        metrics = [self.zeros() for _ in range(9)]
        J = self.zeros()
        _rc(self.lib.f3d_mesh_metrics(self.h, self._pad(x, None, "x"), self._pad(y, None, "y"),
                                      self._pad(z, None, "z"), self._mptrs(metrics),
                                      self._pad(J, None, "J"), self.stream()),
            "f3d_mesh_metrics")
        return metrics, J

    def rk3_stage(self, KRK, U, Un, rhoi, rhon, Q, dt, tag=None):
        if self._metrics_keep is None:
            raise F3DError("set_metrics first")
        Ns = self.Ns
        _rc(self.lib.f3d_rk3_stage(self.h, KRK, self._pad(U, 5, "U"), self._pad(Un, 5, "Un"),
                                   self._pad(rhoi, Ns, "rhoi"), self._pad(rhon, Ns, "rhon"),
                                   self._pad(Q, 7, "Q"), dt, self._tag(tag), self.stream()),
            "f3d_rk3_stage")

    def time_step(self, U, Un, rhoi, rhon, Q, dt, chem=CHEM_NONE, sub_step=1, maxiter=50,
                  rtol=1e-10, ttol=1e-6, tag=None):
        """One Strang time step (solver.jl:240-378); returns the surrogate's
        fallback count (0 for the other chemistry forms)."""
        if self._metrics_keep is None:
            raise F3DError("set_metrics first")
        Ns = self.Ns
        nfb = C.c_int64(0)
        _rc(self.lib.f3d_time_step(self.h, self._pad(U, 5, "U"), self._pad(Un, 5, "Un"),
                                   self._pad(rhoi, Ns, "rhoi"), self._pad(rhon, Ns, "rhon"),
                                   self._pad(Q, 7, "Q"), dt, chem, sub_step, maxiter, rtol, ttol,
                                   C.byref(nfb), self._tag(tag), self.stream()), "f3d_time_step")
        return int(nfb.value)

    # ---- fused float stage pieces (F3D_STAGE_FUSED) ---------------------
    def rhs_zeros(self):
        """The fused stage's right-hand side [5 + Ns][Nz][Ny][Nx], float."""
        return self.torch.zeros((5 + self.Ns, self.Nz, self.Ny, self.Nx),
                                dtype=self.torch.float32, device="cuda")

    def _rhs(self, rhs):
        return self._chk(rhs, (5 + self.Ns, self.Nz, self.Ny, self.Nx), self.torch.float32, "rhs")

    def sweep_conv(self, d, rhs, Q, rhoi, phi, accumulate):
        _rc(self.lib.f3d_sweep_conv(self.h, d, self._rhs(rhs), self._pad(Q, 7, "Q"),
                                    self._pad(rhoi, self.Ns, "rhoi"),
                                    self._rpad(phi, None, "phi"), int(bool(accumulate)),
                                    self.stream()), "f3d_sweep_conv")

    def sweep_diff(self, d, rhs, Q, Yi, lam, mu, D):
        _rc(self.lib.f3d_sweep_diff(self.h, d, self._rhs(rhs), self._pad(Q, 7, "Q"),
                                    self._rpad(Yi, self.Ns, "Yi"), self._rpad(lam, None, "lambda"),
                                    self._rpad(mu, None, "mu"), self._rpad(D, self.Ns, "D"),
                                    self.stream()), "f3d_sweep_diff")

    def sweep_stage_end(self, KRK, U, Un, rhoi, rhon, Q, rhs, dt):
        Ns = self.Ns
        _rc(self.lib.f3d_sweep_stage_end(self.h, KRK, self._pad(U, 5, "U"), self._pad(Un, 5, "Un"),
                                         self._pad(rhoi, Ns, "rhoi"), self._pad(rhon, Ns, "rhon"),
                                         self._pad(Q, 7, "Q"), self._rhs(rhs), dt, self.stream()),
            "f3d_sweep_stage_end")

    def nn_timing(self, enable=None):
        """Enable (True) / disable (False) event timing of f3d_nn_chem; with
        enable None return the last timed call's [network kernel, gate
        epilogue kernel, scan + compaction, read-back + Newton fallback] in ms."""
        if enable is not None:
            _rc(self.lib.f3d_nn_timing(self.h, int(bool(enable)), None), "f3d_nn_timing")
            return None
        ms = (C.c_double * 4)()
        _rc(self.lib.f3d_nn_timing(self.h, 1, ms), "f3d_nn_timing")
        return list(ms)

    def nan_check(self, U) -> bool:
        flag = C.c_int32(0)
        _rc(self.lib.f3d_nan_check(self.h, self._pad(U, 5, "U"), C.byref(flag), self.stream()),
            "f3d_nan_check")
        return bool(flag.value)

    def workspace(self, name: str):
        """torch view of a context-owned intermediate after rk3_stage."""
        t = self.torch
        ptr = self.lib.f3d_workspace(self.h, name.encode())
        if not ptr:
            raise F3DError(f"no workspace {name!r}")
        Ns = self.Ns
        pad = {"Yi": Ns, "lambda": None, "mu": None, "D": Ns, "phi": None, "Fp": 5, "Fm": 5,
               "Fp_i": Ns, "Fm_i": Ns}
        face = {}
        for d, ax in enumerate("xyz"):
            face[f"F{ax}"] = (d, 5)
            face[f"Fv{ax}"] = (d, 4)
            face[f"F{ax}_i"] = (d, Ns)
            face[f"Fd{ax}"] = (d, Ns)
            face[f"Fh{ax}"] = (d, 3)
        if name in pad:
            nv = pad[name]
            shape = self.shape3 if nv is None else (nv,) + self.shape3
        else:
            d, nv = face[name]
            shape = self.face_shape(d, nv)
        n = int(np.prod(shape))
        esz = 8 if self.rdtype == t.float64 else 4
        return _device_view(ptr, n * esz, shape, self.rdtype)


def _device_view(ptr, nbytes, shape, dtype):
    """Wrap a device pointer as a torch tensor (no copy, no ownership)."""
    import torch

    class _Holder:
        def __init__(self):
            self.__cuda_array_interface__ = {
                "shape": tuple(shape), "typestr": "<f8" if dtype == torch.float64 else "<f4",
                "data": (int(ptr), False), "version": 2, "strides": None}
    return torch.as_tensor(_Holder(), device="cuda")
