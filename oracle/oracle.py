"""ctypes binding of the FP64 CPU oracle (oracle/f3d_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, as the checker.  Arrays are numpy float64 in
Fortran order with shape (Nx+2NG, Ny+2NG, Nz+2NG[, NV]) - the reference's
column-major [i, j, k, var] layout - and face arrays (Nx+1, Ny, Nz, NV) etc.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class _Dims(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("Nx", "Ny", "Nz", "NG", "Ns", "Nr")]


class _Thermo(C.Structure):
    _fields_ = [("Ru", C.c_double), ("min_temp", C.c_double), ("max_temp", C.c_double),
                ("mw", _dp), ("coeffs_sep", _dp), ("coeffs_lo", _dp), ("coeffs_hi", _dp),
                ("visc_poly", _dp), ("conduct_poly", _dp), ("binarydiff_poly", _dp)]


class _Kin(C.Structure):
    _fields_ = [("atm", C.c_double), ("reaction_type", _i64p), ("sgm", _i64p),
                ("vf", _i64p), ("vr", _i64p), ("Arr", _dp), ("ef", _dp), ("loP", _dp),
                ("Troe", _dp)]


class _Scheme(C.Structure):
    _fields_ = [("SWeps2", C.c_double), ("WENOeps", C.c_double), ("phi1", C.c_double),
                ("phi2", C.c_double), ("UP7", C.c_double * 7), ("T_criteria", C.c_double),
                ("sensor_kind", C.c_int32), ("recon_kind", C.c_int32)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


def _f(a):
    a = np.asarray(a)
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1), (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(a):
    a = np.asarray(a)
    assert a.dtype == np.int64
    return a.ctypes.data_as(_i64p)


# reference scheme constants, run.jl:51-57 and :12
UP7 = (0.0, 17 / 600, -23 / 120, 22 / 30, 0.5, -0.075, 0.005)
DEFAULT_SCHEME = dict(SWeps2=0.0, WENOeps=1e-26, phi1=1e-3, phi2=0.1, UP7=UP7,
                      T_criteria=500.0)


class Oracle:
    def __init__(self, Nx, Ny, Nz, mech, scheme=None, NG=4):
        build()
        self.lib = C.CDLL(LIB_PATH)
        self.Nx, self.Ny, self.Nz, self.NG = Nx, Ny, Nz, NG
        self.mech = mech
        self.Ns, self.Nr = mech.Ns, mech.Nr
        self.shape = (Nx + 2 * NG, Ny + 2 * NG, Nz + 2 * NG)
        self.dims = _Dims(Nx, Ny, Nz, NG, mech.Ns, mech.Nr)
        m = mech
        self._keep = [np.asfortranarray(x, dtype=np.float64) for x in (
            m.mw, m.coeffs_sep, m.coeffs_lo, m.coeffs_hi, m.visc_poly, m.conduct_poly,
            m.binarydiff_poly)]
        k = self._keep
        self.thermo = _Thermo(m.Ru, m.min_temp, m.max_temp, *[_f(x) for x in k])
        self._keep_k = [np.asfortranarray(x, dtype=np.int64) for x in (
            m.reaction_type if m.Nr else np.zeros(1, np.int64),
            m.sgm if m.Nr else np.zeros(1, np.int64),
            m.vf if m.Nr else np.zeros(1, np.int64),
            m.vr if m.Nr else np.zeros(1, np.int64))]
        self._keep_kf = [np.asfortranarray(x if x.size else np.zeros(1), dtype=np.float64)
                         for x in (m.Arr, m.ef, m.loP, m.Troe)]
        self.kin = _Kin(m.atm, *[_i(x) for x in self._keep_k], *[_f(x) for x in self._keep_kf])
        s = dict(DEFAULT_SCHEME)
        if scheme:
            s.update(scheme)
        self.scheme_dict = s
        self.scheme = _Scheme(s["SWeps2"], s["WENOeps"], s["phi1"], s["phi2"],
                              (C.c_double * 7)(*s["UP7"]), s["T_criteria"],
                              int(s.get("sensor_kind", 0)), int(s.get("recon_kind", 0)))
        L = self.lib
        L.or_get_T.restype = C.c_double
        L.or_exp.restype = C.c_double
        L.or_exp.argtypes = [C.c_double]
        L.or_log.restype = C.c_double
        L.or_log.argtypes = [C.c_double]
        L.or_get_T_from.restype = C.c_double
        L.or_weno7z.restype = C.c_double
        L.or_internal_energy.restype = C.c_double
        L.or_cv.restype = C.c_double
        L.or_num_threads.restype = C.c_int32

    # ------------------------------------------------------------------
    def zeros(self, nv=None):
        sh = self.shape if nv is None else self.shape + (nv,)
        return np.zeros(sh, order="F")

    def face_zeros(self, d, nv):
        sh = [self.Nx, self.Ny, self.Nz]
        sh[d] += 1
        return np.zeros(tuple(sh) + (nv,), order="F")

    def _tag(self, tag):
        return None if tag is None else _i(tag)

    @staticmethod
    def _mptr(metrics):
        arr = (_dp * 9)(*[_f(m) for m in metrics])
        return arr

    # ---- pointwise ---------------------------------------------------
    def internal_energy(self, T, rhoi):
        return self.lib.or_internal_energy(C.byref(self.dims), C.byref(self.thermo),
                                           C.c_double(T), _f(np.ascontiguousarray(rhoi, dtype=np.float64)))

    def cv(self, T, rhoi):
        return self.lib.or_cv(C.byref(self.dims), C.byref(self.thermo), C.c_double(T),
                              _f(np.ascontiguousarray(rhoi, dtype=np.float64)))

    def get_T_from(self, e, rhoi, T_guess, tol=1e-3, maxiter=30):
        r = np.ascontiguousarray(rhoi, dtype=np.float64)
        return self.lib.or_get_T_from(C.byref(self.dims), C.byref(self.thermo), C.c_double(e),
                                      _f(r), C.c_double(T_guess), C.c_double(tol),
                                      C.c_int32(maxiter), C.c_double(self.mech.min_temp),
                                      C.c_double(self.mech.max_temp))

    def get_T(self, e, rhoi, tol=1e-3, maxiter=30, tmin=None, tmax=None):
        it = C.c_int32(0)
        r = np.ascontiguousarray(rhoi, dtype=np.float64)
        T = self.lib.or_get_T(C.byref(self.dims), C.byref(self.thermo), C.c_double(e), _f(r),
                              C.c_double(tol), C.c_int32(maxiter),
                              C.c_double(self.mech.min_temp if tmin is None else tmin),
                              C.c_double(self.mech.max_temp if tmax is None else tmax),
                              C.byref(it))
        return T, it.value

    def h_specs(self, T):
        h = np.zeros(self.Ns)
        self.lib.or_h_specs(C.byref(self.dims), C.byref(self.thermo), C.c_double(T), _f(h))
        return h

    def gibbs(self, T):
        g = np.zeros(self.Ns)
        self.lib.or_gibbs(C.byref(self.dims), C.byref(self.thermo), C.c_double(T), _f(g))
        return g

    def production_rate(self, sc, T, jac_eps=2.220446049250313e-16, jac=True, full_jac=False):
        sc = np.ascontiguousarray(sc, dtype=np.float64)
        w = np.zeros(self.Ns)
        A = np.zeros((self.Ns, self.Ns), order="F") if jac else None
        self.lib.or_production_rate(C.byref(self.dims), C.byref(self.thermo), C.byref(self.kin),
                                    _f(sc), C.c_double(T), C.c_double(jac_eps),
                                    C.c_int32(1 if full_jac else 0), _f(w),
                                    _f(A) if jac else None)
        return (w, A) if jac else w

    def production_rate_dT(self, sc, T):
        """(wdot, d wdot / dT at fixed concentrations) from the analytic derivative."""
        sc = np.ascontiguousarray(sc, dtype=np.float64)
        w = np.zeros(self.Ns)
        dw = np.zeros(self.Ns)
        self.lib.or_production_rate_dT(C.byref(self.dims), C.byref(self.thermo), C.byref(self.kin),
                                       _f(sc), C.c_double(T), _f(w), _f(dw))
        return w, dw

    def gauss_solve(self, A, b):
        A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
        b = np.array(b, dtype=np.float64)
        x = np.zeros_like(b)
        self.lib.or_gauss_solve(C.c_int32(len(b)), _f(A), _f(b), _f(x))
        return x

    def mixture_properties(self, T, P, X, reference_form=False):
        """(lambda, mu, D[Ns]) at one state; reference_form takes the
        mixture-averaged D_k with the reference's literal 1 - X_k numerator."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        D = np.zeros(self.Ns)
        lam, mu = C.c_double(0), C.c_double(0)
        self.lib.or_set_diffusivity_reference_form(C.c_int32(1 if reference_form else 0))
        try:
            self.lib.or_mixture_properties(C.byref(self.dims), C.byref(self.thermo),
                                           C.c_double(T), C.c_double(P), _f(X), _f(D),
                                           C.byref(lam), C.byref(mu))
        finally:
            self.lib.or_set_diffusivity_reference_form(C.c_int32(0))
        return lam.value, mu.value, D

    # ---- operation counts (the chemistry templates run on a counting type)
    _OPS = ("add", "mul", "div", "trans", "sqrt", "iterations")

    def _ops(self, out):
        return dict(zip(self._OPS, (float(x) for x in out)))

    def newton_opcount(self, rhoi, T, h, maxiter=50, rtol=1e-10, ttol=1e-6):
        """Operations of one backward-Euler Newton solve of one reactor."""
        out = np.zeros(6)
        r = np.ascontiguousarray(rhoi, dtype=np.float64)
        self.lib.or_newton_opcount(C.byref(self.dims), C.byref(self.thermo), C.byref(self.kin),
                                   _f(r), C.c_double(T), C.c_double(h), C.c_int32(maxiter),
                                   C.c_double(rtol), C.c_double(ttol), _f(out))
        return self._ops(out)

    def rates_opcount(self, rhoi, T, full_jac=False):
        out = np.zeros(6)
        r = np.ascontiguousarray(rhoi, dtype=np.float64)
        self.lib.or_rates_opcount(C.byref(self.dims), C.byref(self.thermo), C.byref(self.kin),
                                  _f(r), C.c_double(T), C.c_int32(int(full_jac)), _f(out))
        return self._ops(out)

    def getT_update_opcount(self, rhoi, T, ein):
        out = np.zeros(6)
        r = np.ascontiguousarray(rhoi, dtype=np.float64)
        self.lib.or_getT_update_opcount(C.byref(self.dims), C.byref(self.thermo), _f(r),
                                        C.c_double(T), C.c_double(ein), _f(out))
        return self._ops(out)

    def gauss_opcount(self, A, b):
        out = np.zeros(6)
        A = np.asfortranarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        self.lib.or_gauss_opcount(C.c_int32(len(b)), _f(A), _f(b), _f(out))
        return self._ops(out)

    def reactor0d_step(self, rhoi, E, dt, jac_eps, gt_tol, gt_maxiter, gt_tmin, gt_tmax):
        Eb = C.c_double(E)
        T = C.c_double(0)
        it = C.c_int32(0)
        self.lib.or_reactor0d_step(C.byref(self.dims), C.byref(self.thermo), C.byref(self.kin),
                                   _f(rhoi), C.byref(Eb), C.c_double(dt), C.c_double(jac_eps),
                                   C.c_double(gt_tol), C.c_int32(gt_maxiter),
                                   C.c_double(gt_tmin), C.c_double(gt_tmax), C.byref(T),
                                   C.byref(it))
        return Eb.value, T.value, it.value

    # ---- kernels -------------------------------------------------------
    def flux_split(self, Q, Ax, Ay, Az, tag=None):
        Fp, Fm = self.zeros(5), self.zeros(5)
        self.lib.or_flux_split(C.byref(self.dims), _f(Q), _f(Fp), _f(Fm), _f(Ax), _f(Ay),
                               _f(Az), self._tag(tag), C.byref(self.scheme))
        return Fp, Fm

    def species_split(self, rhoi, Q, Ax, Ay, Az, tag=None):
        Fp, Fm = self.zeros(self.Ns), self.zeros(self.Ns)
        self.lib.or_species_split(C.byref(self.dims), _f(rhoi), _f(Q), _f(Fp), _f(Fm),
                                  _f(Ax), _f(Ay), _f(Az), self._tag(tag))
        return Fp, Fm

    def shock_sensor(self, Q, phi=None):
        phi = self.zeros() if phi is None else phi
        self.lib.or_shock_sensor(C.byref(self.dims), _f(phi), _f(Q))
        return phi

    def weno7z(self, v7, eps=1e-26):
        v = np.ascontiguousarray(v7, dtype=np.float64)
        return self.lib.or_weno7z(_f(v), C.c_double(eps))

    def ducros_sensor(self, Q, metrics, J, phi=None):
        phi = self.zeros() if phi is None else phi
        self.lib.or_ducros_sensor(C.byref(self.dims), _f(phi), _f(Q), self._mptr(metrics), _f(J))
        return phi

    def weno(self, d, phi, Fp, Fm, NV, tag=None):
        F = self.face_zeros(d, NV)
        self.lib.or_weno(C.byref(self.dims), C.c_int32(d), _f(F), _f(phi), _f(Fp), _f(Fm),
                         C.c_int32(NV), self._tag(tag), C.byref(self.scheme))
        return F

    def viscous_flux(self, d, Q, metrics, J, lam, mu, Fh, tag=None):
        Fv = self.face_zeros(d, 4)
        self.lib.or_viscous_flux(C.byref(self.dims), C.c_int32(d), _f(Fv), _f(Q),
                                 self._mptr(metrics), _f(J), _f(lam), _f(mu), _f(Fh),
                                 self._tag(tag))
        return Fv

    def spec_viscous_flux(self, d, Q, Yi, metrics, J, D, tag=None):
        Fd = self.face_zeros(d, self.Ns)
        Fh = self.face_zeros(d, 3)
        self.lib.or_spec_viscous_flux(C.byref(self.dims), C.c_int32(d), _f(Fd), _f(Q), _f(Yi),
                                      self._mptr(metrics), _f(J), _f(D), _f(Fh),
                                      C.byref(self.thermo), self._tag(tag))
        return Fd, Fh

    def div(self, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J, tag=None):
        self.lib.or_div(C.byref(self.dims), _f(U), _f(Fx), _f(Fy), _f(Fz), _f(Fvx), _f(Fvy),
                        _f(Fvz), C.c_double(dt), _f(J), self._tag(tag))

    def div_specs(self, rhoi, Fx, Fy, Fz, Fdx, Fdy, Fdz, dt, J, tag=None):
        self.lib.or_div_specs(C.byref(self.dims), _f(rhoi), _f(Fx), _f(Fy), _f(Fz), _f(Fdx),
                              _f(Fdy), _f(Fdz), C.c_double(dt), _f(J), self._tag(tag))

    def mixture(self, Q, rhoi, tag=None):
        Yi, D = self.zeros(self.Ns), self.zeros(self.Ns)
        lam, mu = self.zeros(), self.zeros()
        self.lib.or_mixture(C.byref(self.dims), _f(Q), _f(rhoi), _f(Yi), _f(lam), _f(mu), _f(D),
                            C.byref(self.thermo), self._tag(tag))
        return Yi, lam, mu, D

    def c2prim(self, U, Q, rhoi, tag=None):
        self.lib.or_c2prim(C.byref(self.dims), _f(U), _f(Q), _f(rhoi), C.byref(self.thermo),
                           self._tag(tag))

    def prim2c(self, Q):
        U = self.zeros(5)
        self.lib.or_prim2c(C.byref(self.dims), _f(U), _f(Q))
        return U

    def lincomb(self, U, Un, NV, a, b, tag=None):
        self.lib.or_lincomb(C.byref(self.dims), _f(U), _f(Un), C.c_int32(NV), C.c_double(a),
                            C.c_double(b), self._tag(tag))

    def get_y(self, rhoi, Q, tag=None):
        Yi = self.zeros(self.Ns)
        self.lib.or_get_y(C.byref(self.dims), _f(Yi), _f(rhoi), _f(Q), self._tag(tag))
        return Yi

    def fill(self, kinds, Q, U, rhoi, inflow_Q=None, inflow_rhoi=None):
        k = np.asarray(kinds, dtype=np.int32)
        iq = None if inflow_Q is None else _f(np.asarray(inflow_Q, dtype=np.float64))
        ir = None if inflow_rhoi is None else _f(np.asarray(inflow_rhoi, dtype=np.float64))
        self.lib.or_fill(C.byref(self.dims), k.ctypes.data_as(_i32p), _f(Q), _f(U), _f(rhoi),
                         iq, ir)

    def chem_stiff(self, U, Q, rhoi, dt, tag=None, bug_compat=0):
        self.lib.or_chem_stiff(C.byref(self.dims), _f(U), _f(Q), _f(rhoi), C.c_double(dt),
                               C.byref(self.thermo), C.byref(self.kin), C.byref(self.scheme),
                               self._tag(tag), C.c_int32(bug_compat))

    def chem_explicit(self, U, Q, rhoi, dt, tag=None):
        self.lib.or_chem_explicit(C.byref(self.dims), _f(U), _f(Q), _f(rhoi), C.c_double(dt),
                                  C.byref(self.thermo), C.byref(self.kin), C.byref(self.scheme),
                                  self._tag(tag))

    def chem_newton(self, U, Q, rhoi, dt, maxiter=50, rtol=1e-10, ttol=1e-6, tag=None):
        iters = np.zeros((self.Nx, self.Ny, self.Nz), dtype=np.int32, order="F")
        self.lib.or_chem_newton(C.byref(self.dims), _f(U), _f(Q), _f(rhoi), C.c_double(dt),
                                C.byref(self.thermo), C.byref(self.kin), C.byref(self.scheme),
                                self._tag(tag), C.c_int32(maxiter), C.c_double(rtol),
                                C.c_double(ttol), iters.ctypes.data_as(_i32p))
        return iters

    # ---- stage ---------------------------------------------------------
    def workspace(self):
        Ns = self.Ns
        ws = [self.zeros(Ns), self.zeros(), self.zeros(), self.zeros(Ns), self.zeros(),
              self.zeros(5), self.zeros(5), self.zeros(Ns), self.zeros(Ns)]
        ws += [self.face_zeros(d, 5) for d in range(3)]
        ws += [self.face_zeros(d, 4) for d in range(3)]
        ws += [self.face_zeros(d, Ns) for d in range(3)]
        ws += [self.face_zeros(d, Ns) for d in range(3)]
        ws += [self.face_zeros(d, 3) for d in range(3)]
        return ws

    def rk3_stage(self, KRK, U, Un, rhoi, rhon, Q, ws, metrics, J, dt, fill_kinds,
                  inflow_Q=None, inflow_rhoi=None, tag=None):
        wsp = (_dp * 24)(*[_f(w) for w in ws])
        k = np.asarray(fill_kinds, dtype=np.int32)
        iq = None if inflow_Q is None else _f(np.asarray(inflow_Q, dtype=np.float64))
        ir = None if inflow_rhoi is None else _f(np.asarray(inflow_rhoi, dtype=np.float64))
        self.lib.or_rk3_stage(C.byref(self.dims), C.c_int32(KRK), _f(U), _f(Un), _f(rhoi),
                              _f(rhon), _f(Q), wsp, self._mptr(metrics), _f(J), C.c_double(dt),
                              self._tag(tag), C.byref(self.thermo), C.byref(self.scheme),
                              k.ctypes.data_as(_i32p), iq, ir)

    def num_threads(self):
        return self.lib.or_num_threads()
