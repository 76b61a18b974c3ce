"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the reference it compares against and by the CPU-baseline leg of ``bench.py``.  The product
(``classicaldft_amd``) never imports this module.

The C++ source (oracle/oracle.cpp) restates the reference algorithm; each
function there cites the reference file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

FMT_MODELS = {"Rosenfeld": 0, "RSLT": 1, "WhiteBearI": 2, "WhiteBearII": 3, "esFMT": 4, "none": -1}
POTENTIALS = {"LJ": 0, "tWF": 1, "WRDF": 2}
SPLITS = {"WCA": 0, "BH": 1}
SCHEMES = {"zero": 0, "linearE": 1, "linearF": 2, "quadraticF": 3, "gaussE": 4, "gaussF": 5}

_dp = C.POINTER(C.c_double)
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_long, C.c_long, C.c_long, C.c_double, C.c_int, _dp, C.c_int,
                                 C.POINTER(C.c_int), _dp, C.c_double, C.c_int, C.c_double, C.c_double]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        L.orc_set_serial_pointwise.argtypes = [C.c_void_p, C.c_int]
        L.orc_a_vdw.restype = C.c_double
        L.orc_a_vdw.argtypes = [C.c_void_p, C.c_int]
        for name in ("orc_r2c", "orc_c2r", "orc_weighted_densities"):
            getattr(L, name).argtypes = [C.c_void_p, _dp, _dp]
        for name in ("orc_fmt_weights", "orc_mf_weight_k", "orc_mf_weight_real"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_total.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_fhs_from_fields.restype = C.c_double
        L.orc_fhs_from_fields.argtypes = [C.c_void_p, _dp]
        L.orc_bulk_mu.restype = C.c_double
        L.orc_bulk_mu.argtypes = [C.c_void_p, C.c_double]
        L.orc_phi_dphi.restype = C.c_double
        L.orc_phi_dphi.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_f.restype = C.c_double
        L.orc_f.argtypes = [C.c_int, C.c_int, C.c_double, _dp]
        L.orc_minimize_gp.restype = C.c_int
        L.orc_minimize_gp.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                      C.c_double, _dp, _dp, _dp]
        L.orc_minimize_fixed_mass.restype = C.c_int
        L.orc_minimize_fixed_mass.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                              C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_ddft_run.restype = C.c_int
        L.orc_ddft_run.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_double, _dp, _dp, _dp,
                                   C.POINTER(C.c_ubyte), C.c_double]
        L.orc_ddft_fixed_run.restype = C.c_double
        L.orc_ddft_fixed_run.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, _dp, C.c_double, _dp]
        L.orc_wells_potential.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_smallest_eigenpair.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int, _dp, _dp,
                                             C.c_double, _dp, _dp, _dp]
        L.orc_ddft_split_operator.restype = C.c_double
        L.orc_ddft_split_operator.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_double, _dp]
        L.orc_ddft_crank_nicolson.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_double, _dp, _dp]
        L.orc_picard.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_saddle_search.restype = C.c_int
        L.orc_saddle_search.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double, C.c_ulonglong, _dp, _dp,
                                        _dp]
        L.orc_fire_trace.restype = C.c_int
        L.orc_fire_trace.argtypes = [_dp, C.c_int, _dp, C.c_int, _FIRE_FORCE, C.c_void_p, _dp]
        _lib = L
    return _lib


_FIRE_FORCE = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def fire_trace(x0, force, steps, **fire):
    """Fire::initialize and `steps` Fire::step calls over force(x) -> (energy, forces).  Returns
    (trace, status): trace rows are (x, v, dt, alpha, n_positive, n_negative, energy, rms force) after
    initialisation and after each step; status is the steps taken, or -(k + 1) when step k exceeded the
    uphill limit."""
    f = dict(dt=1e-3, dt_max=1e-2, dt_min=1e-8, alpha_start=0.1, f_inc=1.1, f_dec=0.5, f_alpha=0.99, n_delay=5,
             max_uphill=20, force_tolerance=0.1, max_steps=10000)
    f.update(fire)
    fp = np.array([f[k] for k in ("dt", "dt_max", "dt_min", "alpha_start", "f_inc", "f_dec", "f_alpha", "n_delay",
                                  "max_uphill", "force_tolerance", "max_steps")], dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1)
    n = x0.size

    def cb(_u, xp, fp_out):
        e, g = force(np.ctypeslib.as_array(xp, shape=(n,)).copy())
        np.ctypeslib.as_array(fp_out, shape=(n,))[:] = g
        return float(e)

    fn = _FIRE_FORCE(cb)
    trace = np.zeros((steps + 1, 2 * n + 6))
    st = lib().orc_fire_trace(_p(fp), n, _p(x0), int(steps), fn, None, _p(trace))
    return trace, int(st)


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return int(lib().orc_max_threads())


def phi_dphi(model: str, m: np.ndarray, A: float = 1.0, B: float = 0.0):
    """Phi and dPhi for one measure set (19 doubles: eta,n0,n1,n2,v0[3],v1[3],T[9])."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    dm = np.zeros(19)
    p = lib().orc_phi_dphi(FMT_MODELS[model], A, B, _p(m), _p(dm))
    return p, dm


def f_func(model: str, which: int, eta: float):
    d = C.c_double(0.0)
    v = lib().orc_f(FMT_MODELS[model], which, float(eta), C.byref(d))
    return v, d.value


class Oracle:
    """One grid + species + interactions, mirroring physics::Model + Weights."""

    def __init__(self, shape, dx, diameters=(1.0,), interactions=(), kT=1.0, model="WhiteBearII",
                 A=1.0, B=0.0):
        self.shape = tuple(int(s) for s in shape)
        self.dx = float(dx)
        nx, ny, nz = self.shape
        self.N = nx * ny * nz
        self.M = nx * ny * (nz // 2 + 1)
        self.dv = self.dx ** 3
        self.n_species = len(diameters)
        d = np.asarray(diameters, dtype=np.float64)
        ij = np.zeros(2 * max(1, len(interactions)), dtype=np.int32)
        prm = np.zeros(7 * max(1, len(interactions)))
        for k, it in enumerate(interactions):
            ij[2 * k] = it.get("si", 0)
            ij[2 * k + 1] = it.get("sj", 0)
            prm[7 * k:7 * k + 7] = [POTENTIALS[it.get("potential", "LJ")], it.get("sigma", 1.0),
                                    it.get("epsilon", 1.0), it.get("r_cutoff", -1.0), it.get("alpha", 50.0),
                                    SPLITS[it.get("split", "WCA")], SCHEMES[it.get("scheme", "quadraticF")]]
        self._h = lib().orc_create(nx, ny, nz, self.dx, self.n_species, _p(d), len(interactions),
                                   ij.ctypes.data_as(C.POINTER(C.c_int)), _p(prm), float(kT),
                                   FMT_MODELS[model], float(A), float(B))
        self.n_inter = len(interactions)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    def set_serial_pointwise(self, flag: bool) -> None:
        lib().orc_set_serial_pointwise(self._h, int(bool(flag)))

    def a_vdw(self, k=0) -> float:
        return lib().orc_a_vdw(self._h, k)

    def r2c(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        out = np.zeros(self.M, dtype=np.complex128)
        lib().orc_r2c(self._h, _p(x), out.ctypes.data_as(_dp))
        return out.reshape(self.shape[0], self.shape[1], -1)

    def c2r(self, X):
        X = np.ascontiguousarray(X, dtype=np.complex128).reshape(-1)
        out = np.zeros(self.N)
        lib().orc_c2r(self._h, X.ctypes.data_as(_dp), _p(out))
        return out.reshape(self.shape)

    def fmt_weights(self, s=0):
        out = np.zeros(11 * self.M, dtype=np.complex128)
        lib().orc_fmt_weights(self._h, s, out.ctypes.data_as(_dp))
        return out.reshape(11, self.shape[0], self.shape[1], -1)

    def mf_weight_k(self, k=0):
        out = np.zeros(self.M, dtype=np.complex128)
        lib().orc_mf_weight_k(self._h, k, out.ctypes.data_as(_dp))
        return out.reshape(self.shape[0], self.shape[1], -1)

    def mf_weight_real(self, k=0):
        out = np.zeros(self.N)
        lib().orc_mf_weight_real(self._h, k, _p(out))
        return out.reshape(self.shape)

    def weighted_densities(self, rho):
        """12 fields: eta, n2, v2[x,y,z], T[xx,xy,xz,yy,yz,zz], mean-field convolution."""
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        out = np.zeros(12 * self.N)
        lib().orc_weighted_densities(self._h, _p(rho), _p(out))
        return out.reshape((12,) + self.shape)

    def total(self, rho, mu, vext=None):
        """Return (parts, forces); parts = [F_id, F_ext, F_hs, F_mf, F, Omega]."""
        S = self.n_species
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        assert rho.size == S * self.N
        v = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (S,)))
        parts = np.zeros(6)
        forces = np.zeros(S * self.N)
        lib().orc_total(self._h, _p(rho), _p(v), _p(mu), _p(parts), _p(forces))
        return parts, forces.reshape((S,) + self.shape) if S > 1 else forces.reshape(self.shape)

    def fhs_from_fields(self, wd):
        """F_hs from the 11 FMT weighted-density fields (channel order of weighted_densities)."""
        wd = np.ascontiguousarray(wd, dtype=np.float64).reshape(-1)
        assert wd.size >= 11 * self.N
        return lib().orc_fhs_from_fields(self._h, _p(wd))

    def bulk_mu(self, rho):
        return lib().orc_bulk_mu(self._h, float(rho))

    def minimize_grand_potential(self, rho0, mu, fire=None, param="unbounded", rho_min=1e-18, rho_max=1.0,
                                 homogeneous_boundary=True, vext=None, max_steps=None):
        return self._minimize(rho0, mu, -1.0, fire, param, rho_min, rho_max, homogeneous_boundary, vext, max_steps)

    def minimize_fixed_mass(self, rho0, mu, target_mass, fire=None, param="unbounded", rho_min=1e-18, rho_max=1.0,
                            homogeneous_boundary=True, vext=None, max_steps=None):
        """Minimizer::fixed_mass; the energies are F (mu = 0 inside the loop)."""
        return self._minimize(rho0, mu, float(target_mass), fire, param, rho_min, rho_max, homogeneous_boundary,
                              vext, max_steps)

    def _minimize(self, rho0, mu, target, fire, param, rho_min, rho_max, homogeneous_boundary, vext, max_steps):
        f = dict(dt=1e-3, dt_max=1e-2, dt_min=1e-8, alpha_start=0.1, f_inc=1.1, f_dec=0.5, f_alpha=0.99,
                 n_delay=5, max_uphill=20, force_tolerance=0.1, max_steps=10000, force_max_tolerance=0.0)
        if fire:
            f.update(fire)
        if max_steps is not None:
            f["max_steps"] = max_steps
        fp = np.array([f["dt"], f["dt_max"], f["dt_min"], f["alpha_start"], f["f_inc"], f["f_dec"], f["f_alpha"],
                       f["n_delay"], f["max_uphill"], f["force_tolerance"], f["max_steps"],
                       f["force_max_tolerance"]], dtype=np.float64)
        rho0 = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1)
        v = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        rho_out = np.zeros(self.N)
        out = np.zeros(7)
        energies = np.full(int(f["max_steps"]) + 1, np.nan)
        if target > 0.0:
            rc = lib().orc_minimize_fixed_mass(self._h, _p(fp), 0 if param == "unbounded" else 1, rho_min, rho_max,
                                               int(bool(homogeneous_boundary)), _p(rho0), _p(v), float(mu),
                                               float(target), _p(rho_out), _p(out), _p(energies))
        else:
            rc = lib().orc_minimize_gp(self._h, _p(fp), 0 if param == "unbounded" else 1, rho_min, rho_max,
                                       int(bool(homogeneous_boundary)), _p(rho0), _p(v), float(mu), _p(rho_out),
                                       _p(out), _p(energies))
        if rc != 0:
            raise RuntimeError("fire::step: exceeded max uphill steps")
        res = dict(free_energy=out[0], grand_potential=out[1], rms_force=out[2], iterations=int(out[3]),
                   converged=bool(out[4]), dt=out[5], alpha=out[6])
        return rho_out.reshape(self.shape), res, energies[: res["iterations"] + 1]

    def ddft_run(self, rho0, mu, n_steps, dt=1e-4, dt_max=1.0, fp_tolerance=1e-4, fp_max_iterations=100,
                 min_density=1e-18, vext=None, reservoir_mask=None, reservoir_density=0.0):
        """Simulation::run; reservoir_mask (bool array) applies reservoir_boundary(mask, reservoir_density).
        Raises RuntimeError when a step exceeds the restart limit or its iterate is not finite."""
        cfg = np.array([dt, 1.0, min_density, dt_max, fp_tolerance, fp_max_iterations], dtype=np.float64)
        rho = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1).copy()
        v = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        out = np.zeros(6)
        energies = np.zeros(n_steps + 1)
        mask = None if reservoir_mask is None else np.ascontiguousarray(reservoir_mask, dtype=np.uint8).reshape(-1)
        rc = lib().orc_ddft_run(self._h, _p(cfg), int(n_steps), _p(v), float(mu), _p(rho), _p(out), _p(energies),
                                None if mask is None else mask.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                float(reservoir_density))
        if rc != 0:
            raise RuntimeError("integrating_factor_step: no convergence within the restart limit, or a "
                               "non-finite iterate")
        res = dict(mass_initial=out[0], mass_final=out[1], dt=out[2], time=out[3], energy=out[4],
                   evals=int(out[5]))
        return rho.reshape(self.shape), res, energies

    def ddft_fixed_run(self, rho0, mu, n_steps, dt, fp_iterations, vext=None):
        """Fixed-dt DDFT map (exactly fp_iterations iterations per step); returns (rho, min density)."""
        rho = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1).copy()
        v = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        mn = lib().orc_ddft_fixed_run(self._h, float(dt), int(fp_iterations), int(n_steps), _p(v), float(mu),
                                      _p(rho))
        return rho.reshape(self.shape), mn

    def smallest_eigenpair(self, rho, mu, v0, tolerance=1e-4, max_iterations=300, hessian_eps=1e-6,
                           fixed_boundary=False, vext=None):
        """EigenvalueSolver::solve with forward-difference products; returns (eigenvector, eigenvalue, iterations,
        converged)."""
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1)
        ve = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        vo = np.zeros(self.N)
        out = np.zeros(3)
        lib().orc_smallest_eigenpair(self._h, float(tolerance), int(max_iterations), float(hessian_eps),
                                     int(bool(fixed_boundary)), _p(r), _p(ve), float(mu), _p(v), _p(vo), _p(out))
        return vo.reshape(self.shape), out[0], int(out[1]), bool(out[2])

    def ddft_split_operator(self, rho0, mu, n_steps, dt=1e-4, diffusion_coefficient=1.0, min_density=1e-18,
                            vext=None):
        rho = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1).copy()
        ve = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        cfg = np.array([dt, diffusion_coefficient, min_density])
        e = lib().orc_ddft_split_operator(self._h, _p(cfg), int(n_steps), _p(ve), float(mu), _p(rho))
        return rho.reshape(self.shape), e

    def ddft_crank_nicolson(self, rho0, mu, n_steps, dt=1e-4, diffusion_coefficient=1.0, min_density=1e-18,
                            fp_tolerance=1e-4, fp_max_iterations=100, vext=None):
        rho = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1).copy()
        ve = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        cfg = np.array([dt, diffusion_coefficient, min_density, fp_tolerance, fp_max_iterations], dtype=np.float64)
        out = np.zeros(2)
        lib().orc_ddft_crank_nicolson(self._h, _p(cfg), int(n_steps), _p(ve), float(mu), _p(rho), _p(out))
        return rho.reshape(self.shape), out[0], int(out[1])

    def picard(self, rho0, mu, mixing=0.01, min_density=1e-30, tolerance=1e-6, max_iterations=5000,
               target_mass=None, background=None, vext=None):
        rho = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1).copy()
        ve = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        bg = None if background is None else np.ascontiguousarray(background, dtype=np.float64).reshape(-1)
        cfg = np.array([mixing, min_density, tolerance, max_iterations], dtype=np.float64)
        out = np.zeros(4)
        tgt = None if target_mass is None else np.array([float(target_mass)])
        lib().orc_picard(self._h, _p(cfg), _p(ve), float(mu), _p(tgt), _p(bg), _p(rho), _p(out))
        return rho.reshape(self.shape), dict(grand_potential=out[0], residual=out[1], iterations=int(out[2]),
                                             converged=bool(out[3]))

    def saddle_search(self, rho0, mu, target_mass, seed=0, fire=None, eigen=None, param="unbounded", rho_min=1e-18,
                      rho_max=1.0, omega_tolerance=1e-5, tolerance_start=1e-3, tolerance_factor=10.0,
                      max_outer_iterations=20, eigenvalue_clear_count=2, vext=None):
        """saddle_point::Search::find; returns (rho, eigenvector, result dict)."""
        f = dict(dt=1e-3, dt_max=1.0, dt_min=1e-8, alpha_start=0.01, f_inc=1.1, f_dec=0.5, f_alpha=1.0, n_delay=5,
                 max_uphill=20, force_tolerance=1e-4, max_steps=100000)
        f.update(fire or {})
        e = dict(tolerance=1e-4, max_iterations=300, hessian_eps=1e-6)
        e.update(eigen or {})
        cfg = np.array([f["dt"], f["dt_max"], f["dt_min"], f["alpha_start"], f["f_inc"], f["f_dec"], f["f_alpha"],
                        f["n_delay"], f["max_uphill"], f["force_tolerance"], f["max_steps"], e["tolerance"],
                        e["max_iterations"], e["hessian_eps"], 0 if param == "unbounded" else 1, rho_min, rho_max,
                        omega_tolerance, tolerance_start, tolerance_factor, max_outer_iterations,
                        eigenvalue_clear_count], dtype=np.float64)
        r0 = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1)
        ve = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
        ro, eo, out = np.zeros(self.N), np.zeros(self.N), np.zeros(4)
        rc = lib().orc_saddle_search(self._h, _p(cfg), _p(r0), _p(ve), float(mu), float(target_mass), int(seed),
                                     _p(ro), _p(eo), _p(out))
        if rc != 0:
            raise RuntimeError("fire::step: exceeded max uphill steps")
        return ro.reshape(self.shape), eo.reshape(self.shape), dict(
            grand_potential=out[0], eigenvalue=out[1], outer_iterations=int(out[2]), converged=bool(out[3]))

    def wells_potential(self, params):
        prm = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 5)
        out = np.zeros(self.N)
        lib().orc_wells_potential(self._h, prm.shape[0], _p(prm.reshape(-1)), _p(out))
        return out.reshape(self.shape)
