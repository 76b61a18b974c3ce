"""MI355X (gfx950) free-energy + gradient path of classicalDFT.

Python binding of the C ABI in include/cdft.h (classicaldft_amd/libcdft.so).
The HIP library is the only compute path: if it is not built, every entry
point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# The in-tree build; CDFT_LIBRARY names another build of the same sources.
LIB_PATH = os.environ.get("CDFT_LIBRARY") or os.path.join(HERE, "libcdft.so")

FMT_MODELS = {"none": -1, "Rosenfeld": 0, "RSLT": 1, "WhiteBearI": 2, "WhiteBearII": 3, "esFMT": 4}
POTENTIALS = {"LJ": 0, "tWF": 1, "WRDF": 2}
SPLITS = {"WCA": 0, "BH": 1}
SCHEMES = {"zero": 0, "linearE": 1, "linearF": 2, "quadraticF": 3, "gaussE": 4, "gaussF": 5}
PRECISIONS = {"fp64": 0, "mixed": 1}
SCHEDULES = {"auto": 0, "general": 1}

EXPORTED = [
    "cdft_ctx_create", "cdft_ctx_destroy", "cdft_last_error", "cdft_grid_shape", "cdft_a_vdw",
    "cdft_bulk_chemical_potential", "cdft_total", "cdft_total_device", "cdft_total_device_repeat",
    "cdft_species_count", "cdft_total_species", "cdft_total_species_device", "cdft_time_total_species_device",
    "cdft_bulk_free_energy_density", "cdft_bulk_chemical_potentials", "cdft_bulk_pressure", "cdft_contributions",
    "cdft_bulk_batch",
    "cdft_weighted_densities", "cdft_weighted_densities_noalloc", "cdft_fft_r2c", "cdft_fft_c2r", "cdft_minimize_grand_potential", "cdft_ddft_run",
    "cdft_time_total_device", "cdft_profile_stages", "cdft_stage_name", "cdft_schedule",
    "cdft_minimize_grand_potential_batch", "cdft_hessian_vector", "cdft_wells_potential", "cdft_ddft_fixed_run",
    "cdft_ddft_adjoint_wells", "cdft_autotune", "cdft_minimize_fixed_mass", "cdft_smallest_eigenpair",
    "cdft_saddle_search", "cdft_ddft_split_operator", "cdft_ddft_crank_nicolson", "cdft_picard_solve",
]


class CdftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cdft error {code}: {msg}")
        self.code = code


class InvalidArgument(CdftError, ValueError):
    pass


class _Grid(C.Structure):
    _fields_ = [("dx", C.c_double), ("box", C.c_double * 3)]


class _Interaction(C.Structure):
    _fields_ = [("species_i", C.c_int), ("species_j", C.c_int), ("potential", C.c_int), ("sigma", C.c_double),
                ("epsilon", C.c_double), ("r_cutoff", C.c_double), ("alpha", C.c_double), ("split", C.c_int),
                ("weight_scheme", C.c_int)]


class _Options(C.Structure):
    _fields_ = [("fmt_model", C.c_int), ("esfmt_A", C.c_double), ("esfmt_B", C.c_double), ("precision", C.c_int),
                ("device", C.c_int), ("schedule", C.c_int)]


class _Result(C.Structure):
    _fields_ = [("free_energy", C.c_double), ("grand_potential", C.c_double), ("ideal", C.c_double),
                ("external", C.c_double), ("hard_sphere", C.c_double), ("mean_field", C.c_double),
                ("mass", C.c_double)]


class _Fire(C.Structure):
    _fields_ = [("dt", C.c_double), ("dt_max", C.c_double), ("dt_min", C.c_double), ("alpha_start", C.c_double),
                ("f_inc", C.c_double), ("f_dec", C.c_double), ("f_alpha", C.c_double), ("n_delay", C.c_int),
                ("max_uphill", C.c_int), ("force_tolerance", C.c_double), ("max_steps", C.c_int),
                ("force_max_tolerance", C.c_double)]


class _Param(C.Structure):
    _fields_ = [("kind", C.c_int), ("rho_min", C.c_double), ("rho_max", C.c_double)]


class _MinResult(C.Structure):
    _fields_ = [("free_energy", C.c_double), ("grand_potential", C.c_double), ("rms_force", C.c_double),
                ("iterations", C.c_int), ("converged", C.c_int), ("dt", C.c_double), ("alpha", C.c_double)]


class _StepConfig(C.Structure):
    _fields_ = [("dt", C.c_double), ("diffusion_coefficient", C.c_double), ("min_density", C.c_double),
                ("dt_max", C.c_double), ("fp_tolerance", C.c_double), ("fp_max_iterations", C.c_int)]


class _SimResult(C.Structure):
    _fields_ = [("mass_initial", C.c_double), ("mass_final", C.c_double), ("time", C.c_double),
                ("energy", C.c_double), ("dt", C.c_double), ("evaluations", C.c_long), ("steps", C.c_int)]


_FORCE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))
_PROJECT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double))
_STOP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int))
_SNAPSHOT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double))


class _Hooks(C.Structure):
    _fields_ = [("user", C.c_void_p), ("force", _FORCE_FN), ("project", _PROJECT_FN), ("stop", _STOP_FN),
                ("snapshot", _SNAPSHOT_FN), ("snapshot_interval", C.c_int)]


class _HostCallbacks:
    """Python callables behind a cdft_hooks (the reference's ForceCallback, boundary and stop conditions,
    snapshots and Picard constraints).  An exception inside a callback is kept and re-raised after the
    native call has returned its error status."""

    def __init__(self, shape, force=None, project=None, stop=None, snapshot=None, snapshot_interval=0):
        self.shape, self.n = tuple(shape), int(np.prod(shape))
        self.error = None
        self.fns = {}

        def guard(fn):
            def wrapped(*a):
                try:
                    fn(*a)
                    return 0
                except BaseException as e:  # noqa: BLE001 - handed back to the caller after the native call
                    self.error = e
                    return 1
            return wrapped

        def arr(p):
            return np.ctypeslib.as_array(p, shape=(self.n,))

        if force is not None:
            def _force(_u, rho, e, f):
                en, fr = force(arr(rho).reshape(self.shape).copy())
                fr = np.asarray(fr, dtype=np.float64).reshape(-1)
                if fr.size != self.n:
                    raise ValueError(f"force callback returned {fr.size} values, expected {self.n}")
                e[0] = float(en)
                arr(f)[:] = fr
            self.fns["force"] = _FORCE_FN(guard(_force))
        if project is not None:
            def _project(_u, rho):
                r = arr(rho)
                out = project(r.reshape(self.shape).copy())
                r[:] = np.asarray(out, dtype=np.float64).reshape(-1)
            self.fns["project"] = _PROJECT_FN(guard(_project))
        if stop is not None:
            def _stop(_u, step, t, e, flag):
                flag[0] = 1 if stop(int(step), float(t), float(e)) else 0
            self.fns["stop"] = _STOP_FN(guard(_stop))
        if snapshot is not None:
            def _snapshot(_u, step, t, e, rho):
                snapshot(int(step), float(t), float(e), arr(rho).reshape(self.shape).copy())
            self.fns["snapshot"] = _SNAPSHOT_FN(guard(_snapshot))
        self.struct = _Hooks(None, *[self.fns.get(k, C.cast(None, t)) for k, t in
                                     (("force", _FORCE_FN), ("project", _PROJECT_FN), ("stop", _STOP_FN),
                                      ("snapshot", _SNAPSHOT_FN))], int(snapshot_interval))

    def ref(self):
        return C.byref(self.struct) if self.fns else None

    def check(self, rc):
        if rc != 0 and self.error is not None:
            raise self.error
        _check(rc)


class _EigenConfig(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int), ("hessian_eps", C.c_double),
                ("fixed_boundary", C.c_int)]


class _EigenResult(C.Structure):
    _fields_ = [("eigenvalue", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("hessian_products", C.c_int)]


class _SaddleConfig(C.Structure):
    _fields_ = [("fire", _Fire), ("eigen", _EigenConfig), ("param", _Param), ("omega_tolerance", C.c_double),
                ("tolerance_start", C.c_double), ("tolerance_factor", C.c_double), ("max_outer_iterations", C.c_int),
                ("eigenvalue_clear_count", C.c_int)]


class _SaddleResult(C.Structure):
    _fields_ = [("grand_potential", C.c_double), ("eigenvalue", C.c_double), ("outer_iterations", C.c_int),
                ("converged", C.c_int), ("evaluations", C.c_long)]


class _Picard(C.Structure):
    _fields_ = [("mixing", C.c_double), ("min_density", C.c_double), ("tolerance", C.c_double),
                ("max_iterations", C.c_int)]


class _PicardResult(C.Structure):
    _fields_ = [("grand_potential", C.c_double), ("residual", C.c_double), ("iterations", C.c_int),
                ("converged", C.c_int)]


class _AdjointConfig(C.Structure):
    _fields_ = [("dt", C.c_double), ("fp_iterations", C.c_int), ("n_steps", C.c_int),
                ("checkpoint_interval", C.c_int)]


class _AdjointResult(C.Structure):
    _fields_ = [("loss", C.c_double), ("min_density", C.c_double), ("evaluations", C.c_long),
                ("hessian_products", C.c_long), ("checkpoints", C.c_int), ("store_bytes", C.c_double)]


_dp = C.POINTER(C.c_double)
_lib = None


def lib():
    """Load libcdft.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m classicaldft_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.cdft_last_error.restype = C.c_char_p
        L.cdft_ctx_create.argtypes = [C.POINTER(_Grid), _dp, C.c_int, C.POINTER(_Interaction), C.c_int, C.c_double,
                                      C.POINTER(_Options), C.POINTER(C.c_void_p)]
        L.cdft_ctx_destroy.argtypes = [C.c_void_p]
        L.cdft_ctx_destroy.restype = None
        L.cdft_grid_shape.argtypes = [C.c_void_p, C.POINTER(C.c_long)]
        L.cdft_a_vdw.argtypes = [C.c_void_p, C.c_int, _dp]
        L.cdft_bulk_chemical_potential.argtypes = [C.c_void_p, C.c_double, _dp]
        L.cdft_total.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.POINTER(_Result), _dp]
        L.cdft_species_count.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.cdft_bulk_free_energy_density.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_double)]
        L.cdft_bulk_chemical_potentials.argtypes = [C.c_void_p, _dp, _dp]
        L.cdft_bulk_pressure.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_double)]
        L.cdft_bulk_batch.argtypes = [C.c_void_p, C.c_long, _dp, _dp, _dp, _dp]
        L.cdft_contributions.argtypes = [C.c_void_p, _dp, _dp, _dp, C.POINTER(_Result), _dp, _dp, _dp, _dp]
        L.cdft_total_species.argtypes = [C.c_void_p, _dp, _dp, _dp, C.POINTER(_Result), _dp]
        L.cdft_total_species_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, C.POINTER(_Result),
                                                C.c_void_p]
        L.cdft_time_total_species_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, C.c_int, C.c_void_p,
                                                     C.POINTER(C.c_double)]
        L.cdft_total_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.POINTER(_Result),
                                        C.c_void_p]
        L.cdft_total_device_repeat.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                               C.POINTER(_Result), C.c_void_p]
        L.cdft_weighted_densities.argtypes = [C.c_void_p, _dp, _dp]
        L.cdft_weighted_densities_noalloc.argtypes = [C.c_void_p, _dp, _dp]
        L.cdft_fft_r2c.argtypes = [C.c_void_p, _dp, _dp]
        L.cdft_fft_c2r.argtypes = [C.c_void_p, _dp, _dp]
        L.cdft_minimize_grand_potential.argtypes = [C.c_void_p, C.POINTER(_Fire), C.POINTER(_Param), C.c_int, _dp,
                                                    _dp, C.c_double, _dp, C.POINTER(_MinResult), _dp]
        L.cdft_minimize_grand_potential_batch.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(_Fire),
                                                          C.POINTER(_Param), C.c_int, C.POINTER(_dp), _dp,
                                                          C.POINTER(_dp), C.POINTER(_MinResult), C.POINTER(C.c_int)]
        L.cdft_minimize_fixed_mass.argtypes = [C.c_void_p, C.POINTER(_Fire), C.POINTER(_Param), C.c_int, _dp, _dp,
                                               C.c_double, C.c_double, _dp, C.POINTER(_MinResult), _dp]
        L.cdft_ddft_run.argtypes = [C.c_void_p, C.POINTER(_StepConfig), C.c_int, _dp, C.c_double, C.POINTER(_Hooks),
                                    _dp, C.POINTER(_SimResult), _dp, _dp]
        L.cdft_time_total_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p,
                                             _dp]
        L.cdft_profile_stages.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_int, _dp,
                                          _dp, C.POINTER(C.c_int)]
        L.cdft_stage_name.restype = C.c_char_p
        L.cdft_stage_name.argtypes = [C.c_void_p, C.c_int]
        L.cdft_schedule.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.cdft_autotune.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.cdft_ddft_split_operator.argtypes = [C.c_void_p, C.POINTER(_StepConfig), C.c_int, _dp, C.c_double,
                                               C.POINTER(_Hooks), _dp, _dp]
        L.cdft_ddft_crank_nicolson.argtypes = [C.c_void_p, C.POINTER(_StepConfig), C.c_int, _dp, C.c_double,
                                               C.POINTER(_Hooks), _dp, _dp, C.POINTER(C.c_int)]
        L.cdft_picard_solve.argtypes = [C.c_void_p, C.POINTER(_Picard), _dp, C.c_double, _dp, _dp, C.POINTER(_Hooks),
                                        _dp, C.POINTER(_PicardResult)]
        L.cdft_saddle_search.argtypes = [C.c_void_p, C.POINTER(_SaddleConfig), _dp, _dp, C.c_double, C.c_double,
                                         C.c_ulonglong, _dp, _dp, C.POINTER(_SaddleResult)]
        L.cdft_smallest_eigenpair.argtypes = [C.c_void_p, C.POINTER(_EigenConfig), _dp, _dp, C.c_double, _dp,
                                              C.c_ulonglong, _dp, C.POINTER(_EigenResult)]
        L.cdft_hessian_vector.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.cdft_wells_potential.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.cdft_ddft_fixed_run.argtypes = [C.c_void_p, C.POINTER(_AdjointConfig), _dp, C.c_double, _dp, _dp]
        L.cdft_ddft_adjoint_wells.argtypes = [C.c_void_p, C.POINTER(_AdjointConfig), C.c_int, _dp, C.c_double, _dp,
                                              _dp, C.POINTER(_AdjointResult), _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().cdft_last_error().decode()
        if rc == -1:
            raise InvalidArgument(rc, msg)
        raise CdftError(rc, msg)


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


class Context:
    """One physics::Model (grid, species, interactions, kT) plus its weights on one GPU.

    Mirrors functionals::make_functional (functional.hpp:93-107); `total`
    mirrors functionals::total (functionals.hpp:64-95).
    """

    def __init__(self, shape, dx, diameter=1.0, interactions=(), kT=1.0, model="WhiteBearII", precision="fp64",
                 A=1.0, B=0.0, device=0, schedule="auto", diameters=None):
        """`diameters` (a sequence) makes a context of that many species; a
        mixture is evaluated with `total_species`."""
        g = _Grid(float(dx), (C.c_double * 3)(*[float(n) * float(dx) for n in shape]))
        its = (_Interaction * max(1, len(interactions)))()
        for k, it in enumerate(interactions):
            its[k] = _Interaction(it.get("si", 0), it.get("sj", 0), POTENTIALS[it.get("potential", "LJ")],
                                  it.get("sigma", 1.0), it.get("epsilon", 1.0), it.get("r_cutoff", -1.0),
                                  it.get("alpha", 50.0), SPLITS[it.get("split", "WCA")],
                                  SCHEMES[it.get("scheme", "quadraticF")])
        opt = _Options(FMT_MODELS[model], float(A), float(B), PRECISIONS[precision], int(device),
                       SCHEDULES[schedule])
        d = np.array([float(diameter)] if diameters is None else [float(x) for x in diameters])
        h = C.c_void_p()
        _check(lib().cdft_ctx_create(C.byref(g), _p(d), len(d), its, len(interactions), float(kT), C.byref(opt),
                                     C.byref(h)))
        self._h = h
        sh = (C.c_long * 3)()
        _check(lib().cdft_grid_shape(self._h, sh))
        self.shape = tuple(sh)
        sc = C.c_int()
        _check(lib().cdft_schedule(self._h, C.byref(sc)))
        self.schedule = "fused" if sc.value == 2 else "general"
        self.N = int(np.prod(self.shape))
        self.M = self.shape[0] * self.shape[1] * (self.shape[2] // 2 + 1)
        self.dx = float(dx)
        self.precision = precision
        self.n_inter = len(interactions)
        self.model = model
        self.nf = (11 if model != "none" else 0) + (1 if interactions else 0)
        ns = C.c_int()
        _check(lib().cdft_species_count(self._h, C.byref(ns)))
        self.n_species = ns.value

    def close(self):
        if getattr(self, "_h", None):
            lib().cdft_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _field(self, a, name, species=1):
        """`a` as a contiguous float64 vector of species * N values; ValueError for any other size (the
        native entry points read and write exactly that many values)."""
        v = _f64(a)
        if v.size != species * self.N:
            raise ValueError(f"{name} has {v.size} values, expected {species} x {self.N} (grid {self.shape})")
        return v

    def _opt(self, a, name, species=1):
        return None if a is None else self._field(a, name, species)

    def a_vdw(self, k=0):
        v = C.c_double()
        _check(lib().cdft_a_vdw(self._h, k, C.byref(v)))
        return v.value

    def bulk_mu(self, rho):
        v = C.c_double()
        _check(lib().cdft_bulk_chemical_potential(self._h, float(rho), C.byref(v)))
        return v.value

    def contributions(self, rho, mu, vext=None):
        """The terms of functionals::total separately: (result dict, {"ideal", "external", "hard_sphere",
        "mean_field"} force arrays of shape (n_species, *shape))."""
        S = self.n_species
        rho = self._field(rho, "rho", S)
        v = self._opt(vext, "vext", S)
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (S,)))
        f = {k: np.zeros(S * self.N) for k in ("ideal", "external", "hard_sphere", "mean_field")}
        r = _Result()
        _check(lib().cdft_contributions(self._h, _p(rho), _p(v), _p(m), C.byref(r), _p(f["ideal"]),
                                        _p(f["external"]), _p(f["hard_sphere"]), _p(f["mean_field"])))
        res = {k: getattr(r, k) for k, _ in _Result._fields_}
        return res, {k: a.reshape((S,) + self.shape) for k, a in f.items()}

    def _bulk_rho(self, rho):
        r = np.ascontiguousarray(np.atleast_1d(np.asarray(rho, dtype=np.float64)))
        if r.size != self.n_species:
            raise ValueError(f"expected {self.n_species} bulk densities, got {r.size}")
        return r

    def bulk_free_energy_density(self, rho):
        """BulkThermodynamics::free_energy_density (kT units per volume); rho has n_species values."""
        v = C.c_double()
        _check(lib().cdft_bulk_free_energy_density(self._h, _p(self._bulk_rho(rho)), C.byref(v)))
        return v.value

    def bulk_chemical_potentials(self, rho):
        """BulkThermodynamics::chemical_potentials: one value per species."""
        r = self._bulk_rho(rho)
        mu = np.zeros(self.n_species)
        _check(lib().cdft_bulk_chemical_potentials(self._h, _p(r), _p(mu)))
        return mu

    def bulk_batch(self, rho):
        """The bulk equation of state of many states on the device: rho has shape (n_states, n_species);
        returns (f, mu, p) with shapes (n,), (n, n_species), (n,)."""
        r = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1, self.n_species))
        n = r.shape[0]
        f, mu, p = np.zeros(n), np.zeros((n, self.n_species)), np.zeros(n)
        _check(lib().cdft_bulk_batch(self._h, n, _p(r), _p(f), _p(mu), _p(p)))
        return f, mu, p

    def bulk_pressure(self, rho):
        """BulkThermodynamics::pressure = sum rho mu - f."""
        v = C.c_double()
        _check(lib().cdft_bulk_pressure(self._h, _p(self._bulk_rho(rho)), C.byref(v)))
        return v.value

    def total(self, rho, mu, vext=None, forces=True):
        """Return (result dict, forces or None)."""
        rho = self._field(rho, "rho")
        v = self._opt(vext, "vext")
        f = np.zeros(self.N) if forces else None
        r = _Result()
        _check(lib().cdft_total(self._h, _p(rho), _p(v), float(mu), C.byref(r), _p(f)))
        res = {k: getattr(r, k) for k, _ in _Result._fields_}
        return res, (f.reshape(self.shape) if forces else None)

    def total_species(self, rho, mu, vext=None, forces=True):
        """functionals::total for S species: rho and vext are (S, *shape), mu has S
        values.  Return (result dict, forces (S, *shape) or None)."""
        S = self.n_species
        rho = self._field(rho, "rho", S)
        v = self._opt(vext, "vext", S)
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (S,)))
        f = np.zeros(S * self.N) if forces else None
        r = _Result()
        _check(lib().cdft_total_species(self._h, _p(rho), _p(v), _p(m), C.byref(r), _p(f)))
        res = {k: getattr(r, k) for k, _ in _Result._fields_}
        return res, (f.reshape((S,) + self.shape) if forces else None)

    def total_species_device(self, d_rho, mu, d_vext=None, d_forces=None):
        """Mixture evaluation on device pointers (S * N doubles each)."""
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (self.n_species,)))
        r = _Result()
        _check(lib().cdft_total_species_device(self._h, C.c_void_p(d_rho), C.c_void_p(d_vext) if d_vext else None,
                                               _p(m), C.byref(r), C.c_void_p(d_forces) if d_forces else None))
        return {k: getattr(r, k) for k, _ in _Result._fields_}

    def time_total_species_device(self, d_rho, mu, count, d_vext=None, d_forces=None):
        """GPU milliseconds for `count` back-to-back mixture evaluations."""
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (self.n_species,)))
        ms = C.c_double()
        _check(lib().cdft_time_total_species_device(self._h, C.c_void_p(d_rho),
                                                    C.c_void_p(d_vext) if d_vext else None, _p(m), int(count),
                                                    C.c_void_p(d_forces) if d_forces else None, C.byref(ms)))
        return ms.value

    def total_device(self, d_rho, mu, d_vext=None, d_forces=None, count=1):
        """Evaluate on device pointers (ints, e.g. torch tensor.data_ptr())."""
        r = _Result()
        _check(lib().cdft_total_device_repeat(self._h, C.c_void_p(d_rho), C.c_void_p(d_vext) if d_vext else None,
                                              float(mu), int(count), C.byref(r),
                                              C.c_void_p(d_forces) if d_forces else None))
        return {k: getattr(r, k) for k, _ in _Result._fields_}

    def time_total_device(self, d_rho, mu, count, d_vext=None, d_forces=None):
        """GPU milliseconds for `count` back-to-back evaluations (HIP events on the context stream)."""
        ms = C.c_double()
        _check(lib().cdft_time_total_device(self._h, C.c_void_p(d_rho), C.c_void_p(d_vext) if d_vext else None,
                                            float(mu), int(count), C.c_void_p(d_forces) if d_forces else None,
                                            C.byref(ms)))
        return ms.value

    def profile_stages(self, d_rho, mu, reps=10, d_vext=None, d_forces=None):
        """Per-kernel average ms and algorithmic bytes of one evaluation: list of (name, ms, bytes)."""
        ms = (C.c_double * 16)()
        by = (C.c_double * 16)()
        n = C.c_int()
        _check(lib().cdft_profile_stages(self._h, C.c_void_p(d_rho), C.c_void_p(d_vext) if d_vext else None,
                                         float(mu), C.c_void_p(d_forces) if d_forces else None, int(reps), ms, by,
                                         C.byref(n)))
        return [(lib().cdft_stage_name(self._h, i).decode(), ms[i], by[i]) for i in range(n.value)]

    def weighted_densities(self, rho, noalloc=False):
        """Weighted densities: 11 FMT channels (when an FMT model is set) then the mean-field convolution.
        noalloc: through cdft_weighted_densities_noalloc (no device allocation; the library's workspace)."""
        out = np.zeros(max(self.nf, 1) * self.N)
        fn = lib().cdft_weighted_densities_noalloc if noalloc else lib().cdft_weighted_densities
        _check(fn(self._h, _p(self._field(rho, "rho")), _p(out)))
        return out[: self.nf * self.N].reshape((self.nf,) + self.shape)

    def r2c(self, x):
        out = np.zeros(2 * self.M)
        _check(lib().cdft_fft_r2c(self._h, _p(self._field(x, "x")), _p(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(self.shape[0], self.shape[1], -1)

    def c2r(self, X):
        X = np.ascontiguousarray(X, dtype=np.complex128).reshape(-1)
        if X.size != self.M:
            raise ValueError(f"spectrum has {X.size} values, expected {self.M}")
        buf = np.empty(2 * X.size)
        buf[0::2] = X.real
        buf[1::2] = X.imag
        out = np.zeros(self.N)
        _check(lib().cdft_fft_c2r(self._h, _p(buf), _p(out)))
        return out.reshape(self.shape)

    def minimize_grand_potential(self, rho0, mu, fire=None, param="unbounded", rho_min=1e-18, rho_max=1.0,
                                 homogeneous_boundary=True, vext=None, max_steps=None, record_energies=True):
        f = dict(dt=1e-3, dt_max=1e-2, dt_min=1e-8, alpha_start=0.1, f_inc=1.1, f_dec=0.5, f_alpha=0.99,
                 n_delay=5, max_uphill=20, force_tolerance=0.1, max_steps=10000, force_max_tolerance=0.0)
        if fire:
            f.update(fire)
        if max_steps is not None:
            f["max_steps"] = max_steps
        fs = _Fire(*[f[k] for k, _ in _Fire._fields_])
        prm = _Param(0 if param == "unbounded" else 1, float(rho_min), float(rho_max))
        out = np.zeros(self.N)
        res = _MinResult()
        en = np.full(int(f["max_steps"]) + 1, np.nan) if record_energies else None
        _check(lib().cdft_minimize_grand_potential(self._h, C.byref(fs), C.byref(prm), int(bool(homogeneous_boundary)),
                                                   _p(self._field(rho0, "rho0")), _p(self._opt(vext, "vext")),
                                                   float(mu),
                                                   _p(out), C.byref(res), _p(en)))
        r = {k: getattr(res, k) for k, _ in _MinResult._fields_}
        r["converged"] = bool(r["converged"])
        return out.reshape(self.shape), r, (en[: r["iterations"] + 1] if en is not None else None)

    def minimize_fixed_mass(self, rho0, mu, target_mass, fire=None, param="unbounded", rho_min=1e-18, rho_max=1.0,
                            homogeneous_boundary=True, vext=None, max_steps=None, record_energies=True):
        """Minimizer::fixed_mass: FIRE at fixed total mass; returns (rho, result, F after each step)."""
        f = dict(dt=1e-3, dt_max=1e-2, dt_min=1e-8, alpha_start=0.1, f_inc=1.1, f_dec=0.5, f_alpha=0.99,
                 n_delay=5, max_uphill=20, force_tolerance=0.1, max_steps=10000, force_max_tolerance=0.0)
        if fire:
            f.update(fire)
        if max_steps is not None:
            f["max_steps"] = max_steps
        fs = _Fire(*[f[k] for k, _ in _Fire._fields_])
        prm = _Param(0 if param == "unbounded" else 1, float(rho_min), float(rho_max))
        out = np.zeros(self.N)
        res = _MinResult()
        en = np.full(int(f["max_steps"]) + 1, np.nan) if record_energies else None
        _check(lib().cdft_minimize_fixed_mass(self._h, C.byref(fs), C.byref(prm), int(bool(homogeneous_boundary)),
                                              _p(self._field(rho0, "rho0")), _p(self._opt(vext, "vext")), float(mu),
                                              float(target_mass), _p(out), C.byref(res), _p(en)))
        r = {k: getattr(res, k) for k, _ in _MinResult._fields_}
        r["converged"] = bool(r["converged"])
        return out.reshape(self.shape), r, (en[: r["iterations"] + 1] if en is not None else None)

    def ddft_run(self, rho0, mu, n_steps, dt=1e-4, dt_max=1.0, fp_tolerance=1e-4, fp_max_iterations=100,
                 min_density=1e-18, vext=None, record_energies=True, force=None, boundary=None, stop=None,
                 snapshot=None, snapshot_interval=0):
        """Simulation::run (integrating-factor DDFT).  force(rho) -> (energy, forces) replaces the context's
        functional; boundary(rho) -> rho is applied to the initial state and after every step;
        stop(step, time, energy) -> bool ends the run; snapshot(step, time, energy, rho) receives the density
        at step 0 and every snapshot_interval steps.  Returns (rho, result dict, energies of the steps taken)."""
        cfg = _StepConfig(float(dt), 1.0, float(min_density), float(dt_max), float(fp_tolerance),
                          int(fp_max_iterations))
        rho = self._field(rho0, "rho0").copy()
        res = _SimResult()
        en = np.zeros(n_steps + 1) if record_energies else None
        cb = _HostCallbacks(self.shape, force, boundary, stop, snapshot, snapshot_interval)
        cb.check(lib().cdft_ddft_run(self._h, C.byref(cfg), int(n_steps), _p(self._opt(vext, "vext")),
                                     float(mu), cb.ref(), _p(rho), C.byref(res), _p(en), None))
        r = {k: getattr(res, k) for k, _ in _SimResult._fields_}
        return rho.reshape(self.shape), r, (en[: r["steps"] + 1] if en is not None else None)

    def smallest_eigenpair(self, rho, mu, initial_guess=None, seed=0, tolerance=1e-4, max_iterations=300,
                           hessian_eps=1e-6, fixed_boundary=False, vext=None):
        """EigenvalueSolver::solve: smallest eigenpair of the grand-potential Hessian (hessian_eps = 0: exact
        products); returns (eigenvector, result dict)."""
        cfg = _EigenConfig(float(tolerance), int(max_iterations), float(hessian_eps), int(bool(fixed_boundary)))
        res = _EigenResult()
        vec = np.zeros(self.N)
        _check(lib().cdft_smallest_eigenpair(self._h, C.byref(cfg), _p(self._field(rho, "rho")),
                                             _p(self._opt(vext, "vext")), float(mu),
                                             _p(self._opt(initial_guess, "initial_guess")),
                                             int(seed), _p(vec), C.byref(res)))
        r = {k: getattr(res, k) for k, _ in _EigenResult._fields_}
        r["converged"] = bool(r["converged"])
        return vec.reshape(self.shape), r

    def saddle_search(self, rho0, mu, target_mass, seed=0, fire=None, eigen=None, param="unbounded", rho_min=1e-18,
                      rho_max=1.0, omega_tolerance=1e-5, tolerance_start=1e-3, tolerance_factor=10.0,
                      max_outer_iterations=20, eigenvalue_clear_count=2, vext=None):
        """saddle_point::Search::find (critical cluster at fixed mass); returns (rho, eigenvector, result dict)."""
        f = dict(dt=1e-3, dt_max=1.0, dt_min=1e-8, alpha_start=0.01, f_inc=1.1, f_dec=0.5, f_alpha=1.0, n_delay=5,
                 max_uphill=20, force_tolerance=1e-4, max_steps=100000, force_max_tolerance=0.0)
        f.update(fire or {})
        e = dict(tolerance=1e-4, max_iterations=300, hessian_eps=1e-6)
        e.update(eigen or {})
        cfg = _SaddleConfig(_Fire(*[f[k] for k, _ in _Fire._fields_]),
                            _EigenConfig(float(e["tolerance"]), int(e["max_iterations"]), float(e["hessian_eps"]), 1),
                            _Param(0 if param == "unbounded" else 1, float(rho_min), float(rho_max)),
                            float(omega_tolerance), float(tolerance_start), float(tolerance_factor),
                            int(max_outer_iterations), int(eigenvalue_clear_count))
        res = _SaddleResult()
        ro, eo = np.zeros(self.N), np.zeros(self.N)
        _check(lib().cdft_saddle_search(self._h, C.byref(cfg), _p(self._field(rho0, "rho0")), _p(self._opt(vext, "vext")),
                                        float(mu), float(target_mass), int(seed), _p(ro), _p(eo), C.byref(res)))
        r = {k: getattr(res, k) for k, _ in _SaddleResult._fields_}
        r["converged"] = bool(r["converged"])
        return ro.reshape(self.shape), eo.reshape(self.shape), r

    def ddft_split_operator(self, rho0, mu, n_steps, dt=1e-4, diffusion_coefficient=1.0, min_density=1e-18,
                            vext=None, force=None):
        """split_operator_step, n_steps times (force(rho) -> (energy, forces) replaces the functional);
        returns (rho, energy)."""
        cfg = _StepConfig(float(dt), float(diffusion_coefficient), float(min_density), float(dt), 1e-4, 100)
        rho = self._field(rho0, "rho0").copy()
        e = C.c_double()
        cb = _HostCallbacks(self.shape, force)
        cb.check(lib().cdft_ddft_split_operator(self._h, C.byref(cfg), int(n_steps), _p(self._opt(vext, "vext")),
                                                float(mu), cb.ref(), _p(rho), C.byref(e)))
        return rho.reshape(self.shape), e.value

    def ddft_crank_nicolson(self, rho0, mu, n_steps, dt=1e-4, diffusion_coefficient=1.0, min_density=1e-18,
                            fp_tolerance=1e-4, fp_max_iterations=100, vext=None, force=None):
        """crank_nicholson_step, n_steps times (force as for ddft_split_operator); returns (rho, energy,
        iterations of the last step)."""
        cfg = _StepConfig(float(dt), float(diffusion_coefficient), float(min_density), float(dt), float(fp_tolerance),
                          int(fp_max_iterations))
        rho = self._field(rho0, "rho0").copy()
        e, it = C.c_double(), C.c_int()
        cb = _HostCallbacks(self.shape, force)
        cb.check(lib().cdft_ddft_crank_nicolson(self._h, C.byref(cfg), int(n_steps), _p(self._opt(vext, "vext")),
                                                float(mu), cb.ref(), _p(rho), C.byref(e), C.byref(it)))
        return rho.reshape(self.shape), e.value, it.value

    def picard(self, rho0, mu, mixing=0.01, min_density=1e-30, tolerance=1e-6, max_iterations=5000,
               target_mass=None, background=None, vext=None, force=None, constraint=None):
        """Picard::solve; target_mass (None: unconstrained) applies fixed_mass_constraint with `background`;
        force(rho) -> (energy, forces) replaces the functional and constraint(rho) -> rho is applied after
        every update.  Returns (rho, result dict)."""
        cfg = _Picard(float(mixing), float(min_density), float(tolerance), int(max_iterations))
        res = _PicardResult()
        rho = self._field(rho0, "rho0").copy()
        tgt = None if target_mass is None else C.byref(C.c_double(float(target_mass)))
        cb = _HostCallbacks(self.shape, force, constraint)
        cb.check(lib().cdft_picard_solve(self._h, C.byref(cfg), _p(self._opt(vext, "vext")), float(mu), tgt,
                                         _p(self._opt(background, "background")), cb.ref(), _p(rho), C.byref(res)))
        r = {k: getattr(res, k) for k, _ in _PicardResult._fields_}
        r["converged"] = bool(r["converged"])
        return rho.reshape(self.shape), r

    def autotune(self, reps=3):
        """Tune the general schedule's FFT radix orders and launch shapes for this grid; returns the time of one
        evaluation (microseconds) before and after."""
        a, b = C.c_double(), C.c_double()
        _check(lib().cdft_autotune(self._h, int(reps), C.byref(a), C.byref(b)))
        return a.value, b.value

    # ------------------------------------------------------------ adjoint (FP64 contexts)
    def hessian_vector(self, rho, u):
        """H(rho) u, the Jacobian of the forces of total() applied to u."""
        out = np.zeros(self.N)
        _check(lib().cdft_hessian_vector(self._h, _p(self._field(rho, "rho")), _p(self._field(u, "u")), _p(out)))
        return out.reshape(self.shape)

    def wells_potential(self, params):
        """Sum of Gaussian wells; params: (n_wells, 5) rows of (A, w, cx, cy, cz)."""
        prm = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 5)
        out = np.zeros(self.N)
        _check(lib().cdft_wells_potential(self._h, prm.shape[0], _p(prm.reshape(-1)), _p(out)))
        return out.reshape(self.shape)

    def ddft_fixed_run(self, rho0, mu, n_steps, dt, fp_iterations, vext=None):
        """Fixed-dt DDFT map with exactly fp_iterations iterations per step; returns (rho, min density)."""
        cfg = _AdjointConfig(float(dt), int(fp_iterations), int(n_steps), 0)
        rho = self._field(rho0, "rho0").copy()
        mn = C.c_double(0.0)
        _check(lib().cdft_ddft_fixed_run(self._h, C.byref(cfg), _p(self._opt(vext, "vext")), float(mu),
                                         _p(rho), C.byref(mn)))
        return rho.reshape(self.shape), mn.value

    def ddft_adjoint_wells(self, rho0, target, params, mu, n_steps, dt, fp_iterations, checkpoint_interval=0,
                           want_fields=False):
        """Loss sum (rho_K - target)^2 dV of the fixed-dt DDFT map under a well potential, and its gradient
        with respect to the (n_wells, 5) well parameters.  With want_fields also dL/drho0, dL/dV and rho_K."""
        prm = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 5)
        cfg = _AdjointConfig(float(dt), int(fp_iterations), int(n_steps), int(checkpoint_interval))
        res = _AdjointResult()
        grad = np.zeros(prm.size)
        fields = [np.zeros(self.N) for _ in range(3)] if want_fields else [None, None, None]
        _check(lib().cdft_ddft_adjoint_wells(self._h, C.byref(cfg), prm.shape[0], _p(prm.reshape(-1)), float(mu),
                                             _p(self._field(rho0, "rho0")), _p(self._field(target, "target")),
                                             C.byref(res), _p(grad),
                                             *[_p(f) for f in fields]))
        r = {k: getattr(res, k) for k, _ in _AdjointResult._fields_}
        out = [grad.reshape(prm.shape), r]
        if want_fields:
            out.append(tuple(f.reshape(self.shape) for f in fields))
        return tuple(out)


def minimize_grand_potential_batch(contexts, rho0s, mus, fire=None, param="unbounded", rho_min=1e-18, rho_max=1.0,
                                   homogeneous_boundary=True, max_steps=None):
    """Independent Minimizer::grand_potential runs of several contexts advanced in lockstep on one GPU
    (cdft_minimize_grand_potential_batch).  Returns a list of (rho, result, status) per run; a run whose
    FIRE exceeded max_uphill has status CDFT_ERUNTIME (-2) instead of raising."""
    n = len(contexts)
    if not (len(rho0s) == len(mus) == n):
        raise ValueError("contexts, rho0s and mus differ in length")
    f = dict(dt=1e-3, dt_max=1e-2, dt_min=1e-8, alpha_start=0.1, f_inc=1.1, f_dec=0.5, f_alpha=0.99,
             n_delay=5, max_uphill=20, force_tolerance=0.1, max_steps=10000, force_max_tolerance=0.0)
    if fire:
        f.update(fire)
    if max_steps is not None:
        f["max_steps"] = max_steps
    fs = _Fire(*[f[k] for k, _ in _Fire._fields_])
    prm = _Param(0 if param == "unbounded" else 1, float(rho_min), float(rho_max))
    ins = [_f64(r) for r in rho0s]
    outs = [np.zeros(ctx.N) for ctx in contexts]
    for ctx, r in zip(contexts, ins):
        if r.size != ctx.N:
            raise ValueError("a density does not match its context's grid")
    hs = (C.c_void_p * max(n, 1))(*[ctx._h for ctx in contexts])
    pin = (_dp * max(n, 1))(*[_p(r) for r in ins])
    pout = (_dp * max(n, 1))(*[_p(o) for o in outs])
    mu = np.array([float(m) for m in mus] or [0.0])
    res = (_MinResult * max(n, 1))()
    st = (C.c_int * max(n, 1))()
    _check(lib().cdft_minimize_grand_potential_batch(n, hs, C.byref(fs), C.byref(prm), int(bool(homogeneous_boundary)),
                                                     pin, _p(mu), pout, res, st))
    out = []
    for i, ctx in enumerate(contexts):
        r = {k: getattr(res[i], k) for k, _ in _MinResult._fields_}
        r["converged"] = bool(r["converged"])
        out.append((outs[i].reshape(ctx.shape), r, int(st[i])))
    return out
