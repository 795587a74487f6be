"""Python handle on the device-resident sweep driver (nc_chain_* in the C-ABI),
mirroring the methods of the reference's Chain (includes/nutcracker/chain.hpp)."""

from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import core
from .fixtures import raw_state
from .operators import OperatorSite, compute_bandwidth_dimension_sequence, maximum_bandwidth_dimension

CPLX = np.complex128


class OptimizerFailure(RuntimeError):
    pass


_CHAIN_SIGS = {
    "nc_chain_create": (C.c_void_p, [C.c_uint32] + [C.POINTER(C.c_uint32)] * 5 + [C.c_void_p, C.POINTER(C.c_uint32)]),
    "nc_chain_destroy": (None, [C.c_void_p]),
    "nc_chain_set_options": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_char_p,
                                       C.c_uint32, C.c_int]),
    "nc_chain_set_raw_site": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "nc_chain_get_site": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "nc_chain_reset": (C.c_int, [C.c_void_p]),
    "nc_chain_optimize_site": (C.c_int, [C.c_void_p]),
    "nc_chain_move": (C.c_int, [C.c_void_p, C.c_int]),
    "nc_chain_sweep": (C.c_int, [C.c_void_p]),
    "nc_chain_sweep_until_converged": (C.c_int, [C.c_void_p]),
    "nc_chain_energy": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "nc_chain_expectation": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nc_chain_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "nc_chain_set_mixed": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int]),
    "nc_chain_add_projector": (C.c_int, [C.c_void_p]),
    "nc_chain_batch_create": (C.c_void_p, [C.POINTER(C.c_void_p), C.c_uint32]),
    "nc_chain_batch_destroy": (None, [C.c_void_p]),
    "nc_chain_batch_sweep": (C.c_int, [C.c_void_p]),
    "nc_chain_batch_sweep_until_converged": (C.c_int, [C.c_void_p]),
    "nc_chain_batch_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "nc_chain_batch_reset": (C.c_int, [C.c_void_p]),
    "nc_chain_set_growth": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "nc_chain_increase_bandwidth": (C.c_int, [C.c_void_p, C.c_uint32]),
    "nc_chain_optimize_chain": (C.c_int, [C.c_void_p]),
    "nc_chain_get_bonds": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "nc_chain_state_overlap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nc_chain_clear_projectors": (C.c_int, [C.c_void_p]),
    "nc_chain_projector_overlap": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "nc_chain_randomize": (C.c_int, [C.c_void_p, C.c_uint64]),
    "nc_chain_solve_for_multiple_levels": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.POINTER(C.c_double)]),
    "nc_set_krylov_graphs": (C.c_int, [C.c_int]),
}


def set_krylov_graphs(on: bool) -> bool:
    """Record each Arnoldi extension of a lockstep (ChainBatch) site solve in a HIP
    graph and replay it (default), or launch the kernels one by one.  Returns the
    previous setting."""
    return bool(_lib().nc_set_krylov_graphs(1 if on else 0))


def _lib():
    h = core.lib()
    if not getattr(h, "_chain_sigs", False):
        for name, (res, args) in _CHAIN_SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        h._chain_sigs = True
    return h


def _u32(values) -> "C.Array":
    arr = np.ascontiguousarray(np.asarray(values, dtype=np.uint32))
    return arr, arr.ctypes.data_as(C.POINTER(C.c_uint32))


class Chain:
    """A chain of operator sites solved at a fixed bond dimension ``chi``.

    Parameters follow ChainOptions (chain_options.cpp:21-29); ``seed`` selects the
    raw initial state (nutcracker_amd.fixtures), or pass ``raw_sites``.  ``mixed``
    selects the mixed-precision form (complex64 matvecs, and with ``env32``
    complex64 environment updates; FP64 Krylov vectors, Rayleigh quotient and site
    moves; the Krylov iteration stops at ``tol32``, ``polish`` continues it in
    FP64 to ``site_tol``)."""

    def __init__(self, ops: Sequence[OperatorSite], chi: int, seed: int = 0,
                 raw_sites: Optional[List[np.ndarray]] = None, site_tol: float = 1e-12,
                 sweep_tol: float = 1e-12, sanity: float = 1e-12, max_iterations: int = 10000,
                 which: str = "SR", ncv: int = 3, rethrow: bool = False, mixed: bool = False,
                 env32: bool = False, tol32: float = 1e-7, polish: bool = False, grow: bool = False,
                 chain_tol: float = 1e-12, reset: bool = True):
        self.ops = list(ops)
        self.n = len(self.ops)
        dims = [o.d for o in self.ops]
        chi = min(chi, maximum_bandwidth_dimension(dims))
        self.bonds = compute_bandwidth_dimension_sequence(chi, dims) if self.n > 1 else [1, 1]
        lib = _lib()
        self._keep = []
        d, dp = _u32([o.d for o in self.ops])
        cl, clp = _u32([o.cl for o in self.ops])
        cr, crp = _u32([o.cr for o in self.ops])
        nm, nmp = _u32([o.nmat for o in self.ops])
        idx = np.ascontiguousarray(np.concatenate([np.asarray(o.idx, np.uint32).reshape(-1, order="F") for o in self.ops]))
        mats = np.ascontiguousarray(np.concatenate([np.asarray(o.mats, CPLX).reshape(-1, order="F") for o in self.ops]))
        bonds, bp = _u32(self.bonds)
        self._h = lib.nc_chain_create(self.n, dp, clp, crp, nmp, idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      mats.ctypes.data, bp)
        if not self._h:
            raise RuntimeError(core.last_error())
        self._check(lib.nc_chain_set_options(self._h, site_tol, sweep_tol, sanity, max_iterations, which.encode(),
                                             ncv, 1 if rethrow else 0))
        self.set_mixed(mixed, env32, tol32, polish)
        self._check(lib.nc_chain_set_growth(self._h, 1 if grow else 0, chain_tol))
        if raw_sites is None:
            raw_sites = raw_state(seed, dims, self.bonds)
        for i, a in enumerate(raw_sites):
            a = np.asfortranarray(np.asarray(a, CPLX))
            if a.shape != (self.bonds[i + 1], self.bonds[i], dims[i]):
                raise ValueError("raw site %d has shape %s" % (i, a.shape))
            self._check(lib.nc_chain_set_raw_site(self._h, i, a.ctypes.data))
        if reset:  # reset=False leaves the raw sites for ChainBatch(..., reset=True)
            self.reset()

    def _check(self, rc, failure_ok=False):
        if rc != 0:
            if failure_ok:
                return rc
            msg = core.last_error()
            if rc > 0 or rc == -14:
                raise OptimizerFailure("optimizer failure %d: %s" % (rc, msg))
            raise RuntimeError(msg or ("error %d" % rc))
        return 0

    def close(self):
        if getattr(self, "_h", None):
            _lib().nc_chain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mixed(self, mixed: bool, env32: bool = False, tol32: float = 1e-7, polish: bool = False):
        self._check(_lib().nc_chain_set_mixed(self._h, 1 if mixed else 0, 1 if env32 else 0, tol32,
                                              1 if polish else 0))

    # -- Chain methods -----------------------------------------------------
    def reset(self):
        self._check(_lib().nc_chain_reset(self._h))

    def optimize_site(self) -> int:
        return _lib().nc_chain_optimize_site(self._h)

    def move(self, direction: int):
        self._check(_lib().nc_chain_move(self._h, 1 if direction > 0 else -1))

    def perform_optimization_sweep(self):
        self._check(_lib().nc_chain_sweep(self._h))

    def sweep_until_converged(self):
        self._check(_lib().nc_chain_sweep_until_converged(self._h))

    @property
    def energy(self) -> float:
        e = C.c_double(0)
        cur = C.c_int(0)
        self._check(_lib().nc_chain_energy(self._h, C.byref(e), C.byref(cur)))
        return e.value

    @property
    def current_site(self) -> int:
        e = C.c_double(0)
        cur = C.c_int(0)
        self._check(_lib().nc_chain_energy(self._h, C.byref(e), C.byref(cur)))
        return cur.value

    def expectation(self) -> complex:
        out = np.zeros(1, CPLX)
        self._check(_lib().nc_chain_expectation(self._h, out.ctypes.data))
        return complex(out[0])

    def site(self, i: int) -> np.ndarray:
        d = self.ops[i].d
        out = np.zeros((self.bonds[i + 1], self.bonds[i], d), CPLX, order="F")
        self._check(_lib().nc_chain_get_site(self._h, i, out.ctypes.data))
        return out

    def environments(self, i: int):
        """(left (bl, bl, cl), right (br, br, cr)) environments of the current site i."""
        h = _lib()
        f = h.nc_chain_get_environment
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
        L = np.zeros((self.bonds[i], self.bonds[i], self.ops[i].cl), CPLX, order="F")
        R = np.zeros((self.bonds[i + 1], self.bonds[i + 1], self.ops[i].cr), CPLX, order="F")
        self._check(f(self._h, i, 0, L.ctypes.data))
        self._check(f(self._h, i, 1, R.ctypes.data))
        return L, R

    # -- excited levels (chain.cpp:60-73, 301-310) ----------------------------
    def construct_and_add_projector_from_state(self):
        """Keep the current state (chain at its first site) as a lower level."""
        self._check(_lib().nc_chain_add_projector(self._h))

    def clear_projectors(self):
        self._check(_lib().nc_chain_clear_projectors(self._h))

    def compute_projector_overlap_at_current_site(self) -> float:
        v = C.c_double(0)
        self._check(_lib().nc_chain_projector_overlap(self._h, C.byref(v)))
        return v.value

    def randomize(self, seed: int):
        """New raw site tensors from ``seed``; call reset() afterwards."""
        self._check(_lib().nc_chain_randomize(self._h, seed))
        self._refresh_bonds()

    def restart_level(self, seed: int):
        """reset (chain.cpp:175-237) for the next level: the initial bond
        dimension (with growth) raised to the minimum that leaves room outside the
        lower levels, and a new random state from ``seed``."""
        f = _lib().nc_chain_restart_level
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_uint64]
        self._check(f(self._h, seed))
        self._refresh_bonds()

    def solve_for_multiple_levels(self, levels: int, seed: int = 1) -> List[float]:
        """Converge ``levels`` levels in turn at the chain's bond dimension, each
        orthogonal to the ones before; returns their energies."""
        out = (C.c_double * max(1, levels))()
        self._check(_lib().nc_chain_solve_for_multiple_levels(self._h, levels, seed, out))
        self._refresh_bonds()
        return [out[i] for i in range(levels)]

    solve_for_eigenvalues = solve_for_multiple_levels

    def compute_state_overlap(self, other: "Chain") -> complex:
        """<self|other> (computeStateOverlap, states.hpp:349-382)."""
        out = np.zeros(1, CPLX)
        self._check(_lib().nc_chain_state_overlap(self._h, other._h, out.ctypes.data))
        return complex(out[0])

    # -- bond dimension (base_chain.cpp:64-79, chain.cpp:75-121) -------------
    def _refresh_bonds(self):
        out = (C.c_uint32 * (self.n + 1))()
        self._check(_lib().nc_chain_get_bonds(self._h, out))
        self.bonds = [int(out[i]) for i in range(self.n + 1)]

    def increase_bandwidth_dimension(self, new_chi: int):
        self._check(_lib().nc_chain_increase_bandwidth(self._h, new_chi))
        self._refresh_bonds()

    def optimize_chain(self):
        """Converge, then (with ``grow``) raise chi by one and converge again
        until the energy changes by less than ``chain_tol``."""
        self._check(_lib().nc_chain_optimize_chain(self._h))
        self._refresh_bonds()

    def stats(self) -> dict:
        counts = (C.c_int64 * 7)()
        times = (C.c_double * 4)()
        self._check(_lib().nc_chain_stats(self._h, counts, times))
        return {"matvecs": counts[0], "site_solves": counts[1], "moves": counts[2], "sweeps": counts[3],
                "failures": counts[4], "matvecs32": counts[5], "fp64_refinements": counts[6], "t_optimize_ms": times[0], "t_move_ms": times[1], "t_env_ms": times[2],
                "t_reset_ms": times[3]}


class ChainBatch:
    """Independent chains of the same length sweeping in lockstep, their site
    solves batched into grouped launches (nc_chain_batch_* in the C-ABI).  The
    chains remain usable on their own; keep them alive while the batch is."""

    def __init__(self, chains: Sequence[Chain], reset: bool = False):
        """``reset=True`` resets every chain with grouped launches (the chains
        were made with ``Chain(..., reset=False)`` and hold their raw sites)."""
        self.chains = list(chains)
        arr = (C.c_void_p * len(self.chains))(*[c._h for c in self.chains])
        self._h = _lib().nc_chain_batch_create(arr, len(self.chains))
        if not self._h:
            raise RuntimeError(core.last_error())
        if reset:
            self.reset()

    def reset(self):
        """Chain.reset of every chain at once (nc_chain_batch_reset)."""
        self._check(_lib().nc_chain_batch_reset(self._h))

    def _check(self, rc):
        if rc != 0:
            msg = core.last_error()
            if rc > 0 or rc == -14:
                raise OptimizerFailure("optimizer failure %d: %s" % (rc, msg))
            raise RuntimeError(msg or ("error %d" % rc))

    def perform_optimization_sweep(self):
        self._check(_lib().nc_chain_batch_sweep(self._h))

    def sweep_until_converged(self):
        self._check(_lib().nc_chain_batch_sweep_until_converged(self._h))

    def times(self) -> dict:
        t = (C.c_double * 6)()
        self._check(_lib().nc_chain_batch_times(self._h, t))
        return {"moves_ms": t[0], "site_solves_ms": t[1], "prepare_ms": t[2], "batched_solves_ms": t[3],
                "post_checks_ms": t[4], "krylov_ms": t[5]}

    @property
    def energies(self) -> List[float]:
        return [c.energy for c in self.chains]

    def close(self):
        if getattr(self, "_h", None):
            _lib().nc_chain_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
