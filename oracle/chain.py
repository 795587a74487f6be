"""CPU restatement of the reference's sweep driver (test infrastructure only).

Follows sources/common/chain.cpp (reset 175-237, performOptimizationSweep
147-163), base_chain.cpp (optimizeSite 81-110, sweepUntilConverged 112-123),
includes/nutcracker/chain.hpp (absorb 248-282, move 287-315), states.cpp
(moveSiteCursorLeft/Right 94-164) and optimizer.cpp (optimizeStateSite 201-315),
for a fixed bond dimension, on top of oracle.core.  The random initial state
of the reference (gfortran rand(), core.f95:2617-2671) is replaced by caller-
supplied raw tensors: sites 2..N are right-normalised by their polar factor and
site 1 is divided by its norm, as rand_norm / rand_unnorm do.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may use it.
"""

from __future__ import annotations

from typing import List, Sequence

import numpy as np

from . import core as oc

CPLX = np.complex128


def outside_tolerance(a, b, t):
    """utilities.hpp:418-420."""
    return ((abs(a) + abs(b)) / 2 > t) and (abs(a - b) / (abs(a) + abs(b) + t) > t)


def _regression(which, frm, to, tol):
    """optimizer.cpp:190-200."""
    if which == "SR":
        return to > frm and outside_tolerance(frm, to, tol)
    if which == "LR":
        return to < frm and outside_tolerance(frm, to, tol)
    return abs(to) < abs(frm) and outside_tolerance(abs(frm), abs(to), tol)


class OptimizerFailure(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__("optimizer failure %d %s" % (code, msg))
        self.code = code


def right_normalize(site):
    """Polar factor making sum_{r,s} A(r,l,s) conj A(r,l',s) = delta(l,l')."""
    br, bl, d = site.shape
    M = np.transpose(site, (1, 0, 2)).reshape(bl, br * d, order="F")
    info, u, s, vt = oc.mysvd(bl, br * d, M)
    P = (u @ vt).reshape(bl, br, d, order="F")
    return np.asfortranarray(np.transpose(P, (1, 0, 2)))


class OracleChain:
    def __init__(self, ops, bonds: Sequence[int], raw_sites: List[np.ndarray], site_tol=1e-12, sweep_tol=1e-12,
                 sanity=1e-12, max_iterations=10000, which="SR", rethrow=False):
        self.ops = list(ops)
        self.n = len(self.ops)
        self.bonds = list(bonds)
        self.raw = [np.asfortranarray(np.asarray(a, CPLX)) for a in raw_sites]
        self.site_tol = site_tol
        self.sweep_tol = sweep_tol
        self.sanity = sanity
        self.max_iterations = max_iterations
        self.which = which
        self.rethrow = rethrow
        self.matvecs = 0
        self.failures = 0
        self.reset()

    # -- helpers -----------------------------------------------------------
    def _dims(self, i):
        return self.bonds[i], self.bonds[i + 1], self.ops[i]

    def reset(self):
        """chain.cpp:175-237 for a fixed bond dimension and no projectors."""
        n = self.n
        self.sites = [a.copy(order="F") for a in self.raw]
        triv = np.ones((1, 1, 1), CPLX, order="F")
        self.right_envs = [None] * (n + 1)
        self.left_envs = [None] * (n + 1)
        self.right_envs[n] = triv
        for i in range(n - 1, 0, -1):
            self.sites[i] = right_normalize(self.sites[i])
            bl, br, op = self._dims(i)
            self.right_envs[i] = oc.contract_sos_right(bl, br, op.cl, op.cr, op.d, self.right_envs[i + 1], op.idx,
                                                       op.mats, self.sites[i])
        self.sites[0] = np.asfortranarray(self.sites[0] / np.linalg.norm(self.sites[0]))
        self.left_envs[0] = triv
        self.current = 0
        self.optimized = False
        e = self.expectation()
        if abs(e.imag) / abs(e) > 1e-10:
            raise RuntimeError("InitialChainEnergyNotRealError")
        self.energy = e.real

    def expectation(self, site=None):
        i = self.current
        bl, br, op = self._dims(i)
        psi = self.sites[i] if site is None else site
        return oc.compute_expectation(bl, br, op.cl, op.cr, op.d, self.left_envs[i], psi, op.idx, op.mats,
                                      self.right_envs[i + 1])

    def move_right(self):
        """chain.hpp:287-315 with moveSiteCursorRight (states.cpp:131-164)."""
        i = self.current
        bl, bm, op = self._dims(i)
        br = self.bonds[i + 2]
        info, nl, dn = oc.norm_denorm_going_right(bl, bm, br, op.d, self.ops[i + 1].d, self.sites[i], self.sites[i + 1])
        if info != 0:
            raise RuntimeError("NormalizationError %d" % info)
        dn = dn / np.linalg.norm(dn)
        self.sites[i] = nl
        self.sites[i + 1] = np.asfortranarray(dn)
        self.left_envs[i + 1] = oc.contract_sos_left(bl, bm, op.cl, op.cr, op.d, self.left_envs[i], op.idx, op.mats, nl)
        self.current = i + 1
        self.optimized = False

    def move_left(self):
        """chain.hpp:287-315 with moveSiteCursorLeft (states.cpp:94-129)."""
        i = self.current
        bm, br, op = self._dims(i)
        bll = self.bonds[i - 1]
        info, dn, nr = oc.norm_denorm_going_left(bll, bm, br, self.ops[i - 1].d, op.d, self.sites[i - 1], self.sites[i])
        if info != 0:
            raise RuntimeError("NormalizationError %d" % info)
        dn = dn / np.linalg.norm(dn)
        self.sites[i - 1] = np.asfortranarray(dn)
        self.sites[i] = nr
        self.right_envs[i] = oc.contract_sos_right(bm, br, op.cl, op.cr, op.d, self.right_envs[i + 1], op.idx,
                                                   op.mats, nr)
        self.current = i - 1
        self.optimized = False

    def optimize_site(self):
        """base_chain.cpp:81-110 around optimizeStateSite (optimizer.cpp:201-315)."""
        i = self.current
        bl, br, op = self._dims(i)
        n = bl * br * op.d
        st = {}
        info, iters, res, ev, normal = oc.optimize(bl, br, op.cl, op.cr, op.d, self.left_envs[i], op.idx, op.mats,
                                                   self.right_envs[i + 1], 0, 0, n, None, None, None, self.which,
                                                   self.site_tol, self.max_iterations, self.sites[i], stats=st)
        self.matvecs += st.get("matvecs", 0)
        status = info
        if status == 0:
            ex = self.expectation(res)
            san = self.sanity
            if (abs(ev) + abs(ex)) / 2 > san and abs(ev - ex) / (abs(ev) + abs(ex) + san) > san:
                status = 101
            elif abs(ev) > san and abs(ev.imag) / abs(ev) > san:
                status = 102
            elif normal < 1 - san:
                status = 103
        if status == 0:
            e = ev.real
            if _regression(self.which, self.energy, e, self.sanity):
                status = 104
            else:
                if (self.energy >= 0 and e >= 0) or (self.energy <= 0 and e <= 0) or \
                        outside_tolerance(abs(self.energy), abs(e), self.sanity):
                    self.energy = e
                    self.sites[i] = res
                self.optimized = True
        if status != 0:
            self.failures += 1
            if self.rethrow:
                raise OptimizerFailure(status)
        return status

    def sweep(self):
        """chain.cpp:147-163."""
        start = self.current
        if not self.optimized:
            self.optimize_site()
        while self.current + 1 < self.n:
            self.move_right()
            self.optimize_site()
        while self.current > 0:
            self.move_left()
            self.optimize_site()
        while self.current < start:
            self.move_right()
            self.optimize_site()

    def sweep_until_converged(self):
        """base_chain.cpp:112-123."""
        prev = self.energy
        self.sweep()
        cur = self.energy
        while outside_tolerance(prev, cur, self.sweep_tol):
            self.sweep()
            prev, cur = cur, self.energy
