"""Parity criteria shared by the GPU tests, smoke() and bench.py.

Per-iteration cost traces: the GPU and the oracle evaluate the same FP32 formulas with
different rounding (fused multiply-add, root-to-leaf vs leaf-to-root pose composition), so
a trace is compared entry by entry with relative tolerance ``rtol`` (1e-4) while both costs
are above ``floor``; below it a residual norm is FP32 round-off (pose errors ~1e-6 give
costs ~1e-12) and only "both below the floor" is required.  LM is a branching iteration: an
accept/reject decision taken on a rounding-level difference of rho near rho_min forks the
two trajectories.  The comparison of a problem therefore stops at its first accept-decision
mismatch; such forks are counted and must stay rare (``max_fork_frac``), and the final
converged/failed classification must agree for every problem regardless.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class TraceReport:
    n: int
    compared: int  # trace entries compared
    max_rel: float
    bad: int  # entries outside tolerance
    forks: int  # problems whose accept sequence diverged
    first_fork_iters: list

    def ok(self, max_fork_frac: float) -> bool:
        return self.bad == 0 and self.forks <= max_fork_frac * self.n


def compare_traces(gpu_cost, ora_cost, gpu_acc, ora_acc, rtol: float = 1e-4, floor: float = 1e-8) -> TraceReport:
    gpu_cost = np.asarray(gpu_cost, np.float64)
    ora_cost = np.asarray(ora_cost, np.float64)
    B, T1 = ora_cost.shape
    iters = T1 - 1
    compared = bad = forks = 0
    max_rel = 0.0
    fork_at = []
    for b in range(B):
        mism = np.nonzero(gpu_acc[b] != ora_acc[b])[0]
        stop = iters if mism.size == 0 else int(mism[0])
        if mism.size:
            forks += 1
            fork_at.append(stop)
        # cost_trace[i] is the cost after i steps; entries 0..stop are produced by identical decisions
        g = gpu_cost[b, : stop + 1]
        o = ora_cost[b, : stop + 1]
        both_low = (np.abs(g) < floor) & (np.abs(o) < floor)
        rel = np.abs(g - o) / np.maximum(np.abs(o), floor)
        rel[both_low] = 0.0
        compared += int((~both_low).sum())
        bad += int((rel > rtol).sum())
        max_rel = max(max_rel, float(rel.max(initial=0.0)))
    return TraceReport(B, compared, max_rel, bad, forks, fork_at)


@dataclass
class ShadowReport:
    steps: int  # (problem, iteration) pairs checked
    norm_bad: int  # residual norms at the GPU iterate outside tolerance
    norm_max_excess: float  # largest |n_gpu - n64| / tolerance
    prop_bad: int  # proposal-norm error quantiles (max, p99) above k_f32 x the FP32 oracle's
    prop_max_excess: float
    flips_gpu: int  # accept decisions differing from the FP64 twin
    flips_f32: int  # the FP32 oracle's own flips against the FP64 twin
    worst: tuple = ()  # (step index, n_gpu, n_f32, n_f64) of the largest norm excess

    def ok(self) -> bool:
        # accept flips against FP64 happen at decisions inside the rounding floor (converged
        # problems, rho at its threshold); two FP32 evaluations each flip a random subset of
        # them, so the counts differ by counting noise as well: 3 sigma (sqrt of the FP32 count)
        return (self.norm_bad == 0 and self.prop_bad == 0
                and self.flips_gpu <= 1.25 * self.flips_f32 + 3.0 * self.flips_f32 ** 0.5 + 0.005 * self.steps)


def shadow_check(gpu_trace, gpu_acc, step32, step64, rtol: float = 1e-4, atol: float = 1e-6,
                 k_f32: float = 4.0, step64_perturbed=None) -> ShadowReport:
    """Check every GPU step from the GPU's own iterate (q_i, lam_i).

    ``step32`` / ``step64`` are the FP32 oracle's and its FP64 twin's one-step results from the
    GPU iterates, flattened over (problem, iteration) row-major.  The compared quantity is the
    residual norm |r| = sqrt(2 cost).  A GPU norm n_g passes against the FP64 value n_64 when
        |n_g - n_64| <= max(rtol * n_64 + atol, k_f32 * |n_32 - n_64|),
    i.e. within 1e-4 relative (plus the FP32 resolution atol of a pose residual) wherever the
    reference algorithm's own FP32 evaluation is, and never worse than k_f32 times that FP32
    error where the formula itself is ill-conditioned (the log map's branch cut at theta = pi,
    costs at the round-off floor).  The same test applies to the proposal's norm, which also
    carries the damped solve's conditioning.  Accept decisions: the GPU may not flip against the
    FP64 twin more often than the FP32 oracle does (x 1.25, plus three standard deviations of that
    count and 0.5 % of steps).

    ``step64_perturbed`` (optional): the FP64 step from each iterate moved at FP32 rounding scale
    (q (1 + u 2^-22), u = +-1).  Where the step is discontinuous in q -- a residual with a kink,
    such as the surface-distance collision penalty |sd| whose minimisers sit on sd = 0 -- an FP32
    evaluation lands on either side, and the FP32 oracle alone may not show it; the proposal
    yardstick then takes the larger of the FP32 oracle's and the perturbed step's deviation.
    """
    B, T1 = gpu_trace.shape
    iters = T1 - 1
    n_g = np.sqrt(2 * np.maximum(np.asarray(gpu_trace[:, :iters], np.float64).reshape(-1), 0))
    p_g = np.sqrt(2 * np.maximum(np.asarray(gpu_trace[:, 1:], np.float64).reshape(-1), 0))
    a_g = np.asarray(gpu_acc).reshape(-1)
    nrm = lambda c: np.sqrt(2 * np.maximum(np.asarray(c, np.float64), 0))  # noqa: E731
    n32, n64 = nrm(step32["cost"]), nrm(step64["cost"])
    p32, p64 = nrm(step32["cost_prop"]), nrm(step64["cost_prop"])
    a32, a64 = np.asarray(step32["acc"]), np.asarray(step64["acc"])

    tol = np.maximum(rtol * n64 + atol, k_f32 * np.abs(n32 - n64))
    ex = np.abs(n_g - n64) / tol
    # proposals: the damped solve's rounding is amplified by cond(J^T J + lam I), and one FP32
    # sample is a noisy yardstick of that amplification, so the GPU's error distribution is
    # compared with the FP32 oracle's: max and 99th percentile within k_f32 times.
    both = (a_g == 1) & (a32 == 1) & (a64 == 1)
    e_g = (np.abs(p_g - p64) / (p64 + atol))[both]
    e_32 = np.abs(p32 - p64) / (p64 + atol)
    if step64_perturbed is not None:
        pp = nrm(step64_perturbed["cost_prop"])
        e_32 = np.maximum(e_32, np.abs(pp - p64) / (p64 + atol))
    e_32 = e_32[both]
    prop_bad, prop_excess = 0, 0.0
    if e_g.size:
        for q in (100, 99):
            lim = k_f32 * np.percentile(e_32, q) + rtol
            val = np.percentile(e_g, q)
            prop_excess = max(prop_excess, val / lim)
            prop_bad += int(val > lim)
    return ShadowReport(
        steps=int(n_g.size),
        norm_bad=int((ex > 1).sum()),
        norm_max_excess=float(ex.max(initial=0.0)),
        prop_bad=prop_bad,
        prop_max_excess=float(prop_excess),
        flips_gpu=int((a_g != a64).sum()),
        flips_f32=int((a32 != a64).sum()),
        worst=(int(ex.argmax()), float(n_g[ex.argmax()]), float(n32[ex.argmax()]), float(n64[ex.argmax()]))
        if ex.size else (),
    )
