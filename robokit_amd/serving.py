"""The mixed serving workload: term structure of each request kind and a request's inputs.

A trace (``workloads.config4``) mixes five kinds of requests:
  ik6 / ik7 / ik29  FrameTask(log, w 1) per end effector + PositionLimit(abs, w 1)
                    (terms/dense/frame_task.py:505-621, terms/dense/position_limit.py:508-573);
  retarget          one frame of a clip: 12 PositionTasks (w 1) + PositionLimit (w 1) +
                    SmoothnessTask (w 0.1) to the previous frame, seeded there
                    (terms/dense/position_task.py:188-272, terms/dense/smoothness_task.py:350-422);
  trajopt           a T = 64 collision-aware trajectory (the trajectory batch's own term set).
The benchmark's served leg and the served-trace tests build their requests here, so both run the
same term structures.
"""
from __future__ import annotations

from typing import Dict, Optional, Tuple

import numpy as np

from . import native as N

RT_POS_W, RT_SMOOTH_W, RT_LIMIT_W = 1.0, 0.1, 1.0
IK_KINDS = ("ik6", "ik7", "ik29")


def term_specs(trace) -> Dict[str, Tuple[object, Optional[N.TermSpec]]]:
    """(robot spec, TermSpec) per request kind; the trajectory kind has no TermSpec (its terms are
    the trajectory batch's)"""
    out = {}
    for kind in trace.kinds:
        if kind == "trajopt":
            out[kind] = (trace.pools[kind].spec, None)
        elif kind == "retarget":
            clip = trace.pools[kind][0]
            na = clip.spec.num_actuated_joints
            out[kind] = (clip.spec, N.TermSpec(pos_links=clip.links,
                                               pos_weights=np.full((len(clip.links), 3), RT_POS_W, np.float32),
                                               limit_weights=np.full(na, RT_LIMIT_W, np.float32),
                                               smooth_weights=np.full(na, RT_SMOOTH_W, np.float32)))
        else:
            wl = trace.pools[kind]
            na = wl.spec.num_actuated_joints
            out[kind] = (wl.spec, N.TermSpec(pose_links=wl.links, pose_weights=np.ones((len(wl.links), 6), np.float32),
                                             limit_weights=np.ones(na, np.float32)))
    return out


def register(server: N.Server, trace) -> Dict[str, int]:
    """register every kind of the trace with the server; returns the term id per kind"""
    return {k: (server.add_traj(trace.pools[k]) if k == "trajopt" else server.add_terms(spec, ts))
            for k, (spec, ts) in term_specs(trace).items()}


def request_inputs(trace, i: int):
    """(kind, n, q, pose_targets, pos_targets, q_prev) of request i"""
    kind = trace.kinds[int(trace.kind[i])]
    n, o = int(trace.n[i]), int(trace.off[i])
    if kind == "trajopt":
        tb = trace.pools[kind]
        return kind, n, tb.q_seed[o:o + n], tb.goals[o:o + n, None], None, None
    if kind == "retarget":
        clip, prev = trace.pools[kind]
        return kind, n, prev[o:o + n], None, clip.targets[o + 1:o + 1 + n], prev[o:o + n]
    wl = trace.pools[kind]
    return kind, n, wl.q_seed[o:o + n], wl.targets[o:o + n], None, None
