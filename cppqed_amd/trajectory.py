"""The run loop around a device ensemble: streaming, state files and autostop.

Host mirror of ``cppqedutils::trajectory::run`` (CPPQEDutils/Trajectory.h:462-610) for the
ensembles of this package: stream a row at every interval end, append the ensemble's archive
to the state file every ``sdf`` stream points and at the end, continue from the last archive
of an existing state file, and end the run early when ``AutostopHandlerGeneric``
(Trajectory.h:179-222) reports convergence of the streamed averages.
"""
import os
from collections import deque

import numpy as np


class AutostopHandler:
    """AutostopHandlerGeneric (CPPQEDutils/Trajectory.h:179-222).

    Called with each streamed array of averages; returns True where the reference throws
    StoppingCriterionReachedException.  The first call only starts the running average (its
    array is not queued); the queue then fills to ``repetition`` arrays with their running mean;
    from there on an array whose max_i |avg_i - v_i| / (|avg_i| + |v_i|) is below ``epsilon``
    stops the run, otherwise it replaces the oldest queued array in the moving average.
    Entries below ``epsilon_abs`` in modulus are replaced by ``epsilon_abs``.  Defaults are the
    reference's (--autoStopEpsilon = epsRel default, --autoStopEpsilonAbs = epsAbs default,
    --autoStopRepetition 0 = off, Trajectory.h:110-112).
    """

    def __init__(self, epsilon=1e-6, epsilon_abs=1e-12, repetition=0):
        self.eps, self.eps_abs, self.rep = float(epsilon), float(epsilon_abs), int(repetition)
        self.avg = None
        self.queue = deque()

    def __call__(self, streamed):
        buf = np.array(streamed, dtype=np.float64, copy=True)
        buf[np.abs(buf) < self.eps_abs] = self.eps_abs
        if not self.rep:
            return False
        if self.avg is None:
            self.avg = np.zeros_like(buf)
            return False
        if len(self.queue) == self.rep:
            with np.errstate(invalid="ignore", divide="ignore"):
                r = np.abs(self.avg - buf) / (np.abs(self.avg) + np.abs(buf))
            r = r[~np.isnan(r)]
            if (r.max() if r.size else -np.inf) < self.eps:
                return True
            self.avg = self.avg + (buf - self.queue[0]) / float(self.rep)
            self.queue.popleft()
        else:
            k = float(len(self.queue))
            self.avg = (k * self.avg + buf) / (k + 1.0)
        self.queue.append(buf)
        return False


def run(ens, T=None, Dt=0.1, NDt=None, sdf=0, state_file=None, autostop=None, first_state_stream=True,
        continuing=None):
    """Runs an Ensemble in Dt mode up to time T, or for NDt intervals of Dt.

    Returns the streamed rows (time, mean dtDid, processed averages).  ``sdf``: state-dump
    frequency in stream points (0: only the end); ``state_file``: the archive file; when it
    exists and is not empty (or ``continuing`` is True) the run continues from its last archive
    and streams no row at its start (Trajectory.h:484-500, 576-585); ``autostop``: an
    AutostopHandler or a tuple (epsilon, epsilon_abs, repetition).
    """
    if (T is None) == (NDt is None):
        raise ValueError("give either T (end time) or NDt (number of intervals)")
    if autostop is not None and not isinstance(autostop, AutostopHandler):
        autostop = AutostopHandler(*autostop)
    if continuing is None:
        continuing = bool(state_file) and os.path.exists(state_file) and os.path.getsize(state_file) > 0
    if continuing:
        ens.load_state(state_file)
    t = ens.time()
    time_to_reach = T if T is not None else t + NDt * Dt
    if time_to_reach and time_to_reach <= t:
        return np.zeros((0, 2 + ens.system.n_avr))

    rows, saved, state_count = [], False, 0
    step_by_step = autostop is not None and autostop.rep > 0

    def at_stream_point(row, count):
        """State saving, then the row and the handler (Trajectory.h:569-589); True: stop."""
        nonlocal state_count, saved
        if sdf and state_file and not (state_count % sdf) and (state_count or (first_state_stream and not continuing)):
            ens.save_state(state_file)
            saved = True
        state_count += 1
        if count or not continuing:
            rows.append(row)
            if autostop is not None and autostop(row[2:]):
                return True
        return False

    # count 0: the initial stream point (a zero-length advance streams the current averages)
    stop = at_stream_point(None if continuing else ens.advance([0.0])[0], 0)
    # the remaining intervals: Dt mode up to the end time, or NDt intervals
    ivs = []
    if T is not None:
        tt = t
        while tt < time_to_reach:
            ivs.append(min(Dt, time_to_reach - tt))
            tt += ivs[-1]
    else:
        ivs = [Dt] * int(NDt)
    k, count = 0, 0
    while not stop and k < len(ivs):
        # one launch per chunk: one interval while autostop watches, else up to the next state save
        if step_by_step:
            n = 1
        elif sdf and state_file:
            r = state_count % sdf
            n = 1 if r == 0 else sdf - r + 1
        else:
            n = len(ivs) - k
        n = min(n, len(ivs) - k)
        chunk = ens.advance(ivs[k:k + n])
        k += n
        saved = False
        for j in range(n):
            count += 1
            stop = at_stream_point(chunk[j], count)
    if state_file and not saved:
        ens.save_state(state_file)
    return np.array(rows)
