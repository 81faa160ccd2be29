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


def format_row(row, precision=6, groups=None):
    """One streamed row as the reference writes it (trajectory::stream, Trajectory.h:426-434, with
    FormDouble, FormDouble.h / FormDouble.cc): time and dtDid left-aligned in the width of a
    positive number of this precision (exponent digits + precision + 4 = 13 at 6 digits), then
    each element's block of averages after a tab, every average left-aligned in the width of any
    number (14), C++ default float format.  groups: averages per element (default: one block)."""
    p = max(6, precision)            # formdouble::actualPrecision
    wpos = 3 + p + 4                 # formdouble::widthPositive: int(log10(307)) + 1 exponent digits
    wany = wpos + 1                  # formdouble::widthAny
    text = "".join("%-*s" % (wpos, "%.*g" % (p, v)) for v in row[:2])
    av = list(row[2:])
    for n in (groups or [len(av)]):
        text += "\t" + "".join("%-*s" % (wany, "%.*g" % (p, v)) for v in av[:n])
        av = av[n:]
    return text + "\n"


def run(ens, T=None, Dt=0.1, NDt=None, sdf=0, state_file=None, autostop=None, first_state_stream=True,
        continuing=None, trajectory_file=None, precision=6):
    """Runs an Ensemble in Dt mode up to time T, or for NDt intervals of Dt.

    Returns the streamed rows (time, mean dtDid, processed averages).  ``sdf``: state-dump
    frequency in stream points (0: only the end); ``state_file``: the archive file (default
    ``trajectory_file + ".state"``); ``trajectory_file``: the rows are appended to it in the
    reference's text format; the run continues from the last archive of the state file when the
    trajectory file (or, without one, the state file) exists and is not empty, or when
    ``continuing`` is True, and then streams no row at its start (Trajectory.h:484-500,
    576-585); ``autostop``: an AutostopHandler or a tuple (epsilon, epsilon_abs, repetition).
    """
    if (T is None) == (NDt is None):
        raise ValueError("give either T (end time) or NDt (number of intervals)")
    if autostop is not None and not isinstance(autostop, AutostopHandler):
        autostop = AutostopHandler(*autostop)
    if trajectory_file and not state_file:
        state_file = trajectory_file + ".state"
    if continuing is None:
        probe = trajectory_file or state_file
        continuing = bool(probe) and os.path.exists(probe) and os.path.getsize(probe) > 0
    if continuing:
        ens.load_state(state_file)
    t = ens.time()
    time_to_reach = T if T is not None else t + NDt * Dt
    if time_to_reach and time_to_reach <= t:
        return np.zeros((0, 2 + ens.system.n_avr))

    out = open(trajectory_file, "a") if trajectory_file else None
    groups = [4] * (ens.system.n_avr // 4)  # every element with averages streams four
    if out:
        if continuing:
            out.write("# Continuing from time %s up to time %s\n" % ("%.*g" % (max(6, precision), t),
                                                                      "%.*g" % (max(6, precision), time_to_reach)))
        else:
            out.write("# cppqed_amd MCWF ensemble of %d trajectories, dimension %d\n\n# Key to data:\n"
                      "# Trajectory\n#  1. time\n#  2. dtDid\n#  3-%d. averages of the elements, in element order\n"
                      "# Run Trajectory up to time %s -- Stream period: %s\n\n"
                      % (ens.n, ens.system.dim, 2 + ens.system.n_avr, "%.*g" % (max(6, precision), time_to_reach),
                         "%.*g" % (max(6, precision), Dt)))
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
            if out:
                out.write(format_row(row, precision, groups))
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
    if stop and out:
        out.write("# Stopping criterion has been reached\n")
    if state_file and not saved:
        ens.save_state(state_file)
    if out:
        out.close()
    return np.array(rows)
