"""Host side of the time-domain TF/SF stitch (reference planar_beads/fdtd2d_core.py:1586-1863,
planar_3d/fdtd3d_core.py:1785-2084), shared by the 2D and 3D drivers.

Every window runs on the absolute time axis t = n dt.  The engine advances it in calls of
at most TD_ROWS steps; each call ends at the next step after which the host has work: a
transient frame, a peak update (every TD_CHECK steps) or a stop test (TD_CHECK steps after
the last incoming signal, then every TD_CHECK steps), so that frames, peaks and the stop
step are those of the reference's per-step loop.
"""
import numpy as np

TD_CHECK = 200          # steps between the stop tests of a time-domain window
TD_ROWS = 200           # steps per engine call (rows of the plane records / TF/SF series)


def td_next_event(n, n_in_last, frame_steps, check=TD_CHECK):
    """First step >= n after which the host acts in a time-domain window: a frame, a
    multiple of ``check`` (peak update) or a stop test ``check`` steps after the last
    incoming signal (fdtd2d_core.py:1797-1835, fdtd3d_core.py:1987-2004)."""
    e = ((n + check - 1) // check) * check
    if n > n_in_last:
        e = min(e, n_in_last + ((n - n_in_last + check - 1) // check) * check)
    else:
        e = min(e, n_in_last + check)
    later = frame_steps[frame_steps >= n]
    if len(later):
        e = min(e, int(later[0]))
    return e


def signal_span(amp, N):
    """(first, last) step whose amplitude exceeds 1e-7 of the series maximum, (N, 0) for a
    silent series (fdtd2d_core.py:1739-1743, fdtd3d_core.py:1950-1955)."""
    nz = np.nonzero(amp > 1e-7 * max(float(amp.max()) if len(amp) else 0.0, 1e-30))[0]
    return (int(nz[0]), int(nz[-1])) if len(nz) else (N, 0)
