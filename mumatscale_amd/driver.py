"""The reference's run loop around the CA step (loop.c:148-272, grain.c:715-833).

`run()` advances a subblock object (the HIP `Subblock`, or any object with
the same `step`/`download`/`grains` methods) with the reference's
termination rule: the finish time is checked every step, the solid fraction
only on ScreenPFreq multiples (loop.c:229-249), and it prints the screen line
and the final "N grains ... Average" line of output_grains (grain.c:822-826).
Steps between two checks are issued as one `step(n)` call.
"""
import math

import numpy as np

from . import abi


def _finish_step(t, dt, finish_time):
    """First timestep >= t at which timestep * dt >= finish_time (loop.c:154-155)."""
    if not finish_time:
        return None
    s = max(t, int(math.floor(finish_time / dt)) - 2)
    while s * dt < finish_time:
        s += 1
    return s


def grain_stats(gr_interior, num_grains, cell_size):
    """calculateGrainSizes + output_grains summary (grain.c:715-833)."""
    sizes = np.bincount(gr_interior[gr_interior > 0].ravel(), minlength=num_grains)[:num_grains]
    cell_volume = math.pow(cell_size * 1000, 3.0)          # mm^3
    if num_grains <= 1:
        return sizes, "0 grains."
    s = sizes[1:]
    zero = int(np.sum(s == 0))
    avg = float(np.sum(s)) / float((num_grains - 1) - zero) * cell_volume if (num_grains - 1) > zero \
        else float("nan")
    line = "%d grains.  Min: %g mm^3   Max: %g mm^3   Average:  %g mm^3" % (
        num_grains - 1, s.min() * cell_volume, s.max() * cell_volume, avg)
    return sizes, line


def run(cfg, sb, out=print, max_steps=None):
    """loop() (loop.c:186-272) without data output; returns a result dict."""
    dt = cfg.dt
    total_volume = cfg.total_non_mold_volume()
    cell_vol = math.pow(cfg.cell_size, 3.0)
    t = 0
    fs = 0.0
    rep = None
    finish = _finish_step(1, dt, cfg.finish_time)
    while True:
        targets = []
        if cfg.screenpfreq > 0:
            targets.append((t // cfg.screenpfreq + 1) * cfg.screenpfreq)
        if finish is not None:
            targets.append(max(finish, t + 1))
        if max_steps is not None:
            targets.append(max_steps)
        nxt = min(targets) if targets else t + 1000
        want = cfg.screenpfreq > 0 and nxt % cfg.screenpfreq == 0
        rep = sb.step(nxt - t, want_sums=want)
        t = nxt
        if want:
            # computeFS (loop.c:173-182): solid volume / total volume
            fs = rep.solid_sum * cell_vol / total_volume
            out("%10d %12.5e %10.4f%% %8d" % (t, t * dt, fs * 100.0, rep.num_grains - 1))
        done = False
        if cfg.finish_time and t * dt >= cfg.finish_time:
            done = True
        if fs >= cfg.fs_finish:
            done = True
        if max_steps is not None and t >= max_steps:
            done = True
        if done:
            break
    gr = sb.download(abi.FIELD_GR)[1:-1, 1:-1, 1:-1]
    num_grains = rep.num_grains
    sizes, line = grain_stats(gr, num_grains, cfg.cell_size)
    out(line)
    return {"timestep": t, "fs": fs, "num_grains": num_grains, "sizes": sizes, "average_line": line}
