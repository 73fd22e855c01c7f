"""Measured close approaches of test particles to the massive bodies of the force model (test
helper).

The parity bars of the integrator are stated for orbits without close approaches.  An orbit has
one here when, sampled every `step` days along its integration (GPU dense output in FP64, the
real particle only), it comes within max(floor, `hill` Hill radii) of a perturber: Mercury,
Venus, the Earth, the Moon, Mars, Jupiter, Saturn, Uranus, Neptune, Pluto (floor 0.1 au) or one
of the 16 asteroids (floor 0.05 au).  A body's Hill radius is r (GM / 3 GM_sun)^(1/3), r its
distance from the Sun at that time.  An asteroid passage at 0.02 au already deflects a main-belt
orbit by kilometres within a year (2 GM / (b v) ~ 7e-11 au/day for GM = 2e-15 au^3/day^2), so a
metre-level integration error there is the integrator's, at the configured epsilon, and not a
matter of the form (the FP64 oracle itself moves by metres between epsilon = 1e-9 and 1e-11 on
such orbits).
"""
import numpy as np

PERTURBERS = list(range(1, 27))  # ASSIST body indices: Mercury .. Pluto (with the Earth and the Moon), asteroids
FLOOR = np.array([0.1 if b < 11 else 0.05 for b in PERTURBERS])


def approach_ratio(ephem, t0, t1, states, params=None, step=1.0, hill=3.0, chunk=1 << 19):
    """Per system: min over the sampled times and the perturbers of distance / max(floor, hill R_H),
    and the ASSIST body of that minimum.  A ratio below 1 is a close approach."""
    from assist_amd import batch
    t0 = np.broadcast_to(np.asarray(t0, dtype=np.float64), (len(states),))
    t1 = np.broadcast_to(np.asarray(t1, dtype=np.float64), (len(states),))
    n = len(states)
    outs = []
    for i in range(n):
        span = t1[i] - t0[i]
        k = int(np.floor(abs(span) / step))
        ti = t0[i] + np.sign(span) * step * np.arange(1, k + 1)
        if k == 0 or ti[-1] != t1[i]:
            ti = np.append(ti, t1[i])
        outs.append(ti)
    reals = np.ascontiguousarray(states[:, :1])
    pr = None if params is None else np.ascontiguousarray(params[:, :1])
    r = batch.integrate_dense(ephem, t0, outs, reals, pr, config=batch.BatchConfig.default(min_dt=0.001))
    if not (r.status == 0).all():
        raise RuntimeError("dense sampling failed for %d systems" % int((r.status != 0).sum()))
    pos = np.concatenate([o[:, 0, :3] for o in r.outputs])
    tt = np.concatenate(outs)
    owner = np.repeat(np.arange(n), [len(o) for o in outs])
    times, tidx = np.unique(tt, return_inverse=True)
    nb = len(PERTURBERS)
    q, st = batch.get_particles(ephem, np.repeat(PERTURBERS, len(times)), np.tile(times, nb))
    if not (st == 0).all():
        raise RuntimeError("planet queries failed")
    sun, st = batch.get_particles(ephem, np.zeros(len(times), dtype=np.int32), times)
    assert (st == 0).all()
    px = q[:, 1:4].reshape(nb, len(times), 3)
    gm = q[:, 0].reshape(nb, len(times))[:, 0]
    rh = np.linalg.norm(px - sun[None, :, 1:4], axis=-1) * np.cbrt(gm / (3.0 * sun[0, 0]))[:, None]
    scale = np.maximum(FLOOR[:, None], hill * rh)  # (nb, ntimes)
    best = np.full(n, np.inf)
    body = np.full(n, -1)
    for lo in range(0, len(pos), chunk):
        sl = slice(lo, lo + chunk)
        d = np.linalg.norm(pos[sl, None, :] - px[:, tidx[sl]].transpose(1, 0, 2), axis=-1)
        ratio = d / scale[:, tidx[sl]].T
        rmin = ratio.min(axis=1)
        bmin = np.array(PERTURBERS)[ratio.argmin(axis=1)]
        o = owner[sl]
        # per-system minimum over this chunk's rows (rows of one system are contiguous)
        starts = np.flatnonzero(np.r_[True, o[1:] != o[:-1]])
        seg_min = np.minimum.reduceat(rmin, starts)
        seg_arg = starts + np.array([np.argmin(rmin[a:b]) for a, b in zip(starts, np.r_[starts[1:], len(o)])])
        sys_ids = o[starts]
        better = seg_min < best[sys_ids]
        best[sys_ids[better]] = seg_min[better]
        body[sys_ids[better]] = bmin[seg_arg[better]]
    return best, body


def bar_with_measured_exemptions(ephem, label, d, t0, t1, states, params, bar=1.0, max_exempt=0.05):
    """Asserts d <= bar (a scalar, or one bar per system) for every system except those with a
    measured close approach; prints them."""
    bar = np.broadcast_to(np.asarray(bar, dtype=np.float64), d.shape)
    over = np.flatnonzero(d > bar)
    ratio = np.ones(0)
    if len(over):
        ratio, body = approach_ratio(ephem, t0[over], t1[over], states[over],
                                     None if params is None else params[over])
        for i, r, b in sorted(zip(over, ratio, body), key=lambda x: -d[x[0]])[:10]:
            print("  %s system %d: |dr| %.3e m, span %.1f d, closest approach %.3f of max(floor, 3 R_H) to body %d"
                  % (label, i, d[i], t1[i] - t0[i], r, b))
        bad = over[ratio >= 1.0]
        for i in bad[:10]:
            print("  %s system %d: |dr| %.3e m beyond its bar %.2f m, span %.1f d, no close approach"
                  % (label, i, d[i], bar[i], t1[i] - t0[i]))
        assert len(bad) == 0, "%s: systems beyond their bar without a close approach: %s" % (label, bad[:20])
    covered = 1.0 - len(over) / len(d)
    print("%s: %d systems, max |dr| %.3e m, p99 %.3e m; %d beyond the bar (%.1f-%.1f m), all with measured "
          "close approaches; the bar holds for %.2f%% of the systems" % (label, len(d), d.max(), np.percentile(d, 99),
                                                                          len(over), bar.min(), bar.max(),
                                                                          100 * covered))
    assert covered >= 1.0 - max_exempt
