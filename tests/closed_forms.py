"""Tiny gathers whose semblance / stack follow in closed form from the
physics, independent of how the search is coded.

ramp:    s_k[n] = alpha * n + beta on every trace.  Linear interpolation is
         exact on a ramp and the window is centred, so the stack at (t0, p)
         is beta + alpha * mean_k tau_k(t0, p), tau_k the traveltime in
         samples from the physical moveout formula.
const:   s_k[n] = A on every trace: semblance 1 wherever valid.
flip:    two traces +A / -A: numerator 0, semblance 0 (never a winner).
far:     traces whose moveout leaves the trace: invalid (CMP -1, CRS 0)
         once two of them drop out (err < 2 rule).
"""
import numpy as np

GEOM_DTYPE = np.dtype([("cdp", "<i4"), ("scalco", "<i4"), ("sx", "<i4"), ("sy", "<i4"), ("gx", "<i4"),
                       ("gy", "<i4")])


def line_gather(half_offsets, ns, kind="ramp", alpha=0.01, beta=0.25, amp=1.0, cdp=1, mid=(0.0, 0.0),
                azimuths=None, mids=None):
    """One CDP: trace k has half-offset half_offsets[k] (m) around the midpoint
    mids[k] (default `mid`), coordinates in cm (scalco -100)."""
    n = len(half_offsets)
    g = np.zeros(n, GEOM_DTYPE)
    for k, h in enumerate(half_offsets):
        phi = 0.0 if azimuths is None else azimuths[k]
        mx, my = mid if mids is None else mids[k]
        hx, hy = h * np.cos(phi), h * np.sin(phi)
        g[k] = (cdp, -100, round((mx - hx) * 100), round((my - hy) * 100), round((mx + hx) * 100),
                round((my + hy) * 100))
    s = np.zeros((n, ns), np.float32)
    if kind == "ramp":
        s[:] = (alpha * np.arange(ns) + beta).astype(np.float32)
    elif kind == "const":
        s[:] = amp
    elif kind == "flip":
        s[:] = amp
        s[1::2] = -amp
    return g, s


def half_offset_m(g):
    return 0.5 * np.hypot((g["gx"] - g["sx"]) / 100.0, (g["gy"] - g["sy"]) / 100.0)


def midpoint_m(g):
    return 0.5 * (g["gx"] + g["sx"]) / 100.0, 0.5 * (g["gy"] + g["sy"]) / 100.0


def cmp_tau_samples(g, t0, c, dt):
    """Traveltime in samples, t = sqrt(t0^2 + C h^2) with h the half-offset."""
    h = half_offset_m(g)
    return np.sqrt((t0 * dt) ** 2 + c * h * h) / dt


def crs_tau_samples(g, ref, t0, a, b, c, dt):
    """t^2 = (t0 + A |dm|)^2 + B |dm|^2 + C h^2, dm to the reference trace's midpoint."""
    mx, my = midpoint_m(g)
    rx, ry = midpoint_m(ref)
    dm2 = (mx - rx) ** 2 + (my - ry) ** 2
    h = half_offset_m(g)
    tt = (t0 * dt + a * np.sqrt(dm2)) ** 2 + b * dm2 + c * h * h
    return np.where(tt < 0, -1.0 / dt, np.sqrt(np.maximum(tt, 0)) / dt)


def crs_tau_samples_projected(g, ref, t0, a, b, c, dt, azimuth_deg):
    """The same with dm replaced by its signed projection onto the azimuth
    (degrees clockwise from +y): m = dx sin + dy cos, m^2 = m m."""
    mx, my = midpoint_m(g)
    rx, ry = midpoint_m(ref)
    th = np.deg2rad(azimuth_deg)
    m = (mx - rx) * np.sin(th) + (my - ry) * np.cos(th)
    h = half_offset_m(g)
    tt = (t0 * dt + a * m) ** 2 + b * m * m + c * h * h
    return np.where(tt < 0, -1.0 / dt, np.sqrt(np.maximum(tt, 0)) / dt)


def ramp_stack(taus, alpha, beta):
    return beta + alpha * float(np.mean(taus))


def valid_mask(taus, ns, tau):
    return (taus >= tau) & (taus < ns - tau - 1)
