"""Analytical steady-state flow profiles used to validate simulations.

Same functions, signatures and output conventions as the reference
(jyulb/flow_field.pyx:20-123): the plates lie half a lattice spacing outside
the first and last lattice site, results are written into the given arrays.
"""


def poise_plate_steady_state_pgrad(H, dvisc, pgrad, centrel_dist, ana_vel):
    """Poiseuille flow between two plates driven by a pressure gradient
    (flow_field.pyx:20-49): umax = -(H/2)^2 * pgrad / (2 * dvisc)."""
    channel_h = 0.5 * H
    umax = -0.5 * channel_h * channel_h * pgrad / dvisc
    for x in range(H):
        centrel_dist[x] = (0.5 + x) - channel_h
        rel_dist = centrel_dist[x] / channel_h
        ana_vel[x] = umax * (1.0 - rel_dist * rel_dist)


def poise_plate_steady_state_umax(H, umax, centrel_dist, ana_vel):
    """Poiseuille flow between two plates for a given centre-line velocity
    (flow_field.pyx:55-81)."""
    channel_h = 0.5 * H
    for x in range(H):
        centrel_dist[x] = (0.5 + x) - channel_h
        rel_dist = centrel_dist[x] / channel_h
        ana_vel[x] = umax * (1.0 - rel_dist * rel_dist)


def couette_plate_steady_state(H, low_v1, low_v2, upp_v1, upp_v2, centrel_dist, ana_vel):
    """Couette flow between two moving plates (flow_field.pyx:87-123);
    ana_vel[x, 0:2] are the two tangential velocity components."""
    channel_h = 0.5 * H
    grad_v1 = (upp_v1 - low_v1) / float(H)
    grad_v2 = (upp_v2 - low_v2) / float(H)
    for x in range(H):
        dist_from_low = 0.5 + x
        centrel_dist[x] = dist_from_low - channel_h
        ana_vel[x, 0] = low_v1 + dist_from_low * grad_v1
        ana_vel[x, 1] = low_v2 + dist_from_low * grad_v2
