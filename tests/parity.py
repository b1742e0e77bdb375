"""Parity metric between an FP64 catalog and the oracle's.

Per catalog component c (over all receivers and distances):
    err_c = ||x_c - y_c|| / ||y_c||.
A component whose oracle norm is below 1e-8 of the largest component norm
of the same record kind is a round-off zero of the reference (e.g. the
trace terms of toroidal SH strain, u_22 + u_33 = 0); it is compared in
absolute terms instead: ||x_c - y_c|| / max_c' ||y_c'||.
"""
import numpy as np


def component_errors(x, y, floor=1e-8):
    x = np.asarray(x)
    y = np.asarray(y)
    nc = y.shape[-1]
    norms = np.array([np.linalg.norm(y[..., c]) for c in range(nc)])
    top = norms.max() if norms.size else 0.0
    errs = []
    for c in range(nc):
        d = np.linalg.norm(x[..., c] - y[..., c])
        if top == 0.0:
            errs.append(d)
        elif norms[c] >= floor * top:
            errs.append(d / norms[c])
        else:
            errs.append(d / top)
    return np.array(errs)
