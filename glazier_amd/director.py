"""Tissue measurements from a label field: cell shape, nematic order and
disclinations, as the reference measures them (python/glazier/director.py,
bench/engines.py:68-86) to compare engines (README.md:147-182).

The definitions are kept so that a number here compares with the reference's
table:

* each cell's second-moment ellipse from its pixel coordinates as stored
  (x, y of the plane, no periodic unwrapping, so a cell cut by the lattice
  edge reads as long as the distance between its pieces, exactly as the
  reference reads it), cells of fewer than 4 pixels skipped; the major-axis
  angle theta = atan2(2 c_xy, c_xx - c_yy) / 2 in [0, pi) and the
  eccentricity |(c_xx - c_yy, 2 c_xy)| / (c_xx + c_yy);
* the population order |<exp(2 i theta)>| and its director;
* a per-pixel Q = (e/2)(cos 2theta, sin 2theta) painted from the owning
  cell, smoothed by a periodic Gaussian of width sigma, its director angle,
  and half-integer disclinations from the winding of that angle round each
  pixel plaquette (steps taken mod pi into [-pi/2, pi/2), a plaquette
  counted when |winding / 2pi| > 0.2).

A volume is measured on its middle plane.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple

import numpy as np


@dataclass
class CellShape:
    label: int
    area: int
    centre: Tuple[float, float]
    theta: float
    eccentricity: float


def mid_plane(labels: np.ndarray) -> np.ndarray:
    """The middle z-plane of a (depth, height, width) field, or the plane itself."""
    return labels[labels.shape[0] // 2] if labels.ndim == 3 else labels


def cell_shapes(labels: np.ndarray) -> List[CellShape]:
    """Second-moment ellipse of every cell of a plane (labels[y, x])."""
    plane = mid_plane(np.asarray(labels))
    flat = plane.reshape(-1).astype(np.int64)
    w = plane.shape[1]
    x = (np.arange(flat.size) % w).astype(np.float64)
    y = (np.arange(flat.size) // w).astype(np.float64)
    n_lab = int(flat.max()) + 1 if flat.size else 1
    count = np.bincount(flat, minlength=n_lab).astype(np.float64)
    sx = np.bincount(flat, x, n_lab)
    sy = np.bincount(flat, y, n_lab)
    sxx = np.bincount(flat, x * x, n_lab)
    syy = np.bincount(flat, y * y, n_lab)
    sxy = np.bincount(flat, x * y, n_lab)
    out = []
    for label in range(1, n_lab):
        n = count[label]
        if n < 4:
            continue
        mx, my = sx[label] / n, sy[label] / n
        cxx = sxx[label] / n - mx * mx
        cyy = syy[label] / n - my * my
        cxy = sxy[label] / n - mx * my
        theta = float(np.mod(0.5 * np.arctan2(2.0 * cxy, cxx - cyy), np.pi))
        trace = cxx + cyy
        ecc = float(np.hypot(cxx - cyy, 2.0 * cxy) / trace) if trace > 0 else 0.0
        out.append(CellShape(label, int(n), (float(mx), float(my)), theta, ecc))
    return out


def global_order(shapes: List[CellShape]) -> Tuple[float, float]:
    """|<exp(2 i theta)>| over the cells and the director angle in [0, pi)."""
    if not shapes:
        return 0.0, 0.0
    z = np.exp(2j * np.array([s.theta for s in shapes])).mean()
    return float(abs(z)), float(np.mod(0.5 * np.angle(z), np.pi))


def q_field(labels: np.ndarray, shapes: List[CellShape]) -> np.ndarray:
    """Per-pixel (Q_xx, Q_xy) = (e/2)(cos 2theta, sin 2theta) of the owning cell;
    zero on the medium and on cells too small to measure."""
    plane = mid_plane(np.asarray(labels))
    n_lab = int(plane.max()) + 1
    table = np.zeros((n_lab, 2))
    for s in shapes:
        table[s.label] = 0.5 * s.eccentricity * np.array([np.cos(2 * s.theta), np.sin(2 * s.theta)])
    return table[plane]


def coarse_grain(q: np.ndarray, sigma: float) -> np.ndarray:
    """Each component smoothed by a Gaussian of width sigma, periodic."""
    from scipy.ndimage import gaussian_filter

    return np.stack([gaussian_filter(q[..., c], sigma, mode="wrap") for c in range(q.shape[-1])],
                    axis=-1)


def director_angle(q: np.ndarray) -> np.ndarray:
    return 0.5 * np.arctan2(q[..., 1], q[..., 0])


def defect_charges(theta: np.ndarray) -> List[Tuple[int, int, float]]:
    """(x, y, charge) of every plaquette whose director winds by a half-integer:
    the four steps round (x, y) -> (x+1, y) -> (x+1, y+1) -> (x, y+1), each
    taken mod pi into [-pi/2, pi/2), summed and divided by 2 pi."""
    def step(d):
        return np.mod(d + 0.5 * np.pi, np.pi) - 0.5 * np.pi

    right = np.roll(theta, -1, axis=1)
    diag = np.roll(right, -1, axis=0)
    down = np.roll(theta, -1, axis=0)
    winding = step(right - theta) + step(diag - right) + step(down - diag) + step(theta - down)
    charge = winding / (2.0 * np.pi)
    ys, xs = np.nonzero(np.abs(charge) > 0.2)
    return [(int(x), int(y), float(charge[y, x])) for y, x in zip(ys, xs)]


def measure(labels: np.ndarray, target_volume: float, sigma: float = 4.0) -> dict:
    """The engine-comparison row (bench/engines.py:68-86) of one label field."""
    shapes = cell_shapes(labels)
    areas = np.array([s.area for s in shapes], dtype=np.float64)
    order, _ = global_order(shapes)
    defects = defect_charges(director_angle(coarse_grain(q_field(labels, shapes), sigma)))
    return {
        "cells": len(shapes),
        "mean_area": float(areas.mean()) if areas.size else 0.0,
        "area_rms_from_target": float(np.sqrt(((areas - target_volume) ** 2).mean())) if areas.size else 0.0,
        "mean_eccentricity": float(np.mean([s.eccentricity for s in shapes])) if shapes else 0.0,
        "nematic_order": order,
        "defects": len(defects),
        "net_charge": float(sum(c for _, _, c in defects)),
    }
