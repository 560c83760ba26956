"""Projection angles and weights as the MuhRec engine hands them to the
back-projector: BuildFileList (SequentialScan / GoldenSectionScan) +
ComputeWeights (frameworks/tomography/Framework/ReconFramework/src/
ReconHelpers.cpp:17-177, 365-488), the 1/(0.1*pixel) factor and the 6-digit
text transport of ProjectionReader::Read (.../src/ProjectionReader.cpp:907-919).

All arithmetic is float32 in the reference's order.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

f32 = np.float32


def sequential_angles(first: int, last: int, arc0: float, arc1: float, step: int = 1) -> List[float]:
    astep = f32(f32(arc1) - f32(arc0)) / f32(last - first)
    return [float(f32(i - first) * astep + f32(arc0)) for i in range(first, last + 1, step)]


def golden_angles(first: int, last: int, arc: float, start: int = 0) -> List[float]:
    if arc not in (180.0, 360.0):
        raise ValueError("The golden ratio reconstruction requires arc to be 180 or 360 degrees")
    phi = f32(0.5) * (f32(1.0) + np.sqrt(f32(5.0), dtype=np.float32))
    return [float(f32(math.fmod(float(f32(i - start) * phi * f32(arc)), arc)))
            for i in range(first, last + 1)]


def weights(angles: Sequence[float], arc0: float, arc1: float,
            golden: bool = False) -> Tuple[List[float], List[float]]:
    """ComputeWeights: returns (angles, weights) in ascending angle order."""
    n = len(angles)
    keys = [float(f32(math.fmod(a, 180.0))) for a in angles]
    order = sorted(range(n), key=lambda i: (keys[i], i))
    K = [f32(keys[i]) for i in order]
    A = [angles[i] for i in order]
    if n < 3:
        d = {A[i]: float(f32(1.0) / f32(n)) for i in range(n)}
        return sorted(d), [d[a] for a in sorted(d)]
    W = [f32(0.0)] * n
    for q in range(1, n - 1):
        W[q] = (K[q + 1] - K[q - 1]) / f32(360.0) if K[q] != K[q - 1] else f32(0.0)
    if f32(175.0) < f32(arc1) - f32(arc0):
        rep_last = K[n - 1] == K[n - 2]
        W[n - 1] = f32(0.0) if rep_last else ((K[0] + f32(180)) - K[n - 2]) / f32(360.0)
        num = K[1] - (K[n - 1] - f32(180))
        W[0] = num / f32(360.0) if (rep_last or golden) else num / f32(180.0)
    else:
        W[0] = W[1]
        W[n - 1] = W[1]
    q00, q11 = 0, 1
    while q11 < n:
        if W[q11] == 0.0:
            stack = []
            if q11 != q00:
                q00 = q11 - 1
            while True:
                stack.append(q11)
                q11 += 1
                if not (q11 < n and W[q11] == 0.0):
                    break
            share = W[q00] / f32(len(stack) + 1)
            W[q00] = share
            for s in stack:
                W[s] = share
        else:
            q11 += 1
    out = {}
    tot = f32(0.0)
    for i in range(n):
        out[A[i]] = W[i]
        tot = tot + W[i]
    if abs(float(tot) - 0.5) < 0.001:
        out = {a: w * f32(2) for a, w in out.items()}
    keys_sorted = sorted(out)
    return keys_sorted, [float(out[a]) for a in keys_sorted]


def _text6(x) -> np.float32:
    return f32(float("%g" % float(x)))


def engine_angles_weights(n_proj: int, arc0: float, arc1: float, pixel_mm: float,
                          rotation: float = 0.0, first: int = 0):
    """What the back-projector receives for a sequential scan of ``n_proj``
    projections over [arc0, arc1]: (angles, weights) float32 arrays."""
    ang = sequential_angles(first, first + n_proj - 1, arc0, arc1)
    a, w = weights(ang, arc0, arc1)
    rw = f32(1.0) / (f32(pixel_mm) * f32(0.1) if pixel_mm > 0 else f32(1.0))
    angles = np.array([_text6(f32(x) + f32(rotation)) for x in a], np.float32)
    wts = np.array([_text6(f32(x) * rw) for x in w], np.float32)
    return angles, wts
