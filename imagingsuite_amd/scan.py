"""Projection angles and weights as the MuhRec engine hands them to the
back-projector: BuildFileList (SequentialScan / GoldenSectionScan / angle
list) + ComputeWeights (frameworks/tomography/Framework/ReconFramework/src/
ReconHelpers.cpp:17-177, 365-488), the 1/(0.1*pixel) factor and the 6-digit
text transport of ProjectionReader::Read (.../src/ProjectionReader.cpp:907-919).

All arithmetic is float32 in the reference's order.
"""
from __future__ import annotations

import math
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

f32 = np.float32

SEQUENTIAL = "sequential"
GOLDEN = "golden"


def sequential_angles(first: int, last: int, arc0: float, arc1: float, step: int = 1,
                      skip: Iterable[int] = ()) -> List[float]:
    """SequentialScan (ReconHelpers.cpp:77-98): angle_i = (i - first) * step_a
    + arc0 in float with step_a = (arc1 - arc0) / (last - first); an index in
    ``skip`` is stepped over (the next index takes its place)."""
    skip = set(int(s) for s in skip)
    with np.errstate(divide="ignore", invalid="ignore"):   # one projection: x/0 = inf, 0*inf = NaN
        astep = f32(f32(arc1) - f32(arc0)) / f32(last - first)
        out = []
        i = first
        while i <= last:
            while i in skip:
                i += 1
            out.append(float(f32(i - first) * astep + f32(arc0)))
            i += step
    return out


def golden_angles(first: int, last: int, arc: float, start: int = 0,
                  skip: Iterable[int] = ()) -> List[float]:
    """GoldenSectionScan (ReconHelpers.cpp:100-155): angle_i =
    fmod((i - start - skipped) * phi * arc, arc) in float, phi = (1 + sqrt 5)/2,
    for i = first .. last + skipped; arc must be 180 or 360."""
    if f32(arc) not in (f32(180.0), f32(360.0)):
        raise ValueError("The golden ratio reconstruction requires arc to be 180 or 360 degrees")
    skip = set(int(s) for s in skip)
    phi = f32(0.5) * (f32(1.0) + np.sqrt(f32(5.0), dtype=np.float32))
    n_skip = 0
    i = 0
    while i < first:       # skipped indices before the first one (the loop steps twice on a hit)
        if i in skip:
            i += 1
            n_skip += 1
        i += 1
    out = []
    i = first
    while i <= last + n_skip:
        while i in skip:
            i += 1
            n_skip += 1
        k = i - start - n_skip
        if k < 0:          # size_t arithmetic in the reference
            k += 1 << 64
        out.append(float(f32(math.fmod(float(f32(k) * phi * f32(arc)), float(f32(arc))))))
        i += 1
    return out


def weights(angles: Sequence[float], arc0: float, arc1: float,
            golden: bool = False) -> Tuple[List[float], List[float]]:
    """ComputeWeights: returns (angles, weights) in ascending angle order."""
    n = len(angles)
    keys = [float(f32(math.fmod(a, 180.0))) for a in angles]
    order = sorted(range(n), key=lambda i: (keys[i], i))
    K = [f32(keys[i]) for i in order]
    A = [angles[i] for i in order]
    if n < 3:
        # the list is keyed by fmod(angle, 180) here (the map the reader walks
        # in key order); std::map::insert keeps the first of equal keys
        d = {}
        for i in range(n):
            if float(K[i]) not in d:
                d[float(K[i])] = (A[i], float(f32(1.0 / n)))
        items = [d[k] for k in sorted(d)]
        return [a for a, _ in items], [w for _, w in items]
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
    if abs(tot - f32(0.5)) < f32(0.001):
        out = {a: w * f32(2) for a, w in out.items()}
    keys_sorted = sorted(out)
    return keys_sorted, [float(out[a]) for a in keys_sorted]


def _text6(x) -> np.float32:
    return f32(float("%g" % float(x)))


def engine_angles_weights(n_proj: int, arc0: float, arc1: float, pixel_mm: float,
                          rotation: float = 0.0, first: int = 0, scan: str = SEQUENTIAL,
                          golden_start: int = 0, step: int = 1, skip: Iterable[int] = (),
                          angles: Optional[Sequence[float]] = None):
    """What the back-projector receives: (angles, weights) float32 arrays in
    ascending angle order, both through the 6-digit text transport.

    ``scan`` = "sequential" (projections first .. first + n_proj - 1 over
    [arc0, arc1], every ``step``-th, indices in ``skip`` stepped over) or
    "golden" (GoldenSectionScan over arc1 in {180, 360}, index origin
    ``golden_start``); ``angles`` (acquisition order) replaces the generated
    list, as a list file does (the reference then computes the weights with the
    sequential rule).  The weight factor is 1/(0.1 * pixel_mm) (1 when
    pixel_mm <= 0) and ``rotation`` (MatrixInfo.fRotation) is added to every
    angle."""
    last = first + n_proj - 1
    if angles is not None:
        ang = [float(f32(a)) for a in angles]
        golden = False
    elif scan == SEQUENTIAL:
        ang = sequential_angles(first, last, arc0, arc1, step, skip)
        golden = False
    elif scan == GOLDEN:
        ang = golden_angles(first, last, arc1, golden_start, skip)
        golden = True
    else:
        raise ValueError(f"unknown scan type {scan!r}")
    a, w = weights(ang, arc0, arc1, golden)
    rw = f32(1.0) / (f32(pixel_mm) * f32(0.1) if pixel_mm > 0 else f32(1.0))
    out_a = np.array([_text6(f32(x) + f32(rotation)) for x in a], np.float32)
    out_w = np.array([_text6(f32(x) * rw) for x in w], np.float32)
    return out_a, out_w
