#!/usr/bin/env python3
"""Regenerate the Dunavant (1985) symmetric triangle rules 1..20 as float32 data.

D. A. Dunavant, "High degree efficient symmetrical Gaussian quadrature rules
for the triangle", Int. J. Numer. Meth. Eng. 21 (1985) 1129-1148.

Each rule is a set of S3-symmetric orbits (centroid, 3-point (a,a,b),
6-point (a,b,c)) with one weight per orbit. The orbit barycentric
coordinates and weights are the roots of the rule's moment equations
    sum_o w_o * sum_{points of o} x^i y^j = 2 * i! j! / (i+j+2)!,  i+j <= degree
(weights normalised to sum 1 over the unit triangle of area 1/2). This script
checks the published orbit values (read as numbers from the reference's rule
table, which prints them to 15 digits) against those equations, refining them
by Gauss-Newton only where a residual exceeds 1e-13, so that every rule
integrates all monomials of its degree to 1e-13, and writes the float32 rounding of (xi, eta, w) that the
solver's kernels consume, in the point order
    centroid -> (a,b); 3-orbit -> (a,b),(b,c),(c,a);
    6-orbit  -> (a,b),(b,c),(c,a),(b,a),(c,b),(a,c)
so that xi multiplies vertex 1 and eta vertex 2 (mesh_geometry.cuh:211-222).

Usage: gen_dunavant.py <reference triangle_dunavant_rule.cpp> <out header>
The output header is committed; this script only runs when the table is
regenerated.
"""
import math
import re
import sys

import numpy as np


def parse_seeds(text):
    seeds = {}
    for rule in range(1, 21):
        tag = f"{rule:02d}"
        m_xyz = re.search(r"suborder_xy_rule_%s\[[^\]]*\]\s*=\s*\{([^}]*)\}" % tag, text)
        m_w = re.search(r"suborder_w_rule_%s\[[^\]]*\]\s*=\s*\{([^}]*)\}" % tag, text)
        xyz = [float(t) for t in m_xyz.group(1).replace("\n", " ").split(",") if t.strip()]
        w = [float(t) for t in m_w.group(1).replace("\n", " ").split(",") if t.strip()]
        orbits = []
        for s in range(len(w)):
            a, b, c = xyz[3 * s: 3 * s + 3]
            orbits.append(((a, b, c), w[s]))
        seeds[rule] = orbits
    return seeds


def orbit_kind(abc, tol=1e-12):
    a, b, c = abc
    if abs(a - b) < tol and abs(b - c) < tol:
        return 1
    if abs(a - b) < tol or abs(b - c) < tol or abs(a - c) < tol:
        return 3
    return 6


def expand(kind, a, b, c):
    if kind == 1:
        return [(a, b)]
    if kind == 3:
        return [(a, b), (b, c), (c, a)]
    return [(a, b), (b, c), (c, a), (b, a), (c, b), (a, c)]


def pack(orbits):
    """Free parameters: 3-orbits (a,a,1-2a) -> a ; 6-orbits -> a, b ; weights."""
    kinds, x = [], []
    for (abc, w) in orbits:
        k = orbit_kind(abc)
        kinds.append(k)
        a, b, c = abc
        if k == 3:
            # the repeated coordinate
            rep = a if abs(a - b) < 1e-12 or abs(a - c) < 1e-12 else b
            x.append(rep)
        elif k == 6:
            x.extend([a, b])
        x.append(w)
    return kinds, np.array(x)


def unpack(kinds, x, seeds):
    out, p = [], 0
    for k, (abc, _) in zip(kinds, seeds):
        if k == 1:
            a = b = c = 1.0 / 3.0
        elif k == 3:
            r = x[p]; p += 1
            a0, b0, c0 = abc
            # keep the seed's placement of the repeated coordinate
            if abs(a0 - b0) < 1e-12:
                a, b, c = r, r, 1.0 - 2.0 * r
            elif abs(b0 - c0) < 1e-12:
                a, b, c = 1.0 - 2.0 * r, r, r
            else:
                a, b, c = r, 1.0 - 2.0 * r, r
        else:
            a, b = x[p], x[p + 1]; p += 2
            c = 1.0 - a - b
        w = x[p]; p += 1
        out.append(((a, b, c), w, k))
    return out


def monomials(degree):
    return [(i, j) for i in range(degree + 1) for j in range(degree + 1 - i)]


def exact(i, j):
    # integral over the unit right triangle divided by its area 1/2
    return 2.0 * math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2)


def residual(kinds, x, seeds, degree):
    orb = unpack(kinds, x, seeds)
    res = []
    for (i, j) in monomials(degree):
        s = 0.0
        for (abc, w, k) in orb:
            for (xi, eta) in expand(k, *abc):
                s += w * xi ** i * eta ** j
        res.append(s - exact(i, j))
    return np.array(res)


def refine(kinds, x0, seeds, degree):
    """Gauss-Newton on the moment equations; a step is kept only if it lowers
    the largest moment residual, so the result is never worse than the seed."""
    x = x0.copy()
    best = np.max(np.abs(residual(kinds, x, seeds, degree)))
    for _ in range(8):
        r = residual(kinds, x, seeds, degree)
        J = np.empty((r.size, x.size))
        for c in range(x.size):
            h = 1e-6 * max(1.0, abs(x[c]))
            xp = x.copy(); xp[c] += h
            xm = x.copy(); xm[c] -= h
            J[:, c] = (residual(kinds, xp, seeds, degree) - residual(kinds, xm, seeds, degree)) / (2 * h)
        dx, *_ = np.linalg.lstsq(J, -r, rcond=1e-10)
        cand = x + dx
        err = np.max(np.abs(residual(kinds, cand, seeds, degree)))
        if not err < best:
            break
        x, best = cand, err
    return x


def float_hex(v):
    return float.hex(float(np.float32(v)))


def main():
    text = open(sys.argv[1]).read()
    seeds = parse_seeds(text)
    lines = [
        "// Generated by gen_dunavant.py: Dunavant (1985) symmetric triangle rules 1..20,",
        "// checked against their moment equations and rounded to float32.",
        "// Point p of rule r: (xi, eta) = kDunavantXY[offset + 2p .. +1], weight kDunavantW[offset_w + p].",
        "#pragma once",
        "",
    ]
    orders, xy_all, w_all = [], [], []
    for rule in range(1, 21):
        kinds, x0 = pack(seeds[rule])
        # The published 15-digit values already satisfy the moment equations to
        # ~1e-15; Gauss-Newton only runs when they do not. Keeping the published
        # digits keeps their float32 rounding identical to the reference solver's.
        x = x0
        if np.max(np.abs(residual(kinds, x0, seeds[rule], rule))) > 1e-13:
            x = refine(kinds, x0, seeds[rule], rule)
        r = residual(kinds, x, seeds[rule], rule)
        if np.max(np.abs(r)) > 1e-13:
            raise SystemExit(f"rule {rule}: moment residual {np.max(np.abs(r))}")
        orb = unpack(kinds, x, seeds[rule])
        n = 0
        for (abc, w, k) in orb:
            for (xi, eta) in expand(k, *abc):
                xy_all += [xi, eta]
                w_all.append(w)
                n += 1
        orders.append(n)
    offs = np.concatenate([[0], np.cumsum(orders)])
    lines.append("static const int kDunavantOrder[21] = {0, %s};" % ", ".join(str(o) for o in orders))
    lines.append("static const int kDunavantOffset[21] = {0, %s};" % ", ".join(str(int(o)) for o in offs[:-1]))
    lines.append("static const float kDunavantXY[%d] = {" % len(xy_all))
    for p in range(0, len(xy_all), 4):
        lines.append("    " + ", ".join(float_hex(v) + "f" for v in xy_all[p:p + 4]) + ",")
    lines.append("};")
    lines.append("static const float kDunavantW[%d] = {" % len(w_all))
    for p in range(0, len(w_all), 4):
        lines.append("    " + ", ".join(float_hex(v) + "f" for v in w_all[p:p + 4]) + ",")
    lines.append("};")
    open(sys.argv[2], "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
