"""Writes sasa_ties.npz: SASA inputs whose exposed-dot counts depend on the
rounding of the occlusion predicate.

The predicate (resaxslib/src/rsasa_dot_spheres.cl:122-132,
resaxslib/src/rsasa_cpu_alg.cpp:62-68) is (p + r_j)^2 - |dot - x_j|^2 > 1e-4 in
FP32, with |d|^2 = dx*dx + dy*dy + dz*dz, each product and sum rounded on its
own (the reference's CPU and oracle builds do not contract).  A compiler that
fuses the products into FMAs rounds |d|^2 differently by about one ulp, which
flips the test wherever (p + r_j)^2 - |d|^2 lies within an ulp of 1e-4.

Each case is an isolated pair of atoms (i, j), 20 A from every other pair.  Atom
j is placed so that one dot of atom i lies at a near-tie distance from it, and
its coordinates are nudged by single ulps until the unfused predicate disagrees
with one of the fused orderings below (the case's target; the targets cycle).
The unfused result is computed here with numpy FP32 operations, one rounding per
operation; the fused orderings with glibc fmaf (correctly rounded), so the file
does not depend on the oracle.

Stored: bodies [n, 4] (x, y, z, type bits), radii, probe, r2 = (probe + r)^2
per type, dots [n_types, D, 4] (64 dots per type on a Fibonacci sphere of radius
r + probe, laid out as rsasa_algorithm.cpp:96-116 lays them out; any dot set
serves, since the engine and the oracle both take the table from this file),
counts (unfused, expected), counts_fused [n_variants, n] (what each fused
ordering would give), and tie_atom / tie_dot / tie_partner / tie_target per case.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "sasa_ties.npz")

_libm = ctypes.CDLL("libm.so.6")
_libm.fmaf.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
_libm.fmaf.restype = ctypes.c_float


def fmaf(a, b, c):
    return np.float32(_libm.fmaf(float(a), float(b), float(c)))


F = np.float32


def d2_unfused(dx, dy, dz):
    s = F(dx) * F(dx)
    s = F(s + F(dy) * F(dy))
    return F(s + F(dz) * F(dz))


# fused orderings a compiler may emit for dx*dx + dy*dy + dz*dz
VARIANTS = (
    ("fma(dz,dz,fma(dy,dy,dx*dx))", lambda dx, dy, dz: fmaf(dz, dz, fmaf(dy, dy, F(dx) * F(dx)))),
    ("fma(dz,dz,fma(dx,dx,dy*dy))", lambda dx, dy, dz: fmaf(dz, dz, fmaf(dx, dx, F(dy) * F(dy)))),
    ("fma(dx,dx,fma(dy,dy,dz*dz))", lambda dx, dy, dz: fmaf(dx, dx, fmaf(dy, dy, F(dz) * F(dz)))),
    ("fma(dy,dy,dx*dx)+dz*dz", lambda dx, dy, dz: F(fmaf(dy, dy, F(dx) * F(dx)) + F(dz) * F(dz))),
    ("fma(dz,dz,dx*dx+dy*dy)", lambda dx, dy, dz: fmaf(dz, dz, F(F(dx) * F(dx) + F(dy) * F(dy)))),
)


def occluded(R2, dist2):
    return bool(F(F(R2) - F(dist2)) > F(1e-4))


def fibonacci_template(n_half, radius, probe):
    """2 n_half unit-sphere dots scaled by radius + probe (rsasa_algorithm.cpp:96-116):
    for i in [-(n-1), n]: z = i/n, rho = sqrt((n+i)(n-i))/n, theta = 2 pi i / phi."""
    n = n_half
    phi = (1.0 + np.sqrt(5.0)) / 2.0
    out = np.zeros((2 * n, 4), dtype=np.float32)
    s = F(radius) + F(probe)
    for k, i in enumerate(range(-(n - 1), n + 1)):
        z = i / n
        rho = np.sqrt((n + i) * (n - i)) / n
        th = 2.0 * np.pi * i / phi
        out[k] = (F(rho * np.sin(th)) * s, F(rho * np.cos(th)) * s, F(z) * s, F(radius))
    return out


def counts_of(bodies, R2, dots, d2fn):
    n = len(bodies)
    types = bodies[:, 3].view(np.uint32)
    D = dots.shape[1]
    out = np.zeros(n, dtype=np.uint32)
    for i in range(n):
        c = 0
        for d in range(D):
            t = dots[types[i], d]
            p = (F(t[0] + bodies[i, 0]), F(t[1] + bodies[i, 1]), F(t[2] + bodies[i, 2]))
            occ = False
            for j in range(n):
                if abs(bodies[j, 0] - bodies[i, 0]) > 12 or abs(bodies[j, 1] - bodies[i, 1]) > 12 or \
                        abs(bodies[j, 2] - bodies[i, 2]) > 12:
                    continue  # pairs sit 20 A apart: no dot of i is near another pair
                dist2 = d2fn(F(p[0] - bodies[j, 0]), F(p[1] - bodies[j, 1]), F(p[2] - bodies[j, 2]))
                if occluded(R2[types[j]], dist2):
                    occ = True
                    break
            c += 0 if occ else 1
        out[i] = c
    return out


def main():
    rng = np.random.default_rng(20240611)
    radii = np.array([1.7, 1.52, 1.55], dtype=np.float32)
    probe = F(1.4)
    n_half = 32
    dots = np.stack([fibonacci_template(n_half, r, probe) for r in radii])
    R2 = np.array([F(F(probe + r) * F(probe + r)) for r in radii], dtype=np.float32)
    D = dots.shape[1]
    n_cases = 40
    centres = [(F(-30 + 20 * (k % 4)), F(-30 + 20 * ((k // 4) % 4)), F(-10 + 20 * (k // 16))) for k in range(n_cases)]
    bodies = np.zeros((2 * n_cases, 4), dtype=np.float32)
    tie_atom, tie_dot, tie_partner, tie_target = [], [], [], []
    for k in range(n_cases):
        target = k % len(VARIANTS)
        vfn = VARIANTS[target][1]
        ti, tj = int(rng.integers(0, 3)), int(rng.integers(0, 3))
        xi = np.array(centres[k], dtype=np.float32) + rng.uniform(-0.5, 0.5, 3).astype(np.float32)
        d = int(rng.integers(0, D))
        p = np.array([F(dots[ti, d, c] + xi[c]) for c in range(3)], dtype=np.float32)
        out_dir = (p.astype(np.float64) - xi) / np.linalg.norm(p.astype(np.float64) - xi)
        rho = np.sqrt(float(R2[tj]) - 1e-4)
        found = None
        for attempt in range(200000):
            u = out_dir + rng.normal(0, 0.6, 3)
            u /= np.linalg.norm(u)
            xj = (p.astype(np.float64) + rho * u).astype(np.float32)
            for _ in range(8):
                dx, dy, dz = F(p[0] - xj[0]), F(p[1] - xj[1]), F(p[2] - xj[2])
                a = occluded(R2[tj], d2_unfused(dx, dy, dz))
                b = occluded(R2[tj], vfn(dx, dy, dz))
                if a != b:
                    found = xj.copy()
                    break
                c = int(rng.integers(0, 3))
                xj[c] = np.nextafter(xj[c], F(np.inf) if rng.random() < 0.5 else F(-np.inf))
            if found is not None:
                break
        assert found is not None, k
        bodies[2 * k, :3] = xi
        bodies[2 * k, 3] = np.array([ti], dtype=np.uint32).view(np.float32)[0]
        bodies[2 * k + 1, :3] = found
        bodies[2 * k + 1, 3] = np.array([tj], dtype=np.uint32).view(np.float32)[0]
        tie_atom.append(2 * k)
        tie_dot.append(d)
        tie_partner.append(2 * k + 1)
        tie_target.append(target)
    counts = counts_of(bodies, R2, dots, d2_unfused)
    fused = np.stack([counts_of(bodies, R2, dots, v[1]) for v in VARIANTS])
    for v in range(len(VARIANTS)):
        assert not np.array_equal(fused[v], counts), VARIANTS[v][0]
    np.savez_compressed(OUT, bodies=bodies, radii=radii, probe=np.float32(probe), dots=dots, r2=R2,
                        counts=counts, counts_fused=fused, variants=np.array([v[0] for v in VARIANTS]),
                        tie_atom=np.array(tie_atom), tie_dot=np.array(tie_dot),
                        tie_partner=np.array(tie_partner), tie_target=np.array(tie_target))
    print(OUT, "cases", n_cases, "differences per fused ordering:",
          [int(np.sum(fused[v] != counts)) for v in range(len(VARIANTS))])


if __name__ == "__main__":
    main()
