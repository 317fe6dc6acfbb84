"""Seeded synthetic workloads for the five Debye-SAXS configurations.

All inputs are float32 as the engine consumes them.  Generators use numpy's
PCG64 (portable across platforms), seeds 1..5 for configurations 1..5.

Atom types: 0 C, 1 N, 2 O, 3 S, 4 H at 32/9/10/1/48 %; type 5 is the water
row of the factor table (hydration factor, never an atom type).
Form factors: f_t(q) = a_t exp(-b_t q^2) - rho0 V_t exp(-pi V_t^(2/3) (q/4pi)^2)
with a_t ~ U[1,16), b_t ~ U[0.5,20) from the seed and the excluded-volume
Gaussian of resaxslib/include/ff_coef.hpp:78-90 (rho0 = 0.334 e/A^3, V from
resaxslib/include/constants.hpp:61-69).  The water row is the reference's OH2
atomic factor (Cromer-Mann O + 2 H minus excluded volume).
"""
import numpy as np

TYPE_NAMES = ["C", "N", "O", "S", "H"]
TYPE_FRACTIONS = np.array([0.32, 0.09, 0.10, 0.01, 0.48])
EXCL_VOLUME = np.array([16.44, 2.49, 9.13, 19.86, 5.15])  # C N O S H (A^3)
RADII = np.array([1.70, 1.55, 1.52, 1.80, 1.20], dtype=np.float32)  # Bondi, C N O S H
WATER = 5
RHO0 = 0.334
DENSITY = 0.045          # atoms / A^3
PROBE = 1.4
SHELL = 3.0              # hydration shell thickness (A)
CONTRAST = 0.03          # shell contrast (e / A^3)
ELLIPSOID = (1.0, 1.3, 1.8)

# Cromer-Mann (a1..a4, c, b1..b4) of H and O, resaxslib/include/ff_coef.hpp:40,43
_CM_H = (0.4930020, 0.3229120, 0.1401910, 4.0810000e-02, 3.0380001e-03, 10.51090, 26.12570, 3.142360, 57.79970)
_CM_O = (3.048500, 2.286800, 1.546300, 0.8670000, 0.2508000, 13.27710, 5.701100, 0.3239000, 32.90890)


def type_counts(n):
    """Exact per-type counts (largest remainder) for the C/N/O/S/H mix."""
    raw = TYPE_FRACTIONS * n
    cnt = np.floor(raw).astype(int)
    rem = n - cnt.sum()
    order = np.argsort(-(raw - cnt), kind="stable")
    cnt[order[:rem]] += 1
    return cnt


def seeded_types(rng, n):
    t = np.repeat(np.arange(5, dtype=np.uint32), type_counts(n))
    rng.shuffle(t)
    return t


def _cm(coef, k2):
    return coef[4] + sum(coef[i] * np.exp(-coef[5 + i] * k2) for i in range(4))


def water_factor(q):
    """OH2 = (O + 2 H) vacuum factor minus their excluded volumes (FP64)."""
    q = np.asarray(q, dtype=np.float64)
    k2 = (q / (4 * np.pi)) ** 2
    def excl(v):
        return RHO0 * v * np.exp(-np.pi * v ** (2.0 / 3.0) * k2)
    return (_cm(_CM_O, k2) - excl(9.13)) + 2.0 * (_cm(_CM_H, k2) - excl(5.15))


def factor_table(rng, q):
    """[n_q, 6] float32 table: 5 atom types + the water row."""
    q = np.asarray(q, dtype=np.float64)
    a = rng.uniform(1.0, 16.0, size=5)
    b = rng.uniform(0.5, 20.0, size=5)
    k2 = (q / (4 * np.pi)) ** 2
    tab = np.zeros((len(q), 6))
    for t in range(5):
        v = EXCL_VOLUME[t]
        tab[:, t] = a[t] * np.exp(-b[t] * q * q) - RHO0 * v * np.exp(-np.pi * v ** (2.0 / 3.0) * k2)
    tab[:, WATER] = water_factor(q)
    return tab.astype(np.float32)


def water_weight():
    """w with w * s_i * f_w(0) = CONTRAST * shell volume of a fully exposed atom
    of mean radius (4 pi (r + p)^2 * SHELL)."""
    r_mean = float(np.dot(TYPE_FRACTIONS, RADII.astype(np.float64)))
    v_shell = 4.0 * np.pi * (r_mean + PROBE) ** 2 * SHELL
    return float(CONTRAST * v_shell / water_factor([0.0])[0])


def in_ball(rng, n, radius):
    out = np.zeros((0, 3))
    while len(out) < n:
        p = rng.uniform(-radius, radius, size=(2 * n, 3))
        p = p[(p * p).sum(1) <= radius * radius]
        out = np.concatenate([out, p])
    return out[:n]


def in_ellipsoid(rng, n, axes):
    axes = np.asarray(axes, dtype=np.float64)
    u = in_ball(rng, n, 1.0)
    return u * axes


def ellipsoid_axes(n):
    vol = n / DENSITY
    a = (vol / (4.0 / 3.0 * np.pi * np.prod(ELLIPSOID))) ** (1.0 / 3.0)
    return a * np.array(ELLIPSOID)


def morton_order(pos, bits=10):
    """Indices sorting positions along a Z-order curve (contiguous index ranges
    are then spatially compact)."""
    lo, hi = pos.min(0), pos.max(0)
    g = ((pos - lo) / np.maximum(hi - lo, 1e-9) * (2 ** bits - 1)).astype(np.uint64)
    code = np.zeros(len(pos), dtype=np.uint64)
    for b in range(bits):
        for d in range(3):
            code |= ((g[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + d)
    return np.argsort(code, kind="stable")


def bodies_of(pos, types):
    b = np.zeros((len(pos), 4), dtype=np.float32)
    b[:, :3] = np.asarray(pos, dtype=np.float32)
    b[:, 3] = np.asarray(types, dtype=np.uint32).view(np.float32)
    return b


class Workload(dict):
    __getattr__ = dict.__getitem__


def plain_ball_1888(seed=1):
    """1,888 atoms uniform in a ball of radius 18 A; 101 q in [0, 0.5]; plain."""
    rng = np.random.default_rng(seed)
    n = 1888
    pos = in_ball(rng, n, 18.0).astype(np.float32)
    types = seeded_types(rng, n)
    q = np.linspace(0.0, 0.5, 101).astype(np.float32)
    return Workload(name="plain-1888x101", pos=pos, types=types, q=q, table=factor_table(rng, q),
                    water_index=WATER, w=0.0, bodies=bodies_of(pos, types))


def _protein_like(seed, n, q):
    rng = np.random.default_rng(seed)
    pos = in_ellipsoid(rng, n, ellipsoid_axes(n)).astype(np.float32)
    types = seeded_types(rng, n)
    table = factor_table(rng, q)
    return rng, pos, types, table


def sasa_ellipsoid_20000(seed=2, n=20000):
    """20,000 atoms at 0.045 A^-3 in a 1:1.3:1.8 ellipsoid; 256 q in [0, 0.6];
    SASA hydration (256 Fibonacci dots, probe 1.4 A)."""
    q = np.linspace(0.0, 0.6, 256).astype(np.float32)
    rng, pos, types, table = _protein_like(seed, n, q)
    return Workload(name=f"sasa-hydrated-{n}x256", pos=pos, types=types, q=q, table=table, water_index=WATER,
                    w=water_weight(), radii=RADII.copy(), probe=PROBE, n_half=128, bodies=bodies_of(pos, types))


def incremental_50000(seed=3, n=50000, n_moves=2000, seg=500):
    """50,000 atoms (generator of sasa_ellipsoid_20000, Z-order sorted); moves of 500-atom
    segments: rotation <= 5 deg about the segment centroid, translation <= 1 A."""
    q = np.linspace(0.0, 0.6, 256).astype(np.float32)
    rng, pos, types, table = _protein_like(seed, n, q)
    order = morton_order(pos.astype(np.float64))
    pos, types = pos[order], types[order]
    return Workload(name=f"incremental-{n}x256", pos=pos, types=types, q=q, table=table, water_index=WATER,
                    w=water_weight(), radii=RADII.copy(), probe=PROBE, n_half=128, bodies=bodies_of(pos, types),
                    move_rng=np.random.default_rng(seed + 1000), n_moves=n_moves, seg=seg)


def random_move(rng, pos, seg):
    """One trial move: (start, new positions of [start, start+seg))."""
    n = len(pos)
    start = int(rng.integers(0, n - seg + 1))
    p = pos[start:start + seg].astype(np.float64)
    c = p.mean(0)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = np.deg2rad(rng.uniform(0.0, 5.0))
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    t = d * rng.uniform(0.0, 1.0)
    newp = (p - c) @ R.T + c + t
    return start, newp.astype(np.float32)


def ensemble_512(seed=4, n_frames=512, nmin=1000, nmax=15000):
    """512 frames, N_f = round(exp(U[ln 1000, ln 15000])), alternating ball /
    ellipsoid at 0.045 A^-3 with Gaussian-perturbed shape radii; 128 q in
    [0, 0.5]; plain."""
    rng = np.random.default_rng(seed)
    q = np.linspace(0.0, 0.5, 128).astype(np.float32)
    table = factor_table(rng, q)
    sizes = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_frames))).astype(int)
    bodies, offsets = [], [0]
    for f, m in enumerate(sizes):
        if f % 2 == 0:
            r = (m / DENSITY / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
            axes = np.array([r, r, r])
        else:
            axes = ellipsoid_axes(m)
        axes = axes * (1.0 + 0.05 * rng.normal(size=3))
        p = in_ellipsoid(rng, m, axes)
        t = seeded_types(rng, m)
        bodies.append(bodies_of(p, t))
        offsets.append(offsets[-1] + m)
    return Workload(name=f"batched-{n_frames}frames-x128", q=q, table=table, water_index=WATER, w=0.0,
                    bodies=np.concatenate(bodies), offsets=np.array(offsets, dtype=np.int32), sizes=sizes)


def fit_10000(seed=5, n=10000):
    """10,000 atoms (generator of sasa_ellipsoid_20000), 200 q in [0.01, 0.5]; target from the
    reference conformation with 1 % Gaussian noise; start displaced by N(0, 0.5 A)."""
    q = np.linspace(0.01, 0.5, 200).astype(np.float32)
    rng, pos, types, table = _protein_like(seed, n, q)
    start = (pos.astype(np.float64) + rng.normal(scale=0.5, size=pos.shape)).astype(np.float32)
    return Workload(name=f"fit-{n}x200", pos=pos, types=types, q=q, table=table, water_index=WATER, w=0.0,
                    bodies=bodies_of(pos, types), start=start, start_bodies=bodies_of(start, types),
                    noise_rng=np.random.default_rng(seed + 1000))


def pair_terms(n, nq):
    """Unordered pairs x q."""
    return n * (n - 1) / 2.0 * nq


def forward_flops(n, nq):
    """Algorithmic FP32 FLOP of one forward I(q): P (5 n_q + 9) + 2 N n_q."""
    p = n * (n - 1) / 2.0
    return p * (5 * nq + 9) + 2.0 * n * nq
