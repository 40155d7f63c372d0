"""The engine's own f32 exponential and logarithm, and the volume's
connectivity test, pinned outside the bit-exact comparisons.

The kernels and the emulation carry the same exp2/log2 (ghip_common.h,
oracle/emu.c): the acceptance of every float-energy model and the motility
geometric means go through them, so a wrong coefficient would pass every
device-against-emulation test. Here both are held to the C library (numpy's
f64 exp2/log2, rounded to f32) and to each other, the kernels' copy through
a host build of ghip_common.h.

Connectivity (connectivity.rs:43-84): every one of the 2^26 masks of the
volume's Moore ring is answered three ways, by the kernels' cube dilation
(host build), the emulation's breadth-first search and an independent
union-find over the ring's coordinates, and they must agree on all.
"""

import ctypes
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np
import pytest

import oracle

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def engine_host(tmp_path_factory):
    """ghip_common.h built for the host (tests/native/engine_math.cpp), with
    the flags the engine is built with."""
    out = tmp_path_factory.mktemp("native") / "engine_math.so"
    hipcc = "/opt/rocm/bin/hipcc" if Path("/opt/rocm/bin/hipcc").exists() else "hipcc"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(out),
                    str(ROOT / "tests" / "native" / "engine_math.cpp")], check=True)
    lib = ctypes.CDLL(str(out))
    for name in ("engine_exp2_f32", "engine_log2_f32"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
        getattr(lib, name).restype = None
    lib.engine_ring_table.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    lib.engine_ring_table.restype = None
    lib.engine_pick8.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    lib.engine_pick8.restype = None
    lib.engine_length.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_long]
    lib.engine_length.restype = None
    return lib


def _map(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    fn(x.ctypes.data, y.ctypes.data, x.size)
    return y


def _oracle_map(name, x):
    fn = getattr(oracle.library(), name)
    return np.array([fn(float(v)) for v in x], dtype=np.float32)


def ulps(got, exact):
    """|got - exact| in units of the f32 spacing at the correctly rounded result."""
    spacing = np.spacing(np.abs(exact.astype(np.float32))).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / spacing


def test_exp2_within_two_ulp_over_the_engines_range(engine_host):
    # acceptance: 2^(-dH log2(e) / T) for dH > 0 down to the zero cut at -125;
    # motility: 2^(mean log2 activity), activities up to the 2^20 the tests use
    x = np.concatenate([np.linspace(-125.0, 20.0, 2_000_001), np.arange(-125, 21, dtype=np.float64)])
    x = x.astype(np.float32)
    got = _map(engine_host.engine_exp2_f32, x)
    err = ulps(got, np.exp2(x.astype(np.float64)))
    assert err.max() <= 2.0, (err.max(), x[err.argmax()])
    assert np.array_equal(got[np.arange(-125, 21) + 2_000_001 + 125], np.exp2(np.arange(-125, 21)).astype(np.float32))
    sample = x[::997]
    assert np.array_equal(_oracle_map("oracle_exp2_f32", sample), _map(engine_host.engine_exp2_f32, sample))
    assert _oracle_map("oracle_exp2_f32", [-125.5, -1000.0])[0] == 0.0


def test_log2_within_two_ulp_over_the_engines_range(engine_host):
    lo, hi = np.float32(2.0 ** -30).view(np.uint32), np.float32(2.0 ** 30).view(np.uint32)
    wide = np.arange(lo, hi, 257, dtype=np.uint32).view(np.float32)
    # every f32 in [1/2, 2), where the result is small and the spacing fine
    near_one = np.arange(np.float32(0.5).view(np.uint32), np.float32(2.0).view(np.uint32),
                         dtype=np.uint32).view(np.float32)
    for x in (wide, near_one):
        got = _map(engine_host.engine_log2_f32, x)
        err = ulps(got, np.log2(x.astype(np.float64)))
        assert err.max() <= 2.0, (err.max(), x[err.argmax()])
    powers = np.float32(2.0) ** np.arange(-30, 30, dtype=np.float32)
    assert np.array_equal(_map(engine_host.engine_log2_f32, powers), np.arange(-30, 30, dtype=np.float32))
    sample = wide[::499]
    assert np.array_equal(_oracle_map("oracle_log2_f32", sample), _map(engine_host.engine_log2_f32, sample))


def test_the_bound_rejects_plausible_mistakes():
    # the same decompositions in f64, rounded once to f32: as written they
    # sit within half an ulp; a dropped series term, a lost low part of
    # 1/ln 2 or a wrong Taylor coefficient each land far past two ulp
    x = np.linspace(1.05, 1.41, 20001)
    exact = np.log2(x)

    def log2_f64(c2=0.4, inv_ln2=1.0 / np.log(2.0)):
        f = x - 1.0
        s = f / (2.0 + f)
        z = s * s
        r = z * (2.0 / 3.0 + z * (c2 + z * (2.0 / 7.0 + z * (2.0 / 9.0))))
        hfsq = 0.5 * f * f
        return ((f - hfsq + s * (hfsq + r)) * inv_ln2).astype(np.float32)

    assert ulps(log2_f64(), exact).max() <= 0.6
    assert ulps(log2_f64(c2=0.0), exact).max() > 100.0
    assert ulps(log2_f64(inv_ln2=1.44287109375), exact).max() > 100.0

    t = np.linspace(-0.5, 0.5, 20001)

    def exp2_f64(c2=0.24022651):
        coef = [1.5252734e-5, 1.5403530e-4, 1.3333558e-3, 9.6181291e-3, 5.5504109e-2, c2,
                6.9314718e-1, 1.0]
        p = np.zeros_like(t)
        for c in coef:
            p = p * t + c
        return p.astype(np.float32)

    assert ulps(exp2_f64(), np.exp2(t)).max() <= 1.0
    assert ulps(exp2_f64(c2=0.2402), np.exp2(t)).max() > 10.0


def _table(fn, ring, total, chunks=16):
    out = np.empty(total, dtype=np.uint8)
    step = (total + chunks - 1) // chunks

    def part(i):
        lo, hi = i * step, min(total, (i + 1) * step)
        if lo < hi:
            fn(ring, lo, hi, out[lo:hi].ctypes.data)

    with ThreadPoolExecutor(max_workers=8) as pool:
        list(pool.map(part, range(chunks)))
    return out


def test_planar_ring_masks_agree_both_ways():
    t = _table(lambda r, lo, hi, p: oracle.library().oracle_ring_table(r, lo, hi, ctypes.cast(p, ctypes.POINTER(ctypes.c_uint8))), 8, 256, 1)
    assert np.array_equal(t & 1, t >> 1)
    # connectivity.rs:99-132: a full ring, an opposite pair, an L of three
    assert t[0xFF] & 1 and not (t[(1 << 0) | (1 << 7)] & 1) and t[0b00001011] & 1


def test_every_volume_ring_mask_agrees_three_ways(engine_host):
    total = 1 << 26
    lib = oracle.library()
    ref = _table(lambda r, lo, hi, p: lib.oracle_ring_table(r, lo, hi, ctypes.cast(p, ctypes.POINTER(ctypes.c_uint8))), 26, total)
    dev = _table(lambda r, lo, hi, p: engine_host.engine_ring_table(lo, hi, p), 26, total)
    uf, bfs = ref & 1, ref >> 1
    assert np.array_equal(uf, bfs), np.flatnonzero(uf != bfs)[:5]
    assert np.array_equal(uf, dev), np.flatnonzero(uf != dev)[:5]
    # connectivity.rs:134-151: the sites above and below the target are two
    # pieces, joined round the side by one more; opposite corners are apart
    below, above, side = 1 << 4, 1 << 21, 1 << 13
    assert uf[below | above] == 0 and uf[below | above | side] == 1
    assert uf[(1 << 0) | (1 << 25)] == 0
    # of the 325 two-member masks, the pairs within one step are one piece
    pos = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
           if (dx, dy, dz) != (0, 0, 0)]
    for a in range(26):
        for b in range(a + 1, 26):
            near = max(abs(pos[a][i] - pos[b][i]) for i in range(3)) <= 1
            assert uf[(1 << a) | (1 << b)] == near, (a, b)


def _lowbias32(x):
    """Wellons' lowbias32 over a uint32 array, in 64-bit lanes."""
    x = x.astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & m
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & m
    x ^= x >> np.uint64(16)
    return x.astype(np.uint32)


def test_the_plain_pass_pick_is_the_hash_pick(engine_host):
    """site_pick8 takes the hash's second multiply on 24 bits; its pick must
    be the full hash's low three bits for every key and site. Random pairs,
    then every site of a 4096^2 lattice's first rows under extreme keys."""
    rng = np.random.default_rng(5)
    key = rng.integers(0, 2**32, size=1 << 22, dtype=np.uint64).astype(np.uint32)
    site = rng.integers(0, 2**32, size=1 << 22, dtype=np.uint64).astype(np.uint32)
    rows = np.arange(1 << 20, dtype=np.uint32)
    for k in (0, 1, 0x80000000, 0xFFFFFFFF, 0x9E3779B9):
        key = np.concatenate([key, np.full(rows.size, k, dtype=np.uint32)])
        site = np.concatenate([site, rows])
    got = np.empty_like(key)
    engine_host.engine_pick8(key.ctypes.data, site.ctypes.data, got.ctypes.data, key.size)
    want = _lowbias32(key ^ site) & np.uint32(7)
    assert np.array_equal(got, want)
    # and the oracle's hash agrees with the numpy restatement used here
    lib = oracle.library()
    for i in range(64):
        assert lib.oracle_site_pick(int(key[i]), int(site[i])) & 7 == int(want[i])


def _point_sets(rng):
    """Site clouds (integer coordinates about an anchor) whose covariances
    cover the eigenvalue cases: random clouds, cubes (three equal
    eigenvalues), slabs (two equal largest), rods (two equal smallest), the
    same less a site or two (nearly equal), pairs and collinear sites."""
    sets = [rng.integers(-12, 13, size=(int(rng.integers(2, 600)), 3)) for _ in range(300)]
    grid = lambda a, b, c: np.stack(np.meshgrid(np.arange(a), np.arange(b), np.arange(c), indexing="ij"),
                                    -1).reshape(-1, 3)
    cube, slab, rod = grid(8, 8, 8), grid(10, 10, 2), grid(2, 2, 20)
    sets += [cube, cube - 3, cube[1:], cube[2:], slab, slab[1:], rod, rod[1:], grid(5, 7, 11)]
    sets += [np.array([[0, 0, 0], [1, 0, 0]]), np.array([[0, 0, 0], [0, 0, 0]]),
             np.array([[0, 0, 0], [3, 4, 5], [6, 8, 10]])]
    # random ellipsoids of sites, turned by a random rotation
    for _ in range(60):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        pts = rng.normal(size=(400, 3)) * rng.uniform(0.5, 6.0, size=3) @ q.T
        sets.append(np.rint(pts).astype(np.int64))
    return sets


def _sums(pts):
    x, y, z = (pts[:, k].astype(np.int64) for k in range(3))
    return np.array([x.sum(), y.sum(), z.sum(), (x * x).sum(), (y * y).sum(), (z * z).sum(),
                     (x * y).sum(), (x * z).sum(), (y * z).sum()], dtype=np.int64)


def test_a_volumes_length_is_its_largest_covariance_eigenvalue(oracle_lib):
    # the closed form the pricing takes in a volume, 4 sqrt(lambda_max)/n of
    # the scaled covariance, against numpy's eigensolver on the same sites
    lib = oracle.library()
    P64 = ctypes.POINTER(ctypes.c_int64)
    for pts in _point_sets(np.random.default_rng(11)):
        m = _sums(pts)
        n = len(pts)
        got = lib.emu_length(m.ctypes.data_as(P64), n, 1)
        cov = np.cov(pts.T.astype(np.float64), bias=True)
        want = 4.0 * np.sqrt(max(np.linalg.eigvalsh(cov)[-1], 0.0))
        assert abs(got - want) <= 1e-9 * max(want, 1e-3), (n, got, want)


def test_the_kernels_length_is_the_emulations_bit_for_bit(engine_host, oracle_lib):
    # geometry.h built for the host against the emulation's restatement, in
    # a volume and in a plane
    lib = oracle.library()
    P64 = ctypes.POINTER(ctypes.c_int64)
    sets = _point_sets(np.random.default_rng(12))
    m = np.stack([_sums(p) for p in sets])
    counts = np.array([len(p) for p in sets], dtype=np.int64)
    for three_d in (1, 0):
        out = np.empty(len(sets))
        engine_host.engine_length(m.ctypes.data, counts.ctypes.data, three_d, out.ctypes.data, len(sets))
        want = np.array([lib.emu_length(np.ascontiguousarray(m[i]).ctypes.data_as(P64), int(counts[i]), three_d)
                         for i in range(len(sets))])
        assert np.array_equal(out, want), three_d
