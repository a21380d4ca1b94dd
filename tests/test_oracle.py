"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Three kinds of evidence, chosen so a plausible slip (dropped term, wrong sign/index, transposed
operand) fails at least one:
  * the reference engine's own outputs on a matrix of small configs (tests/golden/cases, produced by
    tests/golden/make_golden.py from the reference sources): byte equality of every output file;
  * known answers from the reference's unit tests (engine/tests/*.cpp);
  * closed forms / brute force (SplitMix64 published vectors, mt19937-64's standard-mandated value,
    brute-force colouring validity, union-find vs BFS components).
"""
import gzip
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def case_names():
    with open(os.path.join(GOLD, "manifest.json")) as f:
        return sorted(json.load(f))


def load_case(name):
    d = os.path.join(GOLD, "cases", name)
    with open(os.path.join(d, "config.json")) as f:
        cfg = json.load(f)
    out = {}
    for fn in ("ca_result.bin", "mc_result.bin"):
        p = os.path.join(d, fn + ".gz")
        if os.path.exists(p):
            out[fn] = gzip.decompress(open(p, "rb").read())
    p = os.path.join(d, "mc_stats.csv")
    if os.path.exists(p):
        out["mc_stats.csv"] = open(p).read()
    with open(os.path.join(d, "stats.json"), encoding="utf-8") as f:
        out["stats"] = json.load(f)
    return d, cfg, out


@pytest.mark.parametrize("name", case_names())
def test_oracle_matches_reference_outputs(name):
    d, cfg, ref = load_case(name)
    r = O.run_config(cfg, d)
    if "ca_result.bin" in ref:
        assert r["ca_result_bin"] == ref["ca_result.bin"]
    if "mc_result.bin" in ref:
        assert r["mc_result_bin"] == ref["mc_result.bin"]
        assert r["mc_stats_csv"] == ref["mc_stats.csv"]
    assert r["stats"] == ref["stats"]


# ---- hash ------------------------------------------------------------------------------------

def test_splitmix64_published_vectors():
    # SplitMix64 stream from state 0 (Vigna's reference implementation): first outputs.
    L = O.lib()
    state, outs = 0, []
    for _ in range(3):
        outs.append(L.orc_splitmix64(state))
        state = (state + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_hash_prefix_decomposition():
    # engine/tests/test_rng.cpp:18-31
    L = O.lib()
    for seed in (0, 42, 2 ** 64 - 1):
        for step in (0, 1, 99):
            p = L.orc_hash_prefix(seed, step)
            for i in (0, 7, 123456789):
                h = L.orc_hash_index(p, i)
                for s in (5, 7, 0, 1):
                    salt = L.orc_salt(s)
                    assert L.orc_hash_salt(h, salt) == L.orc_hash_rand(seed, step, i, salt)


def test_mt19937_64_standard_value():
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937_64 (seed 5489).
    import ctypes as C
    g = (C.c_uint64 * 313)()
    L = O.lib()
    L.orc_mt64_seed.argtypes = [C.c_void_p, C.c_uint64]
    L.orc_mt64_next.restype = C.c_uint64
    L.orc_mt64_next.argtypes = [C.c_void_p]
    L.orc_mt64_seed(C.addressof(g), 5489)
    v = None
    for _ in range(10000):
        v = L.orc_mt64_next(C.addressof(g))
    assert v == 9981545732273789042


# ---- neighbourhoods ------------------------------------------------------------------------

def test_neighborhood_counts():
    # engine/tests/test_neighborhood.cpp: sizes and variant counts
    assert [len(v) for v in O.neighborhood(0, 10, 1, 10)] == [4]
    assert [len(v) for v in O.neighborhood(1, 10, 1, 10)] == [8]
    assert [len(v) for v in O.neighborhood(2, 10, 1, 10)] == [5, 5, 5, 5]
    assert [len(v) for v in O.neighborhood(3, 10, 10, 1)] == [6, 6]
    assert [len(v) for v in O.neighborhood(0, 5, 5, 5)] == [6]
    assert [len(v) for v in O.neighborhood(1, 5, 5, 5)] == [26]
    assert [len(v) for v in O.neighborhood(2, 5, 5, 5)] == [17] * 6
    assert O.neighborhood(3, 5, 5, 5) is None
    # 2D offsets lie in the plane of the two axes of size > 1 (x-z here)
    assert all(o[1] == 0 for v in O.neighborhood(1, 10, 1, 10) for o in v)
    # pentagonal 3D variant 0 drops the dx = -1 face, variant 5 the dz = +1 face
    v = O.neighborhood(2, 5, 5, 5)
    assert all(o[0] != -1 for o in v[0]) and all(o[2] != 1 for o in v[5])


# ---- terrain ---------------------------------------------------------------------------------

def test_fill_depressions_known_answer():
    # engine/tests/test_terrain.cpp:137-145
    h = np.array([[5, 5, 2, 2, 4, 1, 0]], np.int32)
    assert O.fill_depressions(h, False, False).tolist() == [[5, 5, 4, 4, 4, 1, 0]]


def test_fill_depressions_is_minimax_fill():
    # Closed form on random maps: the spill level of a column is the minimum over paths to an
    # outlet of the maximum height along the path (here by Bellman-Ford style relaxation).
    rng = np.random.default_rng(1)
    for trial in range(5):
        h = rng.integers(0, 9, size=(7, 9)).astype(np.int32)
        ny, nx = h.shape
        INF = 10 ** 9
        lvl = np.full(h.shape, INF)
        for y in range(ny):
            for x in range(nx):
                if x in (0, nx - 1) or y in (0, ny - 1):
                    lvl[y, x] = h[y, x]
        changed = True
        while changed:
            changed = False
            for y in range(ny):
                for x in range(nx):
                    for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                        sx, sy = x + dx, y + dy
                        if 0 <= sx < nx and 0 <= sy < ny:
                            v = max(h[y, x], lvl[sy, sx])
                            if v < lvl[y, x]:
                                lvl[y, x] = v
                                changed = True
        assert (O.fill_depressions(h, False, False) == lvl).all()


# ---- MC pieces -------------------------------------------------------------------------------

def _mc(dims, nb, per, q=10, j=1.0, kt=0.0, cb=1, seed=42, domain=None, mob=None):
    import ctypes as C
    p = O.McParams(*dims, nb, (C.c_int * 3)(*per), seed, q, j, kt, cb)
    return O.OracleMC(p, domain, mob)


def test_mobility_known_answers():
    # engine/tests/test_mc.cpp:285-330
    t = np.array([3, 3, 3, 1, 0, 0, 0, 2], np.uint8)
    arr = np.array([3, 0, 4, 0, 0, 0, 0, 8], np.uint16)
    sol = np.array([5, 10, 10, 0, 0, 0, 0, 0], np.uint16)
    dom = np.empty(8, np.uint8)
    mob = np.empty(8, np.float32)
    O.lib().orc_mobility_from_cooling(t.ctypes.data, sol.ctypes.data, arr.ctypes.data, 8, 0.0,
                                      dom.ctypes.data, mob.ctypes.data)
    assert dom.tolist() == [1, 1, 1, 0, 0, 0, 0, 1]
    assert mob[0] == pytest.approx(0.0) and mob[1] == pytest.approx(1.0) and mob[2] == pytest.approx(0.5)
    O.lib().orc_mobility_from_cooling(t.ctypes.data, sol.ctypes.data, arr.ctypes.data, 8, 0.2,
                                      dom.ctypes.data, mob.ctypes.data)
    assert mob[0] == pytest.approx(0.2) and mob[2] == pytest.approx(0.6)


def test_grains_known_answers():
    # engine/tests/test_mc.cpp:332-368
    m = np.array([2, 2, 0, 1, 0, 1, 0, 0, 1, 0, 0, 1], np.int32)
    dims = (4, 3, 1)
    assert O.grains(m, dims, False, (0, 0, 0))[0] == 5
    assert O.grains(m, dims, True, (0, 0, 0))[0] == 4
    assert O.grains(m, dims, False, (1, 0, 0))[0] == 4
    assert O.grains(m, dims, False, (1, 1, 0))[0] == 3
    c = np.full(8, 5, np.int32)
    assert O.grains(c, (2, 2, 2), False, (0, 0, 0))[0] == 1
    c[7] = 6
    g, cells, largest, _ = O.grains(c, (2, 2, 2), True, (0, 0, 0))
    assert (g, cells, largest) == (2, 8, 7)


def test_grains_vs_bfs_bruteforce():
    from collections import deque
    rng = np.random.default_rng(3)
    for dims in [(9, 8, 1), (6, 1, 7), (5, 4, 3), (2, 5, 4)]:
        for moore in (False, True):
            for per in [(0, 0, 0), (1, 1, 1)]:
                per = tuple(int(p and n > 1) for p, n in zip(per, dims))
                c = rng.integers(0, 3, size=int(np.prod(dims))).astype(np.int32)
                nx, ny, nz = dims
                offs = O.neighborhood(1 if moore else 0, nx, ny, nz)[0]
                seen = np.zeros(c.size, bool)
                comps = 0
                for s in range(c.size):
                    if c[s] <= 0 or seen[s]:
                        continue
                    comps += 1
                    q = deque([s]); seen[s] = True
                    while q:
                        i = q.popleft()
                        x, y, z = i % nx, (i // nx) % ny, i // (nx * ny)
                        for dx, dy, dz in offs:
                            p = [x + dx, y + dy, z + dz]
                            ok = True
                            for a, n in enumerate(dims):
                                if not 0 <= p[a] < n:
                                    ok = ok and bool(per[a])
                                    p[a] %= n
                            j = p[0] + nx * (p[1] + ny * p[2])
                            if ok and c[j] == c[i] and not seen[j]:
                                seen[j] = True
                                q.append(j)
                assert O.grains(c, dims, moore, per)[0] == comps


def test_energy_known_answers():
    # engine/tests/test_mc.cpp:133-145: 4x4 von Neumann, left half 1, right half 2, J = 1.5 ->
    # 4 unlike pairs (absorbing, energy 6) / 8 (periodic, energy 12)
    for per, expect in (((0, 0, 0), 6.0), ((1, 1, 0), 12.0)):
        m = _mc((4, 4, 1), 0, per, q=5, j=1.5)
        for g in range(16):
            m.set_spin(g, 2 if g % 4 >= 2 else 1)
        assert m.total_energy() == expect
    # random neighbourhood: 0.5 * J * sum of per-cell unlike counts; uniform map -> 0
    m = _mc((6, 5, 1), 2, (1, 1, 0), q=5)
    for g in range(30):
        m.set_spin(g, 3)
    assert m.total_energy() == 0.0


def test_checkerboard_colours_are_independent_sets():
    # engine/tests/test_parallel.cpp:196-237, 337-359: no two cells of one colour are neighbours in any
    # variant (also across periodic edges, odd lengths included); colour counts 2 / 4 / 8 / up to 27.
    import itertools
    for dims in [(7, 6, 5), (6, 1, 9), (9, 7, 1), (4, 4, 4), (3, 5, 7)]:
        for nb in (0, 1, 2) + ((3,) if 1 in dims else ()):
            for periodic in (0, 1):
                per = [int(periodic and n > 1) for n in dims]
                m = _mc(dims, nb, per, cb=1)
                nx, ny, nz = dims
                variants = O.neighborhood(nb, nx, ny, nz)
                colour = np.full(nx * ny * nz, -1)
                for c, cells in enumerate(m.color_lists()):
                    colour[cells] = c
                assert (colour >= 0).all()
                for x, y, z in itertools.product(range(nx), range(ny), range(nz)):
                    g = x + nx * (y + ny * z)
                    for v in variants:
                        for dx, dy, dz in v:
                            p = [x + dx, y + dy, z + dz]
                            ok = True
                            for a in range(3):
                                if not 0 <= p[a] < dims[a]:
                                    ok = ok and bool(per[a])
                                    p[a] %= dims[a]
                            j = p[0] + nx * (p[1] + ny * p[2])
                            if ok and j != g:
                                assert colour[j] != colour[g]
                odd = any(per[a] and dims[a] % 2 for a in range(3))
                expect = 2 if nb == 0 and not odd else int(np.prod(
                    [1 if n == 1 else (3 if per[a] and n % 2 else 2) for a, n in enumerate(dims)]))
                assert m.colors == expect


def test_planted_mistake_is_caught():
    """A mutated rule (lateral cooling applied to free fall) must change the 2D lava outputs,
    so the golden comparison would fail on such a slip."""
    d, cfg, ref = load_case("lava2d_moore_absorbing")
    cfg = json.loads(json.dumps(cfg))
    cfg["ca"]["coolingFall"] = cfg["ca"]["coolingLateral"]
    r = O.run_config(cfg, d)
    assert r["ca_result_bin"] != ref["ca_result.bin"]
