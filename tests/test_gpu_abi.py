"""GPU parity through the C ABI (libvca.so): the HIP CA step and MC sweep against the oracle.

Each golden case's inputs (terrain, spill levels, seeds) are fed to both the HIP path and the CPU
oracle; every state array, step count, stop reason and counter history must be identical (bit-exact,
integer / byte / float32 bits).  The CA result is also compared with the reference engine's own
ca_result.bin, so the HIP path is checked against the reference directly.
"""
import gzip
import json
import os
import struct

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def _cases():
    with open(os.path.join(GOLD, "manifest.json")) as f:
        return sorted(json.load(f))


@pytest.fixture(scope="module")
def ctx():
    from volcano_ca_mc_amd import native
    c = native.Context(0)
    yield c
    c.close()


def _inputs(name):
    d = os.path.join(GOLD, "cases", name)
    with open(os.path.join(d, "config.json")) as f:
        cfg = O.merged(json.load(f))
    return d, cfg


def _gpu_ca(ctx, cfg, hm, spill, vents, dense=False):
    from volcano_ca_mc_amd import native as N
    p = O.ca_params(cfg)
    params = N.CaParams(seed=p.seed, classic=bool(p.classic), max_steps=p.max_steps, eruption_steps=p.eruption_steps,
                        erupt_temp=p.erupt_temp, solidus_temp=p.solidus_temp, cooling_ambient=p.cooling_ambient,
                        cooling_lateral=p.cooling_lateral, cooling_fall=p.cooling_fall, ponding=bool(p.ponding),
                        dense=dense)
    s = cfg["space"]
    return N.LavaCA(ctx, (s["nx"], s["ny"], s["nz"]), params, cfg["neighborhood"], cfg["boundary"] == "periodic",
                    hm, spill, vents.tolist())


def _ca_cases():
    return [c for c in _cases() if not c.startswith("mcfull")]


@pytest.mark.parametrize("name", _ca_cases())
def test_ca_matches_oracle_and_reference(ctx, name):
    d, cfg = _inputs(name)
    hm = spill = None
    if cfg["ca"]["mode"] == "lava":
        hm = O.terrain_for(cfg, d)
        per = cfg["boundary"] == "periodic"
        spill = O.fill_depressions(hm, per, per)
    vents = O.vents_for(cfg, hm)
    print(f"case {name}: dims {cfg['space']}", flush=True)
    orc = O.OracleCA(O.ca_params(cfg), hm, spill, vents)
    orc.run()
    ref = orc.state()
    # front and dense mode with the quiescent fast-forward, front mode stepping every step
    for dense, ff in ((False, True), (True, True), (False, False)):
        gpu = _gpu_ca(ctx, cfg, hm, spill, vents, dense=dense)
        gpu.set_fast_forward(ff)
        steps, stop = gpu.run()
        assert (steps, stop) == (orc.steps, orc.stop_reason), (dense, ff)
        got = gpu.result()
        for k in ("type", "flow", "solid", "arrival"):
            assert np.array_equal(got[k], ref[k]), (k, dense, ff)
        assert np.array_equal(got["temp"].view(np.uint32), ref["temp"].view(np.uint32)), (dense, ff)
        assert gpu.histories() == orc.histories(), (dense, ff)
        if not ff:
            assert gpu.fast_forwarded() == (-1, 0)
        gpu.close()
    # and directly against the reference engine's file
    blob = gzip.decompress(open(os.path.join(d, "ca_result.bin.gz"), "rb").read())
    assert O.ca_result_bin(tuple(cfg["space"][k] for k in ("nx", "ny", "nz")), got) == blob


def _multi_vent_cases():
    return [c for c in _ca_cases() if int(_inputs(c)[1]["vents"]["count"]) >= 2]


@pytest.mark.parametrize("name", _multi_vent_cases())
@pytest.mark.parametrize("ids", ["shifted", "permuted"])
def test_flow_ids_beyond_16_bits(ctx, name, ids):
    """Vent flowIds above 32767 (the rules kernel then holds flowIds in 32 bits of LDS, not 16), run by the HIP path
    and by the oracle with the same ids: every state array, the step count, the stop reason and both histories must
    be identical.  "shifted" moves every id by 40000 (the ids keep their order); "permuted" gives the vents large ids
    in a different order than their list order, which changes which flow a tie between two flows picks."""
    d, cfg = _inputs(name)
    hm = spill = None
    if cfg["ca"]["mode"] == "lava":
        hm = O.terrain_for(cfg, d)
        per = cfg["boundary"] == "periodic"
        spill = O.fill_depressions(hm, per, per)
    vents = O.vents_for(cfg, hm)
    moved = vents.copy()
    if ids == "shifted":
        moved[:, 3] += 40000
    else:
        big = [2000000 - 37 * k * k - 1000 * (k % 3) for k in range(len(vents))]
        moved[:, 3] = np.array(big, dtype=moved.dtype)[::-1]
    assert moved[:, 3].min() > 32767 and len(set(moved[:, 3].tolist())) == len(vents)
    print(f"case {name}: dims {cfg['space']}, flowIds {moved[:, 3].tolist()}", flush=True)
    orc = O.OracleCA(O.ca_params(cfg), hm, spill, moved)
    orc.run()
    ref = orc.state()
    gpu = _gpu_ca(ctx, cfg, hm, spill, moved)
    assert gpu.run() == (orc.steps, orc.stop_reason)
    got = gpu.result()
    for k in ("type", "flow", "solid", "arrival"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(got["temp"].view(np.uint32), ref["temp"].view(np.uint32))
    assert gpu.histories() == orc.histories()
    assert set(np.unique(got["flow"]).tolist()) <= set(moved[:, 3].tolist()) | {0}
    gpu.close()


def _mc_cases():
    return [c for c in _cases() if "classic" not in c]


@pytest.mark.parametrize("name", _mc_cases())
def test_mc_matches_oracle(ctx, name):
    import ctypes as Cc
    from volcano_ca_mc_amd import native as N
    d, cfg = _inputs(name)
    if not cfg["mc"]["enabled"]:
        pytest.skip("no MC phase")
    s = cfg["space"]
    dims = (s["nx"], s["ny"], s["nz"])
    mc = cfg["mc"]
    full = mc["domain"] == "full"
    periodic = cfg["boundary"] == "periodic"
    per = [periodic and dims[0] > 1, periodic and dims[1] > 1, periodic and full and dims[2] > 1]
    domain = mobility = None
    if not full:
        blob = gzip.decompress(open(os.path.join(d, "ca_result.bin.gz"), "rb").read())
        n = dims[0] * dims[1] * dims[2]
        off = 20
        t = np.frombuffer(blob, np.uint8, n, off); off += n
        off += 8 * n
        sol = np.frombuffer(blob, np.uint16, n, off); off += 2 * n
        arr = np.frombuffer(blob, np.uint16, n, off)
        domain = np.empty(n, np.uint8)
        mobility = np.empty(n, np.float32)
        O.lib().orc_mobility_from_cooling(np.ascontiguousarray(t).ctypes.data, np.ascontiguousarray(sol).ctypes.data,
                                          np.ascontiguousarray(arr).ctypes.data, n, float(mc["mobilityMin"]),
                                          domain.ctypes.data, mobility.ctypes.data)
    dyn = bool(mc["mobilityFromCooling"]) and not full
    print(f"case {name}: dims {dims} Q {mc['Q']} update {mc['update']}", flush=True)
    params = O.McParams(*dims, O.NB_KINDS[cfg["neighborhood"]], (Cc.c_int * 3)(*map(int, per)), int(cfg["rngSeed"]),
                        int(mc["Q"]), float(mc["J"]), float(mc["kT"]), int(mc["update"] == "checkerboard"))
    orc = O.OracleMC(params, domain, mobility if dyn else None)
    gpu = N.PottsMC(ctx, dims, cfg["neighborhood"], per, int(cfg["rngSeed"]), int(mc["Q"]), float(mc["J"]),
                    float(mc["kT"]), mc["update"] == "checkerboard", domain, mobility, mobility_dynamics=dyn)
    assert gpu.info() == (orc.colors, orc.domain_cells)
    for k in range(int(mc["mcs"]) + 1):
        if k > 0:
            orc.sweep()
            gpu.sweep(1)
            assert gpu.counters() == orc.counters(), k
        if k % 3 == 0 or k == int(mc["mcs"]):
            so, root = orc.measure(want_root=True)
            sg = gpu.measure()
            assert (sg["grains"], sg["cells"], sg["largest"]) == (so.grains, so.cells, so.largest), k
            assert sg["energy"] == so.energy
            assert sg["mean_grain_size"] == so.mean_grain_size
            assert np.array_equal(gpu.labels(), root)
    assert np.array_equal(gpu.crystals(), orc.spins())
    # and the reference's mc_result.bin
    p = os.path.join(d, "mc_result.bin.gz")
    if os.path.exists(p):
        blob = gzip.decompress(open(p, "rb").read())
        assert O.mc_result_bin(dims, gpu.crystals()) == blob


@pytest.mark.parametrize("name", ["lava3d_moore_periodic", "scen_caldera_3d", "mcfull3d_moore_odd", "mcfull2d_vn_odd",
                                  "lava2d_von_neumann_absorbing"])
@pytest.mark.parametrize("mode", [1, 2])
def test_mc_sweep_modes_match_oracle(ctx, name, mode):
    """One launch per colour pass (mode 1) and one persistent launch per call (mode 2, a device-wide barrier between
    passes), called for 1, 2, 3 ... sweeps at a time, over the whole space (up to 27 colours on odd periodic axes):
    spins and counters equal the oracle's after every call."""
    import ctypes as Cc
    from volcano_ca_mc_amd import native as N
    d, cfg = _inputs(name)
    s = cfg["space"]
    dims = (s["nx"], s["ny"], s["nz"])
    mc = cfg["mc"]
    periodic = cfg["boundary"] == "periodic"
    per = [periodic and dims[0] > 1, periodic and dims[1] > 1, periodic and dims[2] > 1]
    print(f"case {name} mode {mode}", flush=True)
    params = O.McParams(*dims, O.NB_KINDS[cfg["neighborhood"]], (Cc.c_int * 3)(*map(int, per)), int(cfg["rngSeed"]),
                        int(mc["Q"]), float(mc["J"]), float(mc["kT"]), 1)
    orc = O.OracleMC(params, None, None)
    gpu = N.PottsMC(ctx, dims, cfg["neighborhood"], per, int(cfg["rngSeed"]), int(mc["Q"]), float(mc["J"]),
                    float(mc["kT"]), True, None, None, mobility_dynamics=False)
    gpu.set_sweep_mode(mode)
    done, k = 0, 1
    while done < 12:
        for _ in range(k):
            orc.sweep()
        gpu.sweep(k)
        done += k
        k += 1
        assert gpu.counters() == orc.counters(), done
        assert np.array_equal(gpu.crystals(), orc.spins()), done
    gpu.close()


def test_mc_persistent_many_trials_per_thread(ctx):
    """A domain whose colours hold several cells per thread of the persistent grid (full 256 x 256 x 64 space, eight
    colours of 524,288 cells): the persistent launch equals the per-colour launches, including the mobility skip."""
    from volcano_ca_mc_amd import native as N
    dims = (256, 256, 64)
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(5)
    mob = rng.random(n, dtype=np.float32)
    out = []
    for mode in (1, 2):
        print(f"persistent check mode {mode}", flush=True)
        m = N.PottsMC(ctx, dims, "pentagonal", [True, True, False], 77, 128, 1.0, 0.5, True, None, mob,
                      mobility_dynamics=True)
        m.set_sweep_mode(mode)
        m.sweep(3)
        m.sweep(4)
        out.append((m.counters(), m.crystals(), m.measure()))
        m.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_mc_from_ca_matches_host_domain(ctx):
    """vca_mc_create_from_ca derives the basalt domain and cooling mobility from the device CA state."""
    from volcano_ca_mc_amd import native as N
    d, cfg = _inputs("lava3d_moore_periodic")
    hm = O.terrain_for(cfg, d)
    spill = O.fill_depressions(hm, True, True)
    vents = O.vents_for(cfg, hm)
    gpu_ca = _gpu_ca(ctx, cfg, hm, spill, vents)
    gpu_ca.run()
    st = gpu_ca.result()
    dims = gpu_ca.dims
    mc = cfg["mc"]
    per = [True, True, False]
    a = N.PottsMC.from_ca(ctx, gpu_ca, cfg["neighborhood"], per, int(cfg["rngSeed"]), mc["Q"], mc["J"], mc["kT"],
                          True, True, mc["mobilityMin"])
    n = st["type"].size
    dom = np.empty(n, np.uint8)
    mob = np.empty(n, np.float32)
    O.lib().orc_mobility_from_cooling(st["type"].ctypes.data, st["solid"].ctypes.data, st["arrival"].ctypes.data, n,
                                      float(mc["mobilityMin"]), dom.ctypes.data, mob.ctypes.data)
    b = N.PottsMC(ctx, dims, cfg["neighborhood"], per, int(cfg["rngSeed"]), mc["Q"], mc["J"], mc["kT"], True, dom,
                  mob, True)
    a.sweep(5)
    b.sweep(5)
    assert np.array_equal(a.crystals(), b.crystals())
    assert a.counters() == b.counters()
    assert a.measure() == b.measure()


def _flow_sums_numpy(st, dims, vents):
    """flowSums (Simulation.cpp:408-466) restated with numpy: per flowId lava / basalt cells and max squared
    horizontal vent distance."""
    nx, ny, nz = dims
    n = len(vents)
    t, f = st["type"], st["flow"]
    i = np.arange(t.size, dtype=np.int64)
    x, y = i % nx, (i // nx) % ny
    ok = ((t == 2) | (t == 3)) & (f >= 1) & (f <= n)
    la, ba, d2 = np.zeros(n + 1, np.uint64), np.zeros(n + 1, np.uint64), np.zeros(n + 1, np.uint64)
    if n == 0:
        return la, ba, d2, x, y, ok
    fi = f[ok].astype(np.int64)
    la += np.bincount(fi[t[ok] == 2], minlength=n + 1).astype(np.uint64)
    ba += np.bincount(fi[t[ok] == 3], minlength=n + 1).astype(np.uint64)
    v = np.asarray(vents, dtype=np.int64)
    dd = (x[ok] - v[fi - 1, 0]) ** 2 + (y[ok] - v[fi - 1, 1]) ** 2
    np.maximum.at(d2, fi, dd.astype(np.uint64))
    return la, ba, d2, x, y, ok


def _lava_cases():
    return [c for c in _ca_cases() if _inputs(c)[1]["ca"]["mode"] == "lava"]


@pytest.mark.parametrize("name", _lava_cases())
def test_flow_stats_on_device_match_host_sums(ctx, name):
    """vca_ca_flow_stats / vca_ca_flow_points against the same sums over the downloaded final state."""
    d, cfg = _inputs(name)
    print(f"case {name}: dims {cfg['space']}", flush=True)
    hm = O.terrain_for(cfg, d)
    per = cfg["boundary"] == "periodic"
    spill = O.fill_depressions(hm, per, per)
    vents = O.vents_for(cfg, hm)
    gpu_ca = _gpu_ca(ctx, cfg, hm, spill, vents)
    gpu_ca.run()
    st = gpu_ca.result()
    vl = vents.tolist()
    la, ba, d2 = gpu_ca.flow_stats(vl)
    hla, hba, hd2, x, y, ok = _flow_sums_numpy(st, gpu_ca.dims, vl)
    assert np.array_equal(la, hla) and np.array_equal(ba, hba) and np.array_equal(d2, hd2), name
    for f in range(1, len(vl) + 1):
        if la[f] + ba[f] == 0:
            continue
        cap = int(np.sqrt(float(d2[f]))) + 2
        seen = gpu_ca.flow_points(vl[f - 1], f, int(d2[f]), cap)
        sel = ok & (st["flow"] == f)
        dx = x[sel] - vl[f - 1][0]
        dy = y[sel] - vl[f - 1][1]
        far = np.unique(np.abs(dx[dx * dx + dy * dy == int(d2[f])]))
        expect = np.zeros(cap, np.uint8)
        expect[far] = 1
        assert np.array_equal(seen, expect), (name, f)


@pytest.mark.parametrize("name", _lava_cases()[:4])
def test_flow_stats_with_many_flows(ctx, name):
    """More than 256 flowIds (sums gathered per cell, not per block): the real vents plus vents of no cells."""
    d, cfg = _inputs(name)
    print(f"case {name}: dims {cfg['space']}", flush=True)
    hm = O.terrain_for(cfg, d)
    per = cfg["boundary"] == "periodic"
    spill = O.fill_depressions(hm, per, per)
    vents = O.vents_for(cfg, hm)
    gpu_ca = _gpu_ca(ctx, cfg, hm, spill, vents)
    gpu_ca.run()
    st = gpu_ca.result()
    vl = vents.tolist() + [[k % 7, 0, 0, len(vents) + 1 + k] for k in range(300)]
    la, ba, d2 = gpu_ca.flow_stats(vl)
    hla, hba, hd2, _, _, _ = _flow_sums_numpy(st, gpu_ca.dims, vl)
    assert np.array_equal(la, hla) and np.array_equal(ba, hba) and np.array_equal(d2, hd2), name
    assert la[1:len(vents) + 1].sum() + ba[1:len(vents) + 1].sum() > 0, name


def _lava_cases():
    return [c for c in _ca_cases() if _inputs(c)[1]["ca"]["mode"] == "lava"]


def _same_state(gpu, orc, what):
    got, ref = gpu.result(), orc.state()
    assert gpu.status() == (orc.steps, orc.stop_reason), what
    for k in ("type", "flow", "solid", "arrival"):
        assert np.array_equal(got[k], ref[k]), (k, what)
    assert np.array_equal(got["temp"].view(np.uint32), ref["temp"].view(np.uint32)), what
    assert gpu.histories() == orc.histories(), what


@pytest.mark.parametrize("name", _lava_cases())
def test_fast_forward_from_any_step(ctx, name):
    """Runs split at several steps: a first run() of p steps, compared with the oracle stepped p times (a run that
    stops short of the end fast-forwards only over steps in which no cell can be born, and stops exactly at p), then
    a second run() to the end, compared with the oracle's whole run -- so the fast-forward starts from different
    steps, ends at different steps, and hands over to stepping.  State arrays (float32 bits), step count, stop reason
    and both histories must be identical to the oracle's, which steps every step."""
    d, cfg = _inputs(name)
    hm = O.terrain_for(cfg, d)
    per = cfg["boundary"] == "periodic"
    spill = O.fill_depressions(hm, per, per)
    vents = O.vents_for(cfg, hm)
    full = O.OracleCA(O.ca_params(cfg), hm, spill, vents)
    full.run()
    part = O.OracleCA(O.ca_params(cfg), hm, spill, vents)
    for p in sorted({1, full.steps // 3, full.steps // 2, max(1, full.steps - 2)}):
        while part.steps < p and part.stop_reason == 0:
            part.step()
        print(f"case {name}: first run of {p} steps, then to the end", flush=True)
        gpu = _gpu_ca(ctx, cfg, hm, spill, vents)
        gpu.run(p)
        _same_state(gpu, part, f"after {p} steps")
        gpu.run()
        _same_state(gpu, full, f"to the end after {p}")
        start, n = gpu.fast_forwarded()
        if n:
            assert 0 <= start < full.steps and n <= full.steps - start, (start, n)
        gpu.close()


@pytest.mark.parametrize("name", _lava_cases())
def test_hold_with_single_steps(ctx, name):
    """Single steps that alternate between run(1) (the hold on: a step that finds the simulation quiet holds it, and
    the next run fast-forwards from it) and enqueue(1) (the hold off: a held simulation is released and steps on),
    to the stop: state arrays (float32 bits), step count, stop reason and both histories identical to the oracle's
    whole run."""
    d, cfg = _inputs(name)
    hm = O.terrain_for(cfg, d)
    per = cfg["boundary"] == "periodic"
    spill = O.fill_depressions(hm, per, per)
    vents = O.vents_for(cfg, hm)
    full = O.OracleCA(O.ca_params(cfg), hm, spill, vents)
    full.run()
    print(f"case {name}: {full.steps} steps, run(1) and enqueue(1) in turn", flush=True)
    gpu = _gpu_ca(ctx, cfg, hm, spill, vents)
    k = 0
    while gpu.status()[1] == 0:
        assert k <= 2 * full.steps + 2, "no progress"
        if k % 2 == 0:
            gpu.run(1)
        else:
            gpu.enqueue(1)
        k += 1
    _same_state(gpu, full, "single steps")
    gpu.close()


def test_fast_forward_runs_on_quiescent_tails(ctx):
    """The fast-forward is taken: runs that end in steps where no cell can be born (the cooling tails of the NoLava
    cases, erupting vents whose lava cannot flow) report fast-forwarded steps."""
    taken = []
    for name in _lava_cases():
        d, cfg = _inputs(name)
        hm = O.terrain_for(cfg, d)
        per = cfg["boundary"] == "periodic"
        spill = O.fill_depressions(hm, per, per)
        vents = O.vents_for(cfg, hm)
        gpu = _gpu_ca(ctx, cfg, hm, spill, vents)
        steps, stop = gpu.run()
        start, n = gpu.fast_forwarded()
        gpu.close()
        if n:
            taken.append(name)
            assert 0 <= start and n <= steps - start
    print(f"fast-forwarded: {len(taken)} of {len(_lava_cases())}: {taken}", flush=True)
    assert len(taken) >= 10
