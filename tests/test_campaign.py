"""Campaign driver state (ccs_campaign_*, ccs_parse_prefix): the v13 main loop's tile walk,
stop rules and CC18_CUDA_v10 checkpoint (v13.cu:520-664, :3584-3667, :3719-3765), checked
on the CPU through the C ABI over explicit tile ranges."""
import os

import numpy as np
import pytest

import ccs

M64 = (1 << 64) - 1


def xorshift64(s):   # the reference's rng_next (v13.cu:714-721)
    s ^= (s << 13) & M64
    s ^= s >> 7
    s ^= (s << 17) & M64
    return s


def drain(c, record=True):
    out = []
    while True:
        b = c.next()
        if b is None:
            return out
        out.append(b)
        if record:
            c.record(*b, candidates=7 * b[1], survivors=b[1] // 3)


def test_sequential_prefix_sweep_covers_range_once_then_stops():
    c = ccs.Campaign(tile_range=(10 ** 20, 10 ** 20 + 1000), prefix="0b1", walk="sequential",
                     batch_tiles=64, seed=3)
    b = drain(c)
    assert b[0][0] == 10 ** 20 and all(b[i][0] + b[i][1] == b[i + 1][0] for i in range(len(b) - 1))
    assert b[-1][0] + b[-1][1] == 10 ** 20 + 1000 and b[-1][1] == 1000 % 64
    st = c.state()
    assert st["done"] == 1 and st["total_tiles"] == 1000 and st["total_candidates"] == 7000
    assert c.next() is None


def test_sequential_without_prefix_wraps_and_tiles_limit():
    c = ccs.Campaign(tile_range=(100, 250), start=200, batch_tiles=40, max_tiles=200, seed=1)
    b = drain(c)
    assert [x[0] for x in b] == [200, 240, 100, 140, 180]      # clamp at the end, wrap to the start
    assert [x[1] for x in b] == [40, 10, 40, 40, 40]
    assert c.state()["random_mode"] == 0


def test_random_walk_is_the_seeded_xorshift64():
    lo, hi, batch, seed = 5 * 10 ** 18, 5 * 10 ** 18 + 123457, 1000, 0xDEADBEEF
    c = ccs.Campaign(tile_range=(lo, hi), seed=seed, batch_tiles=batch, max_tiles=20 * batch)
    got = [x[0] for x in drain(c)]
    s, want = seed, []
    for _ in range(20):
        s = xorshift64(s)
        h = s
        s = xorshift64(s)
        want.append(lo + ((h << 64) | s) % (hi - lo - batch))
    assert got == want and c.state()["random_mode"] == 1
    again = ccs.Campaign(tile_range=(lo, hi), seed=seed, batch_tiles=batch, max_tiles=20 * batch)
    assert [x[0] for x in drain(again)] == got


def test_mode_rules():
    assert ccs.Campaign(tile_range=(0, 100), seed=1).state()["random_mode"] == 1            # default
    assert ccs.Campaign(tile_range=(0, 100), start=5, seed=1).state()["random_mode"] == 0   # --start
    c = ccs.Campaign(tile_range=(0, 100), start=5, prefix="0b1", walk="random", seed=1)
    assert c.state()["random_mode"] == 1                                                    # --prefix-mode wins
    with pytest.raises(ccs.CcsError):
        ccs.Campaign(tile_range=(0, 100), walk="sequential")                    # mode without prefix
    with pytest.raises(ccs.CcsError):
        ccs.Campaign(tile_range=(10, 100), prefix="0b1", start=5)               # start outside prefix range
    with pytest.raises(ccs.CcsError):
        ccs.Campaign(tile_range=(0, 100), batch_tiles=0)
    with pytest.raises(ccs.CcsError):
        ccs.Campaign(tile_range=(5, 5))


def test_pipelined_records_and_checkpoint_round_trip(tmp_path):
    path = tmp_path / "run.ckpt"
    kw = dict(tile_range=(1 << 70, (1 << 70) + 10 ** 6), bits=89, target=18, depth=18, prefix="0b101",
              walk="sequential", lanes=4, lane_id=2, batch_tiles=1000, seed=9)
    c = ccs.Campaign(**kw)
    b1, b2, b3 = c.next(), c.next(), c.next()               # three batches in flight
    res = np.zeros(3, ccs.CHAIN_RESULT_DTYPE)
    res["flags"] = ccs.F_PRIME
    res["chain_len"] = [3, 18, 5]
    res["flags"][1] |= ccs.F_NONROOT
    c.record(*b1, candidates=10, survivors=3, res=res)
    with pytest.raises(ccs.CcsError):
        c.record(*b3)                                        # out of order
    c.save(path)
    assert not os.path.exists(str(path) + ".tmp")
    lines = open(path).read().splitlines()
    assert lines[0] == "CC18_CUDA_v10"
    kv = dict(l.split("=", 1) for l in lines[1:])
    assert kv["bits"] == "89" and kv["target"] == "18" and kv["depth"] == "18" and kv["coverage"] == "full"
    assert kv["prefix"] == "0b101" and kv["prefix_lanes"] == "4" and kv["prefix_lane_id"] == "2"
    assert (int(kv["tile_hi"], 16) << 64) | int(kv["tile_lo"], 16) == b2[0]     # first unrecorded batch
    assert (kv["total_tiles"], kv["total_primes"], kv["total_chains"], kv["best_chain"]) == ("1000", "3", "3", "18")
    r = ccs.Campaign(**kw)
    r.resume(path)
    st = r.state()
    assert st["current"] == b2[0] and st["random_mode"] == 0 and st["total_tiles"] == 1000
    assert st["chains"]["best_chain"] == 18 and st["chains"]["primes"] == 3
    assert r.next() == b2 and r.next() == b3
    for bad in (dict(bits=90), dict(prefix="0b100"), dict(lanes=8), dict(lane_id=1), dict(depth=17)):
        x = ccs.Campaign(**{**kw, **bad})
        with pytest.raises(ccs.CcsError):
            x.resume(path)
    with open(path, "w") as f:
        f.write("CC17_OTHER\n")
    with pytest.raises(ccs.CcsError):
        ccs.Campaign(**kw).resume(path)


def test_parse_prefix():
    assert ccs.parse_prefix("0b1") == (1, 1)
    assert ccs.parse_prefix("0B1011") == (11, 4)
    # the reference checks the length before each digit: 127 digits pass, 128 do not
    assert ccs.parse_prefix("0b1" + "0" * 126) == (1 << 126, 127)
    for bad in ("101", "0b", "0b0101", "0b12", "0x1", "0b1" + "0" * 127):
        with pytest.raises(ccs.CcsError):
            ccs.parse_prefix(bad)
