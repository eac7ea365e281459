"""Pins the bitset oracle (oracle/bits_oracle.c) before the GPU is compared to it.

1. Known-answer tests transcribed from the reference (tests/pc_bitset_test.cpp).
2. Brute force over all non-empty domains within [0,6) for every opcode:
   soundness, generalized arc consistency (the fixpoint equals the
   projections of the solution set) and monotonicity.
3. Planted slips in copies of the oracle are each caught.
4. The faithful mode (the reference's has_changed assignment in Inequality,
   formula.hpp:803) can stop at a non-fixpoint; the true mode never does.
"""
import ctypes
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bits_kats.json")


def load_bits_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def run_bits_case(case, fixpoint):
    words = np.array([ol.word(d) for d in case["init"]], np.uint64)
    t = case["tell"]
    op = np.array([ol.BITS_OPS[b[0]] for b in t], np.int32)
    x = np.array([b[1] for b in t], np.int32)
    y = np.array([b[2] for b in t], np.int32)
    z = np.array([b[3] for b in t], np.int32)
    ol.bits_tell_embed(op, x, words)
    ol.bits_sort(op, x, y, z)
    failed, changed = fixpoint(op, x, y, z, words)
    errs = []
    if failed:
        errs.append("unexpected failure")
    for i, d in enumerate(case["after"]):
        if ol.values(words[i]) != d:
            errs.append(f"var {i} = {ol.values(words[i])} != {d}")
    if bool(changed) != case["changed"]:
        errs.append(f"has_changed {bool(changed)} != {case['changed']}")
    return errs


@pytest.mark.parametrize("case", load_bits_cases(), ids=lambda c: c["name"])
@pytest.mark.parametrize("faithful", [True, False])
def test_bits_kat(case, faithful):
    def fx(op, x, y, z, w):
        r = ol.bits_fixpoint(op, x, y, z, w, faithful=faithful)
        return r.failed, r.has_changed
    errs = run_bits_case(case, fx)
    assert not errs, f"{case['name']} ({case['ref']}): " + "; ".join(errs)


@pytest.mark.parametrize("name", list(ol.BITS_OPS))
def test_bits_consistency(name):
    buf = ctypes.create_string_buffer(512)
    n = ol.lib().pcxo_bits_check(ol.BITS_OPS[name], 6, 1, buf, len(buf))
    assert n == 0, f"{name}: {buf.value.decode()}"


BITS_MUTANTS = {
    11: "set difference shifts the wrong way",
    12: "LEQ bounds y by min(z) instead of max(z)",
    13: "NEQ shaves without checking that the fixed side is a singleton",
    14: "EQR: x false treated as x true",
}


@pytest.mark.parametrize("k", sorted(BITS_MUTANTS))
def test_bits_mutant_is_caught(k):
    d = tempfile.mkdtemp(prefix="pcx_bmut_")
    out = os.path.join(d, f"libbmut{k}.so")
    src = [os.path.join(ol.ORACLE_DIR, f) for f in ("pir_oracle.c", "pir_checks.c", "bits_oracle.c")]
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", f"-DPCXO_MUTANT={k}",
                           "-o", out] + src + ["-lm", "-pthread"])
    L = ctypes.CDLL(out)
    L.pcxo_bits_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    buf = ctypes.create_string_buffer(256)
    caught = any(L.pcxo_bits_check(op, 5, 1, buf, 256) for op in range(4))
    assert caught, f"bitset mutant {k} ({BITS_MUTANTS[k]}) passed every check"


@pytest.mark.parametrize("seed", range(6))
def test_bits_order_independence_and_faithful_mode(seed):
    rng = np.random.default_rng(seed)
    V, P = 40, 60
    words = np.array([rng.integers(1, 2**16) for _ in range(V)], np.uint64)
    op = rng.integers(0, 4, P).astype(np.int32)
    xyz = np.stack([rng.choice(V, 3, replace=False) for _ in range(P)]).astype(np.int32)
    x, y, z = xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()
    ol.bits_tell_embed(op, x, words)
    ref = None
    for trial in range(5):
        perm = rng.permutation(P)
        w = words.copy()
        r = ol.bits_fixpoint(op[perm].copy(), x[perm].copy(), y[perm].copy(), z[perm].copy(), w)
        got = (bool(r.failed), None if r.failed else w.tolist())
        ref = ref or got
        assert got == ref, f"seed {seed}: permutation {trial} differs"
    # the faithful result is never below the true fixpoint, and is a fixpoint
    # whenever it equals it
    w = words.copy()
    rf = ol.bits_fixpoint(op, x, y, z, w, faithful=True)
    if not ref[0] and not rf.failed:
        assert all((int(a) & int(b)) == int(b) for a, b in zip(w, ref[1]))


def test_faithful_mode_difference_count():
    """Counts the networks where the faithful mode stops short of the fixpoint.

    Shuffled chains of `y <= z` make a later leq in a sweep overwrite the
    has_changed that an earlier one set (formula.hpp:803 assigns instead of
    or-ing), so the reference's Gauss-Seidel loop can end after a sweep that
    did tighten.  Over these 200 networks 46 end early; each such result is
    above the true fixpoint and is not a fixpoint itself (one more true run
    from it tightens), and the other 154 are identical to the true fixpoint.
    """
    differ = 0
    for seed in range(200):
        rng = np.random.default_rng(seed)
        V = 12
        words = np.array([rng.integers(1, 2**20) for _ in range(V)], np.uint64)
        pairs = [(i, i + 1) for i in range(V - 1)]
        rng.shuffle(pairs)
        P = len(pairs)
        op = np.full(P, ol.BITS_OPS["LEQ"], np.int32)
        x = np.zeros(P, np.int32)
        y = np.array([p[0] for p in pairs], np.int32)
        z = np.array([p[1] for p in pairs], np.int32)
        wt = words.copy()
        rt = ol.bits_fixpoint(op, x, y, z, wt)
        wf = words.copy()
        rf = ol.bits_fixpoint(op, x, y, z, wf, faithful=True)
        if bool(rt.failed) == bool(rf.failed) and (rt.failed or np.array_equal(wt, wf)):
            continue
        differ += 1
        assert not rf.failed, f"seed {seed}: the faithful mode failed where the fixpoint did not"
        if not rt.failed:
            assert all((int(a) & int(b)) == int(b) for a, b in zip(wf, wt)), f"seed {seed}"
        again = wf.copy()
        r2 = ol.bits_fixpoint(op, x, y, z, again)
        assert r2.failed or r2.has_changed, f"seed {seed}: an early stop that is a fixpoint"
        assert bool(r2.failed) == bool(rt.failed) and (rt.failed or np.array_equal(again, wt))
    assert differ == 46, f"{differ} of 200 networks differ"


BUCKET_BITS = np.uint64((1 << 62) | (1 << 63))


@pytest.mark.parametrize("seed,V,P", [(1, 5000, 20000), (3, 200000, 800000), (61, 4000, 16000)])
def test_generated_bitset_domains_stay_inside_nbitset_exact_range(seed, V, P):
    """Generated bitset stores, their splits and their fixpoints hold values
    in [0, 61] only: the values NBitset<64> represents exactly (its bits for
    negative values and for values past 61 are buckets,
    tests/pc_bitset_test.cpp:150-157).  Inside that range the plain words of
    the oracle and NBitset describe the same sets."""
    import pcx
    op, x, y, z, w = pcx.gen_bits_network(seed, V, P)
    assert not (w & BUCKET_BITS).any()
    assert (w != 0).all()
    inst = pcx.gen_bits_splits(seed + 1, w, 8, 1, 20)
    assert not (inst & BUCKET_BITS).any()
    if V <= 5000:
        wo = w.copy()
        ol.bits_tell_embed(op, x, wo)
        ol.bits_sort(op, x, y, z)
        ol.bits_fixpoint(op, x, y, z, wo)
        assert not (wo & BUCKET_BITS).any()


def test_planted_bitset_witness_inside_range():
    import pcx
    op, x, y, z, w, wit = pcx.gen_bits_network(7, 3000, 12000, witness=True)
    assert wit.min() >= 0 and wit.max() <= 61
    assert all((int(w[v]) >> int(wit[v])) & 1 for v in range(len(w)))
