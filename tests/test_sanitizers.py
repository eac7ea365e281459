"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize/san_driver links the FlatZinc front-end (pcx_model.cpp), the
workload generators (pcx_gen.cpp) and the CPU oracle (oracle/*.c) built with
-fsanitize=address,undefined (no GPU code).  It parses the reference's test
models and a seeded corpus of malformed variants of them (truncations,
deleted and replaced characters, inserted tokens, deep nesting, huge
constants), and runs the generators and the oracle on small workloads.  Any
sanitizer report aborts the driver, which fails the test."""
import json
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FZN = os.path.join(HERE, "golden", "fzn_kats.json")
TOKENS = [")", "(", "[", "]", ";", ",", "..", ":", "=", "{", "}", "var", "int", "array[1..3] of",
          "constraint", "int_plus(", "2147483648", "-2147483649", "99999999999999999999", "::", "%",
          "bool_clause([],[])", "int_lin_eq([1,2],[x],3)", "\x00", "\xff", "'", '"']


def corpus(seed=5):
    texts = []
    for case in json.load(open(FZN))["cases"]:
        texts += [s["text"] for s in case["steps"] if s["text"]]
    rng = np.random.default_rng(seed)
    out = list(texts)
    for t in texts:
        n = len(t)
        for _ in range(6):
            out.append(t[:int(rng.integers(0, n + 1))])                   # truncation
        for _ in range(4):
            i = int(rng.integers(0, n))
            j = min(n, i + int(rng.integers(1, 8)))
            out.append(t[:i] + t[j:])                                       # deletion
        for _ in range(4):
            i = int(rng.integers(0, n))
            out.append(t[:i] + chr(int(rng.integers(32, 127))) + t[i + 1:])   # replacement
        for _ in range(4):
            i = int(rng.integers(0, n + 1))
            out.append(t[:i] + TOKENS[int(rng.integers(len(TOKENS)))] + t[i:])  # insertion
    depth = "var 0..9: a; constraint int_eq(a, " + "int_plus(1, " * 2000 + "a" + ")" * 2000 + ");"
    out += [depth, depth[:len(depth) // 2], "var 0..9: " + "a" * 100000 + ";",
            "array[1..2147483647] of var int: q;", "array[1..0] of var int: e; constraint bool_clause(e, e);",
            "var 5..4: empty; constraint int_le(empty, empty);", "constraint int_plus(1, 2, 3);"]
    return [s.replace("\x00", " ") for s in out]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("san") / "san_driver")
    r = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize"), f"OUT={out}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=1:allocator_may_return_null=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return env


def test_front_end_on_malformed_models(driver):
    models = corpus()
    data = "\n%%\n".join(models) + "\n%%\n"
    r = subprocess.run([driver, "models"], input=data.encode("utf-8", "surrogateescape"),
                       capture_output=True, env=_env(), timeout=600)
    err = r.stderr.decode(errors="replace")
    assert r.returncode == 0, err[-4000:]
    assert "ERROR" not in err and "runtime error" not in err, err[-4000:]
    assert f"{len(models)} models" in r.stdout.decode()


def test_generators_and_oracle(driver):
    r = subprocess.run([driver, "gen"], capture_output=True, env=_env(), timeout=600)
    err = r.stderr.decode(errors="replace")
    assert r.returncode == 0 and "runtime error" not in err, err[-4000:]
