"""The reference-side adapter (include/lala_pcx_adapter.hpp) on the GPU.

tests/adapter/adapter_kat drives the adapter over a stand-in of PIR's public
surface (tests/adapter/pir_stub.hpp) as a lala-pc caller would: it tells
bytecodes into the PIR, and the adapter reads them back with
load_deduce, propagates through libpcx.so and embeds the fixpoint into the
PIR's store.  Every known-answer test of tests/golden/pir_kats.json runs
through it, step by step; the stores after each fixpoint (or the failure)
must be the reference's."""
import os
import subprocess

import pytest

import kats
import pcx
from oracle_lib import OPS

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "adapter", "adapter_kat")
OPNAME = {v: k for k, v in OPS.items()}


@pytest.fixture(scope="module")
def gpu():
    if pcx.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (make -C tests/adapter, or __graft_entry__.build())")
    return True


def script_of(case):
    """The KAT as an adapter_kat script, and per fixpoint the expected step."""
    names = {n: i for i, n in enumerate(case["vars"])}
    init = [kats._itv(e) for e in case["init"]]
    lines = [f"vars {len(init)}"] + [f"{a} {b}" for a, b in init]
    for step in case["steps"]:
        if "tell" in step:
            t = step["tell"]
            lines.append(f"tell {len(t)}")
            lines += [f"{b[0]} {names[b[1]]} {names[b[2]]} {names[b[3]]}" for b in t]
        for v, lo, hi in step.get("embed", []):
            lines.append(f"embed {names[v]} {kats._bound(lo)} {kats._bound(hi)}")
        lines.append("fix")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("case", kats.load_cases(), ids=lambda c: c["name"])
def test_kat_through_the_adapter(gpu, case):
    out = subprocess.run([BIN], input=script_of(case), capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    pos = 0
    for s, step in enumerate(case["steps"]):
        head = lines[pos].split()
        pos += 1
        assert head[0] == "fix"
        nd, bot = int(head[1]), bool(int(head[3]))
        store = [tuple(map(int, lines[pos + i].split())) for i in range(len(case["init"]))]
        pos += len(case["init"])
        where = f"{case['name']} ({case['ref']}) step {s}"
        assert nd == step["nd"], f"{where}: num_deductions {nd} != {step['nd']}"
        if step["after"] == "bot":
            assert bot, f"{where}: expected failure"
            continue
        assert not bot, f"{where}: unexpected failure"
        for i, e in enumerate(step["after"]):
            assert store[i] == kats._itv(e), f"{where}: var {case['vars'][i]} {store[i]} != {kats._itv(e)}"
