"""CPU-side checks of the C-ABI library: it loads without a GPU, exports every
symbol the public headers declare, validates its inputs, and its synthetic
workload generators are deterministic and well-formed.  No kernel runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as ol
import pcx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(pcxg?_[a-z_]+)\s*\(", src))


def test_library_loads_and_exports_every_declared_symbol():
    L = pcx.lib()
    names = _declared("pcx.h") | _declared("pcx_gen.h") | _declared("pcx_model.h")
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    bound = {n for n, _, _ in pcx.SIGNATURES}
    assert names == bound, f"ctypes bindings differ from the headers: {names ^ bound}"


def test_version_and_config():
    assert b"pcx" in pcx.lib().pcx_version()
    c = pcx.default_config()
    assert c.props_per_wave == 0 and c.max_sweeps == 0


def test_opcodes_match_oracle():
    assert pcx.OPS == ol.OPS


def test_no_device_is_an_error_not_a_fallback():
    if pcx.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(pcx.PcxError) as e:
        pcx.PIR()
    assert e.value.status == pcx.EDEVICE


def test_generator_deterministic():
    p = pcx.gen_params(nvars=2000, nprops=8000, seed=5)
    a = pcx.gen_network(p)
    b = pcx.gen_network(p)
    for u, v in zip(a, b):
        assert (u == v).all()
    c = pcx.gen_network(pcx.gen_params(nvars=2000, nprops=8000, seed=6))
    assert not (a[1] == c[1]).all()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_planted_network_is_satisfied_by_its_witness(seed):
    p = pcx.gen_params(nvars=5000, nprops=20000, seed=seed)
    op, x, y, z, lb, ub, w = pcx.gen_network(p)
    assert ((lb <= w) & (w <= ub)).all()
    wx, wy, wz = w[x].astype(np.int64), w[y].astype(np.int64), w[z].astype(np.int64)
    expect = np.select([op == pcx.ADD, op == pcx.MUL, op == pcx.MIN, op == pcx.MAX,
                        op == pcx.EQ, op == pcx.LEQ],
                       [wy + wz, wy * wz, np.minimum(wy, wz), np.maximum(wy, wz),
                        (wy == wz).astype(np.int64), (wy <= wz).astype(np.int64)], -1)
    assert (expect == wx).all()
    assert ((x != y) & (y != z) & (x != z)).all()
    counts = np.bincount(op, minlength=10)
    assert counts[[pcx.TDIV, pcx.FDIV, pcx.CDIV, pcx.EDIV]].sum() == 0
    for k in (pcx.MUL, pcx.MIN, pcx.MAX, pcx.EQ, pcx.LEQ):
        assert 0.10 < counts[k] / len(op) < 0.18


def test_literal_network_is_uniform():
    p = pcx.gen_params(nvars=1000, nprops=3000, seed=0, planted=0, lb_max=0,
                       width_min=1000, width_max=1000, chain_frac=0.0)
    op, x, y, z, lb, ub, _ = pcx.gen_network(p)
    assert (lb == 0).all() and (ub == 1000).all()
    assert ((x != y) & (y != z) & (x != z)).all()
    assert set(np.unique(op)) == {pcx.ADD, pcx.MUL, pcx.MIN, pcx.MAX, pcx.EQ, pcx.LEQ}


def test_splits_stay_inside_bounds():
    lb = np.array([0, 5, 10, 3], np.int32)
    ub = np.array([9, 5, 20, 3], np.int32)
    lo, hi = pcx.gen_splits(3, lb, ub, 50, 1, 20)
    lo, hi = lo.reshape(50, 4), hi.reshape(50, 4)
    assert ((lo >= lb) & (hi <= ub) & (lo <= hi)).all()
    assert (lo[:, 1] == 5).all() and (hi[:, 3] == 3).all()   # singletons untouched
    assert ((lo != lb) | (hi != ub)).any(axis=1).all()


def test_subsets_sorted_distinct_sized():
    off, idx = pcx.gen_subsets(9, 1000, 40, 20, 700)
    assert off[0] == 0 and (np.diff(off) >= 20).all() and (np.diff(off) <= 700).all()
    for k in range(40):
        s = idx[off[k]:off[k + 1]]
        assert (np.diff(s) > 0).all() and s.min() >= 0 and s.max() < 1000


def test_planted_bits_network_holds_under_witness():
    op, x, y, z, w, wit = pcx.gen_bits_network(4, 3000, 12000, witness=True)
    assert all((int(w[v]) >> int(wit[v])) & 1 for v in range(3000))
    wx, wy, wz = wit[x], wit[y], wit[z]
    assert ((op != pcx.BADD) | (wx == wy + wz)).all()
    assert ((op != pcx.BLEQ) | (wy <= wz)).all()
    assert ((op != pcx.BNEQ) | (wy != wz)).all()
    assert ((op != pcx.BEQR) | (wx == (wy == wz))).all()
    assert ((x != y) & (y != z) & (x != z)).all()
    counts = np.bincount(op, minlength=4)
    assert (counts > 2500).all()


def test_bits_splits_remove_values():
    w = np.array([0b1111, 0b1, 0b110], np.uint64)
    out = pcx.gen_bits_splits(2, w, 40, 1, 10).reshape(40, 3)
    assert ((out & ~w) == 0).all() and (out != 0).all()
    assert (out[:, 1] == 1).all()


def test_tuning_cache_round_trip(tmp_path):
    path = str(tmp_path / "t.json")
    assert pcx.load_tuning(path, "a") is None
    best = {"mode": 1, "props_per_wave": 256, "lds_mode": 0, "compact_tables": 1,
            "stream_tables": 1, "event_driven": 1, "ms": 3.5}
    pcx.save_tuning(path, "a", best, [best])
    pcx.save_tuning(path, "b", dict(best, mode=2))
    assert pcx.load_tuning(path, "a")["best"] == best
    assert pcx.load_tuning(path, "b")["best"]["mode"] == 2


def test_result_record_layout_matches_c():
    import ctypes
    assert pcx.RESULT_DTYPE.itemsize == ctypes.sizeof(pcx.Result)
    for name, _ in pcx.Result._fields_:
        assert pcx.RESULT_DTYPE.fields[name][1] == getattr(pcx.Result, name).offset


# C struct <-> ctypes mirror: every field's offset and every struct's size.
_MIRRORS = {"pcx_config": "Config", "pcx_result": "Result", "pcx_stats": "Stats",
            "pcx_tuning": "Tuning", "pcx_server_config": "ServerConfig",
            "pcx_served": "Served", "pcx_server_stats": "ServerStats", "pcxg_params": "GenParams"}


def test_struct_layouts_match_c(tmp_path):
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pcx.h"', '#include "pcx_gen.h"',
             "int main(void) {"]
    for cname, pyname in _MIRRORS.items():
        cls = getattr(pcx, pyname)
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:-1]): int(l.split()[-1]) for l in out if l.strip()}
    for cname, pyname in _MIRRORS.items():
        cls = getattr(pcx, pyname)
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, f"{cname}.{f}"


def test_reference_side_adapter_compiles_against_pir_surface(tmp_path):
    """include/lala_pcx_adapter.hpp compiles against the stand-in of PIR's
    public surface (tests/adapter/pir_stub.hpp) and links to libpcx.so; with
    no HIP device the adapter reports pcx's error instead of propagating."""
    import subprocess
    import pcx
    here = os.path.dirname(os.path.abspath(__file__))
    out = str(tmp_path / "adapter_kat")
    r = subprocess.run(["make", "-s", "-C", os.path.join(here, "adapter"), f"OUT={out}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if pcx.device_count() > 0:
        return
    r = subprocess.run([out], input="vars 3\n0 10\n0 10\n0 5\ntell 1\nADD 2 0 1\nfix\n",
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "no HIP device" in r.stderr
