"""Writes tests/golden/fzn_kats.json from the reference's tests/pir_test.cpp,
and tests/golden/pc_kats.json from its tests/pc_test.cpp (the same kind of
test over PC, the general propagator completion, include/lala/pc.hpp).

Each case is one TEST of that file whose model is a string literal: the
model text, the constraints told after it, and, for every deduce_and_test /
deduce_and_test_bot call, the expected number of propagators and the expected
stores before and after the Gauss-Seidel fixpoint (or failure).  Expected
stores are positional over the reference's variables; `names` gives the
reference's variable order: the declared variables in declaration order
(array elements as name[i]), or, where the test lists the generated
variables in a comment (pir_test.cpp:193-206, 611-618), that list.

The text is read, nothing of the reference is run.  Run from the repository
root in a tree where the reference sits at /root/reference:
    python tests/golden/make_fzn_kats.py
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
# (source, TEST suite name, output file, the reference's abstract domain)
SOURCES = [("/root/reference/tests/pir_test.cpp", "PIRTest", os.path.join(HERE, "fzn_kats.json"), "PIR"),
           ("/root/reference/tests/pc_test.cpp", "IPCTest", os.path.join(HERE, "pc_kats.json"), "PC")]
# the test helpers that build the abstract element from a model string
CREATE = r"create_and_interpret_and_tell<[^>]*>|interpret_type_and_tell2?<[^>]*>"
INF = 2**31 - 1
MINF = -2**31


def c_strings(s):
    """The C++ string literals of s, in order (with \\-newline continuations joined)."""
    out = []
    for m in re.finditer(r'"((?:[^"\\]|\\.)*)"', s, flags=re.S):
        out.append(re.sub(r"\\\n", "", m.group(1)).replace('\\"', '"'))
    return out


def split_top(s, sep=","):
    """Split s at top-level separators (outside brackets and string literals)."""
    parts, depth, cur, quoted, prev = [], 0, "", False, ""
    for ch in s:
        if ch == '"' and prev != "\\":
            quoted = not quoted
        elif not quoted and ch in "({[":
            depth += 1
        elif not quoted and ch in ")}]":
            depth -= 1
        prev = ch
        if ch == sep and depth == 0 and not quoted:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


def bound(s, lower):
    s = s.strip()
    if re.fullmatch(r"-?\d+", s):
        return int(s)
    if s in ("Itv::LB::top()", "zlb::top()", "Itv::UB::top()", "zub::top()"):
        return MINF if lower else INF
    m = re.fullmatch(r"z[lu]b\((-?\d+)\)", s)
    if m:
        return int(m.group(1))
    raise ValueError(f"bound {s!r}")


def itv(s):
    s = s.strip()
    if s == "Itv::top()":
        return [MINF, INF]
    if s == "Itv::bot()":
        return "bot"
    m = re.fullmatch(r"Itv\((.*)\)", s, flags=re.S)
    if not m:
        raise ValueError(f"interval {s!r}")
    a, b = split_top(m.group(1))
    return [bound(a, True), bound(b, False)]


def itv_list(s):
    s = s.strip()
    assert s[0] == "{" and s[-1] == "}", s
    return [itv(p) for p in split_top(s[1:-1])]


def call_args(body, start):
    """Arguments of the call whose '(' is at body[start]."""
    depth, i = 0, start
    while True:
        ch = body[i]
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
            if depth == 0:
                return split_top(body[start + 1:i]), i
        elif ch == '"':
            i += 1
            while body[i] != '"':
                i += 2 if body[i] == "\\" else 1
        i += 1


def declared_names(model):
    names = []
    for item in model.split(";"):
        item = item.strip()
        m = re.match(r"array\s*\[\s*1\s*\.\.\s*(\d+)\s*\]\s*of\s+var\s+[^:]+:\s*(\w+)", item)
        if m:
            names += [f"{m.group(2)}[{k}]" for k in range(1, int(m.group(1)) + 1)]
            continue
        m = re.match(r"var\s+[^:]+:\s*(\w+)", item)
        if m:
            names.append(m.group(1))
    return names


def extract(path, suite, domain):
    src = open(path).read()
    rel = "tests/" + os.path.basename(path)
    tests = list(re.finditer(r"^TEST\(" + suite + r",\s*(\w+)\)\s*\{", src, flags=re.M))
    cases, skipped = [], []
    for k, t in enumerate(tests):
        name = t.group(1)
        end = tests[k + 1].start() if k + 1 < len(tests) else len(src)
        body = src[t.end():end]
        line0 = src[:t.start()].count("\n") + 1
        line1 = line0 + body.rstrip().count("\n")
        mk = re.search(r"(" + CREATE + r")\(", body)
        if not mk or "std::format" in body:
            skipped.append(name)
            continue
        texts, steps = [], []
        i = 0
        ok = True
        for m in re.finditer(r"(" + CREATE + r"|interpret_must_succeed<[^>]*>|"
                             r"deduce_and_test_bot|deduce_and_test)\(", body):
            args, _ = call_args(body, m.end() - 1)
            fn = m.group(1)
            if not fn.startswith("deduce_and_test"):
                strs = c_strings(args[0])
                if len(strs) != 1:
                    ok = False
                    break
                texts.append(strs[0])
                continue
            if len(steps) >= len(texts):           # a second check without a new tell
                texts.append("")
            try:
                nd = int(args[1])
                if fn == "deduce_and_test_bot":
                    steps.append({"nd": nd, "before": itv_list(args[2]), "after": None, "failed": True})
                else:
                    before = itv_list(args[2])
                    after = before
                    # deduce_and_test(x, nd, before, after, is_ua, expect_changed = true) or
                    # deduce_and_test(x, nd, before_after, is_ua = false) (pir_test.cpp:54-89)
                    flags = args[4:] if len(args) >= 4 and args[3].strip().startswith("{") else args[3:]
                    if len(args) >= 4 and args[3].strip().startswith("{"):
                        after = itv_list(args[3])
                        changed = flags[1].strip() == "true" if len(flags) > 1 else True
                    else:
                        changed = False
                    ua = bool(flags) and flags[0].strip() == "true"
                    # an expected store with an empty interval is a failure (is_bot)
                    empty = any(e == "bot" or e[0] > e[1] for e in after)
                    steps.append({"nd": nd, "before": before, "after": after, "failed": empty,
                                  "ua": ua, "changed": changed})
            except (ValueError, AssertionError):
                ok = False
                break
        if not ok or not steps or len(texts) != len(steps):
            skipped.append(name)
            continue
        gen = re.search(r"Generated variables:(.*?)\*/", body, flags=re.S)
        if gen:
            names = re.findall(r"\(var\s+(\w+):", gen.group(1))
        else:
            names = declared_names(texts[0])
        cases.append({"name": name, "ref": f"{rel}:{line0}-{line1}", "names": names,
                      "steps": [dict(text=x, **s) for x, s in zip(texts, steps)]})
    about = (f"Models and expected stores of the reference's {rel} (its {domain} over interval "
             "stores), extracted as data by tests/golden/make_fzn_kats.py: for each step, the FlatZinc "
             f"text told, the reference's {domain} propagator count after it (nd), and the expected "
             "store before and after the Gauss-Seidel fixpoint over the reference's variables in the "
             "order of `names` (positions past the declared variables belong to the reference's "
             "generated variables), or failed = true; ua and changed are the test's expected "
             "is_extractable() and has_changed. Bounds are int32; -2^31 / 2^31-1 are the store's "
             "infinities; \"bot\" is an empty interval.")
    return {"about": about, "skipped": skipped, "cases": cases}


def main():
    for path, suite, out, domain in SOURCES:
        data = extract(path, suite, domain)
        json.dump(data, open(out, "w"), indent=1)
        print(f"{os.path.basename(out)}: {len(data['cases'])} cases, skipped {data['skipped']}",
              file=sys.stderr)


if __name__ == "__main__":
    main()
