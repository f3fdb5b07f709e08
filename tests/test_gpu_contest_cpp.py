"""The C++ contest surface as a harness uses it: tests/cpp/contest_driver
(compiled against include/contest/plan.h, linked to libgpujoin.so) builds each
plan with the contest's ColumnInserter, calls Contest::build_context /
execute / destroy_context, decodes the result pages itself and lets the
result's ColumnarTable delete them. Its rows must equal the reference's
literal ground truth (tests/unit_tests.cpp:19-664) and the oracle's rows on
random plans."""
import os
import random
import struct
import subprocess

import pytest

import oracle
from gpujoin import table_rows
from gpujoin.pages import FP64, INT32, INT64, VARCHAR
from helpers import canon_rows, case_plan, load_cases, random_plan

TYPE_NAME = {INT32: "INT32", INT64: "INT64", FP64: "FP64", VARCHAR: "VARCHAR"}
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "contest_driver")


def _value(v, t):
    if v is None:
        return "NULL"
    if t == FP64:
        return "%016x" % struct.unpack("<Q", struct.pack("<d", v))[0]
    if t == VARCHAR:
        b = v.encode() if isinstance(v, str) else bytes(v)
        return b.hex() if b else "-"
    return str(int(v))


def plan_text(plan, name):
    """The driver's text form of a plan and its input rows."""
    out = [f"case {name}"]
    for tab in plan.inputs:
        types = [c.type for c in tab.columns]
        out.append("table %d %s" % (tab.num_rows, " ".join(TYPE_NAME[t] for t in types)))
        for row in table_rows(tab):
            out.append("row " + " ".join(_value(v, t) for v, t in zip(row, types)))
    for nd in plan.nodes:
        attrs = " ".join(f"{i}:{TYPE_NAME[t]}" for i, t in nd.output_attrs)
        if nd.kind == 1:  # scan
            out.append(f"scan {nd.base_table_id} {attrs}")
        else:
            out.append(f"join {int(nd.build_left)} {nd.left} {nd.right} {nd.left_attr} {nd.right_attr} {attrs}")
    out.append(f"root {plan.root}")
    out.append("run")
    return "\n".join(out) + "\n"


def _parse(tok, t):
    if tok == "NULL":
        return None
    if t == FP64:
        return struct.unpack("<d", int(tok, 16).to_bytes(8, "little"))[0]
    if t == VARCHAR:
        return b"" if tok == "-" else bytes.fromhex(tok)
    return int(tok)


def run_driver(plans):
    """Runs the plans through the driver in one process (one context); returns
    each plan's sorted rows."""
    text = "".join(plan_text(p, f"p{i}") for i, p in enumerate(plans))
    proc = subprocess.run([DRIVER], input=text.encode(), capture_output=True, timeout=600)
    assert proc.returncode == 0, proc.stderr.decode()
    lines = proc.stdout.decode().splitlines()
    results, k = [], 0
    for p in plans:
        head = lines[k].split()
        assert head[0] == "result", lines[k]
        nrows, ncols = int(head[1]), int(head[2])
        types = [t for _, t in p.nodes[p.root].output_attrs]
        assert ncols == len(types)
        rows = []
        for r in range(nrows):
            toks = lines[k + 1 + r].split(" ") if ncols else []
            rows.append(tuple(_parse(x, t) for x, t in zip(toks, types)))
        k += 1 + nrows
        results.append(canon_rows(rows))
    return results


def test_driver_is_built():
    """The driver links the library's Contest:: symbols (no GPU needed)."""
    assert os.path.exists(DRIVER), "build tests/cpp/contest_driver (make -C tests/cpp)"
    out = subprocess.run(["ldd", DRIVER], capture_output=True, text=True).stdout
    assert "libgpujoin.so" in out and "not found" not in out


def test_driver_round_trips_inputs():
    """No GPU: the driver's ColumnInserter pages, decoded by its own decoder,
    hold the rows that gpujoin's writer and decoder hold, for the golden
    cases and random plans with NULLs, long strings and every type."""
    rng = random.Random(7)
    plans = [case_plan(c) for c in load_cases()]
    plans += [random_plan(rng, ntables=3, max_rows=400, key_type=INT64, domain=50, long_strings=True)
              for _ in range(4)]
    text = "".join(plan_text(p, f"p{i}") for i, p in enumerate(plans))
    proc = subprocess.run([DRIVER, "--echo-inputs"], input=text.encode(), capture_output=True, timeout=120)
    assert proc.returncode == 0, proc.stderr.decode()
    lines = proc.stdout.decode().splitlines()
    k = 0
    for p in plans:
        for tab in p.inputs:
            head = lines[k].split()
            assert head[0] == "input" and int(head[1]) == tab.num_rows
            types = [c.type for c in tab.columns]
            rows = [tuple(_parse(x, t) for x, t in zip(lines[k + 1 + r].split(" "), types))
                    for r in range(tab.num_rows)]
            assert canon_rows(rows) == canon_rows(table_rows(tab))
            k += 1 + tab.num_rows


@pytest.mark.gpu
def test_contest_surface_golden_cases():
    cases = load_cases()
    got = run_driver([case_plan(c) for c in cases])
    for c, rows in zip(cases, got):
        exp = canon_rows([tuple(v for v in r) for r in c["expect_rows"]])
        assert rows == exp, c["name"]


@pytest.mark.gpu
def test_contest_surface_random_plans_vs_oracle():
    rng = random.Random(4242)
    plans = []
    for i in range(24):
        ntables, max_rows = rng.choice([2, 3, 4, 5]), rng.choice([30, 300, 3000])
        # bounded results: with d key values per join column each join
        # multiplies the row count by about rows/d
        domain = max(rng.choice([10, 200]), max_rows if ntables >= 3 else max_rows // 8)
        plans.append(random_plan(rng, ntables=ntables, max_rows=max_rows, key_type=rng.choice([INT32, INT64]),
                                 domain=domain, long_strings=(i % 6 == 0)))
    got = run_driver(plans)
    for i, (p, rows) in enumerate(zip(plans, got)):
        assert rows == canon_rows(table_rows(oracle.execute(p))), i
