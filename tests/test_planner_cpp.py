"""The planner's restriction choice and singleton elimination, checked by a
C++ program linked against the library (tests/cpp/planner_test.cpp; no GPU
work): kinds as the reference's setupRestriction picks them
(engine/storage/RestrictionLogic.cpp:397-447) with exact membership checked
by brute force, and the singleton reductions of
engine/query/QueryPlan.cpp:680-717 on hand-built queries."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "planner_test")


def test_planner_checks():
    assert os.path.exists(BIN), "build tests/cpp/planner_test (make -C tests/cpp)"
    proc = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "planner checks passed" in proc.stdout
