"""The 113 JOB plans (reference plans.json through the JOB adapter) on the GPU
against the CPU oracle, on synthetic inputs shaped like the filtered IMDB
scans at 1/500 of PostgreSQL's row estimates: bit-exact row multisets,
every result page validated. A third of the scans hold one row at this
scale (tiny dimension tables)."""
import os
import sys

import pytest

import oracle
from gpujoin import job
from helpers import check_table

pytestmark = pytest.mark.gpu
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "job_queries.json")
JS = job.load_set(FIXTURE)
SCALE = 2e-3


@pytest.mark.parametrize("qi", range(len(JS.plans)), ids=JS.names)
def test_job_query_vs_oracle(ctx, qi):
    jp = JS.plans[qi]
    print(f"JOB {jp.name}: {len(jp.scans)} scans", flush=True)
    sys.stdout.flush()
    plan = job.with_inputs(jp, job.host_inputs(JS, qi, SCALE))
    exp = check_table(oracle.execute(plan), jp.types)
    got = check_table(ctx.execute(plan), jp.types)
    assert got == exp
