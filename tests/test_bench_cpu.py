"""bench.py's roofline reads each work family's HBM bytes per launch from the
committed counter summary under profiles/, by kernel name.  Every family and
precision mode the bench reports must find its kernels there (a renamed or
re-templated kernel would otherwise leave the line's traffic null), and the
figures must be positive."""
import importlib.util
import os
import re

import pytest

from tests.conftest import ROOT


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_every_family_has_counter_rows(bench):
    from gshtrans_amd.profile_summary import load_table
    table = load_table(bench.TRAFFIC_FILE)
    assert table and table["kernels"]
    names = list(table["kernels"])
    for family, pats in bench.FAMILY_KERNELS.items():
        for mode in ("fp64", "fp32", "bf16x3", "bf16x6"):
            pat = pats.get(mode) if isinstance(pats, dict) else pats
            if pat is None:
                continue  # the family does not run in this mode
            rows = [n for n in names if re.fullmatch(pat, n)]
            # the real and complex batches' instantiations (one kernel serves
            # both where the kind is not a template argument)
            assert 1 <= len(rows) <= 2, (family, mode, pat, rows)
            t = bench.traffic_per_launch(family, mode)
            assert t is not None and t > 1e8, (family, mode, t)
