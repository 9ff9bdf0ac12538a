"""Precision form on the GPU: FP32 rates against FP64 rates on the same seeds
(the 512^3 study of misa-akmc-gpu_amd/akmc/studies.py at a reduced size).

Bars: every FP32 rate within 1e-6 (relative) of the FP64 rate of the same
stencil, and the Cu-cluster statistics of the final lattices the same in
distribution; with the same seeds most final lattices are identical, since a
selection can only differ when u * total falls within ~1e-7 of a boundary."""
import sys

import pytest

from akmc import studies

pytestmark = pytest.mark.gpu


def test_precision_study_small():
    print("test_precision_study_small", flush=True)
    sys.stdout.flush()
    res = studies.precision(L=32, seeds=6, events=20000, hist_cap=32, chunk=10000, log=lambda m: None,
                            error_events=20000)
    assert res["rate_error"]["rates_compared"] > 10000
    assert res["rate_error"]["max_rel"] <= 1e-6
    assert res["identical_final_lattices"] >= 4
    for k, t in res["welch_t"].items():
        assert abs(t) < 4.0, (k, t)


def test_plain_study_in_threads_small():
    # The plain-vs-FP32 study with its exact engines in host threads: the
    # same per-seed results as with one engine at a time.
    print("test_plain_study_in_threads_small", flush=True)
    a = studies.plain(L=24, seeds=4, events=4000, hist_cap=32, log=lambda m: None, threads=1)
    b = studies.plain(L=24, seeds=4, events=4000, hist_cap=32, log=lambda m: None, threads=4)
    assert a["plain_fp64"]["per_seed"] == b["plain_fp64"]["per_seed"]
    for k, t in b["welch_t"].items():
        assert abs(t) < 6.0, (k, t)
