"""The oracle against the reference's own known-answer vectors
(tests/golden/reference_vectors.json, transcribed from t/*.t)."""
import pytest

from harness import check_case, load_cases
from oracle_backend import OracleBackend
from oracle import pdlcpu as O

DOC = load_cases()


@pytest.mark.parametrize("case", DOC["cases"], ids=[c["id"] for c in DOC["cases"]])
def test_golden_case(case):
    err = check_case(OracleBackend(), case)
    assert err is None, "%s (%s): %s" % (case["id"], case["src"], err)


@pytest.mark.parametrize("row", DOC["pthread"]["rows"], ids=[str(r) for r in DOC["pthread"]["rows"]])
def test_pthread_split_table(row):
    """t/pthread.t:12-33: which dim, how many threads pdl_find_max_pthread picks"""
    target, dims, want_thr, want_dim = row
    got_thr, got_dim = O.find_max_pthread(dims, target)
    assert (got_thr, got_dim) == (want_thr, want_dim)
