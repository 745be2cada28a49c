"""The oracle against the known answers held by the reference's own tests."""
import pytest

from tests import cases


@pytest.mark.parametrize("check", cases.ALL, ids=lambda f: f.__name__)
def test_oracle_known_answers(oracle, golden, check):
    check(oracle, golden)
