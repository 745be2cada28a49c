"""The HIP library (through the C ABI) against the reference's known answers."""
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("check", cases.ALL, ids=lambda f: f.__name__)
def test_gpu_known_answers(gpu, golden, check):
    check(gpu, golden)
