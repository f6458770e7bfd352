"""The grid-machine surface (ClassDDEGridMachine put / get) on the GPU: the native calls go to the
HIP library through the drop-in shim; every output must match the reference golden vectors
(relative L2 <= 1e-5, touched cells identical, sumwt / SumJones <= 1e-10)."""
import pytest

from cases import problem
from test_gridmachine import GM_CASES, _golden, run_put_get

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", GM_CASES)
def test_put_get_on_the_gpu_match_reference(name):
    print("case", name, flush=True)
    run_put_get(problem(name), _golden(name), 1e-5)
