"""The row-gather rate probe (roofline calibration; no reference counterpart):
runs for both row distributions and lane layouts, refuses bad arguments."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dist", [0, 1])
@pytest.mark.parametrize("layout", [0, 1])
def test_gather_probe_runs(gpu, dist, layout):
    ms = gpu.gather_probe(3000, 1 << 16, dist=dist, zipf_s=1.1, layout=layout)
    assert ms > 0.0


def test_gather_probe_rejects_bad_arguments(gpu):
    with pytest.raises(gpu.AltoError):
        gpu.gather_probe(0, 1 << 16)
    with pytest.raises(gpu.AltoError):
        gpu.gather_probe(1000, 1 << 16, layout=2)
    with pytest.raises(gpu.AltoError):
        gpu.gather_probe(1000, 1 << 16, dist=1, zipf_s=0.0)
