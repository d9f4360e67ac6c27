"""elfel_hip_copy_bandwidth, the measured copy ceiling the benchmark reads its
roofline fractions beside: a positive rate below the published 8 TB/s peak
(read + write bytes), and the argument checks."""
import pytest

from elfel_hip import _lib as L

pytestmark = pytest.mark.gpu


def test_copy_bandwidth(ctx):
    g = ctx.copy_bandwidth(64 << 20, 4)
    print(f"64 MiB device copy: {g:.0f} GB/s")
    assert 100.0 < g < 8000.0
    for nbytes, reps in ((0, 1), (100, 1), (1 << 20, 0)):
        with pytest.raises(L.ElfelHipError) as e:
            ctx.copy_bandwidth(nbytes, reps)
        assert e.value.code == L.EINVAL
