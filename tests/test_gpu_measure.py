"""The copy-bandwidth measurement that the benchmark line reports beside
the HBM specification."""
import ctypes as C

import pytest

import btrc_amd

pytestmark = pytest.mark.gpu


def test_measure_copy_is_plausible_and_validates(ctx):
    gbs = ctx.measure_copy(512, 4)
    # the device's HBM peak is 8 TB/s; a 512 MiB copy may be helped a
    # little by the 256 MB last-level cache, never by much
    assert 1000.0 < gbs < 12000.0
    for bad in ((0, 4), (512, 0)):
        with pytest.raises(btrc_amd.BtrcException):
            ctx.measure_copy(*bad)
    out = C.c_double()
    assert ctx.lib.btrc_hip_measure_copy(ctx.h, 17, 1, C.byref(out)) != 0   # not a multiple of 16
    assert ctx.lib.btrc_hip_measure_copy(ctx.h, 1 << 20, 1, None) != 0
