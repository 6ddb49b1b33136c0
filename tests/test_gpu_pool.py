"""Stream ordering of the library's device memory cache.

A block freed while a kernel that writes it is still queued on one stream
must not be handed to another stream before that kernel completes; the
self-test inside hpgl.so frees a block behind a ~50 ms kernel on stream A,
allocates the same size class on stream B, and checks the two differ.
"""
import pytest

from hpgl_amd.geo_bsd import hpgl_wrap


@pytest.mark.gpu
def test_freed_block_waits_for_its_stream():
    rc = hpgl_wrap._hpgl_so.hpgl_amd_pool_selftest()
    hpgl_wrap.check_last_error()
    assert rc in (0, 1), "pool handed a pending block to another stream (code %d)" % rc
    if rc == 1:
        pytest.skip("probe kernel finished before the check (inconclusive)")
