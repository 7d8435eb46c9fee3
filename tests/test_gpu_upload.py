"""tvh_upload_dev: host-to-device copies from pageable memory through the
page-locked staging ring.  Every size class (below the staging threshold, one
slot, several slots, more than the whole ring so that slots are reused),
default and caller streams, a source overwritten right after the call, and two
threads copying at once."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIB = 1 << 20


@pytest.mark.parametrize("count", [0, 1, 1000, (4 * MIB) // 8 - 1, (4 * MIB) // 8,
                                   (32 * MIB) // 8 + 1, (300 * MIB) // 8 + 3])
@pytest.mark.parametrize("use_stream", [False, True])
def test_upload_sizes(T, tvhmod, count, use_stream):
    hip_rt = tvhmod.hip_rt
    rng = np.random.default_rng(count + 7)
    src = rng.integers(0, 2**63, size=max(1, count), dtype=np.uint64)[:count]
    d = T.empty(max(1, count))
    st = hip_rt.Stream() if use_stream else None
    T.lib.upload(d, src if count else None, count, st)
    if st:
        st.synchronize()
    T.sync()
    if count:
        assert np.array_equal(d.download(count), src)


def test_upload_source_reusable_on_return(T, tvhmod):
    """The call returns once the data is staged: overwriting the source right
    after it must not change what arrives."""
    hip_rt = tvhmod.hip_rt
    count = (96 * MIB) // 8
    src = np.arange(count, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    want = src.copy()
    d = T.empty(count)
    st = hip_rt.Stream()
    T.lib.upload(d, src, count, st)
    src[:] = 0
    st.synchronize()
    assert np.array_equal(d.download(), want)


def test_upload_two_threads(T, tvhmod):
    hip_rt = tvhmod.hip_rt
    count = (80 * MIB) // 8 + 5
    srcs = [np.full(count, v, dtype=np.uint64) + np.arange(count, dtype=np.uint64)
            for v in (1, 1 << 40)]
    bufs = [T.empty(count) for _ in srcs]
    streams = [hip_rt.Stream() for _ in srcs]
    err = []

    def run(i):
        try:
            for _ in range(3):
                T.lib.upload(bufs[i], srcs[i], count, streams[i])
            streams[i].synchronize()
        except Exception as e:  # re-raised below
            err.append(e)

    th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    if err:
        raise err[0]
    for b, s in zip(bufs, srcs):
        assert np.array_equal(b.download(), s)
