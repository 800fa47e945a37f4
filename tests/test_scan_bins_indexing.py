"""Indexing of k_scan_bins (billardballraytracing3d_amd/csrc/bbrt_hip.hip), emulated on the CPU.

The kernel is one block of 1024 threads that scans the group bins' counts in chunks of 8192: each chunk is
staged in LDS (zeros past the end), thread t sums its 8 consecutive bins, the 1024 sums get a Hillis-Steele
inclusive scan, thread t writes the exclusive starts of its 8 bins from carry + part[t - 1], and the chunk
total carries into the next pass.  This test runs exactly that schedule (every thread's reads and writes, pass
by pass) and checks the starts against an exclusive cumulative sum, for bin counts below, at, between and
above multiples of the chunk.  The GPU tests cover the kernel itself end to end (grouped queue results equal
input-order results at 256, 8384 and 32768 bins).
"""
import numpy as np
import pytest

CHUNK, THREADS, PER = 8192, 1024, 8


def scan_bins_schedule(hist):
    nbins = hist.size
    start = np.full(nbins, -1, np.int64)
    carry = 0
    for base in range(0, nbins, CHUNK):
        m = min(CHUNK, nbins - base)
        buf = np.zeros(CHUNK, np.int64)
        for i in range(CHUNK):                      # thread t loads i = t, t + 1024, ...
            buf[i] = hist[base + i] if i < m else 0
        v = buf.reshape(THREADS, PER).copy()        # thread t: buf[8t .. 8t + 7]
        part = v.sum(axis=1)
        off = 1
        while off < THREADS:                        # read all, barrier, add, barrier
            u = np.concatenate([np.zeros(off, np.int64), part[:-off]])
            part = part + u
            off <<= 1
        for t in range(THREADS):
            run = carry + (part[t - 1] if t else 0)
            for k in range(PER):
                buf[t * PER + k] = run
                run += v[t, k]
        total = part[THREADS - 1]
        for i in range(m):                          # only the chunk's real bins are stored
            assert start[base + i] == -1, "a bin written twice"
            start[base + i] = buf[i]
        carry += total
    return start


@pytest.mark.parametrize("nbins", [1, 7, 64, 8191, 8192, 8193, 8384, 16384, 20000, 32768 + 5])
def test_scan_bins_schedule_is_an_exclusive_scan(nbins):
    rng = np.random.default_rng(nbins)
    hist = rng.integers(0, 50, nbins).astype(np.int64)
    hist[rng.integers(0, nbins, max(1, nbins // 3))] = 0   # empty bins, as most are
    start = scan_bins_schedule(hist)
    expect = np.concatenate([[0], np.cumsum(hist)[:-1]])
    np.testing.assert_array_equal(start, expect)
