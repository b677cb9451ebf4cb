"""The stream-capture rule (CPU only: sbx_capture_check is host code and needs no device).

The step's captures fork the side streams from the capturing stream, let the broad-phase stream wait on
the sort stream (its placement) and join both back on the capturing stream; the sort stream waits on the
broad-phase stream only across captures (an event of the previous graph).  Two captures that made the
sort stream wait on an event the broad-phase stream recorded inside the same capture ended the host
process; the library now logs every record and wait of a capture and refuses one that breaks the rule.
These tests run the check on the capture shapes the step uses and on the two shapes that crashed,
restated as (op, stream, event) triples: op 0 record, 1 wait; stream 0 the capturing stream, 1 the
broad-phase stream, 2 the sort stream.  tests/test_gpu_network.py checks the logs of real captures.
"""
import ctypes as C
import os

import numpy as np
import pytest

from softbox_amd import _lib

REC, WAIT = 0, 1
MAIN, SIDE, SORT = 0, 1, 2
FORK, JOIN0, JOIN1, PLACED0, PLACED1, ROD0, ROD1 = range(7)


def check(ops):
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsoftbox_hip.so not built")
    lib = _lib.load()
    a = np.ascontiguousarray(np.asarray(ops, np.int32).ravel())
    return lib.sbx_capture_check(a.ctypes.data_as(C.c_void_p), len(ops))


def two_step_graph():
    """The 2-step graph of the overlapped step: per step a fork, the sort and broad phase of the step's
    buffer on the side streams (the sort waiting on the previous graph's broad phase of that buffer), the
    mechanics on the capturing stream waiting on this step's placement, and the join at the end."""
    ops = [(REC, MAIN, FORK), (WAIT, SIDE, FORK), (WAIT, SORT, FORK)]
    for step, (join, placed, rod_prev) in enumerate(((JOIN0, PLACED0, ROD1), (JOIN1, PLACED1, ROD0))):
        ops += [(WAIT, SORT, rod_prev),            # the previous k_rod binned this pose
                (WAIT, SORT, join),                # the broad phase two steps back read this CSR (previous graph)
                (REC, SORT, placed),
                (WAIT, SIDE, placed),
                (REC, SIDE, join),
                (WAIT, MAIN, placed),
                (REC, MAIN, ROD0 if step == 0 else ROD1)]
    ops += [(WAIT, MAIN, JOIN0), (WAIT, MAIN, JOIN1)]
    return ops


def test_forks_joins_and_previous_graph_waits_pass():
    assert check(two_step_graph()) == -1
    assert check([]) == -1


def test_side_wait_on_the_other_side_streams_record_in_the_same_capture_is_refused():
    """The form of both crashes: the sort stream waits on a broad phase the side stream ran inside the same
    capture (an 8-step graph's sort two steps on; a replay leaving its last broad phase to the next one)."""
    ops = two_step_graph()
    ops.insert(len(ops) - 2, (WAIT, SORT, JOIN0))   # JOIN0 was recorded by SIDE earlier in this capture
    assert check(ops) == len(ops) - 3
    crash = [(REC, MAIN, FORK), (WAIT, SIDE, FORK), (WAIT, SORT, FORK), (REC, SIDE, JOIN0), (WAIT, SORT, JOIN0)]
    assert check(crash) == 4


def test_same_stream_and_main_stream_waits_pass():
    ops = [(REC, SIDE, JOIN0), (WAIT, SIDE, JOIN0), (REC, SORT, PLACED0), (WAIT, MAIN, PLACED0),
           (REC, MAIN, FORK), (WAIT, SORT, FORK)]
    assert check(ops) == -1


def test_malformed_triples_are_reported():
    assert check([(REC, MAIN, FORK), (7, MAIN, FORK)]) == 1
    assert check([(WAIT, -1, FORK)]) == 0
    assert check([(REC, 3, FORK)]) == 0       # a stream that is not one of the step's
