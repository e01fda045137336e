"""Host-side parts of the serving loop (no GPU): request streams and the latency report."""
import math

import numpy as np

import cfdx
from cfdx import serve


def test_stream_is_seeded_and_poisson():
    a = serve.periodic_stream(400, rate=20.0, seed=7)
    b = serve.periodic_stream(400, rate=20.0, seed=7)
    assert [(r.arrival, r.ni, r.init, r.mach) for r in a] == [(r.arrival, r.ni, r.init, r.mach) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1 / 20.0) < 0.2 / 20.0          # exponential gaps of mean 1/rate
    assert all(0.1 <= r.mach <= 0.6 and 500 <= r.reyl <= 5000 for r in a)
    assert {r.init for r in a} == {"tgv", "vortex"}
    assert {(r.ni, r.nj, r.nk) for r in a} <= {(128, 64, 16), (192, 96, 24), (256, 128, 32), (384, 128, 32)}


def test_report_percentiles():
    reqs = [serve.Request(i, 8, 8, 8, 10, arrival=float(i)) for i in range(100)]
    res = {r.rid: serve.Result(r.rid, r.arrival, r.arrival, r.arrival + 1.0 + 0.01 * r.rid, 1e-3, None) for r in reqs}
    rep = serve.report(reqs, res)
    assert math.isclose(rep["p50_s"], 1.495, rel_tol=1e-9)
    assert math.isclose(rep["p99_s"], 1.9801, rel_tol=1e-9)
    assert rep["cell_updates"] == 100 * 512 * 10
    assert math.isclose(rep["makespan_s"], 99 + 1.99)


def test_request_params():
    r = serve.Request(0, 8, 8, 8, 1, mach=0.3, reyl=1234.0)
    p = r.params()
    assert p.mach == 0.3 and p.reyl == 1234.0 and p.viscous == 1
    e = serve.Request(0, 8, 8, 8, 1, viscous=False).params()
    assert math.isinf(e.reyl) and e.viscous == 0
    assert cfdx.Block.bytes_for(8, 8, 8, cfdx.MIXED) < cfdx.Block.bytes_for(8, 8, 8, cfdx.FP64)


def test_gate_lets_a_waiting_capture_go_before_the_next_call():
    """The server's gate: set-up calls and captures run one at a time, and a capture that waits
    goes before the set-up thread's next call, so it waits for one call at most (a plain lock,
    released and taken again at once between calls, can hold it back for all of them)."""
    import threading
    import time

    gate = serve._Gate()
    order = []
    started = threading.Event()

    def calls():
        for i in range(20):
            with gate.call:
                order.append(("call", i))
                started.set()
                time.sleep(0.005)

    t = threading.Thread(target=calls)
    t.start()
    started.wait()
    time.sleep(0.012)
    with gate.capture:
        order.append(("capture", None))
    t.join()
    k = order.index(("capture", None))
    # the capture came in while calls were running and went before most of them
    assert 1 <= k <= 10, order
    assert [x for x in order if x[0] == "call"] == [("call", i) for i in range(20)]
