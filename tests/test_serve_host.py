"""Admission of the serving loop (momentumed/serve.py), on the host with stand-ins for the
device context and the batch engine: FIFO order, backfilling behind a head that does not fit,
the head-wait rule (also while the head is still in preparation), and requests larger than the
whole budget. The reference has no serving layer; its sweep is a host loop of EDsolve calls
(`examples/landau level torus/topological order/4. cuda support.ipynb:299-301`)."""
import threading
from fractions import Fraction

import pytest

from momentumed import serve as S


class _Ctx:
    def __init__(self):
        self.used = 0

    def memory(self):
        return self.used, 0


class _Batch:
    def __init__(self):
        self.added = []

    def add(self, lm, nev, **kw):
        self.added.append(lm)
        return len(self.added) - 1


def _server(monkeypatch, budget, sizes, head_wait_s=60.0):
    """A SolveServer whose device parts are stand-ins: request r needs sizes[r] bytes, and
    admitting it adds them to the context's use."""
    srv = S.SolveServer.__new__(S.SolveServer)
    srv.budget = budget
    srv.head_wait_s = head_wait_s
    srv.kw = {}
    srv.ctx = _Ctx()
    srv.batch = _Batch()
    srv.estimate_bytes = lambda req: sizes[req.rid]

    def upload(job):
        job.basis, job.op = object(), object()
        srv.ctx.used += sizes[job.req.rid]
    srv._upload = upload
    monkeypatch.setattr(S, "ROCmLinearMap", lambda op, basis: (op, basis))
    return srv


def _jobs(arrivals, prepared):
    out = []
    for rid, (t, ready) in enumerate(zip(arrivals, prepared)):
        req = S.Request(rid, t, (3, 5), Fraction(1, 3), 5, (0, 0), 2.0)
        out.append(S._Job(req, arrays=((), (), ()) if ready else None))
    return out


def _admit(srv, waiting, now):
    running, rep = {}, S.ServeReport("batched")
    srv._admit_batched(waiting, running, rep, threading.Lock(), now)
    return sorted(j.req.rid for j in running.values()), [j.req.rid for j in waiting], rep


def test_fifo_until_the_budget_is_used(monkeypatch):
    srv = _server(monkeypatch, 100, [60, 30, 30, 50])
    adm, left, rep = _admit(srv, _jobs([0, 0, 0, 0], [True] * 4), now=1.0)
    assert adm == [0, 1] and left == [2, 3] and rep.admitted == 2 and rep.rejected == 0


def test_backfill_behind_a_head_that_does_not_fit(monkeypatch):
    srv = _server(monkeypatch, 100, [120, 90, 40, 30, 20])
    srv.ctx.used = 50                      # 50 free
    # request 0 can never fit: rejected; 1 does not fit now; 2 and 3 do not either after 3
    adm, left, rep = _admit(srv, _jobs([0] * 5, [True] * 5), now=1.0)
    assert rep.rejected == 1 and rep.rejected_labels == ["3x5@1/3"]
    assert adm == [2] and left == [1, 3, 4]
    srv.ctx.used = 50
    adm, left, _ = _admit(srv, _jobs([0] * 5, [True] * 5)[1:], now=1.0)
    assert adm == [2] and left == [1, 3, 4]


def test_head_wait_stops_backfilling(monkeypatch):
    srv = _server(monkeypatch, 100, [90, 10, 10], head_wait_s=5.0)
    srv.ctx.used = 50
    adm, left, _ = _admit(srv, _jobs([0, 1, 2], [True] * 3), now=4.0)
    assert adm == [1, 2] and left == [0]   # the head has waited 4 s: others pass it
    srv.ctx.used = 50
    adm, left, _ = _admit(srv, _jobs([0, 1, 2], [True] * 3), now=6.0)
    assert adm == [] and left == [0, 1, 2]   # 6 s: nothing passes it


@pytest.mark.parametrize("need, expect", [(90, []), (40, [1, 2])])
def test_head_wait_holds_while_the_head_is_in_preparation(monkeypatch, need, expect):
    srv = _server(monkeypatch, 100, [need, 10, 10], head_wait_s=5.0)
    srv.ctx.used = 50
    adm, left, _ = _admit(srv, _jobs([0, 1, 2], [False, True, True]), now=6.0)
    # a head that will not fit when prepared keeps its claim; one that fits does not block
    assert adm == expect
    assert left[0] == 0


def test_unprepared_requests_are_passed_over(monkeypatch):
    srv = _server(monkeypatch, 100, [10, 10, 10])
    adm, left, _ = _admit(srv, _jobs([0, 0, 0], [False, True, False]), now=1.0)
    assert adm == [1] and left == [0, 2]
