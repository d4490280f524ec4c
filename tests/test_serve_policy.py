"""The serving runtime's admission policy (einspline_server_plan_admission)
on the CPU: refusal, deferral and eviction decisions, the serving
workload's three tables under a 4.3 GB budget, and every decision on small
random cases against an exhaustive restatement of the rule (the set of
tables with the fewest bytes that frees enough room; evict it if all of it
is idle, otherwise wait for its busy members)."""
import itertools

import numpy as np
import pytest

import einspline_amd as ea
from einspline_amd.serve import ADMIT_DEFER, ADMIT_RESIDENT, ADMIT_UPLOAD, plan_admission

T1, T2, T3 = 135_834_624, 307_981_312, 3_974_344_704     # 48^3x256 F32, 64^3x512 BF16, 96^3x1024 F32
BUDGET = 4_300_000_000


def tab(b, state=0, busy=0, pending=0, last_use=0):
    return {"bytes": b, "state": state, "busy": busy, "pending": pending, "last_use": last_use}


def used_of(ts):
    return sum(t["bytes"] for t in ts if t["state"])


def test_refuse_resident_and_free_room():
    ts = [tab(T1), tab(5 * 10**9)]
    assert plan_admission(ts, 1, 0, BUDGET)[0] == ea.E_BUDGET
    ts = [tab(T1, state=1), tab(T2)]
    assert plan_admission(ts, 0, used_of(ts), BUDGET) == (ADMIT_RESIDENT, [], [])
    assert plan_admission(ts, 1, used_of(ts), BUDGET) == (ADMIT_UPLOAD, [], [])
    ts = [tab(T1, state=2), tab(T2)]                     # already being uploaded
    assert plan_admission(ts, 0, used_of(ts), BUDGET) == (ADMIT_DEFER, [], [])


def test_serving_tables_keep_the_large_one_resident():
    # T2 and T3 resident and idle, T1 wanted: evict T2 (308 MB), not T3 (4 GB)
    ts = [tab(T1, pending=3), tab(T2, state=1, last_use=5), tab(T3, state=1, last_use=1)]
    assert plan_admission(ts, 0, used_of(ts), BUDGET) == (ADMIT_UPLOAD, [1], [])
    # T2 busy: wait for it instead of evicting the idle T3
    ts[1]["busy"] = 1
    assert plan_admission(ts, 0, used_of(ts), BUDGET) == (ADMIT_DEFER, [], [1])
    # T1 and T3 resident, T2 wanted (118 MB short): T1 goes, or is drained when busy
    ts = [tab(T1, state=1), tab(T2, pending=1), tab(T3, state=1)]
    assert plan_admission(ts, 1, used_of(ts), BUDGET) == (ADMIT_UPLOAD, [0], [])
    ts[0]["busy"] = 1
    assert plan_admission(ts, 1, used_of(ts), BUDGET) == (ADMIT_DEFER, [], [0])
    # T1 still uploading: its reservation is waited for, T3 stays
    ts = [tab(T1, state=2), tab(T2, pending=1), tab(T3, state=1)]
    assert plan_admission(ts, 1, used_of(ts), BUDGET) == (ADMIT_DEFER, [], [0])
    # T3 wanted with T1 and T2 resident: T1 alone frees enough
    ts = [tab(T1, state=1), tab(T2, state=1), tab(T3, pending=1)]
    assert plan_admission(ts, 2, used_of(ts), BUDGET) == (ADMIT_UPLOAD, [0], [])


def test_ties_prefer_tables_without_queued_requests_then_least_recent():
    ts = [tab(100), tab(100, state=1, pending=2, last_use=1), tab(100, state=1, last_use=9)]
    assert plan_admission(ts, 0, 200, 250)[1] == [2]
    ts[2]["pending"] = 1
    assert plan_admission(ts, 0, 200, 250)[1] == [1]     # both queued: the older one


def test_greedy_past_sixteen_candidates():
    ts = [tab(10)] + [tab(10, state=1, last_use=100 - i, busy=i % 3 == 0) for i in range(1, 21)]
    used = used_of(ts)
    rc, victims, drain = plan_admission(ts, 0, used, used - 15)
    # 25 bytes short: the three least recently used idle tables
    idle = sorted((t["last_use"], i) for i, t in enumerate(ts) if i and not t["busy"])
    assert (rc, victims, drain) == (ADMIT_UPLOAD, sorted(i for _, i in idle[:3]), [])


def test_bad_arguments():
    ts = [tab(10), tab(10, state=3)]
    assert plan_admission(ts, 0, 0, 100)[0] == ea.E_INVAL
    assert plan_admission(ts[:1], 1, 0, 100)[0] == ea.E_INVAL


def _expected(ts, want, used, budget):
    """Exhaustive restatement of the rule."""
    w = ts[want]
    if w["state"] == 1:
        return ADMIT_RESIDENT, [], []
    if w["state"] == 2:
        return ADMIT_DEFER, [], []
    if w["bytes"] > budget:
        return ea.E_BUDGET, [], []
    if used + w["bytes"] <= budget:
        return ADMIT_UPLOAD, [], []
    short = used + w["bytes"] - budget
    cand = [i for i, t in enumerate(ts) if i != want and t["state"]]
    best = None
    for r in range(1, len(cand) + 1):
        for sub in itertools.combinations(cand, r):
            b = sum(ts[i]["bytes"] for i in sub)
            if b < short:
                continue
            key = (b, sum(ts[i]["pending"] > 0 for i in sub), max(ts[i]["last_use"] for i in sub))
            if best is None or key < best[0]:
                best = (key, sorted(sub))
    if best is None:
        return ADMIT_DEFER, [], []
    drain = [i for i in best[1] if ts[i]["busy"] or ts[i]["state"] == 2]
    return (ADMIT_DEFER, [], drain) if drain else (ADMIT_UPLOAD, best[1], [])


@pytest.mark.parametrize("seed", range(40))
def test_random_cases_against_exhaustive_rule(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 8))
    ts = []
    for i in range(n):
        st = int(rng.choice([0, 1, 1, 2]))
        ts.append(tab(int(rng.integers(1, 50)) * 16, st, int(st and rng.random() < 0.3),
                      int(rng.integers(0, 3)), int(rng.permutation(1000)[i])))
    used = used_of(ts)
    budget = int(used * rng.uniform(0.6, 1.2)) + 16
    for want in range(n):
        got = plan_admission(ts, want, used, budget)
        exp = _expected(ts, want, used, budget)
        if exp[0] == ADMIT_UPLOAD and exp[1]:
            # equal-key ties may pick another set of the same bytes, queue count and age
            assert got[0] == ADMIT_UPLOAD
            key = lambda s: (sum(ts[i]["bytes"] for i in s), sum(ts[i]["pending"] > 0 for i in s),
                             max(ts[i]["last_use"] for i in s))
            assert key(got[1]) == key(exp[1]) and sum(ts[i]["bytes"] for i in got[1]) >= used + ts[want]["bytes"] - budget
        else:
            assert got == exp, (seed, want, ts, used, budget)
