"""Admission order of the serving loop (host logic, no GPU): shortest work
first, with a waiting-time bound so a long request is not starved by a
stream of short arrivals."""
from wake_amd import serve


def _req(rid, n, steps, arrival, kind="bluff3d"):
    return serve.SimRequest(rid, kind, n, steps, re=200.0, arrival=arrival)


def test_least_work_orders_by_node_steps_then_arrival():
    q = [_req(0, 192, 20000, 0.0), _req(1, 64, 2000, 0.5), _req(2, 64, 2000, 0.2), _req(3, 128, 5000, 0.1)]
    order = [r.rid for r in serve.admission_order(q, now=1.0, policy="least_work", max_wait_s=60.0)]
    assert order == [2, 1, 3, 0]


def test_long_request_is_not_starved_by_short_arrivals():
    long_req = _req(0, 192, 20000, 0.0)
    # short requests keep arriving; before the bound the long one stays last
    for t in (5.0, 30.0, 59.0):
        shorts = [_req(i, 64, 2000, t - 0.01 * i) for i in range(1, 6)]
        order = serve.admission_order([long_req] + shorts, now=t, policy="least_work", max_wait_s=60.0)
        assert order[-1].rid == 0
    # past the bound it goes ahead of every request that has not waited as long
    shorts = [_req(i, 64, 2000, 60.5 - 0.01 * i) for i in range(1, 6)]
    order = serve.admission_order([long_req] + shorts, now=61.0, policy="least_work", max_wait_s=60.0)
    assert order[0].rid == 0
    # several overdue requests: oldest first
    old2 = _req(9, 160, 10000, 0.5)
    order = serve.admission_order([old2, long_req] + shorts, now=61.0, policy="least_work", max_wait_s=60.0)
    assert [r.rid for r in order[:2]] == [0, 9]


def test_arrival_policy_keeps_arrival_order():
    q = [_req(0, 64, 2000, 0.3), _req(1, 192, 20000, 0.1), _req(2, 96, 3000, 0.2)]
    assert [r.rid for r in serve.admission_order(q, now=1.0, policy="arrival")] == [1, 2, 0]
