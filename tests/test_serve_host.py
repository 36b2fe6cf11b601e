"""Host logic of the server that needs no GPU: which running jobs a queued job preempts."""
from types import SimpleNamespace

from opensol_amd.serve import choose_victims


def _job(name, admitted, size, preemptions=0, arrival=0.0):
    return SimpleNamespace(name=name, admitted=admitted, bytes=size, preemptions=preemptions, arrival=arrival)


def test_latest_admitted_jobs_are_preempted_first_until_the_job_fits():
    running = [_job('a', 0.0, 40), _job('b', 1.0, 30), _job('c', 2.0, 20)]
    assert [j.name for j in choose_victims(45, 10, running)] == ['c', 'b']   # 10 + 20 < 45 <= 10 + 20 + 30
    assert [j.name for j in choose_victims(30, 10, running)] == ['c']
    assert choose_victims(5, 10, running) == []                               # fits already


def test_a_job_is_preempted_at_most_once_and_partial_sets_are_not_taken():
    running = [_job('a', 0.0, 40), _job('b', 1.0, 30, preemptions=1), _job('c', 2.0, 20)]
    assert [j.name for j in choose_victims(60, 5, running)] == ['c', 'a']   # b was preempted before
    assert choose_victims(100, 5, running) == []                            # even all would not make room


def test_jobs_admitted_together_give_way_in_reverse_arrival_order():
    running = [_job('early', 1.0, 30, arrival=0.0), _job('late', 1.0, 30, arrival=0.5)]
    assert [j.name for j in choose_victims(35, 10, running)] == ['late']
