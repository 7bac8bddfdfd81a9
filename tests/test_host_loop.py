"""The host-side run! loop of the Python mirror (CPU only): the stopping
criteria and finalize_beginning follow src/schemes/tnrscheme.jl:33-59 and
src/utility/stopping.jl:1-42.  A stand-in scheme replaces the device calls,
so this checks the loop logic alone; the device path of the same loop is
compared with tnr_run in tests/test_gpu_parity.py."""
import pytest


class Fake:
    def __init__(self):
        self.steps = 0


def _patch(tk, monkeypatch):
    sch = tk.schemes

    def fake_step(scheme, trunc):
        scheme.steps += 1
        return scheme

    monkeypatch.setattr(sch, "step", fake_step)
    # norms 1/(steps+1): a convergent sequence for convcrit
    return tk.Finalizer(lambda s: 1.0 / (s.steps + 1))


def test_maxiter_counts_and_finalize_beginning(tk, monkeypatch):
    fin = _patch(tk, monkeypatch)
    s = Fake()
    d = tk.run(s, tk.truncrank(4), tk.maxiter(3), fin, verbosity=0)
    assert d == [1.0, 1 / 2, 1 / 3, 1 / 4] and s.steps == 3
    s = Fake()
    d = tk.run(s, tk.truncrank(4), tk.maxiter(3), fin, finalize_beginning=False, verbosity=0)
    assert d == [1 / 2, 1 / 3, 1 / 4]


def test_maxiter_zero_still_steps_once(tk, monkeypatch):
    # `while crit` evaluates the criterion after the first step
    fin = _patch(tk, monkeypatch)
    s = Fake()
    d = tk.run(s, tk.truncrank(4), tk.maxiter(0), fin, verbosity=0)
    assert s.steps == 1 and len(d) == 2


def test_convcrit_stops_when_delta_reached(tk, monkeypatch):
    fin = _patch(tk, monkeypatch)
    s = Fake()
    crit = tk.convcrit(0.2, lambda steps, data: data[-1])    # continue while 0.2 < last
    d = tk.run(s, tk.truncrank(4), crit, fin, verbosity=0)
    assert d[-1] <= 0.2 < d[-2] and s.steps == 4


def test_multiple_criteria_stop_at_the_first_failing(tk, monkeypatch):
    fin = _patch(tk, monkeypatch)
    s = Fake()
    crit = tk.maxiter(2) & tk.convcrit(1e-9, lambda steps, data: data[-1])
    d = tk.run(s, tk.truncrank(4), crit, fin, verbosity=0)
    assert s.steps == 2 and len(d) == 3
    assert "Maximum amount of iterations" in crit.info(2, d)
