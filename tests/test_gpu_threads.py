"""Two host threads integrating different batches at once (ctypes releases the GIL, so the calls
overlap in the library): each thread's results equal the serial ones bit for bit.  The
integration entry points hold the library's launch lock from the choice of launch shape to the
kernel's completion, so that the calls do not share one work queue or scratch buffer."""
import threading

import numpy as np
import pytest

from conftest import population

pytestmark = pytest.mark.gpu


def test_concurrent_integrations_match_serial(gpu, ephem, model):
    from assist_amd import batch
    jobs = []
    for k, (n, n_var, span) in enumerate([(300, 6, 120.0), (500, 0, 200.0), (200, 2, -90.0), (400, 0, 60.0)]):
        states, params, _ = population(model, n, seed=500 + k, n_var=n_var, ng_fraction=0.2)
        jobs.append((np.zeros(n), np.full(n, span), states, params))
    serial = [batch.integrate(ephem, *j) for j in jobs]
    results = [None] * len(jobs)
    errors = []

    def work(i):
        try:
            for _ in range(3):
                results[i] = batch.integrate(ephem, *jobs[i])
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(jobs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not errors, errors
    for r, s in zip(results, serial):
        assert (r.status == 0).all()
        np.testing.assert_array_equal(r.states, s.states)
        np.testing.assert_array_equal(r.steps, s.steps)
