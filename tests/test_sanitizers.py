"""Host sanitizers over the C++ test infrastructure (CPU).

The oracle and the synthetic generator, built with AddressSanitizer +
UndefinedBehaviorSanitizer and with ThreadSanitizer (oracle/sanitize.py), run
a workload that exercises their worker threads; no finding is allowed.  A
planted mistake (a read past a heap buffer, a data race between two threads)
run through the same harness must be reported, which shows that the
sanitizer runtime is really active there."""
import pytest

from oracle import sanitize


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_oracle_and_generator_are_clean(kind):
    rc, out, err = sanitize.run(kind)
    assert rc == 0, err[-3000:]
    assert "sanitized workload ok" in out
    for mark in ("AddressSanitizer", "ThreadSanitizer", "runtime error:"):
        assert mark not in err, err[-3000:]


@pytest.mark.parametrize("kind,mark", [("asan", "heap-buffer-overflow"), ("tsan", "data race")])
def test_planted_mistake_is_reported(kind, mark):
    rc, err = sanitize.run_planted(kind)
    assert rc != 0
    assert mark in err, err[-3000:]
