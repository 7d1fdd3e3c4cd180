"""Threads that ask for the same lens's device handle at once share one
DeviceLens: a second one would replace the first and free its native lens
while the first thread's call still uses the handle (host-only: creating a
device lens copies the prescription, no device work)."""
import threading

from ghostlight_amd import runtime, synthetic


def test_concurrent_device_lens_is_created_once(monkeypatch):
    lens = synthetic.config1_lens()
    made = []
    real = runtime.DeviceLens

    class Counting(real):
        def __init__(self, l):
            made.append(1)
            super().__init__(l)

    monkeypatch.setattr(runtime, "DeviceLens", Counting)
    start = threading.Barrier(8)
    got = [None] * 8

    def work(i):
        start.wait()
        got[i] = lens.device_lens()

    ts = [threading.Thread(target=work, args=(i,)) for i in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert len(made) == 1
    assert all(g is got[0] for g in got)
    assert lens.device_lens() is got[0]
