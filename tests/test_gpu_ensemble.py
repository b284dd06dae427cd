"""Ensembles (slevolve_amd.ensemble): every member of an ensemble run on concurrent streams
equals the same member run alone, bit for bit (iteration counts, final topography), with the
fixed kernel rule and with per-lmax tuned kernel variants; the stream-count autotuner returns one
of its candidates."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_ensemble_members_equal_solo_runs():
    from slevolve_amd import ensemble as E

    dev = torch.device("cuda", 0)
    specs = E.make_ensemble(8, lmax_choices=(64, 128), steps=(20, 60))
    ices = E.prepare(specs, dev)
    solo_slot = E._Stream(dev)
    solo = {}
    for s in specs:
        print(f"[ensemble] solo member {s.id}: lmax {s.lmax}, {s.nsteps} steps, dt {s.dt:g}", flush=True)
        solo[s.id] = E.run_one(s, ices[s.id], solo_slot)
    wall, out, _ = E.run_ensemble(specs, ices, nstreams=4, dev=dev)
    print(f"[ensemble] 8 members on 4 streams: {wall:.3f} s (alone: {sum(v[0] for v in solo.values()):.3f} s)",
          flush=True)
    for s in specs:
        a, b = solo[s.id], out[s.id]
        assert [r["numiter"] for r in a[1]] == [r["numiter"] for r in b[1]], s.id
        assert torch.equal(a[2], b[2]), s.id
    # tuned variants: every handle of an lmax takes the tuned kernels, members still equal their
    # solo runs on handles with those kernels
    var = E.tune_variants({s.lmax for s in specs}, dev)
    print(f"[ensemble] tuned variants {var}", flush=True)
    tslot = E._Stream(dev, var)
    tsolo = {s.id: E.run_one(s, ices[s.id], tslot) for s in specs[:4]}
    _, tout, _ = E.run_ensemble(specs[:4], ices, nstreams=2, dev=dev, variants=var)
    for s in specs[:4]:
        assert torch.equal(tsolo[s.id][2], tout[s.id][2]), s.id
    best, rates = E.autotune_streams(specs, ices, candidates=(1, 2, 4), probe_steps=5, dev=dev)
    print(f"[ensemble] stream autotune: best {best}, rates {rates}", flush=True)
    assert best in (1, 2, 4) and set(rates) == {1, 2, 4}


_CONCURRENT = r"""
import sys, torch
sys.path.insert(0, sys.argv[1])
from slevolve_amd import ensemble as E
dev = torch.device("cuda", 0)
specs = E.make_ensemble(12, lmax_choices=(64, 128), steps=(10, 30))
ices = E.prepare(specs, dev)
slot = E._Stream(dev)
solo = {s.id: E.run_one(s, ices[s.id], slot) for s in specs}
_, out, _ = E.run_ensemble(specs, ices, nstreams=8, dev=dev)
bad = [s.id for s in specs if [r["numiter"] for r in solo[s.id][1]] != [r["numiter"] for r in out[s.id][1]]
       or not torch.equal(solo[s.id][2], out[s.id][2])]
print("members differing from their solo runs:", bad, flush=True)
sys.exit(1 if bad else 0)
"""


def test_members_on_separate_hardware_queues():
    """With 8 hardware queues per process each of 8 streams runs truly concurrently with the others
    (the default 4 queues pair streams up): solver and transform set-up on the null stream and the
    step tables must be complete before another stream uses them.  Regression test of that
    ordering, in a child process since the queue count is read when the runtime starts."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, GPU_MAX_HW_QUEUES="8")
    print("[ensemble] child process: 12 members on 8 streams, 8 hardware queues", flush=True)
    r = subprocess.run([sys.executable, "-c", _CONCURRENT, root], env=env, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
