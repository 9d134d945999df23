"""Checkpoint and restart: a run restored from a checkpoint continues bitwise as
the uninterrupted run (particles in their order with ids, the ion charge and the
potential after two more steps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_restart_continues_bitwise(tmp_path):
    from mcstreamer_amd import checkpoint, streamer
    cfg = streamer.make_config1(n=(32, 32, 64), npairs=20000, seed=3)
    sim = streamer.Streamer(cfg["popl"], cfg["fields"], cfg["ws"], cfg["eb"], cfg["dt"], cfg["ntarget"],
                            cfg["nmax"], scheme="cic")
    for _ in range(2):
        sim.step()
    path = str(tmp_path / "run.npz")
    checkpoint.save(sim, path)
    for _ in range(2):
        sim.step()
    a = sim.popl.download()
    ua, qa = sim.fields.get("u"), sim.fields.get("qfixed")
    sim2 = checkpoint.load(path)
    assert sim2.i == 2
    for _ in range(2):
        sim2.step()
    b = sim2.popl.download()
    print(f"particles {len(a)} / {len(b)}", flush=True)
    assert len(a) == len(b) and a.next_id == b.next_id
    for k in ("x", "v", "w", "s", "active", "id"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    np.testing.assert_array_equal(ua, sim2.fields.get("u"))
    np.testing.assert_array_equal(qa, sim2.fields.get("qfixed"))
