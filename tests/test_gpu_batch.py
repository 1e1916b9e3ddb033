"""Batched run of many uneven sequences (persistent launches, one graph per
frame step) against single-sequence handles with the same tuning: the
results must be bit-identical, frame by frame, also across repeated calls."""
import numpy as np
import pytest

from _util import say, wrapped_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import glsu  # noqa: E402

SIZES = [(96, 80), (256, 256), (300, 211), (512, 384), (130, 400), (256, 256), (77, 53)]
RRS = [1, 2, 1, 3, 1, 1, 2]


def _inputs(sizes, T, seed0=100):
    frames, phis = [], []
    for i, (W, H) in enumerate(sizes):
        amp0 = 2.0 + 3.0 * ((i * 0.37) % 1.0)
        frames.append(glsu.gen_frames(W, H, T, seed0 + i, amp0=amp0, noise=5.0 + 4.0 * i))
        phis.append(glsu.synth.phase(W, H, amp0, 0.05))
    return frames, phis


@pytest.mark.parametrize("graph", [True, False])
def test_batch_is_bitwise_equal_to_single_runs(graph):
    T = 5
    frames, phis = _inputs(SIZES, T)
    b = glsu.DPRABatch(phis, RRS)
    say(f"[batch] {len(SIZES)} sequences, info={b.info()}, graph={graph}")
    outs = b(frames, graph=graph)
    outs2 = b(frames, graph=graph)  # a second call continues from the state left
    for i, (W, H) in enumerate(SIZES):
        d = glsu.DPRA(phis[i], RRS[i], tuning=b.tuning(i))
        ref = d(frames[i]).cpu().numpy()
        ref2 = d(frames[i]).cpu().numpy()
        got, got2 = outs[i].cpu().numpy(), outs2[i].cpu().numpy()
        same = np.array_equal(got, ref) and np.array_equal(got2, ref2)
        say(f"   seq {i} {W}x{H} rr={RRS[i]} tuning={b.tuning(i)}: bitwise equal={same}")
        assert same, (i, np.abs(got - ref).max(), np.abs(got2 - ref2).max())
        p1, d1 = b.get_state(i)
        p2, d2 = d.get_state()
        assert np.array_equal(p1, p2) and np.array_equal(d1, d2)


def test_batch_with_explicit_tunings_and_oracle(oracle):
    sizes = [(200, 150), (128, 128)]
    T = 3
    frames, phis = _inputs(sizes, T, seed0=300)
    plans = {p["id"]: p for p in glsu.fft_plans()}
    narrow = [i for i, p in plans.items() if p["T"] != 64 and p["N"] >= 150 + 30][0]
    tunings = [dict(plany=narrow, cols_waves=2), dict(rows_waves=2)]
    b = glsu.DPRABatch(phis, 1, tunings=tunings)
    assert b.tuning(0)["plany"] == narrow and b.info()["classes"] == 2
    outs = b(frames)
    for i, (W, H) in enumerate(sizes):
        host = frames[i].cpu().numpy().astype(np.float64)
        ref = oracle.DPRA(phis[i]).run(host)
        for t in range(T):
            rms, mx = wrapped_err(outs[i][t].cpu().numpy(), ref[t])
            say(f"[batch vs oracle] seq {i} frame {t}: rms={rms:.2e} max={mx:.2e}")
            assert rms < 1e-3


def test_batch_long_transform_plans_are_bitwise_equal_to_single_runs():
    # plans with more than 24 elements per lane (64 x 36 = 2304 points) take
    # the kernels' read-ahead branches; forced here on small images so the
    # batched and single-frame kernels run them on the same inputs
    sizes = [(160, 120), (96, 200)]
    T = 3
    frames, phis = _inputs(sizes, T, seed0=700)
    plans = {p["id"]: p for p in glsu.fft_plans()}
    long_plan = [i for i, p in plans.items() if p["T"] == 64 and p["E"] > 24][0]
    tuning = dict(planx=long_plan, plany=long_plan, cols_waves=2, rows_waves=2)
    b = glsu.DPRABatch(phis, 1, tunings=[tuning, tuning])
    outs = b(frames)
    for i, (W, H) in enumerate(sizes):
        assert b.tuning(i)["plany"] == long_plan and b.tuning(i)["planx"] == long_plan
        ref = glsu.DPRA(phis[i], 1, tuning=b.tuning(i))(frames[i]).cpu().numpy()
        got = outs[i].cpu().numpy()
        say(f"[batch long plan {plans[long_plan]['N']}] seq {i} {W}x{H}: "
            f"bitwise equal={np.array_equal(got, ref)}")
        assert np.array_equal(got, ref), np.abs(got - ref).max()


def test_batch_uneven_workload_shape():
    """Sizes drawn from the uneven set (scaled down to 3 sizes, 12 sequences,
    3 frames): runs, and three sequences equal their single-handle runs."""
    rng = np.random.default_rng(5)
    sizes = [(int(s), int(s)) for s in rng.choice([256, 512, 768], 12)]
    frames, phis = _inputs(sizes, 3, seed0=500)
    b = glsu.DPRABatch(phis, 1)
    outs = b(frames, timed=True)
    say(f"[batch12] sizes={[s[0] for s in sizes]} classes={b.info()['classes']} "
        f"{b.last_ms:.2f} ms for 3 frames")
    for i in (0, 5, 11):
        ref = glsu.DPRA(phis[i], 1, tuning=b.tuning(i))(frames[i]).cpu().numpy()
        assert np.array_equal(outs[i].cpu().numpy(), ref)
