"""The particle tiling of the push launches and the autotuner that picks it:
the state, the field, the diagnostics and the sweep counts of a run are the
same bit for bit under every tiling (FP64 and mixed), the autotuner leaves the
simulation where it found it, and bad tilings are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T = 6
PICKED = (0, 5, 11)


def run_with(precision, tiling=None, autotune_at=None, n=12):
    import picgolf
    from picgolf import configs
    cfgs, states, _ = configs.ensemble(n, precision=precision, seed=3)
    with picgolf.Context(cfgs) as c:
        for i, (x, v) in enumerate(states):
            c.upload(x, v, inst=i)
        if tiling is not None:
            c.set_tiling(*tiling)
        rows = []
        if autotune_at is None:
            rows.append(c.run(T))
        else:
            rows.append(c.run(autotune_at))
            tuned = c.autotune(trial_steps=2)
            assert tuned["candidates"] >= 8
            assert c.tiling()[:2] == (tuned["tile_particles"], tuned["block"])
            rows.append(c.run(T - autotune_at))
        D = np.concatenate([r[0] for r in rows], axis=1)
        S = np.concatenate([r[1] for r in rows], axis=1)
        return D, S, [c.download(inst=i) for i in PICKED], c.tiling()


def identical(a, b):
    Da, Sa, sa, _ = a
    Db, Sb, sb, _ = b
    assert np.array_equal(Da, Db) and np.array_equal(Sa, Sb)
    for (xa, va, Ea), (xb, vb, Eb) in zip(sa, sb):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(Ea, Eb)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_results_do_not_depend_on_the_tiling(precision):
    ref = run_with(precision)
    for tiling in ((64, 64), (1000, 256), (5000, 512), (1 << 20, 1024)):
        got = run_with(precision, tiling)
        assert got[3][0] == max(64, -(-min(tiling[0], 1 << 20) // 64) * 64)
        identical(ref, got)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("at", [0, 3])
def test_autotune_leaves_the_run_unchanged(precision, at):
    identical(run_with(precision), run_with(precision, autotune_at=at))


def test_bad_tilings_are_refused():
    import picgolf
    from picgolf import configs
    cfgs, states, _ = configs.ensemble(2, precision="mixed")
    with picgolf.Context(cfgs) as c:
        before = c.tiling()
        for t, b in ((0, 256), (4096, 48), (4096, 2048), (4096, 384)):
            with pytest.raises(picgolf.PicError) as e:
                c.set_tiling(t, b)
            assert e.value.code == -1
        assert c.tiling() == before
        with pytest.raises(picgolf.PicError):
            c.autotune(trial_steps=0)
