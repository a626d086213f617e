"""The batched ensemble (many independent two-stream instances with uneven N,
P, W and beam speed in one context): FP64 instances bit for bit against
independent oracle runs, and the mixed batch against the FP64 batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ens():
    import picgolf
    from picgolf import configs
    out = {}
    T = 24
    for prec in ("fp64", "mixed"):
        cfgs, states, _ = configs.ensemble(512, precision=prec)
        with picgolf.Context(cfgs) as c:
            for i, (x, v) in enumerate(states):
                c.upload(x, v, inst=i)
            D, S = c.run(T)
            picked = {i: c.download(inst=i) for i in (0, 1, 77, 200, 311, 405, 510, 511)}
        out[prec] = (cfgs, states, D, S, picked)
    return T, out


def test_fp64_instances_match_independent_oracle_runs(ens, orc):
    T, out = ens
    cfgs, states, D, S, picked = out["fp64"]
    for i, (x, v, E) in picked.items():
        k = cfgs[i]
        xr, vr, Er, Dr, Sr = orc.run(k.N, 7, 1, k.W, k.w, k.dt, states[i][0], states[i][1], T, rtol=k.rtol,
                                     nthreads=16)
        assert np.array_equal(x, xr) and np.array_equal(v, vr) and np.array_equal(E, Er), i
        assert np.array_equal(D[i], Dr) and np.array_equal(S[i], Sr), i


def test_mixed_ensemble_against_fp64(ens):
    T, out = ens
    Df, Sf = out["fp64"][2], out["fp64"][3]
    Dm, Sm = out["mixed"][2], out["mixed"][3]
    rel = np.abs(Dm[:, :, 0] - Df[:, :, 0]) / np.abs(Df[:, :, 0])
    print(f"ensemble mixed vs FP64 over {T} steps: field energy max rel err {rel.max():.3e}, "
          f"median {np.median(rel.max(axis=1)):.3e}; mean sweeps {Sm.mean():.2f} vs {Sf.mean():.2f}")
    assert rel.max() < 1e-6
    # momentum of every instance: the gather's FP32 weights and the deposit's
    # 2^-30 quanta differ by ~1e-9 of a particle's charge, so action and
    # reaction cancel to about that (the FP64 path conserves it to roundoff)
    assert np.abs(Dm[:, :, 3] - Dm[:, :1, 3]).max() < 2e-9
    assert np.abs(Df[:, :, 3] - Df[:, :1, 3]).max() < 1e-13
    assert abs(Sm.mean() - Sf.mean()) < 0.1
