"""Run files in the drivers' layout (pendulum_reduced_hybrid.m:178-180,
250-268, 278-293): level-5 MAT files with the drivers' variable names, f as
.npy above the level-5 size limit, the stat/MFG structs with the time index
last, and resume from the last f<k>.  Host only."""
import numpy as np
import scipy.io

import hpup.snapshot as S


def _stat(rng, dim=2):
    return {"ER": rng.standard_normal((3, 3)), "Ex": rng.standard_normal(dim), "Varx": rng.standard_normal((dim, dim)),
            "EvR": rng.standard_normal(3), "ExvR": rng.standard_normal((dim, 3)), "EvRvR": rng.standard_normal((3, 3)),
            "U": rng.standard_normal((3, 3)), "S": np.diag(rng.standard_normal(3)), "V": rng.standard_normal((3, 3)),
            "Miu": rng.standard_normal(dim), "Sigma": rng.standard_normal((dim, dim)), "P": rng.standard_normal((dim, 3)),
            "niter": 3, "fit_ok": True}


def test_f_round_trip_and_resume(tmp_path, monkeypatch):
    rng = np.random.default_rng(0)
    f1 = np.asfortranarray(rng.random((4, 4, 4, 6, 6)))
    assert S.resume_index(tmp_path) == 0
    name = S.save_f(tmp_path, 1, f1)
    assert name.endswith("f1.mat")
    # the drivers' variable name, readable by any level-5 MAT reader
    assert np.array_equal(scipy.io.loadmat(name)["f"], f1)
    # above the level-5 limit: numpy format, same reader
    monkeypatch.setattr(S, "MAT5_LIMIT", 1024)
    f2 = f1 * 2
    assert S.save_f(tmp_path, 2, f2).endswith("f2.npy")
    assert S.resume_index(tmp_path) == 2
    g = S.load_f(tmp_path, 2)
    assert np.array_equal(np.asarray(g), f2) and g.shape == f2.shape
    assert np.array_equal(S.load_f(tmp_path, 1), f1)


def test_history_structs(tmp_path):
    rng = np.random.default_rng(1)
    h = S.History()
    sts = [_stat(rng) for _ in range(5)]
    for st in sts:
        h.append(st)
    h.save(tmp_path)
    stat = S.load_struct(tmp_path, "stat")
    mfg = S.load_struct(tmp_path, "MFG")
    assert set(stat) == set(S.STAT_FIELDS) and set(mfg) == set(S.MFG_FIELDS)
    # shapes of the drivers' arrays (pendulum_reduced_hybrid.m:187-199)
    assert stat["ER"].shape == (3, 3, 5) and stat["Ex"].shape == (2, 5) and stat["ExvR"].shape == (2, 3, 5)
    assert mfg["P"].shape == (2, 3, 5) and mfg["Miu"].shape == (2, 5) and mfg["S"].shape == (3, 3, 5)
    for t, st in enumerate(sts):
        assert np.array_equal(stat["ER"][:, :, t], st["ER"])
        assert np.array_equal(stat["Ex"][:, t], st["Ex"])
        assert np.array_equal(mfg["Sigma"][:, :, t], st["Sigma"])


def test_c_fx_files(tmp_path):
    c = np.arange(12.0).reshape(4, 3)
    fx = np.arange(36.0).reshape(6, 6)
    S.save_c(tmp_path, 3, c)
    S.save_fx(tmp_path, 3, fx)
    assert np.array_equal(scipy.io.loadmat(tmp_path / "c3.mat")["c"], c)
    assert np.array_equal(scipy.io.loadmat(tmp_path / "fx3.mat")["fx"], fx)


def test_interrupted_write_leaves_no_partial_file(tmp_path, monkeypatch):
    """A write that dies midway (here: the MAT writer raises after writing part
    of the file) leaves neither a truncated f<k> under the final name nor its
    temporary, so resume_index still points at the last complete step."""
    rng = np.random.default_rng(3)
    f1 = np.asfortranarray(rng.random((4, 4, 4, 3, 3)))
    S.save_f(tmp_path, 1, f1)
    real = scipy.io.savemat

    def dying(fh, *a, **k):
        fh.write(b"MATLAB 5.0 MAT-file, truncated")
        raise OSError("disk full")
    monkeypatch.setattr(scipy.io, "savemat", dying)
    try:
        S.save_f(tmp_path, 2, f1 * 2)
    except OSError:
        pass
    monkeypatch.setattr(scipy.io, "savemat", real)
    assert sorted(p.name for p in tmp_path.iterdir()) == ["f1.mat"]
    assert S.resume_index(tmp_path) == 1
    assert np.array_equal(S.load_f(tmp_path, 1), f1)
