"""Files of a propagation run, in the layout the reference's drivers write
(pendulum_reduced_hybrid.m:178-180, 250-268, 278-293; pendulum.m likewise):

    f1, f2, ...   grid density after each step (f1 = the initial density)
    c<k>, fx<k>   plot colours and the omega marginal of step k
    stat          struct ER, Ex, Varx, EvR, ExvR, EvRvR, time index last
    MFG           struct U, S, V, Miu, Sigma, P, time index last

c, fx, stat and MFG are MATLAB level-5 MAT files with the drivers' variable
names (what their save(...) without '-v7.3' writes).  The drivers write f with
'-v7.3', which is HDF5; this image has no HDF5 library, so f is written as a
level-5 MAT file, which MATLAB's load reads under the same variable name, when
it fits that format's 2^31-byte limit per variable, and as f<k>.npy (numpy
format, memory-mappable) otherwise.  Readers here take either form.
A run resumes from the last f<k> on disk (resume_index, load_f).
"""
import os
import re

import numpy as np

STAT_FIELDS = ("ER", "Ex", "Varx", "EvR", "ExvR", "EvRvR")
MFG_FIELDS = ("U", "S", "V", "Miu", "Sigma", "P")
MAT5_LIMIT = 2 ** 31 - 1  # bytes of one variable in a level-5 MAT file


def _sio():
    import scipy.io
    return scipy.io


def save_f(path, k, f):
    """Writes the grid density of step k as f<k> (MAT level 5, or .npy when
    f exceeds the level-5 limit); returns the file name."""
    f = np.asarray(f)
    if f.nbytes < MAT5_LIMIT:
        name = os.path.join(path, "f%d.mat" % k)
        _sio().savemat(name, {"f": f}, do_compression=False)
    else:
        name = os.path.join(path, "f%d.npy" % k)
        np.save(name, f)
    return name


def load_f(path, k, mmap=True):
    """Reads f<k> written by save_f (or by MATLAB's level-5 save)."""
    mat = os.path.join(path, "f%d.mat" % k)
    if os.path.exists(mat):
        return _sio().loadmat(mat)["f"]
    return np.load(os.path.join(path, "f%d.npy" % k), mmap_mode="r" if mmap else None)


def resume_index(path):
    """The largest k with f<k> on disk (0 if none): a run restarts from load_f(path, k)."""
    ks = [int(m.group(1)) for n in os.listdir(path) for m in [re.fullmatch(r"f(\d+)\.(mat|npy)", n)] if m]
    return max(ks) if ks else 0


def save_c(path, k, c):
    _sio().savemat(os.path.join(path, "c%d.mat" % k), {"c": np.asarray(c, dtype=np.float64)})


def save_fx(path, k, fx):
    _sio().savemat(os.path.join(path, "fx%d.mat" % k), {"fx": np.asarray(fx, dtype=np.float64)})


class History:
    """Accumulates get_stat outputs (dicts from hpup.get_stat / Plan.get_stat)
    and writes them as the drivers' stat and MFG structs: a per-step matrix
    becomes (rows, cols, Nt), a per-step vector (len, Nt)."""

    def __init__(self):
        self.steps = []

    def append(self, st):
        self.steps.append({k: np.asarray(st[k], dtype=np.float64) for k in STAT_FIELDS + MFG_FIELDS})

    def _stack(self, key):
        a = [s[key] for s in self.steps]
        return np.stack(a, axis=-1)

    def struct(self, fields):
        return {k: self._stack(k) for k in fields}

    def save(self, path):
        sio = _sio()
        sio.savemat(os.path.join(path, "stat.mat"), {"stat": self.struct(STAT_FIELDS)})
        sio.savemat(os.path.join(path, "MFG.mat"), {"MFG": self.struct(MFG_FIELDS)})


def load_struct(path, name):
    """Reads stat.mat or MFG.mat back into a dict of arrays."""
    m = _sio().loadmat(os.path.join(path, name + ".mat"), squeeze_me=False)[name]
    return {k: np.asarray(m[k][0, 0]) for k in m.dtype.names}
