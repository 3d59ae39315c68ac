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
Every file is written under a temporary name and renamed into place, so a
run resumes from the last complete f<k> on disk (resume_index, load_f).
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


def _atomic(name, write):
    """Writes through a temporary name in the same directory and renames it
    into place, so a run killed mid-write leaves no truncated file under the
    final name (resume_index only ever sees complete files)."""
    d, base = os.path.split(name)
    tmp = os.path.join(d, ".%s.tmp%d" % (base, os.getpid()))
    try:
        with open(tmp, "wb") as fh:
            write(fh)
            fh.flush()
            os.fsync(fh.fileno())
        os.replace(tmp, name)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)
    return name


def save_f(path, k, f):
    """Writes the grid density of step k as f<k> (MAT level 5, or .npy when
    f exceeds the level-5 limit); returns the file name."""
    f = np.asarray(f)
    if f.nbytes < MAT5_LIMIT:
        return _atomic(os.path.join(path, "f%d.mat" % k), lambda fh: _sio().savemat(fh, {"f": f},
                                                                                  do_compression=False))
    return _atomic(os.path.join(path, "f%d.npy" % k), lambda fh: np.save(fh, f))


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
    _atomic(os.path.join(path, "c%d.mat" % k), lambda fh: _sio().savemat(fh, {"c": np.asarray(c, dtype=np.float64)}))


def save_fx(path, k, fx):
    _atomic(os.path.join(path, "fx%d.mat" % k),
            lambda fh: _sio().savemat(fh, {"fx": np.asarray(fx, dtype=np.float64)}))


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
        _atomic(os.path.join(path, "stat.mat"), lambda fh: sio.savemat(fh, {"stat": self.struct(STAT_FIELDS)}))
        _atomic(os.path.join(path, "MFG.mat"), lambda fh: sio.savemat(fh, {"MFG": self.struct(MFG_FIELDS)}))


def load_struct(path, name):
    """Reads stat.mat or MFG.mat back into a dict of arrays."""
    m = _sio().loadmat(os.path.join(path, name + ".mat"), squeeze_me=False)[name]
    return {k: np.asarray(m[k][0, 0]) for k in m.dtype.names}
