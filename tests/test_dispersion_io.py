"""Dispersion-file writer and reader of the library (host only, no device): byte equality with
printf formatting (ModeSolver::writeDispersion, ModeSolver.cpp:569-589), the fscanf record
semantics of ModalBroadbandPropagator::read_dispersion_file (ModalBroadbandPropagator.cpp:98-138)
on the reference's own dispersion files, multi-threaded parsing, format errors."""
import ctypes as C
import os

import numpy as np
import pytest

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def lib():
    from ncpamodes import _lib
    return _lib.load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def py_parse(path):
    """fscanf-style reference parse in Python (float() and the record rule)."""
    v = open(path).read().split()
    recs, i = [], 0
    while i < len(v):
        f, n, rs, rr = float(v[i]), int(v[i + 1]), float(v[i + 2]), float(v[i + 3])
        q = np.array([float(x) for x in v[i + 4:i + 4 + 4 * n]]).reshape(n, 4)
        recs.append((f, n, rs, rr, q))
        i += 4 + 4 * n
    return recs


def py_format(recs):
    out = []
    for f, n, rs, rr, q in recs:
        s = "%.7e   %d    %.7e   %.7e" % (f, n, rs, rr)
        for j in range(n):
            s += "   %.12e   %.12e" % (q[j, 0], q[j, 1])
            s += "   %.12e   %.12e" % (q[j, 2], q[j, 3])
        out.append(s + "\n")
    return "".join(out)


def lib_read(lib, path, threads=0):
    h = C.c_void_p()
    rc = lib.ncpa_dispersion_read(path.encode(), threads, C.byref(h))
    if rc != 0:
        return rc, None
    nrec, nm = C.c_int(), C.c_size_t()
    lib.ncpa_dispersion_info(h, C.byref(nrec), C.byref(nm))
    n, m = nrec.value, nm.value
    f, rs, rr = np.zeros(n), np.zeros(n), np.zeros(n)
    cnt = np.zeros(n, dtype=np.int32)
    q = np.zeros((4, max(m, 1)))
    lib.ncpa_dispersion_arrays(h, _dp(f), _ip(cnt), _dp(rs), _dp(rr), _dp(q[0]), _dp(q[1]), _dp(q[2]), _dp(q[3]))
    lib.ncpa_dispersion_free(h)
    recs, o = [], 0
    for r in range(n):
        recs.append((f[r], int(cnt[r]), rs[r], rr[r], q[:, o:o + cnt[r]].T.copy()))
        o += cnt[r]
    return 0, recs


def lib_write(lib, path, recs, append=False, threads=0):
    f = np.array([r[0] for r in recs])
    cnt = np.array([r[1] for r in recs], dtype=np.int32)
    rs = np.array([r[2] for r in recs])
    rr = np.array([r[3] for r in recs])
    q = np.concatenate([r[4] for r in recs]) if recs else np.zeros((0, 4))
    cols = [np.ascontiguousarray(q[:, c]) if q.size else np.zeros(1) for c in range(4)]
    return lib.ncpa_dispersion_write(path.encode(), 1 if append else 0, len(recs), _dp(f), _ip(cnt), _dp(rs),
                                     _dp(rr), _dp(cols[0]), _dp(cols[1]), _dp(cols[2]), _dp(cols[3]), threads)


def same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x[0] == y[0] and x[1] == y[1] and x[2] == y[2] and x[3] == y[3]
        assert np.array_equal(x[4], y[4])


@pytest.mark.parametrize("name", ["disp250_subset.dat", "disp256_subset.dat"])
def test_reference_files_round_trip(lib, tmp_path, name):
    """The reference's dispersion files (an older spacing and %e width) parse to the values
    Python's float() gives; rewriting them produces printf's bytes of the current format."""
    path = os.path.join(G, name)
    rc, recs = lib_read(lib, path)
    assert rc == 0
    same(recs, py_parse(path))
    out = str(tmp_path / "w.dat")
    assert lib_write(lib, out, recs) == 0
    assert open(out).read() == py_format(recs)
    rc, back = lib_read(lib, out)
    same(back, py_parse(out))


def test_large_file_threads_and_append(lib, tmp_path):
    """A broadband-sized file (> 1 MiB, parsed on several threads): the same records on 1 and
    8 threads, equal to the %.12e-rounded values; "a" appends after existing records."""
    rng = np.random.default_rng(3)
    recs = []
    for i in range(300):
        n = int(rng.integers(0, 400))
        q = rng.standard_normal((n, 4)) * np.array([1e-2, 1e-7, 1e-2, 1e-2])
        recs.append(((i + 1) / 512.0, n, 1.225 + 1e-3 * i, 1.1 - 1e-4 * i, q))
    out = str(tmp_path / "big.dat")
    assert lib_write(lib, out, recs[:100]) == 0
    assert lib_write(lib, out, recs[100:], append=True, threads=5) == 0
    text = open(out).read()
    assert len(text) > (1 << 20)
    assert text == py_format(recs)
    rc1, r1 = lib_read(lib, out, threads=1)
    rc8, r8 = lib_read(lib, out, threads=8)
    assert rc1 == rc8 == 0
    same(r1, r8)
    same(r1, py_parse(out))


def test_format_errors(lib, tmp_path):
    p = tmp_path / "bad.dat"
    p.write_text("1.0e-02 2 1.2 1.2  1 2 3 4  5 6 7\n")          # record cut short
    assert lib_read(lib, str(p))[0] != 0
    p.write_text("1.0e-02 2.5 1.2 1.2  1 2 3 4  5 6 7 8\n")      # count not a whole number
    assert lib_read(lib, str(p))[0] != 0
    p.write_text("")
    rc, recs = lib_read(lib, str(p))
    assert rc == 0 and recs == []
    p.write_text("1.0e-02 1 1.2 1.3  1 2 3 4\nend of data\n")    # reading stops at a non-number
    rc, recs = lib_read(lib, str(p))
    assert rc == 0 and len(recs) == 1 and recs[0][3] == 1.3
    assert lib_read(lib, str(tmp_path / "missing.dat"))[0] != 0
    assert lib.ncpa_dispersion_read(None, 0, C.byref(C.c_void_p())) != 0
