"""The HDF5 export of the grid fields (MCStreamer/src/plot.jl:210-240 h5export,
SavedGridFields of io.jl:9-44) on the host: the writer's files are read back
by the HDF5 library itself (h5py, h5dump) value for value, and h5diff finds no
difference from the same content written by the library; the export's group,
names, dimensions (HDF5.jl's (N, M) for a Julia M x N matrix) and types; the
grid's LinRanges against their closed forms; and the oracle's resample!
counts p and wtotal against a brute-force restatement and closed forms."""
import os
import subprocess

import numpy as np
import pytest

from mcstreamer_amd import axisym, h5
from mcstreamer_amd import constants as co
from oracle import refcpu as R
from tests.common import H5DIFF, H5DUMP, h5read, h5write_lib, have_h5py

needs_h5 = pytest.mark.skipif(not have_h5py(), reason="the HDF5 library (h5py under Python 3.9) is absent")


def _tree(seed=0):
    rng = np.random.default_rng(seed)
    return {
        "grid": {"R": np.float64(1.5e-3), "L": np.float64(2.0), "M": np.int64(5), "N": np.int64(7),
                 "rc": rng.normal(size=5), "deeper": {"x": rng.normal(size=(2, 3, 4))}},
        "u": rng.normal(size=(7, 5)),
        "p": rng.integers(-2**62, 2**62, size=(7, 5)),
        "tiny": np.array([np.finfo(float).tiny, -0.0, np.inf, -np.inf, 5e-324]),
        "empty": np.zeros((0, 3)),
        "scalar_int": np.int64(-3),
    }


def _flat(tree, prefix=""):
    out = {}
    for k, v in tree.items():
        if isinstance(v, dict):
            out.update(_flat(v, prefix + k + "/"))
        else:
            out[prefix + k] = np.asarray(v)
    return out


@needs_h5
def test_writer_read_back_by_the_hdf5_library(tmp_path):
    t = _tree()
    path = str(tmp_path / "t.h5")
    h5.write(path, t)
    data, meta = h5read(path)
    want = _flat(t)
    assert sorted(data) == sorted(want)
    for k, v in want.items():
        got = data[k]
        assert got.shape == v.shape, k
        assert got.dtype == (np.int64 if v.dtype.kind in "iu" else np.float64), k
        np.testing.assert_array_equal(got, v, err_msg=k)
        assert np.array_equal(np.signbit(got), np.signbit(v)), k
    assert meta["grid/M"] == ("<i8", "()")
    assert meta["u"] == ("<f8", "(7, 5)")
    if os.path.exists(H5DUMP):
        r = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True)
        assert r.returncode == 0 and "H5T_STD_I64LE" in r.stdout and "SCALAR" in r.stdout, r.stderr


@needs_h5
def test_writer_equals_the_library_file_under_h5diff(tmp_path):
    t = _tree(1)
    ours, theirs = str(tmp_path / "ours.h5"), str(tmp_path / "lib.h5")
    h5.write(ours, t)
    h5write_lib(theirs, t)
    if not os.path.exists(H5DIFF):
        pytest.skip("h5diff is absent")
    r = subprocess.run([H5DIFF, "-r", ours, theirs], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_writer_refuses_what_it_cannot_write(tmp_path):
    with pytest.raises(ValueError):
        h5.write(str(tmp_path / "a.h5"), {"a/b": 1.0})
    with pytest.raises(ValueError):
        h5.write(str(tmp_path / "a.h5"), {f"d{i:02d}": 1.0 for i in range(33)})
    with pytest.raises(TypeError):
        h5.write(str(tmp_path / "a.h5"), {"s": np.array(["x"])})


def test_grid_ranges_are_the_linranges():
    for (Rr, L, M, N) in [(1e-3, 2e-3, 8, 16), (4e-3, 4e-2, 1536, 10240), (3e-3, 1e-2, 300, 1000)]:
        g = axisym.Grid(Rr, L, M, N)
        rf, zf, rc, zc = g.rf, g.zf, g.rc, g.zc
        assert rf[0] == 0.0 and rf[-1] == Rr and zf[0] == 0.0 and zf[-1] == L
        assert len(rf) == M + 1 and len(zf) == N + 1 and len(rc) == M and len(zc) == N
        np.testing.assert_allclose(rf, np.arange(M + 1) * (Rr / M), rtol=4e-16, atol=4e-16 * Rr)
        np.testing.assert_allclose(zf, np.arange(N + 1) * (L / N), rtol=4e-16, atol=4e-16 * L)
        # centres are the midpoints of the faces (0.5 (rf[2:end] + rf[1:end-1]) as a LinRange)
        np.testing.assert_allclose(rc, 0.5 * (rf[1:] + rf[:-1]), rtol=0, atol=8e-16 * Rr)
        np.testing.assert_allclose(zc, 0.5 * (zf[1:] + zf[:-1]), rtol=0, atol=8e-16 * L)
        assert rc[0] == 0.5 * (rf[1] + rf[0]) and rc[-1] == 0.5 * (rf[-1] + rf[-2])
        assert np.all(np.diff(rc) > 0) and np.all(np.diff(zc) > 0)


class _FakeFields:
    def __init__(self, grid):
        self.grid = grid


@needs_h5
def test_h5export_layout(tmp_path):
    M, N = 6, 9
    grid = axisym.Grid(1e-3, 3e-3, M, N)
    rng = np.random.default_rng(3)
    cen = {k: rng.normal(size=(N + 2, M + 2)) for k in ("qfixed", "qpart", "q", "u", "wtotal")}
    saved = dict(cen)
    saved["q0"] = ((cen["qfixed"] + cen["qpart"]) - 0.0) * co.elementary_charge
    saved["p"] = rng.integers(0, 70, size=(N + 2, M + 2))
    saved["er"] = rng.normal(size=(N + 3, M + 3))
    saved["ez"] = rng.normal(size=(N + 3, M + 3))
    path = str(tmp_path / "0000.h5")
    axisym.h5export(_FakeFields(grid), path, saved=saved)
    data, meta = h5read(path)
    names = {"grid/R", "grid/L", "grid/M", "grid/N", "grid/rc", "grid/rf", "grid/zc", "grid/zf",
             "qfixed", "qpart", "q0", "q", "u", "er", "ez", "p", "wtotal"}
    assert set(data) == names
    assert meta["grid/R"] == ("<f8", "()") and meta["grid/M"] == ("<i8", "()")
    assert int(data["grid/M"]) == M and int(data["grid/N"]) == N and float(data["grid/R"]) == 1e-3
    np.testing.assert_array_equal(data["grid/rc"], grid.rc)
    np.testing.assert_array_equal(data["grid/zf"], grid.zf)
    # a Julia M x N matrix is an HDF5 (N, M) dataset
    for k in ("qfixed", "qpart", "q0", "q", "u", "wtotal"):
        assert meta[k] == ("<f8", f"({N}, {M})"), k
        np.testing.assert_array_equal(data[k], saved[k][1:N + 1, 1:M + 1], err_msg=k)
    assert meta["p"] == ("<i8", f"({N}, {M})")
    np.testing.assert_array_equal(data["p"], saved["p"][1:N + 1, 1:M + 1])
    assert meta["er"] == ("<f8", f"({N}, {M + 1})")      # er[1:M+1, 1:N]
    np.testing.assert_array_equal(data["er"], saved["er"][1:N + 1, 1:M + 2])
    assert meta["ez"] == ("<f8", f"({N + 1}, {M})")      # ez[1:M, 1:N+1]
    np.testing.assert_array_equal(data["ez"], saved["ez"][1:N + 2, 1:M + 1])


def _counts_brute(w_sorted, starts, ntarget, nmax):
    """resample.jl:16-30 and 63-72 written out: p and wtotal per cell."""
    pc, wc = [], []
    for c in range(len(starts) - 1):
        ws = [float(x) for x in w_sorted[starts[c]:starts[c + 1]]]
        p = len(ws)
        wt = 0.0
        for x in ws:
            wt += x
        if p <= nmax:
            for x in ws:
                if p < ntarget and x > 2.0:
                    p += 1
        pc.append(p)
        wc.append(wt)
    return np.array(pc), np.array(wc)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_resample_counts(seed):
    rng = np.random.default_rng(seed)
    M, N = 7, 5
    G = R.CGrid(1e-3, 1e-3, M, N)
    n = 900
    # clustered positions so that cells span empty, sparse (split) and crowded (merged)
    rr = np.abs(rng.normal(0.0, 3e-4, n)) % 1.1e-3
    th = rng.uniform(0, 2 * np.pi, n)
    x = np.stack([rr * np.cos(th), rr * np.sin(th), rng.normal(5e-4, 2.5e-4, n)], axis=1)
    w = np.where(rng.uniform(size=n) < 0.5, rng.integers(1, 3, n), rng.integers(3, 40, n)).astype(float)
    act = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    P = R.Pop(x, np.zeros((n, 3)), w, np.ones(n), act, cap=4 * n)
    ntarget, nmax = 6, 20
    pc, wc = R.cyl_resample_counts(P, G, ntarget, nmax)
    v = P.view()
    bp, bw = _counts_brute(v["w"], G.cell_start, ntarget, nmax)
    np.testing.assert_array_equal(pc.ravel(), bp[:(M + 2) * (N + 2)])
    np.testing.assert_array_equal(wc.ravel(), bw[:(M + 2) * (N + 2)])
    # closed forms: the counts cover the active particles inside the ghosted cells;
    # a cell of k particles all heavier than 2 and k < ntarget ends at min(2k, ntarget)
    cnt = np.diff(G.cell_start)[:(M + 2) * (N + 2)]
    assert cnt.sum() + (G.cell_start[-1] - G.cell_start[-2]) == act.sum()
    assert np.all(pc.ravel() >= cnt)
    heavy = np.array([np.all(v["w"][G.cell_start[c]:G.cell_start[c + 1]] > 2.0) for c in range(len(cnt))])
    sel = heavy & (cnt > 0) & (cnt < ntarget)
    assert sel.any()
    np.testing.assert_array_equal(pc.ravel()[sel], np.minimum(2 * cnt[sel], ntarget))
    light = np.array([np.all(v["w"][G.cell_start[c]:G.cell_start[c + 1]] <= 2.0) for c in range(len(cnt))])
    np.testing.assert_array_equal(pc.ravel()[light], cnt[light])
    assert np.any(cnt > nmax)
    np.testing.assert_array_equal(pc.ravel()[cnt > nmax], cnt[cnt > nmax])
    np.testing.assert_allclose(wc.sum(), v["w"][:G.cell_start[(M + 2) * (N + 2)]].sum(), rtol=1e-13)
