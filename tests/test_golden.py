"""Committed golden vectors (tests/golden/make_golden.py): the oracle must
reproduce them exactly (a regression pin on the restatement), and they hold
the properties the reference fixes -- a tie is a step of length 0 that is not
pushed, the MLEM update conserves the number of projected LORs."""
import os

import numpy as np

import oracle_ffi as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rows():
    return np.load(os.path.join(GOLDEN, "oracle_rows.npz"))


def _lors(a):
    return np.ascontiguousarray(a, np.float32).view(O.LOR_DTYPE).reshape(-1)


def test_oracle_reproduces_row_fixture():
    r = _rows()
    lors = _lors(r["lors"])
    off, toff = r["offsets"], r["tof_offsets"]
    tof = tuple(float(v) for v in r["tof"])
    tof_pos = {int(k): j for j, k in enumerate(r["tof_lors"])}
    for k, l in enumerate(lors):
        g = int(r["grid"][k])
        full, n = r["grids_full"][g], r["grids_n"][g]
        i, w, st = O.row(full, n, l)
        assert np.array_equal(i, r["idx"][off[k]:off[k + 1]]), k
        assert np.array_equal(w, r["w"][off[k]:off[k + 1]]), k
        assert st == r["steps"][k]
        if k in tof_pos:
            j = tof_pos[k]
            i2, w2, _ = O.row(full, n, l, tof=tof)
            assert np.array_equal(i2, r["tof_idx"][toff[j]:toff[j + 1]])
            assert np.array_equal(w2, r["tof_w"][toff[j]:toff[j + 1]])


def test_row_fixture_properties():
    """Every weight is positive; a walk's steps are its pushed elements plus
    its ties; the diagonal and boundary lines of the fixture produce ties;
    TOF rows are sub-rows of the plain rows (same voxels, in order)."""
    r = _rows()
    off = r["offsets"]
    assert (r["w"] > 0).all()
    lens = np.diff(off)
    hit = r["steps"] > 0
    assert (r["steps"][hit] == lens[hit] + r["ties"][hit]).all()
    assert (r["ties"] > 0).sum() >= 20 and (lens == 0).sum() >= 20
    for j, k in enumerate(r["tof_lors"]):
        plain = list(r["idx"][off[k]:off[k + 1]])
        sub = list(r["tof_idx"][r["tof_offsets"][j]:r["tof_offsets"][j + 1]])
        if sub:
            s0 = plain.index(sub[0])
            assert plain[s0:s0 + len(sub)] == sub


def test_oracle_reproduces_mlem_fixture():
    m = np.load(os.path.join(GOLDEN, "oracle_mlem8.npz"))
    lors = _lors(m["lors"])
    full, n = m["full"], m["n"]
    tof = tuple(float(v) for v in m["tof"])
    for name, t in (("notof", None), ("tof", tof)):
        for f64 in (False, True):
            x = np.ones(512, np.float64 if f64 else np.float32)
            ref = m[f"{name}_{'f64' if f64 else 'f32'}"]
            for it in range(3):
                st = O.mlem_update(full, n, lors, x, tof=t, threads=1, f64=f64)
                assert np.array_equal(x, ref[it]), (name, f64, it)
            if f64 and t is None:
                # sum of x after an update without sensitivity = LORs projected (p > 0)
                assert abs(x.sum() - st["n_lors"] + st["n_missed"] + st["n_skipped"]) < 1e-9 * x.sum()
    # binary32 and binary64 agree closely after three iterations
    assert np.linalg.norm(m["tof_f32"][-1] - m["tof_f64"][-1]) <= 1e-5 * np.linalg.norm(m["tof_f64"][-1])
