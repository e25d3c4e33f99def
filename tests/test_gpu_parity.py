"""Parity of the HIP path against the CPU oracle, through the C ABI.

Bars: voxel-index sequences identical; binary32 geometric weights
bit-identical (the walk is the same IEEE operation sequence); TOF weights
within 8 ulp (device expf vs the host's); images within 1e-4 relative L2
after 10 iterations (the atomic backprojection sums in another order).
"""
import json
import os

import numpy as np
import pytest

import oracle_ffi as O
import petalo_amd as P
from helpers import max_rel, proptest_lors, random_lors, rel_l2, rel_l2_finite, ulps

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def oracle_rows(full, n, lors, tof=None):
    return [O.row(full, n, l, tof=tof) for l in lors]


def test_rows_hand_picked(gpu):
    kats = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))["hand_picked"]["cases"]
    for c in kats:
        _say(f"hand_picked {c['p1']} -> {c['p2']} n={c['n']}")
        fov = P.FOV([c["size"][0], c["size"][1], 1.0], [c["n"][0], c["n"][1], 1])
        lor = P.make_lors([c["p1"][0], c["p1"][1], 0.0], [c["p2"][0], c["p2"][1], 0.0])
        idx, w, cnt, steps = P.system_matrix_rows(fov, lor)
        m = int(cnt[0])
        vox = [[int(i) % c["n"][0], (int(i) // c["n"][0]) % c["n"][1]] for i in idx[0, :m]]
        assert vox == c["voxels"]
        total = np.float32(0)
        for x in w[0, :m]:
            total = np.float32(total + x)
        assert ulps(total, np.float32(c["length"])) <= 1


@pytest.mark.parametrize("tof", [None, (42.47, 3.0)])
def test_rows_random_match_oracle(gpu, tof):
    rng = np.random.default_rng(7)
    cases = proptest_lors(rng, 300)
    n_elems = n_ties = 0
    worst = 0
    for k, (full, nv, p1, p2) in enumerate(cases):
        lor = P.make_lors(p1, p2)
        if tof:
            lor["dt"] = np.float32(rng.normal(0, 200))
        fov = P.FOV(full, nv)
        idx, w, cnt, steps = P.system_matrix_rows(fov, lor, tof=P.Tof(*tof) if tof else None)
        oi, ow, osteps = O.row(full, nv, lor[0], tof=tof)
        m = int(cnt[0])
        assert m == len(oi), f"case {k}: element count {m} vs oracle {len(oi)}"
        assert int(steps[0]) == osteps, f"case {k}: steps"
        assert np.array_equal(idx[0, :m].astype(np.int64), oi), f"case {k}: voxel sequence"
        if tof is None:
            assert np.array_equal(w[0, :m], ow), f"case {k}: weights not bit-identical"
        else:
            u = int(ulps(w[0, :m], ow).max(initial=0))
            worst = max(worst, u)
            assert u <= 8, f"case {k}: TOF weights {u} ulp"
        n_elems += m
        n_ties += int(steps[0]) - m if tof is None else 0
    _say(f"rows: {n_elems} elements, {n_ties} zero-length tie steps, identical indices; "
         f"worst TOF weight difference {worst} ulp")


def test_rows_phantom_bitexact(gpu):
    """Scanner-geometry LORs on the 60^3 FOV: every row identical to the oracle's."""
    lors = P.synth_phantom_lors(11, 2000, 0.0)
    full, n = [180.0] * 3, [60] * 3
    fov = P.FOV(full, n)
    idx, w, cnt, steps = P.system_matrix_rows(fov, lors)
    mism = 0
    for i, l in enumerate(lors):
        oi, ow, os_ = O.row(full, n, l)
        m = int(cnt[i])
        ok = m == len(oi) and np.array_equal(idx[i, :m], oi) and np.array_equal(w[i, :m], ow) and steps[i] == os_
        mism += 0 if ok else 1
    assert mism == 0, f"{mism} of {len(lors)} rows differ"


def _run_gpu(fov, lors, iters, tof=None, sens=None, subsets=1, **kw):
    it = P.mlem(fov, lors, tof=tof, sensitivity=sens, n_subsets=subsets, **kw)
    imgs = []
    for _ in range(iters * subsets):
        img, osem = next(it)
        imgs.append(img.data.copy())
    it.close()
    return imgs


def _run_oracle(full, n, lors, iters, tof=None, sens=None, subsets=1, threads=16, f64=False):
    nv = int(np.prod(n))
    x = np.ones(nv, np.float64 if f64 else np.float32)
    s = None if sens is None else np.asarray(sens, np.float64 if f64 else np.float32)
    size = len(lors) // subsets
    out = []
    for it in range(iters):
        for sub in range(subsets):
            O.mlem_update(full, n, lors[sub * size:(sub + 1) * size], x, sens=s, tof=tof, threads=threads,
                          f64=f64)
            out.append(x.copy())
    return out


SCHEMES = [P.SCHEME_ATOMIC, P.SCHEME_TILE]


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("tof", [None, (42.47, 3.0)])
def test_mlem_small(gpu, tof, scheme):
    full, n = [48.0, 48.0, 40.0], [12, 12, 10]
    rng = np.random.default_rng(3)
    lors = random_lors(rng, 20000, dt_sigma_ps=84.93 if tof else 0.0)
    fov = P.FOV(full, n)
    g = _run_gpu(fov, lors, 5, tof=P.Tof(*tof) if tof else None, scheme=scheme)
    o = _run_oracle(full, n, lors, 5, tof=tof)
    for k in range(5):
        e = rel_l2(g[k], o[k])
        _say(f"small tof={tof} iter {k + 1}: relL2 {e:.2e}")
        assert e <= 1e-5


@pytest.mark.parametrize("scheme", SCHEMES)
def test_mlem_config1_full(gpu, scheme):
    """60^3 x 3 mm, 1e6 phantom LORs, non-TOF, 10 iterations: <= 1e-4 relative L2."""
    full, n = [180.0] * 3, [60] * 3
    lors = P.synth_phantom_lors(1, 1_000_000, 0.0)
    fov = P.FOV(full, n)
    _say(f"config1 scheme {scheme}: GPU 10 iterations")
    g = _run_gpu(fov, lors, 10, scheme=scheme)
    _say("config1: oracle 10 iterations")
    o = _run_oracle(full, n, lors, 10)
    o64 = _run_oracle(full, n, lors, 10, f64=True)
    e, m, e64 = rel_l2(g[-1], o[-1]), max_rel(g[-1], o[-1]), rel_l2(g[-1], o64[-1])
    _say(f"config1: relL2 vs f32 oracle {e:.2e}, max {m:.2e}; vs f64 oracle {e64:.2e}")
    assert e <= 1e-4


@pytest.mark.parametrize("scheme", SCHEMES)
def test_mlem_tof_120(gpu, scheme):
    """120^3 x 1.5 mm TOF (200 ps FWHM data, sigma 42.47 ps, cutoff 3) with 1e6 LORs, 10 iterations."""
    full, n = [180.0] * 3, [120] * 3
    lors = P.synth_phantom_lors(2, 1_000_000, 200.0)
    tof = (42.47, 3.0)
    fov = P.FOV(full, n)
    _say(f"tof120 scheme {scheme}: GPU")
    g = _run_gpu(fov, lors, 10, tof=P.Tof(*tof), scheme=scheme)
    _say("tof120: oracle")
    o = _run_oracle(full, n, lors, 10, tof=tof)
    e = rel_l2(g[-1], o[-1])
    _say(f"tof120: relL2 {e:.2e}, max {max_rel(g[-1], o[-1]):.2e}")
    assert e <= 1e-4


@pytest.mark.parametrize("scheme", SCHEMES)
def test_osem_subsets(gpu, scheme):
    full, n = [180.0] * 3, [30] * 3
    lors = P.synth_phantom_lors(5, 100_003, 0.0)  # remainder LORs are dropped
    fov = P.FOV(full, n)
    it = P.mlem(fov, lors, n_subsets=4, scheme=scheme)
    ids = []
    g = []
    for _ in range(8):
        img, osem = next(it)
        ids.append((osem.iteration, osem.subset))
        g.append(img.data.copy())
    assert ids == [(1, 1), (1, 2), (1, 3), (1, 4), (2, 1), (2, 2), (2, 3), (2, 4)]
    o = _run_oracle(full, n, lors, 2, subsets=4)
    for k in range(8):
        assert rel_l2(g[k], o[k]) <= 1e-5


def test_sensitivity_matches_oracle(gpu):
    att = P.FOV([180.0] * 3, [30] * 3)
    rho = P.synth_density(att)
    mu = P.density_to_attenuation(rho, 0.096)
    n, seed, L, D = 200_000, 99, 300.0, 700.0
    gen = P.sensitivity_lors(att, n, seed, L, D)
    img, st = P.make_sensitivity_image(mu, n, L, D, seed=seed, return_stats=True)
    img2 = P.make_sensitivity_image(mu, n, lors=gen)
    img3 = P.make_sensitivity_image(mu, n, L, D, seed=seed)  # no work counters: the tile jump bins
    ref, ost = O.sensitivity(att.full_size, att.n, mu.data, gen, threads=16)
    ref64, _ = O.sensitivity(att.full_size, att.n, mu.data, gen, threads=16, f64=True)
    assert st["n_steps"] == ost["n_steps"] and st["n_nz"] == ost["n_nz"]
    e1, e2, e3, e64 = rel_l2(img.data, ref), rel_l2(img2.data, ref), rel_l2(img3.data, ref), rel_l2(img.data, ref64)
    _say(f"sensitivity: relL2 gen {e1:.2e} list {e2:.2e} gen without counters {e3:.2e} vs f64 {e64:.2e}")
    assert e1 <= 1e-5 and e2 <= 1e-5 and e3 <= 1e-5 and e64 <= 1e-5


def test_sensitivity_two_fovs(gpu):
    att = P.FOV([180.0] * 3, [30] * 3)
    out = P.FOV([150.0, 150.0, 120.0], [25, 25, 20])
    mu = P.Image(att, np.full(att.nvox, 0.0096, np.float32))
    n, seed, L, D = 50_000, 5, 300.0, 700.0
    gen = P.sensitivity_lors(att, n, seed, L, D)
    img = P.make_sensitivity_image(mu, n, L, D, output_fov=out, seed=seed)
    ref, _ = O.sensitivity(att.full_size, att.n, mu.data, gen, out_full=out.full_size, out_n=out.n, threads=16)
    assert rel_l2(img.data, ref) <= 1e-5


def test_sensitivity_no_attenuation(gpu):
    """mu = 0 everywhere (the tiled path skips the forward half: p = 0
    exactly): image and counters against the binary32 oracle."""
    att = P.FOV([180.0] * 3, [60] * 3)
    mu = P.Image(att, np.zeros(att.nvox, np.float32))
    n, seed, L, D = 300_000, 3, 300.0, 700.0
    gen = P.sensitivity_lors(att, n, seed, L, D)
    img, st = P.make_sensitivity_image(mu, n, L, D, seed=seed, return_stats=True)
    img2 = P.make_sensitivity_image(mu, n, L, D, seed=seed)
    ref, ost = O.sensitivity(att.full_size, att.n, mu.data, gen, threads=16)
    assert st["n_steps"] == ost["n_steps"] and st["n_nz"] == ost["n_nz"] and st["n_lors"] == n
    e, e2 = rel_l2(img.data, ref), rel_l2(img2.data, ref)
    _say(f"sensitivity mu=0: relL2 {e:.2e}, without work counters {e2:.2e}")
    assert e <= 1e-5 and e2 <= 1e-5


def test_large_lor_upload(gpu):
    """LOR arrays of 32 MB and more reach the device through the pinned
    upload ring (16 MB chunks filled by host threads): 1.3e6 given LORs
    (41.6 MB, three chunks, the last partial) give the image of the same
    LORs generated on the device; and a context of 1.1e6 LORs that miss the
    FOV followed by 2e5 phantom LORs (the latter all in the last chunks)
    gives the update of a context of the 2e5 alone (6.4 MB, a plain copy) --
    equal to 1e-9."""
    att = P.FOV([180.0] * 3, [30] * 3)
    mu = P.density_to_attenuation(P.synth_density(att), 0.096)
    n, seed, L, D = 1_300_000, 23, 300.0, 700.0
    gen = P.sensitivity_lors(att, n, seed, L, D)
    a = P.make_sensitivity_image(mu, n, L, D, seed=seed).data
    b = P.make_sensitivity_image(mu, n, lors=gen).data
    e = rel_l2(b, a)
    fov = P.FOV([180.0] * 3, [40] * 3)
    real = P.synth_phantom_lors(5, 200_000, 200.0)
    miss = np.zeros(1_100_000, P.LOR_DTYPE)
    miss["p1"] = (1000.0, 1000.0, 1000.0)
    miss["p2"] = (1000.0, 1001.0, 1000.0)
    miss["ac"] = 1.0
    c1 = P.Mlem(fov, real, tof=P.Tof(42.47, 3.0))
    c2 = P.Mlem(fov, np.concatenate([miss, real]), tof=P.Tof(42.47, 3.0))
    i1, _ = c1.step()
    i2, _ = c2.step()
    e2 = rel_l2(i2.data, i1.data)
    c1.close()
    c2.close()
    _say(f"large uploads: sensitivity given vs generated {e:.1e}, contexts {e2:.1e}")
    assert e <= 1e-9 and e2 <= 1e-9


def test_sensitivity_unit_weights_fixed_point_bound(gpu):
    """Without attenuation the tiles sum the weights in 32-bit fixed point,
    each step rounded to at most 2^-17 of a voxel diagonal.  Against the same
    walks' weights summed in binary64 (the rows, bit-identical with the
    oracle's), every voxel lies within (its step count) x diagonal x 2^-17 / N
    plus the binary32 output's rounding; with 18 steps per voxel on
    average here, the image is within 3e-6 of those sums (the rounding
    averages out as the steps per voxel grow)."""
    att = P.FOV([180.0] * 3, [60] * 3)
    mu = P.Image(att, np.zeros(att.nvox, np.float32))
    n = 40_000
    lors = P.sensitivity_lors(att, n, 11, 300.0, 700.0)
    img = P.make_sensitivity_image(mu, n, lors=lors).data.astype(np.float64)
    idx, w, count, _ = P.system_matrix_rows(att, lors)
    assert int(count.max()) <= idx.shape[1]
    keep = np.arange(idx.shape[1])[None, :] < count[:, None]
    vi, vw = idx[keep], w[keep].astype(np.float64)
    exact = np.bincount(vi, weights=vw, minlength=att.nvox) / n
    steps = np.bincount(vi, minlength=att.nvox)
    diag = float(np.sqrt(np.sum((np.asarray(att.full_size) / np.asarray(att.n)) ** 2)))
    bound = steps * diag * 2.0 ** -17 / n + np.abs(exact) * 2.0 ** -23 + 1e-30
    worst = float(np.max(np.abs(img - exact) / bound))
    _say(f"unit weights: relL2 to the binary64 sums {rel_l2(img, exact):.2e}, worst voxel at {worst:.2f} of its bound")
    assert worst <= 1.0
    assert rel_l2(img, exact) <= 3e-6


def test_sensitivity_irregular_lors(gpu):
    """Given LORs mixing ordinary ones with face-grazing, axis-parallel,
    degenerate and missing ones: image and counters (skips, misses) match the oracle."""
    att = P.FOV([30.0] * 3, [6] * 3)
    mu = P.Image(att, np.random.default_rng(4).uniform(0.0, 0.02, att.nvox).astype(np.float32))
    rng = np.random.default_rng(12)
    normal = random_lors(rng, 5000, radius=60.0, zhalf=20.0)
    odd = P.make_lors(
        [[-100, 5, 0], [5, 5, 5], [-100, 0, 0], [15.0, -100, 1.0], [-100, -100, -100], [-100, 2, 3],
         [15.0, 0.0, 0.0], [-15.0, -100.0, -15.0], [0.0, 0.0, -100.0], [-100, 50, 0]],
        [[100, 5, 0], [5, 5, 5], [100, 0, 0], [15.0, 100, 1.0], [100, 100, 100], [100, -2, -3],
         [100.0, 0.0, 0.0], [-15.0, 100.0, -15.0], [0.0, 0.0, 100.0], [100, 50, 0]])
    lors = np.concatenate([normal[:2000], odd, normal[2000:]])
    img, st = P.make_sensitivity_image(mu, len(lors), lors=lors, return_stats=True)
    ref, ost = O.sensitivity(att.full_size, att.n, mu.data, lors, threads=16)
    _say(f"irregular sensitivity: gpu {st} oracle {ost}")
    for k in ("n_steps", "n_nz", "n_skipped", "n_missed", "n_runaway"):
        assert st[k] == ost[k], k
    assert rel_l2(img.data, ref) <= 1e-5


def test_sensitivity_many_unbounded_walks(gpu):
    """500 000 given LORs in planes of constant z (walks the binning cannot
    bound, projected one lane per LOR; about a fifth of them cross the FOV)
    among 100 000 ordinary ones: more than the 2^16 that the buffer of such
    walks holds at first, so it grows and they are taken again.  Image and counters agree with the one-lane
    path (selected by an explicit output FOV)."""
    att = P.FOV([180.0] * 3, [30] * 3)
    mu = P.density_to_attenuation(P.synth_density(att), 0.096)
    rng = np.random.default_rng(19)
    flat = random_lors(rng, 500_000, radius=350.0, zhalf=80.0)
    flat["p2"][:, 2] = flat["p1"][:, 2]
    lors = np.concatenate([random_lors(rng, 100_000, radius=350.0, zhalf=80.0), flat])
    for m in (mu, P.Image(att, np.zeros(att.nvox, np.float32))):
        a, sa = P.make_sensitivity_image(m, len(lors), lors=lors, return_stats=True)
        b, sb = P.make_sensitivity_image(m, len(lors), lors=lors, output_fov=att, return_stats=True)
        e = rel_l2(a.data, b.data)
        _say(f"unbounded walks: tiled vs one-lane relL2 {e:.2e}; steps {sa['n_steps']} {sb['n_steps']}, "
             f"missed {sa['n_missed']} of {len(lors)}")
        for k in ("n_steps", "n_nz", "n_skipped", "n_missed"):
            assert sa[k] == sb[k], k
        assert e <= 1e-6


def test_sensitivity_batches_agree_with_one_lane_path(gpu):
    """More LORs than one batch of the tiled path: it agrees with the
    one-lane kernel (selected by passing the output FOV explicitly)."""
    att = P.FOV([180.0] * 3, [30] * 3)
    rho = P.synth_density(att)
    mu = P.density_to_attenuation(rho, 0.096)
    n, seed, L, D = (1 << 25) + 4099, 21, 300.0, 700.0
    a, sa = P.make_sensitivity_image(mu, n, L, D, seed=seed, return_stats=True)
    b, sb = P.make_sensitivity_image(mu, n, L, D, output_fov=att, seed=seed, return_stats=True)
    e = rel_l2(a.data, b.data)
    _say(f"sensitivity {n} LORs: tiled vs one-lane relL2 {e:.2e}; steps {sa['n_steps']} {sb['n_steps']}")
    assert sa["n_steps"] == sb["n_steps"] and sa["n_nz"] == sb["n_nz"]
    assert e <= 1e-6


def test_sensitivity_kept_workspace(gpu):
    """The sensitivity calls keep their device buffers between calls: a call
    after a larger one (buffers kept, bigger than needed), after a smaller
    one on another FOV (buffers grown), and after release_workspaces (buffers
    made anew) gives the image of a first call, and each matches the oracle."""
    att = P.FOV([180.0] * 3, [30] * 3)
    small = P.FOV([120.0, 120.0, 90.0], [20, 20, 15])
    mu = P.density_to_attenuation(P.synth_density(att), 0.096)
    mu0 = P.Image(small, np.zeros(small.nvox, np.float32))
    n, seed, L, D = 150_000, 8, 300.0, 700.0
    P.release_workspaces()
    first = P.make_sensitivity_image(mu, n, L, D, seed=seed).data.copy()
    imgs = []
    for k in range(3):
        P.make_sensitivity_image(mu0, [400_000, 20_000, 20_000][k], L, D, seed=k)
        if k == 2:
            P.release_workspaces()
        imgs.append(P.make_sensitivity_image(mu, n, L, D, seed=seed).data.copy())
    ref, _ = O.sensitivity(att.full_size, att.n, mu.data, P.sensitivity_lors(att, n, seed, L, D), threads=16)
    errs = [rel_l2(x, first) for x in imgs]
    _say(f"kept workspace: relL2 to the first call {['%.1e' % e for e in errs]}, first vs oracle {rel_l2(first, ref):.1e}")
    assert max(errs) <= 1e-7
    assert rel_l2(first, ref) <= 1e-5


@pytest.mark.parametrize("attenuated", [False, True])
def test_sensitivity_many_tiles(gpu, attenuated):
    """A 480x480x240 image has 12x12x20 = 2880 sensitivity tiles (40x40x12),
    more than the binning into per-(tile, workgroup) segments keeps cursors
    for: its runs are binned LOR-major and sorted by tile instead.  Image and counters match the oracle,
    with and without attenuation."""
    att = P.FOV([480.0, 480.0, 240.0], [480, 480, 240])
    mu = P.Image(att, np.full(att.nvox, 0.0005 if attenuated else 0.0, np.float32))
    n, seed, L, D = 20_000, 17, 300.0, 700.0
    gen = P.sensitivity_lors(att, n, seed, L, D)
    img, st = P.make_sensitivity_image(mu, n, L, D, seed=seed, return_stats=True)
    ref, ost = O.sensitivity(att.full_size, att.n, mu.data, gen, threads=16)
    assert st["n_steps"] == ost["n_steps"] and st["n_nz"] == ost["n_nz"]
    e = rel_l2(img.data, ref)
    _say(f"sensitivity 480x480x240 (attenuated {attenuated}): relL2 {e:.2e}, {st['n_steps']} steps")
    assert e <= 1e-5


@pytest.mark.parametrize("adjust", [P.ADJUST_CENTRE, P.ADJUST_RANDOM_Z_PHI, P.ADJUST_RANDOM])
def test_discrete_sensitivity_matches_oracle(gpu, adjust):
    """All pairs of detector elements (make_sensitivity_image/discrete.rs):
    the device generates the element pairs, adjusts both ends and backprojects;
    the oracle does the same on the host.  Normalisation n(n-1)/2 of the
    element count n, not of the pair count."""
    att = P.FOV([120.0] * 3, [24] * 3)
    mu = P.density_to_attenuation(P.synth_density(att), 0.096)
    geom = dict(r_min=100.0, dr=10.0, dz=10.0, da=10.0)
    L, seed = 100.0, 5
    img, st = P.make_sensitivity_image_discrete(mu, L, adjust=adjust, seed=seed, return_stats=True, **geom)
    img2 = P.make_sensitivity_image_discrete(mu, L, adjust=adjust, seed=seed, **geom)
    lors, n = O.discrete_lors(geom["r_min"], geom["dr"], geom["dz"], geom["da"], L, adjust, seed=seed)
    ref, ost = O.sensitivity(att.full_size, att.n, mu.data, lors, threads=16)
    ref = ref.astype(np.float64) * len(lors) / (n * (n - 1) / 2)
    assert st["n_lors"] == len(lors) == n * (n + 1) // 2
    assert rel_l2(img2.data, img.data) <= 1e-6  # without work counters: the same runs
    e = rel_l2(img.data, ref)
    ds = abs(st["n_steps"] - ost["n_steps"]) / ost["n_steps"]
    _say(f"discrete sensitivity adjust {adjust}: {n} elements, {len(lors)} pairs, relL2 {e:.2e}, steps {ds:.1e}")
    # element centres and adjusted ends use the device's cosf/sinf: ends may
    # move by an ulp, so a LOR may gain or lose a corner voxel
    assert ds <= 1e-4 and e <= 1e-4


def test_discrete_sensitivity_two_fovs_and_batches(gpu):
    """More pairs than one batch of the tiled path (2^25): the tiled image
    agrees with the one-lane path (selected by an explicit output FOV)."""
    att = P.FOV([120.0] * 3, [24] * 3)
    mu = P.density_to_attenuation(P.synth_density(att), 0.096)
    geom = dict(r_min=100.0, dr=10.0, dz=2.0, da=4.0)
    a, sa = P.make_sensitivity_image_discrete(mu, 100.0, return_stats=True, seed=3, **geom)
    b, sb = P.make_sensitivity_image_discrete(mu, 100.0, output_fov=att, return_stats=True, seed=3, **geom)
    e = rel_l2(a.data, b.data)
    _say(f"discrete sensitivity {sa['n_lors']} pairs: tiled vs one-lane relL2 {e:.2e}")
    assert sa["n_lors"] > (1 << 25) and sa["n_lors"] == sb["n_lors"]
    assert sa["n_steps"] == sb["n_steps"] and sa["n_nz"] == sb["n_nz"]
    assert e <= 1e-6


def test_walk_stats_match_oracle(gpu):
    full, n = [180.0] * 3, [120] * 3
    lors = P.synth_phantom_lors(8, 200_000, 200.0)
    tof = (42.47, 3.0)
    ctx = P.Mlem(P.FOV(full, n), lors, tof=P.Tof(*tof))
    st = ctx.walk_stats()
    ost = O.walk_stats(full, n, lors, tof=tof, threads=16)
    _say(f"walk stats gpu {st} oracle {ost}")
    assert st["n_steps"] == ost["n_steps"] and st["n_missed"] == ost["n_missed"]
    assert abs(st["n_nz"] - ost["n_nz"]) <= 1e-4 * ost["n_nz"]  # TOF cut edge under expf ulps


@pytest.mark.parametrize("scheme", SCHEMES)
def test_edge_cases(gpu, scheme):
    """Empty rows, degenerate LORs, out-of-range skips, zero projections,
    s <= 0 voxels, more subsets than LORs (projector.rs:103-118, mlem.rs:45-52,90-95)."""
    full, n = [30.0, 30.0, 30.0], [3, 3, 3]
    fov = P.FOV(full, n)
    lors = P.make_lors(
        [[-100, 50, 0], [5, 5, 5], [-100, 0, 0], [15.0, -100, 1.0], [-100, -100, -100], [-100, 2, 3],
         [15.0, 0.0, 0.0]],
        [[100, 50, 0], [5, 5, 5], [100, 0, 0], [15.0, 100, 1.0], [100, 100, 100], [100, -2, -3],
         [100.0, 0.0, 0.0]])
    idx, w, cnt, steps = P.system_matrix_rows(fov, lors)
    for i, l in enumerate(lors):
        oi, ow, os_ = O.row(full, n, l)
        m = int(cnt[i])
        assert m == len(oi) and np.array_equal(idx[i, :m], oi) and np.array_equal(w[i, :m], ow)
        assert (steps[i] == os_) or (os_ < 0 and steps[i] == 0xffffffff)
    # MLEM with a sensitivity image containing zeros and negatives
    sens = np.ones(fov.nvox, np.float32)
    sens[0], sens[5] = 0.0, -1.0
    g = _run_gpu(fov, lors, 2, sens=sens, scheme=scheme)
    o = _run_oracle(full, n, lors, 2, sens=sens)
    for a, b in zip(g, o):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin)
        assert np.allclose(a[fin], b[fin], rtol=1e-6, atol=0)
    # more subsets than LORs: empty subsets zero the image (x * 0 * s)
    g = _run_gpu(fov, lors, 1, subsets=10, scheme=scheme)
    assert all(np.all(x == 0) for x in g)


@pytest.mark.parametrize("tof", [None, (31.85, 3.0)])
def test_tile_scheme_odd_shapes(gpu, tof):
    """Images whose dimensions are not multiples of the tile shape, with
    oblique and axis-parallel LORs: tiled and atomic schemes agree with the oracle."""
    full, n = [100.0, 130.0, 170.0], [37, 51, 29]
    rng = np.random.default_rng(17)
    lors = random_lors(rng, 30000, radius=200.0, zhalf=120.0, dt_sigma_ps=63.7 if tof else 0.0)
    lors["p2"][::11, 0] = lors["p1"][::11, 0]
    lors["p2"][::7, 2] = lors["p1"][::7, 2]
    fov = P.FOV(full, n)
    t = P.Tof(*tof) if tof else None
    o = _run_oracle(full, n, lors, 3, tof=tof)
    for scheme in SCHEMES:
        g = _run_gpu(fov, lors, 3, tof=t, scheme=scheme)
        for k in range(3):
            e = rel_l2(g[k], o[k])
            _say(f"odd shapes tof={tof} scheme={scheme} iter {k + 1}: relL2 {e:.2e}")
            assert e <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [(8, 16, 4), (5, 7, 3), (64, 64, 4), (40, 40, 6)])
def test_tile_shapes(gpu, tile):
    """Explicit tile shapes (even and odd row pitch, tiles wider than the
    image, tiles that cut it unevenly) give the oracle's TOF images."""
    full, n = [100.0, 130.0, 170.0], [37, 51, 29]
    tof = (42.47, 3.0)
    rng = np.random.default_rng(23)
    lors = random_lors(rng, 20000, radius=200.0, zhalf=120.0, dt_sigma_ps=63.7)
    lors["p2"][::13, 1] = lors["p1"][::13, 1]
    fov = P.FOV(full, n)
    o = _run_oracle(full, n, lors, 2, tof=tof)
    g = _run_gpu(fov, lors, 2, tof=P.Tof(*tof), scheme=P.SCHEME_TILE, tile=tile)
    for k in range(2):
        e = rel_l2(g[k], o[k])
        _say(f"tile {tile} iter {k + 1}: relL2 {e:.2e}")
        assert e <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [(4, 4, 4), (15, 15, 15), (30, 30, 20)])
def test_reduction_window_forms(gpu, tile):
    """The reduction's partial windows (4096 LORs each) at about 22, 6 and 3
    runs per LOR (60^3 phantom LORs): over 2^15 partials per window (full
    positions, gathered from memory), between 20480 and 2^15 (16-bit window
    offsets, gathered from memory) and under 20480 (16-bit offsets, staged in
    LDS) all give the oracle's images, with and without TOF."""
    full, n = [180.0] * 3, [60] * 3
    fov = P.FOV(full, n)
    lors = P.synth_phantom_lors(5, 30000, 200.0)
    for tof in (None, (42.47, 3.0)):
        o = _run_oracle(full, n, lors, 2, tof=tof)
        g = _run_gpu(fov, lors, 2, tof=P.Tof(*tof) if tof else None, scheme=P.SCHEME_TILE, tile=tile)
        for k in range(2):
            e = rel_l2(g[k], o[k])
            _say(f"reduction windows tile {tile} tof={tof} iter {k + 1}: relL2 {e:.2e}")
            assert e <= 1e-5


def test_rows_match_golden_fixture(gpu):
    """The committed 1000-LOR row fixture (tests/golden/oracle_rows.npz): voxel
    sequences, step counts and ties identical, weights bit-identical without
    TOF and within 8 ulp with TOF (200 LORs)."""
    r = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_rows.npz"))
    lors = np.ascontiguousarray(r["lors"], np.float32).view(P.LOR_DTYPE).reshape(-1)
    off, toff = r["offsets"], r["tof_offsets"]
    tof = P.Tof(*(float(v) for v in r["tof"]))
    matched = ties = 0
    for g in range(len(r["grids_n"])):
        fov = P.FOV(list(r["grids_full"][g]), list(r["grids_n"][g]))
        sel = np.flatnonzero(r["grid"] == g)
        idx, w, cnt, steps = P.system_matrix_rows(fov, lors[sel])
        for j, k in enumerate(sel):
            n = int(cnt[j])
            assert np.array_equal(idx[j, :n], r["idx"][off[k]:off[k + 1]]), k
            assert np.array_equal(w[j, :n], r["w"][off[k]:off[k + 1]]), k
            assert int(steps[j]) == r["steps"][k]
            ties += int(steps[j]) - n if steps[j] > 0 else 0
            matched += 1
        tsel = np.array([k for k in r["tof_lors"] if r["grid"][k] == g], np.int64)
        if len(tsel):
            tidx, tw, tcnt, _ = P.system_matrix_rows(fov, lors[tsel], tof=tof)
            pos = {int(k): j for j, k in enumerate(r["tof_lors"])}
            for j, k in enumerate(tsel):
                q = pos[int(k)]
                n = int(tcnt[j])
                assert np.array_equal(tidx[j, :n], r["tof_idx"][toff[q]:toff[q + 1]])
                assert ulps(tw[j, :n], r["tof_w"][toff[q]:toff[q + 1]]).max(initial=0) <= 8
    assert ties == int(r["ties"].sum())
    _say(f"golden rows: {matched} of 1000 voxel sequences identical, {ties} ties")


@pytest.mark.parametrize("scheme", [P.SCHEME_TILE, P.SCHEME_ATOMIC])
def test_mlem_matches_golden_fixture(gpu, scheme):
    """Three MLEM iterations on the committed 8^3 fixture, without and with
    TOF: within 1e-5 of the binary32 oracle images, and reported against the
    binary64 ones."""
    m = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_mlem8.npz"))
    lors = np.ascontiguousarray(m["lors"], np.float32).view(P.LOR_DTYPE).reshape(-1)
    fov = P.FOV(list(m["full"]), list(m["n"]))
    for name, tof in (("notof", None), ("tof", P.Tof(*(float(v) for v in m["tof"])))):
        it = P.Mlem(fov, lors, tof=tof, scheme=scheme)
        for k in range(3):
            g, _ = it.step()
            e32 = rel_l2(g.data, m[f"{name}_f32"][k])
            e64 = rel_l2(g.data, m[f"{name}_f64"][k])
            _say(f"golden MLEM {name} scheme {scheme} iteration {k + 1}: relL2 vs f32 {e32:.2e}, vs f64 {e64:.2e}")
            assert e32 <= 1e-5
        it.close()


def test_autotuned_context_matches_oracle(gpu):
    """scheme AUTO with autotune: the chosen configuration (atomic, or a tile
    shape) reports its shape and reconstructs as the oracle does."""
    full, n = [90.0, 90.0, 60.0], [45, 45, 30]
    rng = np.random.default_rng(17)
    lors = random_lors(rng, 60000, radius=150.0, zhalf=80.0, dt_sigma_ps=84.93)
    fov = P.FOV(full, n)
    tof = P.Tof(42.47, 3.0)
    m = P.Mlem(fov, lors, tof=tof, n_subsets=2, autotune=True)
    scheme = m.info()["scheme"]
    shape = m.tile_shape()
    steps, n_items = m.work_items()
    _say(f"autotune chose scheme {scheme}, tile {shape}, {steps} steps per work item ({n_items} items)")
    assert steps == 0 or (1 << 14) <= steps <= (1 << 20)
    assert (scheme == P.SCHEME_ATOMIC and shape == (0, 0, 0)) or (scheme == P.SCHEME_TILE and min(shape) > 0)
    o = _run_oracle(full, n, lors, 4, tof=(42.47, 3.0), subsets=2)
    for k in range(4):
        g, _ = m.step()
        assert rel_l2_finite(g.data, o[k]) <= 1e-5
    m.close()
    d = P.Mlem(fov, lors, tof=tof, n_subsets=2, tile=(16, 16, 0))
    assert d.tile_shape()[:2] == (16, 16) and d.tile_shape()[2] > 0
    d.close()


def test_work_item_granularity(gpu):
    """The tiled scheme's work items re-cut to 2^14 and 2^19 replayed steps
    each (and the automatic rule) give the same images: the cut changes which
    workgroup replays a run, not what it adds (binary64 sums in another order,
    rounded to binary32)."""
    full, n = [120.0, 120.0, 90.0], [60, 60, 45]
    fov = P.FOV(full, n)
    lors = P.synth_phantom_lors(23, 400_000, 200.0)
    tof = P.Tof(42.47, 3.0)
    imgs, counts = {}, {}
    for steps in (0, 1 << 14, 1 << 19):
        m = P.Mlem(fov, lors, tof=tof, n_subsets=4, item_steps=steps)
        got, n_items = m.work_items()
        assert got == steps
        counts[steps] = n_items
        for _ in range(3):
            img, _ = m.step()
        imgs[steps] = img.data
        m.close()
    _say(f"work items per 4 subsets: {counts}")
    assert counts[1 << 14] > counts[1 << 19]
    # without a sensitivity image, voxels no LOR of a subset crosses become
    # 0 and later 0 * inf = NaN, as in the reference: NaN-aware comparisons
    for steps in (1 << 14, 1 << 19):
        assert rel_l2_finite(imgs[steps], imgs[0]) <= 1e-6
    o = _run_oracle(full, n, lors, 1, tof=(42.47, 3.0), subsets=4)
    live = int(np.count_nonzero(np.isfinite(o[2]) & (o[2] != 0)))
    _say(f"after 3 subset updates: {live} finite non-zero voxels of {fov.nvox}")
    assert live >= fov.nvox // 4
    assert rel_l2_finite(imgs[1 << 14], o[2]) <= 1e-5


def test_graph_replays_match_direct_launches(gpu):
    """Subset updates replayed as captured graphs (the default) give the
    images of direct launches, replay after replay, with host reads between."""
    full, n = [60.0, 60.0, 48.0], [20, 20, 16]
    rng = np.random.default_rng(31)
    lors = random_lors(rng, 40000, radius=120.0, zhalf=60.0, dt_sigma_ps=84.93)
    fov = P.FOV(full, n)
    tof = P.Tof(42.47, 3.0)
    a = P.Mlem(fov, lors, tof=tof, n_subsets=3)
    b = P.Mlem(fov, lors, tof=tof, n_subsets=3)
    b.set_graphs(False)
    o = _run_oracle(full, n, lors, 3, tof=(42.47, 3.0), subsets=3)
    for k in range(9):
        ga, ia = a.step()
        gb, ib = b.step()
        assert (ia.iteration, ia.subset) == (ib.iteration, ib.subset)
        e, eo = rel_l2_finite(ga.data, gb.data), rel_l2_finite(ga.data, o[k])
        _say(f"graph vs direct update {k + 1}: relL2 {e:.2e}; vs oracle {eo:.2e}; "
             f"{int(np.sum(~np.isfinite(ga.data)))} non-finite voxels")
        assert e <= 1e-6 and eo <= 1e-5
    a.close()
    b.close()


def test_timed_updates_back_to_back(gpu):
    """Timed updates drained after a loop: one timing per update, oldest
    first, the last equal to last_kernel_timing; images unchanged by timing;
    more updates than the ring keeps yields the most recent 256."""
    full, n = [60.0, 60.0, 48.0], [20, 20, 16]
    rng = np.random.default_rng(33)
    lors = random_lors(rng, 20000, radius=120.0, zhalf=60.0, dt_sigma_ps=84.93)
    fov = P.FOV(full, n)
    a = P.Mlem(fov, lors, tof=P.Tof(42.47, 3.0), n_subsets=2)
    b = P.Mlem(fov, lors, tof=P.Tof(42.47, 3.0), n_subsets=2)
    a.set_timing(True)
    for _ in range(5):
        a.step(copy_image=False)
        b.step(copy_image=False)
    last = a.last_kernel_timing()
    t = a.drain_kernel_timings(16)
    assert len(t) == 5 and t[-1] == last
    assert all(v >= 0 for d in t for v in d.values()) and all(d["forward_tiles"] > 0 for d in t)
    assert a.drain_kernel_timings(16) == []
    assert np.array_equal(a.image().data, b.image().data, equal_nan=True)
    for _ in range(300):
        a.step(copy_image=False)
    assert len(a.drain_kernel_timings(512)) == 256
    a.close()
    b.close()


def test_batch_matches_single_reconstructions(gpu):
    """A batch of reconstructions of different shapes, TOF settings, LOR
    counts (one empty) and sensitivity images gives each one's own images."""
    rng = np.random.default_rng(32)
    specs = [([48.0, 48.0, 40.0], [12, 12, 10], (42.47, 3.0), 20000),
             ([60.0, 60.0, 60.0], [30, 30, 30], None, 30000),
             ([40.0, 56.0, 32.0], [17, 23, 11], (31.85, 3.0), 15000),
             ([30.0, 30.0, 30.0], [6, 6, 6], None, 0),
             ([90.0, 90.0, 60.0], [45, 45, 30], (42.47, 3.0), 40000)]
    fovs, lors, tofs, sens = [], [], [], []
    for k, (full, n, tof, m) in enumerate(specs):
        fovs.append(P.FOV(full, n))
        lors.append(random_lors(rng, m, radius=150.0, zhalf=80.0, dt_sigma_ps=84.93 if tof else 0.0))
        tofs.append(P.Tof(*tof) if tof else None)
        nv = int(np.prod(n))
        sens.append(rng.uniform(0.5, 2.0, nv).astype(np.float32) if k % 2 == 0 else None)
    b = P.Batch(fovs, lors, tofs=tofs, sensitivities=sens, n_subsets=2)
    singles = [P.Mlem(fovs[k], lors[k], tof=tofs[k], sensitivity=sens[k], n_subsets=2) for k in range(len(specs))]
    for step in range(6):
        osem = b.step()
        assert (osem.iteration, osem.subset) == (step // 2 + 1, step % 2 + 1)
        for k, s in enumerate(singles):
            g, _ = s.step()
            e = rel_l2_finite(b.image(k).data, g.data)
            _say(f"batch step {step + 1} reconstruction {k}: relL2 vs single {e:.2e}")
            assert e <= 1e-6
    b.close()
    for s in singles:
        s.close()


def test_rows_large_transfer(gpu):
    """System-matrix rows of 1e5 LORs in one call (150 MB of indices and as
    much of weights come down through the pinned transfer ring) equal the
    same rows computed 5000 LORs at a time (plain copies); entries past a
    row's count are not written by the library and are not compared."""
    fov = P.FOV([180.0] * 3, [60] * 3)
    lors = P.synth_phantom_lors(8, 100_000, 200.0)
    tof = P.Tof(42.47, 3.0)
    idx, w, cnt, st = P.system_matrix_rows(fov, lors, tof=tof)
    parts = [P.system_matrix_rows(fov, lors[k:k + 5000], tof=tof, cap=idx.shape[1]) for k in range(0, 100_000, 5000)]
    assert idx.nbytes >= 32 << 20
    assert np.array_equal(cnt, np.concatenate([q[2] for q in parts]))
    assert np.array_equal(st, np.concatenate([q[3] for q in parts]))
    valid = np.arange(idx.shape[1])[None, :] < np.minimum(cnt, idx.shape[1])[:, None]
    assert np.array_equal(idx[valid], np.concatenate([q[0] for q in parts])[valid])
    assert np.array_equal(w.view(np.uint32)[valid], np.concatenate([q[1] for q in parts]).view(np.uint32)[valid])


def test_ingest_large_transfers(gpu):
    """Ingestion of 1.2e6 records in one call (52.8 MB of records up, about
    38 MB of LORs down: both through the pinned transfer ring) equals the
    same records ingested as four event ranges of 3e5 (plain copies),
    concatenated: every draw is keyed by the record's index."""
    rng = np.random.default_rng(78)
    n = 1_200_000
    rec = np.zeros(n, P.H5LOR_DTYPE)
    a1, a2 = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    rec["x1"], rec["y1"], rec["z1"] = 360 * np.cos(a1), 360 * np.sin(a1), rng.uniform(-160, 160, n)
    rec["x2"], rec["y2"], rec["z2"] = 360 * np.cos(a2), 360 * np.sin(a2), rng.uniform(-160, 160, n)
    rec["dt"] = np.arange(n, dtype=np.float32) * np.float32(1e-3)
    rec["q1"], rec["q2"] = rng.uniform(0, 3000, n), rng.uniform(0, 3000, n)
    rec["E1"], rec["E2"] = rng.normal(511, 5, n), rng.normal(511, 5, n)

    def params(lo, hi):
        return P.ingest_params(events=(lo, hi), smear_energy_fwhm=0.12, z_max=150.0,
                               discretize=(350.0, 20.0, 3.0, 3.0), adjust=P.ADJUST_RANDOM_Z_PHI, seed=7)

    whole, nw = P.ingest_lors(rec, params(0, n))
    parts = [P.ingest_lors(rec, params(k, k + 300_000))[0] for k in range(0, n, 300_000)]
    joined = np.concatenate(parts)
    _say(f"ingest in one call: {len(whole)} LORs ({len(whole) * 32 / 2**20:.1f} MiB), in four: {len(joined)}")
    assert nw == n and len(whole) * 32 >= 32 << 20
    assert whole.tobytes() == joined.tobytes()


def test_ingest_matches_oracle(gpu):
    """LOR ingestion (cuts, energy smearing, RandomZPhi and Random element
    adjustment, scattergram) on the device against the oracle's restatement
    with the same counter-based draws.  Records carry a unique dt, which
    passes through unchanged (x 1000), so LORs are matched by it; the device's
    and the host's atan2/sin/cos/log differ by an ulp, so positions agree to
    1e-3 mm except where an end sits within an ulp of an element boundary,
    and a record within an ulp of an energy bound could differ (none do
    here)."""
    rng = np.random.default_rng(77)
    n = 200_000
    rec = np.zeros(n, P.H5LOR_DTYPE)
    a1, a2 = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    r1, r2 = rng.uniform(350, 370, n), rng.uniform(350, 370, n)
    rec["x1"], rec["y1"], rec["z1"] = r1 * np.cos(a1), r1 * np.sin(a1), rng.uniform(-160, 160, n)
    rec["x2"], rec["y2"], rec["z2"] = r2 * np.cos(a2), r2 * np.sin(a2), rng.uniform(-160, 160, n)
    rec["dt"] = np.arange(n, dtype=np.float32) * np.float32(1e-3)
    rec["q1"], rec["q2"] = rng.uniform(0, 3000, n), rng.uniform(0, 3000, n)
    rec["E1"], rec["E2"] = rng.normal(500, 25, n), rng.normal(500, 25, n)
    spec = P.scattergram_spec(phi=12, r=(10, 120.0), z=(8, 300.0))
    for adjust in (P.ADJUST_RANDOM_Z_PHI, P.ADJUST_RANDOM, P.ADJUST_CENTRE):
        p = P.ingest_params(events=(1000, 190_000), charge=(100.0, 2900.0), energy=(430.0, 600.0),
                            smear_energy_fwhm=0.12, z_max=150.0, discretize=(350.0, 20.0, 3.0, 3.0),
                            adjust=adjust, seed=2024)
        g, ng = P.ingest_lors(rec, p, spec)
        o, no = O.ingest(rec, p, spec)
        _say(f"ingest adjust={adjust}: read {ng}, kept gpu {len(g)} oracle {len(o)}")
        assert ng == no == 189_000
        assert len(g) == len(o) and len(g) > 50_000
        assert np.array_equal(g["dt"], o["dt"])
        # an end whose azimuth or z sits within an ulp of an element boundary
        # may round to the neighbouring element on one side only
        d = np.maximum(np.abs(g["p1"] - o["p1"]).max(axis=1), np.abs(g["p2"] - o["p2"]).max(axis=1))
        off = d > 1e-3
        _say(f"  ends off by more than 1e-3 mm: {int(off.sum())} (max {d.max():.3g} mm)")
        assert off.mean() <= 1e-4 and d.max() <= 10.0
        # a LOR whose axis coordinate (the distance from the axis cancels
        # catastrophically near the axis) lands across a bin edge by an ulp
        # moves one count between two bins, and every LOR of those bins sees
        # a slightly different (s + t) / t
        # (with Centre adjustment many LORs join identical element centres,
        # whole groups share a coordinate, and LORs along 0/90/180/270 degrees
        # sit exactly on phi bin edges, where the ulp of sin/cos decides the
        # side: their bin values are not compared)
        same = np.isclose(g["ac"], o["ac"], rtol=1e-6)
        _say(f"  additive corrections equal for {same.mean() * 100:.4f}% of LORs")
        if adjust != P.ADJUST_CENTRE:
            assert same.mean() >= 0.995
