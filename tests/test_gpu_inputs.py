"""GPU parity of the hot path's per-LOR inputs and generators, through the C ABI.

* additive corrections (projector.rs:67-69: p <- p * additive_correction,
  fed by the scattergram, hdf5.rs:208-214): drawn from [1, 3] with a few
  FLT_MAX entries (empty scattergram bins, lorogram.rs:65-74), for both
  schemes, several subsets (the corrections are reordered with the LORs
  inside each subset) and a batch;
* the sensitivity-LOR generator (continuous.rs:20-47) against its host
  restatement in the oracle: same endpoints, radius D/2, |z| <= L/2, every
  LOR through the FOV; a FOV the detector's LORs cannot find is an error;
* the shifted-walk counter against the oracle's.
"""
import numpy as np
import pytest

import oracle_ffi as O
import petalo_amd as P
from helpers import random_lors, rel_l2_finite

pytestmark = pytest.mark.gpu

FLT_MAX = np.float32(3.4028234663852886e38)


def _say(msg):
    print(msg, flush=True)


def _with_corrections(lors, seed, n_max=5):
    rng = np.random.default_rng(seed)
    out = lors.copy()
    out["ac"] = rng.uniform(1.0, 3.0, len(out)).astype(np.float32)
    out["ac"][rng.choice(len(out), n_max, replace=False)] = FLT_MAX
    return out


def _oracle_osem(full, n, lors, iters, subsets, tof=None, threads=16):
    x = np.ones(int(np.prod(n)), np.float32)
    size = len(lors) // subsets
    out = []
    for _ in range(iters):
        for s in range(subsets):
            O.mlem_update(full, n, lors[s * size:(s + 1) * size], x, tof=tof, threads=threads)
            out.append(x.copy())
    return out


@pytest.mark.parametrize("scheme", [P.SCHEME_ATOMIC, P.SCHEME_TILE])
@pytest.mark.parametrize("subsets", [1, 3])
@pytest.mark.parametrize("tof", [None, (42.47, 3.0)])
def test_mlem_additive_corrections(gpu, scheme, subsets, tof):
    full, n = [60.0, 60.0, 48.0], [15, 15, 12]
    rng = np.random.default_rng(21)
    lors = _with_corrections(random_lors(rng, 30_001, dt_sigma_ps=84.93 if tof else 0.0), 22)
    fov = P.FOV(full, n)
    it = P.mlem(fov, lors, tof=P.Tof(*tof) if tof else None, n_subsets=subsets, scheme=scheme, sort_lors=True)
    o = _oracle_osem(full, n, lors, 3, subsets, tof=tof)
    for k in range(3 * subsets):
        img, _ = next(it)
        e = rel_l2_finite(img.data, o[k])
        _say(f"ac in [1,3] + FLT_MAX, scheme {scheme}, {subsets} subsets, tof {tof}: update {k + 1} relL2 {e:.2e}")
        assert e <= 1e-5
    it.close()


def test_mlem_additive_corrections_phantom_tiled(gpu):
    """The 60^3 phantom workload with corrections, tiled scheme, 5 iterations."""
    full, n = [180.0] * 3, [60] * 3
    lors = _with_corrections(P.synth_phantom_lors(4, 300_000, 0.0), 23, n_max=20)
    it = P.mlem(P.FOV(full, n), lors, n_subsets=2, scheme=P.SCHEME_TILE)
    o = _oracle_osem(full, n, lors, 5, 2)
    for k in range(10):
        img, _ = next(it)
        e = rel_l2_finite(img.data, o[k])
        _say(f"phantom 60^3 with corrections: update {k + 1} relL2 {e:.2e}")
        assert e <= 1e-5 if k < 4 else e <= 1e-4
    it.close()


def test_batch_additive_corrections(gpu):
    """Three reconstructions of different shapes, each with its own corrections,
    4 subsets (the batch reads each subset's corrections at its own offset)."""
    specs = [([60.0, 60.0, 48.0], [15, 15, 12], None), ([90.0] * 3, [18] * 3, (42.47, 3.0)),
             ([72.0, 60.0, 60.0], [12, 10, 10], (31.85, 3.0))]
    rng = np.random.default_rng(31)
    fovs, lors, tofs = [], [], []
    for k, (full, n, tof) in enumerate(specs):
        fovs.append(P.FOV(full, n))
        lors.append(_with_corrections(random_lors(rng, 20_003 + 1000 * k, radius=80.0,
                                                  dt_sigma_ps=84.93 if tof else 0.0), 40 + k))
        tofs.append(P.Tof(*tof) if tof else None)
    b = P.Batch(fovs, lors, tofs=tofs, n_subsets=4)
    refs = [_oracle_osem(full, n, lors[k], 2, 4, tof=tof) for k, (full, n, tof) in enumerate(specs)]
    for step in range(8):
        b.step()
        for k in range(len(specs)):
            e = rel_l2_finite(b.image(k).data, refs[k][step])
            assert e <= 1e-5, (step, k, e)
    _say("batch with corrections: 8 subset updates of 3 reconstructions within 1e-5 of the oracle")
    b.close()


def test_sensitivity_lor_stream_matches_restatement(gpu):
    full, n = [180.0] * 3, [30] * 3
    fov = P.FOV(full, n)
    N, seed, L, D = 200_000, 17, 300.0, 700.0
    dev = P.sensitivity_lors(fov, N, seed, L, D)
    host, att = O.sensitivity_lors(full, n, N, seed, L, D)
    assert np.all(att > 0)
    d = np.maximum(np.abs(dev["p1"] - host["p1"]).max(1), np.abs(dev["p2"] - host["p2"]).max(1))
    same = d <= 1e-4
    _say(f"sensitivity LORs: {same.sum()} of {N} within 1e-4 mm of the restatement (largest {d[same].max():.1e} mm); "
         f"mean draws per LOR {att.mean():.2f}")
    # a LOR whose binary32 cosine or sine differs by an ulp between device and
    # host can graze the FOV in one and miss it in the other, and then takes
    # another attempt's endpoints
    assert same.mean() >= 0.9999
    for p in ("p1", "p2"):
        r = np.hypot(dev[p][:, 0].astype(np.float64), dev[p][:, 1].astype(np.float64))
        assert np.all(np.abs(r - D / 2) <= 1e-4 * D / 2)
        assert np.all(np.abs(dev[p][:, 2]) <= L / 2)
    assert np.all(dev["dt"] == 0.0) and np.all(dev["ac"] == 1.0)
    hits = O.entry_shifted(full, n, dev[:20000])
    assert np.all(hits >= 0), "every generated LOR enters the FOV"
    # the acceptance rate agrees with a binary64 Monte Carlo of uniform endpoints
    rng = np.random.default_rng(5)
    m = 400_000
    t = rng.uniform(0, 2 * np.pi, (2, m))
    z = rng.uniform(-L / 2, L / 2, (2, m))
    p1 = np.stack([D / 2 * np.cos(t[0]), D / 2 * np.sin(t[0]), z[0]], 1)
    p2 = np.stack([D / 2 * np.cos(t[1]), D / 2 * np.sin(t[1]), z[1]], 1)
    dd = p2 - p1
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (-90.0 - p1) / dd
        b = (90.0 - p1) / dd
    lo = np.max(np.minimum(a, b), 1)
    hi = np.min(np.maximum(a, b), 1)
    rate_mc = float(np.mean(np.maximum(lo, 0.0) <= np.minimum(hi, 1.0)))
    rate = N / float(att.sum())
    _say(f"acceptance: stream {rate:.4f}, Monte Carlo {rate_mc:.4f}")
    assert abs(rate - rate_mc) <= 4 * np.sqrt(rate_mc * (1 - rate_mc) / m) + 4 * np.sqrt(rate * (1 - rate) / N)


def test_sensitivity_unreachable_fov_is_an_error(gpu):
    fov = P.FOV([0.001] * 3, [2, 2, 2])
    mu = P.Image(fov, np.zeros(fov.nvox, np.float32))
    with pytest.raises(P.PetaloError) as e:
        P.make_sensitivity_image(mu, 1000, 300.0, 700.0, seed=3)
    assert e.value.code == 1 and "4096" in str(e.value)
    with pytest.raises(P.PetaloError):
        P.sensitivity_lors(fov, 1000, 3, 300.0, 700.0)


@pytest.mark.parametrize("nv", [60, 120])
def test_shifted_walks_counted(gpu, nv):
    full, n = [180.0] * 3, [nv] * 3
    lors = P.synth_phantom_lors(1, 200_000, 200.0)
    tof = (42.47, 3.0)
    ctx = P.Mlem(P.FOV(full, n), lors, tof=P.Tof(*tof))
    g = ctx.walk_stats()
    ctx.close()
    o = O.walk_stats(full, n, lors, tof=tof, threads=16)
    _say(f"{nv}^3: shifted walks GPU {g['n_shifted']} oracle {o['n_shifted']} of {g['n_lors']}")
    assert g["n_shifted"] == o["n_shifted"] and g["n_steps"] == o["n_steps"]
    # the sensitivity path counts them too (generated LORs, checked on the oracle)
    fov = P.FOV(full, n)
    mu = P.Image(fov, np.zeros(fov.nvox, np.float32))
    gen = P.sensitivity_lors(fov, 100_000, 3, 300.0, 700.0)
    _, st = P.make_sensitivity_image(mu, len(gen), lors=gen, return_stats=True)
    so = O.walk_stats(full, n, gen, threads=16)
    assert st["n_shifted"] == so["n_shifted"]
