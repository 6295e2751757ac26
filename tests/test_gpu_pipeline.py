"""GPU parity of the fused pipeline against the oracle, stage by stage:
matcher volumes (bit-exact), per-pixel statistics, the 20 features (float32,
as scikit-learn sees them), the forest's cost, and the post-processing chain
(winner-take-all and final disparity agreement)."""
import numpy as np
import pytest

import oracle
from cbmv_amd import metrics, synth

pytestmark = pytest.mark.gpu

RM = oracle.RAND_MAX
H, W, D = 72, 112, 32
BATCH = W * 20          # 20-row feature batches, so batch fills are exercised


@pytest.fixture(scope="module")
def tiny_forest(golden_dir):
    from cbmv_amd import forest
    f = forest.load(f"{golden_dir}/rf_tiny.npz")
    return f, forest.Forest(f)


@pytest.fixture(scope="module")
def run(tiny_forest):
    from cbmv_amd.pipeline import Pipeline
    p = synth.fronto_pair(31, H=H, W=W, ndisp=D, dmin=4, dmax=26)
    pl = Pipeline(H, W, D, forest=tiny_forest[1], test_batch=BATCH)
    disp = pl.run_host(p.left, p.right)
    vols = oracle.get_costs(p.left, p.right, D)
    return p, pl, disp, vols


def test_volumes_bit_exact(run):
    p, pl, disp, vols = run
    c = pl.buffer("census").transpose(0, 2, 1)
    np.testing.assert_array_equal(np.where(c == 255, RM, c), vols[0])
    np.testing.assert_array_equal(pl.buffer("ncc").transpose(0, 2, 1), vols[1])
    s = pl.buffer("sob").transpose(0, 2, 1)
    np.testing.assert_array_equal(np.where(s == 65535, RM, s), vols[2])
    np.testing.assert_array_equal(pl.buffer("sad").transpose(0, 2, 1), vols[3])


def _batch_stats(v, rows):
    """min and sum of exp per pixel of the left and (batch-filled) right views."""
    out_l, out_r = [], []
    for b in range(0, v.shape[0], rows):
        blk = v[b:b + rows]
        for src, out in ((blk, out_l), (oracle.get_right_cost(blk), out_r)):
            m = src.min(axis=2)
            m = np.where(m < RM, m, RM)
            out.append(m)
    return np.concatenate(out_l), np.concatenate(out_r)


def test_statistics(run):
    p, pl, disp, vols = run
    sl, sr = pl.buffer("stats_left"), pl.buffer("stats_right")
    rows = BATCH // W
    for k, v in enumerate(vols):
        ml, mr = _batch_stats(v, rows)
        np.testing.assert_array_equal(sl[k], ml)
        np.testing.assert_array_equal(sr[k], mr)
        e = np.exp(-((v - ml[..., None]) ** 2) / oracle.SIGMAS[k]).sum(axis=2)
        np.testing.assert_allclose(sl[4 + k], e, rtol=1e-12)


def _right_sums(v, mr, rows, sigma):
    """Sum of exp over d of the (batch-filled) right view against its minima."""
    out = []
    for b in range(0, v.shape[0], rows):
        r = oracle.get_right_cost(v[b:b + rows])
        out.append(np.exp(-((r - mr[b:b + rows, :, None]) ** 2) / sigma).sum(axis=2))
    return np.concatenate(out)


@pytest.mark.parametrize("h,w,d,rows", [(20, 1100, 129, 6), (16, 90, 150, 4), (18, 600, 64, 9),
                                        (18, 40, 12, 3), (16, 64, 256, 4)])
def test_statistics_odd_shapes(tiny_forest, h, w, d, rows):
    """Both views' minima (exact) and sums (1e-12) against the oracle, with
    both views' minima from the matcher kernels: W off the 256-pixel tiles
    and below one wavefront's 64 columns, D not a multiple of the 8-step
    unroll and barely above the window kernels' double group, D > W (every
    right view runs past the row into the batch fill), batches of a few
    rows."""
    from cbmv_amd.pipeline import Pipeline
    p = synth.fronto_pair(35, H=h, W=w, ndisp=d, dmin=4, dmax=min(d - 6, w // 2))
    a = Pipeline(h, w, d, forest=tiny_forest[1], test_batch=w * rows)
    a.run_host(p.left, p.right)
    vols = oracle.get_costs(p.left, p.right, d)
    sl, sr = a.buffer("stats_left"), a.buffer("stats_right")
    for k, v in enumerate(vols):
        ml, mr = _batch_stats(v, rows)
        np.testing.assert_array_equal(sl[k], ml)
        np.testing.assert_array_equal(sr[k], mr)
        e = np.exp(-((v - ml[..., None]) ** 2) / oracle.SIGMAS[k]).sum(axis=2)
        np.testing.assert_allclose(sl[4 + k], e, rtol=1e-12)
        np.testing.assert_allclose(sr[4 + k], _right_sums(v, mr, rows, oracle.SIGMAS[k]), rtol=1e-12)


@pytest.mark.parametrize("wins,h,w,d,rows", [((5, 5, 3, 3), 20, 333, 97, 5),
                                              ((9, 7, 7, 5), 17, 150, 200, 3)])
def test_statistics_window_sizes(tiny_forest, wins, h, w, d, rows):
    """Both views' minima (exact) and sums (1e-12) against the oracle for
    other window sizes (census, NCC, zero-mean SAD, SAD of Sobel): the
    specialised window kernels' right-view minima at other group lengths
    (NCC 5, zero-mean SAD 3 and 7, SAD of Sobel 3) and, for a size the
    generic kernels take (NCC 7), the statistics kernel's own pass over the
    right view. (The pipeline takes SAD-of-Sobel windows up to 5, whose
    sums fit its 16-bit volume.)"""
    from cbmv_amd.pipeline import Pipeline
    cw, nw, zw, sw = wins
    p = synth.fronto_pair(36, H=h, W=w, ndisp=d, dmin=3, dmax=min(d - 6, w // 2))
    a = Pipeline(h, w, d, forest=tiny_forest[1], test_batch=w * rows, census_w=cw, ncc_w=nw,
                 sad_w=zw, sob_w=sw)
    a.run_host(p.left, p.right)
    L, R = p.left, p.right
    vols = (oracle.census(L, R, d, cw), oracle.swap_axes(oracle.nccNister(L, R, d, nw)),
            oracle.swap_axes(oracle.sadsob(oracle.sobel(L), oracle.sobel(R), d, sw)),
            oracle.swap_axes(oracle.zsad(L, R, d, zw)))
    sl, sr = a.buffer("stats_left"), a.buffer("stats_right")
    for k, v in enumerate(vols):
        ml, mr = _batch_stats(v, rows)
        np.testing.assert_array_equal(sl[k], ml)
        np.testing.assert_array_equal(sr[k], mr)
        e = np.exp(-((v - ml[..., None]) ** 2) / oracle.SIGMAS[k]).sum(axis=2)
        np.testing.assert_allclose(sl[4 + k], e, rtol=1e-12)
        np.testing.assert_allclose(sr[4 + k], _right_sums(v, mr, rows, oracle.SIGMAS[k]), rtol=1e-12)


def test_features_equal_reference_float32(run):
    p, pl, disp, vols = run
    ref = oracle.test_features(vols, test_batch=BATCH).astype(np.float32).reshape(H, W, D, 20)
    n_eq = n_tot = 0
    worst = 0.0
    for y in (0, 5, 19, 20, 21, 39, 40, 55, 71):
        got = pl.row_features(y).transpose(1, 0, 2)
        eq = got == ref[y]
        n_eq += eq.sum()
        n_tot += eq.size
        rel = np.abs(got.astype(np.float64) - ref[y]) / np.maximum(np.abs(ref[y]), 1e-30)
        worst = max(worst, float(rel[~eq].max()) if (~eq).any() else 0.0)
    frac = n_eq / n_tot
    print(f"features bit-equal after the float32 cast: {frac:.6f}, worst rel diff {worst:.3g}")
    assert n_eq == n_tot


def test_forest_cost(run, tiny_forest):
    p, pl, disp, vols = run
    ref_f = oracle.test_features(vols, test_batch=BATCH)
    exp = oracle.forest_predict(ref_f, tiny_forest[0]).astype(np.float32).reshape(H, W, D)
    got = pl.buffer("cost")
    frac = float(np.mean(got == exp))
    print(f"forest cost equal: {frac:.6f}, max abs diff {np.abs(got - exp).max():.3g}")
    np.testing.assert_array_equal(got, exp)


def _repeat_forest(f, k):
    """The forest's trees k times over (k * n_trees trees)."""
    n = len(f["left"])
    out = {key: np.concatenate([f[key]] * k) for key in ("feature", "threshold", "value")}
    for key in ("left", "right"):
        out[key] = np.concatenate([np.where(f[key] >= 0, f[key] + i * n, -1) for i in range(k)])
    out["roots"] = np.concatenate([f["roots"] + i * n for i in range(k)]).astype(np.int32)
    out["n_features"] = f["n_features"]
    out["value0"] = np.ascontiguousarray(out["value"][:, 0])
    return out


def test_forest_walks_agree(run, tiny_forest, monkeypatch):
    """The classifier kernel with each forest's top levels in LDS (forests of
    at most 50 trees, the default), the global-only walk (CBMV_FOREST_TOP=0)
    and a 60-tree forest (too many trees for the LDS copy) give the oracle's
    costs bit for bit."""
    from cbmv_amd import forest
    from cbmv_amd.pipeline import Pipeline
    p, pl, disp, vols = run
    ref_f = oracle.test_features(vols, test_batch=BATCH)
    monkeypatch.setenv("CBMV_FOREST_TOP", "0")
    g = Pipeline(H, W, D, forest=tiny_forest[1], test_batch=BATCH)
    monkeypatch.delenv("CBMV_FOREST_TOP")
    g.run_host(p.left, p.right)
    np.testing.assert_array_equal(g.buffer("cost"), pl.buffer("cost"))
    big = _repeat_forest(tiny_forest[0], 6)
    m = Pipeline(H, W, D, forest=forest.Forest(big), test_batch=BATCH)
    m.run_host(p.left, p.right)
    exp = oracle.forest_predict(ref_f, big).astype(np.float32).reshape(H, W, D)
    np.testing.assert_array_equal(m.buffer("cost"), exp)


def test_post_chain_agreement(run):
    p, pl, disp, vols = run
    cost = pl.buffer("cost").astype(np.float64)
    ref_disp, ref_wta = oracle.post_mem(oracle.normalise(p.left), oracle.normalise(p.right), cost)
    wta = pl.buffer("wta_left")
    a_wta = metrics.agreement(wta, ref_wta)
    a_disp = metrics.agreement(disp, ref_disp, tol=0.05)
    print(f"WTA agreement {a_wta:.5f}, final disparity agreement (|d|<=0.05) {a_disp:.5f}")
    assert a_wta >= 0.999 and a_disp >= 0.999


def test_aggregated_volume_error(run):
    """The float32 post-processing volume (SGM path values in double,
    accumulated in float32; aggregation sums per row then per column) against
    the oracle's float64 restatement of the chain on the same cost: the
    error of the winner-take-all's input, reported and bounded."""
    p, pl, disp, vols = run
    cost = pl.buffer("cost").astype(np.float64)
    il, ir = oracle.normalise(p.left), oracle.normalise(p.right)
    prm = oracle.DEFAULTS
    lc = oracle.cross(il, prm["L1"], prm["tau1"])
    rc = oracle.cross(ir, prm["L1"], prm["tau1"])
    ref = oracle._chain_volume(il, ir, lc, rc, cost, -1, prm, 4).transpose(1, 2, 0)
    got = pl.buffer("agg_left").astype(np.float64)
    err = np.abs(got - ref)
    rel = err / np.maximum(np.abs(ref), 1e-6)
    print(f"aggregated volume vs float64 oracle: max abs {err.max():.3g}, "
          f"max rel {rel.max():.3g}, mean rel {rel.mean():.3g} (values up to {np.abs(ref).max():.3g})")
    assert rel.max() < 1e-5


def _cbca_f32(vol, lc, rc, direction, L1=6):
    """One cross-based aggregation pass on an [H, W, D] float32 volume in
    float32, in the order the aggregation kernel adds: each row's sum over
    its horizontal range in ascending x (post.cu:420-426), those sums added
    in ascending y with their pixel counts, then sum / count rounded once;
    voxels whose match x + d * direction leaves the image keep their cost.
    Ranges: columns xs < x' < xt with xs = max(lc0(x), rc0(x_r) - d dir),
    xt = min(lc1(x), rc1(x_r) - d dir) within x +- (L1 - 1), rows likewise
    from the arms at (y, x) and (y, x_r) (post.cu:330-396)."""
    H, W, D = vol.shape
    R = L1 - 1
    x = np.arange(W)[:, None]
    dd = np.arange(D)[None, :] * direction
    xr = x + dd
    inside = (xr >= 0) & (xr < W)                      # [W, D]
    xrc = np.clip(xr, 0, W - 1)
    xs = np.maximum(lc[0][:, :, None], rc[0][:, xrc] - dd)   # [H, W, D]
    xt = np.minimum(lc[1][:, :, None], rc[1][:, xrc] - dd)
    hs = np.zeros((H, W, D), np.float32)
    hc = np.zeros((H, W, D), np.int64)
    for q in range(2 * R + 1):
        t = x - R + q                                   # [W, 1]
        v = vol[:, np.clip(t[:, 0], 0, W - 1), :]
        inc = (xs < t) & (t < xt) & (t >= 0) & (t < W)
        hs = hs + np.where(inc, v, np.float32(0))
        hc += inc
    ys = np.maximum(lc[2][:, :, None], rc[2][:, xrc])
    yt = np.minimum(lc[3][:, :, None], rc[3][:, xrc])
    y = np.arange(H)[:, None, None]
    s = np.zeros((H, W, D), np.float32)
    c = np.zeros((H, W, D), np.int64)
    for q in range(2 * R + 1):
        r = y - R + q
        rr = np.clip(r[:, 0, 0], 0, H - 1)
        inc = (ys < r) & (r < yt) & (r >= 0) & (r < H)
        s = s + np.where(inc, hs[rr], np.float32(0))
        c += np.where(inc, hc[rr], 0)
    mean = s / np.maximum(c, 1).astype(np.float32)
    return np.where(inside[None], mean, vol).astype(np.float32)


def test_aggregation_pass_bit_exact(tiny_forest):
    """One aggregation pass (no SGM) of the left view's cost against its
    float32 restatement in the kernel's order: equal bit for bit."""
    from cbmv_amd.pipeline import Pipeline
    p = synth.fronto_pair(35, H=H, W=W, ndisp=D, dmin=4, dmax=26)
    pl = Pipeline(H, W, D, forest=tiny_forest[1], sgm_iters=0, cbca_i2=1)
    pl.run_host(p.left, p.right)
    cost = pl.buffer("cost")
    il, ir = oracle.normalise(p.left), oracle.normalise(p.right)
    prm = oracle.DEFAULTS
    lc = oracle.cross(il, prm["L1"], prm["tau1"]).astype(np.float32)
    rc = oracle.cross(ir, prm["L1"], prm["tau1"]).astype(np.float32)
    ref = _cbca_f32(cost, lc, rc, -1, prm["L1"])
    got = pl.buffer("agg_left")
    print(f"aggregation pass equal: {np.mean(got == ref):.6f}")
    np.testing.assert_array_equal(got, ref)


def test_wta_is_first_minimum_of_aggregated_volume(run):
    """The winner-take-all map is the first minimum over d (test.py:409)
    of the final aggregated volume, exactly: the last aggregation pass
    merges its 64-disparity blocks' minima by 64-bit atomics."""
    p, pl, disp, vols = run
    agg = pl.buffer("agg_left")
    np.testing.assert_array_equal(pl.buffer("wta_left"), np.argmin(agg, axis=2).astype(np.float32))


def _chain_agreement(pl, disp, ref):
    """Agreement of every map the oracle's chain returns: (final, left WTA)
    for the _mem chain, plus (right WTA, outlier codes) with the LR check."""
    a = {"final (0.05 px)": metrics.agreement(disp, ref[0], tol=0.05),
         "WTA left": metrics.agreement(pl.buffer("wta_left"), ref[1])}
    if len(ref) == 4:
        a["WTA right"] = metrics.agreement(pl.buffer("wta_right"), ref[2])
        a["outlier"] = metrics.agreement(pl.buffer("outlier"), ref[3])
    return a


@pytest.mark.parametrize("lr_check,paths", [(1, 4), (1, 8), (0, 8)])
def test_chain_variants(tiny_forest, lr_check, paths):
    from cbmv_amd.pipeline import Pipeline
    p = synth.fronto_pair(32, H=H, W=W, ndisp=D, dmin=4, dmax=26)
    pl = Pipeline(H, W, D, forest=tiny_forest[1], lr_check=lr_check, sgm_paths=paths)
    disp = pl.run_host(p.left, p.right)
    assert np.isfinite(disp).all()
    cost = pl.buffer("cost").astype(np.float64)
    fn = oracle.post_interp if lr_check else oracle.post_mem
    ref = fn(oracle.normalise(p.left), oracle.normalise(p.right), cost, npaths=paths)
    b = metrics.bad_tau(disp, p.gt, 2.0)["all"]
    a = _chain_agreement(pl, disp, ref)
    print(f"lr_check={lr_check} paths={paths}: " + ", ".join(f"{k} {v:.5f}" for k, v in a.items())
          + f", bad-2.0 {b:.2f}%")
    for k, v in a.items():
        assert v >= 0.999, k


@pytest.mark.parametrize("h,w,d", [(37, 70, 100), (50, 130, 160), (23, 300, 64)])
def test_odd_shapes(tiny_forest, h, w, d):
    """Sizes off every tile: D not a multiple of 64 (partial lanes in the SGM
    scan, WTA and aggregation), W off the 8-column aggregation strips and
    the 64-pixel tiles, H of a few row chunks. Volumes and statistics exact,
    the LR-check chain against the oracle."""
    from cbmv_amd.pipeline import Pipeline
    rows = 8
    p = synth.fronto_pair(34, H=h, W=w, ndisp=d, dmin=4, dmax=min(d - 6, w // 2))
    pl = Pipeline(h, w, d, forest=tiny_forest[1], test_batch=w * rows, lr_check=1, sgm_paths=8)
    disp = pl.run_host(p.left, p.right)
    assert np.isfinite(disp).all()
    vols = oracle.get_costs(p.left, p.right, d)
    c = pl.buffer("census").transpose(0, 2, 1)
    np.testing.assert_array_equal(np.where(c == 255, RM, c), vols[0])
    np.testing.assert_array_equal(pl.buffer("ncc").transpose(0, 2, 1), vols[1])
    s = pl.buffer("sob").transpose(0, 2, 1)
    np.testing.assert_array_equal(np.where(s == 65535, RM, s), vols[2])
    np.testing.assert_array_equal(pl.buffer("sad").transpose(0, 2, 1), vols[3])
    sl, sr = pl.buffer("stats_left"), pl.buffer("stats_right")
    for k, v in enumerate(vols):
        ml, mr = _batch_stats(v, rows)
        np.testing.assert_array_equal(sl[k], ml)
        np.testing.assert_array_equal(sr[k], mr)
        e = np.exp(-((v - ml[..., None]) ** 2) / oracle.SIGMAS[k]).sum(axis=2)
        np.testing.assert_allclose(sl[4 + k], e, rtol=1e-12)
    cost = pl.buffer("cost").astype(np.float64)
    ref = oracle.post_interp(oracle.normalise(p.left), oracle.normalise(p.right), cost, npaths=8)
    a = _chain_agreement(pl, disp, ref)
    print(f"{h}x{w}x{d}: " + ", ".join(f"{k} {v:.5f}" for k, v in a.items()))
    for k, v in a.items():
        assert v >= 0.999, k
    # partial disparity blocks: the WTA is still the first minimum, exactly
    agg = pl.buffer("agg_left")
    np.testing.assert_array_equal(pl.buffer("wta_left"), np.argmin(agg, axis=2).astype(np.float32))


@pytest.mark.parametrize("D", [32, 300])
def test_right_view_cost_reindex(golden_dir, D):
    """The LR chain's right-view cost is P[y][x+d][d], or P[0,0,0] where x+d
    leaves the image (get_right_cost, featextract.cpp:115-152), exactly;
    D = 300 takes the narrower LDS tile."""
    from cbmv_amd import forest, synth
    from cbmv_amd.pipeline import Pipeline
    H, W = 24, 96 if D == 32 else 340
    p = synth.fronto_pair(77, H=H, W=W, ndisp=D, dmin=2, dmax=min(D - 8, 120))
    pl = Pipeline(H, W, D, forest=forest.Forest(forest.load(f"{golden_dir}/rf_tiny.npz")), lr_check=1)
    pl.run_host(p.left, p.right)
    P = pl.buffer("cost")
    PR = pl.buffer("cost_right")
    y, x, d = np.meshgrid(np.arange(H), np.arange(W), np.arange(D), indexing="ij")
    xs = x + d
    ref = np.where(xs < W, P[y, np.minimum(xs, W - 1), d], P[0, 0, 0])
    np.testing.assert_array_equal(PR, ref)

