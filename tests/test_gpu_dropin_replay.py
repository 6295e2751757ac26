"""The reference's call sequences replayed through the drop-in modules
(cbmv_amd.compat: libmatchers, libfeatextract, post) argument by argument,
as test.py makes them, and compared with the oracle's compositions:

* Testing.__get_costs (test.py:122-138): the four volumes, with the
  .astype(float64) and swap_axes calls, bit for bit;
* test_model's batch loop (test.py:597-626) over __extract_features_lr
  (test.py:521-595): every feature column bit for bit except the eight
  likelihoods (device exp against glibc exp, 1e-13 relative), then
  predict_proba on the drop-in forest, and eval_prob (test.py:143-149);
* __postprocessing_mem (test.py:341-449) and __postprocessing_mem_interp
  (test.py:153-338) on the device arrays with their shape mutations
  (swap_axis*, CUDA_copy_double), the [W, D] tmp array, set_to_zero
  before each SGM and CUDA_divide_double(., 4): every map the chain
  produces bit for bit against the oracle's chain on the same cost.
"""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

H, W, D = 44, 96, 24
TEST_BATCH = W * 12          # r_s = 12 rows: batches of 12, 12, 12 and 8 rows


class Replay:
    """Testing's methods, in Python 3, calling the drop-in modules."""

    censw, nccw, sadw = 11, 3, 5
    cens_sigma, ncc_sigma, sad_sigma = 2 * 8 ** 2, 0.02, 2 * 100 ** 2
    pi1, pi2, sgm_q1, sgm_q2, alpha1, tau_so, sgm_i = 1.3, 10, 2.2, 3, 1.4, .12, 2
    L1, tau1, cbca_i1, cbca_i2, median_i, median_w, blur_sigma, blur_t = 6, .07, 0, 4, 1, 3, 2.2, 1.3

    def __init__(self, imgl, imgr, d, test_batch):
        from cbmv_amd.compat import libfeatextract, libmatchers, post
        self.mtc, self.fte, self.post = libmatchers, libfeatextract, post
        self.imgl, self.imgr = imgl, imgr
        self.h, self.w = imgl.shape
        self.d = d
        self.test_batch = test_batch

    def get_costs(self, iml, imr):
        mtc, fte = self.mtc, self.fte
        costcensus = mtc.census(iml, imr, self.d, self.censw).astype(np.float64)
        costncc = fte.swap_axes(mtc.nccNister(iml, imr, self.d, self.nccw))
        sobl, sobr = mtc.sobel(iml), mtc.sobel(imr)
        costsob = fte.swap_axes(mtc.sadsob(sobl, sobr, self.d, 5).astype(np.float64))
        costsad = fte.swap_axes(mtc.zsad(iml, imr, self.d, self.sadw).astype(np.float64))
        return costcensus, costncc, costsob, costsad

    def extract_features_lr(self, census, ncc, sobel, sad):
        fte = self.fte
        dims = census.shape
        n2, n3 = dims[0] * dims[1], dims[0] * dims[1] * dims[2]
        eps = (.01, 1.01, .01, .01)
        sig = (self.cens_sigma, self.ncc_sigma, self.sad_sigma, self.sad_sigma)
        features = np.empty((n3, 20))
        for k, v in enumerate((census, ncc, sobel, sad)):
            vr = np.reshape(fte.get_right_cost(v), [n2, dims[2]])
            v = np.reshape(v, [n2, dims[2]])
            features[:, k] = np.reshape(v, [n3])
            features[:, 4 + k] = np.reshape(fte.extract_ratio(v, eps[k]), [n3])
            features[:, 8 + k] = np.reshape(fte.extract_likelihood(v, sig[k]), [n3])
            r_pkrn = np.reshape(fte.extract_ratio(vr, eps[k]), [dims[0], dims[1], dims[2]])
            features[:, 12 + k] = np.reshape(fte.get_left_cost(r_pkrn), [n3])
            r_aml = np.reshape(fte.extract_likelihood(vr, sig[k]), [dims[0], dims[1], dims[2]])
            features[:, 16 + k] = np.reshape(fte.get_left_cost(r_aml), [n3])
        return features

    def test_model(self, rf, keep_features=False):
        census, ncc, sobel, sad = self.get_costs(self.imgl, self.imgr)
        proba = np.empty((0, 2))
        feats = []
        r_s = int(math.floor(self.test_batch / self.w))
        batch_index = 0
        while batch_index < self.h:
            features = self.extract_features_lr(census[batch_index:batch_index + r_s],
                                                ncc[batch_index:batch_index + r_s],
                                                sobel[batch_index:batch_index + r_s],
                                                sad[batch_index:batch_index + r_s])
            if keep_features:
                feats.append(features)
            proba = np.append(proba, rf.predict_proba(features), axis=0)
            batch_index += r_s
        return proba, feats

    def eval_prob(self, prob):
        prob = np.reshape(prob, [self.h, self.w, self.d, 2])
        prob = self.fte.get_cost(prob)
        probr = self.fte.get_right_cost(prob)
        return prob, probr

    def normalised(self):
        imgl = self.imgl.astype(np.float32)
        imgr = self.imgr.astype(np.float32)
        return (imgl - np.mean(imgl)) / np.std(imgl), (imgr - np.mean(imgr)) / np.std(imgr)

    def _cross(self, imgl, imgr):
        post = self.post
        cross_array_init = np.zeros((4, imgl.shape[0], imgl.shape[1])).astype(np.float32)
        imgl_d, imgr_d = post.CUDA_float_array(imgl), post.CUDA_float_array(imgr)
        leftc_d = post.CUDA_float_array(cross_array_init)
        rightc_d = post.CUDA_float_array(cross_array_init)
        post.cross(imgl_d, leftc_d, self.L1, self.tau1)
        post.cross(imgr_d, rightc_d, self.L1, self.tau1)
        return imgl_d, imgr_d, leftc_d, rightc_d

    def _aggregate(self, imgl_d, imgr_d, ac_d, bc_d, cost, direction):
        """The shared volume sequence of test.py:176-228 / 236-286 / 353-405."""
        post = self.post
        cost_d = post.CUDA_double_array(cost)
        temp_cost_d = post.CUDA_double_array(cost)
        post.swap_axis_back(cost_d, temp_cost_d)
        post.CUDA_copy_double(temp_cost_d, cost_d)
        for _ in range(self.cbca_i1):
            post.cbca(ac_d, bc_d, cost_d, temp_cost_d, direction)
            post.CUDA_copy_double(temp_cost_d, cost_d)
        post.swap_axis(cost_d, temp_cost_d)
        post.CUDA_copy_double(temp_cost_d, cost_d)
        tmp = np.zeros((cost.shape[1], cost.shape[2]))
        tmp_d = post.CUDA_double_array(tmp)
        for _ in range(self.sgm_i):
            temp_cost_d.set_to_zero()
            tmp_d.set_to_zero()
            post.sgm(imgl_d, imgr_d, cost_d, temp_cost_d, tmp_d, self.pi1, self.pi2, self.tau_so,
                     self.alpha1, self.sgm_q1, self.sgm_q2, direction)
            post.CUDA_divide_double(temp_cost_d, 4)
            post.CUDA_copy_double(temp_cost_d, cost_d)
        post.swap_axis_back(cost_d, temp_cost_d)
        post.CUDA_copy_double(temp_cost_d, cost_d)
        for _ in range(self.cbca_i2):
            post.cbca(ac_d, bc_d, cost_d, temp_cost_d, direction)
            post.CUDA_copy_double(temp_cost_d, cost_d)
        return cost_d

    def _filters(self, disp_d, disp_temp):
        post = self.post
        for _ in range(self.median_i):
            post.median2d(disp_d, disp_temp, self.median_w)
            post.CUDA_copy_float(disp_temp, disp_d)
        post.mean2d(disp_d, disp_temp, oracle.gaussian(self.blur_sigma), self.blur_t)
        post.CUDA_copy_float(disp_temp, disp_d)
        return disp_d.get_array()

    def postprocessing_mem(self, imgl, imgr, cost, direction):
        post = self.post
        assert post.InitCUDA()
        imgl_d, imgr_d, leftc_d, rightc_d = self._cross(imgl, imgr)
        cost_d = self._aggregate(imgl_d, imgr_d, leftc_d, rightc_d, cost, direction)
        cc = cost_d.get_array()
        disp_l = np.argmin(cc, axis=0).astype(np.float32)
        disp_d = post.CUDA_float_array(disp_l)
        disp_temp = post.CUDA_float_array(disp_l)
        post.subpixel_enchancement(disp_d, cost_d, disp_temp)
        post.CUDA_copy_float(disp_temp, disp_d)
        return self._filters(disp_d, disp_temp), disp_l, cc.shape

    def postprocessing_mem_interp(self, imgl, imgr, cost, direction):
        post = self.post
        r_cost = self.fte.get_right_cost(cost)
        assert post.InitCUDA()
        imgl_d, imgr_d, leftc_d, rightc_d = self._cross(imgl, imgr)
        r_cost_d = self._aggregate(imgl_d, imgr_d, rightc_d, leftc_d, r_cost, -1 * direction)
        r_disp = np.argmin(r_cost_d.get_array(), axis=0).astype(np.float32)
        r_disp_d = post.CUDA_float_array(r_disp)
        cost_d = self._aggregate(imgl_d, imgr_d, leftc_d, rightc_d, cost, direction)
        cc = cost_d.get_array()
        disp_l = np.argmin(cc, axis=0).astype(np.float32)
        out_s = cc.shape[0]
        disp_d = post.CUDA_float_array(disp_l)
        disp_temp = post.CUDA_float_array(disp_l)
        post.subpixel_enchancement(disp_d, cost_d, disp_temp)
        outlier = np.zeros((r_disp.shape[0], r_disp.shape[1])).astype(np.float32)
        outlier_d = post.CUDA_float_array(outlier)
        post.outlier_detection(disp_d, r_disp_d, outlier_d, out_s)
        post.interpolate_mismatch(disp_d, outlier_d, disp_temp)
        post.CUDA_copy_float(disp_temp, disp_d)
        post.interpolate_occlusion(disp_d, outlier_d, disp_temp)
        post.CUDA_copy_float(disp_temp, disp_d)
        post.CUDA_copy_float(disp_temp, disp_d)
        outl = outlier_d.get_array()
        return self._filters(disp_d, disp_temp), disp_l, r_disp, outl


@pytest.fixture(scope="module", params=["fronto", "slanted"])
def replay(request):
    """A fronto-parallel pair and a slanted one whose LR check finds
    mismatched pixels (about 2.7 %), so interpolate_mismatch is exercised."""
    from cbmv_amd import synth
    if request.param == "fronto":
        p = synth.fronto_pair(41, H=H, W=W, ndisp=D, dmin=3, dmax=18)
    else:
        p = synth.slanted_pair(41, H=H, W=W, ndisp=D)
    return Replay(p.left, p.right, D, TEST_BATCH), p, request.param


@pytest.fixture(scope="module")
def tiny(golden_dir):
    from cbmv_amd import forest
    f = forest.load(f"{golden_dir}/rf_tiny.npz")
    return f, forest.Forest(f)


@pytest.fixture(scope="module")
def model_run(replay, tiny):
    rp, p, kind = replay
    proba, feats = rp.test_model(tiny[1], keep_features=True)
    return proba, feats


def test_get_costs_replay(replay):
    rp, p, kind = replay
    got = rp.get_costs(p.left, p.right)
    ref = oracle.get_costs(p.left, p.right, D)
    for g, r in zip(got, ref):
        assert g.shape == (H, W, D) and g.dtype == np.float64
        np.testing.assert_array_equal(g, r)


def test_features_and_predict_replay(replay, tiny, model_run):
    rp, p, kind = replay
    proba, feats = model_run
    vols = oracle.get_costs(p.left, p.right, D)
    r_s = int(math.floor(TEST_BATCH / W))
    assert len(feats) == math.ceil(H / r_s) == 4
    exact = [0, 1, 2, 3, 4, 5, 6, 7, 12, 13, 14, 15]
    like = [8, 9, 10, 11, 16, 17, 18, 19]
    for b, got in enumerate(feats):
        ref = oracle.features_lr(*(v[b * r_s:(b + 1) * r_s] for v in vols))
        np.testing.assert_array_equal(got[:, exact], ref[:, exact])
        np.testing.assert_allclose(got[:, like], ref[:, like], rtol=1e-13, atol=1e-300)
    ref_p = oracle.forest_predict(oracle.test_features(vols, TEST_BATCH), tiny[0])
    assert proba.shape == (H * W * D, 2)
    np.testing.assert_array_equal(proba[:, 0], ref_p)
    np.testing.assert_allclose(proba.sum(axis=1), 1.0, rtol=1e-12)
    prob, probr = rp.eval_prob(proba)
    np.testing.assert_array_equal(prob, ref_p.reshape(H, W, D))
    np.testing.assert_array_equal(probr, oracle.get_right_cost(ref_p.reshape(H, W, D)))


def test_postprocessing_mem_replay(replay, model_run):
    rp, p, kind = replay
    prob, _ = rp.eval_prob(model_run[0])
    il, ir = rp.normalised()
    disp, wta, shape = rp.postprocessing_mem(il, ir, prob, -1)
    assert shape == (D, H, W)
    ref_disp, ref_wta = oracle.post_mem(oracle.normalise(p.left), oracle.normalise(p.right), prob)
    np.testing.assert_array_equal(wta, ref_wta)
    np.testing.assert_array_equal(disp, ref_disp)


def test_postprocessing_mem_interp_replay(replay, model_run):
    rp, p, kind = replay
    prob, _ = rp.eval_prob(model_run[0])
    il, ir = rp.normalised()
    disp, wta, r_disp, outl = rp.postprocessing_mem_interp(il, ir, prob, -1)
    ref = oracle.post_interp(oracle.normalise(p.left), oracle.normalise(p.right), prob)
    np.testing.assert_array_equal(wta, ref[1])
    np.testing.assert_array_equal(r_disp, ref[2])
    np.testing.assert_array_equal(outl, ref[3])
    np.testing.assert_array_equal(disp, ref[0])
    print(f"replayed _mem_interp: {np.mean(outl == 0):.3f} matched, {np.mean(outl == 1):.3f} occluded, "
          f"{np.mean(outl == 2):.3f} mismatched pixels")
    assert np.mean(outl == 1) > 0.05
    if kind == "slanted":
        assert np.mean(outl == 2) > 0.01
