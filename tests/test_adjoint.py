"""Adjoint of the pair sums and the chi^2 gradient.

No reference counterpart exists; pinned by (a) the oracle's FP64 adjoint
(oracle_xi_adjoint), (b) central finite differences in FP64 (step 1e-3 on the
weight-model parameters), (c) torch autograd of the per-forest maps."""
import numpy as np
import pytest

import oracle
from lya2pcf_amd import adjoint, forests, synth


def test_weight_model_and_projection_cpu():
    import torch
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 30, seed=4))
    wm = adjoint.WeightModel(s.z, s.sigma2_pip, s.config.z_nodes, "cpu")
    w = wm(torch.tensor(1.0, dtype=torch.float64), torch.full((10,), 0.05, dtype=torch.float64))
    np.testing.assert_allclose(w.numpy(), s.we, rtol=1e-15)
    pr = adjoint.Projection(s.pix_off, s.loglam, "cpu")
    dw = pr(torch.from_numpy(s.delta), w)
    np.testing.assert_allclose(dw.numpy(), s.dw, rtol=1e-9, atol=1e-12)
    # projection annihilates a + b * loglam
    z = pr(3.0 + 5.0 * torch.from_numpy(s.loglam), w)
    assert float(z.abs().max()) < 1e-9


def _setup(n=80, seed=17):
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], n, seed=seed, length_range=(40, 80)))
    fs = forests.ForestSet.from_survey(s, float32_inputs=True)
    return s, fs, forests.angmax_from(fs, 200.0)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_adjoint_kernel_matches_oracle(mode):
    import torch
    s, fs, am = _setup()
    eng = adjoint.AdjointEngine(fs, am, mode=mode)
    rng = np.random.default_rng(1)
    a = rng.normal(size=2500)
    b = rng.normal(size=2500)
    g_dw, g_w = eng.adjoint(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    torch.cuda.synchronize()
    e_dw, e_w = oracle.xi_adjoint(fs, eng.fwd.nb_off, eng.fwd.nb_idx, a, b)
    s_dw, s_w = oracle.xi_adjoint(fs, eng.fwd.nb_off, eng.fwd.nb_idx, np.abs(a), np.abs(b),
                                  we=np.abs(fs.we), dw=np.abs(fs.dw))
    assert np.all(np.abs(g_dw.cpu().numpy() - e_dw) <= 1e-12 * s_dw + 1e-300)
    assert np.all(np.abs(g_w.cpu().numpy() - e_w) <= 1e-12 * s_w + 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_chi2_gradient_vs_finite_differences(mode):
    import torch
    s, fs, am = _setup()
    dev = "cuda"
    eng = adjoint.AdjointEngine(fs, am, mode=mode)
    ref = adjoint.AdjointEngine(fs, am, mode="fp64", pairs=(eng.fwd.nb_off, eng.fwd.nb_idx))
    wm = adjoint.WeightModel(s.z, s.sigma2_pip, s.config.z_nodes, dev)
    pr = adjoint.Projection(s.pix_off, s.loglam, dev)
    m = torch.from_numpy(adjoint.model_xi()).to(dev)
    delta = torch.from_numpy(s.delta.astype(np.float32).astype(np.float64)).to(dev)

    def chi2(eta, s2, e):
        w = wm(eta, s2)
        dw = pr(delta, w)
        W, DW = adjoint.pair_sums(dw, w, e)
        return adjoint.chi2(W, DW, m, 1e-4)

    eta = torch.tensor(1.0, dtype=torch.float64, device=dev, requires_grad=True)
    s2 = torch.full((10,), 0.05, dtype=torch.float64, device=dev, requires_grad=True)
    L = chi2(eta, s2, eng)
    L.backward()
    grads = np.concatenate([[float(eta.grad)], s2.grad.cpu().numpy()])
    # central differences with step h = 1e-3, Richardson-extrapolated with h/2:
    # the plain step-1e-3 difference carries an O(h^2) truncation error of order
    # (h / (sigma2_pip + sigma2_lss))^2 ~ 1e-3 here, the extrapolated one O(h^4).
    h = 1e-3
    fd = []
    with torch.no_grad():
        for k in range(11):
            def f(step):
                e = torch.tensor(1.0, dtype=torch.float64, device=dev)
                v = torch.full((10,), 0.05, dtype=torch.float64, device=dev)
                if k == 0:
                    e = e + step
                else:
                    v[k - 1] += step
                return float(chi2(e, v, ref))
            d1 = (f(h) - f(-h)) / (2 * h)
            d2 = (f(h / 2) - f(-h / 2)) / h
            fd.append((4 * d2 - d1) / 3)
    fd = np.array(fd)
    print(mode, "chi2", float(L.detach()), "grad", grads, "fd", fd, flush=True)
    used = np.abs(fd) > 1e-8 * np.abs(fd).max()
    rel = np.abs(grads - fd)[used] / np.abs(fd[used])
    assert np.max(rel) < 1e-5, rel


@pytest.mark.gpu
def test_gradient_wrt_delta_and_w_pixels():
    """dchi2/ddelta_k and dchi2/dw_k at sampled pixels against central differences."""
    import torch
    s, fs, am = _setup(60, 23)
    dev = "cuda"
    eng = adjoint.AdjointEngine(fs, am, mode="fp64")
    pr = adjoint.Projection(s.pix_off, s.loglam, dev)
    m = torch.from_numpy(adjoint.model_xi()).to(dev)
    delta0 = torch.from_numpy(s.delta).to(dev)
    w0 = torch.from_numpy(s.we).to(dev)

    def chi2(delta, w):
        dw = pr(delta, w)
        W, DW = adjoint.pair_sums(dw, w, eng)
        return adjoint.chi2(W, DW, m, 1e-4)

    d = delta0.clone().requires_grad_(True)
    w = w0.clone().requires_grad_(True)
    chi2(d, w).backward()
    rng = np.random.default_rng(5)
    ks = rng.choice(fs.n_pix, 12, replace=False)
    with torch.no_grad():
        for k in ks:
            for which, base, grad, h in (("delta", delta0, d.grad, 1e-3), ("w", w0, w.grad, 1e-3)):
                def f(step):
                    x = base.clone()
                    x[k] += step
                    return float(chi2(x, w0) if which == "delta" else chi2(delta0, x))
                d1 = (f(h) - f(-h)) / (2 * h)
                d2 = (f(h / 2) - f(-h / 2)) / h
                fdv = (4 * d2 - d1) / 3
                g = float(grad[k])
                assert abs(g - fdv) <= 1e-5 * abs(fdv) + 1e-9 * float(chi2(delta0, w0)), (which, k, g, fdv)


@pytest.mark.gpu
def test_projection_device_matches_host():
    """The device projection (HIP segment sums) against torch's own segment sums on
    the host: values and gradients w.r.t. delta and w."""
    import torch
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 60, seed=8))
    out = {}
    for dev in ("cpu", "cuda"):
        pr = adjoint.Projection(s.pix_off, s.loglam, dev)
        delta = torch.from_numpy(s.delta).to(dev).requires_grad_(True)
        w = torch.from_numpy(s.we).to(dev).requires_grad_(True)
        dw = pr(delta, w)
        (dw * torch.cos(torch.arange(len(dw), device=dev, dtype=torch.float64))).sum().backward()
        out[dev] = [t.detach().cpu().numpy() for t in (dw, delta.grad, w.grad)]
    for a, b in zip(out["cpu"], out["cuda"]):
        np.testing.assert_allclose(b, a, rtol=1e-11, atol=1e-13 * np.abs(a).max())


def test_weight_model_gradient_cpu():
    """The node interpolation's hand-written backward (per-node sums over pixels
    sorted by interval) against torch autograd of the same expression."""
    import torch
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 40, seed=6))
    wm = adjoint.WeightModel(s.z, s.sigma2_pip, s.config.z_nodes, "cpu")
    g = torch.cos(torch.arange(len(s.z), dtype=torch.float64))
    v = torch.linspace(0.03, 0.08, 10, dtype=torch.float64).requires_grad_(True)
    eta = torch.tensor(1.1, dtype=torch.float64, requires_grad=True)
    (wm(eta, v) * g).sum().backward()
    v2 = v.detach().clone().requires_grad_(True)
    e2 = eta.detach().clone().requires_grad_(True)
    s2 = (1.0 - wm.f) * v2[wm.k0] + wm.f * v2[wm.k0 + 1]
    (g / (e2 * wm.s2pip + s2)).sum().backward()
    np.testing.assert_allclose(v.grad.numpy(), v2.grad.numpy(), rtol=1e-13, atol=1e-14 * float(v2.grad.abs().max()))
    assert abs(float(eta.grad - e2.grad)) <= 1e-13 * abs(float(e2.grad))
