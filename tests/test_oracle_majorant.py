"""Pins of the majorant oracle (oracle/majorant.py) on the CPU: BF16 rounding against torch's
conversion, the float64 network against torch, the backward pass against torch autograd and finite
differences, Adam against torch.optim.Adam, and the training-order shuffle."""
import numpy as np
import pytest

from oracle import majorant as M


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed).astype(np.float64)
    p += scale*0.1*rng.standard_normal(p.shape)
    return p.astype(np.float32)


def rand_data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, M.F)).astype(np.float32)*np.float32(2.0) + np.float32(0.5)
    lt = (0.3*X[:, 0] - 0.2*X[:, 3] + 0.1*rng.standard_normal(n)).astype(np.float32)
    norm = np.concatenate([np.full(M.F, 0.5), np.full(M.F, 2.0)]).astype(np.float32)
    return X, lt, norm


def torch_net(p, norm, X):
    import torch
    q = {k: torch.tensor(np.asarray(v, np.float64), dtype=torch.float64) for k, v in M.unpack(p).items()}
    x = (torch.tensor(X, dtype=torch.float64) - torch.tensor(norm[:M.F], dtype=torch.float64)) \
        / torch.tensor(norm[M.F:], dtype=torch.float64)
    h1 = torch.nn.functional.silu(x @ q["W1"] + q["b1"])
    h2 = torch.nn.functional.silu(h1 @ q["W2"] + q["b2"])
    z = h2 @ q["w3"] + q["b3"]
    return torch.nn.functional.softplus(z)


def test_bf16_matches_torch_round_to_nearest_even():
    import torch
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32)*np.float32(100),
                        np.array([1.0, -1.0, 0.0, 1.0 + 2**-8, 1.0 + 3*2**-8, 1.0 + 2**-9, 3.0e38, -7.5e-39],
                                 np.float32)])
    # exact ties: halfway between two BF16 values, both even and odd lower neighbours
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32) + np.float32(0)
    x = np.concatenate([x, ties, -ties])
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf16(x), ref)
    assert M.bf16(np.float32(1.0 + 2**-8)) == np.float32(1.0)                 # tie to even
    assert M.bf16(np.float32(1.0 + 3*2**-8)) == np.float32(1.0 + 2**-6)       # tie to even, upward


def test_exact_forward_matches_torch():
    p = rand_params(3)
    X, lt, norm = rand_data(4, 257)
    y = M.predict(p, norm, X, exact=True)
    np.testing.assert_allclose(y, torch_net(p, norm, X).numpy(), rtol=1e-13, atol=0)


def test_exact_gradient_matches_autograd_and_finite_differences():
    import torch
    p = rand_params(5)
    X, lt, norm = rand_data(6, 96)
    g, loss = M.loss_grad(p, norm, X, lt, exact=True)
    # torch autograd of the same float64 loss
    pt = torch.tensor(np.asarray(p, np.float64), requires_grad=True)
    q = {"W1": pt[M.P_W1:M.P_B1].view(M.F, M.H), "b1": pt[M.P_B1:M.P_W2], "W2": pt[M.P_W2:M.P_B2].view(M.H, M.H),
         "b2": pt[M.P_B2:M.P_W3], "w3": pt[M.P_W3:M.P_B3], "b3": pt[M.P_B3]}
    x = (torch.tensor(X, dtype=torch.float64) - torch.tensor(norm[:M.F], dtype=torch.float64)) \
        / torch.tensor(norm[M.F:], dtype=torch.float64)
    h1 = torch.nn.functional.silu(x @ q["W1"] + q["b1"])
    h2 = torch.nn.functional.silu(h1 @ q["W2"] + q["b2"])
    y = torch.nn.functional.softplus(h2 @ q["w3"] + q["b3"])
    L = ((torch.log(y) - torch.tensor(lt, dtype=torch.float64))**2).mean()
    L.backward()
    assert abs(L.item() - loss) <= 1e-12*abs(loss)
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-9, atol=1e-13)
    # central differences on a spread of parameters of every block
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(np.arange(lo, hi), 6, replace=False) for lo, hi in
                          ((M.P_W1, M.P_B1), (M.P_B1, M.P_W2), (M.P_W2, M.P_B2), (M.P_B2, M.P_W3),
                           (M.P_W3, M.P_B3))] + [[M.P_B3]])
    p64 = np.asarray(p, np.float64)
    for k in idx:
        h = 1e-6
        a, b = p64.copy(), p64.copy()
        a[k] += h
        b[k] -= h
        fd = (M.loss_value(a, norm, X, lt) - M.loss_value(b, norm, X, lt))/(2*h)
        assert abs(fd - g[k]) <= 1e-6*max(1.0, abs(g[k])), (k, fd, g[k])


def test_rounded_form_stays_near_exact_form():
    p = rand_params(8)
    X, lt, norm = rand_data(9, 512)
    ye = M.predict(p, norm, X, exact=True)
    yr = M.predict(p, norm, X, exact=False)
    rel = np.abs(yr - ye)/ye
    assert np.median(rel) < 5e-3 and rel.max() < 5e-2
    ge, _ = M.loss_grad(p, norm, X, lt, exact=True)
    gr, _ = M.loss_grad(p, norm, X, lt, exact=False)
    assert np.linalg.norm(gr - ge) < 0.05*np.linalg.norm(ge)


def test_adam_matches_torch():
    import torch
    rng = np.random.default_rng(10)
    p0 = rng.standard_normal(50).astype(np.float32)
    grads = rng.standard_normal((5, 50)).astype(np.float32)
    pt = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50, np.float32), np.zeros(50, np.float32)
    for t in range(5):
        opt.zero_grad()
        pt.grad = torch.tensor(grads[t])
        opt.step()
        p, m, v = M.adam_step(p, m, v, grads[t], t + 1)
    np.testing.assert_allclose(p, pt.detach().numpy(), rtol=0, atol=2e-7)


def test_shuffle_is_a_permutation_and_advances_the_state():
    perm, s = M.shuffle_order(1000, 42)
    assert sorted(perm.tolist()) == list(range(1000))
    perm2, s2 = M.shuffle_order(1000, s)
    assert s2 != s and not np.array_equal(perm, perm2)
    assert not np.array_equal(perm, np.arange(1000))


def test_init_params_ranges():
    p = M.init_params(0)
    q = M.unpack(p)
    assert np.abs(q["W1"]).max() <= 1/np.sqrt(8) and np.abs(q["W2"]).max() <= 1/8
    assert abs(float(q["b3"]) - np.log(np.e - 1)) < 1e-6      # softplus(b3) = 1
    assert q["W1"].std() > 0.15           # uniform(+-1/sqrt(8)) has standard deviation 0.204


@pytest.mark.parametrize("seed", [0, 11])
def test_training_reduces_loss(seed):
    X, lt, norm = rand_data(12 + seed, 2048)
    p0 = M.init_params(seed)
    p, losses = M.train(p0, norm, X, lt, epochs=3, batch=256, lr=3e-3, seed=seed)
    assert losses[-1] < losses[0]
