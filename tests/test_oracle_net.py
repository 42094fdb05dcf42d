"""Pins the network oracle (oracle/netref.py), CPU only:
* forward(): every layer against torch's conv2d with the stated weight
  layout, and one pixel against a nested loop over the 3 x 3 window;
* bf16(): torch's bfloat16 rounding (nearest even, ties both ways);
* loss_grad(): central finite differences in FP64;
* logq(): numerical integration over the sphere;
* coons() / patch(): the interpolation reproduces bilinear fields and the
  ring, and the block's own spins never enter the network input;
* seed_site(): the window, the score A(kappa) mu.m and the tie rule;
* block_charge(): the C oracle's lattice charge of the same textures;
* swap_proposal(): an involution (a rotation by pi per site) that maps
  texture A onto B, B a skyrmion more than A by the C oracle's charge, and
  both textures stationary points of the C oracle's energy E_s."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def netref():
    from oracle import netref
    return netref


def random_case(netref, rng, n=2):
    p = netref.init_params(rng)
    patches = np.zeros((n, 24, 24, 4))
    s = rng.normal(size=(n, 24, 24, 3))
    s /= np.linalg.norm(s, axis=-1, keepdims=True)
    patches[..., :3] = s
    patches[..., 3] = 1.0
    patches[:, 4:20, 4:20, :] = 0.0
    t = rng.normal(size=(n, 256, 3))
    t /= np.linalg.norm(t, axis=-1, keepdims=True)
    return p, patches, t


def test_gradient_matches_finite_differences(netref):
    rng = np.random.default_rng(0)
    p, patches, t = random_case(netref, rng)
    loss, g = netref.loss_grad(p, patches, t)
    assert g.size == netref.n_params() == p.size
    # probe coordinates in every layer's weights and biases
    offs, idx = 0, []
    for l, c in enumerate(netref.CIN):
        K = 9 * c
        for _ in range(6):
            k = int(rng.integers(K))
            if l == 0:
                k = (k // c) * c + int(rng.integers(4))
            n = int(rng.integers(netref.NOUT[l]))
            idx.append(offs + k * 32 + n)
        idx.append(offs + K * 32 + int(rng.integers(netref.NOUT[l])))
        offs += K * 32 + 32
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (netref.loss_grad(pp, patches, t)[0] - netref.loss_grad(pm, patches, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_vmf_density_normalised(netref):
    # integral over S^2 of q = 1, and the mean cosine = coth k - 1/k
    for kappa in (0.05, 0.7, 5.0, 40.0):
        th = np.linspace(0, np.pi, 200001)
        mu = np.array([0.0, 0.0, 1.0])
        x = np.stack([np.sin(th), np.zeros_like(th), np.cos(th)], -1)
        lq = netref.logq(mu[None, :], np.full(th.shape, kappa), x)
        f = np.exp(lq) * 2 * np.pi * np.sin(th)
        assert abs(np.trapezoid(f, th) - 1) < 1e-6
        mc = np.trapezoid(f * np.cos(th), th)
        assert abs(mc - (1 / math.tanh(kappa) - 1 / kappa)) < 1e-6


def test_forward_is_a_textbook_convolution(netref):
    # each layer of forward() against torch's conv2d (NCHW, weight
    # [Cout][Cin][dy][dx] = W[(dy * 3 + dx) * Cin + ci][Cout]) + bias + ReLU,
    # with dy along the patch's first axis; and one output pixel of the first
    # layer against a nested loop over the 3 x 3 window
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(11)
    p, patches, _ = random_case(netref, rng, n=3)
    p = p + rng.normal(scale=0.05, size=p.shape)    # also the unused channels of layer 0
    xs, out = netref.forward(p, patches)
    X = np.zeros((3, 24, 24, 16))
    X[..., :4] = patches
    assert np.array_equal(xs[0], X)
    layers = netref.unpack(p)
    cur = torch.from_numpy(X).permute(0, 3, 1, 2)
    for l, (W, b) in enumerate(layers):
        cin = netref.CIN[l]
        w = torch.from_numpy(W.reshape(3, 3, cin, 32).transpose(3, 2, 0, 1).copy())
        y = F.conv2d(cur, w, torch.from_numpy(b))
        if l < 3:
            y = torch.relu(y)
            assert np.allclose(xs[l + 1], y.permute(0, 2, 3, 1).numpy(), rtol=1e-12, atol=1e-12), l
        cur = y
    assert np.allclose(out, cur.permute(0, 2, 3, 1).numpy(), rtol=1e-12, atol=1e-12)
    W0, b0 = layers[0]
    for (n, u, v, o) in ((0, 0, 0, 0), (1, 21, 3, 7), (2, 10, 17, 31)):
        acc = b0[o]
        for dy in range(3):
            for dx in range(3):
                for ci in range(16):
                    acc += X[n, u + dy, v + dx, ci] * W0[(dy * 3 + dx) * 16 + ci, o]
        assert abs(xs[1][n, u, v, o] - max(acc, 0.0)) < 1e-12


def test_bf16_rounding_matches_torch(netref):
    # round to nearest even, including exact ties in both directions,
    # subnormals, signed zeros and overflow to infinity
    import torch
    rng = np.random.default_rng(5)
    vals = [rng.normal(size=4000).astype(np.float32), (rng.normal(size=1000) * 1e-39).astype(np.float32),
            np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 3.3895314e38, -3.3895314e38], dtype=np.float32)]
    # exact ties: bf16 value + half an ulp, with even and odd kept mantissas
    base = rng.normal(size=500).astype(np.float32).view(np.uint32) & np.uint32(0xFFFF0000)
    vals.append((base | np.uint32(0x8000)).view(np.float32))
    vals.append(((base | np.uint32(0x10000)) | np.uint32(0x8000)).view(np.float32))
    x = np.concatenate(vals)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = netref.bf16(x)
    assert np.array_equal(got, want)
    assert np.array_equal(np.signbit(got), np.signbit(want))
    ties = np.concatenate(vals[-2:])
    assert np.any(netref.bf16(ties) > ties.astype(np.float64)) and np.any(netref.bf16(ties) < ties.astype(np.float64))


def _swap_case(netref, orc, L=32, bx=5, by=9, D=0.3, h=0.06, iters=80, seed=3):
    from tests.fixtures import box_locs, neel2d
    rng = np.random.default_rng(seed)
    s = rng.normal(size=(L * L, 3))
    s[:, 2] += 2.0                       # a ferromagnetic background, as at low T
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    d, Jv, Dv = neel2d(D)
    J6, D6 = np.zeros(6), np.zeros((6, 3))
    slot = {(-1, 0): 2, (1, 0): 3, (0, -1): 4, (0, 1): 5}
    for k in range(4):
        g = slot[(int(d[k][0]), int(d[k][1]))]
        J6[g], D6[g] = Jv[k], Dv[k]
    out = np.zeros((16, 16, 4))          # network outputs along the background, kappa pre-activation 1
    out[..., :3] = rng.normal(scale=0.2, size=(16, 16, 3))
    out[..., 2] += 1.0
    out[..., 3] = 1.0
    out[6, 9, :3] = [0.1, 0.0, -1.0]     # the site where a core is predicted
    f = orc.Field(box_locs(2, L), 2, L, 1, d, Jv, Dv)
    return s, J6, D6, out, f, (L, bx, by, D, h, iters)


def test_coons_reproduces_bilinear_fields_and_the_ring(netref):
    rng = np.random.default_rng(1)
    a, b, c, d = rng.normal(size=(4, 3))

    def f(u, v):
        return a + b * u + c * v + d * u * v

    ring = np.array([[f(i - 1, j - 1) for j in range(18)] for i in range(18)])
    want = np.array([[f(u, v) for v in range(16)] for u in range(16)])
    assert np.abs(netref.coons(ring, normalise=False) - want).max() < 1e-12
    n = netref.coons(ring)
    assert np.allclose(n, want / np.linalg.norm(want, axis=-1, keepdims=True), atol=1e-12)
    # a uniform ring gives the uniform block; a vanishing interpolation gives +z
    e = rng.normal(size=3)
    e /= np.linalg.norm(e)
    assert np.allclose(netref.coons(np.broadcast_to(e, (18, 18, 3))), e, atol=1e-12)
    assert np.array_equal(netref.coons(np.zeros((18, 18, 3))), np.broadcast_to([0.0, 0.0, 1.0], (16, 16, 3)))


def test_patch_holds_the_halo_and_the_interpolated_ring_only(netref):
    rng = np.random.default_rng(2)
    L, bx, by = 40, 30, 7                # wraps in x
    s = rng.normal(size=(L * L, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    P = netref.patch(s, L, bx, by)
    g = s.reshape(L, L, 3)
    for pu in (0, 3, 12, 23):
        for pv in (0, 3, 20, 23):
            if 4 <= pu < 20 and 4 <= pv < 20:
                continue
            assert np.array_equal(P[pu, pv, :3], g[(bx - 4 + pu) % L, (by - 4 + pv) % L]) and P[pu, pv, 3] == 1.0
    ring = g[np.ix_(np.arange(bx - 1, bx + 17) % L, np.arange(by - 1, by + 17) % L)]
    assert np.array_equal(P[4:20, 4:20, :3], netref.coons(ring)) and not P[4:20, 4:20, 3].any()
    # the block's own spins do not change the input
    s2 = s.copy()
    blk = (np.arange(bx, bx + 16)[:, None] % L) * L + (np.arange(by, by + 16)[None, :] % L)
    s2[blk.reshape(-1)] = -s2[blk.reshape(-1)]
    assert np.array_equal(netref.patch(s2, L, bx, by), P)


def test_seed_site_window_score_and_ties(netref):
    m = np.array([0.0, 0.6, 0.8])
    out = np.zeros((16, 16, 4))
    out[..., :3] = m
    # constant outputs: every window site ties; the first in row order wins
    assert netref.seed_site(out, m)[0] == (4, 4)
    # the score is the vMF mean alignment A(kappa) mu.m: a low-kappa site
    # along m beats a high-kappa one; a site against m beats both; sites
    # outside [4, 12) are never chosen
    out[..., 3] = 5.0
    out[9, 5, 3] = -3.0
    assert netref.seed_site(out, m)[0] == (9, 5)
    out[7, 10, :3] = -m
    assert netref.seed_site(out, m)[0] == (7, 10)
    out[2, 2, :3] = -m
    out[2, 2, 3] = 9.0
    (cu, cv), score = netref.seed_site(out, m)
    assert (cu, cv) == (7, 10) and np.isinf(score[2, 2])
    # a site's score is the mean of s.m under its vMF law, by quadrature
    kap = np.log1p(np.exp(5.0)) + 0.05
    th = np.linspace(0, np.pi, 200001)
    x = np.stack([np.sin(th), np.zeros_like(th), np.cos(th)], -1)
    q = np.exp(netref.logq(np.array([[0.0, 0.0, 1.0]]), np.full(th.shape, kap), x)) * 2 * np.pi * np.sin(th)
    assert abs(score[4, 4] - np.trapezoid(q * np.cos(th), th)) < 1e-6


def test_block_charge_is_the_lattice_charge_of_an_isolated_skyrmion(netref, orc):
    # a uniform +z lattice with one relaxed-template skyrmion inside a block:
    # the block charge (17 x 17 plaquettes over block and ring) equals the C
    # oracle's periodic plaquette charge of the whole lattice, and the
    # uniform texture has charge 0
    s, J6, D6, out, f, (L, bx, by, D, h, iters) = _swap_case(netref, orc)
    s = np.zeros_like(s)
    s[:, 2] = 1.0
    _, A, B, info = netref.swap_proposal(s, L, bx, by, out, J6, D6, h, iters, 5.0)
    blk = (np.arange(bx, bx + 16)[:, None] % L) * L + (np.arange(by, by + 16)[None, :] % L)
    ring = s.reshape(L, L, 3)[np.ix_(np.arange(bx - 1, bx + 17) % L, np.arange(by - 1, by + 17) % L)].copy()
    for T in (A, B):
        st = s.copy()
        st[blk] = T
        R = ring.copy()
        R[1:17, 1:17] = T
        assert abs(netref.block_charge(R) - f.tcharge(st, wrap=True)[0]) < 1e-9
    assert abs(netref.block_charge(np.broadcast_to([0.0, 0.0, 1.0], (18, 18, 3)))) < 1e-15
    assert abs(abs(info["dq"]) - 1.0) < 1e-9


def test_swap_proposal_is_an_involution_mapping_a_to_b(netref, orc):
    s, J6, D6, out, f, (L, bx, by, D, h, iters) = _swap_case(netref, orc)
    xp, A, B, info = netref.swap_proposal(s, L, bx, by, out, J6, D6, h, iters, 5.0)
    assert info["seed"] == (6, 9)                               # the site predicted against the background
    block = (np.arange(bx, bx + 16)[:, None] % L) * L + (np.arange(by, by + 16)[None, :] % L)
    x = s[block]
    # a rotation by pi about a unit axis: unit spins stay unit, and applying
    # the move twice returns the block
    assert np.allclose(np.linalg.norm(xp, axis=-1), 1.0, atol=1e-12)
    s2 = s.copy()
    s2[block] = xp
    xpp, A2, B2, _ = netref.swap_proposal(s2, L, bx, by, out, J6, D6, h, iters, 5.0)
    assert np.array_equal(A2, A) and np.array_equal(B2, B)     # textures depend on the surroundings only
    assert np.allclose(xpp, x, atol=1e-12)
    # the move maps texture A onto texture B and back
    s3 = s.copy()
    s3[block] = A
    assert np.allclose(netref.swap_proposal(s3, L, bx, by, out, J6, D6, h, iters, 5.0)[0], B, atol=1e-12)
    s3[block] = B
    assert np.allclose(netref.swap_proposal(s3, L, bx, by, out, J6, D6, h, iters, 5.0)[0], A, atol=1e-12)
    # the two textures differ by one skyrmion (the lattice's plaquette charge
    # with the block set to B exceeds the one with A by one, by the oracle's
    # calc_TC with periodic wrap) and both keep unit length
    s3[block] = A
    qa = f.tcharge(s3, wrap=True)[0]
    s3[block] = B
    qb = f.tcharge(s3, wrap=True)[0]
    assert abs(abs(qb - qa) - 1.0) < 1e-9, (qa, qb)
    assert abs(info["dq"] - (qb - qa)) < 1e-9
    assert np.allclose(np.linalg.norm(A, axis=-1), 1.0) and np.allclose(np.linalg.norm(B, axis=-1), 1.0)
    # unrelaxed, B is the Coons start with the template's core (-m) on the seed site
    _, A0, B0, _ = netref.swap_proposal(s, L, bx, by, out, J6, D6, h, 0, 5.0)
    ring = s.reshape(L, L, 3)[np.ix_(np.arange(bx - 1, bx + 17) % L, np.arange(by - 1, by + 17) % L)]
    m = ring[0, 1:17].sum(0) + ring[17, 1:17].sum(0) + ring[1:17, 0].sum(0) + ring[1:17, 17].sum(0)
    m /= np.linalg.norm(m)
    assert np.allclose(A0, netref.coons(ring), atol=1e-15) and np.allclose(B0[6, 9], -m, atol=1e-12)
    far = np.hypot(*np.meshgrid(np.arange(16) - 6, np.arange(16) - 9, indexing="ij")) >= 5.0
    assert np.array_equal(B0[far], A0[far])


def test_swap_textures_are_relaxed_to_the_oracle_energy_minimum(netref, orc):
    # with the block's one-site ring fixed, A and B are relaxed to a stationary
    # point of the sampler's energy E_s: every block spin is parallel to its
    # local field -grad E_s, computed here by central differences of the C
    # oracle's energy (independent of the relaxation's own field)
    s, J6, D6, out, f, (L, bx, by, D, h, iters) = _swap_case(netref, orc, iters=400)
    _, A, B, _ = netref.swap_proposal(s, L, bx, by, out, J6, D6, h, iters, 5.0)
    block = (np.arange(bx, bx + 16)[:, None] % L) * L + (np.arange(by, by + 16)[None, :] % L)
    H = [0.0, 0.0, h]
    for T in (A, B):
        st = s.copy()
        st[block] = T
        worst = 0.0
        for u in range(0, 16, 3):
            for v in range(0, 16, 2):
                i = block[u, v]
                F = np.zeros(3)
                for a in range(3):
                    sp, sm = st.copy(), st.copy()
                    sp[i, a] += 1e-5
                    sm[i, a] -= 1e-5
                    F[a] = -(f.energy(sp, H, use_d=True, convention=1) - f.energy(sm, H, use_d=True, convention=1)) / 2e-5
                worst = max(worst, np.linalg.norm(np.cross(T[u, v], F)) / np.linalg.norm(F))
        assert worst < 1e-4, worst
