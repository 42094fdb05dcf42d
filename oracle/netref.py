"""FP64 NumPy restatement of the learned block-proposal network -- TEST
INFRASTRUCTURE ONLY (the checker of the device's forward and backward
kernels; never imported by the nanomagmc_amd package).

Input (patch()): the 24 x 24 window around a 16 x 16 block, channels
(spin x, y, z, halo flag); the block itself carries the Coons interpolation
of its one-site ring (flag 0), so every block site sees its surroundings
through four 3 x 3 layers.

Network: four valid 3x3 convolutions on a 24 x 24 patch, input channels
(spin x, y, z, halo flag) padded to 16, then 32, 32, 32 channels with ReLU,
last layer 4 outputs per block site: mu (unnormalised) and the kappa
pre-activation.  Per site a von Mises-Fisher law on S^2:
    kappa = softplus(z) + 0.05,  mu = m / (|m| + 1e-6),
    log q(x) = log kappa - log(4 pi) - log sinh(kappa) + kappa mu.x.
Loss = mean over sites of -log q(target).

Parameter vector order (the device's): per layer W as [9 * Cin][32]
(row = (dy * 3 + dx) * Cin + ci) then b[32].  With bf16=True the arrays the
device keeps in bf16 (input patch, weights, hidden activations, output and
hidden gradients) are rounded to bf16 (round to nearest even) at the same
points, so the comparison isolates accumulation-order differences.
"""
import numpy as np

CIN = [16, 32, 32, 32]
NOUT = [32, 32, 32, 4]
HIN = [24, 22, 20, 18]


def n_params():
    return sum(9 * c * 32 + 32 for c in CIN)


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, off = [], 0
    for c in CIN:
        K = 9 * c
        W = p[off:off + K * 32].reshape(K, 32)
        off += K * 32
        b = p[off:off + 32]
        off += 32
        out.append((W, b))
    return out


def im2col(X):
    """(n, H, H, C) -> (n, H-2, H-2, 9C) with k = (dy*3+dx)*C + c."""
    n, H, _, C = X.shape
    Ho = H - 2
    cols = np.empty((n, Ho, Ho, 9 * C))
    for dy in range(3):
        for dx in range(3):
            t = dy * 3 + dx
            cols[..., t * C:(t + 1) * C] = X[:, dy:dy + Ho, dx:dx + Ho, :]
    return cols


def col2im(dcols, C):
    """adjoint of im2col: (n, Ho, Ho, 9C) -> (n, Ho+2, Ho+2, C)."""
    n, Ho, _, _ = dcols.shape
    dX = np.zeros((n, Ho + 2, Ho + 2, C))
    for dy in range(3):
        for dx in range(3):
            t = dy * 3 + dx
            dX[:, dy:dy + Ho, dx:dx + Ho, :] += dcols[..., t * C:(t + 1) * C]
    return dX


def forward(params, patches, use_bf16=False):
    """patches (n, 24, 24, 4) -> per-layer inputs and the output (n,16,16,32)."""
    rnd = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    n = patches.shape[0]
    X = np.zeros((n, 24, 24, 16))
    X[..., :4] = patches
    X = rnd(X)
    layers = unpack(params)
    xs = []
    for l, (W, b) in enumerate(layers):
        xs.append(X)
        Y = im2col(X) @ rnd(W) + b
        if l < 3:
            Y = rnd(np.maximum(Y, 0.0))
        X = Y
    return xs, X


def head(O):
    m = O[..., :3]
    r = np.sqrt((m * m).sum(-1, keepdims=True)) + 1e-6
    mu = m / r
    z = O[..., 3]
    kappa = np.where(z > 20, z, np.log1p(np.exp(np.minimum(z, 20)))) + 0.05
    return mu, kappa, r


def logq(mu, kappa, x):
    lsinh = kappa + np.log1p(-np.exp(-2 * kappa)) - np.log(2.0)
    return np.log(kappa) - np.log(4 * np.pi) - lsinh + kappa * (mu * x).sum(-1)


def loss_grad(params, patches, targets, use_bf16=False):
    """Mean loss and gradient (device parameter order)."""
    rnd = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    n = patches.shape[0]
    xs, O = forward(params, patches, use_bf16)
    x = targets.reshape(n, 16, 16, 3)
    mu, kappa, r = head(O)
    L = -logq(mu, kappa, x)
    loss = L.mean()
    scale = 1.0 / L.size
    # dL/dkappa, dL/dmu_raw, dL/dz
    e2 = np.exp(-2 * kappa)
    coth = (1 + e2) / (1 - e2)
    mux = (mu * x).sum(-1)
    dk = -(1 / kappa - coth + mux)
    g = -kappa[..., None] * x
    proj = (mu * g).sum(-1, keepdims=True)
    dm = (g - mu * proj) / r
    z = O[..., 3]
    sig = 1 / (1 + np.exp(-z))
    dO = np.zeros_like(O)
    dO[..., :3] = scale * dm
    dO[..., 3] = scale * dk * sig
    dY = rnd(dO)
    layers = unpack(params)
    grads = [None] * 4
    for l in range(3, -1, -1):
        W, b = layers[l]
        cols = im2col(xs[l])
        gW = cols.reshape(-1, cols.shape[-1]).T @ dY.reshape(-1, 32)
        gb = dY.reshape(-1, 32).sum(0)
        grads[l] = (gW, gb)
        if l > 0:
            dX = col2im(dY @ rnd(W).T, CIN[l])
            dX = np.where(xs[l] > 0, dX, 0.0)   # through the ReLU of the layer below
            dY = rnd(dX)
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, flat


def init_params(rng):
    p = []
    for l, c in enumerate(CIN):
        fan = 9 * (4 if l == 0 else 32)
        lim = np.sqrt(6.0 / fan) * (0.1 if l == 3 else 1.0)
        W = np.zeros((9 * c, 32))
        for k in range(9 * c):
            if l == 0 and k % c >= 4:
                continue
            W[k, :NOUT[l]] = rng.uniform(-lim, lim, NOUT[l])
        b = np.zeros(32)
        b[:NOUT[l]] = rng.uniform(-0.1, 0.1, NOUT[l])
        p += [W.reshape(-1), b]
    return np.concatenate(p)


def coons(ring, normalise=True):
    """Transfinite (Coons) interpolation of a 16 x 16 block from its one-site
    ring: ring (18, 18, 3) indexed [u + 1, v + 1] for u, v in [-1, 16]
    (corners included).  With s = (u + 1) / 17, t = (v + 1) / 17,
        P = (1-s) R(-1,v) + s R(16,v) + (1-t) R(u,-1) + t R(u,16)
            - [(1-s)(1-t) R(-1,-1) + (1-s) t R(-1,16) + s (1-t) R(16,-1) + s t R(16,16)],
    then normalised per site (|P|^2 < 1e-12 gives +z).  Returns (16, 16, 3)."""
    R = np.asarray(ring, dtype=np.float64)
    u = np.arange(16)[:, None, None]
    v = np.arange(16)[None, :, None]
    s = (u + 1) / 17.0
    t = (v + 1) / 17.0
    P = ((1 - s) * R[0, 1:17][None] + s * R[17, 1:17][None] + (1 - t) * R[1:17, 0][:, None]
         + t * R[1:17, 17][:, None]
         - ((1 - s) * (1 - t) * R[0, 0] + (1 - s) * t * R[0, 17] + s * (1 - t) * R[17, 0] + s * t * R[17, 17]))
    if not normalise:
        return P
    n2 = (P * P).sum(-1, keepdims=True)
    out = P / np.sqrt(np.maximum(n2, 1e-300))
    out[n2[..., 0] < 1e-12] = [0.0, 0.0, 1.0]
    return out


def patch(spins, L, bx, by):
    """The network input of the 16 x 16 block at (bx, by) of a periodic L x L
    lattice (spins (L*L, 3), site index x * L + y): 24 x 24 x 4, the 4-site
    halo's spins with flag 1, and on the block the Coons interpolation of its
    one-site ring with flag 0.  The block's own spins never enter."""
    s = np.asarray(spins, dtype=np.float64).reshape(L, L, 3)
    P = np.zeros((24, 24, 4))
    P[..., :3] = s[np.ix_(np.arange(bx - 4, bx + 20) % L, np.arange(by - 4, by + 20) % L)]
    P[..., 3] = 1.0
    P[4:20, 4:20, :3] = coons(P[3:21, 3:21, :3])
    P[4:20, 4:20, 3] = 0.0
    return P


SEED_LO, SEED_HI = 4, 12     # seed sites u, v in [4, 12)


def seed_site(out, m):
    """Where the swap seeds its skyrmion: among block sites u, v in [4, 12),
    the one whose predicted mean alignment with the ring direction m,
    E_q[s . m] = A(kappa) (mu . m) with A(k) = coth k - 1/k, is lowest (the
    first in u-major order on ties).  out (16, 16, >=4): the network's raw
    outputs.  Returns ((cu, cv), score (16, 16) with +inf outside the window)."""
    O = np.asarray(out, dtype=np.float64)
    m3 = O[..., :3]
    n = np.linalg.norm(m3, axis=-1, keepdims=True)
    mu = np.where(n > 1e-15, m3 / np.maximum(n, 1e-300), [0.0, 0.0, 1.0])
    z = O[..., 3]
    kappa = np.where(z > 20, z, np.log1p(np.exp(np.minimum(z, 20)))) + 0.05
    A = 1.0 / np.tanh(kappa) - 1.0 / kappa
    score = np.full((16, 16), np.inf)
    w = slice(SEED_LO, SEED_HI)
    score[w, w] = (A * (mu @ np.asarray(m, dtype=np.float64)))[w, w]
    k = int(np.argmin(score))
    return (k // 16, k % 16), score


def solid_angle(a, b, c):
    """Signed solid angle of the spherical triangle (a, b, c):
    2 atan2(a . (b x c), 1 + a.b + b.c + c.a)."""
    return 2.0 * np.arctan2((a * np.cross(b, c)).sum(-1), 1.0 + (a * b).sum(-1) + (b * c).sum(-1) + (c * a).sum(-1))


def block_charge(S):
    """Sum of the solid angles / 4 pi over the 17 x 17 plaquettes of an
    (18, 18, 3) block-plus-ring texture, each plaquette split into the
    triangles (p00, p01, p11) and (p00, p11, p10), p_ab at (x + a, y + b):
    the orientation of the reference's lattice charge.  Two textures with the
    same ring differ by an integer."""
    S = np.asarray(S, dtype=np.float64)
    p00, p10, p11, p01 = S[:-1, :-1], S[1:, :-1], S[1:, 1:], S[:-1, 1:]
    return float((solid_angle(p00, p01, p11) + solid_angle(p00, p11, p10)).sum() / (4 * np.pi))


def swap_proposal(spins, L, bx, by, out, J, D, h, iters, core_radius, seed=None):
    """FP64 restatement of the learned texture swap of one 16 x 16 block of a
    periodic L x L lattice (spins (L*L, 3), site index x * L + y).

    J[g], D[g] (3-vector) per stencil slot g: 2 = (-1, 0), 3 = (+1, 0),
    4 = (0, -1), 5 = (0, +1) (x = first coordinate).  out (16, 16, >=4): the
    network's raw outputs on the block, which choose the seed site
    (seed_site); `seed` = (cu, cv) overrides the choice.  Texture A starts
    from the Coons interpolation of the block's one-site ring, texture B from
    the same with a Neel skyrmion of radius core_radius centred on the seed
    site, its core against the ring's mean direction m and its chirality
    -sign(D_{+x} . y).  Relaxation: `iters` pairs of checkerboard half-steps
    s <- normalise(s + 1.85 (F - s)), F the unit local field, ring fixed.
    Returns (x', A, B, info): the rotated block, the two relaxed textures,
    each (16, 16, 3), and info = {"seed": (cu, cv), "dq": Q(B) - Q(A)} with
    Q the block charge over the block and its ring."""
    s = np.asarray(spins, dtype=np.float64).reshape(L, L, 3)
    xs = np.arange(bx - 1, bx + 17) % L
    ys = np.arange(by - 1, by + 17) % L
    ring = s[np.ix_(xs, ys)].copy()                      # (18, 18, 3); index (u + 1, v + 1)
    edge = np.zeros((18, 18), dtype=bool)
    edge[0, 1:17] = edge[17, 1:17] = edge[1:17, 0] = edge[1:17, 17] = True
    m = ring[edge].sum(axis=0)
    nm = np.linalg.norm(m)
    m = m / nm if nm > 1e-10 else np.array([0.0, 0.0, 1.0])
    if seed is None:
        seed, _ = seed_site(out, m)
    cu, cv = seed
    A = ring.copy()
    A[1:17, 1:17] = coons(ring)
    B = A.copy()
    u, v = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    du, dv = (u - cu).astype(np.float64), (v - cv).astype(np.float64)
    rad = np.sqrt(du * du + dv * dv)
    th = np.pi * (1 - rad / core_radius)
    chir = -1.0 if D[3][1] > 0 else 1.0
    safe = np.where(rad > 0, rad, 1.0)
    cph = np.where(rad > 0, chir * du / safe, chir)
    sph = np.where(rad > 0, chir * dv / safe, 0.0)
    ta, tb, tc = np.sin(th) * cph, np.sin(th) * sph, np.cos(th)
    sgn = 1.0 if m[2] >= 0 else -1.0
    a = -1.0 / (sgn + m[2])
    bb = m[0] * m[1] * a
    e1 = np.array([1 + sgn * m[0] * m[0] * a, sgn * bb, -sgn * m[0]])
    e2 = np.array([bb, sgn + m[1] * m[1] * a, -m[1]])
    tmpl = ta[..., None] * e1 + tb[..., None] * e2 + tc[..., None] * m
    core = rad < core_radius
    B[1:17, 1:17][core] = tmpl[core]
    par = (bx + u + by + v) & 1
    nbrs = [((0, 1), 5), ((0, -1), 4), ((1, 0), 3), ((-1, 0), 2)]

    omega = 1.85

    def half(S, p):
        H = np.zeros((16, 16, 3))
        H[..., 2] = h
        for (a_, b_), g in nbrs:
            n = S[1 + a_:17 + a_, 1 + b_:17 + b_]
            H += J[g] * n + np.cross(n, D[g])
        nH = np.linalg.norm(H, axis=-1, keepdims=True)
        cur = S[1:17, 1:17]
        m = cur + omega * (H / np.maximum(nH, 1e-300) - cur)
        new = m / np.linalg.norm(m, axis=-1, keepdims=True)
        sel = (par == p) & (nH[..., 0] > 1e-30)
        S[1:17, 1:17][sel] = new[sel]

    for _ in range(iters):
        for p in (0, 1):
            half(A, p)
            half(B, p)
    dq = block_charge(B) - block_charge(A)
    Ab, Bb = A[1:17, 1:17], B[1:17, 1:17]
    w = Ab + Bb
    nw = np.linalg.norm(w, axis=-1, keepdims=True)
    anti = nw[..., 0] < 1e-6
    if anti.any():
        e = np.where(np.abs(Ab[..., :1]) < 0.9, [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
        w[anti] = np.cross(Ab, e)[anti]
        nw = np.linalg.norm(w, axis=-1, keepdims=True)
    w = w / nw
    x = s[np.ix_(np.arange(bx, bx + 16) % L, np.arange(by, by + 16) % L)]
    xp = 2 * (w * x).sum(-1, keepdims=True) * w - x
    return xp, Ab, Bb, {"seed": (int(cu), int(cv)), "dq": dq}
