"""Learned rank predictor of libh2w: harvest, training and parameter export.

A 16-64-64-64-1 perceptron (GELU, erf form) maps the geometry features of an ACA
block (computed by the library during assembly, see include/h2w.h) to
log2(1 + rank).  The library evaluates it on the GPU with BF16 parameters and
FP32 accumulation and uses it only to size workspaces and batch blocks by rank;
the adaptive cross still decides every pivot and rank.

    harvest(specs)          assemble point sets on the GPU, return (features, ranks, set ids)
    train(F, R, ...)        fit the perceptron (PyTorch), return BF16 parameters (uint16)
    forward(params, F)      the kernel's arithmetic in numpy (reference for its tests)
    save / load             parameters as .npz (uint16 array + metadata, no pickles)

The shipped model `rank_mlp_3d_rbf_tol10.npz` is trained for 3D, RBF kernel
(choice 1), ACA tolerance 1e-10, n_max 64 -- a model is specific to dimension,
kernel and tolerance.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
F_DIM, H_DIM = 16, 64
N_PARAMS = F_DIM * H_DIM + H_DIM + 2 * (H_DIM * H_DIM + H_DIM) + H_DIM + 1


def default_model_path(dim, choice, tol_pow):
    p = os.path.join(HERE, f"rank_mlp_{dim}d_{'rbf' if choice == 1 else 'ie'}_tol{tol_pow}.npz")
    return p if os.path.exists(p) else None


def points(kind, N, seed, dim=3):
    """U([-1,1]^d), or N(0, 0.2^2) per axis rejection-sampled into [-1,1]^d."""
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.uniform(-1.0, 1.0, (dim, N))
    if kind == "gauss":
        out = np.empty((dim, 0))
        while out.shape[1] < N:
            P = rng.normal(0.0, 0.2, (dim, 2 * N))
            out = np.concatenate([out, P[:, np.all(np.abs(P) <= 1.0, axis=0)]], axis=1)
        return np.ascontiguousarray(out[:, :N])
    raise ValueError(kind)


def harvest(specs, tol_pow=10, n_max=64, choice=1, log=print):
    """specs: iterable of (kind, N, seed).  Returns features [n, 16], ranks [n], set index [n]."""
    import h2weak_amd as h2w
    Fs, Rs, Ss = [], [], []
    for q, (kind, N, seed) in enumerate(specs):
        P = points(kind, N, seed)
        g = h2w.Initialisation(3, N, n_max, 1.0, tol_pow, choice, h2w.NESTED, points=P)
        g.tree_initialisation()
        F, R = g.rank_samples()
        Fs.append(F)
        Rs.append(R)
        Ss.append(np.full(len(R), q, dtype=np.int32))
        log(f"harvest {kind} N={N} seed={seed}: {len(R)} blocks, assembly {g.stats()['assembly_seconds']:.1f} s")
        del g
    return np.concatenate(Fs), np.concatenate(Rs), np.concatenate(Ss)


def to_bf16(a):
    """Round-to-nearest-even FP32 -> BF16 bit patterns (uint16)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(w):
    return (np.asarray(w, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def _unpack(params):
    p = from_bf16(params)
    o = 0

    def take(n, shape):
        nonlocal o
        v = p[o:o + n].reshape(shape)
        o += n
        return v
    W1 = take(F_DIM * H_DIM, (F_DIM, H_DIM))
    b1 = take(H_DIM, (H_DIM,))
    W2 = take(H_DIM * H_DIM, (H_DIM, H_DIM))
    b2 = take(H_DIM, (H_DIM,))
    W3 = take(H_DIM * H_DIM, (H_DIM, H_DIM))
    b3 = take(H_DIM, (H_DIM,))
    w4 = take(H_DIM, (H_DIM,))
    b4 = take(1, (1,))
    return W1, b1, W2, b2, W3, b3, w4, b4


def forward(params, F):
    """log2(1 + rank) from BF16 parameters, FP32 arithmetic (the GPU kernel's, up to summation order)."""
    from scipy.special import erf
    W1, b1, W2, b2, W3, b3, w4, b4 = _unpack(params)
    gelu = lambda x: (0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))).astype(np.float32)
    x = np.asarray(F, dtype=np.float32)
    h = gelu(x @ W1 + b1)
    h = gelu(h @ W2 + b2)
    h = gelu(h @ W3 + b3)
    return (h @ w4 + b4[0]).astype(np.float32)


def train(F, R, epochs=300, lr=2e-3, seed=0, val_frac=0.1, log=print):
    """Fit the perceptron to log2(1 + R) (mean squared error, Adam, cosine decay).  Returns BF16
    parameters (uint16, library layout) and a dict of training metrics."""
    import torch
    torch.manual_seed(seed)
    X = torch.tensor(np.asarray(F, dtype=np.float32))
    Y = torch.tensor(np.log2(1.0 + np.asarray(R, dtype=np.float64)).astype(np.float32))
    n = X.shape[0]
    perm = torch.randperm(n)
    nv = max(1, int(val_frac * n))
    iv, it = perm[:nv], perm[nv:]
    model = torch.nn.Sequential(torch.nn.Linear(F_DIM, H_DIM), torch.nn.GELU(),
                                torch.nn.Linear(H_DIM, H_DIM), torch.nn.GELU(),
                                torch.nn.Linear(H_DIM, H_DIM), torch.nn.GELU(),
                                torch.nn.Linear(H_DIM, 1))
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, epochs)
    bs = 1024
    for ep in range(epochs):
        model.train()
        order = it[torch.randperm(len(it))]
        for a in range(0, len(order), bs):
            idx = order[a:a + bs]
            loss = torch.mean((model(X[idx]).squeeze(1) - Y[idx]) ** 2)
            opt.zero_grad()
            loss.backward()
            opt.step()
        sched.step()
        if ep % 50 == 0 or ep == epochs - 1:
            with torch.no_grad():
                mv = torch.mean(torch.abs(model(X[iv]).squeeze(1) - Y[iv])).item()
            log(f"epoch {ep}: train mse {loss.item():.4f}, validation mean |log2 error| {mv:.4f}")
    lin = [m for m in model if isinstance(m, torch.nn.Linear)]
    parts = [lin[0].weight.detach().T, lin[0].bias.detach(), lin[1].weight.detach().T, lin[1].bias.detach(),
             lin[2].weight.detach().T, lin[2].bias.detach(), lin[3].weight.detach()[0], lin[3].bias.detach()]
    flat = np.concatenate([p.contiguous().numpy().astype(np.float32).ravel() for p in parts])
    assert flat.size == N_PARAMS
    params = to_bf16(flat)
    pv = forward(params, F[iv.numpy()])
    yv = Y[iv].numpy()
    metrics = {"samples": int(n), "validation": int(nv), "val_mae_log2_bf16": float(np.mean(np.abs(pv - yv))),
               "val_under_frac_at_1.2x+4": float(np.mean(
                   np.asarray(R)[iv.numpy()] > np.ceil(1.2 * np.maximum(0, np.exp2(pv) - 1)) + 4))}
    return params, metrics


def save(path, params, meta):
    np.savez(path, params=np.asarray(params, dtype=np.uint16), meta=np.array(json.dumps(meta)))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        params = z["params"]
        meta = json.loads(str(z["meta"]))
    if params.size != N_PARAMS:
        raise ValueError(f"{path}: {params.size} parameters, expected {N_PARAMS}")
    return params, meta
