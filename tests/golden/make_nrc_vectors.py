"""Writes nrc_torch_vectors.json: golden vectors for the radiance-cache
MLP and optimiser, computed with torch's CPU autograd and torch.optim.Adam
in float64 (library routines, independent of the oracle's hand-written
back-propagation).

Inputs are closed-form (no random state), so the test regenerates them:
    weight i of layer l: sqrt(6 / fan_in_l) * sin(0.731 * i + 0.3 + l)
    x[b][k] = 0.5 * cos(1.3 * b + 0.29 * k) for k < 43 (k = 39 set to 1), 0 after
    target[b][c] = 0.2 + 0.1 * sin(b + 2 * c)
Loss: mean over samples and channels of (y - t)^2 / (stop_grad(y)^2 + 0.01).

Run: python tests/golden/make_nrc_vectors.py
"""
import json
import os

import numpy as np
import torch

SHAPES = [(64, 48), (64, 64), (64, 64), (64, 64), (64, 64), (3, 64)]
FAN_IN = [43, 64, 64, 64, 64, 64]
N = 8


def weights():
    ws = []
    for l, ((r, c), fi) in enumerate(zip(SHAPES, FAN_IN)):
        i = np.arange(r * c, dtype=np.float64)
        w = np.sqrt(6.0 / fi) * np.sin(0.731 * i + 0.3 + l)
        if l == 0:
            w = w.reshape(r, c)
            w[:, 43:] = 0.0
            w = w.reshape(-1)
        ws.append(w)
    return ws


def inputs():
    b = np.arange(N, dtype=np.float64)[:, None]
    k = np.arange(48, dtype=np.float64)[None, :]
    x = 0.5 * np.cos(1.3 * b + 0.29 * k)
    x[:, 43:] = 0.0
    x[:, 39] = 1.0
    t = 0.2 + 0.1 * np.sin(np.arange(N)[:, None] + 2 * np.arange(3)[None, :])
    return x, t


def main():
    ws = weights()
    x, t = inputs()
    W = [torch.tensor(w.reshape(s), dtype=torch.float64, requires_grad=True) for w, s in zip(ws, SHAPES)]
    X = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    T = torch.tensor(t, dtype=torch.float64)
    h = X
    for l in range(5):
        h = torch.relu(h @ W[l].T)
    y = h @ W[5].T
    loss = ((y - T) ** 2 / (y.detach() ** 2 + 0.01)).mean()
    loss.backward()
    flat = torch.cat([w.grad.reshape(-1) for w in W]).numpy()
    sample_idx = list(range(0, flat.size, 97))
    out = {
        "n": N,
        "y": y.detach().numpy().tolist(),
        "loss": float(loss.detach()),
        "grad_x": X.grad.numpy().tolist(),
        "grad_w_norm_per_layer": [float(torch.linalg.norm(w.grad)) for w in W],
        "grad_w_sample_idx": sample_idx,
        "grad_w_sample": flat[sample_idx].tolist(),
    }
    # Adam: 16 parameters, three steps, betas (0.9, 0.99), eps 1e-15, lr 1e-2.
    i = np.arange(16, dtype=np.float64)
    p = torch.tensor(np.sin(i), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.9, 0.99), eps=1e-15)
    steps = []
    for s in range(3):
        opt.zero_grad()
        p.grad = torch.tensor(np.cos(1.7 * i + s) * 10.0 ** (-s), dtype=torch.float64)
        opt.step()
        steps.append(p.detach().numpy().tolist())
    out["adam"] = {"p0": np.sin(i).tolist(), "steps": steps}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nrc_torch_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
