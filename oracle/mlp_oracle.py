"""Float64 numpy restatement of the planner's cost-model MLP (forward pass,
mean-squared-error loss and its gradient), for checking the GPU training
kernels. TEST INFRASTRUCTURE ONLY.

Layout (as in include/ket_amd.h): W1[H][D], b1[H], W2[H][H], b2[H], W3[H], b3;
y = W3 . relu(W2 relu(W1 x + b1) + b2) + b3; L = mean (y - t)^2.
"""
import numpy as np

D, H = 16, 64


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * D].reshape(H, D); o += H * D
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, D)
    a1 = np.maximum(x @ W1.T + b1, 0.0)
    a2 = np.maximum(a1 @ W2.T + b2, 0.0)
    return a2 @ W3 + b3


def loss_and_grad(p, x, t):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, D)
    t = np.asarray(t, np.float64)
    n = len(x)
    z1 = x @ W1.T + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0.0)
    y = a2 @ W3 + b3
    loss = np.mean((y - t) ** 2)
    dy = 2.0 * (y - t) / n
    gW3 = a2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, W3) * (z2 > 0)
    gW2 = d2.T @ a1
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g
