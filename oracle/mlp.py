"""FP64 CPU restatement of the per-cycle power surrogate (an MLP over toggle bits).

TEST INFRASTRUCTURE ONLY: the oracle for the device surrogate's forward and
backward kernels.  Product code never imports it.

The surrogate has no counterpart in the reference; this module restates the
network the device kernels implement, in FP64, with the device's BF16 rounding
points optionally applied (``bf16=True``):

    W1b = bf16(W1), W2b = bf16(W2)
    z1  = X W1b^T + b1,          h1 = bf16(GELU(z1))
    z2  = h1 W2b^T + b2,         h2 = bf16(GELU(z2))
    y^  = h2 . w3 + b3
    L   = mean((y^ - y)^2)
    dy  = 2 (y^ - y) / M
    dw3 = h2^T dy, db3 = sum(dy)
    dz2 = bf16(dy w3^T * GELU'(z2)),  dW2 = dz2^T h1,  db2 = colsum(dz2)
    dz1 = bf16((dz2 W2b) * GELU'(z1)), dW1 = dz1^T X,  db1 = colsum(dz1)

GELU(x) = x Phi(x) (the exact, erf-based form).  With ``bf16=False`` the
functions are the exact FP64 network and its exact gradient, which the tests
pin with central finite differences.
"""
import math

import numpy as np

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(x):
    """Round to bfloat16 (round to nearest, ties to even) through float32."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def keep_mask(seed, rows, words):
    """The training-time input-dropout keep bits of (seed, row, word), uint32
    arithmetic: h = seed ^ row*0x9E3779B1 ^ word*0x85EBCA77, then the
    murmur3 finaliser steps (xor-shift 16, * 0x7FEB352D, xor-shift 15,
    * 0x846CA68B, xor-shift 16)."""
    r = np.asarray(rows, dtype=np.uint64)
    w = np.asarray(words, dtype=np.uint64)
    M = np.uint64(0xFFFFFFFF)
    h = (np.uint64(seed) ^ ((r * np.uint64(0x9E3779B1)) & M) ^ ((w * np.uint64(0x85EBCA77)) & M)) & M
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7FEB352D)) & M
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x846CA68B)) & M
    h ^= h >> np.uint64(16)
    return h.astype(np.uint32)


def dropped_inputs(bits, n_in, seed):
    """Inputs of a training step with input dropout: kept bits count 2."""
    b = np.asarray(bits, dtype=np.uint32)
    R, W = b.shape
    km = keep_mask(seed, np.arange(R)[:, None], np.arange(W)[None, :])
    return 2.0 * unpack_bits(b & km, n_in)


def unpack_bits(bits, n_in):
    """[M][n_in/32] uint32 rows -> [M][n_in] 0/1 float64 (bit k = bit k%32 of word k/32)."""
    b = np.asarray(bits, dtype=np.uint32)
    M = b.shape[0]
    out = ((b[:, :, None] >> np.arange(32, dtype=np.uint32)[None, None, :]) & 1).reshape(M, -1)
    return out[:, :n_in].astype(np.float64)


def split(params, n_in, H):
    """Parameter vector -> (W1[H][n_in], b1, W2[H][H], b2, w3, b3)."""
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * n_in].reshape(H, n_in); o += H * n_in
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o:o + 1]; o += 1
    assert o == p.size, "parameter count does not match the shape"
    return W1, b1, W2, b2, w3, b3


def forward(params, X, n_in, H, bf16_points=True):
    W1, b1, W2, b2, w3, b3 = split(params, n_in, H)
    q = bf16 if bf16_points else (lambda a: a)
    z1 = X @ q(W1).T + b1
    h1 = q(gelu(z1))
    z2 = h1 @ q(W2).T + b2
    h2 = q(gelu(z2))
    out = h2 @ w3 + b3[0]
    return out, (z1, h1, z2, h2)


def loss_and_grad(params, X, y, n_in, H, bf16_points=True):
    """Mean squared error and its gradient (same layout as params)."""
    W1, b1, W2, b2, w3, b3 = split(params, n_in, H)
    q = bf16 if bf16_points else (lambda a: a)
    out, (z1, h1, z2, h2) = forward(params, X, n_in, H, bf16_points)
    M = X.shape[0]
    d = out - y
    loss = float(np.mean(d * d))
    dy = 2.0 * d / M
    dw3 = h2.T @ dy
    db3 = np.array([dy.sum()])
    dz2 = q(dy[:, None] * w3[None, :] * gelu_grad(z2))
    dW2 = dz2.T @ h1
    db2 = dz2.sum(axis=0)
    dz1 = q((dz2 @ q(W2)) * gelu_grad(z1))
    dW1 = dz1.T @ X
    db1 = dz1.sum(axis=0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, db3])
    return loss, g
