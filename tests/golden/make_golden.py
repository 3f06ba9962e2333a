"""Generate the tiny-grid golden fixtures in tests/golden/p1_tiny.npz.

The fixtures are computed WITHOUT the oracle: a textbook P1 assembly
(element stiffness (1/(4|T|)) [b c] Dt [b c]^T, element mass |T|/12 (1+delta),
Robin edge |e|/6 [[2,1],[1,2]]) over the "right" triangulation of a rectangle,
Dirichlet rows replaced by identity, solved with numpy.linalg.solve.  The
oracle (an FFC-quadrature restatement) and the GPU stencil are checked
against these.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

NEU, DIR, ROB = 0, 1, 2
LEFT, RIGHT, TOP, BOTTOM = 0, 1, 2, 3


def mesh(nx, ny, W, H):
    xs = np.array([(j * W) / nx for j in range(nx + 1)])
    ys = np.array([(i * H) / ny for i in range(ny + 1)])
    X, Y = np.meshgrid(xs, ys)
    tris = []
    for i in range(ny):
        for j in range(nx):
            v0 = j + i * (nx + 1)
            v1, v2 = v0 + 1, v0 + nx + 1
            v3 = v2 + 1
            tris.append((v0, v1, v3))  # diagonal bottom-left -> top-right
            tris.append((v0, v2, v3))
    return X.ravel(), Y.ravel(), tris


def wall_nodes(nx, ny, w):
    idx = []
    for i in range(ny + 1):
        for j in range(nx + 1):
            on = {LEFT: j == 0, RIGHT: j == nx, TOP: i == ny, BOTTOM: i == 0}[w]
            if on:
                idx.append(j + i * (nx + 1))
    return idx


def assemble(nx, ny, W, H, dt, D, gamma, f, walls, Dt=None):
    X, Y, tris = mesh(nx, ny, W, H)
    n = len(X)
    A = np.zeros((n, n))
    M = np.zeros((n, n))
    area_sum = np.zeros(n)
    for t in tris:
        x = X[list(t)]
        y = Y[list(t)]
        area = 0.5 * abs((x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]))
        b = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]])
        c = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]])
        if Dt is None:
            T = np.eye(2)
        else:
            T = np.mean([Dt[v] for v in t], axis=0)  # P1 tensor, exact mean over the triangle
        G = np.stack([b, c], axis=1)  # 3x2
        Ke = D * (G @ T @ G.T) / (4.0 * area)
        Me = area / 12.0 * (np.ones((3, 3)) + np.eye(3))
        for a in range(3):
            area_sum[t[a]] += area / 3.0
            for bb in range(3):
                A[t[a], t[bb]] += (1.0 + dt * gamma) * Me[a, bb] + dt * Ke[a, bb]
                M[t[a], t[bb]] += Me[a, bb]
    rhs_const = dt * f * area_sum
    rob_rhs = np.zeros(n)
    for w, (kind, value, r, s) in walls.items():
        if kind != ROB:
            continue
        nodes = wall_nodes(nx, ny, w)
        for a, bnode in zip(nodes[:-1], nodes[1:]):
            e = np.hypot(X[bnode] - X[a], Y[bnode] - Y[a])
            A[a, a] += dt * r * e / 3.0
            A[bnode, bnode] += dt * r * e / 3.0
            A[a, bnode] += dt * r * e / 6.0
            A[bnode, a] += dt * r * e / 6.0
            rob_rhs[a] += dt * r * s * e / 2.0
            rob_rhs[bnode] += dt * r * s * e / 2.0
    return A, M, rhs_const + rob_rhs


def dirichlet(nx, ny, walls):
    g = {}
    for w in (LEFT, RIGHT, TOP, BOTTOM):  # later walls overwrite corners
        kind, value, r, s = walls.get(w, (NEU, 0, 0, 0))
        if kind == DIR:
            for k in wall_nodes(nx, ny, w):
                g[k] = value
    return g


CASES = [
    # name, nx, ny, W, H, dt, D, gamma, f, walls, anisotropic
    ("neumann_2x1", 2, 1, 2.0, 1.0, 0.1, 4.0, 0.0, 0.0, {}, False),
    ("top_dir_4x2", 4, 2, 2.0, 1.0, 0.1, 400.0, 0.01, 0.0, {TOP: (DIR, 0.0, 0, 0)}, False),
    ("mixed_8x4", 8, 4, 4.0, 2.0, 0.1, 100.0, 0.05, 0.3,
     {TOP: (DIR, 1.5, 0, 0), LEFT: (ROB, 0, 2.0, 0.7), RIGHT: (ROB, 0, 0.5, 0.0)}, False),
    ("rect_h_6x3", 6, 3, 3.0, 2.1, 0.2, 50.0, 0.0, 0.0,
     {BOTTOM: (DIR, 0.25, 0, 0), LEFT: (DIR, -0.5, 0, 0)}, False),
    ("all_dir_5x4", 5, 4, 2.5, 2.0, 0.1, 10.0, 0.1, 1.0,
     {w: (DIR, 0.0, 0, 0) for w in range(4)}, False),
    ("aniso_6x4", 6, 4, 3.0, 2.0, 0.1, 20.0, 0.0, 0.0, {TOP: (DIR, 0.0, 0, 0)}, True),
]


def main():
    out = {}
    rng = np.random.default_rng(12345)
    for name, nx, ny, W, H, dt, D, gamma, f, walls, aniso in CASES:
        n = (nx + 1) * (ny + 1)
        Dt = None
        tensor = np.zeros((n, 3))
        if aniso:
            th = rng.uniform(0, np.pi, n)
            a, b = 1.0, 0.3
            d11 = a * np.cos(th) ** 2 + b * np.sin(th) ** 2
            d22 = a * np.sin(th) ** 2 + b * np.cos(th) ** 2
            d12 = (a - b) * np.sin(th) * np.cos(th)
            tensor = np.stack([d11, d22, d12], axis=1)
            Dt = [np.array([[t[0], t[2]], [t[2], t[1]]]) for t in tensor]
        else:
            tensor[:, 0] = 1.0
            tensor[:, 1] = 1.0
        A, M, rhs0 = assemble(nx, ny, W, H, dt, D, gamma, f, walls, Dt)
        u0 = rng.uniform(0, 1, n)
        b = M @ u0 + rhs0
        A_bc = A.copy()
        b_bc = b.copy()
        for k, v in dirichlet(nx, ny, walls).items():
            A_bc[k, :] = 0.0
            A_bc[k, k] = 1.0
            b_bc[k] = v
        x = np.linalg.solve(A_bc, b_bc)
        wall_arr = np.zeros((4, 4))
        for w, (kind, value, r, s) in walls.items():
            wall_arr[w] = (kind, value, r, s)
        params = np.array([nx, ny, W, H, dt, D, gamma, f], dtype=np.float64)
        out[f"{name}/params"] = params
        out[f"{name}/walls"] = wall_arr
        out[f"{name}/tensor"] = tensor
        out[f"{name}/u0"] = u0
        out[f"{name}/A"] = A
        out[f"{name}/A_bc"] = A_bc
        out[f"{name}/b_bc"] = b_bc
        out[f"{name}/x"] = x
    out["names"] = np.array([c[0] for c in CASES])
    np.savez_compressed(os.path.join(HERE, "p1_tiny.npz"), **out)
    print("wrote", os.path.join(HERE, "p1_tiny.npz"))


if __name__ == "__main__":
    main()
