"""CPU oracle of the learned PPE warm start / preconditioner network (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this module, as the checker of
the HIP kernels in simulation_project_amd/csrc/mpnn.hip.  It has no counterpart in the reference
(the reference solves no pressure Poisson equation); the network is defined here and in the HIP
source alike:

Graph: the active (non-Dirichlet) rows of the PPE matrix A (symmetric, positive definite, negative
off-diagonals).  Edge weights alpha_ij = -A_ij / A_ii (row-normalised, sum_j alpha_ij <= 1).
Aggregation Agg(u)_i = sum_j alpha_ij u_j; its transpose is AggT(g)_j = sum_i alpha_ij g_i, which
by the symmetry of A equals sum_{i in N(j)} (-A_ji) (g_i / A_ii).

Three message-passing layers, hidden width H = 128, message width M = 32:
    u_l = h_l U_l                      (N x F_l) (F_l x M)
    m_l = Agg(u_l)
    z_l = h_l W_l + m_l V_l + c_l      (F_l x H), (M x H)
    h_{l+1} = relu(z_l)                (layers 1, 2)
    y = z_3                            (layer 3: H -> 2, no activation)
h_1 = x, the 16 input features.  Outputs: warm start x0 = s * y[:, 0] (s = the solve's pressure
scale), diagonal preconditioner weight omega = 2 * sigmoid(y[:, 1]).

Loss of one solve (p* = the plain solver's solution, b the right-hand side, D = diag(A)):
    r~   = b / (D s) - (A y0) / D                 (Jacobi-scaled residual of x0 = s y0)
    L    = mean((y0 - p*/s)^2) + beta * mean(r~^2)
    r    = (b - A x0) / s,  q = D^-1 A r,  e = r - omega q
    Lw   = mean(e^2) / mean(r^2)
    total = L + Lw
Gradients are exact, Lw's dependence on y0 through r included (hand-written backward), and are
pinned against central finite differences in FP64 by the tests.
"""
import numpy as np

H, M, F_IN, N_OUT = 128, 32, 16, 2
LAYERS = ((F_IN, H), (H, H), (H, N_OUT))
PARAM_ORDER = ("U1", "W1", "V1", "c1", "U2", "W2", "V2", "c2", "U3", "W3", "V3", "c3")


def param_shapes():
    out = {}
    for l, (fi, fo) in enumerate(LAYERS, 1):
        out[f"U{l}"] = (fi, M)
        out[f"W{l}"] = (fi, fo)
        out[f"V{l}"] = (M, fo)
        out[f"c{l}"] = (fo,)
    return out


def init_params(seed=1):
    """He-style initialisation from a counter-free numpy generator (seeded)."""
    rng = np.random.default_rng(seed)
    p = {}
    for k, shp in param_shapes().items():
        if k.startswith("c"):
            p[k] = np.zeros(shp)
        else:
            p[k] = rng.standard_normal(shp) * np.sqrt(2.0 / shp[0])
    p["W3"] *= 0.1
    p["V3"] *= 0.1
    return p


def flatten(p):
    return np.concatenate([np.asarray(p[k], np.float64).ravel() for k in PARAM_ORDER])


def unflatten(v):
    p, o = {}, 0
    for k in PARAM_ORDER:
        shp = param_shapes()[k]
        n = int(np.prod(shp))
        p[k] = np.asarray(v[o:o + n], np.float64).reshape(shp)
        o += n
    return p


class Graph:
    """Active-row PPE system in CSR: ptr [n+1], col [nnz] (active-row indices), val [nnz]
    (off-diagonal A_ij < 0), diag [n], b [n]."""

    def __init__(self, ptr, col, val, diag, b):
        self.ptr = np.asarray(ptr, np.int64)
        self.col = np.asarray(col, np.int64)
        self.val = np.asarray(val, np.float64)
        self.diag = np.asarray(diag, np.float64)
        self.b = np.asarray(b, np.float64)
        self.n = len(self.diag)
        self.row = np.repeat(np.arange(self.n), np.diff(self.ptr))

    def spmv_off(self, x):
        """sum_j A_ij x_j over the off-diagonals (x may be [n] or [n, k])."""
        x = np.asarray(x, np.float64)
        prod = self.val[:, None] * x[self.col] if x.ndim == 2 else self.val * x[self.col]
        out = np.zeros((self.n,) + x.shape[1:])
        np.add.at(out, self.row, prod)
        return out

    def matvec(self, x):
        return self.diag * x + self.spmv_off(x)

    def agg(self, u):
        return -self.spmv_off(u) / self.diag[:, None]

    def agg_t(self, g):
        return -self.spmv_off(g / self.diag[:, None])


def forward(p, x, g):
    """Returns y [n, 2] and the cache for backward."""
    cache = []
    h = np.asarray(x, np.float64)
    for l in (1, 2, 3):
        u = h @ p[f"U{l}"]
        m = g.agg(u)
        z = h @ p[f"W{l}"] + m @ p[f"V{l}"] + p[f"c{l}"]
        cache.append((h, m, z))
        h = np.maximum(z, 0.0) if l < 3 else z
    return h, cache


def backward(p, cache, dy, g):
    """Gradients of sum(dy * y) with respect to every parameter."""
    grads = {}
    dz = dy
    for l in (3, 2, 1):
        h, m, z = cache[l - 1]
        if l < 3:
            dz = dz * (z > 0)
        grads[f"W{l}"] = h.T @ dz
        grads[f"V{l}"] = m.T @ dz
        grads[f"c{l}"] = dz.sum(axis=0)
        dm = dz @ p[f"V{l}"].T
        du = g.agg_t(dm)
        grads[f"U{l}"] = h.T @ du
        if l > 1:
            dz = dz @ p[f"W{l}"].T + du @ p[f"U{l}"].T
    return grads


def omega_of(y1):
    return 2.0 / (1.0 + np.exp(-y1))


def loss_and_dy(y, g, s, p_star, beta=1.0):
    """Loss of one solve and dL/dy (both outputs)."""
    n = g.n
    y0, y1 = y[:, 0], y[:, 1]
    t = p_star / s
    rt = g.b / (g.diag * s) - g.matvec(y0) / g.diag
    lx = np.mean((y0 - t) ** 2)
    lr = np.mean(rt ** 2)
    d0 = 2.0 * (y0 - t) / n - beta * (2.0 / n) * g.matvec(rt / g.diag)
    r = g.b / s - g.matvec(y0)
    q = g.matvec(r) / g.diag
    om = omega_of(y1)
    e = r - om * q
    rr = np.mean(r ** 2) + 1e-300
    ee = np.mean(e ** 2)
    lw = ee / rr
    dom = 2.0 * (om * q - r) * q / (n * rr)
    d1 = dom * om * (1.0 - om / 2.0)
    # dLw/dr = (2/n) [(e - A D^-1 (omega e)) / rr - ee / rr^2 r];  dr/dy0 = -A
    dr = (2.0 / n) * ((e - g.matvec(om * e / g.diag)) / rr - ee / rr ** 2 * r)
    d0 = d0 - g.matvec(dr)
    return lx + beta * lr + lw, np.stack([d0, d1], axis=1), {"lx": lx, "lr": lr, "lw": lw}


def features(g, x_prev, rel, dist, vel, div, wall, nnz_scale=128.0):
    """The 16 input features of the active rows, given the per-row geometric moments computed
    from the particles (rel = sum_j alpha_ij (r_j - r_i)/l0 [n,3], dist = sum_j alpha_ij
    |r_j - r_i|/l0 [n], vel = v* dt/l0 [n,3], div = sum_j alpha_ij (v_j - v_i).(r_j - r_i)/|r|^2 dt
    [n], wall = min(wall distance / re, 1) [n]).  Returns (x [n,16], s)."""
    jac = g.b / g.diag
    s = np.sqrt(np.mean(jac ** 2)) + 1e-30
    alpha_sum = -g.spmv_off(np.ones(g.n)) / g.diag
    f = np.zeros((g.n, F_IN))
    f[:, 0] = jac / s
    f[:, 1] = x_prev / s
    f[:, 2] = 1.0 - alpha_sum
    f[:, 3] = np.diff(g.ptr) / nnz_scale
    f[:, 4] = g.agg(jac[:, None])[:, 0] / s
    f[:, 5] = g.agg(np.asarray(x_prev, np.float64)[:, None])[:, 0] / s
    f[:, 6:9] = rel
    f[:, 9] = dist
    f[:, 10:13] = vel
    f[:, 13] = div
    f[:, 14] = wall
    f[:, 15] = 1.0
    return f, s
