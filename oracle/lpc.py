"""CPU restatement (float64 numpy) of the learned sparse-approximate-inverse
preconditioner: features, message-passing network, factor, Kaporin loss and
its hand-written backward.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP path.

The reference has no learned preconditioner; this replaces its preconditioner
step z = M^-1 r (block-diagonal inverse, `src/PhysicsHelper.cu:494-499`, used
by the PCG loop `src/Cloth.cu:766-809`).  The network emits a block-lower-
triangular factor G on the static pattern of A (the pattern of
`src/Cloth.cu:709-744`), and the preconditioner is M^-1 = G^T G.

Definitions (every array float64; blocks are 3x3):
  R_i = chol(Minv_i) (lower), S_i = R_i^T, so S_i A_ii S_i^T = I.
  Ahat_ij = S_i A_ij S_j^T on the static pattern (Ahat_ii = I).
  node input xn (16): [sum_j |Ahat_ij|_F^2 (static, j != i), degree / 8,
                       sum_j |S_i Ac_ij S_j^T|_F^2 (contact blocks, j != i), 1, 0 ...]
  edge input xe (16): [Ahat_ij row-major (9), |Ahat_ij|_F, 1, 0 ...] per directed edge i -> j
  h = relu(xn Wn + bn), e = relu(xe We + be)
  per layer: P = h Wa, Q = h Wb; z_ij = relu(P_i + Q_j + e_ij Wc + bc);
             m_ij = z_ij W2 + b2; e_ij += m_ij; agg_i = sum_j m_ij;
             h_i += relu([h_i, agg_i] U1 + bu1) U2 + bu2
  o_i = relu(h_i Dn1 + bdn1) Dn2 + bdn2 (16; 6 used)
  q_ij = relu(e_ij De1 + bde1) De2 + bde2 (16; 9 used), lower edges i > j only
  Ghat_ii = [[exp o0, 0, 0], [o3, exp o1, 0], [o4, o5, exp o2]], Ghat_ij = q_ij[0:9] (3x3 row-major)
  G_ij = Ghat_ij S_j;  M^-1 = G^T G  (Ghat = I gives block Jacobi)
  loss = log(T / n) - (2 / n) sum_i (o_i0 + o_i1 + o_i2),  T = trace(Ghat Ahat Ghat^T), n = 3 N
  (Kaporin's condition number of Ghat Ahat Ghat^T, up to the constant log det Ahat / n),
  plus, when the system's solution x* (A x* = b) is given, lam (1 - cos^2_A(x*, M^-1 b)):
  how far the first preconditioned direction is from the solution, in the energy norm.

With bf16=True the forward rounds the matrix-product operands (weights and
activations) and the stored edge states to bfloat16 (round to nearest even),
as the device kernels do; accumulation stays float64 here.
"""
import numpy as np

H = 64
NL = 3
IN = 16


def _params():
    p = [("Wn", (IN, H)), ("bn", (H,)), ("We", (IN, H)), ("be", (H,))]
    for l in range(NL):
        p += [("Wa%d" % l, (H, H)), ("Wb%d" % l, (H, H)), ("Wc%d" % l, (H, H)), ("bc%d" % l, (H,)),
              ("W2%d" % l, (H, H)), ("b2%d" % l, (H,)), ("U1%d" % l, (2 * H, H)), ("bu1%d" % l, (H,)),
              ("U2%d" % l, (H, H)), ("bu2%d" % l, (H,))]
    p += [("Dn1", (H, H)), ("bdn1", (H,)), ("Dn2", (H, IN)), ("bdn2", (IN,)),
          ("De1", (H, H)), ("bde1", (H,)), ("De2", (H, IN)), ("bde2", (IN,))]
    return p


PARAMS = _params()
COUNT = sum(int(np.prod(s)) for _, s in PARAMS)


def unpack(flat):
    out, off = {}, 0
    for name, shape in PARAMS:
        k = int(np.prod(shape))
        out[name] = np.asarray(flat[off:off + k], dtype=np.float64).reshape(shape)
        off += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in PARAMS])


def init_weights(seed=0):
    """Uniform(+-1/sqrt(fan_in)) weights and biases; the two output layers start at zero
    (Ghat = I: the untrained network is block Jacobi)."""
    rng = np.random.default_rng(seed)
    d = {}
    fan_in = 1
    for name, shape in PARAMS:
        if len(shape) == 2:
            fan_in = shape[0]  # a bias follows its matrix in the list
        if name in ("Dn2", "bdn2", "De2", "bde2"):
            d[name] = np.zeros(shape)
            continue
        b = 1.0 / np.sqrt(fan_in)
        d[name] = rng.uniform(-b, b, shape)
    return pack(d)


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def relu(x):
    return np.maximum(x, 0.0)


class Graph:
    """Static pattern (CSR over nodes, sorted columns, diagonal included) with
    A's blocks on it, the block-Jacobi inverse and contact off-diagonal blocks."""

    def __init__(self, rp, ci, vals, Minv, contact=None):
        self.rp = np.asarray(rp, dtype=np.int64)
        self.ci = np.asarray(ci, dtype=np.int64)
        self.vals = np.asarray(vals, dtype=np.float64).reshape(-1, 3, 3)
        self.Minv = np.asarray(Minv, dtype=np.float64).reshape(-1, 3, 3)
        self.n = len(self.rp) - 1
        self.row = np.repeat(np.arange(self.n), np.diff(self.rp))
        off = self.row != self.ci
        self.eidx = np.nonzero(off)[0]           # pattern entries that are edges
        self.src = self.row[off]
        self.dst = self.ci[off]
        self.low = np.nonzero(self.src > self.dst)[0]   # edge ids (into src/dst) with i > j
        self.diag = np.nonzero(~off)[0]
        # contact: (rows, cols, blocks) of off-diagonal contact blocks
        self.contact = contact


def features(g):
    """S, Ahat (per edge), xn, xe."""
    R = np.linalg.cholesky(g.Minv)
    S = np.transpose(R, (0, 2, 1))
    Ah = np.einsum("eab,ebc,edc->ead", S[g.src], g.vals[g.eidx], S[g.dst])
    nrm2 = (Ah ** 2).sum(axis=(1, 2))
    xn = np.zeros((g.n, IN))
    np.add.at(xn[:, 0], g.src, nrm2)
    np.add.at(xn[:, 1], g.src, 1.0)
    xn[:, 1] /= 8.0
    if g.contact is not None and len(g.contact[0]):
        cr, cc, cb = g.contact
        cb = np.asarray(cb, dtype=np.float64).reshape(-1, 3, 3)
        m = cr != cc
        Ac = np.einsum("eab,ebc,edc->ead", S[cr[m]], cb[m], S[cc[m]])
        np.add.at(xn[:, 2], cr[m], (Ac ** 2).sum(axis=(1, 2)))
    xn[:, 3] = 1.0
    xe = np.zeros((len(g.src), IN))
    xe[:, :9] = Ah.reshape(-1, 9)
    xe[:, 9] = np.sqrt(nrm2)
    xe[:, 10] = 1.0
    return S, Ah, xn, xe


def forward(flat, g, bf=False):
    """Returns (o, q, cache); o: N x 16, q: lower edges x 16."""
    W = unpack(flat)
    r = bf16 if bf else (lambda a: a)
    Wr = {k: (r(v) if v.ndim == 2 else v) for k, v in W.items()}
    S, Ah, xn, xe = features(g)
    src, dst = g.src, g.dst
    c = {"S": S, "Ah": Ah, "xn": r(xn), "xe": r(xe)}
    h = relu(c["xn"] @ Wr["Wn"] + W["bn"])
    e = relu(c["xe"] @ Wr["We"] + W["be"])
    e = r(e)
    c["h0"], c["e0"] = h, e
    for l in range(NL):
        hb = r(h)
        P = hb @ Wr["Wa%d" % l]
        Q = hb @ Wr["Wb%d" % l]
        zpre = P[src] + Q[dst] + e @ Wr["Wc%d" % l] + W["bc%d" % l]
        z = relu(zpre)
        zb = r(z)
        m = zb @ Wr["W2%d" % l] + W["b2%d" % l]
        agg = np.zeros((g.n, H))
        np.add.at(agg, src, m)
        hin = np.concatenate([hb, r(agg)], axis=1)
        u = relu(hin @ Wr["U1%d" % l] + W["bu1%d" % l])
        ub = r(u)
        c["layer%d" % l] = dict(h=h, hb=hb, e=e, z=z, zb=zb, agg=agg, hin=hin, u=u, ub=ub)
        h = h + ub @ Wr["U2%d" % l] + W["bu2%d" % l]
        e = r(e + m)
    c["h3"], c["e3"] = h, e
    hb = r(h)
    dn = relu(hb @ Wr["Dn1"] + W["bdn1"])
    dnb = r(dn)
    o = dnb @ Wr["Dn2"] + W["bdn2"]
    el = e[g.low]
    de = relu(el @ Wr["De1"] + W["bde1"])
    deb = r(de)
    q = deb @ Wr["De2"] + W["bde2"]
    c.update(hb3=hb, dn=dn, dnb=dnb, el=el, de=de, deb=deb)
    return o, q, c


def ghat_blocks(o, q):
    """Ghat diagonal blocks (N x 3 x 3) and lower edge blocks (n_low x 3 x 3)."""
    n = len(o)
    D = np.zeros((n, 3, 3))
    D[:, 0, 0] = np.exp(o[:, 0])
    D[:, 1, 1] = np.exp(o[:, 1])
    D[:, 2, 2] = np.exp(o[:, 2])
    D[:, 1, 0] = o[:, 3]
    D[:, 2, 0] = o[:, 4]
    D[:, 2, 1] = o[:, 5]
    return D, q[:, :9].reshape(-1, 3, 3)


def factor(g, o, q, S):
    """G blocks: diagonal G_ii = Ghat_ii S_i, lower G_ij = Ghat_ij S_j."""
    D, L = ghat_blocks(o, q)
    return np.einsum("nab,nbc->nac", D, S), np.einsum("eab,ebc->eac", L, S[g.dst[g.low]])


def apply(g, Gd, Gl, r):
    """z = G^T (G r) for r of shape (N, 3)."""
    i, j = g.src[g.low], g.dst[g.low]
    y = np.einsum("nab,nb->na", Gd, r)
    np.add.at(y, i, np.einsum("eab,eb->ea", Gl, r[j]))
    z = np.einsum("nba,nb->na", Gd, y)
    np.add.at(z, j, np.einsum("eba,eb->ea", Gl, y[i]))
    return z


def kaporin(g, o, q, Ah):
    """loss, T and the gradients d loss / d o (N x 16), d loss / d q (n_low x 16)."""
    n3 = 3 * g.n
    D, L = ghat_blocks(o, q)
    # Ghat rows as dicts: row i -> {col: block}; Ahat lookup by (k, j)
    ahat = {}
    for eid, (a, b) in enumerate(zip(g.src, g.dst)):
        ahat[(a, b)] = Ah[eid]
    rows = [dict() for _ in range(g.n)]
    for i in range(g.n):
        rows[i][i] = D[i]
    for t, eid in enumerate(g.low):
        rows[g.src[eid]][g.dst[eid]] = L[t]
    T = 0.0
    Z = [dict() for _ in range(g.n)]
    eye = np.eye(3)
    for i in range(g.n):
        Pi = rows[i]
        for j in Pi:
            acc = np.zeros((3, 3))
            for k, Gik in Pi.items():
                A = eye if k == j else ahat.get((k, j))
                if A is not None:
                    acc += Gik @ A
            Z[i][j] = acc
            T += float((acc * Pi[j]).sum())
    loss = np.log(T / n3) - 2.0 / n3 * o[:, :3].sum()
    do = np.zeros_like(o)
    dq = np.zeros_like(q)
    s = 2.0 / T
    for i in range(g.n):
        Zi = Z[i][i]
        do[i, 0] = s * Zi[0, 0] * np.exp(o[i, 0]) - 2.0 / n3
        do[i, 1] = s * Zi[1, 1] * np.exp(o[i, 1]) - 2.0 / n3
        do[i, 2] = s * Zi[2, 2] * np.exp(o[i, 2]) - 2.0 / n3
        do[i, 3] = s * Zi[1, 0]
        do[i, 4] = s * Zi[2, 0]
        do[i, 5] = s * Zi[2, 1]
    for t, eid in enumerate(g.low):
        dq[t, :9] = s * Z[g.src[eid]][g.dst[eid]].reshape(9)
    return loss, T, do, dq


def _csr(g, blocks):
    import scipy.sparse as sp
    return sp.bsr_matrix((blocks, g.ci, g.rp), shape=(3 * g.n, 3 * g.n)).tocsr()


def solution_term(g, o, q, S, b, x):
    """L_sol = 1 - (b^T M b)^2 / ((u^T A u)(b^T x)), u = M b, M = G^T G, and its gradients
    d/do, d/dq through Ghat (G_ij = Ghat_ij S_j)."""
    Gd, Gl = factor(g, o, q, S)
    A = _csr(g, g.vals)
    b = np.asarray(b, dtype=np.float64).reshape(-1, 3)
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    i_l, j_l = g.src[g.low], g.dst[g.low]

    def Gmul(v):
        y = np.einsum("nab,nb->na", Gd, v)
        np.add.at(y, i_l, np.einsum("eab,eb->ea", Gl, v[j_l]))
        return y

    def GTmul(y):
        z = np.einsum("nba,nb->na", Gd, y)
        np.add.at(z, j_l, np.einsum("eba,eb->ea", Gl, y[i_l]))
        return z

    y = Gmul(b)
    u = GTmul(y)
    p = (A @ u.ravel()).reshape(-1, 3)
    w = Gmul(p)
    alpha = float((y * y).sum())
    beta = float((u * p).sum())
    gamma = float((b * x).sum())
    L = 1.0 - alpha * alpha / (beta * gamma)
    # dL = -(2 alpha / (beta gamma)) d alpha + (alpha^2 / (beta^2 gamma)) d beta with
    # d alpha / dG_ij = 2 y_i b_j^T and d beta / dG_ij = 2 (y_i p_j^T + w_i b_j^T)
    c1 = -4.0 * alpha / (beta * gamma)
    c2 = 2.0 * alpha * alpha / (beta * beta * gamma)

    def dG(i, j):
        return c1 * np.outer(y[i], b[j]) + c2 * (np.outer(y[i], p[j]) + np.outer(w[i], b[j]))
    # chain to Ghat: G_ij = Ghat_ij S_j  ->  dGhat_ij = dG_ij S_j^T
    do = np.zeros_like(o)
    dq = np.zeros_like(q)
    for i in range(g.n):
        D = dG(i, i) @ S[i].T
        do[i, 0] = D[0, 0] * np.exp(o[i, 0])
        do[i, 1] = D[1, 1] * np.exp(o[i, 1])
        do[i, 2] = D[2, 2] * np.exp(o[i, 2])
        do[i, 3] = D[1, 0]
        do[i, 4] = D[2, 0]
        do[i, 5] = D[2, 1]
    for t in range(len(g.low)):
        dq[t, :9] = (dG(i_l[t], j_l[t]) @ S[j_l[t]].T).reshape(9)
    return L, do, dq


def probe(n, seed):
    """The training probe of the smoothing term: v_i = (+-1, +-1, +-1) from a hash of (i, seed),
    as the device forms it (uint32 arithmetic)."""
    i = np.arange(n, dtype=np.uint32)
    with np.errstate(over="ignore"):
        h = i * np.uint32(2654435761) + np.uint32(seed) * np.uint32(2246822519)
        h = h ^ (h >> np.uint32(15))
        h = h * np.uint32(2246822519)
        h = h ^ (h >> np.uint32(13))
    return np.stack([np.where(h & np.uint32(k), 1.0, -1.0) for k in (1, 2, 4)], axis=1)


def smoothing_term(g, o, q, S, v, omega=1.0):
    """L_sm = |(I - omega M A) v|_A^2 / |v|_A^2, M = G^T G: how much of a probe's error in the
    energy norm one smoothing sweep leaves (for a white probe the energy is mostly in the high
    frequencies, the ones a smoother has to remove); and its gradients d/do, d/dq through Ghat."""
    Gd, Gl = factor(g, o, q, S)
    A = _csr(g, g.vals)
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    i_l, j_l = g.src[g.low], g.dst[g.low]

    def Gmul(x):
        y = np.einsum("nab,nb->na", Gd, x)
        np.add.at(y, i_l, np.einsum("eab,eb->ea", Gl, x[j_l]))
        return y

    def GTmul(y):
        z = np.einsum("nba,nb->na", Gd, y)
        np.add.at(z, j_l, np.einsum("eba,eb->ea", Gl, y[i_l]))
        return z

    u = (A @ v.ravel()).reshape(-1, 3)
    y = Gmul(u)
    e = v - omega * GTmul(y)
    p = (A @ e.ravel()).reshape(-1, 3)
    w = Gmul(p)
    num = float((e * p).sum())
    den = float((v * u).sum())
    L = num / den
    # dL/ds = -2 omega A e / den with s = G^T G u;  ds = dG^T y + G^T dG u, so
    # dL/dG_ij = c2 (y_i p_j^T + w_i u_j^T), c2 = -2 omega / den
    c2 = -2.0 * omega / den

    def dG(i, j):
        return c2 * (np.outer(y[i], p[j]) + np.outer(w[i], u[j]))
    do = np.zeros_like(o)
    dq = np.zeros_like(q)
    for i in range(g.n):
        D = dG(i, i) @ S[i].T
        do[i, 0] = D[0, 0] * np.exp(o[i, 0])
        do[i, 1] = D[1, 1] * np.exp(o[i, 1])
        do[i, 2] = D[2, 2] * np.exp(o[i, 2])
        do[i, 3] = D[1, 0]
        do[i, 4] = D[2, 0]
        do[i, 5] = D[2, 1]
    for t in range(len(g.low)):
        dq[t, :9] = (dG(i_l[t], j_l[t]) @ S[j_l[t]].T).reshape(9)
    return L, do, dq


def backward(flat, g, c, do, dq):
    """Gradient of the loss w.r.t. the flat weight vector, given d loss / d o and d loss / d q
    (the same rounding points as the forward when it ran with bf16, treated as identity)."""
    W = unpack(flat)
    gr = {k: np.zeros_like(v) for k, v in W.items()}
    src, dst = g.src, g.dst
    # decoders
    gr["Dn2"] += c["dnb"].T @ do
    gr["bdn2"] += do.sum(0)
    ddn = (do @ W["Dn2"].T) * (c["dn"] > 0)
    gr["Dn1"] += c["hb3"].T @ ddn
    gr["bdn1"] += ddn.sum(0)
    dh = ddn @ W["Dn1"].T
    gr["De2"] += c["deb"].T @ dq
    gr["bde2"] += dq.sum(0)
    dde = (dq @ W["De2"].T) * (c["de"] > 0)
    gr["De1"] += c["el"].T @ dde
    gr["bde1"] += dde.sum(0)
    de = np.zeros((len(src), H))
    de[g.low] = dde @ W["De1"].T
    for l in reversed(range(NL)):
        L = c["layer%d" % l]
        # h_out = h + ub U2 + bu2
        gr["U2%d" % l] += L["ub"].T @ dh
        gr["bu2%d" % l] += dh.sum(0)
        du = (dh @ W["U2%d" % l].T) * (L["u"] > 0)
        gr["U1%d" % l] += L["hin"].T @ du
        gr["bu1%d" % l] += du.sum(0)
        dhin = du @ W["U1%d" % l].T
        dh = dh + dhin[:, :H]
        dagg = dhin[:, H:]
        # e_out = e + m, agg = sum m
        dm = de + dagg[src]
        gr["W2%d" % l] += L["zb"].T @ dm
        gr["b2%d" % l] += dm.sum(0)
        dz = (dm @ W["W2%d" % l].T) * (L["z"] > 0)
        gr["Wc%d" % l] += L["e"].T @ dz
        gr["bc%d" % l] += dz.sum(0)
        de = de + dz @ W["Wc%d" % l].T
        dP = np.zeros((g.n, H))
        dQ = np.zeros((g.n, H))
        np.add.at(dP, src, dz)
        np.add.at(dQ, dst, dz)
        gr["Wa%d" % l] += L["hb"].T @ dP
        gr["Wb%d" % l] += L["hb"].T @ dQ
        dh = dh + dP @ W["Wa%d" % l].T + dQ @ W["Wb%d" % l].T
    # encoders
    dh0 = dh * (c["h0"] > 0)
    gr["Wn"] += c["xn"].T @ dh0
    gr["bn"] += dh0.sum(0)
    de0 = de * (c["e0"] > 0)
    gr["We"] += c["xe"].T @ de0
    gr["be"] += de0.sum(0)
    return pack(gr)


def loss_and_grad(flat, g, bf=False, b=None, x=None, lam=1.0, v=None, lam_s=1.0, omega=1.0):
    o, q, c = forward(flat, g, bf)
    loss, T, do, dq = kaporin(g, o, q, c["Ah"])
    if b is not None:
        ls, dos, dqs = solution_term(g, o, q, c["S"], b, x)
        loss = loss + lam * ls
        do = do + lam * dos
        dq = dq + lam * dqs
    if v is not None:
        lsm, dos, dqs = smoothing_term(g, o, q, c["S"], v, omega)
        loss = loss + lam_s * lsm
        do = do + lam_s * dos
        dq = dq + lam_s * dqs
    return loss, backward(flat, g, c, do, dq)
