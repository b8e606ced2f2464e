"""CPU oracle of the path-guiding network — TEST INFRASTRUCTURE ONLY.

Restates, in numpy FP64, the guiding model the GPU trains and queries: frequency encoding of
(position, normal, outgoing direction), a 64-64-64-64-64-41 ReLU MLP, a mixture of 8
von Mises-Fisher lobes over the incident direction (outputs 0-39) and the lobes' selection
probability a = 0.05 + 0.9 sigmoid(output 40), with the weighted negative log-likelihood used
for online training, -w log(a q + (1 - a) b) (q = the lobes' density, b = the BSDF density, a =
the selection probability of the one-sample MIS mixture, a record's fixed value or the
network's own), and its analytic gradients.  The reference renderer
has no path guiding; this is the specification the GPU kernels are checked against.  The
analytic gradients are themselves pinned by central finite differences in the tests.

Optionally rounds the layer inputs and weights to bfloat16 (round to nearest even) at the
points where the GPU feeds the matrix cores, so that forward values can be compared tightly.
"""
from __future__ import annotations

import numpy as np

N_IN = 64
N_HID = 64
N_LAYERS = 5            # 4 hidden layers of 64, then the output layer
N_LOBES = 8
N_LOBE_VALS = 5 * N_LOBES   # per lobe: mean direction (3, unnormalised), log kappa, weight logit
N_OUT = N_LOBE_VALS + 1      # + the logit of the lobes' selection probability
KAPPA_LOG_MIN, KAPPA_LOG_MAX = -8.0, 7.0
SEL_MIN, SEL_MAX = 0.05, 0.95
LAYER_SHAPES = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]


def bf16_round(x):
    """Round float values to bfloat16 (nearest, ties to even), returned as float64."""
    a = np.asarray(x, np.float32).copy()
    u = a.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    u = ((u + rounding) >> 16) << 16
    out = u.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    out[nan] = np.nan
    return out.astype(np.float64)


def features(P, N, V, bbox_min, bbox_max):
    """9 scalars in [0, 1]: position in the scene box, normal and outgoing direction."""
    P, N, V = (np.asarray(a, np.float64).reshape(-1, 3) for a in (P, N, V))
    lo, hi = np.asarray(bbox_min, np.float64), np.asarray(bbox_max, np.float64)
    p = (P - lo) / np.maximum(hi - lo, 1e-30)
    return np.concatenate([p, 0.5 * (N + 1.0), 0.5 * (V + 1.0)], axis=1)


def encode(x9):
    """63 values (9 raw + sin/cos at 3 frequency bands) padded with a constant 1 -> 64."""
    x = np.asarray(x9, np.float64).reshape(-1, 9)
    n = x.shape[0]
    e = np.zeros((n, N_IN))
    e[:, 0:9] = x
    for k in range(3):
        a = (2.0 ** k) * np.pi * x
        e[:, 9 + 18 * k + 0:9 + 18 * k + 18:2] = np.sin(a)
        e[:, 9 + 18 * k + 1:9 + 18 * k + 18:2] = np.cos(a)
    e[:, 63] = 1.0
    return e


def init_params(seed):
    """He-uniform weights (stored as [in, out]) and zero biases from a numpy stream; tests
    hand these arrays to the GPU so that both sides start from the same parameters."""
    rng = np.random.default_rng(seed)
    params = []
    for fan_in, fan_out in LAYER_SHAPES:
        lim = np.sqrt(6.0 / fan_in)
        params.append((rng.uniform(-lim, lim, (fan_in, fan_out)), np.zeros(fan_out)))
    return params


def forward(params, e, bf16=False):
    """Returns the list of layer inputs (e, h1..h4) and the raw outputs."""
    acts = []
    h = np.asarray(e, np.float64)
    for li, (W, b) in enumerate(params):
        x = bf16_round(h) if bf16 else h
        Wq = bf16_round(W) if bf16 else W
        acts.append(x)
        z = x @ Wq + b
        h = np.maximum(z, 0.0) if li < N_LAYERS - 1 else z
    return acts, h


def selection(o):
    """The lobes' selection probability from raw outputs (output 40)."""
    z = np.asarray(o, np.float64).reshape(-1, N_OUT)[:, N_LOBE_VALS]
    return SEL_MIN + (SEL_MAX - SEL_MIN) / (1.0 + np.exp(-z))


def decode(o):
    """Mixture parameters from raw outputs: mu (n,8,3), kappa (n,8), weights (n,8)."""
    o = np.asarray(o, np.float64).reshape(-1, N_OUT)[:, :N_LOBE_VALS].reshape(-1, N_LOBES, 5)
    r = o[:, :, 0:3]
    rl = np.linalg.norm(r, axis=2, keepdims=True)
    mu = np.where(rl > 0, r / np.where(rl > 0, rl, 1), np.array([0.0, 0.0, 1.0]))
    kappa = np.exp(np.clip(o[:, :, 3], KAPPA_LOG_MIN, KAPPA_LOG_MAX))
    logits = o[:, :, 4]
    w = np.exp(logits - logits.max(axis=1, keepdims=True))
    w /= w.sum(axis=1, keepdims=True)
    return mu, kappa, w


def vmf_norm(kappa):
    """kappa / (2 pi (1 - exp(-2 kappa))), with the kappa -> 0 limit 1 / (4 pi)."""
    k = np.asarray(kappa, np.float64)
    small = k < 1e-6
    ks = np.where(small, 1.0, k)
    return np.where(small, 1.0 / (4 * np.pi), ks / (2 * np.pi * (-np.expm1(-2 * ks))))


def mixture_pdf(o, omega):
    mu, kappa, w = decode(o)
    om = np.asarray(omega, np.float64).reshape(-1, 1, 3)
    c = (mu * om).sum(axis=2)
    p = vmf_norm(kappa) * np.exp(kappa * (c - 1.0))
    return (w * p).sum(axis=1)


def _mix_prob(o, q, sel, samp_pdf):
    """The selection probability a of each record: the network's own where samp_pdf > 0 (the
    record's probability is learned) and the record's sel > 0, else sel (None: 0)."""
    if sel is None:
        return np.zeros_like(q)
    a = np.broadcast_to(np.asarray(sel, np.float64), q.shape).copy()
    if samp_pdf is not None:
        use = (np.broadcast_to(np.asarray(samp_pdf, np.float64), q.shape) > 0) & (a > 0)
        a = np.where(use, selection(o), a)
    return a


def _sel_share(q, bsdf_pdf, a):
    """a q / (a q + (1 - a) b): the lobes' share of the one-sample MIS density (1 where b or a is 0)."""
    if bsdf_pdf is None:
        return np.ones_like(q)
    b = np.broadcast_to(np.asarray(bsdf_pdf, np.float64), q.shape)
    use = (a > 0) & (b > 0)
    den = np.where(use, a * q + (1 - a) * b, 1.0)
    return np.where(use, a * q / den, 1.0)


def record_density(q, bsdf_pdf, a):
    """Density whose log the record loss takes: a q + (1 - a) b, or q where b or a is 0."""
    if bsdf_pdf is None:
        return q
    b = np.broadcast_to(np.asarray(bsdf_pdf, np.float64), q.shape)
    use = (a > 0) & (b > 0)
    return np.where(use, a * q + (1 - a) * b, q)


def loss(params, e, omega, weight, bf16=False, bsdf_pdf=None, sel=None, samp_pdf=None):
    """Mean of -weight * log(a q + (1 - a) b) (a q's constant factor dropped where b = 0): the
    objective of the lobe outputs 0..39."""
    _, o = forward(params, e, bf16)
    q = mixture_pdf(o, omega)
    a = _mix_prob(o, q, sel, samp_pdf)
    return -(np.asarray(weight, np.float64) * np.log(record_density(q, bsdf_pdf, a))).mean()


def selection_objective(o, omega, weight, bsdf_pdf, sel, samp_pdf):
    """Per record, weight^2 samp_pdf / (a q + (1 - a) b) with a the network's own selection
    probability: for a direction drawn with density samp_pdf and weight = integrand /
    samp_pdf, the one-sample estimate of the second moment of integrand / (a q + (1 - a) b),
    i.e. of the one-sample MIS estimator's variance plus a constant.  Records not learned
    (samp_pdf = 0), or with a = 0 or b = 0: 0.  Its derivative in output 40 is output_grad's
    column 40."""
    o = np.asarray(o, np.float64).reshape(-1, N_OUT)
    q = mixture_pdf(o, omega)
    a = _mix_prob(o, q, sel, samp_pdf)
    b = np.broadcast_to(np.asarray(bsdf_pdf, np.float64), q.shape)
    ps = np.broadcast_to(np.asarray(samp_pdf, np.float64), q.shape)
    lrn = (ps > 0) & (a > 0) & (b > 0)
    w = np.broadcast_to(np.asarray(weight, np.float64), q.shape)
    dens = a * q + (1 - a) * b
    return np.where(lrn, w * w * ps / np.where(lrn, dens, 1.0), 0.0)


def dlogp_dkappa(k):
    """d/dkappa of log(kappa / (2 pi (1 - exp(-2 kappa)))) + kappa (mu.w - 1), minus the last term."""
    k = np.asarray(k, np.float64)
    small = k < 1e-2
    series = 1.0 - k / 3.0 + k ** 3 / 45.0 - 1.0  # 1/k - 2/expm1(2k) - 1, series near 0, minus 1
    ks = np.where(small, 1.0, k)
    with np.errstate(over="ignore"):  # expm1 overflows to inf for large kappa: 2 / inf = 0
        exact = 1.0 / ks - 2.0 / np.expm1(2.0 * ks) - 1.0
    return np.where(small, series, exact)


def output_grad(o, omega, weight, bsdf_pdf=None, sel=None, samp_pdf=None):
    """Output gradients per sample (not yet divided by the batch size).  Outputs 0..39:
    d(-weight * log(a q + (1 - a) b)); without b and a (or where either is 0) d(-weight * log q).
    a is the network's selection probability where samp_pdf > 0 (and sel > 0); then output 40
    gets -weight^2 samp_pdf (q - b) / (a q + (1 - a) b)^2 da/dz, the derivative of
    selection_objective; otherwise a = sel and output 40 gets 0."""
    oa = np.asarray(o, np.float64).reshape(-1, N_OUT)
    n = oa.shape[0]
    o = oa[:, :N_LOBE_VALS].reshape(-1, N_LOBES, 5)
    r = o[:, :, 0:3]
    rl = np.linalg.norm(r, axis=2, keepdims=True)
    mu, kappa, w = decode(oa)
    om = np.asarray(omega, np.float64).reshape(-1, 1, 3)
    c = (mu * om).sum(axis=2)
    p = vmf_norm(kappa) * np.exp(kappa * (c - 1.0))
    q = (w * p).sum(axis=1, keepdims=True)
    a = _mix_prob(oa, q[:, 0], sel, samp_pdf)
    share = _sel_share(q[:, 0], bsdf_pdf, a).reshape(-1, 1)
    scale = -np.asarray(weight, np.float64).reshape(-1, 1) * share / q  # d loss / d q
    g = np.zeros_like(o)
    wp = w * p
    # mean direction through the normalisation
    g[:, :, 0:3] = (scale * wp * kappa)[:, :, None] * (om - c[:, :, None] * mu) / np.where(rl > 0, rl, 1)
    # log kappa (zero gradient where clamped)
    inside = (o[:, :, 3] > KAPPA_LOG_MIN) & (o[:, :, 3] < KAPPA_LOG_MAX)
    dk = wp * (dlogp_dkappa(kappa) + c) * kappa   # dlogp/dkappa = 1/k - 2/expm1(2k) + (c - 1)
    g[:, :, 3] = scale * np.where(inside, dk, 0.0)
    # softmax logits
    g[:, :, 4] = scale * (wp - w * q)
    out = np.zeros((n, N_OUT))
    out[:, :N_LOBE_VALS] = g.reshape(n, N_LOBE_VALS)
    if samp_pdf is not None and bsdf_pdf is not None and sel is not None:
        b = np.broadcast_to(np.asarray(bsdf_pdf, np.float64), (n,))
        ps = np.broadcast_to(np.asarray(samp_pdf, np.float64), (n,))
        lrn = (ps > 0) & \
            (np.broadcast_to(np.asarray(sel, np.float64), (n,)) > 0) & (b > 0)
        dens = a * q[:, 0] + (1 - a) * b
        dadz = (a - SEL_MIN) * (SEL_MAX - a) / (SEL_MAX - SEL_MIN)
        wt = np.broadcast_to(np.asarray(weight, np.float64), (n,))
        out[:, N_LOBE_VALS] = np.where(lrn, -wt * wt * ps * (q[:, 0] - b) / np.where(lrn, dens, 1.0) ** 2 * dadz,
                                       0.0)
    return out


def backward(params, e, omega, weight, bf16=False, bsdf_pdf=None, sel=None, samp_pdf=None):
    """Gradients of the mean record objectives w.r.t. every (W, b), in the parameter layout.
    Output 40 (the selection logit) is a head on the last hidden layer: its gradient reaches
    its own weights and bias but not the hidden layers."""
    acts, o = forward(params, e, bf16)
    n = acts[0].shape[0]
    g = output_grad(o, omega, weight, bsdf_pdf, sel, samp_pdf) / n
    grads = [None] * N_LAYERS
    for li in range(N_LAYERS - 1, -1, -1):
        W, b = params[li]
        Wq = bf16_round(W) if bf16 else W
        x = acts[li]
        grads[li] = (x.T @ g, g.sum(axis=0))
        if li > 0:
            if li == N_LAYERS - 1:
                Wq = Wq.copy()
                Wq[:, N_LOBE_VALS] = 0.0
            gx = g @ Wq.T
            g = gx * (acts[li] > 0)
    return grads


def adam_step(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) in place; t is the 1-based step count."""
    out = []
    for i, ((W, b), (gW, gb)) in enumerate(zip(params, grads)):
        new = []
        for j, (x, g) in enumerate(((W, gW), (b, gb))):
            m[i][j] = b1 * m[i][j] + (1 - b1) * g
            v[i][j] = b2 * v[i][j] + (1 - b2) * g * g
            mh = m[i][j] / (1 - b1 ** t)
            vh = v[i][j] / (1 - b2 ** t)
            new.append(x - lr * mh / (np.sqrt(vh) + eps))
        out.append(tuple(new))
    return out
