"""Learned subgrid closure, host side: the 8-32-32-1 network's parameters and normalisation, the
training datasets the solver generates from fine runs (2x box filter onto the coarse grid), GPU
training through the library's own kernels, and attaching a trained network to a solver
(params.closure = 'mlp'; Smagorinsky stays the fallback). See include/wt.h for the inputs and
the target.
"""
import numpy as np

from . import _native as N

NPARAM = N.MLP_NPARAM


class Network:
    """FP32 master parameters (W1[32][8], b1, W2[32][32], b2, w3, b3) + input normalisation."""

    def __init__(self, params=None, mu=None, sigma=None, nu_scale=1e-3, smag_floor=1.0):
        self.params = np.zeros(NPARAM, np.float32) if params is None else np.asarray(params, np.float32).copy()
        self.mu = np.zeros(8, np.float32) if mu is None else np.asarray(mu, np.float32).copy()
        self.sigma = np.ones(8, np.float32) if sigma is None else np.asarray(sigma, np.float32).copy()
        self.nu_scale = float(nu_scale)
        self.smag_floor = float(smag_floor)

    @classmethod
    def initial(cls, seed=20261203, nu_scale=1e-3, smag_floor=1.0):
        """He-normal hidden layers, small output layer, zero biases."""
        rng = np.random.default_rng(seed)
        p = np.zeros(NPARAM, np.float32)
        p[0:256] = rng.normal(0, np.sqrt(2 / 8), 256)
        p[288:1312] = rng.normal(0, np.sqrt(2 / 32), 1024)
        p[1344:1376] = rng.normal(0, np.sqrt(1 / 32), 32)
        return cls(p, nu_scale=nu_scale, smag_floor=smag_floor)

    def to_struct(self):
        m = N.Mlp()
        N.C.memmove(m.params, self.params.ctypes.data, NPARAM * 4)
        for i in range(8):
            m.mu[i] = float(self.mu[i])
            m.sigma[i] = float(self.sigma[i])
        m.nu_scale = self.nu_scale
        m.smag_floor = self.smag_floor
        return m

    def save(self, path):
        np.savez(path, params=self.params, mu=self.mu, sigma=self.sigma, nu_scale=np.float64(self.nu_scale),
                 smag_floor=np.float64(self.smag_floor))

    @classmethod
    def load(cls, path):
        z = np.load(path)
        return cls(z["params"], z["mu"], z["sigma"], float(z["nu_scale"]), float(z["smag_floor"]))


class Dataset:
    """Device-resident training records (8 raw inputs, target nu_t, coarse cell index)."""

    def __init__(self, capacity):
        self._lib = N.load()
        h = N.C.c_void_p()
        N.check(self._lib.wt_dataset_create(int(capacity), N.C.byref(h)))
        self._h = h
        N.track(self)

    def add(self, fine, coarse):
        """Append the samples of `fine`'s current state, box-filtered onto `coarse`'s grid."""
        a = N.C.c_int64()
        N.check(self._lib.wt_dataset_add(self._h, fine._h, coarse._h, N.C.byref(a)))
        return a.value

    def append(self, records):
        r = np.ascontiguousarray(records, np.float32).reshape(-1, 10)
        N.check(self._lib.wt_dataset_append(self._h, N.ptr(r), len(r)))

    @property
    def count(self):
        n = N.C.c_int64()
        N.check(self._lib.wt_dataset_count(self._h, N.C.byref(n)))
        return n.value

    def read(self, first=0, count=None):
        count = self.count - first if count is None else count
        out = np.empty((count, 10), np.float32)
        N.check(self._lib.wt_dataset_read(self._h, int(first), int(count), N.ptr(out)))
        return out

    def destroy(self):
        if getattr(self, "_h", None):
            self._lib.wt_dataset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _sample(dataset, samples, seed):
    n = dataset.count
    if n <= samples:
        return dataset.read(0, n)
    idx = np.sort(np.random.default_rng(seed).choice(n, samples, replace=False))
    # read in a few large slices rather than the whole set at once
    out = np.empty((samples, 10), np.float32)
    step = 1 << 22
    k = 0
    for a in range(0, n, step):
        sel = idx[(idx >= a) & (idx < a + step)]
        if len(sel):
            blk = dataset.read(a, min(step, n - a))
            out[k:k + len(sel)] = blk[sel - a]
            k += len(sel)
    return out


def fit_normalization(net, dataset, samples=1_000_000, seed=1, quantile=99.9, headroom=8.0):
    """Input mean / std and the output scale from a sample of the dataset. The output is capped at
    MLP_TARGET_CLIP = 10 nu_scale, so nu_scale is set to the `quantile` of the targets over
    `headroom`: the cap lies above the tail of the targets, where the eddy viscosity matters (wake,
    beside the body), while the bulk of the (uniform-flow) targets near zero still resolves."""
    recs = _sample(dataset, samples, seed)
    x = recs[:, :8].astype(np.float64)
    net.mu = x.mean(0).astype(np.float32)
    sd = x.std(0)
    net.sigma = np.where(sd > 1e-12, sd, 1.0).astype(np.float32)
    t = recs[:, 8].astype(np.float64)
    scale = float(np.percentile(t, quantile)) / headroom if np.any(t > 0) else 1e-3
    net.nu_scale = max(scale, 1e-7)
    return net


def wall_weight(d):
    """Weight of the near-wall term: 1 within two cells of the body, falling linearly to 0 at four
    (the wall-distance input saturates at four cells)."""
    return np.clip((4.0 - np.asarray(d, np.float64)) / 2.0, 0.0, 1.0)


def smagorinsky_nu(raw, cs):
    """The coarse grid's Smagorinsky eddy viscosity (Cs Delta)^2 |S| (Delta = one cell) from records'
    |S| input (x5 = log(1 + |S| / 1e-3))."""
    return cs * cs * 1e-3 * np.expm1(np.abs(np.asarray(raw, np.float64)[:, 5]))


def near_wall_dataset(dataset, cs, chunk=1 << 22):
    """A copy of `dataset` whose targets carry, besides the filtered subgrid stress, the coarse
    Smagorinsky eddy viscosity weighted toward the body: t + wall_weight(d) (Cs Delta)^2 |S|. The
    2x-filtered fine field holds little subgrid stress (its scales are resolved on the coarse grid),
    so the error of the coarse run sits at the body, where the coarse grid under-resolves the
    boundary layer and the separation; this term gives the network a near-wall eddy viscosity whose
    size the output calibration sets."""
    n = dataset.count
    out = Dataset(max(n, 1))
    for a in range(0, n, chunk):
        r = dataset.read(a, min(chunk, n - a))
        r[:, 8] = (r[:, 8].astype(np.float64) + wall_weight(r[:, 6]) * smagorinsky_nu(r, cs)).astype(np.float32)
        out.append(r)
    return out


def train(dataset, net, epochs=10, batch=4096, lr=2e-3, seed=20261203, max_samples=0, beta1=0.9, beta2=0.999, eps=1e-8):
    """Train `net` in place on the GPU; returns the loss per epoch."""
    lib = N.load()
    m = net.to_struct()
    o = N.TrainOpts(int(epochs), int(batch), float(lr), float(beta1), float(beta2), float(eps), int(seed), int(max_samples))
    loss = np.zeros(epochs, np.float64)
    N.check(lib.wt_mlp_train(dataset._h, N.C.byref(m), N.C.byref(o), N.ptr(loss)))
    net.params = np.ctypeslib.as_array(m.params).astype(np.float32).copy()
    return loss


def gradient(dataset, net, first, count):
    """(loss, gradient) of one mini-batch of records [first, first + count)."""
    lib = N.load()
    m = net.to_struct()
    g = np.zeros(NPARAM, np.float32)
    l = N.C.c_double()
    N.check(lib.wt_mlp_grad(dataset._h, N.C.byref(m), int(first), int(count), N.ptr(g), N.C.byref(l)))
    return l.value, g


def infer(net, raw):
    """nu_t from raw inputs (n, 8) on the GPU (the step kernel's MFMA path)."""
    lib = N.load()
    r = np.ascontiguousarray(raw, np.float32).reshape(-1, 8)
    out = np.empty(len(r), np.float32)
    m = net.to_struct()
    N.check(lib.wt_mlp_infer(N.C.byref(m), N.ptr(r), len(r), N.ptr(out)))
    return out


def attach(solver, net):
    """Load `net` into `solver` and switch its closure to the network."""
    m = net.to_struct()
    N.check(solver._lib.wt_mlp_load(solver._h, N.C.byref(m)))
    solver.setParams({"closure": "mlp"})


def wall_distance(solver):
    """Wall distance of every cell in cells (capped at 4), from the solver's flags."""
    out = np.empty(solver.n, np.uint8)
    N.check(solver._lib.wt_read_wall_distance(solver._h, N.ptr(out)))
    return out.astype(np.float64) / 32.0
