"""The learned closure of the coarse concentration update (no counterpart in
the reference).

The coarse solver misses what the fine grid resolves: a fine run's state,
box-filtered onto cells twice as large and advanced k steps by the coarse
solver, lands away from the filtered fine state k steps later.  That
difference per unit time is the target (k = 1, or the horizon given to
collect); a 3-layer 3x3x3 convolutional network (libpllcart3d_amd's BF16
kernels, trained on the GPU) predicts it from the coarse cells' c, u, v, w
and mu, and the coarse run adds the prediction to N in every concentration
update (Hc_op's explicit term, pllcart3d.c:1539-1644), so that a step adds
dt times it to c; the prediction is made anew every `every` steps
(set_closure).  A guarded run falls back to plain WENO where c leaves
[0, 1] or the correction injects mass (pllc_sim_run_guarded).  The
mixing-zone length, from the y-z average of c along x (the profile
conc_average writes to cavgyz.bin, pllcart3d.c:3691-3724), measures the
coarse run against the fine one.
"""
import ctypes as C
import time

import numpy as np

from . import Ensemble, Sim, ic_fill
from ._lib import FIELDS, IC_MULTIMODE, Params, last_error, load
from .configs import make_params

NIN = 5
XMARGIN = 4           # cells at the inlet and the outlet without correction
OFFSETS = 27


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ---------------------------------------------------------------- network

def nparams():
    return load().pllc_cnn_nparams()


def flat_from_dict(P):
    """{'W1' [27][16][16], 'b1' [16], 'W2', 'b2', 'W3' [27][16][1], 'b3' [1]}
    -> the library's flat parameter vector (W3 and b3 padded to 16 outputs)."""
    W3 = np.zeros((OFFSETS, 16, 16))
    W3[:, :, 0] = P["W3"][:, :, 0]
    b3 = np.zeros(16)
    b3[0] = P["b3"][0]
    return np.concatenate([P["W1"].ravel(), P["b1"], P["W2"].ravel(), P["b2"], W3.ravel(), b3])


def dict_from_flat(v):
    w = OFFSETS * 256
    P = {"W1": v[:w].reshape(OFFSETS, 16, 16), "b1": v[w:w + 16],
         "W2": v[w + 16:2 * w + 16].reshape(OFFSETS, 16, 16), "b2": v[2 * w + 16:2 * w + 32]}
    W3 = v[2 * w + 32:3 * w + 32].reshape(OFFSETS, 16, 16)
    P["W3"] = W3[:, :, :1].copy()
    P["b3"] = v[3 * w + 32:3 * w + 33].copy()
    return P


class Network:
    """The closure network on a coarse grid of nx x ny x nz cells (device
    parameters, Adam state, activations and training samples)."""

    def __init__(self, nx, ny, nz):
        self.L = load()
        self.shape = (nx, ny, nz)
        self.n = nx * ny * nz
        self.h = self.L.pllc_cnn_create(nx, ny, nz)
        if not self.h:
            raise RuntimeError("pllc_cnn_create: " + last_error())

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(what + ": " + last_error())

    def close(self):
        if self.h:
            self.L.pllc_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float64)
        assert flat.size == nparams()
        self._chk(self.L.pllc_cnn_set_params(self.h, _dp(flat)), "pllc_cnn_set_params")

    def get_params(self):
        out = np.zeros(nparams())
        self._chk(self.L.pllc_cnn_get_params(self.h, _dp(out)), "pllc_cnn_get_params")
        return out

    def set_norm(self, shift, scale, out_scale):
        shift = np.ascontiguousarray(shift, dtype=np.float64)
        scale = np.ascontiguousarray(scale, dtype=np.float64)
        self._chk(self.L.pllc_cnn_set_norm(self.h, _dp(shift), _dp(scale), float(out_scale)), "pllc_cnn_set_norm")

    def forward(self, x):
        """x: [nx][ny][nz][5] (shifted and scaled) -> y [nx][ny][nz]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.shape == self.shape + (NIN,)
        y = np.zeros(self.shape)
        self._chk(self.L.pllc_cnn_forward(self.h, _fp(x), _dp(y)), "pllc_cnn_forward")
        return y

    def add_sample(self, x, t, mask):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        assert x.shape == self.shape + (NIN,) and t.shape == self.shape and m.shape == self.shape
        i = self.L.pllc_cnn_add_sample(self.h, _fp(x), _fp(t), m.ctypes.data_as(C.POINTER(C.c_ubyte)))
        if i < 0:
            raise RuntimeError("pllc_cnn_add_sample: " + last_error())
        return i

    def loss_grad(self, sample):
        L = C.c_double(0.0)
        g = np.zeros(nparams())
        self._chk(self.L.pllc_cnn_loss_grad(self.h, sample, C.byref(L), _dp(g)), "pllc_cnn_loss_grad")
        return L.value, g

    def loss(self, sample):
        L = C.c_double(0.0)
        self._chk(self.L.pllc_cnn_loss_grad(self.h, sample, C.byref(L), None), "pllc_cnn_loss_grad")
        return L.value

    def train(self, nsteps, order, lr):
        order = np.ascontiguousarray(order, dtype=np.int32)
        losses = np.zeros(nsteps)
        self._chk(self.L.pllc_cnn_train(self.h, nsteps, order.ctypes.data_as(C.POINTER(C.c_int)), len(order),
                                        float(lr), _dp(losses)), "pllc_cnn_train")
        return losses


def init_params(seed):
    """Seeded initial parameters (1/sqrt(fan-in) scaling, a small last layer)."""
    rng = np.random.default_rng(seed)
    P = {"W1": rng.normal(size=(OFFSETS, 16, 16)) / np.sqrt(OFFSETS * NIN),
         "W2": rng.normal(size=(OFFSETS, 16, 16)) / np.sqrt(OFFSETS * 16),
         "W3": 0.1 * rng.normal(size=(OFFSETS, 16, 1)) / np.sqrt(OFFSETS * 16),
         "b1": np.zeros(16), "b2": np.zeros(16), "b3": np.zeros(1)}
    P["W1"][:, NIN:, :] = 0.0
    return P


# ------------------------------------------------------------ box filter

def coarse_params(fp):
    """The 2x coarser grid of the same run."""
    assert fp.nx % 2 == 0 and fp.ny % 2 == 0 and fp.nz % 2 == 0
    cp = Params()
    C.memmove(C.byref(cp), C.byref(fp), C.sizeof(Params))
    cp.nx, cp.ny, cp.nz = fp.nx // 2, fp.ny // 2, fp.nz // 2
    return cp


def _pairs(a, axis):
    """Mean of consecutive pairs along axis."""
    s = list(a.shape)
    s[axis:axis + 1] = [s[axis] // 2, 2]
    return a.reshape(s).mean(axis=axis + 1)


def filter_cells(f, fp):
    """Cell field (ghosted [nx+6][ny+2][nz+2]) -> coarse interior cells:
    the mean of the 8 fine cells of each coarse cell."""
    a = f[3:3 + fp.nx, 1:1 + fp.ny, 1:1 + fp.nz]
    return _pairs(_pairs(_pairs(a, 0), 1), 2)


def filter_faces(f, fp, axis):
    """Face-centred velocity -> coarse faces: a coarse face is every second
    fine face along `axis`, its value the mean over the 4 fine faces it
    covers (the flux through it is kept).  Returns the coarse interior
    block with the closing face: [nx+1][ny][nz] for u (axis 0), [nx][ny+1][nz]
    for v, [nx][ny][nz+1] for w."""
    nx, ny, nz = fp.nx, fp.ny, fp.nz
    lo = [3, 1, 1]
    hi = [3 + nx, 1 + ny, 1 + nz]
    hi[axis] += 1
    a = f[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    a = np.take(a, np.arange(0, a.shape[axis], 2), axis=axis)
    for ax in range(3):
        if ax != axis:
            a = _pairs(a, ax)
    return a


def _full(cp, interior, axis=None):
    """Coarse ghosted array with the interior (plus closing face) placed."""
    out = np.zeros((cp.nx + 6, cp.ny + 2, cp.nz + 2))
    sh = interior.shape
    out[3:3 + sh[0], 1:1 + sh[1], 1:1 + sh[2]] = interior
    return out


def coarse_c0(fp, c0_fine):
    """The coarse run's initial concentration: the filtered fine one on the
    interior, the coarse grid's own erf profile on the x-ghosts, Neumann
    copies on the y and z ghosts."""
    cp = coarse_params(fp)
    from ._lib import IC_ERF
    full = ic_fill(cp, IC_ERF)
    full[3:3 + cp.nx, 1:1 + cp.ny, 1:1 + cp.nz] = filter_cells(c0_fine, fp)
    full[:, 0, :] = full[:, 1, :]
    full[:, cp.ny + 1, :] = full[:, cp.ny, :]
    full[:, :, 0] = full[:, :, 1]
    full[:, :, cp.nz + 1] = full[:, :, cp.nz]
    return full


def features(c, u, v, w, mu, cp):
    """The network's raw inputs on coarse interior cells: c, the
    cell-centred velocities and mu (ghosted coarse arrays in)."""
    s = (slice(3, 3 + cp.nx), slice(1, 1 + cp.ny), slice(1, 1 + cp.nz))
    uc = 0.5 * (u[3:3 + cp.nx] + u[4:4 + cp.nx])[:, 1:1 + cp.ny, 1:1 + cp.nz]
    vc = 0.5 * (v[:, 1:1 + cp.ny] + v[:, 2:2 + cp.ny])[3:3 + cp.nx, :, 1:1 + cp.nz]
    wc = 0.5 * (w[:, :, 1:1 + cp.nz] + w[:, :, 2:2 + cp.nz])[3:3 + cp.nx, 1:1 + cp.ny]
    return np.stack([c[s], uc, vc, wc, mu[s]], axis=-1)


def loss_mask(cp):
    m = np.zeros((cp.nx, cp.ny, cp.nz), np.uint8)
    m[XMARGIN:cp.nx - XMARGIN] = 1
    return m


def filtered_state(fine, fp):
    """The fine simulation's current state box-filtered onto the coarse grid:
    c (ghosted, coarse_c0's ghost rules), the face velocities and their
    pressure gradients (ghost rows left to the coarse ghost routines)."""
    cp = coarse_params(fp)
    f = {n: fine.download(n) for n in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "U_PG", "V_PG", "W_PG")}
    st = {"C": coarse_c0(fp, f["C_OLD"])}
    for q, ax in (("U", 0), ("V", 1), ("W", 2)):
        st[q] = _full(cp, filter_faces(f[q + "_OLD"], fp, ax))
        st[q + "PG"] = _full(cp, filter_faces(f[q + "_PG"], fp, ax))
    return st


def load_coarse(coarse, st):
    """A coarse simulation set to a filtered state (old = new, the
    viscosity of c, the velocity ghosts of update_velo_boundary)."""
    for n, q in (("C_OLD", "C"), ("C_NEW", "C"), ("U_NEW", "U"), ("V_NEW", "V"), ("W_NEW", "W"),
                 ("U_PG", "UPG"), ("V_PG", "VPG"), ("W_PG", "WPG")):
        coarse.upload(n, st[q])
    coarse.op("VISCOSITY")
    coarse.op("VELO_BOUNDARY")
    coarse.op("UPDATE_OLD")


def sample_one_step(fine, fp, horizon=1):
    """(raw features, target) at the fine simulation's current state: the
    coarse solver's error over `horizon` steps.  The filtered state is
    advanced on the coarse grid (mixed precision) step by step with the fine
    run's dt (the coarse Timestep set to each fine step's dt, which the
    coarse CFL limit does not bind); the target is
    (filter(c_fine after) - c_coarse after) / (time advanced), the rate the
    correction of N must add for the coarse run to land on the filtered fine
    one.  Advances `fine` by `horizon` steps."""
    cp = coarse_params(fp)
    st = filtered_state(fine, fp)
    dts = [fine.run(1)[1] for _ in range(horizon)]
    fc1 = fine.download("C_OLD")
    cc = None
    for n, dt in enumerate(dts):
        cpx = coarse_params(fp)
        cpx.Timestep = dt
        coarse = Sim(cpx, "mixed")
        coarse.init(st["C"])
        if n == 0:
            load_coarse(coarse, st)
            mu = coarse.download("MUO_CELL")
        else:
            for f, a in cc.items():
                coarse.upload(f, a)
        _, dtc = coarse.run(1)
        if abs(dtc - dt) > 1e-9 * dt:
            raise RuntimeError("coarse step dt %.9g differs from the fine step's %.9g" % (dtc, dt))
        cc = {f: coarse.download(f) for f in FIELDS}
        coarse.close()
    x = features(st["C"], st["U"], st["V"], st["W"], mu, cp)
    t = (filter_cells(fc1, fp) - cc["C_OLD"][3:3 + cp.nx, 1:1 + cp.ny, 1:1 + cp.nz]) / sum(dts)
    return x, t


# ------------------------------------------------------ mixing-zone length

def yz_average(c, p):
    """The y-z average of c over interior cells per x cell (cavgyz)."""
    return c[3:3 + p.nx, 1:1 + p.ny, 1:1 + p.nz].mean(axis=(1, 2))


def mixing_length(c, p, lo=0.05, hi=0.95):
    """Distance along x between the first points where the y-z averaged
    concentration of the displaced fluid reaches lo and hi (linear
    interpolation between cell centres)."""
    prof = yz_average(c, p)
    if prof[0] > prof[-1]:
        prof = 1.0 - prof   # the displacing fluid (c = 1 here) enters at X1
    dx = (p.X2 - p.X1) / p.nx
    xc = p.X1 + (np.arange(p.nx) + 0.5) * dx

    def first(level):
        i = int(np.argmax(prof >= level))
        if i == 0:
            return xc[0]
        f = (level - prof[i - 1]) / (prof[i] - prof[i - 1])
        return xc[i - 1] + f * dx

    return first(hi) - first(lo)




def set_closure(sim, net, every=0):
    """The simulation adds net's correction to N from now on (None: off);
    with every = k >= 1 the correction is made at the first substage of every
    k-th step and kept until the next one (0: at every substage)."""
    if load().pllc_sim_set_closure(sim.h, net.h if net is not None else None, int(every)) != 0:
        raise RuntimeError("pllc_sim_set_closure: " + last_error())


def run_guarded(sim, nsteps, check_every=2, tol_bound=1e-2, tol_mass=1e-4):
    """nsteps guarded steps with the simulation's closure; returns the number
    of steps that fell back to plain WENO and the largest bound excess and
    mass ratio seen."""
    fb = C.c_int(0)
    worst = np.zeros(2)
    if load().pllc_sim_run_guarded(sim.h, nsteps, check_every, tol_bound, tol_mass, C.byref(fb), _dp(worst)) != 0:
        raise RuntimeError("pllc_sim_run_guarded: " + last_error())
    return fb.value, worst


def sim_time(sim):
    return sim.run(0)[0]


# -------------------------------------------------------------- pipeline

def draw_training_runs(n, seed0=500):
    """Fine runs of the training set: R ~ U[2, 4], Pe ~ U[1000, 3000], the
    16x16-mode front (amplitude 0.01, width 0.02), seed seed0 + r."""
    out = []
    for r in range(n):
        rng = np.random.default_rng(seed0 + r)
        out.append((float(rng.uniform(2.0, 4.0)), float(rng.uniform(1000.0, 3000.0)), seed0 + r))
    return out


def collect(fine_shape, runs, steps, every, warmup, log=print, horizon=1):
    """(features, targets) of box-filtered fine runs: one sample every
    `every` steps from step `warmup` to `steps` of each run, each target the
    coarse solver's error over `horizon` steps.  The fine runs advance
    together as one ensemble (plain FP64, each member a captured graph
    replayed as a branch of one graph; equal to the runs done alone bit for
    bit); each sample advances its member by `horizon` steps, so the members
    stay in step.  Samples are returned run by run."""
    fps, c0s = [], []
    for R, Pe, seed in runs:
        fp = make_params(*fine_shape, Vratio=R, Pe=Pe, Inf_thick=0.02)
        fps.append(fp)
        c0s.append(ic_fill(fp, IC_MULTIMODE, seed, 0.01))
    ens = Ensemble(fps, "plain")
    try:
        for m, c0 in zip(ens.members, c0s):
            m.init(c0)
        Xr = [[] for _ in runs]
        Tr = [[] for _ in runs]
        done = 0
        for target_step in range(warmup, steps + 1, every):
            ens.run(target_step - done)
            for r, (fp, m) in enumerate(zip(fps, ens.members)):
                x, t = sample_one_step(m, fp, horizon)
                Xr[r].append(x.astype(np.float32))
                Tr[r].append(t.astype(np.float32))
            done = target_step + horizon
        for (R, Pe, seed), Tl in zip(runs, Tr):
            rms = float(np.sqrt(np.mean(Tl[-1].astype(np.float64) ** 2)))
            log("[closure data] R=%.2f Pe=%.0f seed %d: samples up to step %d, last target rms %.3e"
                % (R, Pe, seed, done, rms))
    finally:
        ens.close()
    return [x for xs in Xr for x in xs], [t for ts in Tr for t in ts]


def normalisation(X, T, mask):
    """Input shifts and scales (mean and 1/std over the masked cells) and the
    output scale (rms of the targets)."""
    m = mask.astype(bool)
    allx = np.concatenate([x[m] for x in X]).astype(np.float64)
    shift = allx.mean(axis=0)
    scale = 1.0 / np.maximum(allx.std(axis=0), 1e-12)
    out_scale = float(np.sqrt(np.mean(np.concatenate([t[m] for t in T]).astype(np.float64) ** 2)))
    return shift, scale, out_scale


def mixing_length_at(sim, p, t_end):
    """Step sim until its time passes t_end; the mixing-zone length at t_end,
    interpolated linearly between the two steps around it."""
    t0 = sim_time(sim)
    L0 = mixing_length(sim.download("C_OLD"), p)
    while True:
        t1, _ = sim.run(1)
        L1 = mixing_length(sim.download("C_OLD"), p)
        if t1 >= t_end:
            return L0 + (L1 - L0) * (t_end - t0) / (t1 - t0)
        t0, L0 = t1, L1


def train(X, T, mask, shape, steps=3000, seed=1, log=print):
    """A network for a coarse grid of `shape` trained with Adam on the
    samples (X, T): normalisation from the samples, then `steps` steps at
    learning rate 1e-3 and steps // 3 at 3e-4, samples in seeded random
    order.  Returns (network, (shift, scale, out_scale), losses)."""
    shift, scale, osc = normalisation(X, T, mask)
    net = Network(*shape)
    net.set_norm(shift, scale, osc)
    idx = [net.add_sample((x - shift) * scale, t / osc, mask) for x, t in zip(X, T)]
    net.set_params(flat_from_dict(init_params(seed)))
    rng = np.random.default_rng(seed)
    losses = []
    for n, lr in ((steps, 1e-3), (steps // 3, 3e-4)):
        order = rng.permutation(np.repeat(idx, 1 + n // len(idx)))[:n]
        losses.append(net.train(n, order, lr))
    losses = np.concatenate(losses)
    log("[closure train] %d steps: loss %.4f -> %.4f" % (len(losses), losses[:50].mean(), losses[-50:].mean()))
    return net, (shift, scale, osc), losses


def validation_gain(net, norm, X, T, mask):
    """Mean masked loss of the trained network and of no correction (a zero
    output) on held-out samples; the gain is 1 - trained / none."""
    shift, scale, osc = norm
    idx = [net.add_sample((x - shift) * scale, t / osc, mask) for x, t in zip(X, T)]
    trained = float(np.mean([net.loss(i) for i in idx]))
    m = mask.astype(bool)
    none = float(np.mean([np.mean((t[m] / osc) ** 2) for t in T]))
    return trained, none, 1.0 - trained / none


def coarse_run(fp, c0_fine, t_end, net=None, timestep=None, guard_every=10, tol_bound=1e-2, tol_mass=1e-4,
               chunk=20, every=1):
    """The coarse (2x) mixed-precision run of the fine run (fp, c0_fine) up to
    t_end, with the closure `net` (guarded, a new correction every `every`
    steps) or without it; returns the mixing-zone length at t_end, the wall time (s),
    the steps, the fallback steps and the guard's worst bound excess and mass
    ratio."""
    cp = coarse_params(fp)
    if timestep is not None:
        cp.Timestep = timestep

    def make():
        s = Sim(cp, "mixed")
        s.init(coarse_c0(fp, c0_fine))
        if net is not None:
            set_closure(s, net, every)
        return s

    warm = make()   # first use: graph capture and kernel loading
    warm.run(2)
    warm.synchronize()
    warm.close()
    s = make()
    t0 = time.perf_counter()
    n = nfb = 0
    worst = np.zeros(2)
    while True:
        t, dt = s.run(0)
        if t + 3 * max(dt, 1e-6) >= t_end:
            break
        if net is not None:
            f, w = run_guarded(s, chunk, guard_every, tol_bound, tol_mass)
            nfb += f
            worst = np.maximum(worst, w)
        else:
            s.run(chunk)
        n += chunk
    s.synchronize()
    wall = time.perf_counter() - t0
    L = mixing_length_at(s, cp, t_end)
    s.close()
    return L, wall, n, nfb, worst
