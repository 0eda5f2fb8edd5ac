"""The learned sigma(B) surrogate and the surrogate-guided B sweep.

* :class:`MLP` — ctypes handle of the ensemble MLP of include/es_surrogate.h
  (6 -> 128 -> 128 -> 2, tanh; BF16 MFMA forward/backward, Adam).
* :class:`Surrogate` — the MLP over physical inputs: the dimensionless
  groups of (B, T, density, radius, tau, alpha) (:func:`groups`) scaled to
  [-1, 1] over fixed ranges, standardised targets (:func:`net_targets`):
  asinh(sigma_xx / S0), log-like over the decades sigma_xx spans and finite
  where a material conducts nothing (every start position inside an impurity
  gives sigma = 0 exactly, at every B), and the Hall ratio sigma_xy / sigma_xx,
  which the error measure below sees directly and which crosses zero smoothly
  where sigma_xy changes sign.
* :func:`solver_runs` — sigma = coherent + incoherent pass
  (SimulationRunner.cpp:161-173) from the batched GPU solver, for material
  draws; the surrogate's training data and the sweep's solver calls.
* :func:`adaptive_sweep` — a B curve per material from few solver points:
  bisection of the gaps between solver points, each new point a check of the
  interpolant, random solver checks before a curve is accepted; the
  surrogate, when given, resolves gaps where it predicts the interpolant is
  already right.  With ``surrogate=None`` it is the method without the
  learned part.

Error of a curve point: max(|d sigma_xx|, |d sigma_xy|) / |sigma|, with
|sigma| = sqrt(sigma_xx^2 + sigma_xy^2) of the FP64 solver (the relative
error of the conductivity tensor; sigma_xy alone can sit near zero, where
the solver's own jitter exceeds any relative bound on it).
"""
import ctypes as C
import math
import time

import numpy as np

from . import ES_OK, Batch, EsError, load_library

NPARAM = 20752
N_IN = 6

# Feature ranges: B 0.1-60 T, T 0-150 K, density 1e14-1.5e16 m^-2,
# radius 0.5e-8-2e-8 m, tau 1e-13-1e-12 s, alpha 0.2-pi/4.
RANGES = dict(B=(0.1, 60.0), T=(0.0, 150.0), density=(1e14, 1.5e16), radius=(0.5e-8, 2e-8),
              tau=(1e-13, 1e-12), alpha=(0.2, math.pi / 4))


class MlpDesc(C.Structure):
    _fields_ = [("members", C.c_int32), ("device", C.c_int32), ("seed", C.c_uint64)]


class MlpTrainDesc(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("lr_final", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("out_weight", C.c_float * 2),
                ("shuffle_seed", C.c_uint64)]


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """Ensemble MLP on the GPU (include/es_surrogate.h)."""

    def __init__(self, members=4, seed=1, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        d = MlpDesc(int(members), int(device), int(seed))
        rc = self.lib.es_mlp_create(C.byref(d), C.byref(h))
        if rc != ES_OK:
            raise EsError(rc, self.lib.es_mlp_last_error(None).decode())
        self.h = h
        self.members = int(members)

    def _check(self, rc):
        if rc != ES_OK:
            raise EsError(rc, self.lib.es_mlp_last_error(self.h).decode())

    @property
    def params(self):
        p = np.empty((self.members, NPARAM), dtype=np.float32)
        self._check(self.lib.es_mlp_get_params(self.h, _fp(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(self.members, NPARAM)
        self._check(self.lib.es_mlp_set_params(self.h, _fp(p)))

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        assert x.shape[1] == N_IN and t.shape == (x.shape[0], 2)
        self._check(self.lib.es_mlp_set_data(self.h, _fp(x), _fp(t), x.shape[0]))

    def gradients(self, rows, out_weight=(1.0, 1.0)):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        ow = np.ascontiguousarray(out_weight, dtype=np.float32)
        g = np.empty((self.members, NPARAM), dtype=np.float32)
        loss = np.empty(self.members, dtype=np.float32)
        self._check(self.lib.es_mlp_gradients(self.h, rows.ctypes.data_as(C.POINTER(C.c_int32)), len(rows), _fp(ow),
                                              _fp(g), _fp(loss)))
        return g, loss

    def train(self, epochs, batch=256, lr=2e-3, lr_final=1e-5, out_weight=(1.0, 1.0), shuffle_seed=7,
              beta1=0.9, beta2=0.999, eps=1e-8):
        d = MlpTrainDesc(int(epochs), int(batch), lr, lr_final, beta1, beta2, eps, (C.c_float * 2)(*out_weight),
                         int(shuffle_seed))
        hist = np.empty((int(epochs), self.members), dtype=np.float32)
        self._check(self.lib.es_mlp_train(self.h, C.byref(d), _fp(hist)))
        return hist

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((self.members, x.shape[0], 2), dtype=np.float32)
        self._check(self.lib.es_mlp_predict(self.h, _fp(x), x.shape[0], _fp(y)))
        return y

    def last_train_time(self):
        s = C.c_double()
        n = C.c_int64()
        self._check(self.lib.es_mlp_last_train_time(self.h, C.byref(s), C.byref(n)))
        return s.value, n.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.es_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- features

# The solver's constants (es_abi.hip): electron charge, effective mass 5 m0,
# hbar / k_B, and the particle speed of solver_runs.
E_CHARGE = 1.60217646e-19
M_EFF = 5.0 * 9.10938188e-31
HBAR_OVER_KB = 1.05457148e-34 / 1.380649e-23
SPEED = 7e5
T_FLOOR = 1.0  # temperatures below 1 K are taken as 1 K (draw_materials floors T there)


def groups(B, T, density, radius, tau, alpha, speed=SPEED):
    """The six inputs as the dimensionless groups the physics runs on:
    log(omega tau) and log(omega tau_T) (cyclotron frequency omega = e B / m
    against the coherent lifetime tau and the thermal one tau_T = hbar / (k_B
    T), Simulation.h:90), log(l / R_c) (mean free path between disks l = 1 /
    (2 r n) against the orbit radius R_c = v / omega), log(n r^2) (the share of
    the plane the disks cover, which sets the free start positions), log r and
    alpha.  Invertible for B, T, n, r, tau > 0."""
    B, T, n, r, tau, al = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64)
                                                 for v in (B, T, density, radius, tau, alpha)))
    w = E_CHARGE * B / M_EFF
    tT = HBAR_OVER_KB / np.maximum(T, T_FLOOR)
    ell_over_rc = w / (2.0 * r * n * speed)
    return np.stack([np.log(w * tau), np.log(w * tT), np.log(ell_over_rc), np.log(n * r * r), np.log(r), al], axis=-1)


def _group_ranges():
    """Each group's range over the RANGES box (every group is monotone in each input)."""
    import itertools

    R = RANGES
    corners = np.array(list(itertools.product(R["B"], (T_FLOOR, R["T"][1]), R["density"], R["radius"], R["tau"],
                                              R["alpha"])))
    g = groups(*corners.T)
    return g.min(axis=0), g.max(axis=0)


_GROUP_LO, _GROUP_HI = _group_ranges()


def features(B, T, density, radius, tau, alpha):
    """The network's inputs: groups() scaled to [-1, 1] over the RANGES box."""
    g = groups(B, T, density, radius, tau, alpha)
    return 2.0 * (g - _GROUP_LO) / (_GROUP_HI - _GROUP_LO) - 1.0


S0 = 1e-7  # target scale: asinh is logarithmic well above it, linear below


def targets(sig, scale=S0):
    """asinh(sigma / scale) of sigma [..., 2] (both components)."""
    return np.arcsinh(np.asarray(sig, dtype=np.float64) / scale)


def from_targets(y, scale=S0):
    return scale * np.sinh(np.asarray(y, dtype=np.float64))


def net_targets(sig):
    """The network's targets of sigma [..., 2]: (asinh(sigma_xx / S0), sigma_xy / sigma_xx),
    the ratio 0 where sigma_xx = 0.  The tensor error of a prediction is
    about max(|d sigma_xx| / sigma_xx, |d ratio|) where sigma_xy << sigma_xx."""
    sig = np.asarray(sig, dtype=np.float64)
    xx, xy = sig[..., 0], sig[..., 1]
    ok = xx != 0.0
    return np.stack([np.arcsinh(xx / S0), np.where(ok, xy / np.where(ok, xx, 1.0), 0.0)], axis=-1)


def net_sigma(t):
    """sigma [..., 2] of network targets (the inverse of net_targets where sigma_xx != 0)."""
    t = np.asarray(t, dtype=np.float64)
    xx = S0 * np.sinh(t[..., 0])
    return np.stack([xx, xx * t[..., 1]], axis=-1)


# A curve is interpolated in asinh(sigma / s) with s this fraction of the
# smallest |sigma| solved on it (at least S0): a component smaller than that
# (sigma_xy crossing zero) is interpolated linearly, as the error measure
# (relative to |sigma|) sees it, a larger one logarithmically.
LINEAR_FRACTION = 1e-3


# ------------------------------------------------------------- solver runs

def draw_materials(n, seed, alpha=None):
    """n materials over RANGES (density and tau log-uniform, T floored at 1 K)."""
    rng = np.random.default_rng(seed)
    R = RANGES
    mats = []
    for i in range(n):
        mats.append(dict(
            density=math.exp(rng.uniform(math.log(R["density"][0]), math.log(R["density"][1]))),
            radius=rng.uniform(*R["radius"]),
            tau=math.exp(rng.uniform(math.log(R["tau"][0]), math.log(R["tau"][1]))),
            T=max(rng.uniform(*R["T"]), 1.0),
            alpha=rng.uniform(*R["alpha"]) if alpha is None else float(alpha),
            seed=int(seed) * 100003 + i + 1,
        ))
    return mats


def solver_runs(materials, fields, temperatures=None, positions_per_row=128, particles_per_quadrant=65,
                side=2.0e-6, clockwise=1, speed=7e5, device=0, with_free=False, precision="fp64", reuse=None):
    """sigma [n_mat][n_B][n_T][2] = coherent + incoherent pass, all materials in one batch.

    ``fields[i]`` / ``temperatures[i]`` are material i's B values and
    temperatures (default: its own T).  Returns (sigma list, device seconds,
    traced work items[, start positions outside every impurity per material,
    escaped fraction per material and field]); the escaped fraction is the
    share of the real particles of both passes whose trace ended without a hit
    (left the grid or closed a full revolution: lifetime 0, lifetime.h:131-146).
    With a dict ``reuse``, the batch is kept in it and cleared for the next
    call (its drawn fields and buffers stay; the caller closes
    ``reuse["batch"]``); otherwise a batch is made and closed per call.
    """
    rs, ext = side / 2.0, side / 4.0
    if reuse is None:
        with Batch(device) as b:
            return _solver_batch(b, materials, fields, temperatures, positions_per_row, particles_per_quadrant,
                                 rs, ext, clockwise, speed, with_free, precision)
    b = reuse.get("batch")
    if b is None:
        b = reuse["batch"] = Batch(device)
    b.clear()
    return _solver_batch(b, materials, fields, temperatures, positions_per_row, particles_per_quadrant, rs, ext,
                         clockwise, speed, with_free, precision)


def _solver_batch(b, materials, fields, temperatures, positions_per_row, particles_per_quadrant, rs, ext, clockwise,
                  speed, with_free, precision):
    b.set_precision(precision)
    for i, m in enumerate(materials):
        Ts = [m["T"]] if temperatures is None else list(temperatures[i])
        kw = dict(seed=m["seed"], region_size=rs, region_extends=ext, impurity_radius=m["radius"],
                  density=m["density"], target_cell_population=4, positions_per_row=positions_per_row,
                  particles_per_quadrant=particles_per_quadrant, fields=list(fields[i]), temperatures=Ts,
                  tau=m["tau"], alpha=m["alpha"], is_clockwise=clockwise, particle_speed=speed)
        j = b.add(coherent=True, **kw)
        b.add(coherent=False, share_field_of=j, **kw)
    items = b.counts()["traced_items"]
    sig, mets, secs = b.run(metrics=with_free)
    out = [sig[2 * i] + sig[2 * i + 1] for i in range(len(materials))]
    if with_free:
        free = [int(mets[2 * i + 1][0]["real_particles"]) // (4 * particles_per_quadrant)
                for i in range(len(materials))]
        esc = [(mets[2 * i]["particles_escaped"] + mets[2 * i + 1]["particles_escaped"]) /
               np.maximum(mets[2 * i]["real_particles"] + mets[2 * i + 1]["real_particles"], 1)
               for i in range(len(materials))]
        return out, secs, items, free, esc
    return out, secs, items


def training_set(n_materials, fields_per_material, temps_per_material, seed, **kw):
    """(raw inputs [N][6] = B, T, density, radius, tau, alpha; sigma [N][2]; device s; items)."""
    rng = np.random.default_rng(seed + 1)
    mats = draw_materials(n_materials, seed)
    lo, hi = (math.log(v) for v in RANGES["B"])
    fields = [np.exp(rng.uniform(lo, hi, fields_per_material)) for _ in mats]
    temps = [np.maximum(rng.uniform(*RANGES["T"], temps_per_material), 1.0) for _ in mats]
    sig, secs, items = solver_runs(mats, fields, temps, **kw)
    X, Y = [], []
    for m, Bs, Ts, s in zip(mats, fields, temps, sig):
        for bi, B in enumerate(Bs):
            for ti, T in enumerate(Ts):
                X.append((B, T, m["density"], m["radius"], m["tau"], m["alpha"]))
                Y.append(s[bi, ti])
    return np.array(X), np.array(Y), secs, items


# --------------------------------------------------------------- surrogate

class Surrogate:
    """Ensemble MLP over physical inputs; predicts sigma and its spread."""

    def __init__(self, members=4, seed=1, device=0):
        self.mlp = MLP(members, seed, device)
        self.mean = np.zeros(2)
        self.std = np.ones(2)

    def fit(self, X, sig, epochs=300, batch=256, lr=2e-3, lr_final=1e-5, shuffle_seed=7):
        f = features(*X.T)
        y = net_targets(sig)
        self.mean = y.mean(axis=0)
        self.std = y.std(axis=0) + 1e-12
        self.mlp.set_data(f, (y - self.mean) / self.std)
        return self.mlp.train(epochs, batch, lr, lr_final, shuffle_seed=shuffle_seed)

    def predict_targets(self, X):
        """Per-member network targets [members][n][2] (net_targets)."""
        y = self.mlp.predict(features(*np.asarray(X).T)).astype(np.float64)
        return y * self.std + self.mean

    def predict_members(self, X):
        """Per-member sigma [members][n][2]."""
        return net_sigma(self.predict_targets(X))

    def predict(self, X):
        """(mean sigma [n][2], spread of the targets [n][2]) over the ensemble."""
        y = self.predict_targets(X)
        return net_sigma(y.mean(axis=0)), y.std(axis=0)

    def close(self):
        self.mlp.close()


def tensor_error(approx, exact):
    """max(|d xx|, |d xy|) / |sigma| per point (0 where both are exactly 0)."""
    approx = np.asarray(approx, dtype=np.float64)
    exact = np.asarray(exact, dtype=np.float64)
    norm = np.hypot(exact[..., 0], exact[..., 1])
    d = np.max(np.abs(approx - exact), axis=-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = d / norm
    return np.where(d == 0.0, 0.0, np.where(norm > 0, e, np.inf))


# ---------------------------------------------------------- adaptive sweep

def _pchip_end(h0, h1, m0, m1):
    """End slope: the three-point formula, set to 0 against the secant's sign
    and capped at 3 m0 where the secants change sign."""
    d = ((2 * h0 + h1) * m0 - h0 * m1) / (h0 + h1)
    flip = np.sign(d) != np.sign(m0)
    cap = (np.sign(m0) != np.sign(m1)) & (np.abs(d) > 3 * np.abs(m0))
    return np.where(flip, 0.0, np.where(cap, 3 * m0, d))


def pchip(x, y, xe):
    """Piecewise cubic Hermite interpolation with Fritsch-Butland slopes, the
    method of scipy.interpolate.PchipInterpolator: at each interior knot the
    weighted harmonic mean of the two secants (0 where they differ in sign or
    one is 0), at the ends the three-point rule of _pchip_end; cubics
    extended past the ends.  x [k] ascending (k >= 3), y [k][...], xe [n]."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    sh = (-1,) + (1,) * (y.ndim - 1)
    h = np.diff(x).reshape(sh)
    m = np.diff(y, axis=0) / h
    d = np.empty_like(y)
    w1 = 2 * h[1:] + h[:-1]
    w2 = h[1:] + 2 * h[:-1]
    flat = (np.sign(m[1:]) != np.sign(m[:-1])) | (m[1:] == 0) | (m[:-1] == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        hm = (w1 / m[:-1] + w2 / m[1:]) / (w1 + w2)
        d[1:-1] = np.where(flat, 0.0, 1.0 / hm)
    d[0] = _pchip_end(h[0], h[1], m[0], m[1])
    d[-1] = _pchip_end(h[-1], h[-2], m[-1], m[-2])
    i = np.clip(np.searchsorted(x, xe, side="right") - 1, 0, len(x) - 2)
    t = (np.asarray(xe, dtype=np.float64) - x[i]).reshape(sh)
    hi, mi, d0, d1 = h[i], m[i], d[i], d[i + 1]
    c2 = (3 * mi - 2 * d0 - d1) / hi
    c3 = (d0 + d1 - 2 * mi) / (hi * hi)
    return y[i] + t * (d0 + t * (c2 + t * c3))


def _curve_from(u, idx, vals, at=None, scale=S0):
    """PCHIP through the solved points in (log B, asinh(sigma / scale)) at the
    grid indices `at` (all by default); vals [len(idx)][...][2]."""
    ue = u if at is None else u[at]
    y = targets(vals, scale)
    if len(idx) == 1:
        return np.repeat(vals[:1], len(ue), axis=0)
    if len(idx) == 2:
        t = (ue - u[idx[0]]) / (u[idx[1]] - u[idx[0]])
        t = np.clip(t, 0.0, 1.0).reshape((-1,) + (1,) * (y.ndim - 1))
        est = y[0] + t * (y[1] - y[0])
    else:
        est = pchip(u[idx], y, ue)
    return from_targets(est, scale)


def _prior_errors(u, idx, prior, gaps, scale=S0):
    """Per gap (a, b) of consecutive solved indices, the largest error over its
    points and the ensemble members of interpolating each member's own curve
    through the solved indices (one interpolation for all members and gaps)."""
    at = np.concatenate([np.arange(a, b + 1) for a, b in gaps])
    starts = np.cumsum([0] + [b - a + 1 for a, b in gaps[:-1]])
    p = np.moveaxis(prior, 0, 1)  # [n][members][2]
    err = tensor_error(_curve_from(u, idx, p[idx], at, scale), p[at]).max(axis=1)
    return np.maximum.reduceat(err, starts)


class _Curve:
    def __init__(self, mat, u, prior):
        self.mat = mat
        self.u = u
        self.prior = prior      # sigma [members][n][2] of the surrogate, or None
        self.solved = {}        # index -> sigma [2]
        self.active = set()     # gaps (a, b) of consecutive solved indices still to resolve
        self.dense = set()      # indices to solve next round where the curve is rough at grid scale
        self.phase = "init"     # init -> refine -> check -> done
        self.pending = []       # [(index, gap or None)] asked of the solver this round
        self.checks = []        # (index, error) of the acceptance checks
        self.accepted = False
        self.skipped = 0        # gaps the surrogate resolved without a solver point
        self.trust = 0          # gaps the surrogate called resolved that the solver then confirmed
        self.vouched = set()    # gaps of this round the surrogate called resolved (while not yet trusted)
        self.bisect_passed = 0  # midpoint checks that found the interpolant already within safety * tol
        self.rough = 0          # points found rough at grid scale
        self.escaped = {}       # index -> escaped fraction of the solver run
        self.free = 0           # start positions outside every impurity (solver work per point)
        self.few_free = False   # solved at every field (fewer than full_below free positions)
        self.reasons = {}       # why points were solved: init, bisect, fill, dense, few_free, check

    def count(self, reason, k):
        if k:
            self.reasons[reason] = self.reasons.get(reason, 0) + k

    def scale(self):
        """The interpolation scale of this curve (LINEAR_FRACTION)."""
        if not self.solved:
            return S0
        m = min(float(np.hypot(*v)) for v in self.solved.values())
        return max(S0, LINEAR_FRACTION * m)

    def estimate(self):
        idx = np.array(sorted(self.solved))
        vals = np.array([self.solved[i] for i in idx])
        est = _curve_from(self.u, idx, vals, scale=self.scale())
        est[idx] = vals
        return est

    def grow(self, i, width, n):
        """Solve every unsolved index within `width` of i next round."""
        self.dense |= {k for k in range(max(0, i - width), min(n, i + width + 1)) if k not in self.solved}


def adaptive_sweep(materials, B, surrogate=None, tol=1e-3, n_initial=13, safety=0.15, kappa=0.1, n_checks=3,
                   dense=2, grow=8, rough=0.15, top_probe=True, full_below=2000, escape_max=0.01, trust=2,
                   max_rounds=60, seed=11, vouch_width=20, **solver_kw):
    """Builds every material's sigma curve on the grid B; returns (curves [n_mat][n_B][2], report).

    Each curve starts from n_initial solver points spread evenly over the
    grid and is interpolated with PCHIP in (log B, asinh(sigma / s)), s per curve (LINEAR_FRACTION).  Every
    gap between consecutive solver points is resolved by a solver run at its
    midpoint, which is also a check: the gap is done when the interpolant
    missed that point by at most safety * tol, otherwise both halves go on; a
    gap of at most 2 * dense grid steps is solved at every point instead.
    With a surrogate, a gap is done without a solver run when the curve of
    every ensemble member, interpolated through the same points, stays within
    kappa * tol of that member over the gap: the surrogate predicts where
    interpolation fails, and members that disagree on the curve's shape (its
    uncertainty) keep the solver run.  It does so only where the solver runs
    at both ends of the gap had at most escape_max of their particles end
    without a hit: such orbits make sigma step between neighbouring fields,
    which a smooth surrogate cannot foresee.  And only on gaps of at most
    `vouch_width` grid steps: the first gaps (about 30 steps) are always
    bisected, since the solver's sigma can step by about tol between two
    neighbouring fields even where no orbit escapes, and only a solver point
    inside a gap sees such a step.  And only on a curve where the
    solver has already confirmed `trust` of its verdicts: until then the gaps
    it calls resolved are bisected all the same, each midpoint that then lies
    within safety * tol counts for it, and one that does not ends its use on
    that curve (a network whose curves are smooth but of the wrong shape is
    caught so; a feature narrower than the gaps of the first rounds is not).

    The solver's sigma is not smooth everywhere: with few free start
    positions, or where many orbits close a full revolution without a hit
    (lifetime 0, lifetime.h:131-146), one grid step of B can change it by
    more than tol.  No interpolant predicts that, so such stretches are solved
    at every point: a solved point whose two grid neighbours are solved and
    which the local interpolant through the solved points beside it misses by
    more than rough * tol of |sigma| marks the curve rough there, and every unsolved point within `grow` steps of it is solved
    next round (which repeats until the stretch's ends are smooth); the
    surrogate's curves are smooth, so it no longer resolves gaps of a rough
    curve.  The first round also solves the two fields below the last one
    (`top_probe`): on sparse fields the roughness sits at high B, where orbits
    are small.  A material with fewer than `full_below` free start positions
    is solved at every field: its steps are large (few particles carry sigma)
    and its runs cheap (work is proportional to the free positions).

    A curve is accepted once n_checks solver runs at randomly drawn unsolved
    points all lie within tol of it; a failed check reopens the gaps beside it
    and solves its neighbourhood.  Each round runs the solver once, in one
    batch, for every point any unfinished curve asks for.  A material whose
    start positions all lie inside impurities has sigma = 0 at every B (that
    test does not depend on B): its curve is zero after the first round.
    """
    rng = np.random.default_rng(seed)
    B = np.asarray(B, dtype=np.float64)
    u = np.log(B)
    n = len(B)
    priors = [None] * len(materials)
    if surrogate is not None and len(materials):
        # every material's curve from every member, in one inference launch
        X = np.concatenate([np.stack([B, np.full(n, m["T"]), np.full(n, m["density"]), np.full(n, m["radius"]),
                                      np.full(n, m["tau"]), np.full(n, m["alpha"])], axis=-1) for m in materials])
        P = surrogate.predict_members(X)  # [members][n_mat * n][2]
        priors = [P[:, k * n:(k + 1) * n] for k in range(len(materials))]
    curves = [_Curve(m, u, pr) for m, pr in zip(materials, priors)]
    init = [int(i) for i in np.unique(np.round(np.linspace(0, n - 1, max(2, n_initial))).astype(int))]
    reuse = solver_kw.pop("reuse", None)
    own = reuse is None
    reuse = {} if own else reuse
    if top_probe and n >= 4:  # two more points beside the last: a second difference at the high-field end
        init = sorted(set(init) | {n - 3, n - 2})
    opts = dict(tol=tol, safety=safety, kappa=kappa, n_checks=n_checks, dense=dense, grow=grow, rough=rough,
                full_below=full_below, escape_max=escape_max, trust=trust, vouch_width=vouch_width)
    try:
        return _sweep_rounds(curves, B, u, n, init, rng, max_rounds, opts, dict(solver_kw, reuse=reuse))
    finally:
        if own and reuse.get("batch") is not None:
            reuse["batch"].close()


def _pchip_rows(x, y, xe):
    """pchip for many small interpolation problems of the same size at once:
    row r interpolates (x[r], y[r]) at xe[r]; x [m][k] ascending (k >= 3),
    y [m][k][2], xe [m].  The same operations as pchip, row by row."""
    h = np.diff(x, axis=1)[:, :, None]
    m = np.diff(y, axis=1) / h
    d = np.empty_like(y)
    w1 = 2 * h[:, 1:] + h[:, :-1]
    w2 = h[:, 1:] + 2 * h[:, :-1]
    flat = (np.sign(m[:, 1:]) != np.sign(m[:, :-1])) | (m[:, 1:] == 0) | (m[:, :-1] == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        hm = (w1 / m[:, :-1] + w2 / m[:, 1:]) / (w1 + w2)
        d[:, 1:-1] = np.where(flat, 0.0, 1.0 / hm)
    d[:, 0] = _pchip_end(h[:, 0], h[:, 1], m[:, 0], m[:, 1])
    d[:, -1] = _pchip_end(h[:, -1], h[:, -2], m[:, -1], m[:, -2])
    r = np.arange(x.shape[0])
    i = np.clip((x <= xe[:, None]).sum(axis=1) - 1, 0, x.shape[1] - 2)
    t = (xe - x[r, i])[:, None]
    hi, mi, d0, d1 = h[r, i], m[r, i], d[r, i], d[r, i + 1]
    c2 = (3 * mi - 2 * d0 - d1) / hi
    c3 = (d0 + d1 - 2 * mi) / (hi * hi)
    return y[r, i] + t * (d0 + t * (c2 + t * c3))


def _rough_points(c, points, n, tol_r):
    """Indices among `points` (and their grid neighbours) with both grid
    neighbours solved whose value the local interpolant (PCHIP through up to 3
    solved points on each side, the point itself left out) misses by more than
    tol_r of |sigma|: smooth curvature is followed at this spacing, a step of
    the solver's sigma between neighbouring fields is not.  The candidates'
    interpolations run together, grouped by the size of their windows."""
    cand = sorted({k for i in points for k in (i - 1, i, i + 1)
                   if 0 < k < n - 1 and k in c.solved and k - 1 in c.solved and k + 1 in c.solved})
    if not cand:
        return []
    idx = np.array(sorted(c.solved))
    vals = np.array([c.solved[i] for i in idx])
    sc = c.scale()
    y = targets(vals, sc)
    cand = np.array(cand)
    pos = np.searchsorted(idx, cand)
    nl = pos - np.maximum(pos - 3, 0)
    nr = np.minimum(pos + 4, len(idx)) - pos - 1
    est = np.empty((len(cand), 2))
    for a, b in set(zip(nl.tolist(), nr.tolist())):
        sel = (nl == a) & (nr == b)
        win = pos[sel][:, None] + np.concatenate([np.arange(-a, 0), np.arange(1, b + 1)])[None, :]
        X, Y, xe = c.u[idx[win]], y[win], c.u[cand[sel]]
        if a + b == 2:  # as _curve_from for two points: linear, clamped
            t = np.clip((xe - X[:, 0]) / (X[:, 1] - X[:, 0]), 0.0, 1.0)[:, None]
            e = Y[:, 0] + t * (Y[:, 1] - Y[:, 0])
        else:
            e = _pchip_rows(X, Y, xe)
        est[sel] = from_targets(e, sc)
    err = tensor_error(est, vals[pos])
    return [int(k) for k, e in zip(cand, err) if e > tol_r]


def _sweep_rounds(curves, B, u, n, init, rng, max_rounds, o, solver_kw):
    """The rounds of adaptive_sweep; every solver call shares one batch (solver_kw["reuse"])."""
    tol, safety, kappa = o["tol"], o["safety"], o["kappa"]
    solver_s = 0.0
    items = 0
    rounds = 0
    t0 = time.time()
    while rounds < max_rounds:
        # What every unfinished curve asks of the solver this round.
        ask = []
        for ci, c in enumerate(curves):
            c.pending = []
            if c.phase == "done":
                continue
            if c.phase == "init":
                c.pending = [(i, None) for i in init]
                c.count("init", len(init))
            elif c.phase == "refine":
                gaps = sorted(c.active)
                # the surrogate may close a gap only where few orbits end without a hit
                calm = [g for g in gaps if max(c.escaped.get(g[0], 1.0), c.escaped.get(g[1], 1.0)) <= o["escape_max"]
                        and g[1] - g[0] <= o["vouch_width"]]
                if c.prior is not None and calm:
                    idx = np.array(sorted(c.solved))
                    fine = _prior_errors(u, idx, c.prior, calm, c.scale()) <= kappa * tol
                    if c.trust >= o["trust"]:
                        c.skipped += int(np.count_nonzero(fine))
                        closed = {g for g, f in zip(calm, fine) if f}
                        gaps = [g for g in gaps if g not in closed]
                    else:  # not yet trusted on this curve: its verdicts are checked by the solver first
                        c.vouched = {g for g, f in zip(calm, fine) if f}
                want = {i: None for i in c.dense if i not in c.solved}
                c.count("few_free" if c.few_free else "dense", len(want))
                for a, b in gaps:
                    if b - a <= 2 * o["dense"]:
                        fill = [k for k in range(a + 1, b) if k not in want]
                        want.update({k: None for k in fill})
                        c.count("fill", len(fill))
                    else:
                        if (a + b) // 2 not in want:
                            c.count("bisect", 1)
                        want[(a + b) // 2] = (a, b)
                c.pending = sorted(want.items())
                c.active, c.dense = set(), set()
                if not c.pending:
                    c.phase = "check"
            if c.phase == "check":
                free = np.array([i for i in range(n) if i not in c.solved])
                if free.size == 0:
                    c.phase, c.accepted = "done", True
                    continue
                pts = rng.choice(free, size=min(o["n_checks"], free.size), replace=False)
                c.pending = [(int(i), None) for i in sorted(pts)]
                c.count("check", len(pts))
            if c.pending:
                ask.append(ci)
        if not ask:
            break
        rounds += 1
        sig, secs, it, free_pos, escaped = solver_runs([curves[ci].mat for ci in ask],
                                                       [B[[i for i, _ in curves[ci].pending]] for ci in ask],
                                                       with_free=True, **solver_kw)
        solver_s += secs
        items += it
        for ci, s, nf, ef in zip(ask, sig, free_pos, escaped):
            c = curves[ci]
            vals = s[:, 0, :]
            c.escaped.update({i: float(e) for (i, _), e in zip(c.pending, ef)})
            if c.phase == "init":
                for (i, _), v in zip(c.pending, vals):
                    c.solved[i] = v
                if nf == 0:
                    c.phase, c.accepted = "done", True
                    continue
                c.free = int(nf)
                if nf < o["full_below"]:
                    # few free start positions: large steps, cheap runs; every field is solved
                    c.dense = {k for k in range(n) if k not in c.solved}
                    c.few_free = True
                    c.prior = None
                s_idx = sorted(c.solved)
                c.active = {(a, b) for a, b in zip(s_idx[:-1], s_idx[1:]) if b - a > 1}
                c.phase = "refine"
                for k in _rough_points(c, s_idx, n, o["rough"] * tol):
                    c.rough += 1
                    c.grow(k, o["grow"], n)
                    c.prior = None
                continue
            est = c.estimate()
            errs = tensor_error(est[[i for i, _ in c.pending]], vals)
            for (i, _), v in zip(c.pending, vals):
                c.solved[i] = v
            new = [i for i, _ in c.pending]
            rough = _rough_points(c, new, n, o["rough"] * tol)
            c.rough += len(rough)
            for k in rough:
                c.grow(k, o["grow"], n)
            if rough:
                c.prior = None  # the surrogate's curves are smooth: it cannot vouch for a rough one
            if c.phase == "refine":
                for (i, gap), e in zip(c.pending, errs):
                    if gap is not None and e <= tol * safety:
                        c.bisect_passed += 1
                    if gap in c.vouched and c.prior is not None:
                        # the surrogate said this gap needed no solver point: right or wrong?
                        if e <= tol * safety:
                            c.trust += 1
                        else:
                            c.prior = None
                    if gap is not None and e > tol * safety:
                        a, b = gap
                        c.active |= {g for g in ((a, i), (i, b)) if g[1] - g[0] > 1}
                c.vouched = set()
                if not c.active and not c.dense:
                    c.phase = "check"
            else:  # check
                c.checks += [(i, float(e)) for (i, _), e in zip(c.pending, errs)]
                if np.all(errs <= tol) and not c.dense:
                    c.phase, c.accepted = "done", True
                else:
                    s_idx = sorted(c.solved)
                    for (i, _), e in zip(c.pending, errs):
                        if e > tol:
                            c.grow(i, o["grow"], n)
                            k = s_idx.index(i)
                            for g in ((s_idx[k - 1], i) if k > 0 else None,
                                      (i, s_idx[k + 1]) if k + 1 < len(s_idx) else None):
                                if g and g[1] - g[0] > 1:
                                    c.active.add(g)
                    c.phase = "refine" if (c.active or c.dense) else "check"
    est = np.stack([c.estimate() for c in curves])
    report = dict(
        rounds=rounds, curves=len(curves), accepted=int(sum(c.accepted for c in curves)),
        solver_points=int(sum(len(c.solved) for c in curves)),
        points_per_curve=float(np.mean([len(c.solved) for c in curves])),
        gaps_skipped=int(sum(c.skipped for c in curves)),
        bisect_passed=int(sum(c.bisect_passed for c in curves)),
        surrogate_trusted=int(sum(c.trust >= o["trust"] and c.prior is not None for c in curves)),
        rough_points=int(sum(c.rough for c in curves)),
        max_check_error=float(max((e for c in curves for _, e in c.checks), default=0.0)),
        solver_seconds=solver_s, traced_items=int(items), wall_seconds=time.time() - t0,
    )
    # Solver points and work by why they were asked for (work per point is
    # proportional to the material's free start positions).
    names = ("init", "bisect", "fill", "dense", "few_free", "check")
    pts = {r: int(sum(c.reasons.get(r, 0) for c in curves)) for r in names}
    wk = {r: float(sum(c.reasons.get(r, 0) * c.free for c in curves)) for r in names}
    tot = sum(wk.values()) or 1.0
    report["points_by_reason"] = pts
    report["work_share_by_reason"] = {r: wk[r] / tot for r in names}
    return est, report
