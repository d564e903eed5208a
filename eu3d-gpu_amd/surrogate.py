"""Learned chemistry surrogate: training data, training and use in the solver.

No reference counterpart: the reference's DNN chemistry mode
(include/Global.hpp:16-21, src/Euler.cpp:150-156) dispatches to Euler::DNN
(src/Euler.cpp:1473-1480), which runs no chemistry.  Here the network
(include/eu3d_surrogate.h, csrc/eu3d_mlp.hip) is trained on the device from
states integrated by the library's own FP64 adaptive trapezoid, and it
replaces that integrator only in stiff cells, with per-cell acceptance checks
and an FP64 fallback.

    python surrogate.py train --out input/surrogate-H2_JA919.npz   (needs the GPU)

Everything here drives the C ABI; there is no CPU path for any of it.
"""
import argparse
import ctypes
import os
import re
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(HERE))
    import importlib
    _pkg = importlib.import_module("eu3d-gpu_amd")
else:
    from importlib import import_module
    _pkg = import_module(__package__)

load, ptr, check, EU3DError = _pkg.load, _pkg.ptr, _pkg.check, _pkg.EU3DError
_vp = ctypes.c_void_p

Y_EPS = 1e-12          # ln(Y + Y_EPS) feature
ATOMIC_MASS = {"H": 1.00794, "O": 15.9994, "N": 14.0067, "C": 12.011, "AR": 39.948, "HE": 4.002602}
DEFAULT_MODEL = os.path.join(HERE, "input", "surrogate-H2_JA919.npz")


# ---------------------------------------------------------------------------
# mechanism helpers
# ---------------------------------------------------------------------------
def species_names(mech_path):
    """Species names in file order (the line after NS/NR/TB, as ReactionRead
    reads it, src/Reaction.cpp:156-200)."""
    lines = [l.split() for l in open(mech_path) if l.strip() and not l.lstrip().startswith("!")]
    ns = int(lines[1][0])
    return lines[2][:ns]


def parse_formula(name):
    atoms, s = {}, name.upper()
    for el, n in re.findall(r"(AR|HE|[HONC])(\d*)", s):
        atoms[el] = atoms.get(el, 0) + (int(n) if n else 1)
    if sum(len(el) + len(n) for el, n in re.findall(r"(AR|HE|[HONC])(\d*)", s)) != len(s):
        raise ValueError(f"cannot parse species {name}")
    return atoms


def element_matrix(names):
    """A[e, s] = mass fraction of element e in species s (columns sum to 1)."""
    comp = [parse_formula(n) for n in names]
    elems = sorted({e for c in comp for e in c})
    A = np.zeros((len(elems), len(names)))
    for s, c in enumerate(comp):
        m = sum(ATOMIC_MASS[e] * k for e, k in c.items())
        for e, k in c.items():
            A[elems.index(e), s] = ATOMIC_MASS[e] * k / m
    return A, elems


def projector(A):
    """Orthogonal projector onto {dY : A dY = 0} (element conservation)."""
    ns = A.shape[1]
    return np.eye(ns) - A.T @ np.linalg.solve(A @ A.T, A)


def mixture(phi, names, mw):
    """Unburnt H2-air (O2 : N2 = 1 : 3.76) mass fractions at equivalence ratio phi."""
    idx = {n.upper(): i for i, n in enumerate(names)}
    phi = np.atleast_1d(phi).astype(float)
    Y = np.zeros((phi.size, len(names)))
    Y[:, idx["H2"]] = 2.0 * phi * mw[idx["H2"]]
    Y[:, idx["O2"]] = 1.0 * mw[idx["O2"]]
    Y[:, idx["N2"]] = 3.76 * mw[idx["N2"]]
    return Y / Y.sum(axis=1, keepdims=True)


def burnt(phi, names, mw):
    """Complete-combustion products of mixture(phi) (H2O, N2, leftover H2 or O2)."""
    idx = {n.upper(): i for i, n in enumerate(names)}
    phi = np.atleast_1d(phi).astype(float)
    n = np.zeros((phi.size, len(names)))
    h2, o2 = 2.0 * phi, np.ones_like(phi)
    water = np.minimum(h2, 2.0 * o2)
    n[:, idx["H2O"]] = water
    n[:, idx["H2"]] = h2 - water
    n[:, idx["O2"]] = o2 - water / 2.0
    n[:, idx["N2"]] = 3.76
    Y = n * mw
    return Y / Y.sum(axis=1, keepdims=True)


def sample_states(n, names, mw, seed, T_range=(1000.0, 3000.0), p_atm=(0.05, 50.0), phi=(0.5, 2.0),
                  radical_max=1e-2):
    """Seeded 0-D states: T uniform, p and phi log-uniform, composition a
    random point between unburnt and burnt with trace radicals (H, O, OH,
    log-uniform up to radical_max)."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(*T_range, size=n)
    p = 101325.0 * np.exp(rng.uniform(np.log(p_atm[0]), np.log(p_atm[1]), size=n))
    ph = np.exp(rng.uniform(np.log(phi[0]), np.log(phi[1]), size=n))
    c = rng.uniform(0.0, 1.0, size=n) ** 2
    Y = (1 - c)[:, None] * mixture(ph, names, mw) + c[:, None] * burnt(ph, names, mw)
    idx = {s.upper(): i for i, s in enumerate(names)}
    for r in ("H", "O", "OH"):
        Y[:, idx[r]] += 4 * c * (1 - c) * np.exp(rng.uniform(np.log(1e-7), np.log(radical_max), size=n))
    Y /= Y.sum(axis=1, keepdims=True)
    return T, p, Y


def features(T, p, Y, dt):
    """Raw features as the device builds them: T, ln p, ln(Y + 1e-12), ln dt."""
    n = len(T)
    f = np.empty((n, Y.shape[1] + 3), dtype=np.float32)
    f[:, 0] = T
    f[:, 1] = np.log(p)
    f[:, 2:-1] = np.log(Y + Y_EPS)
    f[:, -1] = np.log(np.broadcast_to(dt, (n,)))
    return f


# ---------------------------------------------------------------------------
# the network
# ---------------------------------------------------------------------------
class Network:
    """eu3d_mlp (include/eu3d_surrogate.h): n_in -> H -> H -> H -> n_out, GELU,
    BF16 MFMA with FP32 accumulation, FP32 master weights, Adam on the device."""

    def __init__(self, n_in, n_out, hidden=256, max_batch=65536, seed=1):
        self.L = load()
        h = _vp()
        rc = self.L.eu3d_mlp_create(n_in, hidden, n_out, max_batch, seed, ctypes.byref(h))
        if rc != 0:
            msg = self.L.eu3d_mlp_last_error(h).decode() if h else ""
            if h:
                self.L.eu3d_mlp_destroy(h)
            raise EU3DError(rc, f"eu3d_mlp_create: {msg}")
        self.h = h
        self.n_in, self.n_out, self.hidden, self.max_batch = n_in, n_out, hidden, max_batch
        self.hull_lo = np.full(n_in, -np.inf, np.float32)
        self.hull_hi = np.full(n_in, np.inf, np.float32)

    def _chk(self, rc, what):
        if rc != 0:
            raise EU3DError(rc, f"{what}: {self.L.eu3d_mlp_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.eu3d_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def param_count(self):
        return int(self.L.eu3d_mlp_param_count(self.h))

    def get_params(self):
        p = np.zeros(self.param_count(), np.float32)
        self._chk(self.L.eu3d_mlp_get_params(self.h, ptr(p)), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.param_count()
        self._chk(self.L.eu3d_mlp_set_params(self.h, ptr(p)), "set_params")

    def set_norm(self, in_mean, in_std, out_mean, out_std):
        a = [np.ascontiguousarray(v, np.float32) for v in (in_mean, in_std, out_mean, out_std)]
        self._chk(self.L.eu3d_mlp_set_norm(self.h, *[ptr(v) for v in a]), "set_norm")

    def get_norm(self):
        a = [np.zeros(self.n_in, np.float32), np.zeros(self.n_in, np.float32), np.zeros(self.n_out, np.float32),
             np.zeros(self.n_out, np.float32)]
        self._chk(self.L.eu3d_mlp_get_norm(self.h, *[ptr(v) for v in a]), "get_norm")
        return a

    def fit_norm(self, x, t):
        """Per-column mean and standard deviation of a training set."""
        x, t = np.asarray(x, np.float64), np.asarray(t, np.float64)
        self.set_norm(x.mean(0), np.maximum(x.std(0), 1e-6), t.mean(0), np.maximum(t.std(0), 1e-12))

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        self._chk(self.L.eu3d_mlp_predict(self.h, ptr(x), x.shape[0], ptr(y)), "predict")
        return y

    def gradients(self, x, t):
        x, t = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.param_count(), np.float32)
        loss = ctypes.c_double()
        self._chk(self.L.eu3d_mlp_gradients(self.h, ptr(x), ptr(t), x.shape[0], ptr(g), ctypes.byref(loss)),
                  "gradients")
        return loss.value, g

    def set_dataset(self, x, t):
        x, t = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(t, np.float32)
        self._chk(self.L.eu3d_mlp_set_dataset(self.h, ptr(x), ptr(t), x.shape[0]), "set_dataset")

    def train(self, epochs, batch, lr=1e-3, lr_end=1e-5, seed=7):
        losses = np.zeros(epochs)
        self._chk(self.L.eu3d_mlp_train(self.h, epochs, batch, lr, lr_end, seed, ptr(losses)), "train")
        return losses

    def save(self, path):
        im, isd, om, osd = self.get_norm()
        np.savez(path, params=self.get_params(), in_mean=im, in_std=isd, out_mean=om, out_std=osd,
                 hull_lo=self.hull_lo, hull_hi=self.hull_hi,
                 shape=np.array([self.n_in, self.hidden, self.n_out], np.int64))

    @classmethod
    def load_file(cls, path, max_batch=65536):
        z = np.load(path, allow_pickle=False)
        n_in, hidden, n_out = (int(v) for v in z["shape"])
        net = cls(n_in, n_out, hidden=hidden, max_batch=max_batch)
        net.set_params(z["params"])
        net.set_norm(z["in_mean"], z["in_std"], z["out_mean"], z["out_std"])
        net.hull_lo, net.hull_hi = z["hull_lo"].astype(np.float32), z["hull_hi"].astype(np.float32)
        return net


# ---------------------------------------------------------------------------
# solver side
# ---------------------------------------------------------------------------
def attach(ctx, net, names, min_nchem=2, tol_T=25.0, tol_elem=0.05, y_floor=1e-4, hull=None):
    """eu3d_set_surrogate with the element projector of the species list and
    the network's training box (or hull=(lo, hi))."""
    L = load()
    A, _ = element_matrix(names)
    P = projector(A)
    ns = len(names)
    prm = _pkg.SurrogateParams()
    prm.min_nchem, prm.tol_T, prm.tol_elem, prm.y_floor = min_nchem, tol_T, tol_elem, y_floor
    lo, hi = hull if hull is not None else (net.hull_lo, net.hull_hi)
    for q in range(32):
        prm.hull_lo[q] = float(lo[q]) if q < len(lo) else 0.0
        prm.hull_hi[q] = float(hi[q]) if q < len(hi) else 0.0
    for s in range(ns):
        for q in range(ns):
            prm.proj[s * 16 + q] = P[s, q]
    check(L.eu3d_set_surrogate(ctx, net.h, ctypes.byref(prm)), ctx, "eu3d_set_surrogate")


def detach(ctx):
    check(load().eu3d_set_surrogate(ctx, None, None), ctx, "eu3d_set_surrogate")


def stats(ctx):
    s = _pkg.SurrogateStats()
    check(load().eu3d_surrogate_stats_get(ctx, ctypes.byref(s)), ctx, "eu3d_surrogate_stats_get")
    return {n: int(getattr(s, n)) for n, _ in s._fields_}


def set_cells(ctx, T, p, Y):
    T, p, Y = (np.ascontiguousarray(a, np.float64) for a in (T, p, Y))
    check(load().eu3d_set_cells(ctx, len(T), ptr(T), ptr(p), ptr(Y)), ctx, "eu3d_set_cells")


def harvest_call(ctx, ns, dt, flags, min_nchem, max_rows):
    """One FP64 chemistry call recording (features, targets) of its stiff cells."""
    x = np.zeros((max_rows, ns + 3), np.float32)
    t = np.zeros((max_rows, ns + 1), np.float32)
    rows = ctypes.c_longlong()
    check(load().eu3d_chemistry_harvest(ctx, dt, flags, min_nchem, max_rows, ptr(x), ptr(t), ctypes.byref(rows)),
          ctx, "eu3d_chemistry_harvest")
    w = min(rows.value, max_rows)
    return x[:w], t[:w], rows.value


def write_box_case(tmpdir, ni, nj, nk, name="samples"):
    """ctrl/grid/input files of a box of ni x nj x nk cells (reference formats)
    used as a container of independent 0-D chemistry samples."""
    inp_dir = os.path.join(HERE, "input")
    os.makedirs(tmpdir, exist_ok=True)
    grid = os.path.join(tmpdir, f"grid-{name}.txt")
    inp = os.path.join(tmpdir, f"input-{name}.txt")
    ctrl = os.path.join(tmpdir, f"ctrl-{name}.txt")
    h = 1e-4
    with open(grid, "w") as f:
        f.write(f"total_ni {ni}\nxa 0.0\nxb {(ni - 1) * h!r}\ntotal_nj {nj}\nya 0.0\nyb {(nj - 1) * h!r}\n"
                f"total_nk {nk}\nza 0.0\nzb {(nk - 1) * h!r}\nbc 2\n")
    with open(inp, "w") as f:
        f.write("P_pre 101325.0\nT_pre 300.0\nM_wave 2.0\nNS 9\nNR 19\n")
    with open(ctrl, "w") as f:
        f.write(f"gridname {grid}\ninitialization {inp}\n"
                f"reaction-model {os.path.join(inp_dir, 'H2_JA919.txt')}\n"
                f"Thermofile {os.path.join(inp_dir, 'thermo.dat')}\n"
                "Final-Time 1.0\ncfl 0.0\ndt 1e-8\ncount 1\ndifference MUSCL_2\ntime-adv EE\n"
                "reaction Trapezoid-Adaptive\nm_block_x 1\nm_block_y 1\nm_block_z 1\nnum_thread 1\n"
                "DLB_step 5\nDLB_tol 0.001\n")
    return ctrl


def ignition_dataset(solver, names, mw, n_states, calls, seed, dt_range=(1e-9, 1e-6), min_nchem=2,
                     T_range=(1000.0, 3000.0), log=None):
    """Harvest (features, targets) from 0-D trajectories: the solver's interior
    cells are filled with seeded states (sample_states), then advanced by
    `calls` FP64 chemistry calls, each with its own log-uniform dt; every call
    records its cells with Nchem >= min_nchem."""
    ctx, ns = solver.ctx, len(names)
    cap = solver.ni * solver.nj * solver.nk
    rng = np.random.default_rng(seed + 1)
    xs, ts = [], []
    for b0 in range(0, n_states, cap):
        n = cap  # every cell holds a sample (n_states rounds up to whole boxes)
        T, p, Y = sample_states(n, names, mw, seed + b0, T_range=T_range)
        set_cells(ctx, T, p, Y)
        for c in range(calls):
            dt = float(np.exp(rng.uniform(np.log(dt_range[0]), np.log(dt_range[1]))))
            x, t, rows = harvest_call(ctx, ns, dt, _pkg.CHEM_COMPACT, min_nchem, cap)
            xs.append(x)
            ts.append(t)
            if log:
                log(f"harvest states {b0}+{n} call {c} dt {dt:.3e} rows {rows}")
    return np.concatenate(xs), np.concatenate(ts)


def train_recipe(out, n_states=1 << 18, calls=16, epochs=60, batch=65536, seed=20240601, log=print):
    """The committed model's recipe: 0-D ignition data, normalization from the
    set, Adam with a cosine learning rate, training box as the hull."""
    t0 = time.time()
    with tempfile.TemporaryDirectory() as d:
        side = 64
        ctrl = write_box_case(d, side, side, max(3, (min(n_states, 1 << 18) + side * side - 1) // (side * side)))
        solver = _pkg.Solver(ctrl, base_dir=d)
        names = species_names(os.path.join(HERE, "input", "H2_JA919.txt"))
        mw = np.array(solver.field("mw"), dtype=float)
        x, t = ignition_dataset(solver, names, mw, n_states, calls, seed, log=None)
        solver.close()
    log(f"harvested {len(x)} rows in {time.time() - t0:.1f} s")
    net = Network(x.shape[1], t.shape[1], max_batch=batch, seed=seed)
    net.fit_norm(x, t)
    net.hull_lo, net.hull_hi = x.min(0), x.max(0)
    net.set_dataset(x, t)
    t1 = time.time()
    losses = net.train(epochs, batch, lr=2e-3, lr_end=2e-5, seed=seed)
    log(f"trained {epochs} epochs in {time.time() - t1:.1f} s, loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    net.save(out)
    return net, losses


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--out", default=DEFAULT_MODEL)
    tr.add_argument("--states", type=int, default=1 << 18)
    tr.add_argument("--calls", type=int, default=16)
    tr.add_argument("--epochs", type=int, default=60)
    a = ap.parse_args()
    if a.cmd == "train":
        train_recipe(a.out, n_states=a.states, calls=a.calls, epochs=a.epochs)


if __name__ == "__main__":
    main()
