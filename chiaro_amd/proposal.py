"""Learned band proposal (chd_proposal): dataset building, training and the
band sweep it drives. No counterpart in the reference, which sweeps every
hypothesis; its output is scored against the plain full sweep.

Network (matching chiaro_amd/csrc/proposal.hip): seven 3x3 convolutions on
a grid of one pixel per coarse node (the reference plane in cells of half the
grid step: the plane at half size, the raster's "quarter resolution", for a
4 px grid): 3->32, 32->32 stride 2, 32->32, 32->32 stride 2, 32->32, up x2 +
32->32, up x2 + 32->3; ReLU after all but the last; BF16 weights and
activations, FP32 accumulation. Outputs per pixel: centre label =
(L - 1) / 2 + (L - 1) / 8 o0 for L labels (64 + 16 o0 at 129), log band
half-width o1, confidence sigmoid(o2).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _abi
from .depth import Context, DepthOptions, Session


# ---------------------------------------------------------------- bf16

def to_bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> BF16 bits (uint16), round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def from_bf16(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32)


def quarter_means(planes, factor: int = 2) -> np.ndarray:
    """(h, w, len(planes)) f32 means over factor x factor cells of up to three
    planes (edges clamped), summed row by row from 0 as the device does."""
    ph, pw = planes[0].shape
    h, w = ph // factor, pw // factor
    out = np.zeros((h, w, min(len(planes), 3)), np.float32)
    for c, p in enumerate(planes[:3]):
        p = np.asarray(p, np.float32)
        s = np.zeros((h, w), np.float32)
        for dy in range(factor):
            ys = np.minimum(factor * np.arange(h) + dy, ph - 1)
            for dx in range(factor):
                xs = np.minimum(factor * np.arange(w) + dx, pw - 1)
                s = s + p[np.ix_(ys, xs)]
        out[:, :, c] = np.float32(1.0 / (factor * factor)) * s
    return out


def quarter_input(planes, offset: float = 0.0, scale: float = 0.125,
                  factor: int = 2) -> np.ndarray:
    """(h, w, 16) BF16 network input from up to three planes (reference
    first): (cell mean - offset) * scale in f32, as the device computes it
    (k_quarter_input); factor 2 gives the plane at half size."""
    m = quarter_means(planes, factor)
    out = np.zeros(m.shape[:2] + (16,), np.float32)
    out[:, :, :m.shape[2]] = (m - np.float32(offset)) * np.float32(scale)
    return to_bf16(out)


# ---------------------------------------------------------------- layers

def layer_specs(width: int, height: int):
    """(cin, cout, hin, win, hout, wout, stride, upsample, relu) per layer."""
    w2, h2 = (width + 1) // 2, (height + 1) // 2
    w4, h4 = (w2 + 1) // 2, (h2 + 1) // 2
    return [
        (3, 32, height, width, height, width, 1, 0, 1),
        (32, 32, height, width, h2, w2, 2, 0, 1),
        (32, 32, h2, w2, h2, w2, 1, 0, 1),
        (32, 32, h2, w2, h4, w4, 2, 0, 1),
        (32, 32, h4, w4, h4, w4, 1, 0, 1),
        (32, 32, h4, w4, h2, w2, 1, 1, 1),
        (32, 3, h2, w2, height, width, 1, 1, 0),
    ]


def _pad16(c):
    return (c + 15) // 16 * 16


# ---------------------------------------------------------------- reference

class ReferenceNet:
    """FP64 numpy restatement of the network's forward and backward passes
    (the checker of the GPU kernels). Parameters use the device layout:
    per layer W [9 * cin_p][cout_p] (k = tap * cin_p + c), b [cout_p]. With
    bf16=True activations are rounded to BF16 after each layer and weights
    are taken at BF16 precision, as on the device."""

    def __init__(self, width, height, params, bf16=True):
        self.specs = layer_specs(width, height)
        self.bf16 = bf16
        self.W, self.b = [], []
        at = 0
        for (cin, cout, *_rest) in self.specs:
            k, cp = 9 * _pad16(cin), _pad16(cout)
            w = params[at:at + k * cp].reshape(k, cp).astype(np.float64)
            at += k * cp
            self.b.append(params[at:at + cp].astype(np.float64))
            at += cp
            if bf16:
                w = from_bf16(to_bf16(w.astype(np.float32))).astype(np.float64)
                mask = np.zeros_like(w)
                for t in range(9):
                    mask[t * _pad16(cin):t * _pad16(cin) + cin, :cout] = 1.0
                w = w * mask
            self.W.append(w)

    def _round(self, x):
        return from_bf16(to_bf16(x.astype(np.float32))).astype(np.float64) if self.bf16 else x

    @staticmethod
    def _source(spec, oy, ox, ky, kx):
        cin, cout, hin, win, hout, wout, stride, up, relu = spec
        if up:
            uy, ux = oy + ky - 1, ox + kx - 1
            ok = (uy >= 0) & (ux >= 0) & (uy < 2 * hin) & (ux < 2 * win)
            return ok, np.clip(uy, 0, None) >> 1, np.clip(ux, 0, None) >> 1
        y, x = oy * stride + ky - 1, ox * stride + kx - 1
        ok = (y >= 0) & (x >= 0) & (y < hin) & (x < win)
        return ok, np.clip(y, 0, hin - 1), np.clip(x, 0, win - 1)

    def _im2col(self, spec, x):
        """x: (n, hin, win, cin_p) -> (n * hout * wout, 9 * cin_p)"""
        cin, cout, hin, win, hout, wout, stride, up, relu = spec
        cp = x.shape[-1]
        oy, ox = np.meshgrid(np.arange(hout), np.arange(wout), indexing="ij")
        cols = np.zeros((x.shape[0], hout, wout, 9, cp))
        for t in range(9):
            ok, iy, ix = self._source(spec, oy, ox, t // 3, t % 3)
            cols[:, :, :, t, :] = x[:, np.minimum(iy, hin - 1), np.minimum(ix, win - 1), :] * ok[None, :, :, None]
        return cols.reshape(-1, 9 * cp)

    def forward(self, x):
        """x: (n, h, w, 16) float -> list of activations; the last is the
        f32 output (n, h, w, 16)."""
        acts = [x.astype(np.float64)]
        for i, spec in enumerate(self.specs):
            cin, cout, hin, win, hout, wout, stride, up, relu = spec
            col = self._im2col(spec, acts[-1])
            y = col @ self.W[i] + self.b[i][None, :]
            y[:, cout:] = 0.0
            if relu:
                y = np.maximum(y, 0.0)
            y = y.reshape(x.shape[0], hout, wout, -1)
            acts.append(self._round(y) if i < 6 else y)
        return acts

    @staticmethod
    def loss_grad(out, labels, label_count: int = 129):
        """Loss (sum) and d loss / d out, normalised by the pixel count, as
        k_proposal_loss computes them (exp/log in f64 here)."""
        o = out.reshape(-1, out.shape[-1])
        t = labels.reshape(-1).astype(np.float64)
        m = t > 0
        d = np.zeros_like(o)
        half, k = (label_count - 1) * 0.5, (label_count - 1) * 0.125
        centre = half + k * o[:, 0]
        s = np.clip(o[:, 1], -3.0, 5.0)
        z = o[:, 2]
        e = np.exp(-s)
        diff = centre - t
        ad = np.abs(diff)
        al = np.minimum(ad, 32.0)
        y = (ad <= 6.0).astype(np.float64)
        sig = 1.0 / (1.0 + np.exp(-z))
        loss = ad * 0.125 + al * e + s + (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
        n = o.shape[0]
        d[:, 0] = np.sign(diff) * (k * 0.125) / n
        d[:, 1] = np.where((o[:, 1] > -3.0) & (o[:, 1] < 5.0), (1.0 - al * e) / n, 0.0)
        d[:, 2] = (sig - y) / n
        d[~m] = 0.0
        return float(loss[m].sum()), d.reshape(out.shape)

    def backward(self, acts, dout):
        """Parameter gradients (same order as the parameters) for
        d loss / d out = dout."""
        grads_w, grads_b = [None] * 7, [None] * 7
        g = dout
        for i in range(6, -1, -1):
            spec = self.specs[i]
            cin, cout, hin, win, hout, wout, stride, up, relu = spec
            g2 = g.reshape(-1, g.shape[-1]).copy()
            if relu:
                g2 = g2 * (acts[i + 1].reshape(-1, g.shape[-1]) > 0)
            if self.bf16:
                g2 = from_bf16(to_bf16(g2.astype(np.float32))).astype(np.float64)
            col = self._im2col(spec, acts[i])
            grads_w[i] = col.T @ g2
            grads_b[i] = g2.sum(axis=0)
            if i > 0:
                dcol = g2 @ self.W[i].T  # (m, 9 cin_p)
                cp = acts[i].shape[-1]
                dx = np.zeros_like(acts[i])
                oy, ox = np.meshgrid(np.arange(hout), np.arange(wout), indexing="ij")
                dcol = dcol.reshape(acts[i].shape[0], hout, wout, 9, cp)
                for t in range(9):
                    ok, iy, ix = self._source(spec, oy, ox, t // 3, t % 3)
                    for n in range(acts[i].shape[0]):
                        np.add.at(dx[n], (iy[ok], ix[ok]), dcol[n][ok][:, t, :])
                dx[..., cin:] = 0.0
                g = dx
        out = []
        for i in range(7):
            out.append(grads_w[i].reshape(-1))
            out.append(grads_b[i])
        return np.concatenate(out)


def adam_reference(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in f32 arithmetic, as k_adam computes it."""
    f = np.float32
    m = f(b1) * m + f(1 - b1) * g
    v = f(b2) * v + f(1 - b2) * g * g
    bc1 = f(1 - np.float32(b1) ** np.float32(step))
    bc2 = f(1 - np.float32(b2) ** np.float32(step))
    w = w - f(lr) * (m / bc1) / (np.sqrt(v / bc2) + f(eps))
    return w.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


# ---------------------------------------------------------------- device net

class Proposal:
    """chd_proposal on one context."""

    def __init__(self, context: Context, width: int, height: int, max_batch: int = 32,
                 seed: int = 0):
        self.context = context
        self.lib = context.lib
        self.width, self.height = width, height
        handle = C.c_void_p()
        _abi.check(self.lib.chd_proposal_create(context.handle, width, height, max_batch, seed,
                                                C.byref(handle)), "proposal create")
        self.handle = handle
        n = C.c_size_t()
        _abi.check(self.lib.chd_proposal_parameters(handle, None, None, C.byref(n)), "parameters")
        self.parameter_count = n.value
        self.input_offset, self.input_scale = 0.0, 0.125
        self.label_count = 129
        context._proposals.add(self)

    def set_label_count(self, label_count: int):
        """Labels of the sessions served (centre = (L-1)/2 + (L-1)/8 o0)."""
        _abi.check(self.lib.chd_proposal_set_label_count(self.handle, label_count), "label count")
        self.label_count = label_count

    def set_input_transform(self, offset: float, scale: float):
        """Inputs are (2x2 mean - offset) * scale; capture_inputs must use the
        same pair for the data given to this network."""
        _abi.check(self.lib.chd_proposal_set_input_transform(self.handle, offset, scale),
                   "input transform")
        o, s = C.c_float(), C.c_float()
        _abi.check(self.lib.chd_proposal_get_input_transform(self.handle, C.byref(o), C.byref(s)),
                   "input transform")
        self.input_offset, self.input_scale = o.value, s.value

    def close(self):
        if self.handle:
            self.lib.chd_proposal_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def parameters(self) -> np.ndarray:
        out = np.empty(self.parameter_count, np.float32)
        n = C.c_size_t(self.parameter_count)
        _abi.check(self.lib.chd_proposal_parameters(self.handle, out.ctypes.data, None, C.byref(n)),
                   "parameters")
        return out

    def set_parameters(self, values: np.ndarray):
        v = np.ascontiguousarray(values, np.float32)
        n = C.c_size_t(self.parameter_count)
        _abi.check(self.lib.chd_proposal_parameters(self.handle, None, v.ctypes.data, C.byref(n)),
                   "set parameters")

    def session_input(self, session) -> np.ndarray:
        """The network input the device builds from a session's planes
        (what apply() feeds the network): (height, width, 16) BF16 bits."""
        out = np.empty((self.height, self.width, 16), np.uint16)
        _abi.check(self.lib.chd_session_network_input(session.handle, self.handle,
                                                      out.ctypes.data), "network input")
        return out

    def optimizer_state(self):
        """Adam moments (first, then second, in parameter order) and the step."""
        n, step = C.c_size_t(), C.c_uint64()
        _abi.check(self.lib.chd_proposal_optimizer_state(self.handle, None, None, C.byref(n),
                                                         C.byref(step)), "optimizer state")
        out = np.empty(n.value, np.float32)
        _abi.check(self.lib.chd_proposal_optimizer_state(self.handle, out.ctypes.data, None,
                                                         C.byref(n), C.byref(step)),
                   "optimizer state")
        return out, int(step.value)

    def set_optimizer_state(self, moments: np.ndarray, step: int):
        v = np.ascontiguousarray(moments, np.float32)
        n, s = C.c_size_t(v.size), C.c_uint64(step)
        _abi.check(self.lib.chd_proposal_optimizer_state(self.handle, None, v.ctypes.data,
                                                         C.byref(n), C.byref(s)),
                   "set optimizer state")

    def save(self, path):
        """Checkpoint: parameters, Adam state, step, input transform and
        label count, as an uncompressed .npz (no pickled objects)."""
        moments, step = self.optimizer_state()
        np.savez(path, parameters=self.parameters(), moments=moments,
                 step=np.array(step, np.uint64), shape=np.array([self.width, self.height]),
                 transform=np.array([self.input_offset, self.input_scale], np.float32),
                 labels=np.array(self.label_count, np.uint32))

    def load(self, path):
        with np.load(path, allow_pickle=False) as f:
            if tuple(f["shape"]) != (self.width, self.height):
                raise ValueError("checkpoint is for another network size")
            self.set_parameters(f["parameters"])
            self.set_optimizer_state(f["moments"], int(f["step"]))
            offset, scale = (float(v) for v in f["transform"])
            self.set_input_transform(offset, scale)
            self.set_label_count(int(f["labels"]))

    def set_dataset(self, inputs: np.ndarray, labels: np.ndarray):
        x = np.ascontiguousarray(inputs, np.uint16)
        y = np.ascontiguousarray(labels, np.uint16)
        _abi.check(self.lib.chd_proposal_set_dataset(self.handle, x.ctypes.data, y.ctypes.data,
                                                     x.shape[0]), "set dataset")

    def train(self, epochs: int, batch: int = 32, lr: float = 1e-3, seed: int = 0,
              chunk: int = 0, log=None) -> list:
        """`epochs` epochs of Adam; with `chunk` > 0 in calls of that many
        epochs (the same run: shuffles and schedule of the whole), calling
        log(epochs done, losses so far) after each."""
        losses = np.zeros(epochs, np.float32)
        step = chunk if chunk > 0 else epochs
        for first in range(0, epochs, step):
            n = min(step, epochs - first)
            part = np.zeros(n, np.float32)
            _abi.check(self.lib.chd_proposal_train_epochs(self.handle, first, n, epochs, batch,
                                                          lr, seed, part.ctypes.data), "train")
            losses[first:first + n] = part
            if log is not None:
                log(first + n, losses[:first + n])
        return [float(v) for v in losses]

    def predict(self, inputs: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(inputs, np.uint16)
        out = np.empty((x.shape[0], self.height, self.width, 3), np.float32)
        _abi.check(self.lib.chd_proposal_predict(self.handle, x.ctypes.data, x.shape[0],
                                                 out.ctypes.data), "predict")
        return out

    def gradients(self, inputs: np.ndarray, labels: np.ndarray):
        x = np.ascontiguousarray(inputs, np.uint16)
        y = np.ascontiguousarray(labels, np.uint16)
        g = np.empty(self.parameter_count, np.float32)
        loss = C.c_float()
        _abi.check(self.lib.chd_proposal_gradients(self.handle, x.ctypes.data, y.ctypes.data,
                                                   x.shape[0], g.ctypes.data, C.byref(loss)),
                   "gradients")
        return g, loss.value

    def adam_step(self, gradients: np.ndarray, lr: float):
        g = np.ascontiguousarray(gradients, np.float32)
        _abi.check(self.lib.chd_proposal_adam_step(self.handle, g.ctypes.data, lr), "adam step")

    def apply(self, session: Session, threshold: float = 0.5, half_width: int = 8,
              tile: int = 8, report: bool = True, max_half_width: int = 0,
              fallback_half_width: int = 0):
        """Band windows for `session` from its planes (centre +- half_width,
        or +- ceil(4 b) clamped to [half_width, max_half_width] when
        max_half_width is larger; tiles below `threshold` sweep every label,
        or centre +- fallback_half_width when that is > 0); returns
        (fallback tiles, tiles) when report, else None (no
        synchronisation)."""
        full, tiles = C.c_uint32(), C.c_uint32()
        _abi.check(self.lib.chd_session_set_proposal(
            session.handle, self.handle, threshold, half_width, max_half_width,
            fallback_half_width, tile, C.byref(full) if report else None, C.byref(tiles)),
            "set proposal")
        return (full.value, tiles.value) if report else None


    def attach(self, session: Session, threshold: float = 0.5, half_width: int = 8,
               tile: int = 8, max_half_width: int = 0, fallback_half_width: int = 0):
        """apply() now, and keep this network and setting with the session:
        runs and batches with CHD_STAGE_PROPOSE set the session's windows
        from its current planes first (inside a batch's graph). Returns
        (fallback tiles, tiles)."""
        full, tiles = C.c_uint32(), C.c_uint32()
        _abi.check(self.lib.chd_session_attach_proposal(
            session.handle, self.handle, threshold, half_width, max_half_width,
            fallback_half_width, tile, C.byref(full), C.byref(tiles)), "attach proposal")
        return full.value, tiles.value


def constant_proposal(context: Context, width: int, height: int, centre_label: float,
                      label_count: int = 129) -> Proposal:
    """The method without the network at the same hypothesis budget: zero
    weights, so every pixel gets the output biases - one centre label for the
    whole image at full confidence (every tile banded)."""
    p = Proposal(context, width, height, max_batch=1)
    p.set_label_count(label_count)
    params = np.zeros(p.parameter_count, np.float32)
    last = params[-16:]               # biases of the output layer (cout_p = 16)
    last[0] = (centre_label - (label_count - 1) * 0.5) / ((label_count - 1) * 0.125)
    last[2] = 30.0                    # sigmoid(30) = 1 in f32
    p.set_parameters(params)
    return p


def clear(session: Session):
    _abi.check(session.lib.chd_session_clear_proposal(session.handle), "clear proposal")


# ---------------------------------------------------------------- data

def grid_factor(cap) -> int:
    """Plane pixels per network pixel: one network pixel per coarse node."""
    return max(cap.options.grid_step, 4) // 2


def _input_planes(cap):
    planes = [cap.modules[cap.reference_index].luminance]
    planes += [m.luminance for i, m in enumerate(cap.modules) if i != cap.reference_index][:2]
    return planes


def capture_inputs(cap, offset: float = 0.0, scale: float = 0.125) -> np.ndarray:
    """Network input of a capture: reference and first two targets, one
    pixel per coarse node."""
    return quarter_input(_input_planes(cap), offset, scale, grid_factor(cap))


def input_transform(caps):
    """(offset, scale) that standardise the 2x2 means of the captures' input
    planes: mean 0, standard deviation 1 (the planes hold log luminance with
    a large common offset, which BF16 would otherwise spend its precision
    on)."""
    values = []
    for cap in caps:
        values.append(quarter_means(_input_planes(cap), grid_factor(cap)).astype(np.float64).ravel())
    v = np.concatenate(values)
    return float(np.float32(v.mean())), float(np.float32(1.0 / max(v.std(), 1e-6)))


def full_sweep_labels(ctx: Context, cap, width: int, height: int):
    """The plain full sweep's WTA label per network pixel (0 where the node
    was not tested), and the session's grid."""
    sess = Session(ctx, cap.modules, cap.reference_index, cap.options)
    try:
        sess.run(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM)
        field, fillable, best = sess.coarse()
        _, _, _, tested = sess.cost_volume()
        g = sess.grid
        labels = np.where(tested > 0, best, 0).astype(np.uint16).reshape(g.coarse_rows, g.coarse_columns)
        return np.ascontiguousarray(labels[:height, :width])
    finally:
        sess.close()


def build_dataset(ctx: Context, seeds, make_capture, offset: float = 0.0, scale: float = 0.125):
    """(inputs, labels) for the given seeds; make_capture(seed) -> Capture;
    inputs through the input transform (offset, scale)."""
    xs, ys = [], []
    for s in seeds:
        cap = make_capture(s)
        x = capture_inputs(cap, offset, scale)
        h, w = x.shape[:2]
        xs.append(x)
        ys.append(full_sweep_labels(ctx, cap, w, h))
    return np.stack(xs), np.stack(ys)


def agreement(ctx: Context, cap, proposal: Proposal, threshold: float, half_width: int = 8,
              tile: int = 8, steps: int = 3, max_half_width: int = 0):
    """Full sweep vs band sweep on one capture: the fraction of nodes tested
    by the full sweep whose band label is within +-1, the agreement of the
    tested masks, the fallback-tile fraction, and both times (cost + SGM,
    the band time including the network and the window proposal)."""
    stages = _abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM
    sess = Session(ctx, cap.modules, cap.reference_index, cap.options)
    try:
        sess.run(stages)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sess.run(stages)
        ctx.synchronize()
        full_ms = (time.perf_counter() - t0) * 1000 / steps
        _, _, full_best = sess.coarse()
        full_tested = sess.cost_volume()[3]
        full_tiles, tiles = proposal.apply(sess, threshold, half_width, tile,
                                           max_half_width=max_half_width)
        sess.run(stages)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            proposal.apply(sess, threshold, half_width, tile, report=False,
                           max_half_width=max_half_width)
            sess.run(stages)
        ctx.synchronize()
        band_ms = (time.perf_counter() - t0) * 1000 / steps
        _, _, band_best = sess.coarse()
        band_tested = sess.cost_volume()[3]
        clear(sess)
    finally:
        sess.close()
    valid = full_tested > 0
    close = np.abs(band_best.astype(np.int32) - full_best.astype(np.int32)) <= 1
    stats = {
        "agreement": float(close[valid].mean()) if valid.any() else 1.0,
        "tested_agreement": float((band_tested == full_tested).mean()),
        "fallback_fraction": full_tiles / max(tiles, 1),
        "full_ms": full_ms,
        "band_ms": band_ms,
    }
    # where the band disagrees: is the full sweep's label inside the node's
    # window (then the difference comes from SGM paths crossing windows) or
    # outside it (the centre missed)?
    g = sess.grid
    pred = proposal.predict(capture_inputs(cap, proposal.input_offset, proposal.input_scale)[None])[0]
    rows, cols = g.coarse_rows, g.coarse_columns
    r = np.minimum(np.arange(rows), pred.shape[0] - 1)
    c = np.minimum(np.arange(cols), pred.shape[1] - 1)
    centre = np.clip(np.rint(pred[np.ix_(r, c)][..., 0]), 1, g.labels - 1).reshape(-1)
    conf = pred[np.ix_(r, c)][..., 2]
    scale = np.exp(np.clip(pred[np.ix_(r, c)][..., 1], -3.0, 5.0)).reshape(-1)
    width = np.clip(np.ceil(4.0 * scale), half_width, max(half_width, max_half_width))
    tr, tc = np.arange(rows) // tile, np.arange(cols) // tile
    tile_id = (tr[:, None] * ((cols + tile - 1) // tile) + tc[None, :]).reshape(-1)
    means = np.bincount(tile_id, conf.reshape(-1)) / np.bincount(tile_id)
    banded = means[tile_id] >= threshold
    inside = np.abs(full_best.astype(np.int64) - centre) <= width
    wrong = valid & ~close & banded
    stats["banded_nodes"] = float((valid & banded).mean())
    stats["window_covers_label"] = float(inside[valid & banded].mean()) if (valid & banded).any() else 1.0
    stats["disagree_inside_window"] = float((wrong & inside).sum() / max(wrong.sum(), 1))
    return stats


def batch_agreement(ctx: Context, caps, proposal: Proposal, settings, steps: int = 3):
    """batch_agreement_sessions over fresh sessions of `caps`."""
    sessions = [Session(ctx, c.modules, c.reference_index, c.options) for c in caps]
    try:
        return batch_agreement_sessions(ctx, sessions, proposal, settings, steps)
    finally:
        for s in sessions:
            s.close()


def batch_agreement_sessions(ctx: Context, sessions, proposal: Proposal, settings,
                             steps: int = 3, stages: int = _abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM):
    """The sessions run as one batch (chd_batch): the full sweep against the
    band sweep for each (threshold, tile, max_half_width[, half_width[,
    fallback_half_width]]) in `settings` (half_width 8 and full-sweep
    fallback tiles unless given). Per setting: the fraction of
    nodes tested by the full sweep whose band label is within +-1, mask
    agreement, fallback-tile fraction and both times per batch (the band
    time includes the network and the windows of every capture: the
    proposal is attached and runs as the batch's CHD_STAGE_PROPOSE, inside
    its graph). With the
    direct and target stages in `stages` the rows also compare the end
    products: the fraction of fine nodes whose u8 provenance (the
    consistency mask, depth.rs:548-579) and whose per-target decision codes
    (depth.rs:420-487) equal the full sweep's (consistency_mask_agreement,
    decision_agreement). window_coverage is the fraction of the tested nodes
    whose full-sweep label lies in their window (label 0 is always swept),
    mean_window the mean label count of their windows."""
    from .depth import Batch
    end_to_end = bool(stages & _abi.CHD_STAGE_TARGETS)
    batch = Batch(ctx, sessions)
    rows = []

    def products():
        out = []
        for s in sessions:
            dmap, targets = s.results()
            out.append((dmap.provenance.copy(), {t: r.decisions.copy() for t, r in targets.items()}))
        return out

    try:
        for s in sessions:
            clear(s)
        batch.run(stages)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.run(stages)
        ctx.synchronize()
        full_ms = (time.perf_counter() - t0) * 1000 / steps
        full_best = [s.coarse()[2] for s in sessions]
        full_tested = [s.tested() for s in sessions]
        full_products = products() if end_to_end else None
        for setting in settings:
            threshold, tile, max_half = setting[:3]
            half = setting[3] if len(setting) > 3 else 8
            wide = setting[4] if len(setting) > 4 else 0
            fallback = tiles = 0
            for s in sessions:
                f, t = proposal.attach(s, threshold, half, tile, max_half_width=max_half,
                                       fallback_half_width=wide)
                fallback += f
                tiles += t
            # every step runs the network and the windows of each capture
            # inside the batch's graph, before its cost volume
            band_stages = stages | _abi.CHD_STAGE_PROPOSE
            batch.run(band_stages)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                batch.run(band_stages)
            ctx.synchronize()
            band_ms = (time.perf_counter() - t0) * 1000 / steps
            close, valid, mask, covered, width = 0, 0, 0.0, 0, 0
            for s, fb, ft in zip(sessions, full_best, full_tested):
                bb = s.coarse()[2]
                bt = s.tested()
                v = ft > 0
                close += int((np.abs(bb.astype(np.int32) - fb.astype(np.int32)) <= 1)[v].sum())
                valid += int(v.sum())
                mask += float((bt == ft).mean()) / len(sessions)
                # the full sweep's label inside the node's window (label 0
                # is always swept)
                lo, hi = s.band()
                f = fb.astype(np.int64)
                covered += int(((f == 0) | ((f >= lo) & (f <= hi)))[v].sum())
                width += int((hi.astype(np.int64) - lo + 1)[v].sum())
            row = {"threshold": threshold, "tile": tile, "max_half_width": max_half,
                   "half_width": half, "fallback_half_width": wide,
                   "agreement": close / max(valid, 1), "tested_agreement": mask,
                   "window_coverage": covered / max(valid, 1),
                   "mean_window": width / max(valid, 1),
                   "fallback_fraction": fallback / max(tiles, 1), "full_ms": full_ms,
                   "band_ms": band_ms, "speedup": full_ms / band_ms}
            if end_to_end:
                same_prov = total_prov = same_dec = total_dec = 0
                for (fp, fd), (bp, bd) in zip(full_products, products()):
                    same_prov += int((fp == bp).sum())
                    total_prov += fp.size
                    for t in fd:
                        same_dec += int((fd[t] == bd[t]).sum())
                        total_dec += fd[t].size
                row["consistency_mask_agreement"] = same_prov / max(total_prov, 1)
                row["decision_agreement"] = same_dec / max(total_dec, 1)
            rows.append(row)
    finally:
        for s in sessions:
            clear(s)
        batch.close()
    return rows
