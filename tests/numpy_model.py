"""Independent float64 numpy models of the epoch (test helpers): one
single-beat epoch (epoch_f64) and run_epoch over beats, batches and epochs
(EpochModel).

Written from the mathematics (all-pass filter per connection, gains, Biot-Savart
projection, truncated gradients), vectorised over connections with the filter
state kept once per (voxel, connection, out). It shares no code or loop
structure with the C++ oracle, so agreement between the two pins the oracle's
FP64 twin.
"""
import numpy as np


def epoch_f64(p, lr, mse_strength=1.0, reg_strength=1.0, threshold=1.01, diff_strength=0.0,
              simple=False, freeze_gains=False, freeze_delays=False):
    # configuration scalars are f32 in the reference (config/algorithm.rs:25-58)
    lr, mse_strength, reg_strength, threshold, diff_strength = (
        float(np.float32(v)) for v in (lr, mse_strength, reg_strength, threshold, diff_strength))
    N, S, T = p.num_voxels, p.num_sensors, p.num_steps
    n = 3 * N
    nb = p.geometry.neighbours.astype(np.int64)  # (N,26)
    has = nb >= 0
    a = p.coefs.astype(np.float64)
    D = p.delays.astype(np.int64)
    G = p.gains.astype(np.float64).reshape(N, 3, 26, 3)  # [v, in, d, out]
    Gm = G * has[:, None, :, None]
    H = p.measurement_matrix[0].astype(np.float64)
    mhat = p.measurements[0].astype(np.float64)
    u = p.control_function.astype(np.float64)
    b = p.control_matrix.astype(np.float64)
    cols = 3 * np.where(has, nb, 0)[:, :, None] + np.arange(3)[None, None, :]  # (N,26,3)
    x = np.zeros((T, n))
    y = np.zeros((N, 26, 3))
    fir = np.zeros_like(y)
    iir = np.zeros_like(y)
    dG = np.zeros((N, 3, 26, 3))
    dA = np.zeros((N, 26))
    m = np.zeros((T, S))
    loss_mse = np.zeros(T)
    loss_mr = np.zeros(T)
    mr_sum = 0.0
    s_mse = mse_strength / S
    D0 = p.initial_delays.astype(np.float64)
    delta5 = (D0 - (D + (1 - a) / (1 + a))) ** 5
    for t in range(T):
        x0 = np.where((D <= t)[:, :, None] & has[:, :, None], x[np.clip(t - D, 0, None)[:, :, None], cols], 0.0)
        x1 = np.where((D < t)[:, :, None] & has[:, :, None], x[np.clip(t - D - 1, 0, None)[:, :, None], cols], 0.0)
        ylast = y
        y = np.where(has[:, :, None], a[:, :, None] * (x0 - y) + x1, 0.0)
        x[t] = np.einsum("vido,vdo->vi", Gm, y).reshape(n) + u[t] * b
        m[t] = H @ x[t]
        r = m[t] - mhat[t]
        e = (H.T @ r).reshape(N, 3)
        xv = x[t].reshape(N, 3)
        ssum = np.abs(xv).sum(axis=1)
        on = ssum > threshold
        f = np.where(on, ssum - threshold, 0.0)
        mr_sum += float(np.sum(f * f))
        mr = f[:, None] * np.where(np.signbit(xv), -1.0, 1.0)
        if not freeze_gains:
            c = e * s_mse + mr * reg_strength  # (N,3)
            dG += c[:, :, None, None] * y[:, None, :, :]
        if not freeze_delays:
            if not simple:
                fir = np.where(has[:, :, None] & (t >= D)[:, :, None], -a[:, :, None] * fir + x0, fir)
                iir = np.where((t >= D)[:, :, None], -a[:, :, None] * iir + ylast, iir)
                sens = fir - iir
                dA += np.einsum("vido,vdo,vi->vd", G, sens, e) * s_mse + 9 * diff_strength * delta5
            else:
                ok = (has & (t >= D))[:, :, None]
                sens = np.where(ok, x0 - ylast, 0.0)
                dA += np.einsum("vido,vdo,vi->vd", G, sens, e) * s_mse + \
                    diff_strength * delta5 * 9 * (has & (t >= D))
        loss_mse[t] = np.sum(r * r) / S
        loss_mr[t] = mr_sum
    loss = reg_strength * loss_mr + loss_mse
    bs = T
    gains = p.gains.astype(np.float64).reshape(N, 3, 26, 3)
    coefs = a.copy()
    if not freeze_gains:
        gains = gains - lr / bs * dG
    delays = D.copy()
    if not freeze_delays:
        coefs = coefs - lr / bs * dA
        mg = float(np.float32(1e-4))
        hi, lo = coefs > 1 - mg, coefs < mg
        coefs = np.where(hi, np.where(delays > 1, 2 * mg, 1 - mg),
                         np.where(lo, np.where(delays < 1000, 1 - 2 * mg, mg), coefs))
        delays = np.where(hi & (delays > 1), delays - 1, np.where(lo & (delays < 1000), delays + 1, delays))
    return dict(delays=delays, system_states=x, measurements=m, d_gains=dG.reshape(n, 78), d_coefs=dA,
                gains=gains.reshape(n, 78), coefs=coefs, loss_mse=loss_mse,
                loss_maximum_regularization=loss_mr, loss=loss,
                ap_outputs_now=np.repeat(y.reshape(N, 78), 3, axis=0))


def _delay_offsets():
    """(26, 3) voxel offsets (dx, dy, dz) of the connection index d: index
    k = 9 (dx+1) + 3 (dy+1) + (dz+1) over the 27 cells of the 3x3x3 block with
    the centre (k = 13) left out."""
    k = np.array([d if d < 13 else d + 1 for d in range(26)])
    return np.stack([k // 9 - 1, (k // 3) % 3 - 1, k % 3 - 1], axis=1)


class EpochModel:
    """Independent float64 numpy model of run_epoch over several beats and
    epochs: beats in an explicit order, update batches of `batch_size` beats
    (0 = all), SGD or Adam with bias correction at the running step, the
    average-delay smoothness term and the difference regulariser.

    What it states, from the mathematics of the reference's epoch:
      * every update batch starts with zero derivatives, zero FIR / IIR filter
        sensitivities and a zero running maximum-regularisation sum; inside a
        batch all three carry over from beat to beat;
      * each beat starts its propagation from rest (states and filter outputs
        zero) and projects with its own measurement matrix and measurements;
      * the update divides the learning rate by T * batch_size (or T * B);
      * the per-step losses are those of the last beat visited, and a batch
        loss is their mean over the steps;
      * the Adam moments and step survive across updates and epochs.
    """

    def __init__(self, p, optimizer=0):
        self.p = p
        N = p.num_voxels
        self.N, self.S, self.T, self.B = N, p.num_sensors, p.num_steps, p.num_beats
        self.nb = p.geometry.neighbours.astype(np.int64)
        self.has = self.nb >= 0
        self.gains = p.gains.astype(np.float64).reshape(N, 3, 26, 3).copy()
        self.coefs = p.coefs.astype(np.float64).copy()
        self.delays = p.delays.astype(np.int64).copy()
        self.D0 = p.initial_delays.astype(np.float64)
        self.optimizer = optimizer
        self.moments = {k: np.zeros_like(v) for k, v in
                        (("gains", self.gains), ("coefs", self.coefs))} if optimizer else None
        self.moments2 = {k: np.zeros_like(v) for k, v in
                         (("gains", self.gains), ("coefs", self.coefs))} if optimizer else None
        self.step = 1
        self.batch_index = 0
        self.loss_batch, self.loss_mse_batch, self.loss_mr_batch = [], [], []

    # -- one beat ----------------------------------------------------------
    def _beat(self, beat, cfg, c):
        p, N, S, T = self.p, self.N, self.S, self.T
        n = 3 * N
        nb, has = self.nb, self.has
        a, D, G = self.coefs, self.delays, self.gains
        Gm = G * has[:, None, :, None]
        H = p.measurement_matrix[beat].astype(np.float64)
        mhat = p.measurements[beat].astype(np.float64)
        u = p.control_function.astype(np.float64)
        b = p.control_matrix.astype(np.float64)
        s_mse = f32(cfg.mse_strength) / S
        s_reg, thr = f32(cfg.maximum_regularization_strength), f32(cfg.maximum_regularization_threshold)
        lam = f32(cfg.difference_regularization_strength)
        cols = 3 * np.where(has, nb, 0)[:, :, None] + np.arange(3)[None, None, :]
        delta5 = (self.D0 - (D + (1 - a) / (1 + a))) ** 5
        x = np.zeros((T, n))
        y = np.zeros((N, 26, 3))
        m = np.zeros((T, S))
        loss_mse, loss_mr = np.zeros(T), np.zeros(T)
        for t in range(T):
            x0 = np.where(((D <= t) & has)[:, :, None], x[np.clip(t - D, 0, None)[:, :, None], cols], 0.0)
            x1 = np.where(((D < t) & has)[:, :, None], x[np.clip(t - D - 1, 0, None)[:, :, None], cols], 0.0)
            ylast = y
            y = np.where(has[:, :, None], a[:, :, None] * (x0 - y) + x1, 0.0)
            x[t] = np.einsum("vido,vdo->vi", Gm, y).reshape(n) + u[t] * b
            m[t] = H @ x[t]
            r = m[t] - mhat[t]
            e = (H.T @ r).reshape(N, 3)
            xv = x[t].reshape(N, 3)
            f = np.maximum(np.abs(xv).sum(axis=1) - thr, 0.0)
            c["mr_sum"] += float(np.sum(f * f))
            mr = f[:, None] * np.where(np.signbit(xv), -1.0, 1.0)
            if not cfg.freeze_gains:
                c["dG"] += (e * s_mse + mr * s_reg)[:, :, None, None] * y[:, None, :, :]
            if not cfg.freeze_delays:
                if int(cfg.ap_derivative) == 0:
                    on = (t >= D)[:, :, None]
                    c["fir"] = np.where(has[:, :, None] & on, -a[:, :, None] * c["fir"] + x0, c["fir"])
                    c["iir"] = np.where(on, -a[:, :, None] * c["iir"] + ylast, c["iir"])
                    sens = c["fir"] - c["iir"]
                    c["dA"] += np.einsum("vido,vdo,vi->vd", G, sens, e) * s_mse + 9 * lam * delta5
                else:
                    ok = has & (t >= D)
                    sens = np.where(ok[:, :, None], x0 - ylast, 0.0)
                    c["dA"] += np.einsum("vido,vdo,vi->vd", G, sens, e) * s_mse + 9 * lam * delta5 * ok
            loss_mse[t] = np.sum(r * r) / S
            loss_mr[t] = c["mr_sum"]
        self.system_states, self.filter_outputs, self.measurements_beat = x, y, m
        self.loss_mse, self.loss_mr = loss_mse, loss_mr
        self.loss = s_reg * loss_mr + loss_mse

    # -- update batch -------------------------------------------------------
    def average_delays(self):
        """Gain-weighted mean of (D + (1-a)/(1+a)) / sqrt(|dx|+|dy|+|dz|);
        None (valid False) where every gain of the voxel is zero."""
        off = _delay_offsets()
        corr = (self.delays + (1 - self.coefs) / (1 + self.coefs)) / np.sqrt(np.abs(off).sum(axis=1))
        w = np.abs(self.gains).sum(axis=(1, 3))  # (N, 26)
        gsum = w.sum(axis=1)
        valid = gsum != 0
        avg = np.where(valid, (w * corr).sum(axis=1) / np.where(valid, gsum, 1.0), 0.0)
        return avg, valid

    def smoothness(self, strength):
        """dA[v, :] += strength * (mean of the valid average delays of v and
        its neighbours - v's own), for voxels with a valid average delay."""
        avg, valid = self.average_delays()
        ok = self.has & valid[np.where(self.has, self.nb, 0)]
        hood = avg + np.where(ok, avg[np.where(self.has, self.nb, 0)], 0.0).sum(axis=1)
        hood = hood / (1.0 + ok.sum(axis=1))
        return np.where(valid, strength * (hood - avg), 0.0)[:, None] * np.ones((1, 26))

    def _update(self, cfg, c, beats_in_batch):
        T = self.T
        bs = T * (cfg.batch_size if cfg.batch_size else self.B)
        lr = f32(cfg.learning_rate)
        self.avg, self.avg_valid = self.average_delays()
        s = f32(cfg.smoothness_regularization_strength)
        if not cfg.freeze_delays and abs(s) > np.finfo(np.float32).eps:
            c["dA"] = c["dA"] + self.smoothness(s)
        for name, d, frozen in (("gains", c["dG"], cfg.freeze_gains),
                                ("coefs", c["dA"], cfg.freeze_delays)):
            if frozen:
                continue
            v = getattr(self, name)
            if self.optimizer:
                b1, b2 = f32(0.9), f32(0.999)
                m1 = self.moments[name] = self.moments[name] * b1 + (1 - b1) * d
                m2 = self.moments2[name] = self.moments2[name] * b2 + (1 - b2) * d * d
                step = m1 / (1 - b1 ** self.step) / (np.sqrt(m2 / (1 - b2 ** self.step)) + f32(1e-8))
            else:
                step = d
            setattr(self, name, v - lr / bs * step)
        if not cfg.freeze_delays:
            mg = f32(1e-4)
            a, Dl = self.coefs, self.delays
            hi, lo = a > 1 - mg, a < mg
            self.coefs = np.where(hi, np.where(Dl > 1, 2 * mg, 1 - mg),
                                  np.where(lo, np.where(Dl < 1000, 1 - 2 * mg, mg), a))
            self.delays = np.where(hi & (Dl > 1), Dl - 1, np.where(lo & (Dl < 1000), Dl + 1, Dl))
        self.step += 1
        self.loss_batch.append(float(np.mean(self.loss)))
        self.loss_mse_batch.append(float(np.mean(self.loss_mse)))
        self.loss_mr_batch.append(float(np.mean(self.loss_mr)))
        self.batch_index += 1

    def run_epoch(self, cfg, order=None):
        order = list(range(self.B)) if order is None else list(order)
        N = self.N

        def fresh():
            return dict(dG=np.zeros((N, 3, 26, 3)), dA=np.zeros((N, 26)),
                        fir=np.zeros((N, 26, 3)), iir=np.zeros((N, 26, 3)), mr_sum=0.0)

        c = fresh()
        in_batch = 0
        for beat in order:
            self._beat(beat, cfg, c)
            if cfg.batch_size:
                in_batch += 1
                if in_batch == cfg.batch_size:
                    self._update(cfg, c, in_batch)
                    c = fresh()
                    in_batch = 0
        if not cfg.batch_size or in_batch > 0:
            self._update(cfg, c, in_batch or self.B)
        # what the derivative arrays hold after the epoch (zero after a full batch)
        self.d_gains, self.d_coefs = c["dG"].reshape(3 * N, 78), c["dA"]
        self.fir, self.iir, self.mr_sum = c["fir"], c["iir"], c["mr_sum"]

    @property
    def gains_rows(self):
        return self.gains.reshape(3 * self.N, 78)


def f32(v):
    """A configuration scalar as the reference holds it (f32), widened."""
    return float(np.float32(v))
