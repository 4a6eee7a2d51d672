"""Independent float64 numpy model of one single-beat epoch (test helper).

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
