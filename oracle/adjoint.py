"""Float64 CPU reference of the differentiable map and its gradient.

TEST INFRASTRUCTURE: only tests/ and bench.py's cpu_baseline leg may use
this module, as the checker of the GPU adjoint -- never as the product path.

The map is the one the GPU differentiable tracker runs (one turn of the
reference's intensity map, tracker.py:521-615 with impedance.py:113-146 and
impedance.py:457-489), with the profile deposited continuously:

  p_t      = D(dt_t): quadratic B-spline on bin centres; with
             u = (dt - cut_left)*inv_bw - 1/2 = k + d, k = round(u), it puts
             (1/2-d)^2/2, 3/4-d^2, (1/2+d)^2/2 in bins k-1, k, k+1 (bins
             outside [0, n_slices) dropped).  The reference's
             smooth_histogram (histogram.cpp:75-132) instead moves the whole
             weight from bin k-1 to bin k+1 when a particle crosses the centre
             of bin k, so it has no derivative there.
  V_t      = -scale * (g (*) p_t)[:ns], g = irfft(Z, n_fft), circular
  e1       = dE + q * sum_b V_t[b] w_b(dt)      the induced kick, with the same
             spline weights as the deposit (linear_interp_kick.cpp:21-71
             interpolates linearly; its slope jumps at every bin centre,
             which leaves first-order errors in finite differences)
  e2       = e1 + (q * sum_j V_j sin(w_j dt + phi_j) + acc)   kick.cpp:16-35
  dt'      = dt + coeff * e2,  dE' = e2           drift.cpp:34-37 ("simple")

``gradient`` runs the adjoint of that map (all states kept in memory: for
small cases only).  Z and its parameter derivatives follow
fast_resonator.cpp:45-71: Z(f) = R / (1 + i Q (f/f_r - f_r/f)), Z(0) = 0.
"""
from __future__ import annotations

import numpy as np

E_CHARGE = 1.602176634e-19


def resonator_Z(freq, R, fr, Q):
    """Sum of resonators (fast_resonator.cpp:45-71), Z[f = 0] = 0."""
    freq = np.asarray(freq, dtype=np.float64)
    Z = np.zeros(len(freq), dtype=np.complex128)
    f = freq[1:]
    for r, f0, q in zip(np.atleast_1d(R), np.atleast_1d(fr), np.atleast_1d(Q)):
        u = f / f0 - f0 / f
        Z[1:] += r / (1.0 + 1j * q * u)
    return Z


def resonator_dZ(freq, R, fr, Q):
    """dZ/dR, dZ/df_r, dZ/dQ per resonator: arrays [n_res, n_f]."""
    freq = np.asarray(freq, dtype=np.float64)
    R, fr, Q = (np.atleast_1d(np.asarray(a, dtype=np.float64)) for a in (R, fr, Q))
    n = len(R)
    dR = np.zeros((n, len(freq)), np.complex128)
    dfr = np.zeros_like(dR)
    dQ = np.zeros_like(dR)
    f = freq[1:]
    for k in range(n):
        u = f / fr[k] - fr[k] / f
        den = 1.0 + 1j * Q[k] * u
        dR[k, 1:] = 1.0 / den
        dZdu = -R[k] * 1j * Q[k] / den ** 2
        dfr[k, 1:] = dZdu * (-f / fr[k] ** 2 - 1.0 / f)
        dQ[k, 1:] = -R[k] * 1j * u / den ** 2
    return dR, dfr, dQ


class DiffMap:
    """The differentiable map on host arrays.

    state: dict with per-turn programs -- omega [n_rf, T+1], coeff [T+1]
    (drift "simple" coefficient t_rev*length_ratio*eta_0/(beta^2 E) of each
    turn), acc [T+1] -- and charge, cut_left, cut_right, n_slices,
    n_fft, bin_size, scale (= charge * e * ratio)."""

    def __init__(self, st: dict):
        self.st = st
        self.ns = int(st["n_slices"])
        self.N = int(st["n_fft"])
        self.cl = float(st["cut_left"])
        self.ibw_cut = self.ns / (float(st["cut_right"]) - self.cl)
        bc = np.asarray(st["bin_centers"], dtype=np.float64)
        i = np.arange(self.ns)
        self.circ = (i[:, None] - i[None, :]) % self.N  # g index of V[i] <- p[j]
        self.q = float(st["charge"])
        self.scale = float(st["scale"])
        self.freq = np.fft.rfftfreq(self.N, float(st["bin_size"]))

    # -- pieces ------------------------------------------------------------
    def spline(self, x):
        """Bins k-1, k, k+1 with weights and u-derivatives (each [3, n])."""
        u = (x - self.cl) * self.ibw_cut - 0.5
        k = np.floor(u + 0.5)
        d = u - k
        k = np.clip(k, -10, self.ns + 10).astype(np.int64)
        l, r = 0.5 - d, 0.5 + d
        bins = np.stack([k - 1, k, k + 1])
        w = np.stack([0.5 * l * l, 0.75 - d * d, 0.5 * r * r])
        dw = np.stack([-l, -2.0 * d, r])
        return bins, w, dw

    def deposit(self, x):
        bins, w, _ = self.spline(x)
        p = np.zeros(self.ns)
        for q in range(3):
            m = (bins[q] >= 0) & (bins[q] < self.ns)
            np.add.at(p, bins[q][m], w[q][m])
        return p

    def wake(self, R, fr, Q):
        Z = resonator_Z(self.freq, R, fr, Q) / float(self.st["bin_size"])
        return np.fft.irfft(Z, self.N)

    def induced(self, p, g):
        return -(self.scale * (g[self.circ] @ p))

    def kick(self, x, Vi):
        """q * sum_b V[b] w_b(x) and its x-derivative, with the deposit's
        spline on bin centres (bins outside [0, n_slices) have V = 0)."""
        bins, w, dw = self.spline(x)
        out = np.zeros_like(x)
        dout = np.zeros_like(x)
        for q in range(3):
            m = (bins[q] >= 0) & (bins[q] < self.ns)
            v = np.where(m, Vi[np.clip(bins[q], 0, self.ns - 1)], 0.0)
            out += v * w[q]
            dout += v * dw[q]
        return self.q * out, self.q * dout * self.ibw_cut, bins, w

    # -- forward -----------------------------------------------------------
    def forward(self, dt, dE, V, phi, R, fr, Q, turns, keep=False):
        st = self.st
        g = self.wake(R, fr, Q)
        x, e = np.array(dt, dtype=np.float64), np.array(dE, dtype=np.float64)
        V, phi = np.atleast_1d(V), np.atleast_1d(phi)
        p = self.deposit(x)
        hist = []
        for t in range(turns):
            Vi = self.induced(p, g)
            if keep:
                hist.append((x.copy(), e.copy(), p.copy(), Vi))
            e1 = e + self.kick(x, Vi)[0]
            w = st["omega"][:, t]
            s = np.zeros_like(x)
            for j in range(len(V)):
                s = s + V[j] * np.sin(w[j] * x + phi[j])
            e = e1 + (self.q * s + st["acc"][t])
            x = x + st["coeff"][t + 1] * e
            p = self.deposit(x)
        return (p, hist, g) if keep else p

    # -- adjoint -----------------------------------------------------------
    def gradient(self, dt, dE, V, phi, R, fr, Q, turns, dprofile):
        """dL/dV_j, dL/dphi_j, dL/dR_k, dL/dfr_k, dL/dQ_k for L with dL/dp_T = dprofile."""
        st = self.st
        V, phi = np.atleast_1d(np.asarray(V, np.float64)), np.atleast_1d(np.asarray(phi, np.float64))
        pT, hist, g = self.forward(dt, dE, V, phi, R, fr, Q, turns, keep=True)
        lx = np.zeros(len(dt))
        le = np.zeros(len(dt))
        lp = np.asarray(dprofile, dtype=np.float64).copy()
        gV = np.zeros(len(V))
        gphi = np.zeros(len(V))
        corr = np.zeros(self.N)
        for t in range(turns - 1, -1, -1):
            x, e, p, Vi = hist[t]
            kv, dkv, bins0, w0 = self.kick(x, Vi)
            e1 = e + kv
            w = st["omega"][:, t]
            sn = [np.sin(w[j] * x + phi[j]) for j in range(len(V))]
            cs = [np.cos(w[j] * x + phi[j]) for j in range(len(V))]
            s = np.zeros_like(x)
            for j in range(len(V)):
                s = s + V[j] * sn[j]
            e2 = e1 + (self.q * s + st["acc"][t])
            coeff = st["coeff"][t + 1]
            x1 = x + coeff * e2
            # deposit adjoint of p_{t+1}
            bins, _, dw = self.spline(x1)
            s3 = np.zeros_like(x1)
            for q in range(3):
                m = (bins[q] >= 0) & (bins[q] < self.ns)
                s3 += np.where(m, lp[np.clip(bins[q], 0, self.ns - 1)], 0.0) * dw[q]
            lx = lx + s3 * self.ibw_cut
            le2 = le + coeff * lx
            ds = np.zeros_like(x)
            for j in range(len(V)):
                ds = ds + V[j] * w[j] * cs[j]
                gV[j] += np.sum(le2 * self.q * sn[j])
                gphi[j] += np.sum(le2 * self.q * V[j] * cs[j])
            lx = lx + le2 * (self.q * ds) + le2 * dkv
            lv = np.zeros(self.ns)
            for q in range(3):
                m = (bins0[q] >= 0) & (bins0[q] < self.ns)
                np.add.at(lv, bins0[q][m], (le2 * self.q * w0[q])[m])
            le = le2
            # induced-voltage adjoint and the impedance correlation
            lp = -(self.scale * (g[self.circ].T @ lv))
            pp = np.zeros(self.N)
            pp[: self.ns] = p
            for m_ in range(self.N):
                corr[m_] += np.dot(lv, pp[(np.arange(self.ns) - m_) % self.N])
        return dict(V=gV, phi=gphi, **self.impedance_grad(R, fr, Q, corr), corr=corr, profile=pT)

    def impedance_grad(self, R, fr, Q, corr):
        dR, dfr, dQ = resonator_dZ(self.freq, R, fr, Q)
        bs = float(self.st["bin_size"])
        out = {}
        for name, dz in (("R", dR), ("fr", dfr), ("Q", dQ)):
            out[name] = np.array([-(self.scale * np.dot(np.fft.irfft(d / bs, self.N), corr))
                                  for d in dz])
        return out
