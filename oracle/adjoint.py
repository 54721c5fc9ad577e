"""Float64 CPU reference of the differentiable map and its gradient.

TEST INFRASTRUCTURE: only tests/ and bench.py's cpu_baseline leg may use
this module, as the checker of the GPU adjoint -- never as the product path.

The map is the one the GPU differentiable tracker runs (one turn of the
reference's intensity map, tracker.py:521-615 with impedance.py:113-146 and
impedance.py:457-489), with the profile deposited continuously:

  p_t      = D(dt_t): cloud-in-cell on bin centres; (dt - cut_left)*inv_bw
             - 1/2 = k + d puts 1-d in bin k and d in bin k+1 (bins outside
             [0, n_slices) dropped).  The reference's smooth_histogram
             (histogram.cpp:75-132) instead moves the whole weight from bin
             k-1 to bin k+1 when a particle crosses the centre of bin k, so
             it has no derivative there.
  V_t      = -scale * (g (*) p_t)[:ns], g = irfft(Z, n_fft), circular
  e1       = dE + (dt*slope_k + intercept_k)      linear_interp_kick.cpp:33-64
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
    bin_centers, n_fft, bin_size, scale (= charge * e * ratio)."""

    def __init__(self, st: dict):
        self.st = st
        self.ns = int(st["n_slices"])
        self.N = int(st["n_fft"])
        self.cl = float(st["cut_left"])
        self.ibw_cut = self.ns / (float(st["cut_right"]) - self.cl)
        bc = np.asarray(st["bin_centers"], dtype=np.float64)
        self.bc = bc
        self.bc0 = bc[0]
        self.ibw_bc = (self.ns - 1) / (bc[-1] - bc[0])
        i = np.arange(self.ns)
        self.circ = (i[:, None] - i[None, :]) % self.N  # g index of V[i] <- p[j]
        self.q = float(st["charge"])
        self.scale = float(st["scale"])
        self.freq = np.fft.rfftfreq(self.N, float(st["bin_size"]))

    # -- pieces ------------------------------------------------------------
    def cic(self, x):
        xf = (x - self.cl) * self.ibw_cut - 0.5
        f = np.floor(xf)
        return f.astype(np.int64), xf - f

    def deposit(self, x):
        k, d = self.cic(x)
        p = np.zeros(self.ns)
        m0 = (k >= 0) & (k < self.ns)
        np.add.at(p, k[m0], 1.0 - d[m0])
        m1 = (k + 1 >= 0) & (k + 1 < self.ns)
        np.add.at(p, k[m1] + 1, d[m1])
        return p

    def wake(self, R, fr, Q):
        Z = resonator_Z(self.freq, R, fr, Q) / float(self.st["bin_size"])
        return np.fft.irfft(Z, self.N)

    def induced(self, p, g):
        return -(self.scale * (g[self.circ] @ p))

    def table(self, V):
        sl = self.q * (V[1:] - V[:-1]) * self.ibw_bc
        ic = (self.q * V[:-1] - self.bc[:-1] * sl) + 0.0
        return sl, ic

    def interp_bin(self, x):
        f = np.floor((x - self.bc0) * self.ibw_bc)
        inn = (f >= 0) & (f < self.ns - 1)
        return np.where(inn, f, 0).astype(np.int64), inn

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
            sl, ic = self.table(Vi)
            if keep:
                hist.append((x.copy(), e.copy(), p.copy(), sl, ic))
            k, inn = self.interp_bin(x)
            e1 = np.where(inn, e + (x * sl[k] + ic[k]), e)
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
            x, e, p, sl, ic = hist[t]
            k, inn = self.interp_bin(x)
            e1 = np.where(inn, e + (x * sl[k] + ic[k]), e)
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
            kk, _ = self.cic(x1)
            l0 = np.where((kk >= 0) & (kk < self.ns), lp[np.clip(kk, 0, self.ns - 1)], 0.0)
            l1 = np.where((kk + 1 >= 0) & (kk + 1 < self.ns), lp[np.clip(kk + 1, 0, self.ns - 1)], 0.0)
            lx = lx + (l1 - l0) * self.ibw_cut
            le2 = le + coeff * lx
            ds = np.zeros_like(x)
            for j in range(len(V)):
                ds = ds + V[j] * w[j] * cs[j]
                gV[j] += np.sum(le2 * self.q * sn[j])
                gphi[j] += np.sum(le2 * self.q * V[j] * cs[j])
            lx = lx + le2 * (self.q * ds)
            lx = lx + np.where(inn, le2 * sl[k], 0.0)
            wt = (x - self.bc[k]) * self.ibw_bc
            lv = np.zeros(self.ns)
            np.add.at(lv, k[inn], (le2 * self.q * (1.0 - wt))[inn])
            np.add.at(lv, k[inn] + 1, (le2 * self.q * wt)[inn])
            le = le2
            # induced-voltage adjoint and the impedance correlation
            lp = -(self.scale * (g[self.circ].T @ lv))
            pp = np.zeros(self.N)
            pp[: self.ns] = p
            for m in range(self.N):
                corr[m] += np.dot(lv, pp[(np.arange(self.ns) - m) % self.N])
        return dict(V=gV, phi=gphi, **self.impedance_grad(R, fr, Q, corr), corr=corr, profile=pT)

    def impedance_grad(self, R, fr, Q, corr):
        dR, dfr, dQ = resonator_dZ(self.freq, R, fr, Q)
        bs = float(self.st["bin_size"])
        out = {}
        for name, dz in (("R", dR), ("fr", dfr), ("Q", dQ)):
            out[name] = np.array([-(self.scale * np.dot(np.fft.irfft(d / bs, self.N), corr))
                                  for d in dz])
        return out
