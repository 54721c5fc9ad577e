"""Float64 CPU reference of the differentiable map and its gradient.

TEST INFRASTRUCTURE: only tests/ and bench.py's cpu_baseline leg may use
this module, as the checker of the GPU adjoint -- never as the product path.

The map is one turn of the reference's intensity map (tracker.py:521-615
with impedance.py:113-146 and impedance.py:457-489) with a smooth profile,
in one of two forms (``mode``):

``"reference"`` -- the reference's own functions:
  p_t      = smooth_histogram(dt_t) (histogram.cpp:75-132, its quirk kept):
             with f = (dt - cut_left) inv_bw, k = (int) f, d = f - k, bin k
             gets 0.5 - d and bin k2 = (int)(f + 1) (d > 1/2) or (int)(f - 1)
             (d <= 1/2) gets 0.5 + d; particles outside
             [cut_left + bw/2, cut_right - bw/2] are dropped
  V_t      = -scale * (g (*) p_t)[:ns], g = irfft(Z, n_fft), circular
  e1       = linear_interp_kick(dt, dE, V_t, bin_centers, q, 0)
             (linear_interp_kick.cpp:21-71)
  The weights are piecewise linear in dt and jump where a particle crosses a
  bin centre or a bin edge; the interpolated kick is continuous with kinks
  at bin centres.  The gradient uses their derivatives on the pieces
  (d p[k] / d dt = -inv_bw, d p[k2] / d dt = +inv_bw; d kick / d dt = the
  bin's slope, d kick / d V[k], V[k+1] = q (1 - s), q s with
  s = (dt - bc[k]) inv_bw_bc): the pathwise derivative of the loss, exact
  wherever no particle sits on a bin centre or edge.

``"spline"`` -- a variant with continuous first derivatives:
  p_t      = D(dt_t): quadratic B-spline on bin centres; with
             u = (dt - cut_left)*inv_bw - 1/2 = k + d, k = round(u), it puts
             (1/2-d)^2/2, 3/4-d^2, (1/2+d)^2/2 in bins k-1, k, k+1 (bins
             outside [0, n_slices) dropped)
  e1       = dE + q * sum_b V_t[b] w_b(dt), the same spline weights

Both then kick and drift as the reference (kick.cpp:16-35, drift.cpp:34-37
"simple"):
  e2       = e1 + (q * sum_j V_j sin(w_j dt + phi_j) + acc)
  dt'      = dt + coeff * e2,  dE' = e2

The forward pass calls the oracle's C kernels (oracle/liboracle.so) for the
RF kick, the smooth histogram and the interpolated kick, so it is the
reference's arithmetic; the drift coefficient is formed in drift.cpp's
order.  ``gradient`` runs the adjoint of the map (all states kept in memory:
for small cases only).  Z and its parameter derivatives follow
fast_resonator.cpp:45-71: Z(f) = R / (1 + i Q (f/f_r - f_r/f)), Z(0) = 0.
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc

E_CHARGE = 1.602176634e-19
MODES = ("reference", "spline")


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
    turn, formed as drift.cpp:34 forms it), acc [T+1] -- and charge,
    cut_left, cut_right, n_slices, bin_centers, n_fft, bin_size, scale
    (= charge * e * ratio)."""

    def __init__(self, st: dict, mode: str = "reference", lib=None):
        if mode not in MODES:
            raise ValueError(f"mode {mode!r} not in {MODES}")
        self.st = st
        self.mode = mode
        self.lib = lib or orc.oracle_lib()
        self.ns = int(st["n_slices"])
        self.N = int(st["n_fft"])
        self.cl = float(st["cut_left"])
        self.cr = float(st["cut_right"])
        self.ibw_cut = self.ns / (self.cr - self.cl)
        bw = (self.cr - self.cl) / self.ns
        self.lo, self.hi = self.cl + bw * 0.5, self.cr - bw * 0.5  # histogram.cpp:82-85
        self.bc = np.ascontiguousarray(st["bin_centers"], dtype=np.float64)
        self.ibw_bc = (self.ns - 1) / (self.bc[-1] - self.bc[0])  # linear_interp_kick.cpp:33
        i = np.arange(self.ns)
        self.circ = (i[:, None] - i[None, :]) % self.N  # g index of V[i] <- p[j]
        self.q = float(st["charge"])
        self.scale = float(st["scale"])
        self.freq = np.fft.rfftfreq(self.N, float(st["bin_size"]))

    # -- the spline variant's pieces -----------------------------------------
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

    def _spline_kick(self, x, Vi):
        """q * sum_b V[b] w_b(x) and its x-derivative (bins outside
        [0, n_slices) have V = 0)."""
        bins, w, dw = self.spline(x)
        out = np.zeros_like(x)
        dout = np.zeros_like(x)
        for q in range(3):
            m = (bins[q] >= 0) & (bins[q] < self.ns)
            v = np.where(m, Vi[np.clip(bins[q], 0, self.ns - 1)], 0.0)
            out += v * w[q]
            dout += v * dw[q]
        return self.q * out, self.q * dout * self.ibw_cut, bins, w

    # -- the reference's pieces ------------------------------------------------
    def smooth_cells(self, x):
        """histogram.cpp:100-113 per particle: (in range, k, k2) -- bin k gets
        0.5 - d, bin k2 gets 0.5 + d."""
        inside = (x >= self.lo) & (x <= self.hi)
        f = np.where(inside, (x - self.cl) * self.ibw_cut, 1.0)
        k = np.trunc(f).astype(np.int64)
        d = f - k
        k2 = np.where(d > 0.5, np.trunc(f + 1.0), np.trunc(f - 1.0)).astype(np.int64)
        inside &= (k >= 0) & (k < self.ns) & (k2 >= 0) & (k2 < self.ns)
        return inside, k, k2

    def interp_cells(self, x):
        """linear_interp_kick.cpp:56-63 per particle: (kicked, k)."""
        f = np.floor((x - self.bc[0]) * self.ibw_bc)
        hit = (f >= 0) & (f < self.ns - 1)
        return hit, np.where(hit, f, 0).astype(np.int64)

    def cells(self, x):
        """Which piece of the map each particle is on: two of these differ
        exactly when a particle crossed a bin centre or edge (deposit) or a
        bin centre (interpolated kick)."""
        if self.mode == "spline":
            return np.floor((x - self.cl) * self.ibw_cut * 2.0)
        inside, k, k2 = self.smooth_cells(x)
        hit, kk = self.interp_cells(x)
        return np.stack([np.where(inside, 2 * k + (k2 > k), -1), np.where(hit, kk, -1)])

    # -- shared pieces -----------------------------------------------------------
    def deposit(self, x):
        if self.mode == "reference":
            p = np.zeros(self.ns)
            self.lib.slice_smooth(np.ascontiguousarray(x, dtype=np.float64), p, self.cl, self.cr)
            return p
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

    def _induced_kick(self, x, e, Vi):
        """e + the induced kick (in place on a copy)."""
        e = np.array(e, dtype=np.float64, copy=True)
        if self.mode == "reference":
            self.lib.linear_interp_kick(x, e, Vi, self.bc, self.q, 0.0)
            return e
        return e + self._spline_kick(x, Vi)[0]

    def _rf_kick(self, x, e, V, phi, t):
        """kick.cpp:16-35 (the oracle's kernel, the reference's arithmetic)."""
        e = np.ascontiguousarray(e, dtype=np.float64)
        self.lib.kick(x, e, V, self.st["omega"][:, t], phi, self.q, len(V), self.st["acc"][t])
        return e

    # -- forward -----------------------------------------------------------
    def forward(self, dt, dE, V, phi, R, fr, Q, turns, keep=False, coords=False):
        """Final profile after ``turns`` turns; with ``keep`` also the states
        of every turn (x, e, p, V_t) and the wake; with ``coords`` also the
        final coordinates."""
        st = self.st
        g = self.wake(R, fr, Q)
        x = np.array(dt, dtype=np.float64, copy=True)
        e = np.array(dE, dtype=np.float64, copy=True)
        V = np.ascontiguousarray(np.atleast_1d(V), dtype=np.float64)
        phi = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
        p = self.deposit(x)
        hist = []
        for t in range(turns):
            Vi = np.ascontiguousarray(self.induced(p, g))
            if keep:
                hist.append((x.copy(), e.copy(), p.copy(), Vi))
            e = self._induced_kick(x, e, Vi)
            e = self._rf_kick(x, e, V, phi, t)
            x = x + st["coeff"][t + 1] * e
            p = self.deposit(x)
        if keep:
            return p, hist, g
        return (p, x, e) if coords else p

    # -- adjoint -----------------------------------------------------------
    def _induced_adjoint(self, x, Vi):
        """For the reverse turn: d kick / d x per particle, and the particles'
        bins and weights d kick / d V[bin] (rows of [2, n])."""
        if self.mode == "spline":
            _, dkv, bins, w = self._spline_kick(x, Vi)
            return dkv, bins, self.q * w
        hit, k = self.interp_cells(x)
        vk = self.q * (Vi[np.minimum(k + 1, self.ns - 1)] - Vi[k]) * self.ibw_bc
        s = (x - self.bc[k]) * self.ibw_bc
        dkv = np.where(hit, vk, 0.0)
        bins = np.stack([np.where(hit, k, -1), np.where(hit, k + 1, -1)])
        w = np.stack([np.where(hit, self.q * (1.0 - s), 0.0), np.where(hit, self.q * s, 0.0)])
        return dkv, bins, w

    def _deposit_adjoint(self, x1, lp):
        """sum_b lp[b] d p[b] / d x at x1 per particle."""
        if self.mode == "spline":
            bins, _, dw = self.spline(x1)
            s3 = np.zeros_like(x1)
            for q in range(3):
                m = (bins[q] >= 0) & (bins[q] < self.ns)
                s3 += np.where(m, lp[np.clip(bins[q], 0, self.ns - 1)], 0.0) * dw[q]
            return s3 * self.ibw_cut
        inside, k, k2 = self.smooth_cells(x1)
        return np.where(inside, (lp[k2] - lp[k]) * self.ibw_cut, 0.0)

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
            e2 = self._rf_kick(x, self._induced_kick(x, e, Vi), V, phi, t)
            coeff = st["coeff"][t + 1]
            x1 = x + coeff * e2
            w = st["omega"][:, t]
            sn = [np.sin(w[j] * x + phi[j]) for j in range(len(V))]
            cs = [np.cos(w[j] * x + phi[j]) for j in range(len(V))]
            # deposit adjoint of p_{t+1}
            lx = lx + self._deposit_adjoint(x1, lp)
            le2 = le + coeff * lx
            ds = np.zeros_like(x)
            for j in range(len(V)):
                ds = ds + V[j] * w[j] * cs[j]
                gV[j] += np.sum(le2 * self.q * sn[j])
                gphi[j] += np.sum(le2 * self.q * V[j] * cs[j])
            dkv, bins0, w0 = self._induced_adjoint(x, Vi)
            lx = lx + le2 * (self.q * ds) + le2 * dkv
            lv = np.zeros(self.ns)
            for q in range(len(bins0)):
                m = (bins0[q] >= 0) & (bins0[q] < self.ns)
                np.add.at(lv, bins0[q][m], (le2 * w0[q])[m])
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
