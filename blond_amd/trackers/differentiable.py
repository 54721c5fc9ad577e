"""Differentiable tracking on the MI355X.

``DifferentiableTracker`` runs the intensity map of a RingAndRFTracker /
Profile / broadband-resonator impedance on the GPU and returns the final
profile as a torch tensor that is differentiable with respect to the RF
voltages and phases and the resonators' shunt impedances, resonant
frequencies and quality factors.  The gradient is the hand-written adjoint
of the library (``blmi_adjoint_*``): turn checkpoints every
``checkpoint_every`` turns, recomputation of each segment from its
checkpoint, and a reverse pass over the particles.  With
``checkpoint_every=1`` every turn's state stays in HBM (16 B per
macroparticle per turn: 8 GB for 1e6 particles x 500 turns) and the reverse
pass recomputes nothing.

Two maps (``map``):

* ``"reference"`` (default): the reference's own smooth profile and
  interpolated kick -- smooth_histogram (histogram.cpp:75-132, with its
  negative-weight quirk) and linear_interp_kick (linear_interp_kick.cpp:21-71)
  on the profile's bin centres.  Both are piecewise linear in dt: the
  profile jumps when a particle crosses a bin centre or a bin edge and the
  kick has kinks at bin centres.  The adjoint differentiates each piece (the
  pathwise derivative of the loss), which is exact wherever no particle sits
  on a centre or edge; central finite differences agree with it only while
  no particle crosses one within the step.
* ``"spline"``: the profile deposited, and the induced voltage interpolated
  for the kick, with one quadratic B-spline on bin centres (weights
  (1/2-d)^2/2, 3/4-d^2, (1/2+d)^2/2 on bins k-1, k, k+1), which has a
  continuous first derivative, so finite differences converge to the
  adjoint at second order.

The induced voltage is -scale * (g circular-convolved with the profile),
g = irfft(Z, n_fft) (impedance.py:457-489), Z from the library's resonator
kernel (fast_resonator.cpp:45-71) divided by the bin size
(impedance.py:881-890).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from .. import _lib
from ..constants import e
from ..utils import bmath as bm
from .fused import rf_program_array, ring_program_arrays


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


def resonator_derivatives(freq, R, fr, Q):
    """dZ/dR, dZ/df_r, dZ/dQ of Z(f) = R / (1 + iQ (f/f_r - f_r/f)), Z(0) = 0,
    per resonator: complex arrays [n_res, n_f]."""
    freq = np.asarray(freq, dtype=np.float64)
    n = len(R)
    out = [np.zeros((n, len(freq)), np.complex128) for _ in range(3)]
    f = freq[1:]
    for k in range(n):
        u = f / fr[k] - fr[k] / f
        den = 1.0 + 1j * Q[k] * u
        out[0][k, 1:] = 1.0 / den
        out[1][k, 1:] = (-R[k] * 1j * Q[k] / den ** 2) * (-f / fr[k] ** 2 - 1.0 / f)
        out[2][k, 1:] = -R[k] * 1j * u / den ** 2
    return out


MAPS = {"spline": 0, "reference": 1}


class DifferentiableTracker:
    def __init__(self, tracker, profile, n_fft: int | None = None, checkpoint_every: int = 50,
                 map: str = "reference"):
        import torch

        from ..toolbox.next_regular import next_regular

        if tracker.solver != "simple":
            raise NotImplementedError("the adjoint covers the 'simple' drift")
        self._torch = torch
        self.tracker, self.profile = tracker, profile
        rp, beam = tracker.rf_params, tracker.beam
        self.rp, self.beam = rp, beam
        self.n_slices = profile.n_slices
        self.n_fft = int(n_fft or next_regular(self.n_slices))
        self.bin_size = float(profile.bin_size)
        self.freq = np.fft.rfftfreq(self.n_fft, self.bin_size)
        self.scale = float(beam.particle.charge * e * beam.ratio)
        self.n_turns = rp.n_turns
        n = beam.n_macroparticles
        lib = _lib.load()
        h = lib.blmi_adjoint_create(n, rp.n_rf, rp.n_turns, self.n_slices, int(checkpoint_every))
        if not h:
            raise _lib.BlmiError(f"blmi_adjoint_create: {_lib.last_error()}")
        self._h = ct.c_void_p(h)
        dt = beam.dt if isinstance(beam.dt, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(beam.dt, dtype=np.float64)).cuda()
        dE = beam.dE if isinstance(beam.dE, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(beam.dE, dtype=np.float64)).cuda()
        dt, dE = dt.to(torch.float64).contiguous(), dE.to(torch.float64).contiguous()
        _lib.call("blmi_adjoint_set_particles", self._h, ct.c_void_p(dt.data_ptr()),
                  ct.c_void_p(dE.data_ptr()))
        _lib.call("blmi_adjoint_set_slices", self._h, float(profile.cut_left),
                  float(profile.cut_right))
        if map not in MAPS:
            raise ValueError(f"map {map!r} not in {tuple(MAPS)}")
        self.map = map
        bc = profile.bin_centers
        bc = np.ascontiguousarray(bc.detach().cpu().numpy() if hasattr(bc, "detach") else bc,
                                  dtype=np.float64)
        self._bc = bc
        _lib.call("blmi_adjoint_set_map", self._h, MAPS[map], _dptr(bc))
        self._ring = ring_program_arrays(rp)
        self._acc = np.zeros(rp.n_turns + 1)
        self._acc[: len(tracker.acceleration_kick)] = tracker.acceleration_kick
        self._phi0 = np.array(rp.host_array("phi_rf"), dtype=np.float64)

    # -- pieces used by the autograd function ------------------------------
    def _set_rf(self, V: np.ndarray, phi: np.ndarray):
        """Constant voltage and phase programs V_j, phi_j for every turn."""
        rp = self.rp
        nt = rp.n_turns + 1
        phi_prog = np.repeat(np.asarray(phi, np.float64)[:, None], nt, axis=1)
        rf = rf_program_array(rp, phi_prog)
        rf[:, 0, :] = np.asarray(V, np.float64)[None, :]
        self._rf = np.ascontiguousarray(rf)
        r = self._ring
        _lib.call("blmi_adjoint_set_programs", self._h, _dptr(self._rf), _dptr(self._acc),
                  _dptr(r[0]), float(rp.length_ratio), float(rp.alpha_order),
                  *[_dptr(a) for a in r[1:]], float(rp.particle.charge))

    def _set_wake(self, R, fr, Q):
        Z = bm.fast_resonator(np.asarray(R, np.float64), np.asarray(Q, np.float64), self.freq,
                              np.asarray(fr, np.float64))
        g = np.ascontiguousarray(bm.irfft(Z / self.bin_size, self.n_fft), dtype=np.float64)
        self._g = g
        _lib.call("blmi_adjoint_set_wake", self._h, self.n_fft, _dptr(g), self.scale)

    def _forward(self, turns: int):
        torch = self._torch
        out = torch.empty(self.n_slices, dtype=torch.float64, device="cuda")
        _lib.call("blmi_adjoint_forward", self._h, int(turns), ct.c_void_p(out.data_ptr()),
                  ct.c_void_p(torch.cuda.current_stream().cuda_stream))
        return out

    def _run(self, key, turns: int):
        v, p, r, f, q = key
        self._set_rf(v, p)
        self._set_wake(r, f, q)
        out = self._forward(turns)
        self._last = (tuple(a.tobytes() for a in key), turns)
        return out

    def _holds(self, key, turns: int) -> bool:
        return getattr(self, "_last", None) == (tuple(a.tobytes() for a in key), turns)

    def _backward(self, dprofile, R, fr, Q):
        torch = self._torch
        dp = dprofile.to(torch.float64).contiguous()
        grf = np.zeros(2 * self.rp.n_rf)
        corr = np.zeros(self.n_fft)
        _lib.call("blmi_adjoint_backward", self._h, ct.c_void_p(dp.data_ptr()), _dptr(grf),
                  _dptr(corr), ct.c_void_p(torch.cuda.current_stream().cuda_stream))
        dZ = resonator_derivatives(self.freq, R, fr, Q)
        gz = [np.array([-(self.scale * np.dot(np.fft.irfft(d / self.bin_size, self.n_fft), corr))
                        for d in dk]) for dk in dZ]
        return grf[0::2], grf[1::2], gz[0], gz[1], gz[2]

    # -- user entry ---------------------------------------------------------
    def profile_after(self, V, phi, R, fr, Q, turns: int | None = None):
        """Final profile after ``turns`` turns (default n_turns) as a
        function of the parameter tensors (float64, any device)."""
        return _TrackFn.apply(V, phi, R, fr, Q, self, int(turns or self.n_turns))

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().blmi_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64).ravel()


def _make_fn():
    import torch

    class TrackFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, V, phi, R, fr, Q, dtr, turns):
            key = tuple(_np(a) for a in (V, phi, R, fr, Q))
            out = dtr._run(key, turns)
            ctx.dtr = dtr
            ctx.key = key
            ctx.turns = turns
            ctx.like = (V, phi, R, fr, Q)
            return out

        @staticmethod
        def backward(ctx, grad):
            dtr = ctx.dtr
            if not dtr._holds(ctx.key, ctx.turns):
                dtr._run(ctx.key, ctx.turns)  # checkpoints of another call: recompute
            gV, gphi, gR, gfr, gQ = dtr._backward(grad, *ctx.key[2:])
            outs = []
            for g, like in zip((gV, gphi, gR, gfr, gQ), ctx.like):
                outs.append(torch.from_numpy(np.asarray(g, dtype=np.float64)).to(like.device)
                            .reshape(like.shape).to(like.dtype))
            return (*outs, None, None)

    return TrackFn


class _Lazy:
    fn = None

    @classmethod
    def apply(cls, *a):
        if cls.fn is None:
            cls.fn = _make_fn()
        return cls.fn.apply(*a)


_TrackFn = _Lazy
