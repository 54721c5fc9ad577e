"""Writes the golden vectors of the tracking-path kernels.

Inputs are seeded; outputs come from BLonD's own C++ kernels compiled from
the reference sources by oracle/Makefile (oracle/_ref, strict build:
-O3 -fopenmp -ffp-contract=off), run single-threaded so that the smooth
histogram's summation order is the serial one.  Also records, as data, the
reference's profile test fixture: unittests/beam_profile/dt_coordinates.npz
and the expected profiles written in
unittests/beam_profile/test_beam_profile_object.py:169-230.

Run in the build container (needs /root/reference and oracle/_ref):
    OMP_NUM_THREADS=1 python tests/golden/make_golden.py
"""
from __future__ import annotations

import ast
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference"

W1 = 2.5182298872245054e9


def bunch(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(1.25e-9, 0.3e-9, n), rng.normal(0.0, 1e8, n)


def kernels(lib, out):
    n = 2001
    dt, dE = bunch(n, 101)
    out["dt"], out["dE"] = dt, dE
    # kick: one and two harmonics, with an acceleration kick
    for n_rf in (1, 2):
        V = np.array([6e6, 1.5e6])[:n_rf]
        w = np.array([W1, 4 * W1])[:n_rf]
        ph = np.array([0.1, 0.3])[:n_rf]
        e = dE.copy()
        lib.kick(dt, e, V, w, ph, 1.0, n_rf, -1234.5)
        out[f"kick_rf{n_rf}_V"], out[f"kick_rf{n_rf}_w"], out[f"kick_rf{n_rf}_phi"] = V, w, ph
        out[f"kick_rf{n_rf}"] = e
    # drift: three solvers, alpha orders 0..2
    args = dict(t_rev=8.892465516509709e-05, length_ratio=1.0, eta_0=3.172867586042721e-04,
                eta_1=2e-6, eta_2=-3e-7, alpha_0=3.2e-4, alpha_1=1e-5, alpha_2=1e-6,
                beta=0.9999978262922387, energy=450.0009780e9)
    out["drift_args"] = np.array(list(args.values()))
    for solver in ("simple", "legacy", "exact"):
        for ao in (0, 1, 2):
            t = dt.copy()
            lib.drift(t, dE, solver, args["t_rev"], args["length_ratio"], ao, args["eta_0"],
                      args["eta_1"], args["eta_2"], args["alpha_0"], args["alpha_1"],
                      args["alpha_2"], args["beta"], args["energy"])
            out[f"drift_{solver}_ao{ao}"] = t
    # histograms with edge cases appended
    cl, cr = 0.0, 2.495080111254127e-09
    for ns in (1, 100, 4096):
        bw = (cr - cl) / ns
        extra = np.array([cl, cr, cl - 1e-20, cr - 1e-25, cl + bw, np.nan, np.inf, -np.inf,
                          1e30, -1e30, cl + 0.5 * bw])
        x = np.concatenate([dt, extra])
        h = np.zeros(ns)
        lib.slice_beam(x, h, cl, cr)
        out[f"hist_{ns}_dt"], out[f"hist_{ns}"] = x, h
    for ns in (8, 150):
        h = np.zeros(ns)
        lib.slice_smooth(dt, h, cl, cr)
        out[f"smooth_{ns}"] = h
    # interpolated kick
    for ns in (2, 100):
        bc = np.linspace(0, 2.5e-9, ns + 1)[:-1] + 2.5e-9 / (2 * ns)
        V = np.random.default_rng(7 + ns).normal(0, 1e3, ns)
        e = dE.copy()
        lib.linear_interp_kick(dt, e, V, bc, 1.0, 12.5)
        out[f"interp_{ns}_bc"], out[f"interp_{ns}_V"], out[f"interp_{ns}"] = bc, V, e
    # RF voltage on bin centres
    bc = np.linspace(0, 2.5e-9, 257)[:-1]
    out["rfv_bc"] = bc
    out["rfv"] = lib.rf_volt_comp(np.array([6e6, 1.5e6]), np.array([W1, 4 * W1]),
                                  np.array([0.0, 0.2]), bc)
    # resonator impedance
    f = np.fft.rfftfreq(2048, 2.5e-9 / 512)
    R, fr, Q = np.array([5e3, 1e4, 3e5]), np.array([1e9, 2e8, 5e7]), np.array([1.0, 100.0, 5e3])
    out["res_f"], out["res_R"], out["res_fr"], out["res_Q"] = f, R, fr, Q
    out["res_Z"] = lib.fast_resonator(R, Q, f, fr)


def kernels_single(lib, out):
    n = 2001
    dt, dE = bunch(n, 202)
    dt, dE = dt.astype(np.float32), dE.astype(np.float32)
    out["f32_dt"], out["f32_dE"] = dt, dE
    e = dE.copy()
    lib.kick(dt, e, np.array([6e6, 1.5e6], np.float32), np.array([W1, 4 * W1], np.float32),
             np.array([0.1, 0.3], np.float32), 1.0, 2, -12.5)
    out["f32_kick"] = e
    h = np.zeros(100, np.float32)
    lib.slice_beam(dt, h, 0.0, 2.5e-9)
    out["f32_hist"] = h


def profile_fixture():
    """The reference's profile test data: dt coordinates and the expected
    profiles written in test_beam_profile_object.py."""
    src = os.path.join(REF, "unittests/beam_profile")
    shutil.copyfile(os.path.join(src, "dt_coordinates.npz"), os.path.join(HERE, "dt_coordinates.npz"))
    text = open(os.path.join(src, "test_beam_profile_object.py")).read()
    # the literal lists following each "self.profileK.n_macroparticles.tolist()," line
    expected = {}
    for k in (1, 2, 3):
        key = f"self.profile{k}.n_macroparticles.tolist(),"
        i = text.index(key) + len(key)
        j = text.index("[", i)
        depth, e = 0, j
        while True:
            c = text[e]
            depth += c == "["
            depth -= c == "]"
            e += 1
            if depth == 0:
                break
        expected[f"profile{k}"] = ast.literal_eval(text[j:e])
    expected["t_rev0"] = 5.71753954209e-07
    expected["source"] = "unittests/beam_profile/test_beam_profile_object.py:169-230"
    with open(os.path.join(HERE, "profile_expected.json"), "w") as fh:
        json.dump(expected, fh)


def main():
    from oracle import oracle as orc

    if os.environ.get("OMP_NUM_THREADS") != "1":
        raise SystemExit("run with OMP_NUM_THREADS=1 (serial summation order)")
    lib = orc.ref_lib("strict", "double")
    libs = orc.ref_lib("strict", "single")
    if lib is None or libs is None:
        raise SystemExit("oracle/_ref missing: make -C oracle ref")
    out = {}
    kernels(lib, out)
    kernels_single(libs, out)
    np.savez_compressed(os.path.join(HERE, "kernels_golden.npz"), **out)
    profile_fixture()
    print("wrote", sorted(out)[:5], "...", len(out), "arrays")


if __name__ == "__main__":
    main()
