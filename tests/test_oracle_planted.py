"""Planted mistakes in the oracle must fail the conductance pins.

The oracle's sources are copied to a temporary directory, one reversal-potential term is given the
wrong sign there (E - V written as V - E), and the copy is compiled with the oracle's own flags. The
closed-form checks of the conductance terms (the same ones test_oracle_pins.py runs on the real
oracle) must accept the real oracle and reject each planted copy: a pin that a sign error passes
pins nothing. CPU only; the planted library is never used outside this file.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ORC = os.path.join(os.path.dirname(HERE), "oracle")
SOURCES = ("nbx_oracle.c", "nbx_oracle.h", "orc_model.inc")
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared"]

# (name, file, text of the right term, its sign-flipped form)
PLANTS = [
    ("lif_exp_e_rev", "orc_model.inc", "g * ((REAL)m->e_rev - V)", "g * (V - (REAL)m->e_rev)"),
    ("izh_ampa_e_ampa", "orc_model.inc", "tot += ga * ((REAL)m->e_ampa - V);", "tot += ga * (V - (REAL)m->e_ampa);"),
]


def _load(path):
    L = C.CDLL(path)
    L.orc_model_prepare.argtypes = [C.POINTER(O.OrcModel)]
    L.orc_update_f64.restype = C.c_int32
    L.orc_update_f64.argtypes = [C.POINTER(O.OrcModel), C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    return L


def _model(L, **kw):
    base = dict(model=O.LIF_VJ, integrator=O.EULER, n=1, t_ref=50, dt=1.0, v_rest=0.0, tau=10.0, v_th=20.0, r=1.0,
                v_reset=-5.0, g_tau=12.0, e_rev=0.0, izh_v_th=30.0, izh_r=1.0, e_ampa=0.0, e_gaba=-80.0,
                ampa_tau=5.0, gaba_tau=10.0)
    base.update(kw)
    keep = []
    m = O.OrcModel(**{k: v for k, v in base.items() if k not in "abcd"})
    for k in "abcd":
        if k in base:
            arr = np.ascontiguousarray(base[k], dtype=np.float32)
            keep.append(arr)
            setattr(m, k, O.ptr(arr, C.c_float))
    L.orc_model_prepare(C.byref(m))
    m._keep = keep
    return m


def _step(L, m, x):
    last = np.array([10000], np.int32)
    sp = np.zeros(1, np.int32)
    L.orc_update_f64(C.byref(m), O.ptr(x, C.c_double), O.ptr(last, C.c_int32), None, O.ptr(sp, C.c_int32))


def lif_exp_closed_form_holds(L, integ):
    # tau dV/dt = R I + g (E - V) + (V_rest - V) with g held: V_n = V_inf + (V_0 - V_inf) f^n
    dt, tau, r, E, v_rest, I, g = 0.125, 20.0, 1.5, 0.0, -65.0, 4.0, 0.35
    m = _model(L, model=O.LIF_EXP, integrator=integ, dt=dt, tau=tau, r=r, e_rev=E, v_rest=v_rest, v_th=100.0,
               g_tau=0.0)
    x = np.array([-70.0, g, I], np.float64)
    got = []
    for _ in range(100):
        _step(L, m, x)
        got.append(float(x[0]))
    v_inf = (r * I + g * E + v_rest) / (1.0 + g)
    h = (1.0 + g) * dt / tau
    f = 1.0 - h if integ == O.EULER else 1.0 - h + h * h / 2.0 - h ** 3 / 6.0 + h ** 4 / 24.0
    want = v_inf + (-70.0 - v_inf) * f ** np.arange(1, 101)
    return np.allclose(got, want, rtol=1e-12, atol=0.0)


def izh_ampa_sign_holds(L):
    # the AMPA current g (E_ampa - v) is zero at E_ampa and depolarises below it
    m = _model(L, model=O.IZH, integrator=O.EULER, dt=0.5, v_rest=-65.0, izh_r=1.0, e_ampa=0.0, e_gaba=-80.0,
               ampa_tau=0.0, gaba_tau=0.0, izh_v_th=1000.0, a=[0.02], b=[0.2], c=[-65.0], d=[8.0])

    def v_after(v, ga):
        x = np.array([v, -13.0, ga, 0.0, 0.0, 0.0], np.float64)
        _step(L, m, x)
        return x[0]
    return v_after(-65.0, 0.6) > v_after(-65.0, 0.0) and v_after(0.0, 0.6) == v_after(0.0, 0.0)


def _build(tmp, plant):
    for f in SOURCES:
        shutil.copy(os.path.join(ORC, f), tmp)
    if plant is not None:
        _, fname, right, wrong = plant
        p = os.path.join(tmp, fname)
        text = open(p).read()
        assert text.count(right) == 1, f"planted term {right!r} not found once in {fname}"
        open(p, "w").write(text.replace(right, wrong))
    so = os.path.join(tmp, "liborc_planted.so")
    subprocess.run(["gcc", *CFLAGS, "-o", so, os.path.join(tmp, "nbx_oracle.c"), "-lm"], check=True)
    return _load(so)


def test_unplanted_copy_passes(tmp_path):
    L = _build(str(tmp_path), None)
    assert lif_exp_closed_form_holds(L, O.EULER) and lif_exp_closed_form_holds(L, O.RK4)
    assert izh_ampa_sign_holds(L)


@pytest.mark.parametrize("plant", PLANTS, ids=[p[0] for p in PLANTS])
def test_planted_reversal_sign_is_caught(tmp_path, plant):
    L = _build(str(tmp_path), plant)
    if plant[0].startswith("lif"):
        assert not lif_exp_closed_form_holds(L, O.EULER)
        assert not lif_exp_closed_form_holds(L, O.RK4)
    else:
        assert not izh_ampa_sign_holds(L)
