"""ctypes access to the two CPU checkers under oracle/.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg; never by the product package.

  * ``Oracle``    — oracle/_build/libpforacle.so, the FP64 C restatement of
                    the reference CPU Newton-Raphson path (pforacle.c).
  * ``Reference`` — oracle/_ref/libgapref.so, the reference's own CPU
                    sources compiled where they lie (ref_harness.cpp); only
                    present where the reference tree was available at build time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "_build" / "libpforacle.so"
REF_SO = HERE / "_ref" / "libgapref.so"

_i = C.POINTER(C.c_int)
_d = C.POINTER(C.c_double)


class PfNet(C.Structure):
    _fields_ = [
        ("n_bus", C.c_int), ("bus_id", _i), ("u_rated", _d), ("bus_type", _i), ("u_pu", _d),
        ("p_w", _d), ("q_var", _d),
        ("n_branch", C.c_int), ("br_id", _i), ("br_from", _i), ("br_to", _i), ("br_status", _i),
        ("br_r", _d), ("br_x", _d), ("br_g", _d), ("br_b", _d),
        ("n_app", C.c_int), ("app_id", _i), ("app_node", _i), ("app_status", _i), ("app_type", _i),
        ("app_p", _d), ("app_q", _d), ("app_g", _d), ("app_b", _d),
    ]


def _ptr(a, t):
    return a.ctypes.data_as(t)


def make_pfnet(net):
    """Returns (struct, keepalive) for a gap_amd.network.Network-like object."""
    keep = []

    def I(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return _ptr(a, _i)

    def D(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _ptr(a, _d)

    s = PfNet(
        int(len(net.bus_id)), I(net.bus_id), D(net.u_rated), I(net.bus_type), D(net.u_pu),
        D(net.p_w), D(net.q_var),
        int(len(net.br_id)), I(net.br_id), I(net.br_from), I(net.br_to), I(net.br_status),
        D(net.br_r), D(net.br_x), D(net.br_g), D(net.br_b),
        int(len(net.app_id)), I(net.app_id), I(net.app_node), I(net.app_status), I(net.app_type),
        D(net.app_p), D(net.app_q), D(net.app_g), D(net.app_b))
    return s, keep


def build(with_ref: bool | None = None) -> None:
    """Compile the C restatement, and the reference harness when the tree is present."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if with_ref is None:
        with_ref = Path(os.environ.get("GAP_REFERENCE", "/root/reference")).joinpath("src").is_dir()
    if with_ref:
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


class _Solver:
    prefix = ""

    def __init__(self, path: Path):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built")
        self.lib = C.CDLL(str(path))
        p = self.prefix
        self._ybus = getattr(self.lib, p + "build_ybus")
        self._ybus.restype = C.c_int
        self._solve = getattr(self.lib, p + "solve")
        self._solve.restype = C.c_int
        self._mis = getattr(self.lib, p + "mismatches")
        self._mis.restype = C.c_int
        self._lu = getattr(self.lib, p + "lu_solve")
        self._lu.restype = C.c_int

    def build_ybus(self, net):
        s, keep = make_pfnet(net)
        cap = len(net.bus_id) + 2 * len(net.br_id) + 1
        rp = np.zeros(len(net.bus_id) + 1, np.int32)
        ci = np.zeros(cap, np.int32)
        v = np.zeros(2 * cap)
        nnz = self._ybus(C.byref(s), _ptr(rp, _i), _ptr(ci, _i), _ptr(v, _d), cap)
        if nnz < 0:
            raise RuntimeError("Y-bus capacity")
        return rp, ci[:nnz].copy(), (v[0:2 * nnz:2] + 1j * v[1:2 * nnz:2])

    def solve(self, net, tol=1e-8, max_it=100, base_power=100e6, j_call=0, traj=True):
        s, keep = make_pfnet(net)
        n = len(net.bus_id)
        v = np.zeros(2 * n)
        it = C.c_int(0)
        conv = C.c_int(0)
        fm = C.c_double(0)
        tr = np.zeros(max(max_it, 1) + 1)
        ntr = C.c_int(0)
        nvar_cap = 2 * n + 1
        jcap = 8 * (len(net.br_id) * 2 + n) + 16
        jn = C.c_int(0)
        jrp = np.zeros(nvar_cap + 1, np.int32)
        jci = np.zeros(jcap, np.int32)
        jv = np.zeros(jcap)
        jnnz = C.c_int(0)
        rc = self._solve(C.byref(s), C.c_double(tol), int(max_it), C.c_double(base_power),
                         _ptr(v, _d), C.byref(it), C.byref(conv), C.byref(fm),
                         _ptr(tr, _d) if traj else None, len(tr), C.byref(ntr), int(j_call),
                         C.byref(jn), _ptr(jrp, _i), _ptr(jci, _i), _ptr(jv, _d), jcap, C.byref(jnnz))
        if rc != 0:
            raise RuntimeError(f"{type(self).__name__}.solve failed rc={rc}")
        out = dict(v=v[0::2] + 1j * v[1::2], iterations=it.value, converged=bool(conv.value),
                   final_mismatch=fm.value, traj=tr[:min(ntr.value, len(tr))].copy())
        if j_call is not None and j_call >= 0 and jn.value > 0:
            m = jnnz.value
            out["J"] = (jrp[:jn.value + 1].copy(), jci[:m].copy(), jv[:m].copy())
        return out

    def mismatches(self, net, v, y):
        s, keep = make_pfnet(net)
        rp, ci, vals = y
        vri = np.ascontiguousarray(np.stack([np.real(v), np.imag(v)], 1).ravel())
        yri = np.ascontiguousarray(np.stack([np.real(vals), np.imag(vals)], 1).ravel())
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        out = np.zeros(2 * len(net.bus_id) + 1)
        k = self._mis(C.byref(s), _ptr(vri, _d), _ptr(yri, _d), _ptr(rp, _i), _ptr(ci, _i),
                      len(ci), _ptr(out, _d), len(out))
        if k < 0:
            raise RuntimeError("mismatch capacity")
        return out[:k].copy()

    def lu_solve(self, n, rp, ci, vals, rhs):
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        vals = np.asarray(vals, np.complex128)
        vri = np.ascontiguousarray(np.stack([vals.real, vals.imag], 1).ravel())
        rhs = np.asarray(rhs, np.complex128)
        bri = np.ascontiguousarray(np.stack([rhs.real, rhs.imag], 1).ravel())
        x = np.zeros(2 * n)
        rc = self._lu(int(n), _ptr(rp, _i), _ptr(ci, _i), _ptr(vri, _d), len(ci), _ptr(bri, _d),
                      _ptr(x, _d))
        return rc, x[0::2] + 1j * x[1::2]


class Reference(_Solver):
    prefix = "gapref_"

    def __init__(self):
        super().__init__(REF_SO)


class Oracle(_Solver):
    prefix = "pfo_"

    def __init__(self):
        super().__init__(ORACLE_SO)
        self._jac = self.lib.pfo_jacobian
        self._jac.restype = C.c_int

    def jacobian(self, net, v, ypu):
        """Jacobian (reference layout, |J| > 1e-12) of the per-unit Y at voltages v."""
        s, keep = make_pfnet(net)
        rp, ci, vals = ypu
        vri = np.ascontiguousarray(np.stack([np.real(v), np.imag(v)], 1).ravel())
        yri = np.ascontiguousarray(np.stack([np.real(vals), np.imag(vals)], 1).ravel())
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        n = len(net.bus_id)
        cap = 8 * (len(ci) + n) + 16
        jn = C.c_int(0)
        jrp = np.zeros(2 * n + 2, np.int32)
        jci = np.zeros(cap, np.int32)
        jv = np.zeros(cap)
        nnz = self._jac(C.byref(s), _ptr(vri, _d), _ptr(yri, _d), _ptr(rp, _i), _ptr(ci, _i), C.byref(jn),
                        _ptr(jrp, _i), _ptr(jci, _i), _ptr(jv, _d), cap)
        if nnz < 0:
            raise RuntimeError("Jacobian capacity")
        return jrp[:jn.value + 1].copy(), jci[:nnz].copy(), jv[:nnz].copy()


def reference_available() -> bool:
    return REF_SO.exists()
