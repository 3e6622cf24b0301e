"""ctypes binding of libuvart_emu.so, the CPU emulation of the island tasks
(the kernels' run_task code compiled for the host, one thread per lane).
Test infrastructure: checks kernel logic on machines without a GPU."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent.parent / "ultimate_uv_artist_amd" / "libuvart_emu.so"

_i64p = C.POINTER(C.c_int64)
NO_RESIDENT = 1 << 8   # emulator plan flag: the ARAP CG streams instead of keeping rows resident
_dp = C.POINTER(C.c_double)
_lib = None


def _stale():
    if not LIB.exists():
        return True
    csrc = LIB.parent / "csrc"
    newest = max(p.stat().st_mtime for p in csrc.iterdir() if p.suffix in (".hip", ".hpp", ".cpp", ".inc", ".h"))
    return newest > LIB.stat().st_mtime


def lib():
    global _lib
    if _lib is None:
        from ultimate_uv_artist_amd import _lib as L
        from ultimate_uv_artist_amd import build

        if _stale():
            build.build_emulator()
        _lib = C.CDLL(str(LIB))
        _lib.uva_emu_create.restype = C.c_void_p
        _lib.uva_emu_create.argtypes = [C.c_int64, _i64p, _i64p, _dp, _i64p, _i64p, C.c_uint32, C.c_int32, C.c_int32]
        _lib.uva_emu_destroy.argtypes = [C.c_void_p]
        _lib.uva_emu_lscm.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int32, C.c_int32]
        _lib.uva_emu_arap.argtypes = [C.c_void_p, _dp, C.POINTER(L.ArapOpts)]
        _lib.uva_emu_cg.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int32, C.c_int32]
        _lib.uva_emu_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _dp]
        _lib.uva_emu_stats2.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int32), _dp]
        _lib.uva_emu_slim.argtypes = [C.c_void_p, _dp, C.POINTER(L.SlimOpts)]
        _lib.uva_emu_local_rhs.argtypes = [C.c_void_p, _dp, _dp]
        _lib.uva_emu_energy_fwd.argtypes = [C.c_void_p, _dp, _dp, C.c_int32, _dp]
        _lib.uva_emu_energy_bwd.argtypes = [C.c_void_p, _dp, _dp, C.c_int32, _dp, _dp, _dp]
        _lib.uva_emu_step_bwd.argtypes = [C.c_void_p, _dp, _dp, _dp]
        _lib.uva_emu_adjoint_cg.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int32]
        _lib.uva_emu_step_apply.argtypes = [C.c_void_p, _dp, _dp, C.c_int32, C.c_int32]
        _lib.uva_emu_plan_pcg.argtypes = [C.c_void_p, _dp, _dp, C.c_int32, C.c_int32, C.c_double, C.c_int32]
        _lib.uva_emu_vperm.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        _lib.uva_emu_lscm_coarse.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_float), C.c_int64,
                                             C.POINTER(C.c_int32)]
        _lib.uva_emu_set_pins.argtypes = [C.c_void_p, _i64p, C.c_int64, C.c_double]
        _lib.uva_emu_overlap_flags.argtypes = [_dp, C.c_int64, _i64p, C.c_int64, C.c_void_p, C.c_double, C.c_int32,
                                               C.c_int32, C.c_void_p]
        _lib.uva_emu_csr_pcg.argtypes = [C.c_int64, _i64p, _i64p, _dp, _dp, _dp, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_double, C.c_int32, C.POINTER(C.c_int32)]
        _lib.uva_emu_raster_last.argtypes = [_dp, _i64p, C.c_int64, C.c_int32, C.c_void_p]
        _lib.uva_emu_team_guard.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int32)]
    return _lib


def team_guard(members, absent, barriers, spin):
    """(largest ticks a member spent, fault flag) of `barriers` team barriers
    where member `absent` (-1: none) never arrives (team_wait on the host)."""
    t, f = C.c_int64(0), C.c_int32(0)
    lib().uva_emu_team_guard(members, absent, barriers, spin, C.byref(t), C.byref(f))
    return t.value, f.value


def overlap_flags(U, F, skip, rel, strict, grid):
    """uva_overlap_flags' algorithm run on the host (same pair test code)."""
    U = np.ascontiguousarray(U, np.float64)
    F = np.ascontiguousarray(F, np.int64)
    sk = None if skip is None else np.ascontiguousarray(skip, np.uint8)
    out = np.zeros(len(F), np.uint8)
    lib().uva_emu_overlap_flags(U.ctypes.data_as(_dp), len(U), F.ctypes.data_as(_i64p), len(F),
                                None if sk is None else sk.ctypes.data_as(C.c_void_p), float(rel), int(strict),
                                int(grid), out.ctypes.data_as(C.c_void_p))
    return out.astype(bool)


def jacobi_pcg(A, b, x0, max_iters, check_every=50, tol2=1e-12, strict=True):
    """uva_csr_pcg's iteration run on the host (same row code), for b and x0
    of shape (n,) or (n, 2): returns (x, iterations run)."""
    A = A.tocsr()
    ip = np.ascontiguousarray(A.indptr, np.int64)
    ix = np.ascontiguousarray(A.indices, np.int64)
    a = np.ascontiguousarray(A.data, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, np.float64, copy=True, order="C")
    nc = 1 if b.ndim == 1 else b.shape[1]
    it = C.c_int32(0)
    rc = lib().uva_emu_csr_pcg(A.shape[0], ip.ctypes.data_as(_i64p), ix.ctypes.data_as(_i64p),
                               a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), x.ctypes.data_as(_dp), int(nc),
                               int(max_iters), int(check_every), float(tol2), int(bool(strict)), C.byref(it))
    if rc != 0:
        raise ValueError(f"uva_emu_csr_pcg returned {rc}")
    return x, it.value


class EmuPlan:
    """A batch of islands run by the CPU emulation with teams of
    ``members`` x ``waves`` wavefronts of 64 lane threads."""

    def __init__(self, charts, members=1, waves=1, batch=False, pins=None, resident=True, keep_order=False,
                 no_coarse=False):
        from ultimate_uv_artist_amd.plan import concat_charts

        self.vo, self.to, P, T = concat_charts(charts)
        self.n = int(self.vo[-1])
        self.k = len(charts)
        pa = None if pins is None else np.ascontiguousarray(np.asarray(pins, np.int64))
        self.h = lib().uva_emu_create(self.k, self.vo.ctypes.data_as(_i64p), self.to.ctypes.data_as(_i64p),
                                      P.ctypes.data_as(_dp), T.ctypes.data_as(_i64p),
                                      None if pa is None else pa.ctypes.data_as(_i64p),
                                      (1 if batch else 0) | (0 if resident else NO_RESIDENT)
                                      | (8 if keep_order else 0) | (16 if no_coarse else 0), members, waves)
        assert self.h, "emulator plan creation failed"

    def __del__(self):
        if getattr(self, "h", None):
            lib().uva_emu_destroy(self.h)

    def lscm(self, tol=1e-8, max_iters=0, precision=0):
        U = np.zeros((self.n, 2))
        rc = lib().uva_emu_lscm(self.h, U.ctypes.data_as(_dp), tol, max_iters, precision)
        if rc != 0:
            raise ValueError("island without boundary")
        return U

    def arap(self, U, iters, line_search=True, batch=False, tol=0.0, cg_tol=1e-6, cg_max_iters=0, precision=0,
             mode=None, iters_folded=0):
        from ultimate_uv_artist_amd import _lib as L

        U = np.ascontiguousarray(np.array(U, np.float64))
        mode = (1 if batch else 0) if mode is None else mode
        o = L.ArapOpts(iters, int(line_search), mode, precision, tol, cg_tol, cg_max_iters, 0, iters_folded)
        lib().uva_emu_arap(self.h, U.ctypes.data_as(_dp), C.byref(o))
        return U

    def slim(self, U, iters, tol=1e-5, cg_tol=1e-10, cg_max_iters=0):
        from ultimate_uv_artist_amd import _lib as L

        U = np.ascontiguousarray(np.array(U, np.float64))
        o = L.SlimOpts(iters, cg_max_iters, tol, cg_tol)
        lib().uva_emu_slim(self.h, U.ctypes.data_as(_dp), C.byref(o))
        return U

    def plan_pcg(self, b, x0, max_iters, check_every=1, tol2=0.0, strict=False):
        """uva_plan_pcg's iteration on the host: (x, iterations run)."""
        b = np.ascontiguousarray(np.array(b, np.float64))
        x = np.ascontiguousarray(np.array(x0, np.float64))
        its = lib().uva_emu_plan_pcg(self.h, b.ctypes.data_as(_dp), x.ctypes.data_as(_dp), int(max_iters),
                                     int(check_every), float(tol2), int(bool(strict)))
        return x, its

    def step_apply(self, U, Un, mode=1, line_search=True):
        """uva_step_apply on the host: the updated layout."""
        U = np.ascontiguousarray(np.array(U, np.float64))
        Un = np.ascontiguousarray(np.array(Un, np.float64))
        lib().uva_emu_step_apply(self.h, U.ctypes.data_as(_dp), Un.ctypes.data_as(_dp), int(mode),
                                 int(bool(line_search)))
        return U

    def vperm(self):
        out = np.zeros(self.n, np.int32)
        lib().uva_emu_vperm(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def lscm_coarse(self, isl):
        """(rows of [Pc | Qc] as an nc x 2nc array or None, slices per aggregate)."""
        cap = 2 * 128 * 128
        buf = np.zeros(cap, np.float32)
        g = C.c_int32(0)
        nc = lib().uva_emu_lscm_coarse(self.h, int(isl), buf.ctypes.data_as(C.POINTER(C.c_float)), cap, C.byref(g))
        return (buf[:2 * nc * nc].reshape(nc, 2 * nc) if nc > 0 else None), g.value

    def set_pins(self, pins, w):
        pins = np.ascontiguousarray(np.asarray(pins, np.int64))
        lib().uva_emu_set_pins(self.h, pins.ctypes.data_as(_i64p), len(pins), float(w))

    def cg(self, b, x0, tol=1e-6, max_iters=0, precision=0):
        b = np.ascontiguousarray(np.array(b, np.float64))
        x = np.ascontiguousarray(np.array(x0, np.float64))
        lib().uva_emu_cg(self.h, b.ctypes.data_as(_dp), x.ctypes.data_as(_dp), tol, max_iters, precision)
        return x

    @staticmethod
    def _p(a):
        return None if a is None else a.ctypes.data_as(_dp)

    def local_rhs(self, U):
        U = np.ascontiguousarray(np.array(U, np.float64))
        b = np.zeros_like(U)
        lib().uva_emu_local_rhs(self.h, self._p(U), self._p(b))
        return b

    def energy_fwd(self, U, w=None, kind=0):
        U = np.ascontiguousarray(np.array(U, np.float64))
        w = None if w is None else np.ascontiguousarray(np.asarray(w, np.float64))
        e = np.zeros(self.k)
        lib().uva_emu_energy_fwd(self.h, self._p(U), self._p(w), kind, self._p(e))
        return e

    def energy_bwd(self, U, ge, w=None, kind=0):
        U = np.ascontiguousarray(np.array(U, np.float64))
        w = None if w is None else np.ascontiguousarray(np.asarray(w, np.float64))
        ge = np.ascontiguousarray(np.asarray(ge, np.float64))
        gU = np.zeros_like(U)
        gw = np.zeros(int(self.to[-1]))
        lib().uva_emu_energy_bwd(self.h, self._p(U), self._p(w), kind, self._p(ge), self._p(gU), self._p(gw))
        return gU, gw

    def step_bwd(self, U, lam):
        U = np.ascontiguousarray(np.array(U, np.float64))
        lam = np.ascontiguousarray(np.array(lam, np.float64))
        gU = np.zeros_like(U)
        lib().uva_emu_step_bwd(self.h, self._p(U), self._p(lam), self._p(gU))
        return gU

    def adjoint_cg(self, g, tol=1e-13, max_iters=0):
        g = np.ascontiguousarray(np.array(g, np.float64))
        lam = np.zeros_like(g)
        lib().uva_emu_adjoint_cg(self.h, self._p(g), self._p(lam), tol, max_iters)
        return lam

    def stats(self):
        it = (C.c_int32 * self.k)()
        cv = (C.c_int32 * self.k)()
        rr = (C.c_double * self.k)()
        lib().uva_emu_stats(self.h, it, cv, rr)
        en = (C.c_double * self.k)()
        ni = (C.c_int32 * self.k)()
        sp = (C.c_double * self.k)()
        lib().uva_emu_stats2(self.h, en, ni, sp)
        return {"cg_iters": np.array(it[:]), "converged": np.array(cv[:]), "rel_resid": np.array(rr[:]),
                "energy": np.array(en[:]), "iters": np.array(ni[:]), "step": np.array(sp[:])}


def raster_last(U, F, res):
    """uva_raster_last's per-texel code run on the host (triangles in order)."""
    U = np.ascontiguousarray(U, np.float64)
    F = np.ascontiguousarray(F, np.int64)
    out = np.empty(res * res, np.int32)
    lib().uva_emu_raster_last(U.ctypes.data_as(_dp), F.ctypes.data_as(_i64p), len(F), int(res),
                              out.ctypes.data_as(C.c_void_p))
    return out.reshape(res, res)
