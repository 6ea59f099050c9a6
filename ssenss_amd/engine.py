"""Python handle on an ssx context.  Device vectors are torch tensors on the
GPU (torch provides the memory and the stream, nothing else); all numerics
run in libssx's HIP kernels."""
import ctypes as C

import numpy as np

from .native import NSParams, ParticlesStruct, Stats, check, load


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    def __init__(self, mesh, stream=None):
        self._lib = load()
        self.mesh = mesh
        self._ms = mesh.struct()
        h = C.c_void_p()
        check(self._lib.ssx_create(C.byref(h), C.byref(self._ms)))
        self.h = h
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if getattr(self, "h", None):
            self._lib.ssx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- info ---
    @property
    def nlevels(self):
        return self._lib.ssx_nlevels(self.h)

    def level_n(self, level):
        return self._lib.ssx_level_n(self.h, level)

    def ntot(self, level=0):
        return self._lib.ssx_level_ntot(self.h, level)

    @property
    def lam(self):
        return self._lib.ssx_lambda(self.h)

    def set_stream(self, stream):
        """Order the library's work on a torch.cuda.Stream (torch's default
        stream is the HIP null stream, handle 0); None returns to the
        context's own non-blocking stream."""
        if stream is None:
            s = self._lib.ssx_own_stream(self.h)
        else:
            s = stream.cuda_stream
        check(self._lib.ssx_set_stream(self.h, C.c_void_p(s) if s else None))

    def stream_handle(self):
        return self._lib.ssx_get_stream(self.h)

    def synchronize(self):
        check(self._lib.ssx_synchronize(self.h))

    def level_data(self, level, which):
        n = self.ntot(level) * (6 if which == 0 else 1)
        out = np.zeros(n)
        check(self._lib.ssx_level_data(
            self.h, level, which, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def dss_groups(self, level=0):
        a = (C.c_longlong * 3)()
        check(self._lib.ssx_dss_groups(self.h, level, a))
        return tuple(a)

    def set_diag_choice(self, choice):
        check(self._lib.ssx_set_diag_choice(self.h, choice))

    def set_option(self, option, value):
        check(self._lib.ssx_set_option(self.h, option, int(value)))

    # --- operator launch shapes ---
    def autotune(self, prec_mask=1):
        check(self._lib.ssx_autotune(self.h, prec_mask))

    def ax_variant(self, level, prec=0):
        """(variant in use, measured us per launch of each variant)"""
        from .native import AX_NVARIANTS
        us = np.zeros(AX_NVARIANTS)
        v = self._lib.ssx_ax_variant(self.h, level, prec, _dp(us))
        return v, us

    def set_ax_variant(self, level, prec, v):
        check(self._lib.ssx_set_ax_variant(self.h, level, prec, v))

    # --- operators (torch device tensors) ---
    def ax(self, level, op, u, au):
        check(self._lib.ssx_ax(self.h, level, op, _ptr(u), _ptr(au)))

    def opeval(self, level, op, var, p, ap):
        check(self._lib.ssx_opeval(self.h, level, op, var, _ptr(p), _ptr(ap)))

    def dss(self, level, u):
        check(self._lib.ssx_dss(self.h, level, _ptr(u)))

    def dss_average(self, level, u):
        check(self._lib.ssx_dss_average(self.h, level, _ptr(u)))

    def dot(self, level, x, y):
        out = C.c_double()
        check(self._lib.ssx_dot(self.h, level, _ptr(x), _ptr(y),
                                C.byref(out)))
        return out.value

    def mask(self, level, var, u):
        check(self._lib.ssx_mask(self.h, level, var, _ptr(u)))

    def nullspace(self, level, var, u):
        check(self._lib.ssx_nullspace(self.h, level, var, _ptr(u)))

    def restrict(self, level, fine, coarse):
        check(self._lib.ssx_restrict(self.h, level, _ptr(fine), _ptr(coarse)))

    def prolong(self, level, coarse, fine):
        check(self._lib.ssx_prolong(self.h, level, _ptr(coarse), _ptr(fine)))

    def jacobi(self, level, var, x, px):
        check(self._lib.ssx_jacobi(self.h, level, var, _ptr(x), _ptr(px)))

    def mg_precon(self, op, var, eps, x, px):
        check(self._lib.ssx_mg_precon(self.h, op, var, eps, _ptr(x), _ptr(px)))

    def cg(self, level, op, var, f, u, eps, max_iter, allow_notconv=False):
        st = Stats()
        rc = self._lib.ssx_cg(self.h, level, op, var, _ptr(f), _ptr(u), eps,
                              max_iter, C.byref(st))
        if not (allow_notconv and rc == 4):
            check(rc)
        return st

    def fcg(self, op, var, precon, f, u, eps, max_iter, allow_notconv=False):
        st = Stats()
        rc = self._lib.ssx_fcg(self.h, op, var, precon, _ptr(f), _ptr(u), eps,
                               max_iter, C.byref(st))
        if not (allow_notconv and rc == 4):
            check(rc)
        return st

    # --- host-array entry points (reference solve_*_cg) ---
    def solve_poisson_cg(self, f, u, var, eps):
        st = Stats()
        dp = C.POINTER(C.c_double)
        check(self._lib.ssx_solve_poisson_cg(self.h, f.ctypes.data_as(dp),
                                             u.ctypes.data_as(dp), var, eps,
                                             C.byref(st)))
        return st

    def solve_helmholtz_cg(self, f, u, var, eps):
        st = Stats()
        dp = C.POINTER(C.c_double)
        check(self._lib.ssx_solve_helmholtz_cg(self.h, f.ctypes.data_as(dp),
                                               u.ctypes.data_as(dp), var, eps,
                                               C.byref(st)))
        return st

    # --- particles (torch device tensors) ---
    def particles_set_velocity(self, u, v, w):
        check(self._lib.ssx_particles_set_velocity(self.h, _ptr(u), _ptr(v),
                                                   _ptr(w)))

    def particles_step(self, p, t_now, dt):
        """p: dict with int32 tensors el, old_el, reinit and float64
        tensors x, y, z, u, v, w, t, tau (all on the GPU)"""
        ps = ParticlesStruct(int(p["x"].numel()),
                             *(p[k].data_ptr() if p.get(k) is not None
                               else None for k in
                               ("el", "old_el", "x", "y", "z", "u", "v", "w",
                                "t", "tau", "reinit")))
        mp = C.c_int()
        check(self._lib.ssx_particles_step(self.h, C.byref(ps), t_now, dt,
                                           C.byref(mp)))
        return mp.value

    def tracers_step(self, p, t_now, dt):
        """RK2 tracers (ssx_tracers_step): p as particles_step, tau
        optional; u, v, w receive the fluid velocity at the end point"""
        ps = ParticlesStruct(int(p["x"].numel()),
                             *(p[k].data_ptr() if p.get(k) is not None
                               else None for k in
                               ("el", "old_el", "x", "y", "z", "u", "v", "w",
                                "t", "tau", "reinit")))
        mp = C.c_int()
        check(self._lib.ssx_tracers_step(self.h, C.byref(ps), t_now, dt,
                                         C.byref(mp)))
        return mp.value

    # --- particles on host arrays (numpy) ---
    def particles_set_velocity_host(self, u, v, w):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (u, v, w)]
        check(self._lib.ssx_particles_set_velocity_host(
            self.h, *(_dp(a) for a in arrs)))

    def particles_pass_counts(self):
        """particles that entered each relocation pass of the last
        pass-by-pass step (ssx_particles_pass_counts)"""
        a = (C.c_int * 12)()
        check(self._lib.ssx_particles_pass_counts(self.h, a))
        return list(a)

    def particles_step_host(self, p, t_now, dt):
        """p: dict of contiguous numpy arrays (int32 el, old_el, reinit;
        float64 x, y, z, u, v, w, t, tau), updated in place"""
        for k in ("el", "old_el", "reinit"):
            assert p[k].dtype == np.int32 and p[k].flags.c_contiguous
        for k in ("x", "y", "z", "u", "v", "w", "t", "tau"):
            assert p[k].dtype == np.float64 and p[k].flags.c_contiguous
        ps = ParticlesStruct(int(p["x"].size), *(p[k].ctypes.data for k in
                             ("el", "old_el", "x", "y", "z", "u", "v", "w",
                              "t", "tau", "reinit")))
        mp = C.c_int()
        check(self._lib.ssx_particles_step_host(self.h, C.byref(ps), t_now,
                                                dt, C.byref(mp)))
        return mp.value

    def particles_step_host_noalloc(self, p, t_now, dt):
        """particles_step_host through ssx_particles_step_host_noalloc (no
        device allocation; at most NOALLOC_MAX_PARTICLES particles)"""
        for k in ("el", "old_el", "reinit"):
            assert p[k].dtype == np.int32 and p[k].flags.c_contiguous
        for k in ("x", "y", "z", "u", "v", "w", "t", "tau"):
            assert p[k].dtype == np.float64 and p[k].flags.c_contiguous
        ps = ParticlesStruct(int(p["x"].size), *(p[k].ctypes.data for k in
                             ("el", "old_el", "x", "y", "z", "u", "v", "w",
                              "t", "tau", "reinit")))
        mp = C.c_int()
        check(self._lib.ssx_particles_step_host_noalloc(
            self.h, C.byref(ps), t_now, dt, C.byref(mp)))
        return mp.value

    def tracers_step_host(self, p, t_now, dt):
        """RK2 tracers on host arrays (ssx_tracers_step_host); p as
        particles_step_host, tau optional"""
        for k in ("el", "old_el", "reinit"):
            assert p[k].dtype == np.int32 and p[k].flags.c_contiguous
        for k in ("x", "y", "z", "u", "v", "w", "t"):
            assert p[k].dtype == np.float64 and p[k].flags.c_contiguous
        ps = ParticlesStruct(int(p["x"].size), *(
            p[k].ctypes.data if p.get(k) is not None else None for k in
            ("el", "old_el", "x", "y", "z", "u", "v", "w", "t", "tau",
             "reinit")))
        mp = C.c_int()
        check(self._lib.ssx_tracers_step_host(self.h, C.byref(ps), t_now,
                                              dt, C.byref(mp)))
        return mp.value


class Batch:
    """Independent Jacobi-PCG problems solved together (ssx_batch_*):
    problems = [(context, op, var), ...]; a context may repeat."""

    def __init__(self, problems):
        self._lib = load()
        self._ctxs = [p[0] for p in problems]     # keep them alive
        n = len(problems)
        hs = (C.c_void_p * n)(*[p[0].h.value for p in problems])
        ops = (C.c_int * n)(*[p[1] for p in problems])
        var = (C.c_int * n)(*[p[2] for p in problems])
        h = C.c_void_p()
        check(self._lib.ssx_batch_create(C.byref(h), hs, ops, var, n))
        self.h = h
        self.n = n

    def close(self):
        if getattr(self, "h", None):
            self._lib.ssx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_stream(self, stream):
        handle = None if stream is None else C.c_void_p(stream.cuda_stream)
        check(self._lib.ssx_batch_set_stream(self.h, handle))

    def cg(self, fs, us, eps, max_iter, use_graph=True, allow_notconv=False,
           sync=True):
        """fs, us: lists of float64 device tensors; returns a list of
        Stats.  sync: synchronise the device first (torch's streams are not
        the contexts'; the batch itself waits for the contexts' streams)."""
        import torch
        if sync:
            torch.cuda.synchronize()
        n = self.n
        fp = (C.c_void_p * n)(*[f.data_ptr() for f in fs])
        up = (C.c_void_p * n)(*[u.data_ptr() for u in us])
        ep = (C.c_double * n)(*[float(e) for e in eps])
        st = (Stats * n)()
        rc = self._lib.ssx_batch_cg(self.h, fp, up, ep, max_iter,
                                    1 if use_graph else 0, st)
        if not (allow_notconv and rc == 4):
            check(rc)
        return list(st)

    def solve(self, precon, fs, us, eps, max_iter, use_graph=True,
              allow_notconv=False, sync=True):
        """ssx_batch_solve: precon[i] PC_MG (FCG + p-multigrid) or
        PC_JACOBI (Jacobi-PCG) per problem; fs, us as cg"""
        import torch
        if sync:
            torch.cuda.synchronize()
        n = self.n
        pc = (C.c_int * n)(*[int(p) for p in precon])
        fp = (C.c_void_p * n)(*[f.data_ptr() for f in fs])
        up = (C.c_void_p * n)(*[u.data_ptr() for u in us])
        ep = (C.c_double * n)(*[float(e) for e in eps])
        st = (Stats * n)()
        rc = self._lib.ssx_batch_solve(self.h, pc, fp, up, ep, max_iter,
                                       1 if use_graph else 0, st)
        if not (allow_notconv and rc == 4):
            check(rc)
        return list(st)


class NavierStokes:
    """The reference's Navier-Stokes time step on a context's mesh
    (ssx_ns_*), every field on the GPU."""

    def __init__(self, ctx, dt, visc, dens=1.0, t=0.0, g=(0.0, 0.0, 0.0),
                 batch_velocity=True):
        self._lib = load()
        self._ctx = ctx                                 # keep it alive
        prm = NSParams(float(dt), float(visc), float(dens), float(t),
                       (C.c_double * 3)(*[float(x) for x in g]),
                       1 if batch_velocity else 0)
        h = C.c_void_p()
        check(self._lib.ssx_ns_create(C.byref(h), ctx.h, C.byref(prm)))
        self.h = h
        self.ntot = ctx.ntot(0)

    def close(self):
        if getattr(self, "h", None):
            self._lib.ssx_ns_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init(self, levels):
        """levels[l][c]: float64 device tensor of velocity component c at
        t - (2 - l) dt (l = 0 oldest)"""
        arr = (C.c_void_p * 9)(*[levels[l][c].data_ptr() for l in range(3)
                                 for c in range(3)])
        check(self._lib.ssx_ns_init(self.h, arr))

    def step(self):
        st = (Stats * 4)()
        check(self._lib.ssx_ns_step(self.h, st))
        return list(st)

    def read(self, which, level=0):
        import torch
        out = torch.empty(self.ntot, dtype=torch.float64, device="cuda")
        check(self._lib.ssx_ns_read(self.h, which, level,
                                    C.c_void_p(out.data_ptr())))
        return out

    @property
    def time(self):
        return self._lib.ssx_ns_time(self.h)
