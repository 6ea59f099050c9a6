"""GPU path (libssx HIP kernels, called through the C ABI) against the golden
vectors of the compiled reference and against the oracle.

Bars (also stated in each assert): FP64 Ax / assembled operator 1e-12
relative; DSS bit-exact (integer and random fields: the add order is the
reference's); multiplicity and mask arrays exact; CG/FCG iteration counts
within one of the reference; solutions within the solve tolerance."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ssenss_amd as S  # noqa: E402
from ssenss_amd.mesh import box_mesh, NOSLIP, PERIODIC  # noqa: E402
from conftest import golden, golden_mesh, rel, CASES  # noqa: E402
from oracle import pyoracle as P  # noqa: E402

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def make_ctx(mesh):
    ctx = S.Context(mesh)
    ctx.set_stream(torch.cuda.current_stream())
    return ctx


@pytest.fixture(scope="module", params=CASES)
def gcase(request):
    Z = golden(request.param)
    mesh = golden_mesh(Z)
    print(f"[gpu] case {request.param} n={mesh.n} nel={mesh.nel}", flush=True)
    ctx = make_ctx(mesh)
    yield request.param, Z, mesh, ctx
    ctx.close()


def test_level_data(gcase):
    name, Z, mesh, ctx = gcase
    assert ctx.nlevels == len(Z["levels"])
    for l in range(ctx.nlevels):
        assert ctx.level_n(l) == Z["levels"][l]
        L = Z[f"ref_level{l}"].reshape(4, -1)
        assert rel(ctx.level_data(l, S.DATA_DIAG_P), L[0]) <= 1e-13
        assert rel(ctx.level_data(l, S.DATA_DIAG_H), L[1]) <= 1e-13
        assert np.array_equal(ctx.level_data(l, S.DATA_SCALE), L[2])
        assert np.array_equal(ctx.level_data(l, S.DATA_MASK), L[3])
    assert rel(ctx.level_data(0, S.DATA_INVMASS), Z["ref_invmass"]) <= 1e-13


def test_ax_poisson_helmholtz(gcase):
    name, Z, mesh, ctx = gcase
    u = dev(Z["in_u"])
    au = torch.zeros_like(u)
    ctx.ax(0, S.OP_POISSON, u, au)
    assert rel(host(au), Z["ref_ax_poisson"]) <= 1e-12       # FP64 bar
    ctx.ax(0, S.OP_HELMHOLTZ, u, au)
    assert rel(host(au), Z["ref_ax_helmholtz"]) <= 1e-12


def test_opeval(gcase):
    name, Z, mesh, ctx = gcase
    u = dev(Z["in_u"])
    au = torch.zeros_like(u)
    ctx.opeval(0, S.OP_POISSON, S.PRESSURE, u, au)
    assert rel(host(au), Z["ref_opeval_pressure"]) <= 1e-12
    ctx.opeval(0, S.OP_HELMHOLTZ, S.X_VELOCITY, u, au)
    assert rel(host(au), Z["ref_opeval_helm_vel"]) <= 1e-12


def test_dss_bit_exact(gcase):
    name, Z, mesh, ctx = gcase
    nt = mesh.ntot
    q = np.arange(nt, dtype=np.uint64)
    iv = ((q * np.uint64(2654435761)) % np.uint64(1000003)).astype(np.float64)
    t = dev(iv)
    ctx.dss(0, t)
    assert np.array_equal(host(t), Z["ref_dss_int"])
    t = dev(Z["in_u"])
    ctx.dss(0, t)
    assert np.array_equal(host(t), Z["ref_dss_u"])
    t = dev(Z["in_u"])
    ctx.dss_average(0, t)
    assert np.array_equal(host(t), Z["ref_dss_avg"])


def test_dot_nullspace_transfers(gcase):
    name, Z, mesh, ctx = gcase
    u = dev(Z["in_u"])
    one = torch.ones_like(u)
    d = Z["ref_dots"]
    assert abs(ctx.dot(0, u, u) / d[0] - 1) <= 1e-13
    assert abs(ctx.dot(0, u, one) - d[1]) <= 1e-12 * np.sqrt(d[0] * d[2])
    assert ctx.dot(0, one, one) == d[2]
    t = dev(Z["in_u"])
    ctx.nullspace(0, S.PRESSURE, t)
    assert rel(host(t), Z["ref_nullspace"]) <= 1e-13
    if ctx.nlevels > 1:
        o = P.Oracle(mesh)
        ref = o.dss_average(1, Z["ref_restrict_raw"])
        c = torch.zeros(ctx.ntot(1), dtype=torch.float64, device="cuda")
        ctx.restrict(1, u, c)
        assert rel(host(c), ref) <= 1e-13
        f = torch.zeros_like(u)
        ctx.prolong(0, dev(Z["in_v1"]), f)
        assert rel(host(f), Z["ref_prolong"]) <= 1e-13


def test_mg_precon_application(gcase):
    name, Z, mesh, ctx = gcase
    eps = Z["ref_scalars"][0]
    b = dev(Z["in_f"])
    pb = torch.zeros_like(b)
    ctx.set_diag_choice(-1)
    ctx.mg_precon(S.OP_POISSON, S.PRESSURE, eps, b, pb)
    assert rel(host(pb), Z["ref_mg_pressure"]) <= 1e-9


def test_cg_helmholtz(gcase):
    name, Z, mesh, ctx = gcase
    sc = Z["ref_scalars"]
    eps = sc[0]
    f = dev(Z["in_f"])
    u = torch.zeros_like(f)
    ctx.set_diag_choice(-1)
    st = ctx.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, f, u, eps, 5000)
    print(f"[gpu] {name} cg helmholtz iters {st.iterations} ref {int(sc[1])}")
    assert abs(st.iterations - int(sc[1])) <= 1 and st.converged
    assert rel(host(u), Z["ref_cg_helm"]) <= 1e-8


def test_fcg_mg_pressure(gcase):
    name, Z, mesh, ctx = gcase
    sc = Z["ref_scalars"]
    eps = sc[0]
    f = dev(Z["in_f"])
    u = torch.zeros_like(f)
    ctx.set_diag_choice(-1)
    st = ctx.fcg(S.OP_POISSON, S.PRESSURE, S.PC_MG, f, u, eps, 5000)
    print(f"[gpu] {name} fcg pressure iters {st.iterations} ref {int(sc[3])}")
    assert abs(st.iterations - int(sc[3])) <= 1 and st.converged
    assert rel(host(u), Z["ref_fcg_pressure"]) <= 1e-8


def test_fcg_mg_dirichlet_poisson(gcase):
    name, Z, mesh, ctx = gcase
    sc = Z["ref_scalars"]
    eps = sc[0]
    f = dev(Z["in_f"])
    u = torch.zeros_like(f)
    ctx.set_diag_choice(0)
    try:
        st = ctx.fcg(S.OP_POISSON, S.X_VELOCITY, S.PC_MG, f, u, eps, 5000)
    finally:
        ctx.set_diag_choice(-1)
    print(f"[gpu] {name} fcg dirichlet iters {st.iterations} ref {int(sc[5])}")
    assert abs(st.iterations - int(sc[5])) <= 1 and st.converged
    assert rel(host(u), Z["ref_fcg_dirichlet"]) <= 1e-8


def test_host_array_entry_points():
    Z = golden("cube2_n8")
    mesh = golden_mesh(Z)
    sc = Z["ref_scalars"]
    with S.Context(mesh) as ctx:
        f = Z["in_f"].copy()
        u = np.zeros(mesh.ntot)
        st = ctx.solve_helmholtz_cg(f, u, S.X_VELOCITY, sc[0])
        assert abs(st.iterations - int(sc[1])) <= 1
        assert rel(u, Z["ref_cg_helm"]) <= 1e-8
        f = Z["in_f"].copy()
        u = np.zeros(mesh.ntot)
        st = ctx.solve_poisson_cg(f, u, S.PRESSURE, sc[0])
        assert abs(st.iterations - int(sc[3])) <= 1
        assert rel(u, Z["ref_fcg_pressure"]) <= 1e-8


@pytest.mark.parametrize("n", list(range(2, 17)))
def test_ax_every_order_against_oracle(n):
    """All compiled orders, ragged element counts, curved elements."""
    mesh = box_mesh(n, (3, 2, 5), bc=(NOSLIP, NOSLIP, NOSLIP, NOSLIP,
                                      PERIODIC, PERIODIC),
                    deform=1, amp=0.07, dt=0.01, visc=0.5)
    o = P.Oracle(mesh)
    rng = np.random.default_rng(100 + n)
    u = rng.standard_normal(mesh.ntot)
    with make_ctx(mesh) as ctx:
        du, au = dev(u), torch.zeros(mesh.ntot, dtype=torch.float64,
                                     device="cuda")
        for op in (S.OP_POISSON, S.OP_HELMHOLTZ):
            ctx.ax(0, op, du, au)
            assert rel(host(au), o.ax(0, op, u)) <= 1e-12, (n, op)
        ctx.opeval(0, S.OP_POISSON, S.X_VELOCITY, du, au)
        assert rel(host(au), o.opeval(0, S.OP_POISSON, S.X_VELOCITY, u)) \
            <= 1e-12
        t = dev(u)
        ctx.dss(0, t)
        assert np.array_equal(host(t), o.dss(0, u))


def test_errors_are_reported():
    Z = golden("box2_n5")
    mesh = golden_mesh(Z)
    with make_ctx(mesh) as ctx:
        f = dev(Z["in_f"])
        u = torch.zeros_like(f)
        st = ctx.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, f, u, 1e-30, 3,
                    allow_notconv=True)
        assert st.iterations == 3 and not st.converged
        with pytest.raises(S.SsxError) as e:
            ctx.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, f, u, 1e-30, 3)
        assert e.value.code == 4
        with pytest.raises(S.SsxError):
            ctx.ax(7, S.OP_POISSON, f, u)


def test_full_size_properties():
    """32^3 elements at n = 8 (16.8M local nodes): size-independent
    properties at the benchmark size."""
    mesh = box_mesh(8, (32, 32, 32), deform=1, amp=0.1)
    print("[gpu] full size 32^3 n=8", flush=True)
    with make_ctx(mesh) as ctx:
        nt = mesh.ntot
        one = torch.ones(nt, dtype=torch.float64, device="cuda")
        assert ctx.dot(0, one, one) == 225 ** 3        # unique nodes
        au = torch.zeros_like(one)
        ctx.ax(0, S.OP_POISSON, one, au)
        g = ctx.level_data(0, S.DATA_GFAC)
        assert host(au.abs().max()) <= 1e-12 * np.abs(g).max() * 64
        # integer-valued DSS equals the host assembly bit for bit
        q = torch.arange(nt, dtype=torch.int64, device="cuda")
        iv = ((q * 2654435761) % 1000003).to(torch.float64)
        ref = iv.cpu().numpy().copy()
        import ctypes as C
        ms = mesh.struct()
        S.native.check(S.load().ssx_dss_host(
            C.byref(ms), 8, ref.ctypes.data_as(C.POINTER(C.c_double))))
        ctx.dss(0, iv)
        assert np.array_equal(host(iv), ref)
        # symmetry of the assembled operator on continuous vectors: the
        # unmasked operator, and the masked one on masked vectors (the
        # solvers keep p masked, solver.c:72)
        rng = np.random.default_rng(7)
        x = dev(rng.standard_normal(nt))
        y = dev(rng.standard_normal(nt))
        ctx.dss_average(0, x)
        ctx.dss_average(0, y)
        ax = torch.zeros_like(x)
        ay = torch.zeros_like(y)
        for var in (S.PRESSURE, S.X_VELOCITY):
            ctx.mask(0, var, x)
            ctx.mask(0, var, y)
            ctx.opeval(0, S.OP_POISSON, var, x, ax)
            ctx.opeval(0, S.OP_POISSON, var, y, ay)
            a, b = ctx.dot(0, y, ax), ctx.dot(0, x, ay)
            assert abs(a - b) <= 1e-10 * abs(a), (var, a, b)


@pytest.mark.parametrize("name", ["cube2_n8", "periodic3_n6", "channel_n6"])
def test_fused_cg_matches_plain(name):
    """The fused six-kernel CG iteration (device-side stopping test) against
    the plain one-kernel-per-operation loop: same iteration counts (within
    one) and solutions within the solve tolerance, for Helmholtz-Jacobi,
    pressure FCG + p-MG and Dirichlet FCG + p-MG."""
    Z = golden(name)
    mesh = golden_mesh(Z)
    sc = Z["ref_scalars"]
    eps = sc[0]
    out = {}
    with make_ctx(mesh) as ctx:
        for fused in (1, 0):
            ctx.set_option(S.OPT_FUSED_CG, fused)
            f = dev(Z["in_f"])
            u = torch.zeros_like(f)
            st1 = ctx.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, f, u, eps, 5000)
            f = dev(Z["in_f"])
            v = torch.zeros_like(f)
            st2 = ctx.fcg(S.OP_POISSON, S.PRESSURE, S.PC_MG, f, v, eps, 5000)
            out[fused] = (st1, host(u), st2, host(v))
        ctx.set_option(S.OPT_FUSED_CG, 1)
    a, b = out[1], out[0]
    print(f"[gpu] {name} fused/plain cg {a[0].iterations}/{b[0].iterations}"
          f" fcg {a[2].iterations}/{b[2].iterations}"
          f" coarse {a[2].coarse_iters}/{b[2].coarse_iters}")
    assert abs(a[0].iterations - b[0].iterations) <= 1
    assert abs(a[2].iterations - b[2].iterations) <= 1
    assert abs(a[0].iterations - int(sc[1])) <= 1
    assert abs(a[2].iterations - int(sc[3])) <= 1
    assert rel(a[1], b[1]) <= 1e-8 and rel(a[3], b[3]) <= 1e-8
    assert rel(a[1], Z["ref_cg_helm"]) <= 1e-8
    assert rel(a[3], Z["ref_fcg_pressure"]) <= 1e-8


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("name", ["cube2_n8", "periodic3_n6", "channel_n6",
                                  "cube2_n12"])
def test_mixed_precision_smoother(name, prec):
    """FP32 (prec 1) / FP32 + BF16-factor (prec 2) multigrid smoother and
    transfers under the FP64 flexible CG: reaches the same FP64 tolerance;
    the solution agrees with the reference's FP64 solution to the level the
    tolerance fixes (both residuals <= 1e-10 ||b||_w), in the reference's
    FP64 iteration count to within one."""
    Z = golden(name)
    mesh = golden_mesh(Z)
    sc = Z["ref_scalars"]
    eps = sc[0]
    with make_ctx(mesh) as ctx:
        ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
        for var, ref, it_ref in ((S.PRESSURE, "ref_fcg_pressure", sc[3]),
                                 (S.X_VELOCITY, "ref_fcg_dirichlet", sc[5])):
            ctx.set_diag_choice(-1 if var == S.PRESSURE else 0)
            f = dev(Z["in_f"])
            u = torch.zeros_like(f)
            st = ctx.fcg(S.OP_POISSON, var, S.PC_MG, f, u, eps, 500)
            err = rel(host(u), Z[ref])
            print(f"[gpu] {name} prec {prec} var {var}: {st.iterations} its "
                  f"(fp64 ref {int(it_ref)}), |x - x_ref|/|x_ref| = {err:.2e}")
            assert st.converged and st.residual <= eps
            assert abs(st.iterations - int(it_ref)) <= 1
            assert err <= 1e-7
        ctx.set_option(S.OPT_SMOOTHER_PRECISION, 0)
        ctx.set_diag_choice(-1)


def _dev_particles(rec):
    n = len(rec)
    ip = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).cuda()
    return {"el": ip(rec[:, 0]), "old_el": ip(rec[:, 0]),
            "x": dev(rec[:, 1]), "y": dev(rec[:, 2]), "z": dev(rec[:, 3]),
            "u": dev(rec[:, 4]), "v": dev(rec[:, 5]), "w": dev(rec[:, 6]),
            "t": dev(rec[:, 7]), "tau": dev(rec[:, 8]),
            "reinit": ip(np.zeros(n))}


def _host_particles(p):
    torch.cuda.synchronize()
    cols = [p["el"].cpu().numpy().astype(np.float64)] + [
        p[k].cpu().numpy() for k in ("x", "y", "z", "u", "v", "w", "t")]
    return np.stack(cols, 1), p["reinit"].cpu().numpy()


def _run_gpu_particles(ctx, vel, scales, rec, t_now, dt):
    nt = ctx.ntot(0)
    for s in scales:
        ctx.particles_set_velocity(dev(s * vel[:nt]), dev(s * vel[nt:2 * nt]),
                                   dev(s * vel[2 * nt:]))
    p = _dev_particles(rec)
    passes = ctx.particles_step(p, t_now, dt)
    out, re = _host_particles(p)
    return out, re, passes


@pytest.mark.parametrize("name", ["box2_n5", "periodic3_n6", "channel_n6"])
def test_particles_against_reference(name):
    """owning elements bit-exact against the compiled reference; states
    within 1e-12 (the engine's GLL basis and J^-1 agree with the
    reference's to rounding)"""
    Z = golden(name)
    h = Z["particle_hdr"]
    with make_ctx(golden_mesh(Z)) as ctx:
        out, re, passes = _run_gpu_particles(ctx, Z["in_vel"], h[2:],
                                             Z["in_particles"], h[0], h[1])
    ref = Z["ref_particles"].reshape(-1, 8)
    print(f"[gpu] particles {name}: passes {passes}, max state diff "
          f"{np.abs(out[:, 1:] - ref[:, 1:]).max():.2e}")
    assert re.sum() == 0
    assert np.array_equal(out[:, 0], ref[:, 0])
    assert np.abs(out[:, 1:] - ref[:, 1:]).max() <= 1e-12


def test_particles_many_hops_against_oracle():
    """20k particles, fast and heavy, several hops per step on a periodic
    4^3 mesh at n = 8 (stencil interpolation): elements bit-exact against
    the oracle"""
    mesh = box_mesh(8, (4, 4, 4), lo=(0, 0, 0), hi=(2 * np.pi,) * 3,
                    bc=(PERIODIC,) * 6)
    rng = np.random.default_rng(31)
    X, Y, Zc = mesh.x, mesh.y, mesh.z
    vel = np.concatenate([(np.sin(X) * np.cos(Y) * np.cos(Zc)).ravel(),
                          (-np.cos(X) * np.sin(Y) * np.cos(Zc)).ravel(),
                          (0.5 * np.sin(X + Zc)).ravel()])
    n = 20000
    rec = np.zeros((n, 9))
    rec[:, 0] = rng.integers(0, mesh.nel, n)
    rec[:, 1:4] = rng.uniform(-0.99, 0.99, (n, 3))
    rec[:, 4:7] = 4.0 * rng.standard_normal((n, 3))
    rec[:, 7] = 1.0 - 2 * 0.2
    rec[:, 8] = np.exp(rng.uniform(np.log(0.1), 0.0, n))
    o = P.Oracle(mesh)
    ref, re_o, passes_o = o.particles(vel, (0.8, 0.9, 1.0), rec, 1.0, 0.2)
    with make_ctx(mesh) as ctx:
        out, re, passes = _run_gpu_particles(ctx, vel, (0.8, 0.9, 1.0), rec,
                                             1.0, 0.2)
    hops = (out[:, 0] != rec[:, 0]).sum()
    print(f"[gpu] 20k particles: {hops} changed element, passes {passes} "
          f"(oracle {passes_o}), reinit {re.sum()}")
    d = np.abs(out[:, 1:] - ref[:, 1:]).max(axis=1)
    print(f"[gpu] 20k particles state diff: median {np.median(d):.1e}, "
          f"99th pct {np.percentile(d, 99):.1e}, max {d.max():.1e}")
    assert hops > 2000
    assert np.array_equal(out[:, 0], ref[:, 0])
    assert np.array_equal(re, re_o)
    # The engine's GLL basis and J^-1 agree with the oracle's to rounding
    # (~1e-15, giving ~1e-13 after one pass); up to 12 linearised steps to
    # the boundary, reflections and re-entries per particle amplify that on
    # a few trajectories, so states are held to 1e-8 for 99% of particles
    # and 1e-4 for all; owning elements and re-seeds stay exact.
    assert np.median(d) <= 1e-11
    assert np.percentile(d, 99) <= 1e-8
    assert d.max() <= 1e-4


def test_particles_bad_element_is_an_error():
    """an element index outside the mesh is reported, not dereferenced;
    the other particles still advance"""
    Z = golden("periodic3_n6")
    h = Z["particle_hdr"]
    rec = Z["in_particles"][:64].copy()
    rec[5, 0] = 10 ** 6
    rec[9, 0] = -3
    with make_ctx(golden_mesh(Z)) as ctx:
        nt = ctx.ntot(0)
        vel = Z["in_vel"]
        for s in h[2:]:
            ctx.particles_set_velocity(dev(s * vel[:nt]),
                                       dev(s * vel[nt:2 * nt]),
                                       dev(s * vel[2 * nt:]))
        p = _dev_particles(rec)
        with pytest.raises(S.SsxError, match="2 particles"):
            ctx.particles_step(p, h[0], h[1])
        out, _ = _host_particles(p)
    assert out[5, 0] == 10 ** 6 and out[5, 1] == rec[5, 1]
    assert out[9, 0] == -3 and out[9, 1] == rec[9, 1]
    assert np.all(out[:, 7][[0, 1, 2]] == h[0])


@pytest.mark.parametrize("grid", [1, 3])
@pytest.mark.parametrize("name", ["cube2_n8", "periodic3_n6", "channel_n6",
                                  "cube2_n12"])
def test_persistent_coarse_cg(name, grid):
    """The coarse-level CG as one cooperative launch (grid 1: blocks as the
    device keeps co-resident; 3: three blocks, so every phase loops) against
    six launches per iteration: Jacobi-PCG on the coarsest level for both
    operators and both variable kinds, and the full FCG + p-MG solves --
    iteration counts within one, solutions within the solve tolerance."""
    Z = golden(name)
    mesh = golden_mesh(Z)
    sc = Z["ref_scalars"]
    eps = sc[0]
    with make_ctx(mesh) as ctx:
        ctx.set_option(S.OPT_COARSE_ASSEMBLED, 0)     # the E-vector CG
        lc = ctx.nlevels - 1
        rng = np.random.default_rng(5)
        fc = rng.standard_normal(ctx.ntot(lc))
        res = {}
        for opt in (0, grid):
            ctx.set_option(S.OPT_PERSISTENT_CG, opt)
            runs = []
            for op, var in ((S.OP_POISSON, S.PRESSURE),
                            (S.OP_HELMHOLTZ, S.X_VELOCITY)):
                f = dev(fc)
                ctx.dss_average(lc, f)
                u = torch.zeros_like(f)
                st = ctx.cg(lc, op, var, f, u, 1e-9 * float(f.norm()), 500)
                runs.append((st.iterations, host(u)))
            ctx.set_diag_choice(-1)
            f = dev(Z["in_f"])
            v = torch.zeros_like(f)
            st = ctx.fcg(S.OP_POISSON, S.PRESSURE, S.PC_MG, f, v, eps, 5000)
            runs.append((st.iterations, host(v), st.coarse_iters))
            ctx.set_diag_choice(0)
            f = dev(Z["in_f"])
            v = torch.zeros_like(f)
            st = ctx.fcg(S.OP_POISSON, S.X_VELOCITY, S.PC_MG, f, v, eps, 5000)
            runs.append((st.iterations, host(v), st.coarse_iters))
            ctx.set_diag_choice(-1)
            res[opt] = runs
        ctx.set_option(S.OPT_PERSISTENT_CG, 1)
    a, b = res[grid], res[0]
    print(f"[gpu] persistent {name} grid {grid}: coarse cg "
          f"{a[0][0]}/{b[0][0]}, {a[1][0]}/{b[1][0]}; fcg {a[2][0]}/{b[2][0]} "
          f"(coarse {a[2][2]}/{b[2][2]}), {a[3][0]}/{b[3][0]}")
    for k in range(4):
        assert abs(a[k][0] - b[k][0]) <= 1, k
    for k in range(2):
        assert rel(a[k][1], b[k][1]) <= 1e-7, k
    assert rel(a[2][1], Z["ref_fcg_pressure"]) <= 1e-8
    assert rel(a[3][1], Z["ref_fcg_dirichlet"]) <= 1e-8
    assert abs(a[2][0] - int(sc[3])) <= 1 and abs(a[3][0] - int(sc[5])) <= 1


def test_persistent_cg_abort_is_reported():
    """the persistent CG's failure path, forced from the host: the launch
    starts with its abort word set, every block leaves at its first grid
    barrier (no wait), the solve reports SSX_EHIP; the next solve on the
    context runs normally"""
    Z = golden("cube2_n8")
    with make_ctx(golden_mesh(Z)) as ctx:
        ctx.set_option(S.OPT_COARSE_ASSEMBLED, 0)
        ctx.set_option(S.OPT_PERSISTENT_CG, 1)
        lc = ctx.nlevels - 1
        assert ctx.level_n(lc) <= 4
        rng = np.random.default_rng(6)
        fc = dev(rng.standard_normal(ctx.ntot(lc)))
        ctx.dss_average(lc, fc)
        ctx.set_option(S.OPT_PERSIST_ABORT_TEST, 1)
        f = fc.clone()
        u = torch.zeros_like(f)
        with pytest.raises(S.SsxError) as e:
            ctx.cg(lc, S.OP_POISSON, S.X_VELOCITY, f, u, 1e-12, 500)
        assert e.value.code == 3 and "barrier" in str(e.value)
        f = fc.clone()
        u = torch.zeros_like(f)
        st = ctx.cg(lc, S.OP_POISSON, S.X_VELOCITY, f, u, 1e-12, 500)
        assert st.converged
        print(f"[gpu] persistent CG after the forced abort: "
              f"{st.iterations} its")


def test_particles_host_entry_matches_device_entry():
    """ssx_particles_*_host (host arrays, the drop-in for the reference's
    host particle list) gives bit for bit what the device entry gives"""
    Z = golden("channel_n6")
    h = Z["particle_hdr"]
    rec = Z["in_particles"]
    vel = Z["in_vel"]
    with make_ctx(golden_mesh(Z)) as ctx:
        out_d, re_d, pass_d = _run_gpu_particles(ctx, vel, h[2:], rec, h[0],
                                                 h[1])
    with make_ctx(golden_mesh(Z)) as ctx:
        nt = ctx.ntot(0)
        for s in h[2:]:
            ctx.particles_set_velocity_host(s * vel[:nt], s * vel[nt:2 * nt],
                                            s * vel[2 * nt:])
        ip = lambda a: np.ascontiguousarray(a, np.int32)
        fp = lambda a: np.ascontiguousarray(a, np.float64)
        p = {"el": ip(rec[:, 0]), "old_el": ip(rec[:, 0]),
             "reinit": np.zeros(len(rec), np.int32)}
        for c, k in enumerate(("x", "y", "z", "u", "v", "w", "t", "tau")):
            p[k] = fp(rec[:, 1 + c])
        pass_h = ctx.particles_step_host(p, h[0], h[1])
    out_h = np.stack([p["el"].astype(np.float64)] + [
        p[k] for k in ("x", "y", "z", "u", "v", "w", "t")], 1)
    assert pass_h == pass_d
    assert np.array_equal(out_h, out_d) and np.array_equal(p["reinit"], re_d)


@pytest.mark.parametrize("name", ["box2_n5", "cube2_n8", "cube2_n12"])
def test_operator_launch_variants_bit_identical(name):
    """every launch shape of the element operator (workgroup size, factor
    prefetch) gives bit-identical results, in FP64 against the reference's
    operator to 1e-12; the autotuner's choice is one of them"""
    Z = golden(name)
    with make_ctx(golden_mesh(Z)) as ctx:
        v0, us = ctx.ax_variant(0, 0)
        print(f"[gpu] {name} autotuned variant {v0}, us per launch "
              f"{np.round(us, 1).tolist()}")
        assert 0 <= v0 < S.AX_NVARIANTS and us.min() > 0
        u = dev(Z["in_u"])
        outs = {}
        for op in (S.OP_POISSON, S.OP_HELMHOLTZ):
            for v in range(S.AX_NVARIANTS):
                ctx.set_ax_variant(0, 0, v)
                au = torch.zeros_like(u)
                ctx.ax(0, op, u, au)
                outs[(op, v)] = host(au)
            for v in range(1, S.AX_NVARIANTS):
                assert np.array_equal(outs[(op, v)], outs[(op, 0)]), (op, v)
        assert rel(outs[(S.OP_POISSON, 0)], Z["ref_ax_poisson"]) <= 1e-12
        assert rel(outs[(S.OP_HELMHOLTZ, 0)], Z["ref_ax_helmholtz"]) <= 1e-12
        # the reduced-precision copies are tuned when first built; forcing
        # each shape leaves the FP64 and the mixed solves (FP32 and BF16
        # factors) unchanged bit for bit (one operator partial per element, whatever the shape), so a
        # solve does not depend on what the autotuner timed
        ctx.set_diag_choice(0)
        for prec in (S.PREC_FP64, S.PREC_FP32, S.PREC_BF16):
            ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
            its = []
            for v in range(S.AX_NVARIANTS):
                for l in range(ctx.nlevels - 1):
                    ctx.set_ax_variant(l, prec, v)
                    ctx.set_ax_variant(l, S.PREC_FP64, v)
                f = dev(Z["in_f"])
                x = torch.zeros_like(f)
                st = ctx.fcg(S.OP_POISSON, S.X_VELOCITY, S.PC_MG, f, x,
                             Z["ref_scalars"][0], 500)
                its.append((st.iterations, host(x)))
            assert all(i == its[0][0] for i, _ in its), prec
            assert all(np.array_equal(x, its[0][1]) for _, x in its), prec


def _batch_problems():
    """golden meshes of different orders and element counts, each with the
    reference's Jacobi-PCG Helmholtz solve, plus a pressure Poisson"""
    names = ["box2_n5", "cube2_n8", "periodic3_n6", "channel_n6",
             "cube2_n12"]
    return [(nm, golden(nm)) for nm in names]


def test_batch_cg_matches_single_solves_and_reference():
    """one launch per CG stage for five meshes of orders 4..11 (Helmholtz,
    x velocity) plus a pressure Poisson problem: iteration counts within
    one of the single-problem solver and of the reference, solutions within
    the solve tolerance of the reference's"""
    cases = _batch_problems()
    ctxs = [make_ctx(golden_mesh(Z)) for _, Z in cases]
    try:
        probs = [(c, S.OP_HELMHOLTZ, S.X_VELOCITY) for c in ctxs]
        probs.append((ctxs[1], S.OP_POISSON, S.PRESSURE))
        rhs = [Z["in_f"] for _, Z in cases] + [cases[1][1]["in_f"]]
        eps = [Z["ref_scalars"][0] for _, Z in cases]
        eps.append(cases[1][1]["ref_scalars"][0])
        batch = S.Batch(probs)
        try:
            fs = [dev(r) for r in rhs]
            us = [torch.zeros_like(f) for f in fs]
            stats = batch.cg(fs, us, eps, 5000)
        finally:
            batch.close()
        for k, (c, op, var) in enumerate(probs):
            f = dev(rhs[k])
            u = torch.zeros_like(f)
            st1 = c.cg(0, op, var, f, u, eps[k], 5000)
            name = cases[k][0] if k < len(cases) else "cube2_n8 pressure"
            print(f"[gpu] batch {name}: {stats[k].iterations} its "
                  f"(single {st1.iterations})")
            assert stats[k].converged
            assert abs(stats[k].iterations - st1.iterations) <= 1
            assert rel(host(us[k]), host(u)) <= 1e-8
            if k < len(cases):
                Z = cases[k][1]
                assert abs(stats[k].iterations - int(Z["ref_scalars"][1])) <= 1
                assert rel(host(us[k]), Z["ref_cg_helm"]) <= 1e-8
    finally:
        for c in ctxs:
            c.close()


def test_batch_many_small_problems_match_single_solves():
    """40 small Helmholtz problems of orders 2..7 on 2^3..4^3-element boxes
    (random deformation and lambda), more than one flag block of the batch:
    each problem's iteration count within one of the single-problem
    solver's, its solution within the solve tolerance, and the final
    solution step (taken after the last iteration) applied exactly once"""
    rng = np.random.default_rng(2026)
    ctxs, rhs = [], []
    try:
        for _ in range(40):
            n = int(rng.choice([3, 4, 5, 6, 8]))
            m = int(rng.integers(2, 5))
            lam = rng.uniform(1, 100)
            mesh = box_mesh(n, (m, m, m), deform=1, amp=rng.uniform(0, 0.15),
                            dt=(11.0 / 6.0) / lam, visc=1.0)
            c = make_ctx(mesh)
            ctxs.append(c)
            b = dev(rng.standard_normal(mesh.ntot)
                    * c.level_data(0, S.DATA_MASS))
            c.dss(0, b)
            rhs.append(host(b))
        eps = [1e-10 * np.linalg.norm(r) for r in rhs]
        batch = S.Batch([(c, S.OP_HELMHOLTZ, S.X_VELOCITY) for c in ctxs])
        try:
            fs = [dev(r) for r in rhs]
            us = [torch.zeros_like(f) for f in fs]
            stats = batch.cg(fs, us, eps, 5000)
        finally:
            batch.close()
        worst = 0
        for k, c in enumerate(ctxs):
            f = dev(rhs[k])
            u = torch.zeros_like(f)
            st1 = c.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, f, u, eps[k], 5000)
            assert stats[k].converged
            assert abs(stats[k].iterations - st1.iterations) <= 1
            assert rel(host(us[k]), host(u)) <= 1e-8
            worst = max(worst, abs(stats[k].iterations - st1.iterations))
        print(f"[gpu] batch of 40 small problems: largest iteration "
              f"difference {worst}")
    finally:
        for c in ctxs:
            c.close()


def test_batch_waits_for_work_on_the_context_stream():
    """the right-hand side is assembled by a DSS still queued on the
    context's own stream when the batch starts on its stream: the batch
    waits for it (its solution is that of the assembled right-hand side)"""
    Z = golden("cube2_n8")
    mesh = golden_mesh(Z)
    rng = np.random.default_rng(5)
    with S.Context(mesh) as ctx:             # its own non-blocking stream
        raw = rng.standard_normal(mesh.ntot) * ctx.level_data(0, S.DATA_MASS)
        ref_f = dev(raw)
        ref_u = torch.zeros_like(ref_f)
        torch.cuda.synchronize()
        ctx.dss(0, ref_f)
        ctx.synchronize()
        eps = 1e-10 * np.linalg.norm(host(ref_f))
        st1 = ctx.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, ref_f, ref_u, eps, 5000)
        batch = S.Batch([(ctx, S.OP_HELMHOLTZ, S.X_VELOCITY)])
        try:
            f = dev(raw)
            u = torch.zeros_like(f)
            torch.cuda.synchronize()
            for _ in range(3):                # queued on the context stream
                ctx.mask(0, S.X_VELOCITY, f)
            ctx.dss(0, f)
            st = batch.cg([f], [u], [eps], 5000, sync=False)
        finally:
            batch.close()
        print(f"[gpu] batch after context-stream work: {st[0].iterations} "
              f"its (single {st1.iterations})")
        assert st[0].converged
        assert abs(st[0].iterations - st1.iterations) <= 1
        assert rel(host(u), host(ref_u)) <= 1e-8


def test_batch_graph_replays_match_eager_launches():
    """the captured-graph iterations give bit for bit what eager launches
    give, replay after replay, with the program's own eager work and
    synchronisations in between; one context three times (three velocity
    components with different right-hand sides)"""
    Z = golden("channel_n6")
    with make_ctx(golden_mesh(Z)) as ctx:
        rng = np.random.default_rng(11)
        rhs = [Z["in_f"]] + [rng.standard_normal(ctx.ntot(0)) for _ in
                             range(2)]
        for r in rhs[1:]:
            t = dev(r)
            ctx.dss(0, t)
            r[:] = host(t)
        probs = [(ctx, S.OP_HELMHOLTZ, v) for v in (S.X_VELOCITY,
                                                     S.Y_VELOCITY,
                                                     S.Z_VELOCITY)]
        eps = [1e-10 * np.linalg.norm(r) for r in rhs]
        batch = S.Batch(probs)
        try:
            res = []
            for use_graph in (False, True, True, True):
                fs = [dev(r) for r in rhs]
                us = [torch.zeros_like(f) for f in fs]
                st = batch.cg(fs, us, eps, 5000, use_graph=use_graph)
                res.append(([s.iterations for s in st],
                            [host(u) for u in us]))
                junk = torch.randn(1 << 20, device="cuda").sum()   # eager
                torch.cuda.synchronize()
                del junk
        finally:
            batch.close()
        its0, sol0 = res[0]
        print(f"[gpu] batch graph replays: iterations {its0}")
        for its, sol in res[1:]:
            assert its == its0
            assert all(np.array_equal(a, b) for a, b in zip(sol, sol0))


@pytest.mark.parametrize("name", ["cube2_n8", "periodic3_n6", "channel_n6"])
def test_last_workgroup_reductions_bit_identical(name):
    """alpha and beta formed by the last workgroup of the DSS and update
    kernels (three launches per CG iteration) give bit for bit what the
    separate reduction launches give: Jacobi-PCG Helmholtz, FCG + p-MG
    pressure, and the FP32 mixed-precision solve"""
    Z = golden(name)
    eps = Z["ref_scalars"][0]
    out = {}
    with make_ctx(golden_mesh(Z)) as ctx:
        for fr in (0, 1):
            ctx.set_option(S.OPT_FUSED_REDUCE, fr)
            runs = []
            f = dev(Z["in_f"])
            u = torch.zeros_like(f)
            st = ctx.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, f, u, eps, 5000)
            runs.append((st.iterations, host(u)))
            for prec in (S.PREC_FP64, S.PREC_FP32):
                ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
                f = dev(Z["in_f"])
                v = torch.zeros_like(f)
                st = ctx.fcg(S.OP_POISSON, S.PRESSURE, S.PC_MG, f, v, eps, 5000)
                runs.append((st.iterations, host(v)))
            ctx.set_option(S.OPT_SMOOTHER_PRECISION, S.PREC_FP64)
            out[fr] = runs
        ctx.set_option(S.OPT_FUSED_REDUCE, 0)
    print(f"[gpu] last-workgroup reductions {name}: iterations "
          f"{[r[0] for r in out[1]]} / {[r[0] for r in out[0]]}")
    for a, b in zip(out[1], out[0]):
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("name", ["cube2_n8", "periodic3_n6", "channel_n6",
                                  "cube2_n12"])
def test_gather_dss_update_bit_identical(name):
    """A p assembled inside the CG update (each copy gathers its partners in
    its own add order) against the DSS launch followed by the update:
    bit-identical iteration counts and solutions for Jacobi-PCG Helmholtz,
    FCG + p-MG pressure in FP64, FP32 and BF16, and p-MG velocity Poisson
    with the E-vector coarsest level"""
    Z = golden(name)
    eps = Z["ref_scalars"][0]
    out = {}
    with make_ctx(golden_mesh(Z)) as ctx:
        for g in (0, 1):
            ctx.set_option(S.OPT_GATHER_DSS, g)
            runs = []
            f = dev(Z["in_f"])
            u = torch.zeros_like(f)
            st = ctx.cg(0, S.OP_HELMHOLTZ, S.X_VELOCITY, f, u, eps, 5000)
            runs.append((st.iterations, host(u)))
            for prec in (S.PREC_FP64, S.PREC_FP32, S.PREC_BF16):
                ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
                f = dev(Z["in_f"])
                v = torch.zeros_like(f)
                st = ctx.fcg(S.OP_POISSON, S.PRESSURE, S.PC_MG, f, v, eps, 5000)
                runs.append((st.iterations, host(v)))
            ctx.set_option(S.OPT_SMOOTHER_PRECISION, S.PREC_FP64)
            ctx.set_option(S.OPT_COARSE_ASSEMBLED, 0)
            f = dev(Z["in_f"])
            v = torch.zeros_like(f)
            st = ctx.fcg(S.OP_POISSON, S.X_VELOCITY, S.PC_MG, f, v, eps, 5000)
            runs.append((st.iterations, host(v)))
            ctx.set_option(S.OPT_COARSE_ASSEMBLED, 1)
            out[g] = runs
        ctx.set_option(S.OPT_GATHER_DSS, 1)
    print(f"[gpu] gather DSS update {name}: iterations "
          f"{[r[0] for r in out[1]]} / {[r[0] for r in out[0]]}")
    for a, b in zip(out[1], out[0]):
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("name", ["cube2_n8", "periodic3_n6", "channel_n6"])
def test_coarse_two_launch_bit_identical(name):
    """the assembled coarse CG with two launches per iteration (each kernel
    replays the reduction it needs) against four launches (separate
    reductions): bit-identical coarse solves for both operators and
    variable kinds, and bit-identical FCG + p-MG solves in FP64 and FP32"""
    Z = golden(name)
    mesh = golden_mesh(Z)
    eps = Z["ref_scalars"][0]
    with make_ctx(mesh) as ctx:
        lc = ctx.nlevels - 1
        rng = np.random.default_rng(11)
        fc = rng.standard_normal(ctx.ntot(lc))
        res = {}
        for opt in (1, 2):
            ctx.set_option(S.OPT_COARSE_ASSEMBLED, opt)
            runs = []
            for op, var in ((S.OP_POISSON, S.PRESSURE),
                            (S.OP_HELMHOLTZ, S.X_VELOCITY)):
                for mi in (500, 3):          # converged, stopped by max_iter
                    f = dev(fc)
                    ctx.dss_average(lc, f)
                    u = torch.zeros_like(f)
                    st = ctx.cg(lc, op, var, f, u, 1e-9 * float(f.norm()), mi,
                                allow_notconv=True)
                    runs.append((st.iterations, st.residual, host(u)))
            for prec in (S.PREC_FP64, S.PREC_FP32):
                ctx.set_option(S.OPT_SMOOTHER_PRECISION, prec)
                f = dev(Z["in_f"])
                v = torch.zeros_like(f)
                st = ctx.fcg(S.OP_POISSON, S.PRESSURE, S.PC_MG, f, v, eps, 5000)
                runs.append((st.iterations, st.coarse_iters, host(v)))
            ctx.set_option(S.OPT_SMOOTHER_PRECISION, S.PREC_FP64)
            res[opt] = runs
        ctx.set_option(S.OPT_COARSE_ASSEMBLED, 1)
    print(f"[gpu] coarse two-launch {name}: {[r[:2] for r in res[1]]}")
    for a, b in zip(res[1], res[2]):
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("name", ["cube2_n8", "periodic3_n6", "channel_n6",
                                  "cube2_n12"])
def test_assembled_coarse_level(name):
    """the coarsest level solved over unique nodes with its operator
    assembled once (CSR) against the E-vector CG with element operator and
    DSS: coarse Jacobi-PCG for both operators and variable kinds (iteration
    counts within one, solutions within the tolerance, the same value on
    every copy of a node), and the full FCG + p-MG solves against the
    reference's iteration counts and solutions, in FP64 and FP32"""
    Z = golden(name)
    mesh = golden_mesh(Z)
    sc = Z["ref_scalars"]
    eps = sc[0]
    with make_ctx(mesh) as ctx:
        lc = ctx.nlevels - 1
        rng = np.random.default_rng(8)
        fc = rng.standard_normal(ctx.ntot(lc))
        res = {}
        for opt in (0, 1):
            ctx.set_option(S.OPT_COARSE_ASSEMBLED, opt)
            runs = []
            for op, var in ((S.OP_POISSON, S.PRESSURE),
                            (S.OP_HELMHOLTZ, S.X_VELOCITY)):
                f = dev(fc)
                ctx.dss_average(lc, f)
                u = torch.zeros_like(f)
                st = ctx.cg(lc, op, var, f, u, 1e-9 * float(f.norm()), 500)
                runs.append((st.iterations, host(u)))
            for var, diag in ((S.PRESSURE, -1), (S.X_VELOCITY, 0)):
                ctx.set_diag_choice(diag)
                f = dev(Z["in_f"])
                v = torch.zeros_like(f)
                st = ctx.fcg(S.OP_POISSON, var, S.PC_MG, f, v, eps, 5000)
                runs.append((st.iterations, host(v), st.coarse_iters))
            ctx.set_diag_choice(-1)
            res[opt] = runs
        ctx.set_option(S.OPT_SMOOTHER_PRECISION, S.PREC_FP32)
        ctx.set_diag_choice(0)
        f = dev(Z["in_f"])
        v = torch.zeros_like(f)
        st32 = ctx.fcg(S.OP_POISSON, S.X_VELOCITY, S.PC_MG, f, v, eps, 5000)
        v32 = host(v)
    a, b = res[1], res[0]
    print(f"[gpu] assembled coarse {name}: coarse cg {a[0][0]}/{b[0][0]}, "
          f"{a[1][0]}/{b[1][0]}; fcg {a[2][0]}/{b[2][0]} (coarse "
          f"{a[2][2]}/{b[2][2]}), {a[3][0]}/{b[3][0]}; fp32 {st32.iterations}")
    for k in range(4):
        assert abs(a[k][0] - b[k][0]) <= 1, k
    for k in range(2):
        assert rel(a[k][1], b[k][1]) <= 1e-7, k
    assert abs(a[2][0] - int(sc[3])) <= 1 and abs(a[3][0] - int(sc[5])) <= 1
    assert rel(a[2][1], Z["ref_fcg_pressure"]) <= 1e-8
    assert rel(a[3][1], Z["ref_fcg_dirichlet"]) <= 1e-8
    assert st32.converged and st32.iterations <= int(sc[5]) + 3
    assert rel(v32, Z["ref_fcg_dirichlet"]) <= 1e-7


@pytest.mark.parametrize("n", range(3, 14))
def test_transfers_every_level_pair(n):
    """Restriction (with the DSS average, precon.c:283-291) and prolongation
    (interpolation.c:136-176) between order n and its p-multigrid level
    (n-1)/2 + 1 against the oracle, for n = 3..13: the wavefront-per-element
    kernels (n <= 12) and the block-per-element ones (n = 13).  18 elements:
    the last block of four elements has two idle.  Bar 1e-13 relative."""
    mesh = box_mesh(n, (3, 2, 3), deform=1, amp=0.1)
    rng = np.random.default_rng(100 + n)
    ctx = make_ctx(mesh)
    o = P.Oracle(mesh)
    try:
        assert ctx.level_n(1) == (n - 1) // 2 + 1
        fine = rng.standard_normal(ctx.ntot(0))
        c = torch.zeros(ctx.ntot(1), dtype=torch.float64, device="cuda")
        ctx.restrict(1, dev(fine), c)
        ref = o.dss_average(1, o.restrict(1, fine))
        assert rel(host(c), ref) <= 1e-13
        coarse = rng.standard_normal(ctx.ntot(1))
        f = torch.full((ctx.ntot(0),), 7.0, dtype=torch.float64, device="cuda")
        ctx.prolong(0, dev(coarse), f)
        assert rel(host(f), o.prolong(0, coarse)) <= 1e-13
    finally:
        o.close()
        ctx.close()
