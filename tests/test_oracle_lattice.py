"""Lattice-level pins of the oracle: initial state, node flags, one-step
streaming + boundary rules (walls, inlet, outlet, solid obstacles) against an
independent numpy construction, conservation, decomposition invariance, and
the duct and pipe Poiseuille profiles."""
import numpy as np
import pytest

import oracle

E = None


def e_vec():
    d, opp, w = oracle.tables()
    return np.stack([d[0], d[2], d[1]], axis=1).astype(int), opp, w


def params(**kw):
    base = dict(nx=9, ny=8, nz=7, tau=0.6, smagorinsky=0.0, enable_les=0, enable_mrt=1, ref_vel=(0.05, 0.0, 0.0),
                init_lid_velocity=0)
    base.update(kw)
    return oracle.default_params(**base)


def lattice(p, steps=0, slices=None):
    L = oracle.Lattice(p, slices)
    L.initialize()
    L.run(steps)
    return L


def pipe_mask(p):
    y = np.arange(p.ny)[:, None, None]
    z = np.arange(p.nz)[None, :, None]
    m = (y - p.pipe_cy) ** 2 + (z - p.pipe_cz) ** 2 > p.pipe_radius ** 2
    return np.broadcast_to(m, (p.ny, p.nz, p.nx)).copy()


def brute_types(p, mask=None):
    if mask is None and p.casetype == oracle.PIPE:
        mask = pipe_mask(p)
    t = np.zeros((p.ny, p.nz, p.nx), dtype=int)
    for y in range(p.ny):
        for z in range(p.nz):
            for x in range(p.nx):
                face = x in (0, p.nx - 1) or z in (0, p.nz - 1)
                if p.casetype == oracle.CAVITY:
                    v = oracle.NODE_FLUID
                    if y == 0:
                        v = oracle.NODE_FIXEDWALL
                    if y == p.ny - 1:
                        v = oracle.NODE_MOVINGWALL
                    if face:
                        v = oracle.NODE_FIXEDWALL
                else:
                    v = oracle.NODE_FLUID
                    if x == 0:
                        v = oracle.NODE_INLET
                    if x == p.nx - 1:
                        v = oracle.NODE_OUTLET
                    if y in (0, p.ny - 1) or z in (0, p.nz - 1):
                        v = oracle.NODE_FIXEDWALL
                if mask is not None and mask[y, z, x]:
                    v = oracle.NODE_STRUCT
                t[y, z, x] = v
    return t


@pytest.mark.parametrize("case", [oracle.CAVITY, oracle.CHANNEL])
def test_node_types_and_direction_masks(case):
    p = params(casetype=case)
    t, d = lattice(p).node_type()
    assert np.array_equal(t, brute_types(p))
    e, _, _ = e_vec()
    for y in range(p.ny):
        for z in range(p.nz):
            for x in range(p.nx):
                bits = 1
                for q in range(1, 19):
                    xx, yy, zz = x + e[q, 0], y + e[q, 1], z + e[q, 2]
                    if 0 <= xx < p.nx and 0 <= yy < p.ny and 0 <= zz < p.nz:
                        bits |= 1 << q
                assert d[y, z, x] == bits


@pytest.mark.parametrize("init_lid", [0, 1])
def test_initial_state(init_lid):
    p = params(init_lid_velocity=init_lid, ref_vel=(0.1, 0.0, 0.0))
    L = lattice(p)
    a = L.read_arF()
    _, _, w = e_vec()
    assert not a[0].any() and not a[-1].any() and not a[:, 0].any() and not a[:, :, 0].any()
    inner = a[1:-1, 1:-1, 1:-1]
    t, _ = L.node_type()
    lid = t == oracle.NODE_MOVINGWALL
    assert np.allclose(inner[~lid], w, atol=0)
    want = oracle.cell_feq(1.0, [0.1, 0, 0]) if init_lid else w
    assert np.allclose(inner[lid], want, atol=1e-16)


def numpy_one_step_tau1(p, arf0, mask=None):
    """Expected arF after one SRT step with tau = 1 (f* = feq), built with
    array shifts: pull from x - e_q where inside, otherwise the boundary rule
    applied to feq at the node itself; a solid donor bounces back at the node,
    and solid nodes stay zero."""
    e, opp, w = e_vec()
    if mask is None and p.casetype == oracle.PIPE:
        mask = pipe_mask(p)
    f = arf0[1:-1, 1:-1, 1:-1]  # [y][z][x][q]
    solid = np.zeros(f.shape[:3], dtype=bool) if mask is None else np.asarray(mask, dtype=bool)
    rho = np.where(solid, 1.0, f.sum(-1))
    j = np.einsum("yzxq,qa->yzxa", f, e)
    u = j / rho[..., None]
    eu = np.einsum("yzxa,qa->yzxq", u, e)
    usq = (u * u).sum(-1)[..., None]
    feq = w * rho[..., None] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * usq)
    ny, nz, nx = rho.shape
    out = np.zeros_like(f)
    t = brute_types(p, mask)
    uin = np.full((ny, nz), p.ref_vel[0]) if p.casetype == oracle.PIPE else oracle.inlet_profile(p)
    for q in range(19):
        ex, ey, ez = e[q]
        for y in range(ny):
            for z in range(nz):
                for x in range(nx):
                    if solid[y, z, x]:
                        continue
                    sx, sy, sz = x - ex, y - ey, z - ez
                    o = opp[q]
                    if 0 <= sx < nx and 0 <= sy < ny and 0 <= sz < nz:
                        out[y, z, x, q] = feq[y, z, x, o] if solid[sy, sz, sx] else feq[sy, sz, sx, q]
                        continue
                    base = feq[y, z, x, o]
                    ty = t[y, z, x]
                    if ty == oracle.NODE_FIXEDWALL:
                        out[y, z, x, q] = base
                    elif ty == oracle.NODE_MOVINGWALL:
                        out[y, z, x, q] = base - 6 * w[o] * p.ref_den * (e[o] @ np.array(list(p.ref_vel)))
                    elif ty == oracle.NODE_INLET:
                        rin = 0.0 if solid[y, z, x + 1] else rho[y, z, x + 1]  # solid nodes hold rho = 0
                        out[y, z, x, q] = base - 6 * w[o] * rin * (e[o][0] * uin[y, z])
                    elif ty == oracle.NODE_OUTLET:
                        ub = u[y, z, x] + 0.5 * (u[y, z, x] - u[y, z, x - 1])
                        eub = e[o] @ ub
                        out[y, z, x, q] = -base + 2 * w[o] * p.ref_den * (1 + 4.5 * eub * eub - 1.5 * ub @ ub)
                    else:
                        raise AssertionError("fluid node with an outside donor")
    return out


@pytest.mark.parametrize("case", [oracle.CAVITY, oracle.CHANNEL])
def test_one_step_streaming_and_boundaries(case):
    p = params(casetype=case, enable_mrt=0, tau=1.0, ref_vel=(0.07, 0.01, -0.02), inlet_peak=0.06, perturb=0.2,
               seed=7)
    rng = np.random.default_rng(8)
    L = lattice(p)
    a = L.read_arF()
    a[1:-1, 1:-1, 1:-1] *= 1.0 + 0.05 * rng.standard_normal(a[1:-1, 1:-1, 1:-1].shape)
    L.write_arF(a)
    L.run(1)
    got = L.read_arF()[1:-1, 1:-1, 1:-1]
    want = numpy_one_step_tau1(p, a)
    assert np.abs(got - want).max() < 1e-15


@pytest.mark.parametrize("case", [oracle.CAVITY, oracle.CHANNEL])
def test_one_step_with_solid_nodes(case):
    p = params(casetype=case, enable_mrt=0, tau=1.0, ref_vel=(0.07, 0.01, -0.02), inlet_peak=0.06, perturb=0.2,
               seed=3)
    rng = np.random.default_rng(11)
    mask = rng.random((p.ny, p.nz, p.nx)) < 0.25
    s = oracle.decomposition(p.nx, p.ny, p.nz, 2, 3, 2)  # donors across tile edges
    L = oracle.Lattice(p, s)
    L.set_mask(mask)
    L.initialize()
    t, _ = L.node_type()
    assert np.array_equal(t, brute_types(p, mask))
    a = L.read_arF()
    assert not a[1:-1, 1:-1, 1:-1][mask].any()
    a[1:-1, 1:-1, 1:-1] *= 1.0 + 0.05 * rng.standard_normal(a[1:-1, 1:-1, 1:-1].shape)
    L.write_arF(a)
    L.run(1)
    got = L.read_arF()[1:-1, 1:-1, 1:-1]
    want = numpy_one_step_tau1(p, a, mask)
    assert np.abs(got - want).max() < 1e-15


def test_pipe_node_types_and_one_step():
    p = params(nx=6, ny=11, nz=10, casetype=oracle.PIPE, pipe_radius=4.2, pipe_cy=5.0, pipe_cz=4.5, enable_mrt=0,
               tau=1.0, ref_vel=(0.04, 0.0, 0.0))
    L = lattice(p)
    t, _ = L.node_type()
    assert np.array_equal(t, brute_types(p))
    m = pipe_mask(p)
    assert 0 < m.sum() < m.size
    rng = np.random.default_rng(5)
    a = L.read_arF()
    a[1:-1, 1:-1, 1:-1] *= 1.0 + 0.05 * rng.standard_normal(a[1:-1, 1:-1, 1:-1].shape)
    L.write_arF(a)
    L.run(1)
    assert np.abs(L.read_arF()[1:-1, 1:-1, 1:-1] - numpy_one_step_tau1(p, a)).max() < 1e-15


def test_obstacle_conserves_mass_and_mirror_symmetry():
    # closed cavity with a block symmetric about the z mid-plane: walls and
    # obstacle only bounce back, so mass is conserved, and the z-mirror
    # symmetry of the set-up keeps the z momentum at zero
    p = params(nx=12, ny=12, nz=12, enable_les=1, smagorinsky=0.17, ref_vel=(0.08, 0.0, 0.0))
    mask = np.zeros((12, 12, 12), dtype=bool)
    mask[4:8, 3:9, 5:8] = True
    mask[7, 1:11, 2] = True
    L = oracle.Lattice(p)
    L.set_mask(mask)
    L.initialize()
    L.run(80)
    m, mom, _ = L.diagnostics()
    assert abs(m - (12 ** 3 - mask.sum())) < 1e-10
    assert abs(mom[2]) < 1e-12
    r, ux, uy, uz = L.read_macro()
    assert not r[mask].any() and not ux[mask].any()
    assert np.abs(ux[:, 3:9, :]).max() > 1e-3  # the lid drives flow past the block


@pytest.mark.parametrize("mrt,les", [(0, 0), (1, 1)])
def test_cavity_conserves_mass(mrt, les):
    p = params(nx=12, ny=12, nz=12, enable_mrt=mrt, enable_les=les, smagorinsky=0.17)
    L = lattice(p, 60)
    m, mom, d = L.diagnostics()
    assert abs(m - 12 ** 3) < 1e-10
    assert abs(mom[2]) < 1e-12  # the z-mirror symmetry of the cavity


def test_results_do_not_depend_on_decomposition():
    p = params(nx=11, ny=9, nz=7, enable_les=1, smagorinsky=0.17)
    ref = lattice(p, 25).read_arF()
    for dec in [(3, 3, 3), (2, 1, 3), (1, 4, 1), (5, 5, 5)]:
        s = oracle.decomposition(11, 9, 7, *dec)
        assert np.array_equal(lattice(p, 25, s).read_arF(), ref), dec


def duct_profile(a, b, y, z, terms=101):
    """Fully developed laminar flow in an a x b rectangular duct (series
    solution, e.g. White, Viscous Fluid Flow, eq. 3-48), up to a constant."""
    u = np.zeros(np.broadcast(y, z).shape)
    for n in range(1, terms, 2):
        k = n * np.pi / a
        u += ((-1) ** ((n - 1) // 2) / n ** 3) * (1 - np.cosh(k * z) / np.cosh(k * b / 2)) * np.cos(k * y)
    return u


def test_channel_develops_the_duct_poiseuille_profile():
    ny = nz = 12
    p = params(nx=48, ny=ny, nz=nz, casetype=oracle.CHANNEL, enable_mrt=1, enable_les=0, tau=0.8,
               inlet_peak=0.03, perturb=0.0)
    L = lattice(p, 2500)
    _, ux, _, _ = L.read_macro()
    sec = ux[:, :, 32]  # [y][z] at x = 32
    yc = np.arange(ny) + 0.5 - ny / 2.0  # half-way walls at -ny/2 and +ny/2
    zc = np.arange(nz) + 0.5 - nz / 2.0
    ana = duct_profile(ny, nz, yc[:, None], zc[None, :])
    inner = (slice(1, -1), slice(1, -1))  # wall nodes carry the wall rule
    ratio = sec[inner] / ana[inner]
    assert np.std(ratio) / np.mean(ratio) < 0.01, np.std(ratio) / np.mean(ratio)  # measured 0.6%


def test_pipe_develops_the_poiseuille_profile():
    # Hagen-Poiseuille flow in a circular pipe: u(r) = u_max (1 - r^2 / R^2),
    # u_max = 2 u_mean; the staircase wall of the solid mask leaves ~1% rms
    n, c, R = 22, 10.5, 9.5
    p = params(nx=40, ny=n, nz=n, casetype=oracle.PIPE, pipe_radius=R, pipe_cy=c, pipe_cz=c, enable_mrt=1,
               enable_les=0, tau=0.8, ref_vel=(0.03, 0.0, 0.0))
    L = lattice(p, 1500)
    _, ux, uy, uz = L.read_macro()
    sec = ux[:, :, 28]
    yy, zz = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r2 = (yy - c) ** 2 + (zz - c) ** 2
    fluid, inner = r2 <= R * R, r2 <= (R - 1.5) ** 2
    a = np.stack([np.ones(inner.sum()), r2[inner]], 1)
    coef, *_ = np.linalg.lstsq(a, sec[inner], rcond=None)
    rms = np.sqrt(((sec[inner] - a @ coef) ** 2).mean()) / sec[inner].mean()
    assert rms < 0.02, rms  # measured 1.2%
    r_eff = np.sqrt(coef[0] / -coef[1])
    assert R - 1.0 < r_eff < R + 0.5, r_eff  # measured 9.27
    assert abs(sec.max() / sec[fluid].mean() - 2.0) < 0.05  # measured 2.03
    assert np.abs(uy[:, :, 28]).max() < 1e-3 * sec.max()
    assert not sec[~fluid].any()
