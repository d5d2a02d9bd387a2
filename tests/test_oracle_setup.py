"""Hand-built cases for the oracle's set-up restatement (oracle/setup.c):
surface detection and its 2nd-order band, kbeg_a, absorber homogenisation,
the layered velocity models, the Cerjan taper and the PML / Zener
coefficients, each against what its definition fixes.
"""
import ctypes as C

import numpy as np

from oracle import oracle as O

OFF = O.OFF


def _halo3(ny, nx, nz):
    return np.zeros((ny + 6, nx + 6, nz + 6), np.float32)


def test_surface_detection_with_topography(orc):
    """kfs = last air cell (lam = 0) above material, kob = last cell with mu =
    0 above solid (m_medium.f90:355-368); the band is the 6x6-window min - 2
    / max + 2 clipped to [1, nz] (:375-382)."""
    L = O.lib(32)
    nx, ny, nz = 14, 12, 20
    jj, ii = np.meshgrid(np.arange(ny + 6) - OFF, np.arange(nx + 6) - OFF, indexing="ij")
    kfs_true = np.where(ii > 7, 5, 3) + (jj > 8)            # a step in x and one in y
    kob_true = kfs_true + np.where(ii % 3 == 0, 4, 0)       # ocean columns of 4 cells here and there
    lam, mu = _halo3(ny, nx, nz), _halo3(ny, nx, nz)
    k = np.arange(nz + 6) - OFF
    solid = k[None, None, :] > kob_true[..., None]
    fluid = (k[None, None, :] > kfs_true[..., None]) & ~solid
    lam[solid | fluid] = 1.5
    mu[solid] = 1.0
    maps = [np.zeros((ny + 6, nx + 6), np.int32) for _ in range(6)]
    L.orc_surface_detection(nx, ny, nz, lam, mu, *maps)
    kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot = maps
    inner = (slice(OFF + 0, OFF + ny + 3), slice(OFF + 0, OFF + nx + 3))   # 0..n+2 are detected
    assert np.array_equal(kfs[inner], kfs_true[inner])
    assert np.array_equal(kob[inner], kob_true[inner])
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            w = (slice(j - 2 + OFF, j + 4 + OFF), slice(i - 2 + OFF, i + 4 + OFF))
            assert kfs_top[j + OFF, i + OFF] == max(kfs[w].min() - 2, 1)
            assert kfs_bot[j + OFF, i + OFF] == min(kfs[w].max() + 2, nz)
            assert kob_top[j + OFF, i + OFF] == max(kob[w].min() - 2, 1)
            assert kob_bot[j + OFF, i + OFF] == min(kob[w].max() + 2, nz)
    # a point next to the x step sees both depths in its window
    assert kfs_top[2 + OFF, 7 + OFF] == 1 and kfs_bot[2 + OFF, 7 + OFF] == 7


def test_kbeg_a_partition(orc):
    """PML columns start at k = 1 at the sides, at nz - na + 1 under the interior (m_global.f90:334-343)"""
    L = O.lib(32)
    nx, ny, nz, na = 30, 26, 22, 6
    kb = np.zeros((ny + 6, nx + 6), np.int32)
    L.orc_kbeg_a(nx, ny, nz, na, kb)
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            side = i <= na or i > nx - na or j <= na or j > ny - na
            assert kb[j + OFF, i + OFF] == (1 if side else nz - na + 1), (i, j)


def test_absorber_homogenisation(orc):
    """absorber cells copy the nearest interior value: x, then y, then the bottom (m_medium.f90:124-190)"""
    L = O.lib(32)
    nx, ny, nz, na = 16, 14, 12, 3
    rng = np.random.default_rng(5)
    arrs = [rng.standard_normal((ny + 6, nx + 6, nz + 6)).astype(np.float32) for _ in range(5)]
    orig = [a.copy() for a in arrs]
    L.orc_homogenize_absorber(nx, ny, nz, na, *arrs)
    i = np.arange(nx + 6) - OFF
    j = np.arange(ny + 6) - OFF
    k = np.arange(nz + 6) - OFF
    ic = np.clip(i, na + 1, nx - na)
    jc = np.clip(j, na + 1, ny - na)
    kc = np.minimum(k, nz - na)
    for a, o in zip(arrs, orig):
        want = o[np.ix_(jc + OFF, ic + OFF, kc + OFF)]
        assert np.array_equal(a, want)


def test_layered_models(orc):
    """lhm: constant layers, air above the first layer top; lgm: linear in z
    between layer tops; vcut lifts slow layers to the next one (m_vmodel_lhm.f90:91-150, m_vmodel_lgm.f90)"""
    L = O.lib(32)
    nx, ny, nz, dz, zbeg = 4, 4, 30, 0.5, -2.0
    layers = np.array([[0.0, 2.0, 3.0, 1.2, 100, 50], [4.0, 2.5, 5.0, 2.8, 200, 100],
                       [9.0, 2.9, 6.5, 3.7, 300, 150]], np.float32)
    out = {}
    for name, fn in (("lhm", L.orc_vmodel_lhm), ("lgm", L.orc_vmodel_lgm)):
        a = [_halo3(ny, nx, nz) for _ in range(5)]
        assert fn(nx, ny, nz, zbeg, dz, 1.5, 0, 3, np.ascontiguousarray(layers.reshape(-1)), *a) == 0
        out[name] = a
    k = np.arange(nz + 6) - OFF
    z = zbeg + (k - 0.5) * dz
    rho, lam, mu, qp, qs = (x[2 + OFF, 2 + OFF] for x in out["lhm"])
    air = z < 0
    assert np.all(rho[air] == np.float32(0.001)) and np.all(lam[air] == 0) and np.all(mu[air] == 0)
    assert np.all(qp[air] == 10)
    # layer 1 (vs 1.2 < vcut 1.5) takes layer 2's values down to z = 4
    l1 = (z >= 0) & (z < 4)
    np.testing.assert_allclose(mu[l1], 2.5 * 2.8**2, rtol=1e-6)
    deep = z >= 9
    np.testing.assert_allclose(lam[deep], 2.9 * (6.5**2 - 2 * 3.7**2), rtol=1e-6)
    # lgm: halfway between the tops of layers 2 and 3
    rho_g, lam_g, mu_g, _, _ = (x[2 + OFF, 2 + OFF] for x in out["lgm"])
    kk = int(np.argmin(np.abs(z - 6.5)))
    f = (z[kk] - 4.0) / 5.0
    r = 2.5 + 0.4 * f
    vs = 2.8 + 0.9 * f
    np.testing.assert_allclose(mu_g[kk], r * vs * vs, rtol=1e-5)


def test_ocean_column(orc):
    """Between sea level and depth(1) the column is seawater: rho 1, mu 0,
    lam = vp^2 with vp = 1.5 km/s, Q = 1e6 (m_vmodel_lhm.f90:111-133,
    m_vmodel_lgm.f90:110-128); Munk's profile (m_seawater.f90:34-47) gives
    1.5 km/s at its sound axis z = 1.3 km and more above and below it."""
    L = O.lib(32)
    nx, ny, nz, dz, zbeg = 4, 4, 30, 0.5, -2.0
    layers = np.array([[3.0, 2.5, 5.0, 2.8, 200, 100], [9.0, 2.9, 6.5, 3.7, 300, 150]], np.float32)
    k = np.arange(nz + 6) - OFF
    z = (zbeg + (k - 0.5) * dz).astype(np.float32)
    for fn in (L.orc_vmodel_lhm, L.orc_vmodel_lgm):
        for munk in (0, 1):
            a = [_halo3(ny, nx, nz) for _ in range(5)]
            assert fn(nx, ny, nz, zbeg, dz, 0.0, munk, 2, np.ascontiguousarray(layers.reshape(-1)), *a) == 0
            rho, lam, mu, qp, qs = (x[2 + OFF, 2 + OFF] for x in a)
            sea = (z >= 0) & (z < 3.0)
            assert sea.sum() == 6
            assert np.all(rho[sea] == 1.0) and np.all(mu[sea] == 0) and np.all(qp[sea] == 1e6)
            assert np.all(qs[sea] == 1e6)
            vp = np.sqrt(lam[sea].astype(np.float64))
            if munk:
                zb = 2 * (z[sea].astype(np.float64) * 1000 - 1300) / 1300
                np.testing.assert_allclose(vp, 1.5 * (1 + 0.00737 * (zb - 1 + np.exp(-zb))), rtol=1e-6)
                assert vp.min() > 1.5
            else:
                assert np.all(lam[sea] == np.float32(2.25))
            air = z < 0
            assert np.all(lam[air] == 0) and np.all(rho[air] == np.float32(0.001))
            assert np.all(mu[z >= 3.0] > 0)
    assert L.orc_seawater_vel(1.3, 1) == np.float32(1.5)
    assert abs(L.orc_seawater_vel(0.0, 1) - 1.5 * (1 + 0.00737 * (-2 - 1 + np.exp(2)))) < 1e-6


def test_ocean_bottom_detection(orc):
    """With a water layer kfs is the last air cell and kob the last water
    cell (m_medium.f90:339-394); the two bands are disjoint and 'obb'
    stations sit on kob + 1 (m_wav.f90:209)."""
    cfg = O.Config(nx=24, ny=22, nz=30, na=6, nm=3, zbeg=-1.0, dz=0.5, dx=0.5, dy=0.5,
                   layers=[[2.0, 2.5, 5.0, 2.8, 200, 100], [6.0, 2.9, 6.5, 3.7, 300, 150]],
                   stations=[(0.4, 0.3, 0.0, "obb"), (0.4, 0.3, 0.0, "fsb")])
    m = O.build_model(cfg)
    cells = (slice(OFF + 1, OFF + cfg.ny + 1), slice(OFF + 1, OFF + cfg.nx + 1))
    # band windows (i-2..i+3, j-2..j+3) inside the detected part 0..n+2
    inner = (slice(OFF + 2, OFF + cfg.ny), slice(OFF + 2, OFF + cfg.nx))
    # z(k) = -1 + (k - 0.5) dz: air for k <= 2, water for k = 3..6, solid below
    assert np.all(m.kfs[cells] == 2) and np.all(m.kob[cells] == 6)
    assert np.all(m.kfs_top[cells] == 1) and np.all(m.kfs_bot[cells] == 4)
    assert np.all(m.kob_top[inner] == 4) and np.all(m.kob_bot[cells] == 8)
    # at the grid edges the window reaches the undetected value kbeg - 1 = 0
    assert m.kob_top[OFF + 1, OFF + 5] == 1 and m.kob_top[OFF + 5, OFF + cfg.nx] == 1
    assert list(m.kst) == [7, 3]
    assert m.vmin > 2.7 and m.vmax > 6.0


def test_cerjan_taper_closed_form(orc):
    """g = exp(-0.09 (1 - d/L)^2) with d the distance of the (staggered) point
    from the outer edge, L = na h; no taper at the top in z (m_absorb_c.f90:50-96)"""
    L = O.lib(32)
    n, na, h = 40, 8, 0.5
    for top in (1, 0):
        gc = np.zeros(n + 6, np.float32)
        gb = np.zeros(n + 6, np.float32)
        L.orc_cerjan_profile_axis(n, na, np.float32(h), top, gc, gb)
        Lw = na * h
        for i in range(1, n + 1):
            if i <= na:
                dc, db = ((i - 0.5) * h, i * h) if top else (None, None)
            elif i >= n - na + 1:
                dc, db = ((n - i + 1) * h, (n - i + 0.5) * h)
            else:
                dc = db = None
            wc = 1.0 if dc is None else np.exp(-0.09 * (1 - dc / Lw) ** 2)
            wb = 1.0 if db is None else np.exp(-0.09 * (1 - db / Lw) ** 2)
            np.testing.assert_allclose(gc[i + OFF], wc, rtol=2e-6, err_msg=f"c {i} top={top}")
            np.testing.assert_allclose(gb[i + OFF], wb, rtol=2e-6, err_msg=f"b {i} top={top}")


def test_pml_profile_dc_limit(orc):
    """The ADE pair  dtilde = g1 d + g2 a,  a <- g3 a + g4 d dt  is the
    Crank-Nicolson discretisation of the CFS stretching 1/s(w),
    s = b + d/(a + i w).  For a constant derivative it settles at
    dtilde = d / (b + d_pml / a_pml), the w = 0 limit, and in the
    interior (no damping) it is the plain derivative."""
    L = O.lib(32)
    n, na, h, fcut, dt = 60, 20, 0.5, 1.0, 0.02
    gc = np.zeros(4 * n, np.float32)
    ge = np.zeros(4 * n, np.float32)
    L.orc_pml_profile_axis(n, na, np.float32(-15.0), np.float32(h), np.float32(fcut), np.float32(dt), gc, ge)
    H = na * h
    r0 = 10.0 ** (-((np.log10(na) - 1) / np.log10(2.0)) - 3.0)
    assert abs(r0 - 1e-4) < 1e-12       # R0 of a 20-cell layer
    d0 = -(1 / (2 * H)) * 2 * 6.0 * np.log(r0)
    for g, off in ((gc, 0.0), (ge, 0.5)):
        g = g.reshape(n, 4).astype(np.float64)
        for i in range(1, n + 1):
            x = -15.0 + (i - 0.5 + off) * h
            xx = max((-15.0 + H) - x, x - (15.0 - H), 0.0)
            r = abs(xx / H)
            d, a, b = d0 * r, np.pi * fcut * (1 - r), 1 + 6.0 * r * r
            g1, g2, g3, g4 = g[i - 1]
            steady = g1 + g2 * g4 * dt / (1 - g3)
            want = 1.0 / (b + d / a) if a > 0 else 0.0
            np.testing.assert_allclose(steady, want, rtol=2e-4, atol=1e-6, err_msg=f"{i} {off}")
            if r == 0:
                # undamped: unit weight on the derivative, the auxiliary never driven
                assert g1 == 1.0 and g4 == 0.0


def test_zener_recursion_coefficients(orc):
    """c1 = the trapezoidal step of exp(-dt/ts); under a constant drive F the
    recursion R <- c1 R - c2 F dt settles at -F/NM for every mechanism"""
    L = O.lib(32)
    nm, dt = 3, 0.02
    ts = np.zeros(nm, np.float32)
    L.orc_visco_set_relaxtime(nm, 0.05, 5.0, ts)
    c1, c2, d1 = (np.zeros(nm, np.float32) for _ in range(3))
    d2 = C.c_float()
    L.orc_kernel_coef(nm, ts, np.float32(dt), c1, c2, d1, C.byref(d2))
    x = dt / ts.astype(np.float64)
    np.testing.assert_allclose(c1, (1 - x / 2) / (1 + x / 2), rtol=1e-6)
    assert np.all(np.abs(c1 - np.exp(-x)) < x**3)
    R = np.zeros(nm)
    for _ in range(20000):
        R = c1 * R - c2 * 1.0 * dt
    np.testing.assert_allclose(R, -1.0 / nm, rtol=1e-4)


def test_cerjan_applied_at_staggered_points(orc):
    """absorb_c__update_stress/_vel (m_absorb_c.f90:113-198) multiply each
    field by the taper at that field's own staggered point: a unit field
    becomes G(x) G(y) G(z) with x at i - 1/2 (cell) or i (edge) cells from
    the outer face, per axis as the field is staggered, and no taper at the
    top.  FP64; a control swaps the profile of one axis."""
    cfg = O.Config(nx=22, ny=20, nz=18, na=5, nm=0, abc="cerjan", dx=0.5, dy=0.4, dz=0.3, zbeg=-1.0, sources=[],
                   layers=[[0.0, 2.7, 5.5, 3.2, 1e6, 1e6]])
    run = O.OracleRun(O.build_model(cfg), 64)
    for name in O.FIELDS_V + O.FIELDS_S:
        run.arr[name][...] = 1.0
    run.L.orc_absorb_c_update_stress(run.ptr)
    run.L.orc_absorb_c_update_vel(run.ptr)

    def taper(n, h, top, edge):
        """G at cells 1..n: distance from the outer face of the point at
        i - 1/2 (cell) or i (edge) cells"""
        Lw, na = cfg.na * h, cfg.na
        out = np.ones(n)
        for i in range(1, n + 1):
            if i <= na and top:
                d = (i if edge else i - 0.5) * h
            elif i >= n - na + 1:
                d = (n - i + (0.5 if edge else 1.0)) * h
            else:
                continue
            out[i - 1] = np.exp(-0.09 * (1 - d / Lw) ** 2)
        return out

    gx = {e: taper(cfg.nx, cfg.dx, 1, e) for e in (False, True)}
    gy = {e: taper(cfg.ny, cfg.dy, 1, e) for e in (False, True)}
    gz = {e: taper(cfg.nz, cfg.dz, 0, e) for e in (False, True)}
    stag = {"Vx": (1, 0, 0), "Vy": (0, 1, 0), "Vz": (0, 0, 1), "Sxx": (0, 0, 0), "Syy": (0, 0, 0), "Szz": (0, 0, 0),
            "Syz": (0, 1, 1), "Sxz": (1, 0, 1), "Sxy": (1, 1, 0)}
    s = (slice(OFF + 1, OFF + 1 + cfg.ny), slice(OFF + 1, OFF + 1 + cfg.nx), slice(OFF + 1, OFF + 1 + cfg.nz))
    for name, (ex, ey, ez) in stag.items():
        want = gy[bool(ey)][:, None, None] * gx[bool(ex)][None, :, None] * gz[bool(ez)][None, None, :]
        got = run.arr[name][s]
        np.testing.assert_allclose(got, want, rtol=2e-6, err_msg=name)
        if name == "Syz":
            # control: the cell-centred y profile in place of the edge one is seen
            bad = gy[False][:, None, None] * gx[False][None, :, None] * gz[True][None, None, :]
            assert not np.allclose(got, bad, rtol=1e-4)
