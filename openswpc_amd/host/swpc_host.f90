!! swpc_host.f90 -- host-side set-up of the 3D time loop (medium, viscoelastic
!! coefficients, absorber profiles, sources, stations) for libswpc_gpu.
!!
!! Every routine is exported with bind(C) so the Fortran driver, the C tests
!! and Python (ctypes) share one implementation.  Arithmetic follows the
!! reference's single-precision semantics (real(SP), with double only where
!! the reference promotes through PI = atan(1.0_DP)*4, src/shared/m_std.f90:11)
!! so that the kernel inputs it produces are the ones the reference computes.
!!
!! 3D arrays use the reference layout: (k,i,j) with halo bounds -2:n+3
!! (src/swpc_3d/m_global.f90:295-300); (i,j) maps likewise.
module swpc_host

    use iso_c_binding
    implicit none
    private

    integer, parameter :: SP = c_float
    integer, parameter :: DP = c_double
    real(DP), parameter :: PI = 3.141592653589793238462643383279502884_DP
    real(SP), parameter :: EPSL = epsilon(1.0_SP)

    public :: swpch_i2x, swpch_x2i, swpch_seawater_vel, swpch_vmodel_layers, swpch_perturb, swpch_medium_setup
    public :: swpch_vmodel_profile, swpch_visco_setup
    public :: swpch_kernel_coef, swpch_pml_profile, swpch_cerjan_profile, swpch_sdr2moment
    public :: swpch_momentrate, swpch_stf_table, swpch_station_index, swpch_stable_dt

contains

    !! cell centre of index i (src/shared/m_fdtool.f90:639-649)
    real(c_float) function swpch_i2x(i, beg, h) bind(C, name='swpch_i2x')
        integer(c_int), value :: i
        real(c_float), value :: beg, h
        swpch_i2x = beg + (i - 0.5_SP) * h
    end function swpch_i2x

    !! cell index of coordinate x (src/shared/m_fdtool.f90:600-636)
    integer(c_int) function swpch_x2i(x, beg, h) bind(C, name='swpch_x2i')
        real(c_float), value :: x, beg, h
        swpch_x2i = ceiling((x - beg) / h)
    end function swpch_x2i

    !! stable time step of the 4th-order scheme (src/shared/m_fdtool.f90:81-96)
    real(c_float) function swpch_stable_dt(dx, dy, dz, vmax) bind(C, name='swpch_stable_dt')
        real(c_float), value :: dx, dy, dz, vmax
        real(SP) :: hh
        hh = 1.0_SP / sqrt(1 / dx**2 + 1 / dy**2 + 1 / dz**2)
        swpch_stable_dt = (6.0_SP / 7.0_SP) * hh / vmax
    end function swpch_stable_dt

    !! Sound speed of seawater at depth z [km] (src/shared/m_seawater.f90:34-47):
    !! 1.5 km/s, or Munk's canonical profile (epsilon 0.00737, sound axis at
    !! 1300 m, scale depth 650 m) when munk is nonzero.
    real(c_float) function swpch_seawater_vel(z, munk) bind(C, name='swpch_seawater_vel')
        real(c_float), value :: z
        integer(c_int), value :: munk
        real(DP) :: e, zb
        e = 0.0_DP
        if (munk /= 0) e = 0.00737_DP
        zb = 2 * (real(z, DP) * 1000.0_DP - 1300.0_DP) / 1300.0_DP
        swpch_seawater_vel = real(1.5_DP * (1.0_DP + e * (zb - 1.0_DP + exp(-zb))), SP)
    end function swpch_seawater_vel

    !! Layered medium as one value per plane k = -2..nz+3: kind 0 = homogeneous
    !! layers (vmodel_lhm, src/swpc_3d/m_vmodel_lhm.f90:20-158), kind 1 = linear
    !! gradient between layer tops (vmodel_lgm, m_vmodel_lgm.f90:20-174).
    !! layers(6,nl) = depth, rho, vp, vs, qp, qs.  Above depth(1): air where
    !! z < 0, seawater (rho 1, mu 0, Q 1e6, sound speed swpch_seawater_vel)
    !! where 0 <= z (m_vmodel_lhm.f90:111-133).  prof(:,1..5) = rho, lam, mu,
    !! qp, qs of each plane.
    subroutine swpch_vmodel_profile(kind, nz, zbeg, dz, vcut, munk, nl, layers, prof, rc) &
        bind(C, name='swpch_vmodel_profile')
        integer(c_int), value :: kind, nz, nl, munk
        real(c_float), value :: zbeg, dz, vcut
        real(c_float), intent(in) :: layers(6, nl)
        real(c_float), intent(out) :: prof(-2:nz + 3, 5)
        integer(c_int), intent(out) :: rc
        real(SP) :: dep(nl), r0(nl), vp0(nl), vs0(nl), q0p(nl), q0s(nl)
        real(SP) :: z, rho1, vp1, vs1, qp1, qs1, h, d
        integer :: k, l

        rc = 0
        if (nl < 1) then
            rc = -1
            return
        end if
        dep = layers(1, :); r0 = layers(2, :); vp0 = layers(3, :); vs0 = layers(4, :)
        q0p = layers(5, :); q0s = layers(6, :)
        ! velocity cut-off: slow layers take the values of the layer below
        do l = nl - 1, 1, -1
            if ((vp0(l) < vcut .or. vs0(l) < vcut) .and. (vp0(l) > 0 .and. vs0(l) > 0)) then
                vp0(l) = vp0(l + 1); vs0(l) = vs0(l + 1); r0(l) = r0(l + 1)
                q0p(l) = q0p(l + 1); q0s(l) = q0s(l + 1)
            end if
        end do
        do k = -2, nz + 3
            z = swpch_i2x(k, zbeg, dz)
            if (z < dep(1)) then
                if (z < 0.0_SP) then
                    rho1 = 0.001_SP; vp1 = 0.0_SP; vs1 = 0.0_SP; qp1 = 10.0_SP; qs1 = 10.0_SP
                else
                    rho1 = 1.0_SP; vp1 = 1.0_SP * swpch_seawater_vel(z, munk); vs1 = 0.0_SP
                    qp1 = 1000000.0_SP; qs1 = 1000000.0_SP
                end if
            else if (kind == 0) then
                do l = 1, nl
                    if (z >= dep(l)) then
                        rho1 = r0(l); vp1 = vp0(l); vs1 = vs0(l); qp1 = q0p(l); qs1 = q0s(l)
                    end if
                end do
            else
                rho1 = r0(nl); vp1 = vp0(nl); vs1 = vs0(nl); qp1 = q0p(nl); qs1 = q0s(nl)
                do l = 1, nl - 1
                    if (dep(l) <= z .and. z < dep(l + 1)) then
                        h = dep(l + 1) - dep(l)
                        d = z - dep(l)
                        rho1 = r0(l) + (r0(l + 1) - r0(l)) / h * d
                        vp1 = 1.0_SP * (vp0(l) + (vp0(l + 1) - vp0(l)) / h * d)
                        vs1 = 1.0_SP * (vs0(l) + (vs0(l + 1) - vs0(l)) / h * d)
                        qp1 = q0p(l) + (q0p(l + 1) - q0p(l)) / h * d
                        qs1 = q0s(l) + (q0s(l + 1) - q0s(l)) / h * d
                        exit
                    end if
                end do
            end if
            prof(k, 1) = rho1
            prof(k, 2) = rho1 * (vp1 * vp1 - 2 * vs1 * vs1)
            prof(k, 3) = rho1 * vs1 * vs1
            prof(k, 4) = qp1
            prof(k, 5) = qs1
        end do
    end subroutine swpch_vmodel_profile

    !! The layered medium of swpch_vmodel_profile on the halo grid.
    subroutine swpch_vmodel_layers(kind, nx, ny, nz, zbeg, dz, vcut, munk, nl, layers, rho, lam, mu, qp, qs, rc) &
        bind(C, name='swpch_vmodel_layers')
        integer(c_int), value :: kind, nx, ny, nz, nl, munk
        real(c_float), value :: zbeg, dz, vcut
        real(c_float), intent(in) :: layers(6, nl)
        real(c_float), intent(out) :: rho(-2:nz + 3, -2:nx + 3, -2:ny + 3), lam(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        real(c_float), intent(out) :: mu(-2:nz + 3, -2:nx + 3, -2:ny + 3), qp(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        real(c_float), intent(out) :: qs(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        integer(c_int), intent(out) :: rc
        real(c_float) :: prof(-2:nz + 3, 5)
        integer :: k

        call swpch_vmodel_profile(kind, nz, zbeg, dz, vcut, munk, nl, layers, prof, rc)
        if (rc /= 0) return
        do k = -2, nz + 3
            rho(k, :, :) = prof(k, 1)
            lam(k, :, :) = prof(k, 2)
            mu(k, :, :) = prof(k, 3)
            qp(k, :, :) = prof(k, 4)
            qs(k, :, :) = prof(k, 5)
        end do
    end subroutine swpch_vmodel_layers

    !! Random-medium perturbation xi of every solid cell: vp, vs scaled by
    !! (1 + xi), rho by (1 + 0.8 xi) (the per-cell intent of
    !! src/swpc_3d/m_vmodel_lgm_rmed.f90:195-203), clipped to [vmin, vmax]
    !! like vcheck (src/shared/m_fdtool.f90:15-48).
    subroutine swpch_perturb(nx, ny, nz, xi, vmin, vmax, rhomin, rho, lam, mu) bind(C, name='swpch_perturb')
        integer(c_int), value :: nx, ny, nz
        real(c_float), intent(in) :: xi(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        real(c_float), value :: vmin, vmax, rhomin
        real(c_float), intent(inout) :: rho(-2:nz + 3, -2:nx + 3, -2:ny + 3), lam(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        real(c_float), intent(inout) :: mu(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        integer :: i, j, k
        real(SP) :: vp, vs, r, e, gamma, xi2

        !$omp parallel do private(i, k, vp, vs, r, e, gamma, xi2)
        do j = -2, ny + 3
            do i = -2, nx + 3
                do k = -2, nz + 3
                    if (mu(k, i, j) <= EPSL) cycle
                    r = rho(k, i, j)
                    vp = sqrt((lam(k, i, j) + 2 * mu(k, i, j)) / r)
                    vs = sqrt(mu(k, i, j) / r)
                    e = xi(k, i, j)
                    r = r * (1 + 0.8_SP * e)
                    vp = vp * (1 + e)
                    vs = vs * (1 + e)
                    gamma = vp / vs
                    if (vp > vmax .or. vs > vmax) then
                        xi2 = (1 + e) * vmax / vp - 1
                        vp = vmax
                        vs = vmax / gamma
                        r = r * (1 + 0.8_SP * xi2) / (1 + 0.8_SP * e)
                    end if
                    if (vp < vmin .or. vs < vmin) then
                        vs = vmin
                        vp = vmin * gamma
                    end if
                    if (r < rhomin) r = rhomin
                    rho(k, i, j) = r
                    mu(k, i, j) = r * vs * vs
                    lam(k, i, j) = r * (vp * vp - 2 * vs * vs)
                end do
            end do
        end do
        !$omp end parallel do
    end subroutine swpch_perturb

    !! medium__setup after the velocity model (src/swpc_3d/m_medium.f90:124-229):
    !! absorber homogenisation, tau-method relaxation times ts and zeta, Q ->
    !! taup/taus, relaxed moduli at fq_ref, free-surface / ocean-bottom
    !! detection with the 2nd-order band limits, and velocity extremes.
    !! On entry taup/taus hold Qp/Qs.
    subroutine swpch_medium_setup(nx, ny, nz, na, nm, fq_min, fq_max, fq_ref, rho, lam, mu, taup, taus, ts, zeta, &
                                  kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot, vmin, vmax) &
        bind(C, name='swpch_medium_setup')
        integer(c_int), value :: nx, ny, nz, na, nm
        real(c_float), value :: fq_min, fq_max, fq_ref
        real(c_float), intent(inout) :: rho(-2:nz + 3, -2:nx + 3, -2:ny + 3), lam(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        real(c_float), intent(inout) :: mu(-2:nz + 3, -2:nx + 3, -2:ny + 3), taup(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        real(c_float), intent(inout) :: taus(-2:nz + 3, -2:nx + 3, -2:ny + 3)
        real(c_float), intent(out) :: ts(*), zeta
        integer(c_int), intent(out) :: kfs(-2:nx + 3, -2:ny + 3), kob(-2:nx + 3, -2:ny + 3)
        integer(c_int), intent(out) :: kfs_top(-2:nx + 3, -2:ny + 3), kfs_bot(-2:nx + 3, -2:ny + 3)
        integer(c_int), intent(out) :: kob_top(-2:nx + 3, -2:ny + 3), kob_bot(-2:nx + 3, -2:ny + 3)
        real(c_float), intent(out) :: vmin, vmax
        integer :: i, j, k
        real(SP) :: rb2, ra2, chi_mu, chi_lam, vp, vs, ccv(2)
        complex(SP) :: cc

        call homogenize(rho)
        call homogenize(lam)
        call homogenize(mu)
        call homogenize(taup)
        call homogenize(taus)

        call swpch_visco_setup(nm, fq_min, fq_max, fq_ref, ts, zeta, ccv)
        cc = cmplx(ccv(1), ccv(2), SP)

        !$omp parallel do private(i, k)
        do j = -2, ny + 3
            do i = -2, nx + 3
                do k = -2, nz + 3
                    taup(k, i, j) = nm * zeta / taup(k, i, j)
                    taus(k, i, j) = nm * zeta / taus(k, i, j)
                end do
            end do
        end do
        !$omp end parallel do

        ! relaxed moduli at the reference frequency (m_medium.f90:231-271)
        if (nm > 0) then
            !$omp parallel do private(i, k, rb2, ra2, chi_mu, chi_lam)
            do j = -2, ny + 3
                do i = -2, nx + 3
                    do k = -2, nz + 3
                        rb2 = mu(k, i, j)
                        ra2 = lam(k, i, j) + 2 * mu(k, i, j)
                        chi_mu = 1.0_SP / real(1.0_SP / sqrt(1.0_SP - cc * taus(k, i, j)))
                        chi_lam = 1.0_SP / real(1.0_SP / sqrt(1.0_SP - cc * taup(k, i, j)))
                        mu(k, i, j) = rb2 / chi_mu**2
                        lam(k, i, j) = ra2 / chi_lam**2 - 2 * mu(k, i, j)
                    end do
                end do
            end do
            !$omp end parallel do
        end if

        ! free surface and ocean bottom (m_medium.f90:339-394)
        kfs = 0
        kob = 0
        !$omp parallel do private(i, k)
        do j = 0, ny + 2
            do i = 0, nx + 2
                do k = 1, nz - 1
                    if (abs(mu(k, i, j)) < EPSL .and. abs(mu(k + 1, i, j)) > EPSL) kob(i, j) = k
                    if (abs(lam(k, i, j)) < EPSL .and. abs(lam(k + 1, i, j)) > EPSL) kfs(i, j) = k
                end do
            end do
        end do
        !$omp end parallel do
        kfs_top = 0; kfs_bot = 0; kob_top = 0; kob_bot = 0
        do j = 1, ny
            do i = 1, nx
                kfs_top(i, j) = max(minval(kfs(i - 2:i + 3, j - 2:j + 3)) - 2, 1)
                kfs_bot(i, j) = min(maxval(kfs(i - 2:i + 3, j - 2:j + 3)) + 2, nz)
                kob_top(i, j) = max(minval(kob(i - 2:i + 3, j - 2:j + 3)) - 2, 1)
                kob_bot(i, j) = min(maxval(kob(i - 2:i + 3, j - 2:j + 3)) + 2, nz)
            end do
        end do

        ! velocity extremes below the free surface (m_medium.f90:396-427)
        vmax = -1
        vmin = 1e30
        do j = 1, ny
            do i = 1, nx
                do k = kfs(i, j) + 1, nz
                    vp = sqrt((lam(k, i, j) + 2 * mu(k, i, j)) / rho(k, i, j))
                    vs = sqrt(mu(k, i, j) / rho(k, i, j))
                    vmax = max(vmax, vp)
                    if (vs < EPSL) cycle
                    vmin = min(vmin, vs)
                end do
            end do
        end do

    contains

        subroutine homogenize(a)
            real(c_float), intent(inout) :: a(-2:nz + 3, -2:nx + 3, -2:ny + 3)
            integer :: ii, jj
            do ii = -2, na
                a(:, ii, :) = a(:, na + 1, :)
            end do
            do ii = nx - na + 1, nx + 3
                a(:, ii, :) = a(:, nx - na, :)
            end do
            do jj = -2, na
                a(:, :, jj) = a(:, :, na + 1)
            end do
            do jj = ny - na + 1, ny + 3
                a(:, :, jj) = a(:, :, ny - na)
            end do
            do jj = -2, ny + 3
                do ii = -2, nx + 3
                    a(nz - na + 1:nz + 3, ii, jj) = a(nz - na, ii, jj)
                end do
            end do
        end subroutine homogenize

    end subroutine swpch_medium_setup

    !! The scalars of the tau method (src/shared/m_fdtool.f90:691-812): the nm
    !! relaxation times ts log-spaced over [1/(2 pi fq_max), 1/(2 pi fq_min)],
    !! zeta, and cc = mean_m i w ts_m / (1 - i w ts_m) at w = 2 pi fq_ref
    !! (real and imaginary part), which relaxes the moduli
    !! (src/swpc_3d/m_medium.f90:231-271).
    subroutine swpch_visco_setup(nm, fq_min, fq_max, fq_ref, ts, zeta, ccv) bind(C, name='swpch_visco_setup')
        integer(c_int), value :: nm
        real(c_float), value :: fq_min, fq_max, fq_ref
        real(c_float), intent(out) :: ts(*), zeta, ccv(2)
        real(SP) :: om_a, om_b, omega
        complex(SP) :: cc
        integer :: m

        om_a = real(2 * PI * fq_min, SP)
        om_b = real(2 * PI * fq_max, SP)
        if (nm == 1) then
            ts(1) = 1.0_SP / sqrt(om_a * om_b)
        else if (nm > 1) then
            do m = 1, nm
                omega = real(om_a * real(om_b / om_a, DP)**(real(m - 1, DP) / real(nm - 1, DP)), SP)
                ts(m) = 1.0_SP / omega
            end do
        end if
        zeta = constq_zeta(nm, om_a, om_b, ts)
        cc = (0.0_SP, 0.0_SP)
        if (nm > 0) then
            omega = real(2 * PI * fq_ref, SP)
            do m = 1, nm
                cc = cc + cmplx((cmplx(0.0_DP, 1.0_DP, DP) * omega * ts(m)) / (1.0_DP - cmplx(0.0_DP, 1.0_DP, DP) &
                                * omega * ts(m)), kind=SP)
            end do
            cc = cc / nm
        end if
        ccv(1) = real(cc)
        ccv(2) = aimag(cc)
    end subroutine swpch_visco_setup

    !! Blanch's tau-method zeta for a constant Q over [om_a, om_b]
    !! (src/shared/m_fdtool.f90:756-812)
    real(SP) function constq_zeta(nm, om_a, om_b, ts)
        integer, intent(in) :: nm
        real(SP), intent(in) :: om_a, om_b, ts(*)
        real(SP) :: i0(nm), i1(nm), i2s, w1, w2, s0, s1, t
        integer :: m, l
        if (nm == 0) then
            constq_zeta = 0.0_SP
            return
        end if
        do m = 1, nm
            t = ts(m)
            i0(m) = (log(1.0_SP + om_b**2 * t**2) - log(1.0_SP + om_a**2 * t**2)) / (2 * t)
            i1(m) = ((atan(om_b * t) - om_b * t / (1 + om_b**2 * t**2)) &
                     - (atan(om_a * t) - om_a * t / (1 + om_a**2 * t**2))) / (2 * t)
        end do
        i2s = 0.0_SP
        do m = 1, nm - 1
            do l = m + 1, nm
                w1 = atan(om_b * ts(m)) / ts(m) - atan(om_b * ts(l)) / ts(l)
                w2 = atan(om_a * ts(m)) / ts(m) - atan(om_a * ts(l)) / ts(l)
                i2s = i2s + ts(m) * ts(l) / (ts(l)**2 - ts(m)**2) * (w1 - w2)
            end do
        end do
        s0 = 0.0_SP
        s1 = 0.0_SP
        do m = 1, nm
            s0 = s0 + i0(m)
        end do
        do m = 1, nm
            s1 = s1 + i1(m)
        end do
        constq_zeta = s0 / (s1 + 2 * i2s)
    end function constq_zeta

    !! Zener recursion coefficients (src/swpc_3d/m_kernel.f90:58-67)
    subroutine swpch_kernel_coef(nm, ts, dt, c1, c2, d1, d2) bind(C, name='swpch_kernel_coef')
        integer(c_int), value :: nm
        real(c_float), intent(in) :: ts(*)
        real(c_float), value :: dt
        real(c_float), intent(out) :: c1(*), c2(*), d1(*), d2
        integer :: m
        real(SP) :: acc
        d2 = 0.0_SP
        if (nm <= 0) return
        do m = 1, nm
            c1(m) = (2 * ts(m) - dt) / (2 * ts(m) + dt)
            c2(m) = (2) / (2 * ts(m) + dt) / nm
        end do
        acc = 0.0_SP
        do m = 1, nm
            acc = acc + dt / (2 * ts(m) - dt)
        end do
        d2 = acc / nm
        do m = 1, nm
            d1(m) = 2 * ts(m) / (2 * ts(m) - dt)
        end do
    end subroutine swpch_kernel_coef

    !! ADE-CFS PML profiles g(4,n) at cell centres and edges along one axis
    !! (src/swpc_3d/m_absorb_p.f90:79-93 with damping_profile :533-573)
    subroutine swpch_pml_profile(n, na, beg, h, fcut, dt, gc, ge) bind(C, name='swpch_pml_profile')
        integer(c_int), value :: n, na
        real(c_float), value :: beg, h, fcut, dt
        real(c_float), intent(out) :: gc(4, n), ge(4, n)
        real(SP) :: hh, fin, x
        integer :: i
        hh = na * h
        fin = beg + n * h
        do i = 1, n
            x = swpch_i2x(i, beg, h)
            call damping(x, gc(:, i))
            call damping(x + h / 2.0_SP, ge(:, i))
        end do
    contains
        subroutine damping(xx0, g)
            real(SP), intent(in) :: xx0
            real(SP), intent(out) :: g(4)
            real(SP) :: r0, d0, a0, b0, xx, r, d, a, b, den
            r0 = 10**(-(log10(real(na, SP)) - 1) / log10(2.0_SP) - 3.0_SP)
            d0 = -(1.0_SP / (2.0_SP * hh)) * 2 * 6.0_SP * log(r0)
            b0 = 7.0_SP
            a0 = real(PI * fcut, SP)
            if (xx0 <= beg + hh) then
                xx = (beg + hh) - xx0
            else if (xx0 >= fin - hh) then
                xx = xx0 - (fin - hh)
            else
                xx = 0.0_SP
            end if
            r = abs(xx / hh)
            d = d0 * r
            a = a0 * (1.0_SP - r)
            b = 1.0_SP + (b0 - 1.0_SP) * r**2
            den = 1.0_SP + (dt / 2.0_SP) * (a + d / b)
            g(1) = ((1.0_SP + (dt / 2.0_SP) * a) / b) / den
            g(2) = (-1.0_SP / b) / den
            g(3) = (1.0_SP - (dt / 2.0_SP) * (a + d / b)) / den
            g(4) = (d / b) / den
        end subroutine damping
    end subroutine swpch_pml_profile

    !! Cerjan sponge exp(-0.09 (1 - d/L)^2) along one axis, halo -2:n+3
    !! (src/swpc_3d/m_absorb_c.f90:27-101); top = 0 leaves the low side
    !! untapered (z: free surface).
    subroutine swpch_cerjan_profile(n, na, h, top, g_c, g_b) bind(C, name='swpch_cerjan_profile')
        integer(c_int), value :: n, na, top
        real(c_float), value :: h
        real(c_float), intent(out) :: g_c(-2:n + 3), g_b(-2:n + 3)
        real(SP), parameter :: alpha = 0.09_SP
        real(SP) :: L
        integer :: i
        L = na * h
        g_c = 1.0_SP
        g_b = 1.0_SP
        do i = 1, n
            if (i <= na) then
                if (top == 0) cycle
                g_c(i) = exp(-alpha * (1.0_SP - (swpch_i2x(i, 0.0_SP, h)) / L)**2)
                g_b(i) = exp(-alpha * (1.0_SP - ((swpch_i2x(i, 0.0_SP, h) + h / 2)) / L)**2)
            else if (i >= n - na + 1) then
                g_c(i) = exp(-alpha * (1.0_SP - (swpch_i2x(i, n * h, -h) + h / 2) / L)**2)
                g_b(i) = exp(-alpha * (1.0_SP - ((swpch_i2x(i, n * h, -h))) / L)**2)
            end if
        end do
    end subroutine swpch_cerjan_profile

    !! double-couple moment tensor from strike/dip/rake [deg]
    !! (src/shared/m_fdtool.f90:307-336); m6 = xx yy zz yz xz xy
    subroutine swpch_sdr2moment(strike, dip, rake, m6) bind(C, name='swpch_sdr2moment')
        real(c_float), value :: strike, dip, rake
        real(c_float), intent(out) :: m6(6)
        real(SP) :: sd, cd, s2d, c2d, sl, cl, sf, cf, s2f, c2f
        sd = sin(d2r(dip)); cd = cos(d2r(dip))
        s2d = sin(d2r(2 * dip)); c2d = cos(d2r(2 * dip))
        sl = sin(d2r(rake)); cl = cos(d2r(rake))
        sf = sin(d2r(strike)); cf = cos(d2r(strike))
        s2f = sin(d2r(2 * strike)); c2f = cos(d2r(2 * strike))
        m6(1) = -(sd * cl * s2f + s2d * sl * sf * sf)
        m6(6) = (sd * cl * c2f + s2d * sl * s2f / 2)
        m6(5) = -(cd * cl * cf + c2d * sl * sf)
        m6(2) = (sd * cl * s2f - s2d * sl * cf * cf)
        m6(4) = -(cd * cl * sf - c2d * sl * cf)
        m6(3) = (s2d * sl)
    contains
        real(SP) function d2r(deg)
            real(SP), intent(in) :: deg
            d2r = real(PI / 180.0_SP * deg, SP)
        end function d2r
    end subroutine swpch_sdr2moment

    !! moment-rate functions (src/shared/m_fdtool.f90:338-497);
    !! stftype 0 boxcar, 1 triangle, 2 herrmann, 3 kupper, 4 cosine, 5 texp
    real(c_float) function swpch_momentrate(t, stftype, ts, tr) bind(C, name='swpch_momentrate')
        real(c_float), value :: t, ts, tr
        integer(c_int), value :: stftype
        real(SP) :: tt, t1, t2
        swpch_momentrate = 0.0_SP
        select case (stftype)
        case (0)
            if (ts <= t .and. t <= ts + tr) swpch_momentrate = 1.0_SP / tr
        case (1)
            if (ts <= t .and. t <= ts + tr / 2) then
                swpch_momentrate = 4 * (t - ts) / (tr * tr)
            else if (ts + tr / 2 < t .and. t <= ts + tr) then
                swpch_momentrate = -4 * (t - ts - tr) / (tr * tr)
            end if
        case (2)
            t1 = ts + tr / 4
            t2 = ts + 3 * tr / 4
            if (ts <= t .and. t < t1) then
                swpch_momentrate = 16 * (t - ts)**2 / (tr**3)
            else if (t1 <= t .and. t < t2) then
                swpch_momentrate = -2 * (8 * (t * t + tr * ts + ts * ts - t * tr - 2 * t * ts) + tr * tr) / (tr**3)
            else if (t2 <= t .and. t <= ts + tr) then
                swpch_momentrate = 16 * (ts + tr - t)**2 / (tr**3)
            end if
        case (4)
            if (ts <= t .and. t <= ts + tr) swpch_momentrate = real((1 - cos(2 * PI * (t - ts) / tr)) / tr, SP)
        case (5)
            if (ts <= t) then
                tt = t - ts
                swpch_momentrate = real((2 * PI)**2 * tt / (tr * tr) * exp(-2 * PI * tt / tr), SP)
            end if
        case default
            if (ts <= t .and. t <= ts + tr) swpch_momentrate = real(3 * PI * (sin(PI * (t - ts) / tr))**3 / (4 * tr), SP)
        end select
    end function swpch_momentrate

    !! table stf(n, it) of every source at t = tbeg + (it - shift) dt: shift
    !! 0.5 for moment tensors (m_source.f90:800), 0 for forces (:873)
    subroutine swpch_stf_table(nt, nsrc, stftype, t0, tr, tbeg, dt, shift, stf) bind(C, name='swpch_stf_table')
        integer(c_int), value :: nt, nsrc, stftype
        real(c_float), intent(in) :: t0(nsrc), tr(nsrc)
        real(c_float), value :: tbeg, dt, shift
        real(c_float), intent(out) :: stf(nsrc, nt)
        integer :: it, n
        real(SP) :: t
        do it = 1, nt
            t = tbeg + (it - shift) * dt
            do n = 1, nsrc
                stf(n, it) = swpch_momentrate(t, stftype, t0(n), tr(n))
            end do
        end do
    end subroutine swpch_stf_table

    !! station grid indices (src/swpc_3d/m_wav.f90:188-240); zsw 0 = 'dep',
    !! 1 = 'fsb' (one cell below the free surface), 2 = 'obb'.  inside = 0 for
    !! stations outside the model.
    subroutine swpch_station_index(nst, x, y, z, zsw, nx, ny, nz, xbeg, ybeg, zbeg, dx, dy, dz, kfs, kob, &
                                   ist, jst, kst, inside) bind(C, name='swpch_station_index')
        integer(c_int), value :: nst, nx, ny, nz
        real(c_float), intent(in) :: x(nst), y(nst), z(nst)
        integer(c_int), intent(in) :: zsw(nst)
        real(c_float), value :: xbeg, ybeg, zbeg, dx, dy, dz
        integer(c_int), intent(in) :: kfs(-2:nx + 3, -2:ny + 3), kob(-2:nx + 3, -2:ny + 3)
        integer(c_int), intent(out) :: ist(nst), jst(nst), kst(nst), inside(nst)
        integer :: n, i, j, k
        do n = 1, nst
            i = swpch_x2i(x(n), xbeg, dx)
            j = swpch_x2i(y(n), ybeg, dy)
            k = swpch_x2i(z(n), zbeg, dz)
            inside(n) = 0
            ist(n) = i; jst(n) = j; kst(n) = k
            if (.not. (swpch_i2x(1, xbeg, dx) < x(n) .and. x(n) < swpch_i2x(nx, xbeg, dx) .and. &
                       swpch_i2x(1, ybeg, dy) < y(n) .and. y(n) < swpch_i2x(ny, ybeg, dy) .and. 1 < k .and. k < nz)) cycle
            select case (zsw(n))
            case (1); k = kfs(i, j) + 1
            case (2); k = kob(i, j) + 1
            end select
            if (k > nz) k = nz - 1
            if (k < 1) k = 2
            kst(n) = k
            inside(n) = 1
        end do
    end subroutine swpch_station_index

    !! GF cells of green__setup for 'xyz' point lists
    !! (src/swpc_3d/m_green.f90:225-296): a point is kept when its horizontal
    !! distance from the virtual source (the station) is within maxdist and its
    !! cell lies in the grid and in the solid part, kob(i,j) <= k <= nz
    subroutine swpch_green_points(ng, x, y, z, xsrc, ysrc, maxdist, nx, ny, nz, xbeg, ybeg, zbeg, dx, dy, dz, kob, &
                                  ig, jg, kg, keep) bind(C, name='swpch_green_points')
        integer(c_int), value :: ng, nx, ny, nz
        real(c_float), intent(in) :: x(ng), y(ng), z(ng)
        real(c_float), value :: xsrc, ysrc, maxdist, xbeg, ybeg, zbeg, dx, dy, dz
        integer(c_int), intent(in) :: kob(-2:nx + 3, -2:ny + 3)
        integer(c_int), intent(out) :: ig(ng), jg(ng), kg(ng), keep(ng)
        integer :: n, ii, jj, kk
        real(SP) :: dd
        do n = 1, ng
            keep(n) = 0
            ig(n) = 0; jg(n) = 0; kg(n) = 0
            dd = sqrt((x(n) - xsrc)**2 + (y(n) - ysrc)**2)
            if (dd > maxdist) cycle
            ii = swpch_x2i(x(n), xbeg, dx)
            jj = swpch_x2i(y(n), ybeg, dy)
            kk = swpch_x2i(z(n), zbeg, dz)
            ig(n) = ii; jg(n) = jj; kg(n) = kk
            if (1 <= ii .and. ii <= nx .and. 1 <= jj .and. jj <= ny) then
                if (kob(ii, jj) <= kk .and. kk <= nz) keep(n) = 1
            end if
        end do
    end subroutine swpch_green_points

end module swpc_host
