!! swpc_fdrive -- Fortran driver of the MI355X time loop through m_swpc_gpu.
!!
!! Follows the reference main program (src/swpc_3d/main.f90:64-139):
!! medium__setup -> kernel__setup -> source__setup -> absorb__setup ->
!! wav__setup, then the time loop, here as one call to swpc_step (HIP-graph
!! replays of whole steps).  The set-up math comes from the host library
!! (module swpc_host); the arrays keep the reference's (k,i,j) halo layout.
!!
!! Model: two homogeneous layers under a 2 km air column (lhm), 8-cell PML,
!! 3 Zener mechanisms, one double-couple source, two free-surface stations.
!!
!! usage: swpc_fdrive [nx ny nz nt]     (defaults 48 44 40 30)
!! prints: vmax <3 hex words>, then seis <station> <hex words of Vz>
program swpc_fdrive
    use iso_c_binding
    use swpc_host
    use m_swpc_gpu
    implicit none

    integer, parameter :: NA = 8, NM = 3, NL = 2, NST = 2
    real(c_float), parameter :: DX = 0.5, DT = 0.02, ZBEG = -2.0
    integer(c_int) :: nx, ny, nz, nt, rc, ntw, n, it
    real(c_float) :: xbeg, ybeg, zeta, vmin, vmax, d2, fcut, M0, UC
    real(c_float), allocatable, dimension(:, :, :) :: rho, lam, mu, taup, taus
    integer(c_int), allocatable, dimension(:, :) :: kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot
    real(c_float) :: layers(6, NL), ts(NM), c1(NM), c2(NM), d1(NM), m6(6), vm(3)
    real(c_float), allocatable :: gxc(:), gxe(:), gyc(:), gye(:), gzc(:), gze(:), stf(:), seis(:, :, :)
    integer(c_int) :: isrc(1), jsrc(1), ksrc(1), ist(NST), jst(NST), kst(NST), inside(NST), zsw(NST)
    real(c_float) :: mo(1), t0(1), tr(1), stx(NST), sty(NST), stz(NST)
    type(swpc_desc) :: desc
    type(c_ptr) :: ctx
    character(len=32) :: arg

    nx = 48; ny = 44; nz = 40; nt = 30
    if (command_argument_count() >= 4) then
        call get_command_argument(1, arg); read (arg, *) nx
        call get_command_argument(2, arg); read (arg, *) ny
        call get_command_argument(3, arg); read (arg, *) nz
        call get_command_argument(4, arg); read (arg, *) nt
    end if
    xbeg = -real(nx, c_float) / 2 * DX
    ybeg = -real(ny, c_float) / 2 * DX

    ! ---- medium__setup (m_medium.f90:36-229) ----
    allocate (rho(-2:nz + 3, -2:nx + 3, -2:ny + 3), lam(-2:nz + 3, -2:nx + 3, -2:ny + 3))
    allocate (mu(-2:nz + 3, -2:nx + 3, -2:ny + 3), taup(-2:nz + 3, -2:nx + 3, -2:ny + 3))
    allocate (taus(-2:nz + 3, -2:nx + 3, -2:ny + 3))
    allocate (kfs(-2:nx + 3, -2:ny + 3), kob(-2:nx + 3, -2:ny + 3), kfs_top(-2:nx + 3, -2:ny + 3))
    allocate (kfs_bot(-2:nx + 3, -2:ny + 3), kob_top(-2:nx + 3, -2:ny + 3), kob_bot(-2:nx + 3, -2:ny + 3))
    layers(:, 1) = [0.0, 2.7, 5.5, 3.2, 320.0, 160.0]
    layers(:, 2) = [5.0, 2.9, 6.5, 3.7, 370.0, 185.0]
    call swpch_vmodel_layers(0, nx, ny, nz, ZBEG, DX, 0.0, 0, NL, layers, rho, lam, mu, taup, taus, rc)
    call check(rc, 'velocity model')
    call swpch_medium_setup(nx, ny, nz, NA, NM, 0.05, 5.0, 1.0, rho, lam, mu, taup, taus, ts, zeta, &
                            kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot, vmin, vmax)

    ! ---- kernel__setup (m_kernel.f90:30-73) ----
    call swpch_kernel_coef(NM, ts, DT, c1, c2, d1, d2)

    ! ---- source__setup (m_source.f90:41-314): one double couple ----
    isrc(1) = swpch_x2i(0.3, xbeg, DX)
    jsrc(1) = swpch_x2i(-0.6, ybeg, DX)
    ksrc(1) = swpch_x2i(5.2, ZBEG, DX)
    t0(1) = 0.0; tr(1) = 0.5; mo(1) = 1.0e15
    call swpch_sdr2moment(30.0, 60.0, 90.0, m6)
    M0 = mo(1)
    fcut = 1.0 / tr(1)
    allocate (stf(nt))
    call swpch_stf_table(nt, 1, 3, t0, tr, 0.0, DT, 0.5, stf)
    UC = 1.0000001095e-15

    ! ---- absorb__setup (m_absorb_p.f90:60-124) ----
    allocate (gxc(4 * nx), gxe(4 * nx), gyc(4 * ny), gye(4 * ny), gzc(4 * nz), gze(4 * nz))
    call swpch_pml_profile(nx, NA, xbeg, DX, fcut, DT, gxc, gxe)
    call swpch_pml_profile(ny, NA, ybeg, DX, fcut, DT, gyc, gye)
    call swpch_pml_profile(nz, NA, ZBEG, DX, fcut, DT, gzc, gze)

    ! ---- wav__setup (m_wav.f90:46-271): stations one cell below the surface ----
    stx = [2.1, -3.2]; sty = [1.3, 2.7]; stz = [0.0, 0.0]; zsw = [1, 1]
    call swpch_station_index(NST, stx, sty, stz, zsw, nx, ny, nz, xbeg, ybeg, ZBEG, DX, DX, DX, kfs, kob, &
                             ist, jst, kst, inside)
    ntw = nt

    ! ---- device context and upload ----
    desc = swpc_desc(nx, ny, nz, NA, NM, SWPC_ABC_PML, real(DX, c_double), real(DX, c_double), &
                     real(DX, c_double), DT, 0)
    call check(swpc_create(desc, ctx), 'swpc_create')
    call check(swpc_set_medium(ctx, rho, lam, mu, taup, taus, kfs_top, kfs_bot, kob_top, kob_bot, kob), 'medium')
    call check(swpc_set_visco(ctx, c1, c2, d1, d2), 'visco')
    call check(swpc_set_pml(ctx, gxc, gxe, gyc, gye, gzc, gze), 'pml')
    call check(swpc_set_sources(ctx, 1, 0, isrc, jsrc, ksrc, mo, m6, M0, nt, stf), 'sources')
    call check(swpc_set_receivers(ctx, NST, ist, jst, kst, SWPC_WAV_V, 1, ntw, M0, UC), 'receivers')

    ! ---- time loop (main.f90:119-139) ----
    it = 1
    call check(swpc_step(ctx, it, nt, 1, 10), 'swpc_step')
    call check(swpc_kernel_vmax(ctx, vm), 'vmax')
    write (*, '(A,3(1X,Z8.8))') 'vmax', (transfer(vm(n), 0_c_int32_t), n=1, 3)

    ! ---- wav__write: velocity traces (ntw, 3, nst) ----
    allocate (seis(ntw, 3, NST))
    call check(swpc_get_seismograms(ctx, SWPC_WAV_V, seis), 'seismograms')
    do n = 1, NST
        write (*, '(A,1X,I0,*(1X,Z8.8))') 'seis', n, (transfer(seis(it, 3, n), 0_c_int32_t), it=1, ntw)
    end do
    call check(swpc_destroy(ctx), 'destroy')

contains

    subroutine check(code, what)
        integer(c_int), intent(in) :: code
        character(*), intent(in) :: what
        if (code /= 0) then
            write (*, '(A,A,A,I0,A,A)') 'swpc_fdrive: ', what, ' failed (', code, '): ', &
                trim(swpc_last_error_f(ctx))
            error stop 1
        end if
    end subroutine check

end program swpc_fdrive
