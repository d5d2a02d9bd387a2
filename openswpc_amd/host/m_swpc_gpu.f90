!! m_swpc_gpu -- iso_c_binding interfaces to libswpc_gpu (include/swpc_gpu.h)
!!
!! A Fortran caller (the reference's own main loop, or swpc_fdrive.f90 here)
!! uses this module to drive the MI355X time loop.  Arrays are passed in the
!! reference's layout: 3D (k,i,j) with bounds (-2:nz+3, -2:nx+3, -2:ny+3),
!! 2D maps (-2:nx+3, -2:ny+3), real(SP) = c_float.  Every function returns 0
!! on success and a negative code otherwise; swpc_last_error_f gives the text.
module m_swpc_gpu
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: SWPC_ABC_PML = 0, SWPC_ABC_CERJAN = 1
    integer(c_int), parameter :: SWPC_VX = 0, SWPC_VY = 1, SWPC_VZ = 2, SWPC_SXX = 3, SWPC_SYY = 4, &
                                 SWPC_SZZ = 5, SWPC_SYZ = 6, SWPC_SXZ = 7, SWPC_SXY = 8
    integer(c_int), parameter :: SWPC_WAV_V = 1, SWPC_WAV_U = 2, SWPC_WAV_STRESS = 4, SWPC_WAV_STRAIN = 8
    integer(c_int), parameter :: SWPC_NTIMER = 10
    integer(c_int), parameter :: SWPC_SNP_XY = 0, SWPC_SNP_YZ = 1, SWPC_SNP_XZ = 2, SWPC_SNP_FS = 3, SWPC_SNP_OB = 4
    integer(c_int), parameter :: SWPC_SNP_PS = 0, SWPC_SNP_V = 1, SWPC_SNP_U = 2, SWPC_SNP_NSECTION = 15
    integer(c_int), parameter :: SWPC_MED_RHO = 0, SWPC_MED_LAM = 1, SWPC_MED_MU = 2, SWPC_MED_TAUP = 3, &
                                 SWPC_MED_TAUS = 4, SWPC_MED_KFS_TOP = 5, SWPC_MED_KFS_BOT = 6, SWPC_MED_KOB_TOP = 7, &
                                 SWPC_MED_KOB_BOT = 8, SWPC_MED_KOB = 9

    type, bind(C) :: swpc_desc
        integer(c_int) :: nx, ny, nz, na, nm, abc
        real(c_double) :: dx, dy, dz
        real(c_float) :: dt
        integer(c_int) :: device
    end type swpc_desc

    !! section id = plane * 3 + quantity
    type, bind(C) :: swpc_snap_desc
        integer(c_int) :: sw(SWPC_SNP_NSECTION)
        integer(c_int) :: k0_xy, i0_yz, j0_xz
        integer(c_int) :: idec, jdec, kdec
        integer(c_int) :: ntdec_s, nframe
        real(c_float) :: M0, UC
    end type swpc_snap_desc

    interface
        integer(c_int) function swpc_create(desc, ctx) bind(C, name='swpc_create')
            import :: c_int, c_ptr, swpc_desc
            type(swpc_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function swpc_destroy(ctx) bind(C, name='swpc_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function swpc_last_error(ctx) bind(C, name='swpc_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_set_medium(ctx, rho, lam, mu, taup, taus, kfs_top, kfs_bot, kob_top, kob_bot, &
                                                kob) bind(C, name='swpc_set_medium')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: rho(*), lam(*), mu(*), taup(*), taus(*)
            integer(c_int), intent(in) :: kfs_top(*), kfs_bot(*), kob_top(*), kob_bot(*), kob(*)
        end function
        !! medium__setup of a layered model built on the device: prof(-2:nz+3, 5)
        !! = rho, lam, mu, Qp, Qs per plane (swpch_vmodel_profile), xi = the
        !! perturbation's address (c_loc of a (k,i,j) array, the staging buffer or
        !! c_null_ptr), clip = vmin, vmax, rhomin, tau = zeta, Re cc, Im cc
        !! (swpch_visco_setup); returns the kfs/kob maps and vmin, vmax
        integer(c_int) function swpc_medium_layered(ctx, prof, xi, clip, tau, kfs, kob, vrange) &
            bind(C, name='swpc_medium_layered')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx, xi
            real(c_float), intent(in) :: prof(*), clip(3), tau(3)
            integer(c_int), intent(out) :: kfs(*), kob(*)
            real(c_float), intent(out) :: vrange(2)
        end function
        integer(c_int) function swpc_get_medium(ctx, which, out) bind(C, name='swpc_get_medium')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, out
            integer(c_int), value :: which
        end function
        type(c_ptr) function swpc_staging_buffer(ctx, nfloats) bind(C, name='swpc_staging_buffer')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: nfloats
        end function
        integer(c_int) function swpc_set_visco(ctx, c1, c2, d1, d2) bind(C, name='swpc_set_visco')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: c1(*), c2(*), d1(*)
            real(c_float), value :: d2
        end function
        integer(c_int) function swpc_set_pml(ctx, gxc, gxe, gyc, gye, gzc, gze) bind(C, name='swpc_set_pml')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: gxc(*), gxe(*), gyc(*), gye(*), gzc(*), gze(*)
        end function
        integer(c_int) function swpc_set_cerjan(ctx, gx_c, gx_b, gy_c, gy_b, gz_c, gz_b) bind(C, name='swpc_set_cerjan')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: gx_c(*), gx_b(*), gy_c(*), gy_b(*), gz_c(*), gz_b(*)
        end function
        integer(c_int) function swpc_set_sources(ctx, nsrc, bf_mode, isrc, jsrc, ksrc, mo, comp, M0, nt, stf) &
            bind(C, name='swpc_set_sources')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsrc, bf_mode, nt
            integer(c_int), intent(in) :: isrc(*), jsrc(*), ksrc(*)
            real(c_float), intent(in) :: mo(*), comp(*), stf(*)
            real(c_float), value :: M0
        end function
        integer(c_int) function swpc_set_receivers(ctx, nst, ist, jst, kst, what, ntdec_w, ntw, M0, UC) &
            bind(C, name='swpc_set_receivers')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: nst, what, ntdec_w, ntw
            integer(c_int), intent(in) :: ist(*), jst(*), kst(*)
            real(c_float), value :: M0, UC
        end function
        integer(c_int) function swpc_kernel_update_stress(ctx) bind(C, name='swpc_kernel_update_stress')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_absorb_update_stress(ctx) bind(C, name='swpc_absorb_update_stress')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_source_stressglut(ctx, it) bind(C, name='swpc_source_stressglut')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function
        integer(c_int) function swpc_kernel_update_vel(ctx) bind(C, name='swpc_kernel_update_vel')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_absorb_update_vel(ctx) bind(C, name='swpc_absorb_update_vel')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_source_bodyforce(ctx, it) bind(C, name='swpc_source_bodyforce')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function
        integer(c_int) function swpc_wav_store(ctx, it) bind(C, name='swpc_wav_store')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function
        integer(c_int) function swpc_update_stress(ctx) bind(C, name='swpc_update_stress')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_update_vel(ctx) bind(C, name='swpc_update_vel')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_step(ctx, it0, nsteps, use_graph, block) bind(C, name='swpc_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it0, nsteps, use_graph, block
        end function
        integer(c_int) function swpc_step_timed(ctx, it0, nsteps, ms) bind(C, name='swpc_step_timed')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: it0, nsteps
            real(c_float), intent(out) :: ms(*)
        end function
        integer(c_int) function swpc_set_snapshots(ctx, desc, kfs) bind(C, name='swpc_set_snapshots')
            import :: c_int, c_ptr, swpc_snap_desc
            type(c_ptr), value :: ctx
            type(swpc_snap_desc), intent(in) :: desc
            integer(c_int), intent(in) :: kfs(*)
        end function
        integer(c_int) function swpc_snap_write(ctx, it) bind(C, name='swpc_snap_write')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function
        integer(c_int) function swpc_get_snapshot(ctx, section, frame, out) bind(C, name='swpc_get_snapshot')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: section, frame
            real(c_float), intent(out) :: out(*)
        end function
        integer(c_int) function swpc_get_snap_max(ctx, section, out) bind(C, name='swpc_get_snap_max')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: section
            real(c_float), intent(out) :: out(*)
        end function
        integer(c_int) function swpc_snap_dims(ctx, section, dims) bind(C, name='swpc_snap_dims')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: section
            integer(c_int), intent(out) :: dims(3)
        end function
        integer(c_int) function swpc_kernel_vmax(ctx, out) bind(C, name='swpc_kernel_vmax')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: out(3)
        end function
        !! kernel__vmax without allocation: the maxima into device memory d_out
        integer(c_int) function swpc_kernel_vmax_noalloc(ctx, d_out) bind(C, name='swpc_kernel_vmax_noalloc')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, d_out
        end function
        integer(c_int) function swpc_get_field(ctx, field, out) bind(C, name='swpc_get_field')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: field
            real(c_float), intent(out) :: out(*)
        end function
        integer(c_int) function swpc_set_field(ctx, field, in) bind(C, name='swpc_set_field')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: field
            real(c_float), intent(in) :: in(*)
        end function
        integer(c_int) function swpc_get_seismograms(ctx, kind, out) bind(C, name='swpc_get_seismograms')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: kind
            real(c_float), intent(out) :: out(*)
        end function
        !! reciprocity Green's-function mode (src/swpc_3d/m_green.f90)
        integer(c_int) function swpc_set_green(ctx, ng, ig, jg, kg, isrc, jsrc, ksrc, cmp, bforce, ntdec_w, ntw, &
                                               nt, stf) bind(C, name='swpc_set_green')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: ng, isrc, jsrc, ksrc, cmp, bforce, ntdec_w, ntw, nt
            integer(c_int), intent(in) :: ig(*), jg(*), kg(*)
            real(c_float), intent(in) :: stf(*)
        end function
        integer(c_int) function swpc_green_store(ctx, it) bind(C, name='swpc_green_store')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function
        integer(c_int) function swpc_green_source(ctx, it) bind(C, name='swpc_green_source')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function
        integer(c_int) function swpc_get_green(ctx, out) bind(C, name='swpc_get_green')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: out(*)
        end function
        !! random media (src/shared/m_rmedia.f90:40-173); media(n1, n2, n3)
        integer(c_int) function swpc_rmedia_3d(device, ptype, n1, n2, n3, d1, d2, d3, a1, a2, a3, epsil, kappa, &
                                               seed, media) bind(C, name='swpc_rmedia_3d')
            import :: c_int, c_float, c_int64_t
            integer(c_int), value :: device, ptype, n1, n2, n3
            real(c_float), value :: d1, d2, d3, a1, a2, a3, epsil, kappa
            integer(c_int64_t), value :: seed
            real(c_float), intent(out) :: media(*)
        end function
        integer(c_int) function swpc_adj_gradient(ctx, nt, obs, ckp, misfit, g_lam, g_mu, g_rho) &
            bind(C, name='swpc_adj_gradient')
            import :: c_int, c_ptr, c_float, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nt, ckp
            real(c_float), intent(in) :: obs(*)
            real(c_double), intent(out) :: misfit
            real(c_float), intent(out) :: g_lam(*), g_mu(*), g_rho(*)
        end function
        integer(c_int) function swpc_adj_plan(ctx, nt, budget, ckp_outer, ckp_inner, bytes, fwd_steps) &
            bind(C, name='swpc_adj_plan')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: nt
            integer(c_int64_t), value :: budget
            integer(c_int), intent(out) :: ckp_outer, ckp_inner
            integer(c_int64_t), intent(out) :: bytes, fwd_steps
        end function
        integer(c_int) function swpc_adj_reset(ctx) bind(C, name='swpc_adj_reset')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_adj_accumulate(ctx, nt, obs, ckp_outer, ckp_inner, misfit) &
            bind(C, name='swpc_adj_accumulate')
            import :: c_int, c_ptr, c_float, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nt, ckp_outer, ckp_inner
            real(c_float), intent(in) :: obs(*)
            real(c_double), intent(out) :: misfit
        end function
        integer(c_int) function swpc_adj_get_gradient(ctx, g_lam, g_mu, g_rho) bind(C, name='swpc_adj_get_gradient')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: g_lam(*), g_mu(*), g_rho(*)
        end function
        integer(c_int) function swpc_adj_release(ctx) bind(C, name='swpc_adj_release')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function swpc_batch_create(members, n, batch) bind(C, name='swpc_batch_create')
            import :: c_int, c_ptr
            type(c_ptr), intent(in) :: members(*)
            integer(c_int), value :: n
            type(c_ptr), intent(out) :: batch
        end function
        integer(c_int) function swpc_batch_destroy(batch) bind(C, name='swpc_batch_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
        end function
        type(c_ptr) function swpc_batch_last_error(batch) bind(C, name='swpc_batch_last_error')
            import :: c_ptr
            type(c_ptr), value :: batch
        end function
        integer(c_int) function swpc_batch_step(batch, it0, nsteps, use_graph, block) bind(C, name='swpc_batch_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: it0, nsteps, use_graph, block
        end function
        integer(c_int) function swpc_batch_synchronize(batch) bind(C, name='swpc_batch_synchronize')
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
        end function
        integer(c_int) function swpc_batch_tune(batch, trial_steps, kchunk, by, width, ms) bind(C, name='swpc_batch_tune')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: batch
            integer(c_int), value :: trial_steps
            integer(c_int), intent(out) :: kchunk, by, width
            real(c_float), intent(out) :: ms
        end function
        integer(c_int) function swpc_batch_set_packing(batch, packing) bind(C, name='swpc_batch_set_packing')
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: packing
        end function
        integer(c_int) function swpc_batch_get_packing(batch) bind(C, name='swpc_batch_get_packing')
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
        end function
        type(c_ptr) function swpc_batch_stream(batch) bind(C, name='swpc_batch_stream')
            import :: c_ptr
            type(c_ptr), value :: batch
        end function
        integer(c_int) function swpc_synchronize(ctx) bind(C, name='swpc_synchronize')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int64_t) function swpc_device_bytes(ctx) bind(C, name='swpc_device_bytes')
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: ctx
        end function
        ! per-cell arrays striped over the HBM regions (call before swpc_create)
        integer(c_int) function swpc_stripe_reserve(device, bytes_per_region, nregions) &
                bind(C, name='swpc_stripe_reserve')
            import :: c_int, c_int64_t
            integer(c_int), value :: device, nregions
            integer(c_int64_t), value :: bytes_per_region
        end function
        integer(c_int) function swpc_stripe_release(device) bind(C, name='swpc_stripe_release')
            import :: c_int
            integer(c_int), value :: device
        end function
        integer(c_int64_t) function swpc_stripe_free_bytes(device, region) bind(C, name='swpc_stripe_free_bytes')
            import :: c_int, c_int64_t
            integer(c_int), value :: device, region
        end function
        integer(c_int) function swpc_set_tuning(ctx, kchunk, by, variant) bind(C, name='swpc_set_tuning')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: kchunk, by, variant
        end function
        type(c_ptr) function swpc_stream(ctx) bind(C, name='swpc_stream')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
    end interface

contains

    !! swpc_last_error as a Fortran string
    function swpc_last_error_f(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=512) :: msg
        character(kind=c_char), pointer :: p(:)
        type(c_ptr) :: cp
        integer :: n
        msg = ''
        cp = swpc_last_error(ctx)
        if (.not. c_associated(cp)) return
        call c_f_pointer(cp, p, [512])
        do n = 1, 512
            if (p(n) == c_null_char) exit
            msg(n:n) = p(n)
        end do
    end function swpc_last_error_f

end module m_swpc_gpu
