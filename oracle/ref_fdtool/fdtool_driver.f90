! fdtool_driver.f90 -- drives the reference's own src/shared/m_fdtool.f90 and
! src/shared/m_seawater.f90 (compiled unmodified from /root/reference by
! oracle/ref_fdtool/Makefile)
! and prints input/output bit patterns used as golden vectors for the oracle.
! Test infrastructure only.
program fdtool_driver
    use m_std
    use m_fdtool
    use m_seawater
    implicit none
    integer :: munk
    integer :: nm, im, k, it, ityp
    real(SP) :: ts(8), zeta, fmin, fmax, t, v, m6(6), dt
    real(SP) :: fmins(3) = [0.05, 0.02, 0.1], fmaxs(3) = [5.0, 2.0, 10.0]
    real(SP) :: angles(3, 5)
    character(8) :: names(6) = [character(8) :: 'boxcar', 'triangle', 'herrmann', 'kupper', 'cosine', 'texp']
    real(SP) :: srcprm(2)

    ! relaxation times and zeta (m_fdtool.f90:691-812)
    do k = 1, 3
        fmin = fmins(k); fmax = fmaxs(k)
        do nm = 1, 5
            call visco_set_relaxtime(nm, ts(1:nm), fmin, fmax)
            zeta = visco_constq_zeta(nm, fmin, fmax, ts(1:nm))
            write (*, '(A,1X,I0,2(1X,Z8.8))', advance='no') 'visco', nm, transfer(fmin, 0), transfer(fmax, 0)
            do im = 1, nm
                write (*, '(1X,Z8.8)', advance='no') transfer(ts(im), 0)
            end do
            write (*, '(1X,Z8.8)') transfer(zeta, 0)
        end do
    end do

    ! source time functions (m_fdtool.f90:338-497), t = tbeg + (it - 0.5) * dt
    dt = 0.02
    srcprm = [0.1, 1.0]
    do ityp = 1, 6
        do it = 1, 80
            t = 0.0 + (it - 0.5) * dt
            v = momentrate(t, names(ityp), 2, srcprm)
            write (*, '(A,1X,I0,3(1X,Z8.8))') 'stf', ityp - 1, transfer(t, 0), transfer(srcprm(1), 0), transfer(v, 0)
        end do
    end do
    srcprm = [0.3, 4.0]
    do ityp = 1, 6
        do it = 1, 300
            t = 0.0 + it * 0.015
            v = momentrate(t, names(ityp), 2, srcprm)
            write (*, '(A,1X,I0,4(1X,Z8.8))') 'stf2', ityp - 1, transfer(t, 0), transfer(srcprm(1), 0), &
                transfer(srcprm(2), 0), transfer(v, 0)
        end do
    end do

    ! sdr2moment (m_fdtool.f90:307-336)
    angles = reshape([0.0, 90.0, 0.0, 30.0, 45.0, 90.0, 215.3, 62.1, -47.9, 359.9, 10.0, 179.0, &
                      123.4, 88.8, -180.0], [3, 5])
    do k = 1, 5
        call sdr2moment(angles(1, k), angles(2, k), angles(3, k), m6(1), m6(2), m6(3), m6(4), m6(5), m6(6))
        write (*, '(A,9(1X,Z8.8))') 'sdr', transfer(angles(1, k), 0), transfer(angles(2, k), 0), &
            transfer(angles(3, k), 0), (transfer(m6(im), 0), im=1, 6)
    end do

    ! index maps (m_fdtool.f90:600-649)
    do k = -3, 40
        v = -32.0 + k * 0.37
        write (*, '(A,3(1X,Z8.8),1X,I0)') 'x2i', transfer(v, 0), transfer(-32.0, 0), transfer(0.5, 0), &
            x2i(v, -32.0, 0.5)
        write (*, '(A,1X,I0,3(1X,Z8.8))') 'i2x', k, transfer(-32.0, 0), transfer(0.25, 0), &
            transfer(i2x(k, -32.0, 0.25), 0)
    end do

    ! stable dt and seismic moment
    call fdm_stable_dt(0.5, 0.5, 0.5, 6.5, v)
    write (*, '(A,5(1X,Z8.8))') 'cfl', transfer(0.5, 0), transfer(0.5, 0), transfer(0.5, 0), transfer(6.5, 0), transfer(v, 0)
    call fdm_stable_dt(0.25, 0.2, 0.3, 7.9, v)
    write (*, '(A,5(1X,Z8.8))') 'cfl', transfer(0.25, 0), transfer(0.2, 0), transfer(0.3, 0), transfer(7.9, 0), transfer(v, 0)
    do k = 1, 8
        t = 2.0 + k * 0.75
        write (*, '(A,2(1X,Z8.8))') 'mw', transfer(t, 0), transfer(seismic_moment(t), 0)
    end do

    ! seawater sound speed, constant and Munk's profile (m_seawater.f90:20-47)
    do munk = 0, 1
        call seawater__init(munk == 1)
        do k = 0, 40
            t = k * 0.173
            write (*, '(A,1X,I0,2(1X,Z8.8))') 'sea', munk, transfer(t, 0), transfer(seawater__vel(t), 0)
        end do
    end do
end program fdtool_driver
