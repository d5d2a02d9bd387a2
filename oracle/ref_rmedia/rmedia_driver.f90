! rmedia_driver.f90 -- drives the reference's own src/shared/m_rmedia.f90 and
! m_fk.f90 (compiled unmodified from /root/reference by
! oracle/ref_rmedia/Makefile) and prints golden vectors for the random-media
! oracle: power spectral density values, FFT wavenumbers and two small
! realisations, as IEEE single bit patterns.  Test infrastructure only.
program rmedia_driver
    use m_std
    use m_fk
    use m_rmedia
    implicit none
    real(SP) :: ms(6) = [0.0, 0.05, 0.3, 1.0, 2.5, 7.0]
    real(SP) :: as(3) = [0.5, 2.0, 5.0]
    real(SP) :: kaps(3) = [0.1, 0.5, 1.0]
    real(SP) :: p, mx, my, mz
    integer :: it, ia, ik, im, n, i
    integer :: ns(4) = [8, 9, 16, 17]
    real(SP) :: ds(4) = [0.5, 0.25, 1.0, 0.1]
    real(SP), allocatable :: a3(:, :, :), kx(:), ky(:), kz(:), x(:), y(:), z(:)
    complex(SP), allocatable :: b3(:, :, :)

    ! isotropic 3D PSDFs (m_rmedia.f90:590-679)
    do im = 1, size(ms)
        do ia = 1, size(as)
            call rmedia__psdf_gauss3d(ms(im), as(ia), 0.05, p)
            write (*, '(A,3(1X,Z8.8))') 'iso 1', transfer(ms(im), 0), transfer(as(ia), 0), transfer(p, 0)
            call rmedia__psdf_exp3d(ms(im), as(ia), 0.05, p)
            write (*, '(A,3(1X,Z8.8))') 'iso 2', transfer(ms(im), 0), transfer(as(ia), 0), transfer(p, 0)
            do ik = 1, size(kaps)
                call rmedia__psdf_vonKarman3d(ms(im), as(ia), 0.05, kaps(ik), p)
                write (*, '(A,4(1X,Z8.8))') 'iso 3', transfer(ms(im), 0), transfer(as(ia), 0), &
                    transfer(kaps(ik), 0), transfer(p, 0)
            end do
        end do
    end do
    ! anisotropic 3D PSDFs (m_rmedia.f90:682-732)
    do it = 1, 3
        do im = 1, size(ms)
            mx = ms(im); my = 0.5 * ms(im) + 0.1; mz = 0.25 * ms(im)
            select case (it)
            case (1); call rmedia__psdf_ani_gauss3d(mx, my, mz, 2.0, 1.0, 0.5, 0.07, p)
            case (2); call rmedia__psdf_ani_exp3d(mx, my, mz, 2.0, 1.0, 0.5, 0.07, p)
            case (3); call rmedia__psdf_ani_vonKarman3d(mx, my, mz, 2.0, 1.0, 0.5, 0.07, 0.3, p)
            end select
            write (*, '(A,I2,4(1X,Z8.8))') 'ani', it, transfer(mx, 0), transfer(my, 0), transfer(mz, 0), transfer(p, 0)
        end do
    end do
    ! wavenumbers of fk__x2k_3d (m_fk.f90)
    do i = 1, size(ns)
        n = ns(i)
        allocate (a3(n, 2, 2), b3(n, 2, 2), kx(n), ky(2), kz(2))
        a3 = 1.0
        call fk__x2k_3d(n, 2, 2, ds(i), 1.0, 1.0, a3, b3, kx, ky, kz)
        write (*, '(A,I3,1X,Z8.8)', advance='no') 'k', n, transfer(ds(i), 0)
        do it = 1, n
            write (*, '(1X,Z8.8)', advance='no') transfer(kx(it), 0)
        end do
        write (*, *)
        deallocate (a3, b3, kx, ky, kz)
    end do
    ! two realisations (rmedia__3dgen_ani, rmedia__3dgen); the FFT of
    ! m_rfft.f90 is radix 2, so every size is a power of two
    allocate (a3(32, 16, 16), x(32), y(16), z(16))
    call rmedia__3dgen_ani(3, 32, 16, 16, 0.5, 0.4, 0.3, 2.0, 1.5, 1.0, 0.05, 0.5, x, y, z, a3, 7)
    write (*, '(A)') 'field ani 3 32 16 16 0.5 0.4 0.3 2.0 1.5 1.0 0.05 0.5'
    write (*, '(8(Z8.8,1X))') transfer(a3, [0], size(a3))
    deallocate (a3, x, y, z)
    allocate (a3(16, 16, 8), x(16), y(16), z(8))
    call rmedia__3dgen(2, 16, 16, 8, 0.5, 0.5, 0.5, 1.5, 0.1, 0.0, x, y, z, a3, 11)
    write (*, '(A)') 'field iso 2 16 16 8 0.5 0.5 0.5 1.5 1.5 1.5 0.1 0.0'
    write (*, '(8(Z8.8,1X))') transfer(a3, [0], size(a3))
end program rmedia_driver
