! sac_driver.f90 -- drives the reference's own src/shared/m_sac.f90 and
! m_tar.f90 (compiled unmodified from /root/reference by
! oracle/ref_sac/Makefile) and writes SAC, CSF and tar files of fixed headers
! and data into the directory given as the first argument.  The files are
! the golden vectors of the SAC/CSF/tar writers (tests/test_sacio.py).
! Test infrastructure only.
program sac_driver
    use m_std
    use m_tar
    use m_sac
    implicit none
    character(256) :: odir
    type(sac__hdr) :: sh, sd, s3(3)
    real(SP) :: d1(37), d3(37, 3), d98(98)
    integer :: i, io

    call get_command_argument(1, odir)

    do i = 1, 37
        d1(i) = real(i, SP) * 0.25 - 3.0
        d3(i, 1) = d1(i)
        d3(i, 2) = real(i, SP) * 1.0e-3
        d3(i, 3) = -real(i * i, SP) * 7.5e-5
    end do
    do i = 1, 98
        d98(i) = real(i, SP) * 0.5 - 20.0
    end do

    ! every header field sac__whdr writes, with values whose rounding to
    ! single precision and truncation (delta) matter
    call sac__init(sh)
    sh%delta = 0.15000000000000002d0
    sh%depmin = -3.0d0; sh%depmax = 6.25d0
    sh%b = -0.1d0; sh%e = 5.3d0; sh%o = 1.0d0/3.0d0; sh%a = 2.7182818284590452d0
    sh%t0 = 0.1d0; sh%t1 = 0.2d0; sh%t2 = 0.3d0; sh%t3 = 0.4d0; sh%t4 = 0.5d0
    sh%t5 = 0.6d0; sh%t6 = 0.7d0; sh%t7 = 0.8d0; sh%t8 = 0.9d0; sh%t9 = 1.1d0
    sh%stla = 35.123456789d0; sh%stlo = 139.87654321d0; sh%stel = 12.5d0; sh%stdp = 1500.0d0
    sh%evla = 41.5d0; sh%evlo = 142.0d0; sh%evel = -3.0d0; sh%evdp = 30.0d0; sh%mag = 6.3333333333d0
    sh%user0 = 1.0d-1; sh%user1 = -2.0d-2; sh%user2 = 3.0d-3; sh%user3 = 4.0d4; sh%user4 = -5.5d0
    sh%user5 = 6.6d0; sh%user6 = 141.4d0; sh%user7 = 41.1d0; sh%user8 = 0.0d0; sh%user9 = 9.9d0
    sh%dist = 123.456d0; sh%az = 45.6789d0; sh%baz = 225.6789d0; sh%gcarc = 1.2345d0
    sh%depmen = 0.125d0; sh%cmpaz = 90.0d0; sh%cmpinc = 90.0d0
    sh%user10 = 10.0d0; sh%user11 = 11.0d0; sh%user12 = 12.0d0; sh%user13 = 13.0d0
    sh%user14 = 14.0d0; sh%user15 = 15.0d0; sh%user16 = 16.0d0
    sh%nzyear = 2025; sh%nzjday = 123; sh%nzhour = 4; sh%nzmin = 5; sh%nzsec = 6; sh%nzmsec = 7
    sh%npts = 37; sh%idep = 7; sh%ievtyp = 40
    sh%iuser0 = 100; sh%iuser1 = 101; sh%iuser2 = 102; sh%iuser3 = 103
    sh%iuser4 = 104; sh%iuser5 = 105; sh%iuser6 = 106; sh%iuser7 = 107
    sh%leven = .true.; sh%lpspol = .true.; sh%lovrok = .false.; sh%lcalda = .true.; sh%luser0 = .true.
    sh%kstnm = 'N.TMRF'; sh%kevnm = 'swpc_example_run'; sh%khole = 'H1'; sh%ko = 'ORIGIN'
    sh%ka = 'P'; sh%kt0 = 'T0'; sh%kt1 = 'T1'; sh%kt2 = 'T2'; sh%kt3 = 'T3'; sh%kt4 = 'T4'
    sh%kt5 = 'T5'; sh%kt6 = 'T6'; sh%kt7 = 'T7'; sh%kt8 = 'T8'; sh%kt9 = 'T9'; sh%kf = 'FIN'
    sh%kuser0 = 'U0'; sh%kuser1 = 'U1'; sh%kuser2 = 'U2'; sh%kcmpnm = 'Vx'; sh%knetwk = 'NET'
    sh%kdatrd = 'DATRD'; sh%kinst = 'SWPC'
    sh%tim = 1735657200
    call sac__write(trim(odir)//'/full.sac', sh, d1, .true.)

    ! defaults of sac__init with the three fields every trace sets
    call sac__init(sd)
    sd%delta = 0.05d0; sd%npts = 37; sd%b = 0.0d0
    call sac__write(trim(odir)//'/default.sac', sd, d1, .true.)

    ! concatenated SAC: three traces of one header set
    do i = 1, 3
        s3(i) = sh
        s3(i)%npts = 37
    end do
    s3(2)%kcmpnm = 'Vy'; s3(2)%cmpaz = 180.0d0
    s3(3)%kcmpnm = 'Vz'; s3(3)%cmpinc = 0.0d0; s3(3)%idep = 6
    call csf__write(trim(odir)//'/three.csf', 3, 37, s3, d3, .true.)

    ! tar: one member with a partial last block, one whose size is a whole
    ! number of 512-byte blocks ((158 + 98) * 4 = 1024)
    open (newunit=io, file=trim(odir)//'/two.tar', action='write', access='stream', status='replace')
    call sac__wtar(io, 'swpc.N.TMRF.3d.Vx.sac', sh, d1)
    sd%npts = 98
    sd%tim = 1700000123
    call sac__wtar(io, 'swpc.ST02.3d.Uz.sac', sd, d98)
    call tar__wend(io)
    close (io)
end program sac_driver
