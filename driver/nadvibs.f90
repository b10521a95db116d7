! nadvibs.x -- drop-in driver for the Lanczos path of NadVibS on one MI355X.
!
! Reads the reference's inputs (basis.in namelist &NADVIBS, nadvibs.in,
! restart.log when restart /= 0), runs the Lanczos loop of Source/main.f90:63-71
! through the C ABI of libnadvibs.so, and writes restart.log, the q1/q2/q3
! restart vectors (nadvibs.35.0 in the -o directory) and the eigenvalue table
! of output.dat in the reference's formats (Source/top.f90:1007-1112).
!
!   nadvibs.x [-m MB] [-o dir]
! Environment: NV_PREC=f32 runs H*v with FP32 couplings (default f64),
! NV_QUIRKS=exact uses the exact symmetric operator (default: the
! reference's conventions), NV_LOOP=1 issues the per-iteration entry points
! one by one instead of the captured-graph run.
program nadvibs_main
    use iso_c_binding
    use nadvibs_abi
    implicit none
    real(c_double), parameter :: AU2WAVE = 219474.6313705d0, AU2EV = 27.211384523d0
    ! basis.in (Source/secondary.f90:826-870)
    integer :: niter, natoms, nmodes, nstates, restart, bconv, idroots, soroots, reorthog, chkorthog, nseg
    integer :: shiftref, nirreps, ordr
    integer :: basis(100), initstate(100), npirr(10)
    real(8) :: ztoler, maxdisk, weights(10)
    namelist /NADVIBS/ niter, natoms, nmodes, nstates, basis, restart, bconv, idroots, soroots, reorthog, &
                       chkorthog, nseg, ztoler, maxdisk, weights, shiftref, nirreps, npirr, ordr, initstate
    ! problem arrays
    integer(c_int64_t), allocatable, target :: nfunc(:), npirrep(:), istate(:)
    real(c_double), allocatable, target :: aomega(:), concoef(:), potterms(:, :, :), statew(:)
    real(c_double), allocatable, target :: bomega(:), dvec(:), tmat(:)
    real(c_double), allocatable, target :: alpha(:), beta(:), omega(:, :)
    integer(c_int64_t), allocatable :: counts(:)
    type(nv_problem), target :: prob
    type(nv_lanczos_opts) :: opts
    type(c_ptr) :: ctx
    character(len=256) :: outdir, arg, env
    character(kind=c_char, len=300), target :: proname
    logical :: keep, have_pro
    integer :: nstblks, ntmax, i, j, k, iiter, nload, ios, prec
    integer(c_int64_t) :: it
    logical :: orthog
    real(8) :: t0, t1

    ! ---- command line (Source/top.f90:6-30): -m MB (unused), -o dir ----
    outdir = './'
    i = 1
    do while (i <= command_argument_count())
        call get_command_argument(i, arg)
        if (trim(arg) == '-o' .and. i < command_argument_count()) then
            call get_command_argument(i + 1, outdir)
            outdir = trim(outdir)//'/'
            i = i + 1
        else if (trim(arg) == '-m') then
            i = i + 1
        end if
        i = i + 1
    end do

    ! ---- basis.in with the reference's defaults ----
    restart = 0; niter = 1; basis = 1; initstate = 0; nstates = 1; ordr = 2; natoms = 2; nmodes = 1
    nirreps = 1; npirr = 0; shiftref = 0; bconv = 0; idroots = 0; soroots = 0; reorthog = 0
    chkorthog = 100; nseg = 1; ztoler = 1d-20; maxdisk = 1000; weights = 1d0
    open(11, file='basis.in', status='old', iostat=ios)
    if (ios /= 0) stop 'nadvibs: basis.in not found'
    read(11, nml=NADVIBS)
    close(11)
    if (nirreps == 1 .and. npirr(1) == 0) npirr(1) = nmodes
    if (sum(npirr(1:nirreps)) /= nmodes) then
        write(*, '(1x,a)') 'Warning: symmetry is disabled due to conflict input'
        nirreps = 1; npirr(1) = nmodes
    end if
    orthog = reorthog > 0
    iiter = 1
    nstblks = nstates * (nstates + 1) / 2
    ntmax = nterms(nmodes, ordr)

    ! ---- nadvibs.in (Source/top.f90:122-135): one skipped record per group ----
    allocate(nfunc(nmodes), aomega(nmodes), concoef(nstblks), potterms(ntmax, ordr, nstblks), statew(nstates))
    allocate(npirrep(nirreps), istate(nmodes))
    nfunc = basis(1:nmodes); statew = weights(1:nstates); npirrep = npirr(1:nirreps); istate = initstate(1:nmodes)
    potterms = 0d0
    open(12, file='nadvibs.in', status='old', iostat=ios)
    if (ios /= 0) stop 'nadvibs: nadvibs.in not found'
    read(12, *); read(12, *) aomega(1:nmodes)
    do i = 1, nstblks
        read(12, *); read(12, *) concoef(i)
        do j = 1, ordr
            read(12, *); read(12, *) potterms(1:nterms(nmodes, j), j, i)
        end do
    end do
    allocate(bomega(nmodes), dvec(nmodes), tmat(nmodes * nmodes))
    if (shiftref /= 0) then
        read(12, *); read(12, *) bomega(1:nmodes)
        read(12, *); read(12, *) dvec(1:nmodes)
        read(12, *); read(12, *) tmat(1:nmodes * nmodes)
    end if
    close(12)

    ! ---- operator on the device ----
    prob%nstates = nstates; prob%nmodes = nmodes; prob%ordr = ordr
    prob%nfunc = c_loc(nfunc); prob%aomega = c_loc(aomega); prob%concoef = c_loc(concoef)
    prob%potterms = c_loc(potterms); prob%ztoler = ztoler
    prob%quirks = NV_QUIRKS_REF
    call get_environment_variable('NV_QUIRKS', env)
    if (trim(env) == 'exact') prob%quirks = NV_QUIRKS_EXACT
    prob%neworigin = merge(1, 0, shiftref /= 0)
    prob%statew = c_loc(statew); prob%bomega = c_loc(bomega); prob%dvec = c_loc(dvec); prob%tmat = c_loc(tmat)
    prob%nirreps = nirreps; prob%npirrep = c_loc(npirrep); prob%istate = c_loc(istate)
    call cpu_time(t0)
    call nv_check(nv_create(prob, 0_c_int32_t, ctx), c_null_ptr, 'nv_create')

    open(13, file='output.dat', status='replace')
    write(13, '(a)') ' NadVibS Lanczos path on MI355X (libnadvibs)'
    write(13, '(a,i0,a,i0,a,i0,a,i0)') '  nstates = ', nstates, '  nmodes = ', nmodes, '  ordr = ', ordr, &
        '  niter = ', niter
    allocate(counts(ordr * nstblks))
    call nv_check(nv_term_counts(ctx, counts), ctx, 'nv_term_counts')
    write(13, '(a)') ' ------- TERM COUNTING FOR VIBRONIC HAMILTONIAN --------'
    k = 0
    do i = 1, nstblks
        do j = 1, ordr
            write(13, '(a,i4,a,i4,a,i20)') '  TOT. NUM.', j, '-INDEX TERMS IN BLK', i, ':', counts((j - 1) * nstblks + i)
            k = k + int(counts((j - 1) * nstblks + i))
        end do
        write(13, '(a)') ''
    end do
    write(13, '(a,i20)') '  TOTAL NUMBER OF TERMS IN ALL BLOCKS:', k

    ! ---- restart (Source/secondary.f90:857-861, :872-892) ----
    if (restart /= 0) then
        open(14, file='restart.log', status='old', iostat=ios)
        if (ios /= 0) stop 'nadvibs: restart requested but restart.log not found'
        read(14, *); read(14, *) i
        read(14, *); read(14, *) nload
        iiter = nload + 1
        niter = niter + iiter - 1
        allocate(alpha(niter), beta(0:niter), omega(2, 0:niter))
        omega = 0d0
        read(14, *); read(14, *) alpha(1:nload)
        read(14, *); read(14, *) beta(0:nload)
        ! omega rows follow when the run re-orthogonalised (Source/top.f90:1032-1037)
        if (orthog) then
            read(14, *, iostat=ios); if (ios == 0) read(14, *, iostat=ios) omega(1, 0:nload)
            if (ios == 0) read(14, *, iostat=ios); if (ios == 0) read(14, *, iostat=ios) omega(2, 0:nload)
            if (ios /= 0) stop 'nadvibs: restart.log holds no omega rows for a re-orthogonalising run'
        end if
        close(14)
    else
        allocate(alpha(niter), beta(0:niter), omega(2, 0:niter))
        omega = 0d0
    end if

    ! ---- Lanczos (Source/main.f90:63-71) ----
    prec = NV_PREC_F64
    call get_environment_variable('NV_PREC', env)
    if (trim(env) == 'f32') prec = NV_PREC_F32
    opts%niter = niter; opts%reorthog = reorthog; opts%prec = prec; opts%chkorthog = chkorthog
    opts%seed = 20240101_c_int64_t; opts%noise_scale = 1d0; opts%pro_eps = 0d0
    ! savevecs (Source/secondary.f90:863): the root analysis reads every Lanczos vector
    opts%savevecs = 0
    opts%checkpoint = 0
    opts%pro_tau = 0d0; opts%pro_eta = 0d0
    if (idroots > 0 .or. soroots > 0) opts%savevecs = 1
    ! every Lanczos vector is kept: the archive nadvibs.34 (Source/secondary.f90:273-308) and,
    ! with re-orthogonalisation, its state (nadvibs.36) are written at the end and reloaded on restart
    keep = orthog .or. opts%savevecs /= 0
    proname = trim(outdir)//'nadvibs.36.0'//c_null_char
    call nv_check(nv_lanczos_init(ctx, opts), ctx, 'nv_lanczos_init')
    if (restart /= 0) then
        it = nload
        if (keep) then
            inquire(file=trim(outdir)//'nadvibs.36.0', exist=have_pro)
            if (orthog .and. have_pro) then
                call nv_check(nv_load_archive(ctx, trim(outdir)//'nadvibs.34.0'//c_null_char, c_loc(proname), it), &
                              ctx, 'nv_load_archive')
            else
                call nv_check(nv_load_archive(ctx, trim(outdir)//'nadvibs.34.0'//c_null_char, c_null_ptr, it), &
                              ctx, 'nv_load_archive')
            end if
        end if
        call nv_check(nv_load_vectors(ctx, trim(outdir)//'nadvibs.35.0'//c_null_char, it, alpha, beta, &
                                      c_loc(omega)), ctx, 'nv_load_vectors')
    end if
    call get_environment_variable('NV_LOOP', env)
    if (trim(env) == '1') then
        do it = iiter, niter
            call nv_check(nv_update_lanczos(ctx, it), ctx, 'update_lanczos')
            call nv_check(nv_hv_iter(ctx, it), ctx, 'Hv')
            call nv_check(nv_lanczos_step(ctx, it), ctx, 'lanczos_step')
            if (orthog .and. it > 1) then
                call nv_check(nv_set_omega(ctx, it), ctx, 'set_omega')
                call nv_check(nv_partial_orthog(ctx, it), ctx, 'partial_orthog')
            end if
        end do
    else
        call nv_check(nv_lanczos_run(ctx, int(iiter, c_int64_t), int(niter, c_int64_t)), ctx, 'nv_lanczos_run')
    end if
    if (orthog) then
        call nv_check(nv_get_coefficients(ctx, c_loc(alpha), c_loc(beta), c_loc(omega)), ctx, 'nv_get_coefficients')
    else
        call nv_check(nv_get_coefficients(ctx, c_loc(alpha), c_loc(beta), c_null_ptr), ctx, 'nv_get_coefficients')
    end if
    call cpu_time(t1)
    write(13, '(a)') '  - Lanczos Iterations Completed - '
    write(13, '(a,f14.5,a)') '  Time Required: ', t1 - t0, ' secs.'

    ! ---- make_restart (Source/top.f90:1007-1049) ----
    open(15, file='restart.log', status='replace')
    write(15, '(a)') ' TOTAL NUMBER OF ITERATIONS IN FILE'
    write(15, *) niter
    write(15, '(a)') ' ITERATION NUMBER TO RESTART FROM'
    write(15, *) niter
    write(15, '(a)') ' ALPHA COEFFICIENTS -- diagonal T matrix elements'
    write(15, *) alpha(1:niter)
    write(15, '(a)') ' BETA COEFFICIENTS -- off-diagonal T matrix elements'
    write(15, *) beta(0:niter)
    if (orthog) then
        write(15, '(a)') ' OMEGA(1) -- ORTHOGONALITY INFORMATION'
        write(15, *) omega(1, 0:niter)
        write(15, '(a)') ' OMEGA(2) -- ORTHOGONALITY INFORMATION'
        write(15, *) omega(2, 0:niter)
    end if
    close(15)
    call nv_check(nv_save_vectors(ctx, trim(outdir)//'nadvibs.35.0'//c_null_char), ctx, 'nv_save_vectors')
    if (orthog) then
        call nv_check(nv_save_archive(ctx, trim(outdir)//'nadvibs.34.0'//c_null_char, c_loc(proname)), &
                      ctx, 'nv_save_archive')
    else if (keep) then
        call nv_check(nv_save_archive(ctx, trim(outdir)//'nadvibs.34.0'//c_null_char, c_null_ptr), &
                      ctx, 'nv_save_archive')
    end if

    call compute_eigenvalues()
    close(13)
    if (idroots > 0 .or. soroots > 0) call identify_roots()
    call nv_check(nv_destroy(ctx), c_null_ptr, 'nv_destroy')
    write(*, *) 'NadVibs terminated.'

contains

    integer function nterms(n, p)
        integer, intent(in) :: n, p
        integer :: q
        nterms = 1
        do q = 1, p
            nterms = nterms * (n - 1 + q) / q
        end do
    end function nterms

    ! Eigenvalues and intensities of T, the kept roots and the table of
    ! Source/top.f90:1051-1112.
    subroutine compute_eigenvalues()
        real(c_double), allocatable :: ev(:), inten(:), bji(:)
        integer(c_int32_t), allocatable :: sgn(:)
        integer, allocatable :: keep(:)
        integer :: ngood, q
        real(8) :: bjiconv, zpve, imax
        allocate(ev(niter), inten(niter), bji(niter), sgn(niter), keep(niter))
        call nv_check(nv_ritz(ctx, int(niter, c_int64_t), alpha, beta, ev, inten, bji, sgn), ctx, 'nv_ritz')
        bjiconv = 10d0**(-bconv)
        ngood = 1
        keep(1) = 1
        do q = 2, niter
            if (bji(q) < bjiconv) then
                ngood = ngood + 1
                keep(ngood) = q
            end if
        end do
        zpve = 0.5d0 * sum(aomega)
        imax = 0d0
        do q = 1, ngood
            imax = max(imax, inten(keep(q)))
        end do
        write(13, '(a)') ''
        write(13, '(a)') 'COMPUTED EIGENVALUES '
        write(13, '(a)') '------------------------------'
        write(13, '(a)') ''
        write(13, '(a,f12.4,a)') '  ZPVE from nadvibs.in (reference state): ', zpve * AU2WAVE, ' cm-1'
        write(13, '(a,f12.4,a)') '  FIRST eigenvalue:                       ', ev(keep(1)) * AU2WAVE, ' cm-1'
        write(13, '(a,f12.8)') '  Intensity Factor (MAX Intensity):       ', imax
        write(13, '(a)') ''
        write(13, '(a)') '  Root  Index  Convergence     E(cm-1)        E(eV)     Intensity'
        do q = 1, ngood
            write(13, '(I6,I7,es13.4,f15.5,f12.5,f12.7)') q, keep(q), bji(keep(q)), &
                (ev(keep(q)) - ev(keep(1))) * AU2WAVE, (ev(keep(q)) - ev(1)) * AU2EV, inten(keep(q)) / imax
        end do
        write(13, '(a)') ''
        write(13, '(a)') '   All requested eigenvalues computed.'
    end subroutine compute_eigenvalues

    ! Ritz vectors of the lowest max(idroots, soroots) roots, rootinfo.dat and
    ! sodata.in (identify_roots, Source/top.f90:1122-1283).
    subroutine identify_roots()
        integer :: nroots, kk, q, j, k, l, cnt, m
        integer(c_int64_t) :: lab
        real(c_double), allocatable :: ev(:), inten(:), mag2(:, :), pct(:, :), ovl(:, :, :, :)
        integer(c_int32_t), allocatable :: sgn(:), st(:, :), ph(:, :)
        integer(c_int64_t), allocatable :: idx(:, :)
        integer, allocatable :: barray(:)
        real(8) :: dp(100)
        character(len=*), parameter :: fmt3 = &
            '(i5,1x,i5,2x,es14.6,1x,es14.6,1x,es14.6,/i5,1x,i5,2x,es14.6,1x,es14.6,1x,es14.6)'
        nroots = max(idroots, soroots)
        kk = max(idroots, 1)
        allocate(ev(nroots), inten(nroots), sgn(nroots), mag2(kk, nroots), idx(kk, nroots), st(kk, nroots), &
                 ph(kk, nroots), pct(nstates, nroots), ovl(nstates, nstates, nroots, nroots), barray(nmodes))
        call nv_check(nv_identify_roots(ctx, int(niter, c_int64_t), alpha, beta, int(nroots, c_int64_t), &
                      int(idroots, c_int64_t), ev, inten, sgn, idx, st, ph, mag2, pct, ovl, c_null_ptr), &
                      ctx, 'nv_identify_roots')
        open(33, file='rootinfo.dat', status='replace')
        write(33, '(a,i4,a)') ' Summary for computing ', nroots, ' eigenvectors corresponding to the lowest eigenvalues'
        do q = 1, nroots
            write(33, '(a,i4,a)') ' Eigenvector ', q, ' has been computed...'
            write(33, '(a)') ''
            write(33, '(a,i4,a)') ' ************ ROOT ', q, ' *************'
            write(33, '(a,f9.2,a,f9.5,a,f6.4/)') ' Energy: ', ev(q) * AU2WAVE, ' cm-1, ', ev(q) * AU2EV, &
                ' eV   Intensity: ', inten(q)
            do j = 1, idroots
                write(33, '(a,i2,a,i12,a,i2,a,f8.6)') ' State vector: ', st(j, q), ' | Index: ', idx(j, q), &
                    ' Phase: ', ph(j, q), ' | Magnitude^2: ', mag2(j, q)
                ! get_barray (Source/secondary.f90:812-824)
                lab = max(idx(j, q) - 1, 0_c_int64_t)
                do m = nmodes, 1, -1
                    barray(m) = int(mod(lab, int(basis(m), c_int64_t)))
                    lab = lab / basis(m)
                end do
                write(33, '(a,25(i3))') ' v(i): ', barray
            end do
            write(33, '(a)') ' '
            if (nstates == 2) write(33, '(a,f6.2,a,f6.2,a)') ' Percent State (1,2): (', pct(1, q), ' %, ', &
                pct(2, q), ' %)'
            if (nstates == 3) write(33, '(a,f6.2,a,f6.2,a,f6.2,a)') ' Percent State (1,2,3): (', pct(1, q), ' %, ', &
                pct(2, q), ' %, ', pct(3, q), ' %)'
            if (nstates == 4) write(33, '(a,f6.2,a,f6.2,a,f6.2,a,f6.2,a)') ' Percent State (1,2,3,4): (', &
                pct(1, q), ' %, ', pct(2, q), ' %, ', pct(3, q), ' %, ', pct(4, q), ' %)'
            if (nstates == 5) write(33, '(a,f6.2,a,f6.2,a,f6.2,a,f6.2,a,f6.2,a)') ' Percent State (1,2,3,4,5): (', &
                pct(1, q), ' %, ', pct(2, q), ' %, ', pct(3, q), ' %, ', pct(4, q), ' %, ', pct(5, q), ' %)'
            write(33, '(a)') ' --------------------------------------------------------------'
        end do
        write(33, '(a)') ' '
        write(33, '(a)') '------------ FINISHED EXECUTION ------------'
        close(33)
        if (soroots <= 0) return
        open(34, file='sodata.in', status='replace')
        write(34, '(i5)') soroots
        write(34, '(a)') ' Root   Energy (eV)      Sqrt(Int.)*sign'
        do q = 1, soroots
            write(34, '(i5,2x,es15.8,2x,es15.8)') q, ev(q) * AU2EV, sqrt(inten(q)) * sgn(q)
        end do
        if (nstates == 2) write(34, '(a)') ' R(i)  R(j)   [S1(i)S2(j)]   [S2(i)S1(j)]'
        if (nstates == 3) then
            write(34, '(a)') ' R(i)  R(j)   [S1(i)S2(j)]   [S1(i)S3(j)]  [S2(i)S1(j)]'
            write(34, '(a)') ' R(i)  R(j)   [S2(i)S3(j)]   [S3(i)S1(j)]  [S3(i)S2(j)]'
        end if
        do q = 1, soroots
            do j = 1, q - 1
                cnt = 0
                do k = 1, nstates
                    do l = 1, nstates
                        if (k /= l) then
                            cnt = cnt + 1
                            dp(cnt) = ovl(l, k, j, q)
                        end if
                    end do
                end do
                if (nstates == 2) write(34, '(i5,1x,i5,2x,es14.6,1x,es14.6)') q, j, dp(1), dp(2)
                if (nstates == 3) write(34, fmt3) q, j, dp(1), dp(2), dp(3), q, j, dp(4), dp(5), dp(6)
            end do
        end do
        close(34)
    end subroutine identify_roots

end program nadvibs_main
