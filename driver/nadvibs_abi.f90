! Fortran interfaces of the C ABI (include/nadvibs.h) through iso_c_binding.
module nadvibs_abi
    use iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: NV_OK = 0
    integer(c_int32_t), parameter :: NV_QUIRKS_REF = 3, NV_QUIRKS_EXACT = 0
    integer(c_int32_t), parameter :: NV_PREC_F64 = 0, NV_PREC_F32 = 1

    type, bind(C) :: nv_problem
        integer(c_int64_t) :: nstates, nmodes, ordr
        type(c_ptr) :: nfunc, aomega, concoef, potterms
        real(c_double) :: ztoler
        integer(c_int32_t) :: quirks, neworigin
        type(c_ptr) :: statew, bomega, dvec, tmat
        integer(c_int64_t) :: nirreps
        type(c_ptr) :: npirrep, istate
    end type nv_problem

    type, bind(C) :: nv_lanczos_opts
        integer(c_int64_t) :: niter
        integer(c_int32_t) :: reorthog, prec
        integer(c_int64_t) :: chkorthog
        integer(c_int64_t) :: seed
        real(c_double) :: noise_scale, pro_eps
        integer(c_int32_t) :: savevecs
        integer(c_int64_t) :: checkpoint
        real(c_double) :: pro_tau
        real(c_double) :: pro_eta
    end type nv_lanczos_opts

    interface
        integer(c_int32_t) function nv_create(prob, device, ctx) bind(C, name="nv_create")
            import :: nv_problem, c_int32_t, c_ptr
            type(nv_problem), intent(in) :: prob
            integer(c_int32_t), value :: device
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int32_t) function nv_destroy(ctx) bind(C, name="nv_destroy")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function nv_last_error(ctx) bind(C, name="nv_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function nv_term_counts(ctx, out) bind(C, name="nv_term_counts")
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: out(*)
        end function
        integer(c_int32_t) function nv_term_summary(ctx, out) bind(C, name="nv_term_summary")
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: out(*)
        end function
        integer(c_int32_t) function nv_lanczos_init(ctx, opts) bind(C, name="nv_lanczos_init")
            import :: c_int32_t, c_ptr, nv_lanczos_opts
            type(c_ptr), value :: ctx
            type(nv_lanczos_opts), intent(in) :: opts
        end function
        integer(c_int32_t) function nv_update_lanczos(ctx, iter) bind(C, name="nv_update_lanczos")
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: iter
        end function
        integer(c_int32_t) function nv_hv_iter(ctx, iter) bind(C, name="nv_hv_iter")
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: iter
        end function
        integer(c_int32_t) function nv_lanczos_step(ctx, iter) bind(C, name="nv_lanczos_step")
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: iter
        end function
        integer(c_int32_t) function nv_set_omega(ctx, iter) bind(C, name="nv_set_omega")
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: iter
        end function
        integer(c_int32_t) function nv_partial_orthog(ctx, iter) bind(C, name="nv_partial_orthog")
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: iter
        end function
        integer(c_int32_t) function nv_lanczos_run(ctx, iiter, niter) bind(C, name="nv_lanczos_run")
            import :: c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: iiter, niter
        end function
        integer(c_int32_t) function nv_get_coefficients(ctx, alpha, beta, omega) bind(C, name="nv_get_coefficients")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: ctx, alpha, beta, omega
        end function
        integer(c_int32_t) function nv_ritz(ctx, n, alpha, beta, eval, inten, bji, sgn) bind(C, name="nv_ritz")
            import :: c_int32_t, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: alpha(*), beta(*)
            real(c_double), intent(out) :: eval(*), inten(*), bji(*)
            integer(c_int32_t), intent(out) :: sgn(*)
        end function
        integer(c_int32_t) function nv_identify_roots(ctx, n, alpha, beta, nroots, idroots, eval, inten, sgn, &
                top_index, top_state, top_phase, top_mag2, percent, overlap, vectors) bind(C, name="nv_identify_roots")
            import :: c_int32_t, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n, nroots, idroots
            real(c_double), intent(in) :: alpha(*), beta(*)
            real(c_double), intent(out) :: eval(*), inten(*), top_mag2(*), percent(*), overlap(*)
            integer(c_int32_t), intent(out) :: sgn(*), top_state(*), top_phase(*)
            integer(c_int64_t), intent(out) :: top_index(*)
            type(c_ptr), value :: vectors
        end function
        integer(c_int32_t) function nv_save_vectors(ctx, path) bind(C, name="nv_save_vectors")
            import :: c_int32_t, c_ptr, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int32_t) function nv_load_vectors(ctx, path, iter, alpha, beta, omega) bind(C, name="nv_load_vectors")
            import :: c_int32_t, c_ptr, c_char, c_int64_t, c_double
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int64_t), value :: iter
            real(c_double), intent(in) :: alpha(*), beta(*)
            type(c_ptr), value :: omega
        end function
        ! pro_path: c_loc of a NUL-terminated name, or c_null_ptr
        integer(c_int32_t) function nv_save_archive(ctx, path, pro_path) bind(C, name="nv_save_archive")
            import :: c_int32_t, c_ptr, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: path(*)
            type(c_ptr), value :: pro_path
        end function
        integer(c_int32_t) function nv_load_archive(ctx, path, pro_path, iter) bind(C, name="nv_load_archive")
            import :: c_int32_t, c_ptr, c_char, c_int64_t
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: path(*)
            type(c_ptr), value :: pro_path
            integer(c_int64_t), value :: iter
        end function
    end interface

contains

    ! Text of the last error (of ctx, or of nv_create when ctx is null).
    function nv_error_text(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=512) :: msg
        character(kind=c_char), pointer :: p(:)
        type(c_ptr) :: cp
        integer :: i
        msg = ' '
        cp = nv_last_error(ctx)
        if (.not. c_associated(cp)) return
        call c_f_pointer(cp, p, [512])
        do i = 1, 512
            if (p(i) == c_null_char) exit
            msg(i:i) = p(i)
        end do
    end function nv_error_text

    subroutine nv_check(rc, ctx, where)
        integer(c_int32_t), intent(in) :: rc
        type(c_ptr), intent(in) :: ctx
        character(len=*), intent(in) :: where
        if (rc /= NV_OK) then
            write(*, '(a,a,a,i0,a,a)') ' nadvibs: ', where, ' failed (status ', rc, '): ', trim(nv_error_text(ctx))
            stop 1
        end if
    end subroutine nv_check

end module nadvibs_abi
