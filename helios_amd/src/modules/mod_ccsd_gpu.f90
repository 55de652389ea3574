! iso_c_binding interface to libhelios_ccsd (include/helios_ccsd.h).
!
! Every hcc_* entry point is bound here with the argument meaning of the C
! header; gpuCheck turns a negative return code into the host's fatal-error
! path (message from hcc_last_error), as the reference does with ierror +
! primaryInformation('error') (src/general/primaryInformation.f90:77-88).
module ccsdGPU
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: HCC_FP64 = 0, HCC_MIXED = 1
    integer(c_int), parameter :: HCC_GUESS_REFERENCE = 0, HCC_GUESS_MP2 = 1, HCC_GUESS_LEARNED = 2
    integer(c_int), parameter :: HCC_MLP_PARAMS = 5313

    interface
        integer(c_int) function hcc_create(ctx, device, precision, diis_steps) bind(C, name='hcc_create')
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: ctx
            integer(c_int), value :: device, precision, diis_steps
        end function hcc_create

        integer(c_int) function hcc_set_options(ctx, t1_step, t2_step, projection_type, diis_enabled) &
                bind(C, name='hcc_set_options')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: t1_step, t2_step
            integer(c_int), value :: projection_type, diis_enabled
        end function hcc_set_options

        integer(c_int) function hcc_set_mixed(ctx, bf16, kappa, max_inner, max_outer) &
                bind(C, name='hcc_set_mixed')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: bf16, max_inner, max_outer
            real(c_double), value :: kappa
        end function hcc_set_mixed

        integer(c_int) function hcc_set_guess_network(ctx, params, count) bind(C, name='hcc_set_guess_network')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: params(*)
            integer(c_int), value :: count
        end function hcc_set_guess_network

        integer(c_int) function hcc_set_molecule(ctx, n, n_occ, C, F_mo, gamma, e_ref) &
                bind(C, name='hcc_set_molecule')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: n, n_occ
            real(c_double), intent(in) :: C(*), F_mo(*), gamma(*)
            real(c_double), value :: e_ref
        end function hcc_set_molecule

        integer(c_int) function hcc_guess(ctx, kind) bind(C, name='hcc_guess')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: kind
        end function hcc_guess

        integer(c_int) function hcc_iterate(ctx, iteration, eps, accuracy) bind(C, name='hcc_iterate')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: iteration
            real(c_double), value :: eps
            real(c_double), intent(out) :: accuracy(5)
        end function hcc_iterate

        integer(c_int) function hcc_energy(ctx, energy) bind(C, name='hcc_energy')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: energy(5)
        end function hcc_energy

        integer(c_int) function hcc_get_amplitudes(ctx, t1, t2) bind(C, name='hcc_get_amplitudes')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: t1(*), t2(*)
        end function hcc_get_amplitudes

        integer(c_int) function hcc_solve(ctx, guess, maxiter, eps, energy, iterations, status) &
                bind(C, name='hcc_solve')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: guess, maxiter
            real(c_double), value :: eps
            real(c_double), intent(out) :: energy(5)
            integer(c_int), intent(out) :: iterations, status
        end function hcc_solve

        ! (T) of the converged amplitudes: energy(5) as energyCC('spin-r-ccsd(t)')
        ! (mod_cc.f90:2028-2111), e_t = standard E[T], E[ST]
        integer(c_int) function hcc_batch_triples(ctx, energy, e_t) bind(C, name='hcc_batch_triples')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: energy(5), e_t(2)
        end function hcc_batch_triples

        subroutine hcc_destroy(ctx) bind(C, name='hcc_destroy')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine hcc_destroy

        type(c_ptr) function hcc_last_error(ctx) bind(C, name='hcc_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function hcc_last_error
    end interface

contains

    function gpuErrorText(ctx) result(text)
        type(c_ptr), intent(in) :: ctx
        character(len=512) :: text
        character(kind=c_char), pointer :: chars(:)
        type(c_ptr) :: p
        integer :: k

        text = ''
        p = hcc_last_error(ctx)
        if (.not. c_associated(p)) return
        call c_f_pointer(p, chars, [512])
        do k = 1, 512
            if (chars(k) == c_null_char) exit
            text(k:k) = chars(k)
        end do
    end function gpuErrorText

    subroutine gpuCheck(rc, what, ctx)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        type(c_ptr), intent(in) :: ctx

        if (rc == 0) return
        write (*, '(A,A,A,I0,A,A)') ' Error: ', what, ' failed (', rc, '): ', trim(gpuErrorText(ctx))
        error stop 'libhelios_ccsd error'
    end subroutine gpuCheck

end module ccsdGPU
