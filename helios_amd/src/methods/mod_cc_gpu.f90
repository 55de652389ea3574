! Method module of the GPU r-ccsd: the reference's method contract
! setCCParameters / initCC / iterationCC / energyCC / finalizeCC
! (src/methods/coupledCluster/mod_cc.f90:59-60, 82-131, 489-603, 820-891,
! 1787-1952, 2119-2129) with the coupled-cluster arithmetic done by
! libhelios_ccsd through the iso_c_binding module ccsdGPU.  The host keeps the
! SCF, the MO Fock build and the iterator.
module coupledClusterGPU
    use iso_c_binding
    use hkinds,    only: rglu, iglu
    use hinput,    only: hcontrol
    use hmolecule, only: hmol
    use hscf,      only: runSCF, moFock
    use ccsdGPU
    implicit none
    private

    type(c_ptr)        :: ctx = c_null_ptr
    integer(kind=iglu) :: N, Nocc
    real(kind=rglu)    :: refeEnergy
    real(kind=rglu), allocatable :: hV(:,:), E(:), F(:,:)

    public :: setCCParameters, initCC, iterationCC, energyCC, finalizeCC, solveCC, getAmplitudes, energyCCT
    public :: refeEnergy, N, Nocc

contains

    subroutine setCCParameters(ctl, mol)
        type(hcontrol), intent(in) :: ctl
        type(hmol),     intent(in) :: mol
        integer(c_int) :: prec, proj

        N = mol%nAtoms
        Nocc = mol%nOcc
        prec = HCC_FP64
        if (ctl%precision == 'mixed') prec = HCC_MIXED
        call gpuCheck(hcc_create(ctx, 0_c_int, prec, int(ctl%diisSteps, c_int)), 'hcc_create', ctx)
        proj = 2
        if (ctl%projType == '1') proj = 1
        call gpuCheck(hcc_set_options(ctx, real(ctl%t1Step, c_double), real(ctl%t2Step, c_double), proj, &
                                      merge(1_c_int, 0_c_int, ctl%diisEnabled)), 'hcc_set_options', ctx)
        if (ctl%guess == 'learned') call loadGuessNetwork(ctl%guessNetwork)
    end subroutine setCCParameters

    ! parameters of the learned guess: HCC_MLP_PARAMS little-endian float32 values
    subroutine loadGuessNetwork(path)
        character(len=*), intent(in) :: path
        real(c_float) :: params(HCC_MLP_PARAMS)
        integer :: u, ios

        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='old', action='read', &
              iostat=ios)
        if (ios /= 0) then
            write (*, '(A,A)') ' Error: cannot open the guess network ', trim(path)
            error stop 'guess network missing'
        end if
        read (u, iostat=ios) params
        close (u)
        if (ios /= 0) then
            write (*, '(A,A)') ' Error: short guess network file ', trim(path)
            error stop 'guess network unreadable'
        end if
        call gpuCheck(hcc_set_guess_network(ctx, params, HCC_MLP_PARAMS), 'hcc_set_guess_network', ctx)
    end subroutine loadGuessNetwork

    integer(c_int) function guessKind(ctl)
        type(hcontrol), intent(in) :: ctl
        guessKind = HCC_GUESS_REFERENCE
        if (ctl%guess == 'mp2') guessKind = HCC_GUESS_MP2
        if (ctl%guess == 'learned') guessKind = HCC_GUESS_LEARNED
    end function guessKind

    ! SCF, MO Fock and reference energy on the host, integrals and guess on the device
    subroutine initCC(ctl, mol)
        type(hcontrol), intent(in) :: ctl
        type(hmol),     intent(in) :: mol
        real(kind=rglu)    :: escf
        integer(kind=iglu) :: st, it
        integer(c_int)     :: kind

        if (allocated(hV)) deallocate (hV, E, F)
        allocate (hV(N,N), E(N), F(N,N))
        call runSCF(mol, ctl%scfMaxIters, ctl%scfAccuracy, ctl%scfStep, hV, E, escf, st, it)
        write (*, '(A,I0,A,I0,A,F24.15,A)') ' RHF: status ', st, ' after ', it, ' iterations, energy ', escf, ' eV'
        call moFock(mol, hV, F, refeEnergy)
        call gpuCheck(hcc_set_molecule(ctx, int(N, c_int), int(Nocc, c_int), hV, F, mol%G, refeEnergy), &
                      'hcc_set_molecule', ctx)
        kind = guessKind(ctl)
        call gpuCheck(hcc_guess(ctx, kind), 'hcc_guess', ctx)
    end subroutine initCC

    subroutine iterationCC(iteration, epsilon, accuracy)
        integer(kind=iglu), intent(in)  :: iteration
        real(kind=rglu),    intent(in)  :: epsilon
        real(kind=rglu),    intent(out) :: accuracy(5)
        call gpuCheck(hcc_iterate(ctx, int(iteration, c_int), real(epsilon, c_double), accuracy), 'hcc_iterate', ctx)
    end subroutine iterationCC

    subroutine energyCC(energy)
        real(kind=rglu), intent(out) :: energy(5)
        call gpuCheck(hcc_energy(ctx, energy), 'hcc_energy', ctx)
    end subroutine energyCC

    ! r-ccsd(t): the reference routes it to the spin-orbital CCSD and its (T)
    ! loop (mod_cc.f90:109-118, 2028-2111); for the closed-shell RHF reference
    ! the same energies follow from the converged closed-shell amplitudes.
    ! energy: CCSD + c1 + c2 + c3, reference, singles, doubles, c1 (the
    ! reference's layout and signs); et: E[T], E[ST]
    subroutine energyCCT(energy, et)
        real(kind=rglu), intent(out) :: energy(5), et(2)
        call gpuCheck(hcc_batch_triples(ctx, energy, et), 'hcc_batch_triples', ctx)
    end subroutine energyCCT

    ! whole solve on the device (iterator rules applied there)
    subroutine solveCC(ctl, energy, iterations, status)
        type(hcontrol),     intent(in)  :: ctl
        real(kind=rglu),    intent(out) :: energy(5)
        integer(kind=iglu), intent(out) :: iterations, status
        integer(c_int) :: it, st, kind
        kind = guessKind(ctl)
        call gpuCheck(hcc_solve(ctx, kind, int(ctl%ccMaxIters, c_int), real(ctl%ccAccuracy, c_double), energy, &
                                it, st), 'hcc_solve', ctx)
        iterations = it
        status = st
    end subroutine solveCC

    subroutine getAmplitudes(t1, t2)
        real(kind=rglu), intent(out) :: t1(:,:), t2(:,:,:,:)
        call gpuCheck(hcc_get_amplitudes(ctx, t1, t2), 'hcc_get_amplitudes', ctx)
    end subroutine getAmplitudes

    subroutine finalizeCC
        if (c_associated(ctx)) call hcc_destroy(ctx)
        ctx = c_null_ptr
        if (allocated(hV)) deallocate (hV, E, F)
    end subroutine finalizeCC

end module coupledClusterGPU
