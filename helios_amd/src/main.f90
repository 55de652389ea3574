! helios_gpu: host program for the GPU r-ccsd path.
!
!   helios_gpu -f input.inp          r-ccsd (or r-ccsd(t)) on the GPU
!   helios_gpu -f input.inp -scf     RHF only (no device needed)
!
! Reads the reference's input blocks, builds the PPP model, runs the RHF and
! the coupled-cluster solve through libhelios_ccsd, and prints the result in
! the reference's "Resulting energy" form (src/calculation/gIterator.f90:271-273).
! Like the reference's getEnergy (src/modules/mod_property.f90:159-177) the
! method runs setCCParameters -> initCC -> iterator(iterationCC, energyCC) ->
! energyCC -> finalizeCC.
program helios_gpu
    use hkinds,            only: rglu, iglu
    use hinput,            only: hcontrol, readInput
    use hmolecule,         only: hmol, readMolecule, buildModel
    use hiterator,         only: iterator
    use coupledClusterGPU
    use hscf,              only: runSCF, moFock
    implicit none

    type(hcontrol)     :: ctl
    type(hmol)         :: mol
    character(len=512) :: arg, fname
    integer(kind=iglu) :: k, status, iterations, m
    real(kind=rglu)    :: energy(5), escf, eref, energyT(5), et(2)
    logical            :: triples
    real(kind=rglu), allocatable :: V(:,:), E(:), Fmo(:,:)
    logical :: scfOnly

    fname = ''
    scfOnly = .false.
    k = 1
    do while (k <= command_argument_count())
        call get_command_argument(k, arg)
        if (trim(arg) == '-f') then
            call get_command_argument(k + 1, fname)
            k = k + 1
        end if
        if (trim(arg) == '-scf') scfOnly = .true.
        k = k + 1
    end do
    if (len_trim(fname) == 0) then
        write (*, '(A)') ' usage: helios_gpu -f input.inp'
        stop 1
    end if

    call readInput(trim(fname), ctl)
    call readMolecule(ctl%molLines, ctl%nMolLines, mol)
    call buildModel(mol, trim(ctl%coulomb))
    write (*, '(A,A,A,I0,A,I0,A,I0)') ' Molecule: ', trim(mol%name), '  atoms ', mol%nAtoms, &
        '  bonds ', mol%nBonds, '  electrons ', mol%nEls

    if (scfOnly) then
        allocate (V(mol%nAtoms, mol%nAtoms), E(mol%nAtoms), Fmo(mol%nAtoms, mol%nAtoms))
        call runSCF(mol, ctl%scfMaxIters, ctl%scfAccuracy, ctl%scfStep, V, E, escf, status, iterations)
        call moFock(mol, V, Fmo, eref)
        write (*, '(A,I0,A,I0)') ' rhf status ', status, '  iterations ', iterations
        write (*, '(A,F24.15,A)') ' Resulting energy ', escf, ' eV'
        write (*, '(A)') ' Orbital energies (eV):'
        write (*, '(4F20.12)') E
        stop
    end if

    do m = 1, max(ctl%nmethods, 1)
        triples = .false.
        if (ctl%nmethods > 0) then
            if (trim(ctl%methods(m)) /= 'r-ccsd' .and. trim(ctl%methods(m)) /= 'r-ccsd(t)') then
                write (*, '(A,A,A)') ' method ', trim(ctl%methods(m)), ' is not provided by helios_gpu; skipped'
                cycle
            end if
            triples = trim(ctl%methods(m)) == 'r-ccsd(t)'
        end if
        call setCCParameters(ctl, mol)
        call initCC(ctl, mol)
        if (ctl%driver == 'device') then
            call solveCC(ctl, energy, iterations, status)
        else
            call iterator(iterationCC, energyCC, ctl%ccMaxIters, ctl%ccAccuracy, ctl%divergence, &
                          ctl%stagnationAfter, ctl%stagnationPercent, status, iterations, energy, .true.)
            call energyCC(energy)
        end if
        write (*, '(/A,I0,A,I0)') ' r-ccsd status ', status, '  iterations ', iterations
        if (triples .and. status == 1) then
            call energyCCT(energyT, et)
            write (*, '(A,F24.15,A)') ' CCSD energy ', energy(1), ' eV'
            write (*, '(A,2F24.15,A)') ' (T) E[T], E[ST] ', et(1), et(2), ' eV'
            energy = energyT
        end if
        write (*, '(A,F24.15,A)') ' Resulting energy ', energy(1), ' eV'
        write (*, '(A,F24.15,A)') ' Reference energy ', energy(2), ' eV'
        write (*, '(A,F24.15,A)') ' Correlation energy ', energy(3) + energy(4), ' eV'
        call finalizeCC
    end do
end program helios_gpu
