! Fixed-point driver of the host with the reference iterator's rules
! (src/calculation/gIterator.f90:88-210): iterate callsub + energysub, stop on
! max(accuracy) < eps (converged, 1), accuracy > divergence threshold
! (diverged, -2), iteration cap (-1), or -- after `stagnationAfter`
! iterations -- a share of improving iterations at or below
! `stagnationPercent` (stagnated, -3).  Statuses: src/modules/mod_helios.f90:119-120.
module hiterator
    use hkinds, only: rglu, iglu
    implicit none
    private
    public :: iterator

    abstract interface
        subroutine callsub_t(iteration, epsilon, accuracy)
            import :: rglu, iglu
            integer(kind=iglu), intent(in)  :: iteration
            real(kind=rglu),    intent(in)  :: epsilon
            real(kind=rglu),    intent(out) :: accuracy(5)
        end subroutine callsub_t
        subroutine energysub_t(energy)
            import :: rglu
            real(kind=rglu), intent(out) :: energy(5)
        end subroutine energysub_t
    end interface

contains

    subroutine iterator(callsub, energysub, maxiter, eps, divergence, stagnationAfter, stagnationPercent, &
                        status, iterations, energy, verbose)
        procedure(callsub_t)               :: callsub
        procedure(energysub_t)             :: energysub
        integer(kind=iglu), intent(in)     :: maxiter, stagnationAfter
        real(kind=rglu),    intent(in)     :: eps, divergence, stagnationPercent
        integer(kind=iglu), intent(out)    :: status, iterations
        real(kind=rglu),    intent(out)    :: energy(5)
        logical,            intent(in)     :: verbose
        real(kind=rglu)    :: accuracy(5), cur, prev
        integer(kind=iglu) :: successful

        successful = 1
        cur = 0
        prev = 10
        iterations = 0
        if (verbose) write (*, '(/A)') '   iter      accuracy                 energy (eV)'
        do
            iterations = iterations + 1
            call callsub(iterations, eps, accuracy)
            call energysub(energy)
            prev = cur
            cur = maxval(accuracy)
            if (cur < prev) successful = successful + 1
            if (verbose) write (*, '(I7,2X,ES14.6,2X,F24.15)') iterations, cur, energy(1)
            if (cur < eps) then
                status = 1
                exit
            end if
            if (cur > divergence) then
                status = -2
                exit
            end if
            if (iterations + 1 > maxiter) then
                status = -1
                exit
            end if
            if (iterations > stagnationAfter) then
                if (100.0_rglu * successful / iterations <= stagnationPercent) then
                    status = -3
                    exit
                end if
            end if
        end do
    end subroutine iterator

end module hiterator
