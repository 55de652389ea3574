! Closed-shell PPP RHF of the host, the reference's algorithm
! (src/methods/mod_scf.f90): Hueckel guess :351-386, damped density step
! D -= step*Fmin with Fmin = [F, 2D-1](2D-1) :446-494 and :106-136, Fock
! F = core - G*D + 2 diag(G diag D) :417-442, energy :166-200; then the MO Fock
! matrix and reference energy prepared for coupled cluster
! (src/methods/coupledCluster/mod_cc.f90:2311-2504).
module hscf
    use hkinds,    only: rglu, iglu
    use hmolecule, only: hmol
    use linalg,    only: symeig
    implicit none
    private
    public :: runSCF, moFock

contains

    subroutine density(V, D, N, Nocc)
        real(kind=rglu),    intent(in)  :: V(:,:)
        real(kind=rglu),    intent(out) :: D(:,:)
        integer(kind=iglu), intent(in)  :: N, Nocc
        D(1:N,1:N) = matmul(V(1:N,1:Nocc), transpose(V(1:N,1:Nocc)))
    end subroutine density

    subroutine fockian(mol, D, F, N)
        type(hmol),         intent(in)  :: mol
        real(kind=rglu),    intent(in)  :: D(:,:)
        real(kind=rglu),    intent(out) :: F(:,:)
        integer(kind=iglu), intent(in)  :: N
        real(kind=rglu)    :: dd(N)
        integer(kind=iglu) :: mu
        do mu = 1, N
            dd(mu) = D(mu, mu)
        end do
        F(1:N,1:N) = mol%core - mol%G * D(1:N,1:N)
        do mu = 1, N
            F(mu, mu) = F(mu, mu) + 2 * dot_product(mol%G(mu, 1:N), dd)
        end do
    end subroutine fockian

    real(kind=rglu) function scfEnergy(mol, D, F, N)
        type(hmol),         intent(in) :: mol
        real(kind=rglu),    intent(in) :: D(:,:), F(:,:)
        integer(kind=iglu), intent(in) :: N
        integer(kind=iglu) :: mu
        real(kind=rglu) :: enuc
        enuc = 0
        do mu = 1, N
            enuc = enuc + (sum(mol%G(mu, :)) - mol%G(mu, mu)) / 2
        end do
        scfEnergy = sum(D(1:N,1:N) * (mol%core + F(1:N,1:N))) + enuc
    end function scfEnergy

    ! returns V (MO coefficients, ascending), E, status (1 converged, -1 cap) and iterations
    subroutine runSCF(mol, maxiter, accuracy, step, V, E, energy, status, iters)
        type(hmol),         intent(in)  :: mol
        integer(kind=iglu), intent(in)  :: maxiter
        real(kind=rglu),    intent(in)  :: accuracy, step
        real(kind=rglu),    intent(out) :: V(:,:), E(:), energy
        integer(kind=iglu), intent(out) :: status, iters
        real(kind=rglu), allocatable    :: D(:,:), F(:,:), R(:,:), Fmin(:,:)
        integer(kind=iglu) :: N, No, mu
        real(kind=rglu)    :: eps

        N = mol%nAtoms
        No = mol%nOcc
        allocate (D(N,N), F(N,N), R(N,N), Fmin(N,N))
        call symeig(mol%huckel, V, E, N)
        call density(V, D, N, No)
        call fockian(mol, D, F, N)
        status = -1
        iters = 0
        do
            iters = iters + 1
            R = 2 * D
            do mu = 1, N
                R(mu, mu) = R(mu, mu) - 1
            end do
            Fmin = matmul(matmul(F, R) - matmul(R, F), R)
            eps = maxval(abs(Fmin))
            D = D - step * Fmin
            call fockian(mol, D, F, N)
            call symeig(F, V, E, N)
            call density(V, D, N, No)
            call fockian(mol, D, F, N)
            if (eps < accuracy) then
                status = 1
                exit
            end if
            if (iters + 1 > maxiter) exit
        end do
        energy = scfEnergy(mol, D, F, N)
        deallocate (D, F, R, Fmin)
    end subroutine runSCF

    ! F_mo = V^T X V with X the Fock matrix of D = V_occ V_occ^T; eref its energy
    subroutine moFock(mol, V, Fmo, eref)
        type(hmol),      intent(in)  :: mol
        real(kind=rglu), intent(in)  :: V(:,:)
        real(kind=rglu), intent(out) :: Fmo(:,:), eref
        real(kind=rglu), allocatable :: D(:,:), X(:,:)
        integer(kind=iglu) :: N

        N = mol%nAtoms
        allocate (D(N,N), X(N,N))
        call density(V, D, N, mol%nOcc)
        call fockian(mol, D, X, N)
        eref = scfEnergy(mol, D, X, N)
        Fmo(1:N,1:N) = matmul(transpose(V(1:N,1:N)), matmul(X, V(1:N,1:N)))
        deallocate (D, X)
    end subroutine moFock

end module hscf
