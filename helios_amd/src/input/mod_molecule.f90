! Molecule block and PPP model of the host, following the reference's
! readMoleculeInformation (src/input/readMoleculeInformation.f90):
! header (name, atoms, bonds, alpha and beta electrons) :20-24, atoms :43-61,
! bonds :87-106, gamma :216-264 with k = 1/(a0*Eh), connectivity :108-117,
! core Hamiltonian H_kk = alpha_k - sum_{l/=k} n_l gamma_kl :140-159.
module hmolecule
    use hkinds, only: rglu, iglu, BohrRadius, HartreeEnergy
    implicit none
    private

    type, public :: hmol
        character(len=128) :: name = ''
        integer(kind=iglu) :: nAtoms = 0, nBonds = 0, nEls = 0, nOcc = 0
        real(kind=rglu), allocatable :: coords(:,:), alpha(:), gam(:), atomEls(:)
        integer(kind=iglu), allocatable :: bondAtoms(:,:)
        real(kind=rglu), allocatable :: beta(:)
        real(kind=rglu), allocatable :: G(:,:), core(:,:), huckel(:,:)
    end type hmol

    public :: readMolecule, buildModel

contains

    subroutine readMolecule(lines, nlines, mol)
        character(len=*),   intent(in)  :: lines(:)
        integer(kind=iglu), intent(in)  :: nlines
        type(hmol),         intent(out) :: mol
        integer(kind=iglu) :: k, na, nb, ios
        character(len=8)   :: sym, kind
        real(kind=rglu)    :: x(3), ip, ci, el

        if (nlines < 5) error stop 'Molecule block. FileFormatError. Incorrect header.'
        mol%name = trim(lines(1))
        read (lines(2), *, iostat=ios) mol%nAtoms
        read (lines(3), *, iostat=ios) mol%nBonds
        read (lines(4), *, iostat=ios) na
        read (lines(5), *, iostat=ios) nb
        if (ios /= 0) error stop 'Molecule block. FileFormatError. Incorrect header.'
        mol%nEls = na + nb
        mol%nOcc = mol%nEls / 2
        if (nlines < 5 + mol%nAtoms + mol%nBonds) error stop 'Molecule block. FileFormatError. Incomplete block.'
        allocate (mol%coords(3, mol%nAtoms), mol%alpha(mol%nAtoms), mol%gam(mol%nAtoms), mol%atomEls(mol%nAtoms))
        allocate (mol%bondAtoms(2, mol%nBonds), mol%beta(mol%nBonds))
        do k = 1, mol%nAtoms
            read (lines(5 + k), *, iostat=ios) sym, x, ip, ci, el
            if (ios /= 0) error stop 'Molecule block. FileFormatError. Incorrect atoms information.'
            mol%coords(:, k) = x
            mol%alpha(k) = ip
            mol%gam(k) = ci
            mol%atomEls(k) = el
        end do
        do k = 1, mol%nBonds
            read (lines(5 + mol%nAtoms + k), *, iostat=ios) mol%bondAtoms(1, k), mol%bondAtoms(2, k), mol%beta(k), kind
            if (ios /= 0) error stop 'Molecule block. FileFormatError. Incorrect bonds information.'
        end do
    end subroutine readMolecule

    subroutine buildModel(mol, coulomb)
        type(hmol),       intent(inout) :: mol
        character(len=*), intent(in)    :: coulomb
        integer(kind=iglu) :: N, i, j, k
        real(kind=rglu)    :: r, med, trConst

        N = mol%nAtoms
        trConst = 1 / (BohrRadius * HartreeEnergy)
        allocate (mol%G(N, N), mol%core(N, N), mol%huckel(N, N))
        mol%G = 0
        do i = 1, N
            do j = i, N
                r = sqrt(sum((mol%coords(:, i) - mol%coords(:, j))**2))
                med = (mol%gam(i) + mol%gam(j)) / 2
                select case (coulomb)
                    case ('ohno-klopman')
                        mol%G(i, j) = med / sqrt(1 + (trConst * med * r)**2)
                    case ('mataga-nishimoto')
                        mol%G(i, j) = med / (1 + trConst * med * r)
                    case ('hubbard')
                        if (i == j) mol%G(i, j) = mol%gam(i)
                    case default
                        error stop 'unknown coulomb-integrals'
                end select
                mol%G(j, i) = mol%G(i, j)
            end do
        end do
        mol%huckel = 0
        do i = 1, N
            mol%huckel(i, i) = mol%alpha(i)
        end do
        do k = 1, mol%nBonds
            i = mol%bondAtoms(1, k)
            j = mol%bondAtoms(2, k)
            mol%huckel(i, j) = mol%beta(k)
            mol%huckel(j, i) = mol%beta(k)
        end do
        mol%core = mol%huckel
        do k = 1, N
            r = 0
            do j = 1, N
                if (j /= k) r = r + mol%atomEls(j) * mol%G(k, j)
            end do
            mol%core(k, k) = mol%core(k, k) - r
        end do
    end subroutine buildModel

end module hmolecule
