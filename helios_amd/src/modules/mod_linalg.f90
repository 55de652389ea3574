! Symmetric eigensolver of the host: Householder reduction to tridiagonal form
! followed by the implicit QL iteration, eigenvalues sorted ascending with
! their vectors -- the Wilkinson-Reinsch scheme the reference's tred4 uses
! (src/modules/mod_math.f90:31-230).  Written from the published algorithm.
module linalg
    use hkinds, only: rglu, iglu
    implicit none
    private
    public :: symeig

contains

    ! A (n x n, symmetric, untouched) -> Z (eigenvectors in columns), D (ascending)
    subroutine symeig(A, Z, D, n)
        integer(kind=iglu), intent(in)  :: n
        real(kind=rglu),    intent(in)  :: A(:,:)
        real(kind=rglu),    intent(out) :: Z(:,:), D(:)
        real(kind=rglu), allocatable    :: E(:)
        integer(kind=iglu)              :: i, j, k

        allocate (E(n))
        Z(1:n,1:n) = A(1:n,1:n)
        call householder(Z, D, E, n)
        call ql_implicit(Z, D, E, n)
        ! ascending selection sort with column swaps
        do i = 1, n - 1
            k = i
            do j = i + 1, n
                if (D(j) < D(k)) k = j
            end do
            if (k /= i) then
                call swap_cols(Z, D, i, k, n)
            end if
        end do
        deallocate (E)
    end subroutine symeig

    subroutine swap_cols(Z, D, i, k, n)
        real(kind=rglu),    intent(inout) :: Z(:,:), D(:)
        integer(kind=iglu), intent(in)    :: i, k, n
        real(kind=rglu) :: t
        integer(kind=iglu) :: r
        t = D(i); D(i) = D(k); D(k) = t
        do r = 1, n
            t = Z(r,i); Z(r,i) = Z(r,k); Z(r,k) = t
        end do
    end subroutine swap_cols

    ! Householder tridiagonalisation; on exit Z holds the accumulated
    ! orthogonal transformation, D the diagonal, E(2:n) the sub-diagonal.
    subroutine householder(Z, D, E, n)
        real(kind=rglu),    intent(inout) :: Z(:,:)
        real(kind=rglu),    intent(out)   :: D(:), E(:)
        integer(kind=iglu), intent(in)    :: n
        integer(kind=iglu) :: i, j, k, l
        real(kind=rglu)    :: f, g, h, hh, scale

        do i = n, 2, -1
            l = i - 1
            h = 0
            if (l > 1) then
                scale = sum(abs(Z(i,1:l)))
                if (scale == 0) then
                    E(i) = Z(i,l)
                else
                    Z(i,1:l) = Z(i,1:l) / scale
                    h = sum(Z(i,1:l)**2)
                    f = Z(i,l)
                    g = -sign(sqrt(h), f)
                    E(i) = scale * g
                    h = h - f * g
                    Z(i,l) = f - g
                    f = 0
                    do j = 1, l
                        Z(j,i) = Z(i,j) / h
                        g = 0
                        do k = 1, j
                            g = g + Z(j,k) * Z(i,k)
                        end do
                        do k = j + 1, l
                            g = g + Z(k,j) * Z(i,k)
                        end do
                        E(j) = g / h
                        f = f + E(j) * Z(i,j)
                    end do
                    hh = f / (h + h)
                    do j = 1, l
                        f = Z(i,j)
                        g = E(j) - hh * f
                        E(j) = g
                        do k = 1, j
                            Z(j,k) = Z(j,k) - (f * E(k) + g * Z(i,k))
                        end do
                    end do
                end if
            else
                E(i) = Z(i,l)
            end if
            D(i) = h
        end do
        D(1) = 0
        E(1) = 0
        do i = 1, n
            l = i - 1
            if (D(i) /= 0) then
                do j = 1, l
                    g = 0
                    do k = 1, l
                        g = g + Z(i,k) * Z(k,j)
                    end do
                    do k = 1, l
                        Z(k,j) = Z(k,j) - g * Z(k,i)
                    end do
                end do
            end if
            D(i) = Z(i,i)
            Z(i,i) = 1
            do j = 1, l
                Z(j,i) = 0
                Z(i,j) = 0
            end do
        end do
    end subroutine householder

    ! QL iteration with implicit Wilkinson shifts on the tridiagonal (D, E),
    ! accumulating the rotations into Z.
    subroutine ql_implicit(Z, D, E, n)
        real(kind=rglu),    intent(inout) :: Z(:,:), D(:), E(:)
        integer(kind=iglu), intent(in)    :: n
        integer(kind=iglu) :: i, k, l, m, iter
        real(kind=rglu)    :: b, c, dd, f, g, p, r, s
        logical            :: underflow

        do i = 2, n
            E(i-1) = E(i)
        end do
        E(n) = 0
        do l = 1, n
            iter = 0
            do
                m = n
                do k = l, n - 1
                    dd = abs(D(k)) + abs(D(k+1))
                    if (abs(E(k)) <= epsilon(dd) * dd) then
                        m = k
                        exit
                    end if
                end do
                if (m == l) exit
                iter = iter + 1
                if (iter > 30 * n) exit
                g = (D(l+1) - D(l)) / (2 * E(l))
                r = hypot(g, 1.0_rglu)
                g = D(m) - D(l) + E(l) / (g + sign(r, g))
                s = 1
                c = 1
                p = 0
                underflow = .false.
                do i = m - 1, l, -1
                    f = s * E(i)
                    b = c * E(i)
                    r = hypot(f, g)
                    E(i+1) = r
                    if (r == 0) then
                        D(i+1) = D(i+1) - p
                        E(m) = 0
                        underflow = .true.
                        exit
                    end if
                    s = f / r
                    c = g / r
                    g = D(i+1) - p
                    r = (D(i) - g) * s + 2 * c * b
                    p = s * r
                    D(i+1) = g + p
                    g = c * r - b
                    do k = 1, n
                        f = Z(k,i+1)
                        Z(k,i+1) = s * Z(k,i) + c * f
                        Z(k,i) = c * Z(k,i) - s * f
                    end do
                end do
                if (underflow) cycle
                D(l) = D(l) - p
                E(l) = g
                E(m) = 0
            end do
        end do
    end subroutine ql_implicit

end module linalg
