! Input reader of the host: the reference's block syntax
! `%name{key=value, key=value}` with `#` comments (src/input/definebd.f90:14-15)
! for the blocks this path reads, with the reference's defaults
! (src/input/definebd.f90:27-157): general, parametrization, scf,
! coupled-cluster, iteration; and the free-format molecule block
! (layout src/input/definebdDescription.f90:218-242).  Accuracies are given as
! exponents, 10**(-value) (potentiate='-', src/input/mod_blockdata.f90:1942-1946).
module hinput
    use hkinds, only: rglu, iglu
    implicit none
    private

    type, public :: hcontrol
        character(len=64)  :: methods(8)
        integer(kind=iglu) :: nmethods = 0
        character(len=32)  :: coulomb = 'ohno-klopman'
        ! scf
        integer(kind=iglu) :: scfMaxIters = 400
        real(kind=rglu)    :: scfAccuracy = 1e-12_rglu
        real(kind=rglu)    :: scfStep = 0.05_rglu
        ! coupled-cluster
        character(len=8)   :: projType = '2-1'
        integer(kind=iglu) :: ccMaxIters = 2000
        real(kind=rglu)    :: ccAccuracy = 1e-12_rglu
        real(kind=rglu)    :: t1Step = 1, t2Step = 1
        logical            :: diisEnabled = .true.
        integer(kind=iglu) :: diisSteps = 20
        character(len=8)   :: device = 'gpu'
        character(len=8)   :: precision = 'fp64'
        character(len=16)  :: guess = 'reference'  ! reference | mp2 | learned
        character(len=256) :: guessNetwork = 'helios_amd/data/guess_mlp.f32'
        character(len=8)   :: driver = 'host'       ! host: the host iterator; device: hcc_solve
        ! iteration
        real(kind=rglu)    :: divergence = 1e3_rglu
        integer(kind=iglu) :: stagnationAfter = 200
        real(kind=rglu)    :: stagnationPercent = 60
        ! molecule block text
        character(len=256), allocatable :: molLines(:)
        integer(kind=iglu) :: nMolLines = 0
    end type hcontrol

    public :: readInput

contains

    subroutine readInput(fname, ctl)
        character(len=*), intent(in)  :: fname
        type(hcontrol),   intent(out) :: ctl
        character(len=:), allocatable :: text
        character(len=1024)           :: line
        integer :: unit, ios, p, q, depth, k0
        character(len=64) :: bname

        text = ''
        open (newunit=unit, file=fname, status='old', action='read', iostat=ios)
        if (ios /= 0) then
            write (*, '(A,A)') ' Error: cannot open input file ', trim(fname)
            error stop 'input'
        end if
        do
            read (unit, '(A)', iostat=ios) line
            if (ios /= 0) exit
            p = index(line, '#')
            if (p > 0) line(p:) = ' '
            text = text // trim(line) // achar(10)
        end do
        close (unit)
        allocate (ctl%molLines(4096))
        ctl%methods = ''
        p = 1
        do
            q = index(text(p:), '%')
            if (q == 0) exit
            p = p + q
            q = index(text(p:), '{')
            if (q == 0) exit
            bname = adjustl(text(p:p+q-2))
            k0 = p + q
            depth = 1
            q = k0
            do while (q <= len(text) .and. depth > 0)
                if (text(q:q) == '{') depth = depth + 1
                if (text(q:q) == '}') depth = depth - 1
                q = q + 1
            end do
            call block(trim(bname), text(k0:q-2), ctl)
            p = q
        end do
    end subroutine readInput

    subroutine block(bname, body, ctl)
        character(len=*), intent(in)    :: bname, body
        type(hcontrol),   intent(inout) :: ctl
        integer :: p, q, e
        character(len=256) :: item, key, val

        if (bname == 'molecule') then
            p = 1
            do while (p <= len(body))
                q = index(body(p:), achar(10))
                if (q == 0) q = len(body) - p + 2
                item = body(p:p+q-2)
                if (len_trim(item) > 0) then
                    ctl%nMolLines = ctl%nMolLines + 1
                    ctl%molLines(ctl%nMolLines) = adjustl(item)
                end if
                p = p + q
            end do
            return
        end if
        p = 1
        do while (p <= len(body))
            q = index(body(p:), ',')
            if (q == 0) q = len(body) - p + 2
            item = body(p:p+q-2)
            e = index(item, '=')
            if (e > 0) then
                key = adjustl(strip(item(1:e-1)))
                val = adjustl(strip(item(e+1:)))
                call setKey(bname, trim(key), trim(val), ctl)
            end if
            p = p + q
        end do
    end subroutine block

    function strip(s) result(r)
        character(len=*), intent(in) :: s
        character(len=len(s)) :: r
        integer :: k
        r = s
        do k = 1, len(r)
            if (iachar(r(k:k)) < 32) r(k:k) = ' '
        end do
    end function strip

    real(kind=rglu) function rval(s)
        character(len=*), intent(in) :: s
        read (s, *) rval
    end function rval

    integer(kind=iglu) function ival(s)
        character(len=*), intent(in) :: s
        read (s, *) ival
    end function ival

    logical function lval(s)
        character(len=*), intent(in) :: s
        lval = (s == 'true' .or. s == 'yes' .or. s == '1')
    end function lval

    subroutine setKey(bname, key, val, ctl)
        character(len=*), intent(in)    :: bname, key, val
        type(hcontrol),   intent(inout) :: ctl
        integer :: p, q

        select case (bname // ':' // key)
            case ('general:method')
                p = 1
                do
                    q = index(val(p:), ' ')
                    ctl%nmethods = ctl%nmethods + 1
                    if (q == 0) then
                        ctl%methods(ctl%nmethods) = val(p:)
                        exit
                    end if
                    ctl%methods(ctl%nmethods) = val(p:p+q-2)
                    p = p + q
                    if (p > len(val)) exit
                end do
            case ('parametrization:coulomb-integrals'); ctl%coulomb = val
            case ('scf:max-iterations');                ctl%scfMaxIters = ival(val)
            case ('scf:accuracy');                      ctl%scfAccuracy = 10.0_rglu**(-rval(val))
            case ('scf:iter-step');                     ctl%scfStep = rval(val)
            case ('coupled-cluster:projection-type');   ctl%projType = val
            case ('coupled-cluster:max-iterations');    ctl%ccMaxIters = ival(val)
            case ('coupled-cluster:accuracy');          ctl%ccAccuracy = 10.0_rglu**(-rval(val))
            case ('coupled-cluster:t1-step');           ctl%t1Step = rval(val)
            case ('coupled-cluster:t2-step');           ctl%t2Step = rval(val)
            case ('coupled-cluster:diis-enabled');      ctl%diisEnabled = lval(val)
            case ('coupled-cluster:diis-steps');        ctl%diisSteps = ival(val)
            case ('coupled-cluster:device');            ctl%device = val
            case ('coupled-cluster:precision');         ctl%precision = val
            case ('coupled-cluster:guess');             ctl%guess = val
            case ('coupled-cluster:guess-network');     ctl%guessNetwork = val
            case ('coupled-cluster:driver');            ctl%driver = val
            case ('iteration:divergence-threshold');    ctl%divergence = 10.0_rglu**rval(val)
            case ('iteration:stagnation-check-after');  ctl%stagnationAfter = ival(val)
            case ('iteration:stagnation-threshold');    ctl%stagnationPercent = rval(val)
            case default
                ! keys of other blocks and tasks are accepted and ignored
        end select
    end subroutine setKey

end module hinput
