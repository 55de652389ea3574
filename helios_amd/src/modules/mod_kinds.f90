! Kinds and constants of the Fortran host (the reference's glob kinds,
! src/modules/mod_global.f90:62-71, and constants, src/modules/mod_helios.f90:41-43).
module hkinds
    use iso_c_binding, only: c_double, c_int
    implicit none
    integer, parameter :: rglu = c_double
    integer, parameter :: iglu = c_int
    real(kind=rglu), parameter :: BohrRadius    = 0.52917721067_rglu   ! Angstrom
    real(kind=rglu), parameter :: HartreeEnergy = 27.21138602_rglu     ! eV
end module hkinds
