! rgf_shim.f90 -- ISO_C_BINDING shim of the MI355X RGF engine (rgf.h) for Fortran drivers.
!
! green_RGF_RS_gpu has exactly the argument list of the reference routine
! green_RGF_RS (src/green_rgf.f90:3950-3958), including the optional
! off-diagonal outputs GRi1/GLi1/GGi1, so a driver replaces
!     call green_RGF_RS(TEMP,nm,nx,E,mu,Hii,H1i,sigL,sigG,sigR,ndens,pdens,ldos,tr,tre,cur,...)
! by the same call to green_RGF_RS_gpu.  green_RGF_RS_energies_gpu replaces a
! whole `do ie` energy loop (src/green_rgf.f90:3174-3180) by one batched solve
! over the caller's (nm,nm,nx,nen) arrays.  Both use the reference's
! hard-coded parameters unless an rgf_params value is passed.
module rgf_shim
  use iso_c_binding
  implicit none
  private
  public :: rgf_params, rgf_params_reference, green_RGF_RS_gpu, green_RGF_RS_energies_gpu, green_RGF_CMS_gpu
  public :: RGF_MODE_COMPAT, RGF_MODE_PHYSICAL, rgf_last_error_f

  integer(c_int), parameter :: RGF_MODE_COMPAT = 0, RGF_MODE_PHYSICAL = 1

  type, bind(C) :: rgf_params
    real(c_double) :: eta_dev
    real(c_double) :: eta_lead
    real(c_double) :: dec_tol
    integer(c_int) :: dec_nmax
    real(c_double) :: boltz
    integer(c_int) :: mode
  end type rgf_params

  interface
    subroutine rgf_params_reference(p) bind(C, name="rgf_params_reference")
      import :: rgf_params
      type(rgf_params), intent(out) :: p
    end subroutine

    function rgf_last_error_c() bind(C, name="rgf_last_error") result(s)
      import :: c_ptr
      type(c_ptr) :: s
    end function

    integer(c_int) function rgf_rs_ex_c(p, temp, nm, nx, E, mu, Hii, H1i, sigL, sigG, sigR, &
        ndens, pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1) bind(C, name="rgf_rs_ex")
      import :: c_int, c_double, c_ptr, rgf_params
      type(rgf_params), intent(in) :: p
      real(c_double), intent(in) :: temp(2), mu(2)
      integer(c_int), value :: nm, nx
      real(c_double), value :: E
      type(c_ptr), value :: Hii, H1i, sigL, sigG, sigR, ndens, pdens, ldos
      real(c_double), intent(inout) :: tr, tre
      type(c_ptr), value :: cur, GRi1, GLi1, GGi1
    end function

    integer(c_int) function rgf_cms_ex_c(p, temp, nm, nx, E, mu, Sii, Hii, H1i, sigL, sigG, sigR, &
        ndens, pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1) bind(C, name="rgf_cms_ex")
      import :: c_int, c_double, c_ptr, rgf_params
      type(rgf_params), intent(in) :: p
      real(c_double), intent(in) :: temp(2), mu(2)
      integer(c_int), value :: nm, nx
      real(c_double), value :: E
      type(c_ptr), value :: Sii, Hii, H1i, sigL, sigG, sigR, ndens, pdens, ldos
      real(c_double), intent(inout) :: tr, tre
      type(c_ptr), value :: cur, GRi1, GLi1, GGi1
    end function

    integer(c_int) function rgf_rs_energies_c(p, temp, nm, nx, nE, E, mu, Hii, H1i, sigL, sigG, sigR, &
        sig_per_energy, ndens, pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1, info) bind(C, name="rgf_rs_energies")
      import :: c_int, c_double, c_ptr, rgf_params
      type(rgf_params), intent(in) :: p
      real(c_double), intent(in) :: temp(2), mu(2)
      integer(c_int), value :: nm, nx, nE, sig_per_energy
      real(c_double), intent(in) :: E(*)
      type(c_ptr), value :: Hii, H1i, sigL, sigG, sigR, ndens, pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1, info
    end function
  end interface

contains

  function rgf_last_error_f() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: n
    type(c_ptr) :: p
    p = rgf_last_error_c()
    call c_f_pointer(p, cs, [4096])
    n = 0
    do while (n < 4096)
      if (cs(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function

  ! Same arguments as green_RGF_RS (src/green_rgf.f90:3950-3958); stops on an
  ! engine error (invalid sizes, device failure).  Numerical failures follow
  ! the reference (zeroed singular blocks, bare-surface fallback).
  subroutine green_RGF_RS_gpu(TEMP, nm, nx, E, mu, Hii, H1i, sigma_lesser_ph, sigma_greater_ph, sigma_r_ph, &
      ndens, pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1, params)
    integer, intent(in) :: nm, nx
    real(8), intent(in) :: E, mu(2), temp(2)
    complex(8), intent(in), dimension(nm, nm, nx), target :: sigma_lesser_ph, sigma_greater_ph, sigma_r_ph
    complex(8), intent(in), dimension(nm, nm, nx), target :: Hii, H1i
    complex(8), intent(inout), dimension(nm, nm, nx), target :: ldos, ndens, pdens
    real(8), intent(inout), dimension(nm, nm, nx), target :: cur
    complex(8), intent(inout), dimension(nm, nm, nx), optional, target :: GRi1, GLi1, GGi1
    real(8), intent(inout) :: tr, tre
    type(rgf_params), intent(in), optional :: params
    type(rgf_params) :: p
    type(c_ptr) :: pr, pl, pg
    integer(c_int) :: rc
    if (present(params)) then
      p = params
    else
      call rgf_params_reference(p)
    end if
    pr = c_null_ptr; pl = c_null_ptr; pg = c_null_ptr
    if (present(GRi1)) pr = c_loc(GRi1)
    if (present(GLi1)) pl = c_loc(GLi1)
    if (present(GGi1)) pg = c_loc(GGi1)
    rc = rgf_rs_ex_c(p, temp, int(nm, c_int), int(nx, c_int), E, mu, c_loc(Hii), c_loc(H1i), &
        c_loc(sigma_lesser_ph), c_loc(sigma_greater_ph), c_loc(sigma_r_ph), &
        c_loc(ndens), c_loc(pdens), c_loc(ldos), tr, tre, c_loc(cur), pr, pl, pg)
    if (rc /= 0) then
      print *, 'green_RGF_RS_gpu: ', rgf_last_error_f()
      stop 1
    end if
  end subroutine

  ! Same arguments as green_RGF_CMS (src/green_rgf.f90:4167-4175): the overlap
  ! variant (Sii before Hii).
  subroutine green_RGF_CMS_gpu(TEMP, nm, nx, E, mu, Sii, Hii, H1i, sigma_lesser_ph, sigma_greater_ph, sigma_r_ph, &
      ndens, pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1, params)
    integer, intent(in) :: nm, nx
    real(8), intent(in) :: E, mu(2), temp(2)
    complex(8), intent(in), dimension(nm, nm, nx), target :: sigma_lesser_ph, sigma_greater_ph, sigma_r_ph
    complex(8), intent(in), dimension(nm, nm, nx), target :: Sii, Hii, H1i
    complex(8), intent(inout), dimension(nm, nm, nx), target :: ldos, ndens, pdens
    real(8), intent(inout), dimension(nm, nm, nx), target :: cur
    complex(8), intent(inout), dimension(nm, nm, nx), optional, target :: GRi1, GLi1, GGi1
    real(8), intent(inout) :: tr, tre
    type(rgf_params), intent(in), optional :: params
    type(rgf_params) :: p
    type(c_ptr) :: pr, pl, pg
    integer(c_int) :: rc
    if (present(params)) then
      p = params
    else
      call rgf_params_reference(p)
    end if
    pr = c_null_ptr; pl = c_null_ptr; pg = c_null_ptr
    if (present(GRi1)) pr = c_loc(GRi1)
    if (present(GLi1)) pl = c_loc(GLi1)
    if (present(GGi1)) pg = c_loc(GGi1)
    rc = rgf_cms_ex_c(p, temp, int(nm, c_int), int(nx, c_int), E, mu, c_loc(Sii), c_loc(Hii), c_loc(H1i), &
        c_loc(sigma_lesser_ph), c_loc(sigma_greater_ph), c_loc(sigma_r_ph), &
        c_loc(ndens), c_loc(pdens), c_loc(ldos), tr, tre, c_loc(cur), pr, pl, pg)
    if (rc /= 0) then
      print *, 'green_RGF_CMS_gpu: ', rgf_last_error_f()
      stop 1
    end if
  end subroutine

  ! Batched replacement of `do ie=1,nen; call green_RGF_RS(...(:,:,:,ie)...); enddo`.
  ! Arrays carry a trailing energy index; sigma arrays are per energy.
  subroutine green_RGF_RS_energies_gpu(TEMP, nm, nx, nen, En, mu, Hii, H1i, sigma_lesser, sigma_greater, sigma_r, &
      ndens, pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1, params, info)
    integer, intent(in) :: nm, nx, nen
    real(8), intent(in) :: En(nen), mu(2), temp(2)
    complex(8), intent(in), dimension(nm, nm, nx, nen), target :: sigma_lesser, sigma_greater, sigma_r
    complex(8), intent(in), dimension(nm, nm, nx), target :: Hii, H1i
    complex(8), intent(inout), dimension(nm, nm, nx, nen), target :: ldos, ndens, pdens
    real(8), intent(inout), dimension(nm, nm, nx, nen), target :: cur
    real(8), intent(inout), dimension(nen), target :: tr, tre
    complex(8), intent(inout), dimension(nm, nm, nx, nen), optional, target :: GRi1, GLi1, GGi1
    type(rgf_params), intent(in), optional :: params
    integer(c_int), intent(out), dimension(4, nen), optional, target :: info
    type(rgf_params) :: p
    type(c_ptr) :: pr, pl, pg, pin
    integer(c_int) :: rc
    if (present(params)) then
      p = params
    else
      call rgf_params_reference(p)
    end if
    pr = c_null_ptr; pl = c_null_ptr; pg = c_null_ptr; pin = c_null_ptr
    if (present(GRi1)) pr = c_loc(GRi1)
    if (present(GLi1)) pl = c_loc(GLi1)
    if (present(GGi1)) pg = c_loc(GGi1)
    if (present(info)) pin = c_loc(info)
    rc = rgf_rs_energies_c(p, temp, int(nm, c_int), int(nx, c_int), int(nen, c_int), En, mu, c_loc(Hii), c_loc(H1i), &
        c_loc(sigma_lesser), c_loc(sigma_greater), c_loc(sigma_r), 1_c_int, c_loc(ndens), c_loc(pdens), c_loc(ldos), &
        c_loc(tr), c_loc(tre), c_loc(cur), pr, pl, pg, pin)
    if (rc /= 0) then
      print *, 'green_RGF_RS_energies_gpu: ', rgf_last_error_f()
      stop 1
    end if
  end subroutine

end module rgf_shim
